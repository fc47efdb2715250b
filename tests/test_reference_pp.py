"""The oracle's projector tables, list vectors and forces (oracle/kukan_oracle_pp.c) against the
reference's own pseudocalc8f.F90 and force8f.F90 compiled unchanged into
oracle/_ref/libkukan_ref.so (one MPI rank, one thread).  Same operand and summation order on
both sides: every output agrees to the last bit.  The error paths (which abort the reference
process) are checked on the oracle alone (tests/test_oracle_pp.py).  Skipped where the reference
library was not built."""
import numpy as np
import pytest

from rspacedist_amd import pp_synth as P

ref = pytest.importorskip("oracle.ref")
pytestmark = pytest.mark.skipif(not ref.available(), reason="reference library not built")


def both(orc, name, kw):
    out = []
    for f in (orc.call, ref.call):
        a = {k: (v.copy(order="F") if isinstance(v, np.ndarray) else v) for k, v in kw.items()}
        r = f(name, **a)
        assert r["rc"] == 0
        out.append((a, r))
    return out


def same(o, r, keys):
    for k in keys:
        x, y = o[k], r[k]
        if x.dtype.kind == "f":
            assert np.array_equal(x.ravel(order="F").view(np.uint64), y.ravel(order="F").view(np.uint64)), k
        else:
            assert np.array_equal(x, y), k


def weights(orc, nfdg, npmesh):
    w = np.zeros(2 * nfdg * npmesh)
    orc.call("pseudocalc_01a", nfdg=nfdg, npmesh=npmesh, wxyz=w)
    return w


@pytest.mark.parametrize("nfdg", range(1, 9))
def test_dg_weights_bitwise(orc, nfdg):
    for npmesh in (1, 2, 3, 4, 5):
        (o, _), (r, _) = both(orc, "pseudocalc_01a", dict(nfdg=nfdg, npmesh=npmesh, wxyz=np.zeros(2 * nfdg * npmesh)))
        s = npmesh % 2  # the first weight is not defined for odd npmesh
        assert np.array_equal(o["wxyz"][s:].view(np.uint64), r["wxyz"][s:].view(np.uint64)), npmesh


LIST_CASES = [  # nperi, nmesh, nfdg, edge, tight box
    (0, 2, 3, True, True), (1, 2, 3, True, True), (2, 2, 4, True, True), (3, 2, 3, True, True),
    (3, 1, 1, False, True), (0, 3, 2, False, False), (3, 2, 2, True, False), (2, 1, 4, False, True)]


def _cell(nperi, nmesh, nfdg, edge, tight, seed=0, **kw):
    d, x = P.cell(seed=seed + 10 * nperi + nfdg, nperi=nperi, nmesh=nmesh, nfdg=nfdg, edge=edge, **kw)
    if tight:  # a box that holds every point whose neighbourhood meets the sphere
        npx = int(np.ceil(d["radial"].max() * d["psctoff"] / d["dx"])) + nfdg + 1
        d.update(npxmax=npx, npymax=npx, npzmax=npx, num_list=(2 * npx) ** 3, num_list_d=(2 * npx * nmesh) ** 3)
    return d, x


@pytest.mark.parametrize("nperi,nmesh,nfdg,edge,tight", LIST_CASES)
def test_list_vectors_bitwise(orc, nperi, nmesh, nfdg, edge, tight):
    d, _ = _cell(nperi, nmesh, nfdg, edge, tight)
    (o, _), (r, _) = both(orc, "pseudocalc_01b", {**d, **P.list_outputs(d)})
    same(o, r, list(P.list_outputs(d)))
    assert o["natinf"].min() > 0 and o["natinfd"].min() > 0
    if edge and nperi == 0:  # the face atom's sphere is cut by the isolated boundary
        assert o["natinf"][0] < o["natinf"].max()


def _lists(orc, d):
    o = {**d, **P.list_outputs(d)}
    orc.call("pseudocalc_01b", **o)
    return o


PP07_CASES = [  # nperi, npmesh, nfdg, nprj per species, nfiltyp
    (3, 2, 3, (18, 1, 4), 1), (0, 2, 4, (2, 5, 7), 1), (3, 3, 2, (8, 9, 10), 0), (1, 1, 1, (12, 13, 14), 1),
    (2, 4, 2, (15, 17, 3), 1), (3, 2, 4, (9, 18, 1), 0)]


@pytest.mark.parametrize("nperi,npmesh,nfdg,nprj,nfiltyp", PP07_CASES)
def test_projector_table_bitwise(orc, nperi, npmesh, nfdg, nprj, nfiltyp):
    d, x = P.cell(seed=nperi + 3 * npmesh, nperi=nperi, nfdg=nfdg, npmesh=npmesh, nprj=nprj,
                  rc=(1.05, 0.85, 0.95), natom=4, nqmx=18)
    if nperi == 3:
        # put atom 0 on a fine-grid point: the r <= eps branch (l = 0 limits of the Bessel sums)
        dp = d["dx"] / npmesh
        d["atx"][0] = -d["xmax"] + 7.5 * dp
        d["aty"][0] = -d["ymax"] + 9.5 * dp
        d["atz"][0] = -d["zmax"] + 8.5 * dp
    L = _lists(orc, d)
    a = P.pp07_args(d, x, L, weights(orc, nfdg, npmesh))
    a["nfiltyp"] = nfiltyp
    (o, _), (r, _) = both(orc, "pseudocalc_07", a)
    same(o, r, ["vnlocp"])
    v = o["vnlocp"]
    assert np.abs(v).max() > 1e-4
    if 3 in nprj:  # nprj values without a case give zero projectors, as in the reference
        sp = nprj.index(3) + 1
        for na in np.nonzero(d["indspe"] == sp)[0]:
            assert not v[:, :, L["naps"][na] - 1].any()


FORCE_CASES = [  # nperi, jellium, nspv, bias
    (0, False, 1, (0.0, 0.0, 0.0)), (1, False, 2, (0.01, 0.0, -0.02)), (2, True, 2, (0.0, 0.03, 0.0)),
    (3, False, 2, (0.02, -0.01, 0.01)), (3, True, 1, (0.0, 0.0, 0.0))]


@pytest.mark.parametrize("nperi,jel,nspv,bias", FORCE_CASES)
def test_force_bitwise(orc, nperi, jel, nspv, bias):
    d, x = P.cell(seed=40 + nperi, nperi=nperi, natom=4, rc=(0.9, 0.8))
    L = _lists(orc, d)
    a = P.force_args(d, x, L, seed=nperi, nspv=nspv, jel=jel, bias=bias)
    (o, ro), (r, rr) = both(orc, "force", a)
    same(o, r, ["fatx", "faty", "fatz"])
    assert np.float64(ro["fjel"]).view(np.uint64) == np.float64(rr["fjel"]).view(np.uint64)
    assert (ro["fjel"] != 0.0) == jel
