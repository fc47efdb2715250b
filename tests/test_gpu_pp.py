"""Projector tables, list vectors and forces through the GPU library (rspacedist_amd/csrc/
projector.hip, force.hip) against the oracle (itself bit-identical to the reference's
pseudocalc8f.F90 / force8f.F90, tests/test_reference_pp.py):

- pseudocalc_01a and pseudocalc_01b bit for bit (weights and integer list vectors), from host
  and from device arrays, including the error messages;
- pseudocalc_07 (vnlocp) within 1e-12 of max|vnlocp| (device sin/cos/exp differ from glibc's in
  the last bits; the filter sums run in the reference's order);
- force() within 1e-10 of the largest force (the sums run in another, fixed order);
- a 128^3 cell with 32 atoms: lists bit for bit, vnlocp and (for four atoms) forces as above."""
import numpy as np
import pytest

from rspacedist_amd import pp_synth as P
from tests.test_reference_pp import FORCE_CASES, LIST_CASES, PP07_CASES, _cell

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def kk():
    from rspacedist_amd import kukan
    kukan.init(0)
    yield kukan


def _dev(kk, a, skip=()):
    return {k: (kk.DeviceArray.from_numpy(v) if isinstance(v, np.ndarray) and k not in skip else v)
            for k, v in a.items()}


def _host(v):
    return v.numpy() if hasattr(v, "numpy") else v


def weights(orc, nfdg, npmesh):
    w = np.zeros(2 * nfdg * npmesh)
    orc.call("pseudocalc_01a", nfdg=nfdg, npmesh=npmesh, wxyz=w)
    return w


@pytest.mark.parametrize("on_device", [False, True])
def test_dg_weights_bitwise(kk, orc, on_device):
    for nfdg in range(1, 9):
        for npmesh in (1, 2, 3, 4, 5):
            w = np.full(2 * nfdg * npmesh, 7.0)
            dw = kk.DeviceArray.from_numpy(w) if on_device else w
            kk.call("pseudocalc_01a", nfdg=nfdg, npmesh=npmesh, wxyz=dw)
            g = _host(dw)
            o = weights(orc, nfdg, npmesh)
            assert np.array_equal(g.view(np.uint64), o.view(np.uint64)), (nfdg, npmesh)


def _lists_both(kk, orc, d, on_device):
    o = {**d, **P.list_outputs(d)}
    orc.call("pseudocalc_01b", **o)
    g = {**d, **P.list_outputs(d)}
    if on_device:
        gd = _dev(kk, g)
        kk.call("pseudocalc_01b", **gd)
        g = {k: _host(v) for k, v in gd.items()}
    else:
        kk.call("pseudocalc_01b", **g)
    return g, o


@pytest.mark.parametrize("case", LIST_CASES)
@pytest.mark.parametrize("on_device", [False, True])
def test_list_vectors_bitwise(kk, orc, case, on_device):
    d, _ = _cell(*case)
    g, o = _lists_both(kk, orc, d, on_device)
    for k in P.list_outputs(d):
        np.testing.assert_array_equal(g[k], o[k], err_msg=k)


def test_list_errors_carry_reference_messages(kk):
    d, _ = P.cell(seed=1, nperi=3)
    for key, msg in (("num_list", "num_list is too small"), ("num_list_d", "num_list_d is too small"),
                     ("num_ppcell", "num_ppcell is too small"), ("num_atcell", "num_atcell is too small"),
                     ("num_ppcell_d", "num_ppcell_d is too small")):
        e = dict(d)
        e[key] = 2 if key in ("num_ppcell", "num_atcell", "num_ppcell_d") else 10
        with pytest.raises(kk.KukanError, match=msg):
            kk.call("pseudocalc_01b", **e, **P.list_outputs(e))
    with pytest.raises(kk.KukanError, match="nfdg should be between 1 and 8"):
        kk.call("pseudocalc_01a", nfdg=0, npmesh=2, wxyz=np.zeros(4))


@pytest.mark.parametrize("case", PP07_CASES)
@pytest.mark.parametrize("on_device", [False, True])
def test_projector_table(kk, orc, case, on_device):
    nperi, npmesh, nfdg, nprj, nfiltyp = case
    d, x = P.cell(seed=nperi + 3 * npmesh, nperi=nperi, nfdg=nfdg, npmesh=npmesh, nprj=nprj,
                  rc=(1.05, 0.85, 0.95), natom=4, nqmx=18)
    if nperi == 3:  # an atom on a fine-grid point (the r <= eps branch)
        dp = d["dx"] / npmesh
        d["atx"][0] = -d["xmax"] + 7.5 * dp
        d["aty"][0] = -d["ymax"] + 9.5 * dp
        d["atz"][0] = -d["zmax"] + 8.5 * dp
    L = {**d, **P.list_outputs(d)}
    orc.call("pseudocalc_01b", **L)
    a = P.pp07_args(d, x, L, weights(orc, nfdg, npmesh))
    a["nfiltyp"] = nfiltyp
    ref = {k: (v.copy(order="F") if isinstance(v, np.ndarray) else v) for k, v in a.items()}
    orc.call("pseudocalc_07", **ref)
    if on_device:
        ad = _dev(kk, a, skip=("vtmp_pp", "vtmp_pp1", "vtmp_pp2", "vnlspd"))
        kk.call("pseudocalc_07", **ad)
        got = ad["vnlocp"].numpy()
    else:
        kk.call("pseudocalc_07", **a)
        got = a["vnlocp"]
    scale = np.abs(ref["vnlocp"]).max()
    err = np.abs(got - ref["vnlocp"]).max()
    print(f"vnlocp {case}: max |diff| / max|vnlocp| = {err / scale:.2e}")
    assert err <= 1e-12 * scale
    # entries beyond natinf are zero in both
    for na in range(d["natom"]):
        col = L["naps"][na] - 1
        assert not got[L["natinf"][na]:, :, col].any()


@pytest.mark.parametrize("case", FORCE_CASES)
@pytest.mark.parametrize("on_device", [False, True])
def test_forces(kk, orc, case, on_device):
    nperi, jel, nspv, bias = case
    d, x = P.cell(seed=40 + nperi, nperi=nperi, natom=4, rc=(0.9, 0.8))
    L = {**d, **P.list_outputs(d)}
    orc.call("pseudocalc_01b", **L)
    a = P.force_args(d, x, L, seed=nperi, nspv=nspv, jel=jel, bias=bias)
    ref = {k: (v.copy(order="F") if isinstance(v, np.ndarray) else v) for k, v in a.items()}
    ro = orc.call("force", **ref)
    if on_device:
        ad = _dev(kk, a)
        rg = kk.call("force", **ad)
        got = {k: ad[k].numpy() for k in ("fatx", "faty", "fatz")}
    else:
        rg = kk.call("force", **a)
        got = a
    F = np.stack([ref["fatx"], ref["faty"], ref["fatz"]])
    G = np.stack([got["fatx"], got["faty"], got["fatz"]])
    scale = np.abs(F).max()
    print(f"force {case}: max |diff| / max|F| = {np.abs(G - F).max() / scale:.2e}, fjel {rg['fjel']} vs {ro['fjel']}")
    assert np.abs(G - F).max() <= 1e-10 * scale
    assert abs(rg["fjel"] - ro["fjel"]) <= 1e-10 * max(abs(ro["fjel"]), 1e-300) + (0 if jel else 0.0)
    assert (rg["fjel"] != 0.0) == jel


def test_full_cell_lists_projectors_forces(kk, orc):
    """128^3 coarse grid (h = 0.2), 32 atoms of two species (nprj 9 and 18), nfdg 4, npmesh 2,
    nmesh 2: the list vectors and vnlocp on device arrays against the oracle, and the forces on
    four of the atoms."""
    d, x = P.cell(seed=77, nperi=3, nxmax=64, nymax=64, nzmax=64, h=0.2, nmesh=2, nfdg=4, npmesh=2,
                  natom=32, nprj=(9, 18), rc=(1.0, 0.8), nqmx=20)
    npx = int(np.ceil(d["radial"].max() * d["psctoff"] / d["dx"])) + d["nfdg"] + 1
    d.update(npxmax=npx, npymax=npx, npzmax=npx, num_list=(2 * npx) ** 3, num_list_d=(2 * npx * 2) ** 3)
    g, o = _lists_both(kk, orc, d, True)
    for k in P.list_outputs(d):
        np.testing.assert_array_equal(g[k], o[k], err_msg=k)
    a = P.pp07_args(d, x, o, weights(orc, d["nfdg"], x["npmesh"]))
    ref = {k: (v.copy(order="F") if isinstance(v, np.ndarray) else v) for k, v in a.items()}
    orc.call("pseudocalc_07", **ref)
    ad = _dev(kk, a, skip=("vtmp_pp", "vtmp_pp1", "vtmp_pp2", "vnlspd"))
    kk.call("pseudocalc_07", **ad)
    v = ad["vnlocp"].numpy()
    scale = np.abs(ref["vnlocp"]).max()
    print(f"full cell vnlocp: max |diff| / max = {np.abs(v - ref['vnlocp']).max() / scale:.2e}")
    assert np.abs(v - ref["vnlocp"]).max() <= 1e-12 * scale
    # forces on 4 of the atoms (the oracle's core-correction loop visits the whole dense grid
    # 27 times per atom and spin)
    d4 = dict(d, natom=4, num_atcell=4, num_ppcell=4, num_ppcell_d=4, indspe=d["indspe"][:4].copy(order="F"),
              atx=d["atx"][:4].copy(), aty=d["aty"][:4].copy(), atz=d["atz"][:4].copy())
    o = {**d4, **P.list_outputs(d4)}
    orc.call("pseudocalc_01b", **o)
    fa = P.force_args(d4, x, o, seed=5, nspv=1, jel=False, bias=(0.0, 0.0, 0.01), new_pw=3, new_rs=1)
    fr = {k: (v.copy(order="F") if isinstance(v, np.ndarray) else v) for k, v in fa.items()}
    orc.call("force", **fr)
    fd = _dev(kk, fa)
    kk.call("force", **fd)
    F = np.stack([fr["fatx"], fr["faty"], fr["fatz"]])
    G = np.stack([fd[k].numpy() for k in ("fatx", "faty", "fatz")])
    print(f"full cell forces: max |diff| / max|F| = {np.abs(G - F).max() / np.abs(F).max():.2e}")
    assert np.abs(G - F).max() <= 1e-10 * np.abs(F).max()
