"""The oracle's complex (Bloch) path against the reference's own routines (CPU).

oracle/_ref/libkukan_ref.so holds the reference's kscg_c, rayleighritz_c,
orthogonalization_c(_01), nonlocaloperation_c_01/_02, kslaplacian_c, ksprecondition_c,
overlap_finitedifference_c and scf_charge_smt_c, compiled unchanged (oracle/Makefile).  Both
are called with the same arguments on the same seeded inputs, at a generic k-point, with
S = 1 and S != 1.  Index work agrees exactly; FP64 results to rounding, the iterative routines
within the growth of rounding over their iterations.  The Hermitian eigensolver is the one
place where the two may legitimately differ: each eigenvector's phase is the solver's choice
(LAPACK zheev in the reference, Jacobi in the oracle), so after a Rayleigh-Ritz step bands are
compared after aligning each one's phase, and quantities that do not depend on it (eigenvalues,
densities) directly.  Skipped where the reference library was not built."""
import numpy as np
import pytest

from rspacedist_amd import synth
from tests import helpers as H
from tests.test_oracle_bloch import K, halo_c

ref = pytest.importorskip("oracle.ref")
pytestmark = pytest.mark.skipif(not ref.available(), reason="reference library not built")

QS = [0.0, 2.0]
ZK = ("sveccm", "ssvcm", "hsvcm", "cspsep")


def both(orc, name, arrays, **kw):
    out = []
    for f in (orc.call, ref.call):
        a = {k: (v.copy(order="F") if isinstance(v, np.ndarray) else v) for k, v in arrays.items()}
        a["_ret"] = f(name, **kw, **a)
        out.append(a)
    return out


def close(o, r, k, rel):
    scale = max(np.abs(r[k]).max(), 1e-300)
    err = np.abs(o[k] - r[k]).max() / scale
    assert err <= rel, f"{k}: relative difference {err:.2e} > {rel:.0e}"


def align(a, b):
    """b's bands multiplied by the unit phase that best matches a's (per column)."""
    ph = np.einsum("ij,ij->j", np.conj(b), a)
    ph = ph / np.maximum(np.abs(ph), 1e-300)
    return b * ph[None, :]


def init_c(orc, s, m=None):
    m = s.nbands if m is None else m
    psi = synth.random_bands_c(s, nbands=m)
    st = dict(sveccm=psi, ssvcm=np.zeros_like(psi),
              cspsep=np.zeros((9, s.natom, m), dtype=np.complex128, order="F"))
    orc.call("orthogonalization_c", **synth.bloch_args(s, K, neigmx=m), nflag=1, **st)
    st.update(hsvcm=np.zeros_like(psi), sval=np.zeros(m), residual_states=np.zeros(m))
    return st


@pytest.mark.parametrize("nperi", [0, 1, 2, 3])
def test_halo_c_bit_exact(orc, nperi):
    rng = np.random.default_rng(nperi)
    v = np.zeros((9 + 8, 8 + 8, 7 + 8), dtype=np.complex128, order="F")
    v[4:13, 4:12, 4:11] = rng.standard_normal((9, 8, 7)) + 1j * rng.standard_normal((9, 8, 7))
    o, r = both(orc, "overlap_finitedifference_c", {"v_cc": v}, nperi=nperi, ncpa=9, ncpb=8, ncpc=7,
                nf=4, nf1=3, nsnd=4, ncol=1)
    np.testing.assert_array_equal(o["v_cc"], r["v_cc"])


@pytest.mark.parametrize("nf", [1, 2, 4, 6, 8])
def test_kslaplacian_c_equals_reference(orc, nf):
    n1, n2, n3, h = 12, 10, 9, (0.3, 0.35, 0.4)
    rng = np.random.default_rng(nf)
    u = rng.standard_normal((n1, n2, n3)) + 1j * rng.standard_normal((n1, n2, n3))
    arr = {"vcm": halo_c(u, nf, nf) if n1 == n2 == n3 else None}
    v = np.zeros((n1 + 2 * nf, n2 + 2 * nf, n3 + 2 * nf), dtype=np.complex128, order="F")
    v[nf:nf + n1, nf:nf + n2, nf:nf + n3] = u
    arr = {"vcm": v, "avcm": np.asfortranarray(rng.standard_normal((n1, n2, n3, 2)) + 0j),
           "workc": np.zeros((n1, n2, n3), dtype=np.complex128, order="F")}
    xm = (n1 * h[0] / 2, n2 * h[1] / 2, n3 * h[2] / 2)
    o, r = both(orc, "kslaplacian_c", arr, ncpx=n1, ncpy=n2, ncpz=n3, neigmx=2, ncol=1, nvef=1, nf=nf,
                dx=h[0], dy=h[1], dz=h[2], xmax=xm[0], ymax=xm[1], zmax=xm[2], skpx=K[0], skpy=K[1],
                skpz=K[2], veff=np.asfortranarray(rng.standard_normal((n1, n2, n3))), l=2)
    close(o, r, "avcm", 1e-14)
    close(o, r, "vcm", 1e-15)  # the input is multiplied by the phase in place


def test_ksprecondition_c_equals_reference(orc):
    rng = np.random.default_rng(4)
    u = rng.standard_normal((10, 10, 10)) + 1j * rng.standard_normal((10, 10, 10))
    o, r = both(orc, "ksprecondition_c", {"vcm": halo_c(u, 4, 1),
                "avcm": np.zeros((10, 10, 10), dtype=np.complex128, order="F")},
                ncpx=10, ncpy=10, ncpz=10, nf=4, ncol=1)
    np.testing.assert_array_equal(o["avcm"], r["avcm"])


@pytest.mark.parametrize("q", QS)
def test_nonlocal_c_equal_reference(orc, q):
    s = synth.make_system(12, 0.45, 3, 2, seed=5, q_scale=q)
    kw = dict(synth.bloch_args(s, K), neigmx=1, l=1)
    rng = np.random.default_rng(7)
    u = np.asfortranarray(rng.standard_normal(s.npts) + 1j * rng.standard_normal(s.npts))
    z = lambda *sh: np.zeros(sh, dtype=np.complex128, order="F")
    o, r = both(orc, "nonlocaloperation_c_01", {"vcm2": u, "cpsep": z(9, s.natom), "vcccm": z(s.num_list)}, **kw)
    close(o, r, "cpsep", 1e-14)
    for nhs, X in ((1, s.dij), (0, s.sss)):
        o2, r2 = both(orc, "nonlocaloperation_c_02", {"cpsep": o["cpsep"], "dijsss": X, "avcm": u.copy(),
                      "vcccm": z(s.num_list), "avc": z(s.num_list)}, **dict(kw, nhs=nhs, nso=0))
        close(o2, r2, "avcm", 1e-14)


@pytest.mark.parametrize("q", QS)
def test_orthogonalization_c_equals_reference(orc, q):
    s = synth.make_system(10, 0.5, 2, 40, seed=6, q_scale=q)  # 40 bands: the block path
    psi = synth.random_bands_c(s, smooth=False)
    arr = {"sveccm": psi, "ssvcm": np.zeros_like(psi),
           "cspsep": np.zeros((9, s.natom, 40), dtype=np.complex128, order="F")}
    o, r = both(orc, "orthogonalization_c", arr, **synth.bloch_args(s, K), nflag=1)
    for k in ("sveccm", "ssvcm", "cspsep"):
        close(o, r, k, 1e-12)


@pytest.mark.parametrize("q", QS)
def test_rayleighritz_c_equals_reference(orc, q):
    s = synth.make_system(10, 0.5, 3, 10, seed=33, q_scale=q)
    st = init_c(orc, s)
    arr = {k: st[k] for k in ("cspsep", "sveccm", "ssvcm", "hsvcm", "sval")}
    o, r = both(orc, "rayleighritz_c", arr, **synth.bloch_args(s, K), l_rrb=1)
    np.testing.assert_allclose(o["sval"], r["sval"], rtol=0, atol=1e-12)
    for k in ("sveccm", "ssvcm", "hsvcm"):
        rr = align(o[k], r[k]) if k == "sveccm" else r[k] * (np.einsum("ij,ij->j", np.conj(r["sveccm"]), o["sveccm"]) /
                                                              np.abs(np.einsum("ij,ij->j", np.conj(r["sveccm"]), o["sveccm"])))[None, :]
        np.testing.assert_allclose(o[k], rr, rtol=0, atol=1e-9 * np.abs(r[k]).max(), err_msg=k)


@pytest.mark.parametrize("q", QS)
@pytest.mark.parametrize("nsdmax", [1, 4, 12])  # 12 passes the reset at iteration 10
def test_kscg_c_iterates_equal_reference(orc, q, nsdmax):
    """Without a Rayleigh-Ritz step in between, the band-sequential complex CG from the same
    start follows the same iterates (no eigensolver phase enters)."""
    s = synth.make_system(12, 0.45, 3, 8, seed=35, q_scale=q)
    st = init_c(orc, s)
    arr = {k: st[k] for k in ("sval", "cspsep", "sveccm", "ssvcm", "residual_states", "hsvcm")}
    o, r = both(orc, "kscg_c", arr, **synth.bloch_args(s, K, nsdmax=nsdmax, epssd=1e-14))
    np.testing.assert_allclose(o["sval"], r["sval"], rtol=1e-11, atol=1e-11)
    for k in ZK:
        close(o, r, k, 1e-10)
    np.testing.assert_allclose(o["residual_states"], r["residual_states"], rtol=1e-7, atol=1e-16)
    assert o["_ret"]["ksitmax"] == r["_ret"]["ksitmax"]
    assert o["_ret"]["ksconv"] == r["_ret"]["ksconv"]


@pytest.mark.parametrize("q", QS)
def test_scf_diag_bloch_steps_equal_reference(orc, q):
    """scf_diag with nrc = 1 (Rayleigh-Ritz, kscg_c, reorder) over three steps: eigenvalues and
    the density sum |psi|^2 agree (band phases are the eigensolvers' choice)."""
    s = synth.make_system(10, 0.5, 3, 8, seed=37, q_scale=q)
    st0 = init_c(orc, s)
    dummies = {k: np.zeros(1) for k in ("rspsep", "svecre", "hsvre", "ssvre")}
    res = []
    for f in (orc.call, ref.call):
        st = {k: v.copy(order="F") for k, v in st0.items()}
        kw = synth.bloch_args(s, K)
        for it in range(1, 4):
            kw["iscf"] = it
            f("scf_diag", **kw, **st, **dummies)
        res.append(st)
    o, r = res
    np.testing.assert_allclose(o["sval"], r["sval"], rtol=1e-10, atol=1e-10)
    rho = [np.zeros(s.npts), np.zeros(s.npts)]
    for st, rh, f in ((o, rho[0], orc.call), (r, rho[1], ref.call)):
        f("scf_charge_smt_c", neigmx=s.nbands, nums=1, ncol=1, nspv=1, numk=1, ncpx=s.n, ncpy=s.n,
          ncpz=s.n, fnele=np.ones(s.nbands), sveccm=st["sveccm"], rhosmt=rh)
    np.testing.assert_allclose(rho[0], rho[1], rtol=0, atol=1e-9 * rho[1].max())


def test_reorderstates_c_equals_reference(orc):
    s = synth.make_system(6, 0.5, 2, 7, seed=34)
    rng = np.random.default_rng(2)
    z = lambda *sh: np.asfortranarray(rng.standard_normal(sh) + 1j * rng.standard_normal(sh))
    arr = {"sval": np.array([0.3, -0.1, 0.3, -0.5, 0.2, 0.9, -0.1]), "residual_states": np.arange(7.0),
           "cspsep": z(9, s.natom, 7), "sveccm": z(s.npts, 7), "ssvcm": z(s.npts, 7),
           "hsvcm": np.zeros((s.npts, 7), dtype=np.complex128, order="F"),
           "rspsep": np.zeros(1), "svecre": np.zeros(1), "ssvre": np.zeros(1), "hsvre": np.zeros(1)}
    o, r = both(orc, "reorderstates", arr, **synth.abi_args(s, neigmx=7, nrc=1), ns1=1)
    for k in ("sval", "residual_states", "cspsep", "sveccm", "ssvcm"):
        np.testing.assert_array_equal(o[k], r[k])


@pytest.mark.parametrize("nspv", [1, 2])
def test_scf_charge_smt_c_spin_mapping_equals_reference(orc, nspv):
    """Two spins and two k-points: with nspv = 1 both spins add into one density."""
    s = synth.make_system(6, 0.5, 1, 3, seed=3)
    rng = np.random.default_rng(5)
    psi = np.asfortranarray(rng.standard_normal((s.npts, 3, 2, 2)) + 1j * rng.standard_normal((s.npts, 3, 2, 2)))
    f = np.asfortranarray(rng.random((3, 2, 2)))
    o, r = both(orc, "scf_charge_smt_c", {"rhosmt": np.full((s.npts, nspv), 7.0, order="F")}, neigmx=3,
                nums=2, ncol=1, nspv=nspv, numk=2, ncpx=s.n, ncpy=s.n, ncpz=s.n, fnele=f, sveccm=psi)
    np.testing.assert_allclose(o["rhosmt"], r["rhosmt"], rtol=1e-14, atol=0)


def _align_cols(o, r):
    """r's bands times the unit phase that best matches o's (per band)."""
    ph = np.einsum("ij,ij->j", np.conj(r), o)
    return r * (ph / np.maximum(np.abs(ph), 1e-300))[None, :], ph / np.maximum(np.abs(ph), 1e-300)


@pytest.mark.parametrize("q", QS)
@pytest.mark.parametrize("l_hsv", [0, 1])
def test_ksdiis_c_equals_reference(orc, q, l_hsv):
    """Two RMM-DIIS sweeps (ndiismax = 3) after four CG iterations per band.  The combination's
    overall phase per band is the eigensolver's choice; every later operation carries a band's
    phase along, so the bands agree after one phase per band."""
    s = synth.make_system(12, 0.45, 3, 8, seed=36, q_scale=q)
    st = init_c(orc, s)
    keys = ("sval", "cspsep", "sveccm", "ssvcm", "residual_states", "hsvcm")
    orc.call("kscg_c", **synth.bloch_args(s, K, nsdmax=4), **{k: st[k] for k in keys})
    o, r = both(orc, "ksdiis_c", {k: st[k] for k in keys},
                **synth.bloch_args(s, K, ndiismax=3, l_hsv=l_hsv))
    np.testing.assert_allclose(o["sval"], r["sval"], rtol=1e-11, atol=1e-11)
    np.testing.assert_allclose(o["residual_states"], r["residual_states"], rtol=1e-7, atol=1e-15)
    _, ph = _align_cols(o["sveccm"], r["sveccm"])
    for k in ("sveccm", "ssvcm", "hsvcm"):
        np.testing.assert_allclose(o[k], r[k] * ph[None, :], rtol=0, atol=1e-9 * np.abs(r[k]).max(), err_msg=k)
    cs = lambda a: a.reshape(-1, s.nbands, order="F")
    np.testing.assert_allclose(cs(o["cspsep"]), cs(r["cspsep"]) * ph[None, :], rtol=0,
                               atol=1e-9 * np.abs(r["cspsep"]).max())
    assert o["_ret"]["ksitmax"] == r["_ret"]["ksitmax"]
    assert o["_ret"]["ksconv"] == r["_ret"]["ksconv"]


@pytest.mark.parametrize("q", QS)
def test_scf_diag_bloch_diis_steps_equal_reference(orc, q):
    """scf_diag with nrc = 1 taking the DIIS branch after two CG steps (ncgscf = 2), with the
    Rayleigh-Ritz step before and the orthonormalisation of l_hsv: eigenvalues and densities."""
    s = synth.make_system(10, 0.5, 3, 8, seed=38, q_scale=q)
    st0 = init_c(orc, s)
    dummies = {k: np.zeros(1) for k in ("rspsep", "svecre", "hsvre", "ssvre")}
    res = []
    for f in (orc.call, ref.call):
        st = {k: v.copy(order="F") for k, v in st0.items()}
        kw = synth.bloch_args(s, K, ncgscf=2, nrrz=2, ndiismax=3)
        for it in range(1, 6):
            kw["iscf"] = it
            f("scf_diag", **kw, **st, **dummies)
        res.append(st)
    o, r = res
    np.testing.assert_allclose(o["sval"], r["sval"], rtol=1e-10, atol=1e-10)
    rho = [np.zeros(s.npts), np.zeros(s.npts)]
    for st, rh, f in ((o, rho[0], orc.call), (r, rho[1], ref.call)):
        f("scf_charge_smt_c", neigmx=s.nbands, nums=1, ncol=1, nspv=1, numk=1, ncpx=s.n, ncpy=s.n,
          ncpz=s.n, fnele=np.ones(s.nbands), sveccm=st["sveccm"], rhosmt=rh)
    np.testing.assert_allclose(rho[0], rho[1], rtol=0, atol=1e-9 * rho[1].max())
