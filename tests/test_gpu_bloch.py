"""The complex (Bloch, k-point) path through the GPU library against the oracle (which the
reference's own routines pin, tests/test_reference_bloch.py) on the same seeded inputs, at a
generic k-point, with S = 1 and S != 1.

Bars: index work (halo copies, reorder) bit-exact; single applications (H_k, K, the projector
gather and expansion, the density) to 1e-13 relative (the GPU's sincos, FMA contraction and
reduction order are the only differences); the Cholesky-QR orthonormalisation against the
oracle's block Gram-Schmidt to 1e-10 of the largest entry (the same Q in exact arithmetic)
with Q^H S Q = I to 1e-12; the band-sequential complex CG with the oracle's iterates (no
eigensolver phase enters) to 1e-9; Rayleigh-Ritz after aligning each band's phase (the
Hermitian eigensolvers fix phases differently); scf_diag (nrc = 1) to the lowest eigenpairs
of the dense eigh(H_k, S_k)."""
import numpy as np
import pytest
import scipy.linalg

from rspacedist_amd import synth
from tests.test_oracle_bloch import K, dense_hk, dense_sk, halo_c, phase_matrix

pytestmark = pytest.mark.gpu

QS = [0.0, 2.0]
ZK = ("sveccm", "ssvcm", "hsvcm", "cspsep")


@pytest.fixture(scope="module")
def kk():
    from rspacedist_amd import kukan
    kukan.init(0)
    yield kukan


def both(kk, orc, name, arrays, **kw):
    out = []
    for f in (kk.call, orc.call):
        a = {k: (v.copy(order="F") if isinstance(v, np.ndarray) else v) for k, v in arrays.items()}
        a["_ret"] = f(name, **kw, **a)
        out.append(a)
    return out


def close(g, o, k, rel):
    scale = max(np.abs(o[k]).max(), 1e-300)
    err = np.abs(g[k] - o[k]).max() / scale
    assert err <= rel, f"{k}: relative difference {err:.2e} > {rel:.0e}"


def init_c(orc, s, m=None):
    m = s.nbands if m is None else m
    psi = synth.random_bands_c(s, nbands=m)
    st = dict(sveccm=psi, ssvcm=np.zeros_like(psi),
              cspsep=np.zeros((9, s.natom, m), dtype=np.complex128, order="F"))
    orc.call("orthogonalization_c", **synth.bloch_args(s, K, neigmx=m), nflag=1, **st)
    st.update(hsvcm=np.zeros_like(psi), sval=np.zeros(m), residual_states=np.zeros(m))
    return st


REAL_DUMMIES = ("rspsep", "svecre", "hsvre", "ssvre")


def init_c_dummy(s):
    z = np.zeros((s.npts, s.nbands), dtype=np.complex128, order="F")
    return dict(sveccm=z, ssvcm=z.copy(order="F"), hsvcm=z.copy(order="F"), sval=np.zeros(s.nbands),
                residual_states=np.zeros(s.nbands),
                cspsep=np.zeros((9, s.natom, s.nbands), dtype=np.complex128, order="F"))


@pytest.mark.parametrize("nperi", [0, 1, 2, 3])
def test_halo_c_bit_exact(kk, orc, nperi):
    rng = np.random.default_rng(nperi)
    v = np.zeros((9 + 8, 8 + 8, 7 + 8, 2), dtype=np.complex128, order="F")
    v[4:13, 4:12, 4:11] = rng.standard_normal((9, 8, 7, 2)) + 1j * rng.standard_normal((9, 8, 7, 2))
    g, o = both(kk, orc, "overlap_finitedifference_c", {"v_cc": v}, nperi=nperi, ncpa=9, ncpb=8,
                ncpc=7, nf=4, nf1=3, nsnd=4, ncol=2)
    np.testing.assert_array_equal(g["v_cc"], o["v_cc"])


@pytest.mark.parametrize("nf", [1, 4, 8])
def test_kslaplacian_c(kk, orc, nf):
    n1, n2, n3, h = 12, 10, 9, (0.3, 0.35, 0.4)
    rng = np.random.default_rng(nf)
    v = np.zeros((n1 + 2 * nf, n2 + 2 * nf, n3 + 2 * nf), dtype=np.complex128, order="F")
    v[:] = rng.standard_normal(v.shape) + 1j * rng.standard_normal(v.shape)
    arr = {"vcm": v, "avcm": np.asfortranarray(rng.standard_normal((n1, n2, n3, 2)) + 0j),
           "workc": np.zeros((n1, n2, n3), dtype=np.complex128, order="F")}
    xm = (n1 * h[0] / 2, n2 * h[1] / 2, n3 * h[2] / 2)
    g, o = both(kk, orc, "kslaplacian_c", arr, ncpx=n1, ncpy=n2, ncpz=n3, neigmx=2, ncol=1, nvef=1,
                nf=nf, dx=h[0], dy=h[1], dz=h[2], xmax=xm[0], ymax=xm[1], zmax=xm[2], skpx=K[0],
                skpy=K[1], skpz=K[2], veff=np.asfortranarray(rng.standard_normal((n1, n2, n3))), l=2)
    close(g, o, "avcm", 1e-13)
    close(g, o, "vcm", 1e-14)
    close(g, o, "workc", 1e-13)


def test_ksprecondition_c(kk, orc):
    rng = np.random.default_rng(4)
    u = rng.standard_normal((10, 10, 10)) + 1j * rng.standard_normal((10, 10, 10))
    g, o = both(kk, orc, "ksprecondition_c", {"vcm": halo_c(u, 4, 1),
                "avcm": np.zeros((10, 10, 10), dtype=np.complex128, order="F")},
                ncpx=10, ncpy=10, ncpz=10, nf=4, ncol=1)
    close(g, o, "avcm", 1e-15)


@pytest.mark.parametrize("q", QS)
def test_nonlocal_c(kk, orc, q):
    s = synth.make_system(12, 0.45, 3, 2, seed=5, q_scale=q)
    kw = dict(synth.bloch_args(s, K), neigmx=2, l=2)
    rng = np.random.default_rng(7)
    u = np.asfortranarray(rng.standard_normal((s.npts, 2)) + 1j * rng.standard_normal((s.npts, 2)))
    z = lambda *sh: np.zeros(sh, dtype=np.complex128, order="F")
    cp0 = np.asfortranarray(rng.standard_normal((9, s.natom)) + 1j * rng.standard_normal((9, s.natom)))
    g, o = both(kk, orc, "nonlocaloperation_c_01", {"vcm2": u, "cpsep": cp0, "vcccm": z(s.num_list)}, **kw)
    close(g, o, "cpsep", 1e-13)  # accumulated onto cp0
    np.testing.assert_allclose(g["cpsep"].ravel(order="F") - cp0.ravel(order="F"),
                               phase_matrix(s, K).T @ u[:, 1], rtol=0, atol=1e-12)
    for nhs, X in ((1, s.dij), (0, s.sss)):
        g2, o2 = both(kk, orc, "nonlocaloperation_c_02", {"cpsep": o["cpsep"], "dijsss": X, "avcm": u.copy(),
                      "vcccm": z(s.num_list), "avc": z(s.num_list)}, **dict(kw, nhs=nhs, nso=0))
        close(g2, o2, "avcm", 1e-13)


@pytest.mark.parametrize("q", QS)
def test_orthogonalization_c_block(kk, orc, q):
    s = synth.make_system(10, 0.5, 2, 40, seed=6, q_scale=q)
    psi = synth.random_bands_c(s, smooth=False)
    arr = {"sveccm": psi, "ssvcm": np.zeros_like(psi),
           "cspsep": np.zeros((9, s.natom, 40), dtype=np.complex128, order="F")}
    g, o = both(kk, orc, "orthogonalization_c", arr, **synth.bloch_args(s, K), nflag=1)
    for k in ("sveccm", "ssvcm", "cspsep"):
        close(g, o, k, 1e-10)
    Sd = dense_sk(s, K)
    Q = g["sveccm"]
    np.testing.assert_allclose(np.conj(Q.T) @ Sd @ Q * s.dv, np.eye(40), rtol=0, atol=1e-12)
    np.testing.assert_allclose(g["ssvcm"], Sd @ Q, rtol=0, atol=1e-11 * np.abs(Sd @ Q).max())


@pytest.mark.parametrize("q", QS)
def test_orthogonalization_c_01_single_band(kk, orc, q):
    s = synth.make_system(10, 0.5, 2, 6, seed=16, q_scale=q)
    st = init_c(orc, s)
    rng = np.random.default_rng(3)
    st["sveccm"][:, 4] += 0.3 * (rng.standard_normal(s.npts) + 1j * rng.standard_normal(s.npts))
    arr = {k: st[k] for k in ("sveccm", "ssvcm", "cspsep")}
    g, o = both(kk, orc, "orthogonalization_c_01", arr, **synth.bloch_args(s, K), l_inp=5, nflag=1)
    for k in ("sveccm", "ssvcm", "cspsep"):
        close(g, o, k, 1e-12)


@pytest.mark.parametrize("q", QS)
def test_rayleighritz_c(kk, orc, q):
    s = synth.make_system(10, 0.5, 3, 10, seed=33, q_scale=q)
    st = init_c(orc, s)
    arr = {k: st[k] for k in ("cspsep", "sveccm", "ssvcm", "hsvcm", "sval")}
    g, o = both(kk, orc, "rayleighritz_c", arr, **synth.bloch_args(s, K), l_rrb=1)
    np.testing.assert_allclose(g["sval"], o["sval"], rtol=0, atol=1e-12)
    ph = np.einsum("ij,ij->j", np.conj(o["sveccm"]), g["sveccm"])
    ph = ph / np.abs(ph)
    for k in ("sveccm", "ssvcm", "hsvcm"):
        np.testing.assert_allclose(g[k], o[k] * ph[None, :], rtol=0, atol=1e-9 * np.abs(o[k]).max(), err_msg=k)
    np.testing.assert_allclose(g["cspsep"].reshape(-1, 10, order="F"),
                               o["cspsep"].reshape(-1, 10, order="F") * ph[None, :], rtol=0,
                               atol=1e-9 * np.abs(o["cspsep"]).max())


@pytest.mark.parametrize("q", QS)
@pytest.mark.parametrize("nsdmax", [1, 4, 12])  # 12 passes the reset at iteration 10
def test_kscg_c_iterates(kk, orc, q, nsdmax):
    s = synth.make_system(12, 0.45, 3, 8, seed=35, q_scale=q)
    st = init_c(orc, s)
    arr = {k: st[k] for k in ("sval", "cspsep", "sveccm", "ssvcm", "residual_states", "hsvcm")}
    g, o = both(kk, orc, "kscg_c", arr, **synth.bloch_args(s, K, nsdmax=nsdmax, epssd=1e-14))
    np.testing.assert_allclose(g["sval"], o["sval"], rtol=1e-11, atol=1e-11)
    for k in ZK:
        close(g, o, k, 1e-9)
    np.testing.assert_allclose(g["residual_states"], o["residual_states"], rtol=1e-6, atol=1e-16)
    assert g["_ret"]["ksitmax"] == o["_ret"]["ksitmax"]
    assert g["_ret"]["ksconv"] == o["_ret"]["ksconv"]


def test_reorderstates_c_bit_exact(kk, orc):
    s = synth.make_system(6, 0.5, 2, 7, seed=34)
    rng = np.random.default_rng(2)
    z = lambda *sh: np.asfortranarray(rng.standard_normal(sh) + 1j * rng.standard_normal(sh))
    arr = {"sval": np.array([0.3, -0.1, 0.3, -0.5, 0.2, 0.9, -0.1]), "residual_states": np.arange(7.0),
           "cspsep": z(9, s.natom, 7), "sveccm": z(s.npts, 7), "ssvcm": z(s.npts, 7),
           "hsvcm": np.zeros((s.npts, 7), dtype=np.complex128, order="F"),
           **{k: np.zeros(1) for k in REAL_DUMMIES}}
    g, o = both(kk, orc, "reorderstates", arr, **synth.abi_args(s, neigmx=7, nrc=1), ns1=1)
    for k in ("sval", "residual_states", "cspsep", "sveccm", "ssvcm"):
        np.testing.assert_array_equal(g[k], o[k])


@pytest.mark.parametrize("q", QS)
def test_scf_diag_bloch_steps(kk, orc, q):
    """Three scf_diag steps (nrc = 1: Rayleigh-Ritz, kscg_c, reorder): eigenvalues and the
    density agree with the oracle's (band phases are the eigensolvers' choice)."""
    s = synth.make_system(10, 0.5, 3, 8, seed=37, q_scale=q)
    st0 = init_c(orc, s)
    res = []
    for f in (kk.call, orc.call):
        st = {k: v.copy(order="F") for k, v in st0.items()}
        kw = synth.bloch_args(s, K)
        for it in range(1, 4):
            kw["iscf"] = it
            f("scf_diag", **kw, **st, **{k: np.zeros(1) for k in REAL_DUMMIES})
        res.append(st)
    g, o = res
    np.testing.assert_allclose(g["sval"], o["sval"], rtol=1e-10, atol=1e-10)
    rho = [np.zeros(s.npts), np.zeros(s.npts)]
    for st, rh, f in ((g, rho[0], kk.call), (o, rho[1], orc.call)):
        f("scf_charge_smt_c", neigmx=s.nbands, nums=1, ncol=1, nspv=1, numk=1, ncpx=s.n, ncpy=s.n,
          ncpz=s.n, fnele=np.ones(s.nbands), sveccm=st["sveccm"], rhosmt=rh)
    np.testing.assert_allclose(rho[0], rho[1], rtol=0, atol=1e-9 * rho[1].max())
    np.testing.assert_allclose(rho[0], (np.abs(g["sveccm"]) ** 2).sum(axis=1), rtol=1e-13)


@pytest.mark.parametrize("q", QS)
def test_scf_diag_bloch_converges_to_dense_eigenpairs(kk, orc, q):
    s = synth.make_system(10, 0.5, 3, 6, seed=8, q_scale=q)
    st = init_c(orc, s)
    kw = synth.bloch_args(s, K, nsdmax=4, epssd=1e-12)
    for it in range(1, 61):
        kw["iscf"] = it
        kk.call("scf_diag", **kw, **st, **{k: np.zeros(1) for k in REAL_DUMMIES})
    Hd, Sd = dense_hk(s, K), dense_sk(s, K)
    w = scipy.linalg.eigh(Hd, Sd, eigvals_only=True)[:6]
    np.testing.assert_allclose(st["sval"], w, rtol=0, atol=1e-8)
    Q = st["sveccm"]
    np.testing.assert_allclose(np.conj(Q.T) @ Sd @ Q * s.dv, np.eye(6), rtol=0, atol=1e-12)
    R = Hd @ Q - (Sd @ Q) * st["sval"][None, :]
    assert np.abs(R).max() < 1e-3


def test_complex_entry_points_reject_noncollinear(kk):
    s = synth.make_system(6, 0.5, 1, 2, seed=1)
    psi = synth.random_bands_c(s)
    with pytest.raises(kk.KukanError, match=r"\(5\)"):
        kk.call("scf_charge_smt_c", neigmx=2, nums=1, ncol=2, nspv=1, numk=1, ncpx=s.n, ncpy=s.n,
                ncpz=s.n, fnele=np.ones(2), sveccm=psi, rhosmt=np.zeros(2 * s.npts))
    with pytest.raises(kk.KukanError, match=r"\(5\)"):
        kk.call("kscg_c", **dict(synth.bloch_args(s, K), nso=1), **init_c_dummy(s))


@pytest.mark.parametrize("q", QS)
@pytest.mark.parametrize("l_hsv", [0, 1])
def test_ksdiis_c(kk, orc, q, l_hsv):
    """Two RMM-DIIS sweeps after four CG iterations per band.  Both sides fix the combination's
    phase the same way (the incoming band's coefficient real and non-negative), so the bands
    compare directly."""
    s = synth.make_system(12, 0.45, 3, 8, seed=36, q_scale=q)
    st = init_c(orc, s)
    keys = ("sval", "cspsep", "sveccm", "ssvcm", "residual_states", "hsvcm")
    orc.call("kscg_c", **synth.bloch_args(s, K, nsdmax=4), **{k: st[k] for k in keys})
    g, o = both(kk, orc, "ksdiis_c", {k: st[k] for k in keys},
                **synth.bloch_args(s, K, ndiismax=3, l_hsv=l_hsv))
    np.testing.assert_allclose(g["sval"], o["sval"], rtol=1e-11, atol=1e-11)
    np.testing.assert_allclose(g["residual_states"], o["residual_states"], rtol=1e-7, atol=1e-15)
    for k in ZK:
        close(g, o, k, 1e-9)
    assert g["_ret"]["ksitmax"] == o["_ret"]["ksitmax"]
    assert g["_ret"]["ksconv"] == o["_ret"]["ksconv"]


@pytest.mark.parametrize("q", QS)
def test_scf_diag_bloch_diis_steps(kk, orc, q):
    """Five scf_diag steps with nrc = 1: CG for two, RMM-DIIS afterwards, Rayleigh-Ritz every
    other step (so l_hsv is set on the steps before one): eigenvalues and density."""
    s = synth.make_system(10, 0.5, 3, 8, seed=38, q_scale=q)
    st0 = init_c(orc, s)
    res = []
    for f in (kk.call, orc.call):
        st = {k: v.copy(order="F") for k, v in st0.items()}
        kw = synth.bloch_args(s, K, ncgscf=2, nrrz=2, ndiismax=3)
        for it in range(1, 6):
            kw["iscf"] = it
            f("scf_diag", **kw, **st, **{k: np.zeros(1) for k in REAL_DUMMIES})
        res.append(st)
    g, o = res
    np.testing.assert_allclose(g["sval"], o["sval"], rtol=1e-10, atol=1e-10)
    rho = [np.zeros(s.npts), np.zeros(s.npts)]
    for st, rh, f in ((g, rho[0], kk.call), (o, rho[1], orc.call)):
        f("scf_charge_smt_c", neigmx=s.nbands, nums=1, ncol=1, nspv=1, numk=1, ncpx=s.n, ncpy=s.n,
          ncpz=s.n, fnele=np.ones(s.nbands), sveccm=st["sveccm"], rhosmt=rh)
    np.testing.assert_allclose(rho[0], rho[1], rtol=0, atol=1e-9 * rho[1].max())


@pytest.mark.parametrize("q", QS)
def test_scf_diag_bloch_diis_converges_to_dense_eigenpairs(kk, orc, q):
    s = synth.make_system(10, 0.5, 3, 6, seed=18, q_scale=q)
    st = init_c(orc, s)
    kw = synth.bloch_args(s, K, nsdmax=4, epssd=1e-12, ncgscf=10, ndiismax=6)
    for it in range(1, 81):
        kw["iscf"] = it
        kk.call("scf_diag", **kw, **st, **{k: np.zeros(1) for k in REAL_DUMMIES})
    Hd, Sd = dense_hk(s, K), dense_sk(s, K)
    w = scipy.linalg.eigh(Hd, Sd, eigvals_only=True)[:6]
    np.testing.assert_allclose(st["sval"], w, rtol=0, atol=1e-7)
    Q = st["sveccm"]
    np.testing.assert_allclose(np.conj(Q.T) @ Sd @ Q * s.dv, np.eye(6), rtol=0, atol=1e-12)
