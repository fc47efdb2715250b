"""S != 1 (PAW overlap, sss = q != 0) through the GPU library: every routine that applies or
orthonormalises in the S metric against the oracle on the same seeded inputs, and the block
solver against the generalised eigenproblem H psi = eps S psi of the dense H and S.

S = 1 + sum_a sum_ij |p_i> q_ij <p_j| / dV with q = synth.seeded_qmat (positive semi-definite,
so S is positive definite).  With q_scale = 2 the generalised eigenvalues differ from the
S = 1 ones by ~1e-2 Ha, so a dropped or misplaced S term fails every check below.
Tolerances as in the S = 1 parity tests (tests/test_gpu_parity.py).
"""
import numpy as np
import pytest
import scipy.linalg

from rspacedist_amd import synth
from tests import helpers as H

pytestmark = pytest.mark.gpu

Q = 2.0


@pytest.fixture(scope="module")
def kk():
    from rspacedist_amd import kukan
    kukan.init(0)
    kukan.set_param("kscg_mode", 0)
    yield kukan
    kukan.set_param("kscg_mode", 0)
    kukan.set_param("precision", 0)
    kukan.set_param("mixed_min_rows", 65536)


def both(kk, orc, name, arrays, **kw):
    out = []
    for f in (kk.call, orc.call):
        a = {k: (v.copy(order="F") if isinstance(v, np.ndarray) else v) for k, v in arrays.items()}
        a["_ret"] = f(name, **kw, **a)
        out.append(a)
    return out


def s_gram(s, st):
    return st["svecre"].T @ H.dense_s(s) @ st["svecre"] * s.dv


def test_orthogonalization_r_block_path_s_metric(kk, orc):
    s = synth.make_system(12, 0.45, 3, 40, seed=31, q_scale=Q)
    psi0 = synth.random_bands(s)
    arr = {"svecre": psi0, "ssvre": np.zeros_like(psi0, order="F"),
           "rspsep": np.zeros((9, s.natom, 40), order="F")}
    g, o = both(kk, orc, "orthogonalization_r", arr, **dict(synth.abi_args(s), nflag=1))
    for k in ("svecre", "ssvre", "rspsep"):
        np.testing.assert_allclose(g[k], o[k], atol=1e-10 * np.abs(o[k]).max(), err_msg=k)
    np.testing.assert_allclose(s_gram(s, g), np.eye(40), atol=1e-12)
    np.testing.assert_allclose(g["ssvre"], H.dense_s(s) @ g["svecre"], atol=1e-11)


def test_orthogonalization_r_01_single_band_s_metric(kk, orc):
    s = synth.make_system(12, 0.45, 3, 8, seed=32, q_scale=Q)
    st = H.init_state(s, orc, smooth=True)
    st["svecre"][:, 5] += 0.3 * np.random.default_rng(1).standard_normal(s.npts)
    arr = {k: st[k] for k in ("svecre", "ssvre", "rspsep")}
    g, o = both(kk, orc, "orthogonalization_r_01", arr, **dict(synth.abi_args(s), l_inp=6, nflag=1))
    for k in arr:
        np.testing.assert_allclose(g[k], o[k], atol=1e-12 * np.abs(o[k]).max(), err_msg=k)
    np.testing.assert_allclose(s_gram(s, g)[:6, :6], np.eye(6), atol=1e-12)


def test_rayleighritz_s_metric(kk, orc):
    s = synth.make_system(12, 0.45, 3, 10, seed=33, q_scale=Q)
    st = H.init_state(s, orc, smooth=True)
    psi0 = st["svecre"].copy()
    arr = {k: st[k] for k in ("rspsep", "svecre", "ssvre", "hsvre", "sval")}
    g, o = both(kk, orc, "rayleighritz_r", arr, **dict(synth.abi_args(s), l_rrb=1))
    np.testing.assert_allclose(g["sval"], o["sval"], rtol=1e-12, atol=1e-12)
    for k in ("svecre", "ssvre", "rspsep", "hsvre"):
        np.testing.assert_allclose(g[k], o[k], atol=1e-9 * np.abs(o[k]).max(), err_msg=k)
    Hd, Sd = H.dense_h(s), H.dense_s(s)
    w = scipy.linalg.eigh(psi0.T @ Hd @ psi0 * s.dv, psi0.T @ Sd @ psi0 * s.dv, eigvals_only=True)
    np.testing.assert_allclose(g["sval"], w, rtol=1e-11, atol=1e-11)


@pytest.mark.parametrize("nsdmax", [4, 12])  # 12 passes the CG reset at iteration 10
def test_kscg_r_sequential_s_metric(kk, orc, nsdmax):
    kk.set_param("kscg_mode", 0)
    s = synth.make_system(12, 0.45, 3, 8, seed=35, q_scale=Q)
    st = H.init_state(s, orc, smooth=True)
    arr = {k: st[k] for k in ("sval", "rspsep", "svecre", "ssvre", "residual_states", "hsvre")}
    g, o = both(kk, orc, "kscg_r", arr, **synth.abi_args(s, nsdmax=nsdmax, epssd=1e-14))
    np.testing.assert_allclose(g["sval"], o["sval"], rtol=1e-10, atol=1e-10)
    for k in ("svecre", "ssvre", "hsvre", "rspsep"):
        np.testing.assert_allclose(g[k], o[k], atol=1e-8 * np.abs(o[k]).max(), err_msg=k)
    np.testing.assert_allclose(g["residual_states"], o["residual_states"], rtol=1e-6, atol=1e-14)
    assert g["_ret"]["ksitmax"] == o["_ret"]["ksitmax"]


@pytest.mark.parametrize("l_hsv", [0, 1])
def test_ksdiis_r_s_metric(kk, orc, l_hsv):
    s = synth.make_system(12, 0.45, 3, 8, seed=36, q_scale=Q)
    st = H.init_state(s, orc, smooth=True)
    arr = {k: st[k] for k in ("rspsep", "sval", "svecre", "ssvre", "residual_states", "hsvre")}
    g, o = both(kk, orc, "ksdiis_r", arr, **synth.abi_args(s, l_hsv=l_hsv))
    np.testing.assert_allclose(g["sval"], o["sval"], rtol=1e-11, atol=1e-11)
    sg = np.sign(np.einsum("ij,ij->j", g["svecre"], o["svecre"]))
    for k in ("svecre", "ssvre", "hsvre", "rspsep"):
        ref = o[k] * sg.reshape((1,) * (o[k].ndim - 1) + (-1,))
        np.testing.assert_allclose(g[k], ref, atol=1e-9 * np.abs(o[k]).max(), err_msg=k)
    np.testing.assert_allclose(g["residual_states"], o["residual_states"], rtol=1e-8)


@pytest.mark.parametrize("ncgscf", [200, 2])  # CG only, then DIIS after two CG steps
def test_scf_diag_sequential_s_metric_over_steps(kk, orc, ncgscf):
    kk.set_param("kscg_mode", 0)
    s = synth.make_system(12, 0.45, 3, 8, seed=37, q_scale=Q)
    st0 = H.init_state(s, orc, smooth=True)
    res = {}
    for name, f in (("gpu", kk.call), ("orc", orc.call)):
        st = {k: v.copy(order="F") for k, v in st0.items()}
        kw = synth.abi_args(s, ncgscf=ncgscf)
        for it in range(1, 5):
            kw["iscf"] = it
            f("scf_diag", **kw, **st)
        res[name] = st
    np.testing.assert_allclose(res["gpu"]["sval"], res["orc"]["sval"], rtol=1e-9, atol=1e-10)
    np.testing.assert_allclose(s_gram(s, res["gpu"]), np.eye(8), atol=1e-11)


def _generalised(s, m):
    Hd, Sd = H.dense_h(s), H.dense_s(s)
    w = scipy.linalg.eigh(Hd, Sd, eigvals_only=True)[:m]
    w1 = scipy.linalg.eigh(Hd, eigvals_only=True)[:m]
    assert np.abs(w - w1).max() > 1e-3  # the S term matters at this q
    return w


@pytest.mark.parametrize("proj", [1, 0])
def test_scf_diag_block_s_metric_converges_to_generalised_eigenvalues(kk, orc, proj):
    """Block CG (kscg_mode 1) with S != 1: the projections use S psi, the closing
    Cholesky-QR the S Gram matrix; the bands converge to eigh(H, S) and stay S-orthonormal."""
    kk.set_param("kscg_mode", 1)
    kk.set_param("bcg_proj", proj)
    try:
        s = synth.make_system(10, 0.5, 3, 8, seed=8, q_scale=Q)
        st = H.init_state(s, orc, smooth=True)
        kw = synth.abi_args(s, nsdmax=4, epssd=1e-12)
        for it in range(1, 41):
            kw["iscf"] = it
            kk.call("scf_diag", **kw, **st)
    finally:
        kk.set_param("bcg_proj", 1)
        kk.set_param("kscg_mode", 0)
    np.testing.assert_allclose(st["sval"], _generalised(s, 8), atol=1e-8)
    np.testing.assert_allclose(s_gram(s, st), np.eye(8), atol=1e-12)
    np.testing.assert_allclose(st["ssvre"], H.dense_s(s) @ st["svecre"], atol=1e-11)


def test_mixed_scf_diag_s_metric_converges_to_generalised_eigenvalues(kk, orc):
    """Split-precision subspace products with S != 1 (the S psi rotation of the closing
    Cholesky-QR and the S-metric projections run as split products)."""
    kk.set_param("kscg_mode", 1)
    kk.set_param("precision", 1)
    kk.set_param("mixed_min_rows", 0)
    try:
        s = synth.make_system(10, 0.5, 3, 8, seed=8, q_scale=Q)
        st = H.init_state(s, orc, smooth=True)
        kw = synth.abi_args(s, nsdmax=4, epssd=1e-12)
        for it in range(1, 41):
            kw["iscf"] = it
            kk.call("scf_diag", **kw, **st)
    finally:
        kk.set_param("precision", 0)
        kk.set_param("mixed_min_rows", 65536)
        kk.set_param("kscg_mode", 0)
    np.testing.assert_allclose(st["sval"], _generalised(s, 8), atol=1e-6)
    g = s_gram(s, st)
    np.testing.assert_allclose(np.diag(g), 1.0, rtol=0, atol=1e-12)
    np.testing.assert_allclose(g, np.eye(8), atol=1e-12)
