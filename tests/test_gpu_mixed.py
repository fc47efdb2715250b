"""Split-FP16 subspace products (kukan_set_param("precision", 1)) against FP64.

The mixed products split each FP64 operand, scaled per column by a power of two, into two
FP16 halves and sum three MFMA products in FP32 (FP64 across 256-deep K-chunks).  Per
element the dropped terms are ~2^-22 relative, so the bound used here is

    |C - C_fp64| <= 2e-6 * (|A|^T |B|)_ij

Measured errors stay below ~4% of it (about 1e-7 of |A|^T |B|), while a dropped cross term
(2^-11 relative) or any indexing or layout mistake exceeds it by orders of magnitude.  The eigenvalue tests hold the headline accuracy target:
eigenvalues within 1e-6 Ha of the FP64 solver.
"""
import numpy as np
import pytest
import scipy.linalg

from rspacedist_amd import synth
from tests import helpers as H

pytestmark = pytest.mark.gpu

TOL = 2e-6


@pytest.fixture(scope="module")
def kk():
    from rspacedist_amd import kukan
    kukan.init(0)
    yield kukan
    kukan.set_param("precision", 0)
    kukan.set_param("mixed_min_rows", 65536)
    kukan.set_param("mixed_terms", 3)
    kukan.set_param("kscg_mode", 0)


@pytest.fixture
def mixed(kk):
    kk.set_param("precision", 1)
    kk.set_param("mixed_min_rows", 0)
    yield kk
    kk.set_param("precision", 0)
    kk.set_param("mixed_min_rows", 65536)


def _tn(kk, A, B, alpha=1.0, beta=0.0, C=None):
    K, m = A.shape
    n = B.shape[1]
    C = np.zeros((m, n), order="F") if C is None else C.copy(order="F")
    kk.subspace_gemm(1, m, n, K, alpha, np.asfortranarray(A), K, np.asfortranarray(B), K, beta, C,
                     m)
    return C


def _nn(kk, X, Cm, alpha=1.0, beta=0.0, Y=None):
    R, K = X.shape
    n = Cm.shape[1]
    Y = np.zeros((R, n), order="F") if Y is None else Y.copy(order="F")
    kk.subspace_gemm(0, R, n, K, alpha, np.asfortranarray(X), R, np.asfortranarray(Cm), K, beta,
                     Y, R)
    return Y


def test_subspace_gemm_fp64(kk):
    kk.set_param("precision", 0)
    rng = np.random.default_rng(1)
    A, B = rng.standard_normal((1000, 37)), rng.standard_normal((1000, 53))
    C0 = rng.standard_normal((37, 53))
    np.testing.assert_allclose(_tn(kk, A, B, 0.7, 0.3, C0), 0.7 * A.T @ B + 0.3 * C0, rtol=0,
                               atol=1e-12)
    X, Cm = rng.standard_normal((1000, 37)), rng.standard_normal((37, 53))
    Y0 = rng.standard_normal((1000, 53))
    np.testing.assert_allclose(_nn(kk, X, Cm, 0.7, 0.3, Y0), 0.7 * X @ Cm + 0.3 * Y0, rtol=0,
                               atol=1e-12)


@pytest.mark.parametrize("m,n,K", [(1, 1, 1), (37, 53, 1000), (130, 129, 70001),
                                   (256, 200, 300000)])
def test_mixed_tn(mixed, m, n, K):
    rng = np.random.default_rng(m + n + K)
    # columns of very different magnitude exercise the per-column scaling
    A = rng.standard_normal((K, m)) * np.exp2(rng.integers(-20, 20, m))
    B = rng.standard_normal((K, n)) * np.exp2(rng.integers(-20, 20, n))
    C0 = rng.standard_normal((m, n))
    ref = 0.5 * A.T @ B - 2.0 * C0
    got = _tn(mixed, A, B, 0.5, -2.0, C0)
    bound = TOL * 0.5 * np.abs(A).T @ np.abs(B) + 1e-14 * np.abs(C0)
    err = np.abs(got - ref)
    print(f"TN {m}x{n}x{K}: max err/bound {np.max(err / bound):.3e}")
    assert np.all(err <= bound)


def test_mixed_tn_same_operand(mixed):
    rng = np.random.default_rng(5)
    A = rng.standard_normal((90000, 70))
    got = _tn(mixed, A, A)
    ref = A.T @ A
    bound = TOL * np.abs(A).T @ np.abs(A)
    assert np.all(np.abs(got - ref) <= bound)
    # hi*lo and lo*hi swap places under transposition, so symmetry holds to rounding only
    assert np.all(np.abs(got - got.T) <= bound)


@pytest.mark.parametrize("R,K,n,beta", [(1, 1, 1, 0.0), (1000, 37, 53, 0.0),
                                        (70001, 130, 129, 0.5), (300000, 256, 300, -1.0)])
def test_mixed_nn(mixed, R, K, n, beta):
    rng = np.random.default_rng(R + K + n)
    X = rng.standard_normal((R, K)) * np.exp2(rng.integers(-20, 20, K))
    Cm = rng.standard_normal((K, n)) * np.exp2(rng.integers(-8, 8, (K, 1)))
    Y0 = rng.standard_normal((R, n))
    ref = 1.5 * X @ Cm + beta * Y0
    got = _nn(mixed, X, Cm, 1.5, beta, Y0)
    bound = TOL * 1.5 * np.abs(X) @ np.abs(Cm) + 1e-14 * np.abs(Y0)
    err = np.abs(got - ref)
    print(f"NN {R}x{K}x{n}: max err/bound {np.max(err / bound):.3e}")
    assert np.all(err <= bound)


@pytest.mark.parametrize("R,K,n,beta", [(1000, 37, 53, 0.0), (70001, 130, 300, 0.5),
                                        (300000, 256, 257, -1.0)])
def test_nn_tile_shapes_bitwise(mixed, R, K, n, beta):
    """128-row work-groups (two per CU) and 256-row work-groups (one per CU) give every
    output the same MFMA sequence: the products are equal bit for bit, ragged rows, columns
    and K included."""
    rng = np.random.default_rng(R + n)
    X = rng.standard_normal((R, K))
    Cm = rng.standard_normal((K, n))
    Y0 = rng.standard_normal((R, n))
    out = []
    try:
        for rows in (256, 128):
            mixed.set_param("nn_rows", rows)
            out.append(_nn(mixed, X, Cm, 0.75, beta, Y0))
    finally:
        mixed.set_param("nn_rows", 256)
    np.testing.assert_array_equal(out[0], out[1])


def test_mixed_terms_four_is_tighter(mixed):
    rng = np.random.default_rng(9)
    A = rng.standard_normal((200000, 64))
    ref = A.T @ A
    e3 = np.abs(_tn(mixed, A, A) - ref).max()
    mixed.set_param("mixed_terms", 4)
    try:
        e4 = np.abs(_tn(mixed, A, A) - ref).max()
    finally:
        mixed.set_param("mixed_terms", 3)
    print(f"3 terms {e3:.3e}, 4 terms {e4:.3e}")
    assert e4 <= e3


def _run_scf(kk, s, st, steps, precision):
    kk.set_param("kscg_mode", 1)
    kk.set_param("precision", precision)
    kk.set_param("mixed_min_rows", 0)
    try:
        kw = synth.abi_args(s, nsdmax=4, epssd=1e-12)
        for it in range(1, steps + 1):
            kw["iscf"] = it
            kk.call("scf_diag", **kw, **st)
    finally:
        kk.set_param("precision", 0)
        kk.set_param("mixed_min_rows", 65536)
        kk.set_param("kscg_mode", 0)
    return st


def test_mixed_scf_diag_matches_fp64_eigenvalues(kk, orc):
    """Same start, same number of block-CG/RR steps in FP64 and in split-FP16: eigenvalues
    agree within 1e-6 Ha; the mixed bands come out of the FP64 closing Cholesky-QR
    orthonormal to FP64 rounding."""
    s = synth.make_system(16, 0.4, 4, 24, seed=41)
    st0 = H.init_state(s, orc, smooth=True)
    copy = lambda st: {k: v.copy(order="F") for k, v in st.items()}
    f64 = _run_scf(kk, s, copy(st0), 12, 0)
    mix = _run_scf(kk, s, copy(st0), 12, 1)
    d = np.abs(mix["sval"] - f64["sval"]).max()
    print(f"max |eig mixed - eig fp64| = {d:.3e} Ha")
    assert d <= 1e-6
    q = mix["svecre"]
    g = q.T @ mix["ssvre"] * s.dv
    np.testing.assert_allclose(np.diag(g), 1.0, rtol=0, atol=1e-12)
    np.testing.assert_allclose(g, np.eye(24), rtol=0, atol=1e-12)


def test_mixed_scf_diag_converges_to_dense(kk, orc):
    s = synth.make_system(10, 0.5, 3, 8, seed=8)
    st = _run_scf(kk, s, H.init_state(s, orc, smooth=True), 40, 1)
    w = scipy.linalg.eigh(H.dense_h(s), eigvals_only=True)[:8]
    np.testing.assert_allclose(st["sval"], w, atol=1e-6)


@pytest.mark.parametrize("n,m", [(16, 24), (15, 20)])
def test_rayleigh_ritz_scales_from_h_psi_pass(kk, orc, n, m):
    """The split-precision Rayleigh-Ritz takes psi's split scale from its column maxima and
    H psi's from a bound on its column maxima (stencil part + projector part), both produced
    by the H psi pass, instead of two column-max passes.  psi's scales are the same; H psi's
    differ by a power of two at most, so the two runs are two split-precision runs of the
    same steps: eigenvalues agree to the split-precision bar (1e-6 Ha) and both return bands
    orthonormal to FP64 rounding (FP64 closing Cholesky-QR)."""
    s = synth.make_system(n, 0.4, 3, m, seed=53 + n)
    st0 = H.init_state(s, orc, smooth=True)
    copy = lambda st: {k: v.copy(order="F") for k, v in st.items()}
    out = []
    try:
        for fused in (1, 0):
            kk.set_param("rr_fused_scales", fused)
            out.append(_run_scf(kk, s, copy(st0), 3, 1))
    finally:
        kk.set_param("rr_fused_scales", 1)
    a, b = out
    print(f"max |d eig| {np.abs(a['sval'] - b['sval']).max():.2e} Ha")
    np.testing.assert_allclose(a["sval"], b["sval"], rtol=0, atol=1e-6)
    for st in (a, b):
        g = st["svecre"].T @ st["ssvre"] * s.dv
        np.testing.assert_allclose(np.diag(g), 1.0, rtol=0, atol=1e-12)
        np.testing.assert_allclose(g, np.eye(m), rtol=0, atol=1e-12)


def test_triangular_products_bitwise(kk, orc):
    """Rayleigh-Ritz and Cholesky-QR compute only the triangle their solver reads, and the
    Cholesky-QR rotation skips the zero half of R^-1.  With 300 bands (two 256-band tiles, the
    second ragged) the steps equal those with full products bit for bit."""
    s = synth.make_system(24, 0.4, 3, 300, seed=61)
    st0 = H.init_state(s, orc, smooth=True)
    copy = lambda st: {k: v.copy(order="F") for k, v in st.items()}
    out = []
    try:
        for tri in (1, 0):
            kk.set_param("tn_triangle", tri)
            kk.set_param("nn_triangle", tri)
            out.append(_run_scf(kk, s, copy(st0), 2, 1))
    finally:
        kk.set_param("tn_triangle", 1)
        kk.set_param("nn_triangle", 1)
    a, b = out
    np.testing.assert_array_equal(a["sval"], b["sval"])
    np.testing.assert_array_equal(a["svecre"], b["svecre"])
