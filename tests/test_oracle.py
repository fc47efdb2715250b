"""Pins of the CPU oracle against closed forms, invariants and brute force.

The reference ships no tests, fixtures or golden vectors, and its Fortran cannot be
built here (it needs MPI and BLAS/LAPACK), so every oracle routine is checked against
mathematics that fixes its result: FD polynomial exactness, plane-wave eigenvectors of
the periodic stencil, a dense projector matrix, an FFT Poisson solve, and dense
eigensolves of the same discrete Hamiltonian.
"""
import math

import numpy as np
import pytest
import scipy.linalg

from rspacedist_amd import synth
from tests import helpers as H


@pytest.mark.parametrize("nf", range(1, 9))
def test_fd_weights_polynomial_exactness(nf):
    a = H.fd_weights(nf)
    m = np.arange(1, nf + 1)
    assert abs(a[0] + 2 * a[1:nf + 1].sum()) < 1e-12
    assert abs(2 * (a[1:nf + 1] * m ** 2).sum() - 2.0) < 1e-12
    for k in range(2, nf + 1):
        assert abs(2 * (a[1:nf + 1] * m ** (2 * k)).sum()) < 1e-9 * math.factorial(2 * k)
    assert np.all(a[nf + 1:] == 0)


@pytest.mark.parametrize("nperi", [0, 1, 2, 3])
@pytest.mark.parametrize("nsnd", [1, 4])
def test_halo_wrap(orc, nperi, nsnd):
    rng = np.random.default_rng(3)
    v = np.asfortranarray(rng.standard_normal((9, 8, 7)))
    nf = 4
    got = np.zeros((9 + 2 * nf, 8 + 2 * nf, 7 + 2 * nf), order="F")
    got[nf:nf + 9, nf:nf + 8, nf:nf + 7] = v
    orc.call("overlap_finitedifference_r", nperi=nperi, ncpa=9, ncpb=8, ncpc=7, nf=nf,
             nf1=nf - 1, nsnd=nsnd, v_cc=got)
    np.testing.assert_array_equal(got, H.halo(v, nf, nsnd, nperi))


@pytest.mark.parametrize("nf", [1, 2, 4, 6, 8])
def test_stencil_plane_wave_eigenvector(orc, nf):
    n, h = 16, 0.3
    kx, ky, kz = 2 * np.pi * np.array([1, 3, 5]) / (n * h)
    psi = H.plane_wave(n, h, (kx, ky, kz))
    veff = np.asfortranarray(np.full((n, n, n), 0.37))
    vre = H.halo(psi, nf, nf)
    avre = np.zeros((n, n, n, 2), order="F")
    avre[..., 1] = 1.5  # accumulate semantics: avre(:,:,:,l) += H psi
    orc.call("kslaplacian_r", ncpx=n, ncpy=n, ncpz=n, neigmx=2, nf=nf, dx=h, dy=h, dz=h,
             vre=vre, veff=veff, l=2, avre=avre)
    lam = sum(H.fd_symbol(k, h, nf) for k in (kx, ky, kz)) + 0.37
    np.testing.assert_allclose(avre[..., 1], 1.5 + lam * psi, atol=1e-9 * abs(lam))
    assert np.all(avre[..., 0] == 0)


def test_stencil_anisotropic_spacing_matches_dense(orc):
    s = synth.make_system(8, 0.45, 1, 4, seed=2)
    rng = np.random.default_rng(1)
    psi = np.asfortranarray(rng.standard_normal((8, 8, 8)))
    dx, dy, dz = 0.4, 0.45, 0.5
    avre = np.zeros((8, 8, 8, 1), order="F")
    orc.call("kslaplacian_r", ncpx=8, ncpy=8, ncpz=8, neigmx=1, nf=4, dx=dx, dy=dy, dz=dz,
             vre=H.halo(psi, 4, 4), veff=s.veff, l=1, avre=avre)
    I = np.eye(8)
    T = (np.kron(I, np.kron(I, H.circulant_d2(8, dx, 4))) + np.kron(I, np.kron(H.circulant_d2(8, dy, 4), I))
         + np.kron(H.circulant_d2(8, dz, 4), np.kron(I, I)))
    ref = T @ psi.ravel(order="F") + s.veff.ravel(order="F") * psi.ravel(order="F")
    np.testing.assert_allclose(avre.ravel(order="F"), ref, rtol=1e-12, atol=1e-10)


def test_preconditioner_symbol(orc):
    n, h = 12, 0.25
    k = 2 * np.pi * np.array([2, 1, 4]) / (n * h)
    psi = H.plane_wave(n, h, k)
    out = np.zeros((n, n, n), order="F")
    orc.call("ksprecondition", ncpx=n, ncpy=n, ncpz=n, nf=4, ncol=1, vre=H.halo(psi, 4, 1), avre=out)
    sym = 0.5 + (1.0 / 6.0) * np.cos(k * h).sum()
    np.testing.assert_allclose(out, sym * psi, atol=1e-13)


def _nl_setup():
    s = synth.make_system(12, 0.4, 3, 4, seed=5)
    rng = np.random.default_rng(9)
    dmat = rng.uniform(-0.5, 0.5, (9, 9))  # non-symmetric on purpose: catches D vs D^T
    s.dij = np.asfortranarray(np.repeat(dmat[:, :, None, None], s.natom, axis=3))
    s.dmat = dmat
    return s, rng


def test_nonlocal_r01_matches_dense(orc):
    s, rng = _nl_setup()
    psi = rng.standard_normal(s.npts)
    rp = np.zeros((9, s.natom), order="F")
    orc.call("nonlocaloperation_r_01", **synth.abi_args(s), vre2=psi, rpsep=rp)
    P = H.dense_projectors(s)
    np.testing.assert_allclose(rp.ravel(order="F"), P.T @ psi, rtol=1e-12, atol=1e-12)


def test_nonlocal_r02_matches_dense_and_s_form(orc):
    s, rng = _nl_setup()
    rp = np.asfortranarray(rng.standard_normal((9, s.natom)))
    avre = np.asfortranarray(rng.standard_normal(s.npts))
    base = avre.copy()
    orc.call("nonlocaloperation_r_02", **synth.abi_args(s), nhs=1, rpsep=rp, dijsss=s.dij,
             avre=avre, avr=np.zeros(s.num_list))
    P = H.dense_projectors(s)
    Dbig = np.kron(np.eye(s.natom), s.dmat)
    coef = Dbig.T @ rp.ravel(order="F") / s.dv
    np.testing.assert_allclose(avre - base, P @ coef, rtol=1e-11, atol=1e-10)
    # S form with q = 0 is the identity
    out = base.copy()
    orc.call("nonlocaloperation_r_02", **synth.abi_args(s), nhs=0, rpsep=rp, dijsss=s.sss,
             avre=out, avr=np.zeros(s.num_list))
    np.testing.assert_array_equal(out, base)


def test_nonlocal_operator_symmetric_for_symmetric_d(orc):
    s = synth.make_system(12, 0.4, 3, 4, seed=5)
    P = H.dense_projectors(s)
    V = P @ np.kron(np.eye(s.natom), s.dmat).T @ P.T
    np.testing.assert_allclose(V, V.T, atol=1e-14)
    assert np.linalg.matrix_rank(V) <= 9 * s.natom


def test_projectors_normalised_and_smooth():
    s = synth.make_system(48, 0.1, 1, 1, seed=4)
    pv = s.vnlocp[: s.natinf[0], :, 0] / s.dv
    norms = (pv ** 2).sum(0) * s.dv
    np.testing.assert_allclose(norms, 1.0, rtol=2e-3)
    gram = pv.T @ pv * s.dv
    np.testing.assert_allclose(gram, np.eye(9), atol=2e-3)


def test_hartree_matches_fft_solution(orc):
    n, h = 16, 0.35
    xmax = n * h / 2
    rng = np.random.default_rng(11)
    rho = np.asfortranarray(rng.random((n, n, n)))
    vh = np.zeros((n, n, n), order="F")
    z = np.zeros(1)
    orc.call("scf_hartree", nd=1, ncpx_d=n, ncpy_d=n, ncpz_d=n, nfh=4, ndisp=6, nperi=3,
             ncgres=801, ncgmin=1, ncgmax=800, xmax=xmax, ymax=xmax, zmax=xmax, epsvh=1e-12,
             rho_aug_dense=rho, boundx=z, boundy=z, boundz=z, vh_dense=vh)
    orc.call("tools_correctpotential", ncpx_d=n, ncpy_d=n, ncpz_d=n, xmax=xmax, ymax=xmax,
             zmax=xmax, ddx=h, ddy=h, ddz=h, vh_dense=vh)
    k = 2 * np.pi * np.fft.fftfreq(n, d=h)
    sym = (H.fd_symbol(k, h, 4)[:, None, None] + H.fd_symbol(k, h, 4)[None, :, None]
           + H.fd_symbol(k, h, 4)[None, None, :])
    bh = np.fft.fftn(2 * np.pi * rho)
    bh[0, 0, 0] = 0
    sym[0, 0, 0] = 1
    ref = np.real(np.fft.ifftn(bh / sym))
    ref -= ref.mean()
    assert abs(vh.mean()) < 1e-12
    np.testing.assert_allclose(vh, ref, atol=1e-9 * np.abs(ref).max())


def test_hartree_isolated_dirichlet_matches_dense(orc):
    """nperi = 0: boundary planes enter b as -sum c_m bound; compare with a dense solve of the
    operator with those planes held fixed."""
    n, h, nfh = 6, 0.5, 4
    xmax = n * h / 2
    rng = np.random.default_rng(2)
    rho = np.asfortranarray(rng.random((n, n, n)))
    bx = np.asfortranarray(rng.random((2 * nfh, n, n)))
    by = np.asfortranarray(rng.random((n, 2 * nfh, n)))
    bz = np.asfortranarray(rng.random((n, n, 2 * nfh)))
    vh = np.zeros((n, n, n), order="F")
    orc.call("scf_hartree", nd=0, ncpx_d=n, ncpy_d=n, ncpz_d=n, nfh=nfh, ndisp=6, nperi=0,
             ncgres=801, ncgmin=1, ncgmax=2000, xmax=xmax, ymax=xmax, zmax=xmax, epsvh=1e-13,
             rho_aug_dense=rho, boundx=bx, boundy=by, boundz=bz, vh_dense=vh)
    # dense: field on an extended grid with fixed boundary planes
    a = H.fd_weights(nfh)
    N = n + 2 * nfh
    full = np.zeros((N, N, N))
    full[:nfh, nfh:nfh + n, nfh:nfh + n] = bx[:nfh]
    full[nfh + n:, nfh:nfh + n, nfh:nfh + n] = bx[nfh:]
    full[nfh:nfh + n, :nfh, nfh:nfh + n] = by[:, :nfh].transpose(0, 1, 2)
    full[nfh:nfh + n, nfh + n:, nfh:nfh + n] = by[:, nfh:]
    full[nfh:nfh + n, nfh:nfh + n, :nfh] = bz[:, :, :nfh]
    full[nfh:nfh + n, nfh:nfh + n, nfh + n:] = bz[:, :, nfh:]
    idx = np.arange(n ** 3).reshape((n, n, n), order="F")
    A = np.zeros((n ** 3, n ** 3))
    b = (2 * np.pi * rho).ravel(order="F").copy()
    for ix in range(n):
        for iy in range(n):
            for iz in range(n):
                r = idx[ix, iy, iz]
                A[r, r] += -0.5 * 3 * a[0] / h ** 2
                for ax in range(3):
                    for m in range(1, nfh + 1):
                        for sg in (-1, 1):
                            p = [ix, iy, iz]
                            p[ax] += sg * m
                            c = -0.5 * a[m] / h ** 2
                            if 0 <= p[ax] < n:
                                A[r, idx[p[0], p[1], p[2]]] += c
                            else:
                                b[r] -= c * full[p[0] + nfh, p[1] + nfh, p[2] + nfh]
    ref = np.linalg.solve(A, b)
    np.testing.assert_allclose(vh.ravel(order="F"), ref, rtol=1e-8, atol=1e-8)


def test_orthogonalization_is_gram_schmidt(orc):
    s = synth.make_system(10, 0.5, 2, 40, seed=6)  # 40 bands: exercises the mstep=32 block path
    psi0 = synth.random_bands(s)
    st = {"svecre": psi0.copy(order="F"), "ssvre": np.zeros_like(psi0, order="F"),
          "rspsep": np.zeros((9, s.natom, 40), order="F")}
    orc.call("orthogonalization_r", **synth.abi_args(s), nflag=1, **st)
    q = st["svecre"]
    np.testing.assert_allclose(q.T @ q * s.dv, np.eye(40), atol=1e-12)
    # same Q as a QR factorisation with positive diagonal (classical GS, exact arithmetic)
    qr, r = np.linalg.qr(psi0)
    qr = qr * np.sign(np.diag(r))[None, :] / math.sqrt(s.dv)
    np.testing.assert_allclose(q, qr, atol=1e-10)
    # S psi and <p|psi> recomputed consistently
    P = H.dense_projectors(s)
    np.testing.assert_allclose(st["ssvre"], q, atol=1e-14)
    np.testing.assert_allclose(st["rspsep"].reshape(9 * s.natom, 40, order="F"), P.T @ q, atol=1e-11)


def test_rayleigh_ritz_matches_dense_projection(orc):
    s = synth.make_system(8, 0.55, 2, 6, seed=7)
    st = H.init_state(s, orc)
    psi0 = st["svecre"].copy()
    orc.call("rayleighritz_r", **synth.abi_args(s), l_rrb=1, **{k: st[k] for k in
             ("rspsep", "svecre", "ssvre", "hsvre", "sval")})
    Hd = H.dense_h(s)
    A = psi0.T @ Hd @ psi0 * s.dv
    w, z = scipy.linalg.eigh(A)
    np.testing.assert_allclose(st["sval"], w, rtol=1e-11, atol=1e-11)
    rot = psi0 @ z
    for j in range(6):  # eigenvectors up to sign
        c = np.sign(rot[:, j] @ st["svecre"][:, j])
        np.testing.assert_allclose(st["svecre"][:, j], c * rot[:, j], atol=1e-9)
    np.testing.assert_allclose(st["hsvre"], st["ssvre"], atol=0)  # hsvre = S psi rotated


def _converge(orc, s, nscf, ncgscf=200, m=None, **over):
    st = H.init_state(s, orc, nbands=m, smooth=True)
    kw = synth.abi_args(s, neigmx=m or s.nbands, **over)
    for it in range(1, nscf + 1):
        kw["iscf"] = it
        kw["ncgscf"] = ncgscf
        orc.call("scf_diag", **kw, **st)
    return st


def test_scf_diag_cg_converges_to_dense_eigenvalues(orc):
    s = synth.make_system(10, 0.5, 3, 8, seed=8)
    st = _converge(orc, s, 15, nsdmax=40, epssd=1e-12)
    w = scipy.linalg.eigh(H.dense_h(s), eigvals_only=True)[:8]
    np.testing.assert_allclose(st["sval"], w, atol=1e-8)
    assert np.all(np.diff(st["sval"]) >= 0)


def test_scf_diag_diis_converges_to_dense_eigenvalues(orc):
    s = synth.make_system(10, 0.5, 3, 8, seed=8)
    st = _converge(orc, s, 30, ncgscf=2, ndiismax=10, epssd=1e-12)
    w = scipy.linalg.eigh(H.dense_h(s), eigvals_only=True)[:8]
    np.testing.assert_allclose(st["sval"], w, atol=1e-8)


def test_s_metric_orthogonalization_is_gram_schmidt_in_s(orc):
    """S != 1 (sss positive semi-definite): orthogonalization_r gives Q^T S Q = I, S psi and
    <p|psi> consistent with the dense S and projectors, and the same Q as Gram-Schmidt in
    the S inner product (the Cholesky factor of X^T S X, orthogonalization8f.F90:144-256)."""
    s = synth.make_system(10, 0.5, 2, 40, seed=6, q_scale=2.0)
    psi0 = synth.random_bands(s)
    st = {"svecre": psi0.copy(order="F"), "ssvre": np.zeros_like(psi0, order="F"),
          "rspsep": np.zeros((9, s.natom, 40), order="F")}
    orc.call("orthogonalization_r", **synth.abi_args(s), nflag=1, **st)
    Sd = H.dense_s(s)
    q = st["svecre"]
    np.testing.assert_allclose(q.T @ Sd @ q * s.dv, np.eye(40), atol=1e-12)
    np.testing.assert_allclose(st["ssvre"], Sd @ q, atol=1e-12)
    P = H.dense_projectors(s)
    np.testing.assert_allclose(st["rspsep"].reshape(9 * s.natom, 40, order="F"), P.T @ q, atol=1e-11)
    # Gram-Schmidt in the S metric: X = Q R with R = chol(dV X^T S X)^T upper triangular
    R = np.linalg.cholesky(psi0.T @ Sd @ psi0 * s.dv).T
    np.testing.assert_allclose(q, scipy.linalg.solve_triangular(R, psi0.T, trans="T").T, atol=1e-10)
    # a dropped S term is visible: the plain Gram-Schmidt Q differs
    assert np.abs(q.T @ q * s.dv - np.eye(40)).max() > 1e-3


def test_s_metric_rayleigh_ritz_matches_dense_projection(orc):
    s = synth.make_system(8, 0.55, 2, 6, seed=7, q_scale=2.0)
    st = H.init_state(s, orc)
    psi0 = st["svecre"].copy()
    orc.call("rayleighritz_r", **synth.abi_args(s), l_rrb=1, **{k: st[k] for k in
             ("rspsep", "svecre", "ssvre", "hsvre", "sval")})
    Hd, Sd = H.dense_h(s), H.dense_s(s)
    w, z = scipy.linalg.eigh(psi0.T @ Hd @ psi0 * s.dv, psi0.T @ Sd @ psi0 * s.dv)
    np.testing.assert_allclose(st["sval"], w, rtol=1e-11, atol=1e-11)
    q = st["svecre"]
    np.testing.assert_allclose(q.T @ Sd @ q * s.dv, np.eye(6), atol=1e-11)
    np.testing.assert_allclose(st["ssvre"], Sd @ q, atol=1e-11)


@pytest.mark.parametrize("ncgscf", [200, 2])  # CG (kscg_r), then DIIS after two CG steps
def test_s_metric_scf_diag_converges_to_generalized_eigenvalues(orc, ncgscf):
    """kscg_r / ksdiis_r with S != 1 converge to the generalized eigenpairs H psi = eps S psi
    of the dense H and S (scipy.linalg.eigh(H, S)); they differ from the S = 1 eigenvalues
    by ~1e-2 Ha here, so a dropped or misplaced S term fails."""
    s = synth.make_system(10, 0.5, 3, 8, seed=8, q_scale=2.0)
    st = _converge(orc, s, 60 if ncgscf > 2 else 100, ncgscf=ncgscf, nsdmax=4, epssd=1e-12)
    Hd, Sd = H.dense_h(s), H.dense_s(s)
    w = scipy.linalg.eigh(Hd, Sd, eigvals_only=True)[:8]
    w1 = scipy.linalg.eigh(Hd, eigvals_only=True)[:8]
    assert np.abs(w - w1).max() > 1e-3
    np.testing.assert_allclose(st["sval"], w, atol=1e-8)
    q = st["svecre"]
    np.testing.assert_allclose(q.T @ Sd @ q * s.dv, np.eye(8), atol=1e-12)


def _dense_precond(n):
    I = np.eye(n)
    S = np.roll(I, 1, axis=0) + np.roll(I, -1, axis=0)
    return (0.5 * np.eye(n ** 3) + (np.kron(I, np.kron(I, S)) + np.kron(I, np.kron(S, I))
                                    + np.kron(S, np.kron(I, I))) / 12.0)


def test_ksdiis_single_step_matches_dense_rmm_diis(orc):
    """One RMM-DIIS pass (ksdiis8f.F90:97-600, nitermax=2) restated on dense matrices: the
    preconditioned residual step with the quadratic line search for lambda, then the
    2-vector residual minimisation (min <R|R>/<phi|phi>, R_j = (H - eps) phi_j)."""
    s = synth.make_system(8, 0.55, 2, 3, seed=13)
    st = H.init_state(s, orc, smooth=True)
    Hd, K, dv = H.dense_h(s), _dense_precond(8), s.dv
    psi0 = st["svecre"].copy()
    kw = synth.abi_args(s, ndiismax=2)
    orc.call("ksdiis_r", **kw, **{k: st[k] for k in ("rspsep", "sval", "svecre", "ssvre",
                                                       "residual_states", "hsvre")})
    for l in range(3):
        p0 = psi0[:, l]
        eps = p0 @ Hd @ p0 * dv
        kr = K @ (Hd @ p0 - eps * p0)
        a1 = 0.5
        aaa = 2.0 * (kr @ Hd @ p0 - eps * (kr @ p0)) * dv
        p1 = p0 + a1 * kr
        bbb = ((p1 @ Hd @ p1) / (p1 @ p1) - eps - aaa * a1) / a1 ** 2
        lam = min(max(0.5 * aaa / bbb, 0.1), 1.0)
        phis = np.stack([p0, p0 + lam * kr], axis=1)
        R = Hd @ phis - eps * phis
        w, v = scipy.linalg.eigh(R.T @ R * dv, phis.T @ phis * dv)
        c = v[:, 0] / math.sqrt(v[:, 0] @ (phis.T @ phis * dv) @ v[:, 0])
        c *= np.sign(c[0])
        new = phis @ c
        assert st["sval"][l] == pytest.approx(eps, abs=1e-12)
        np.testing.assert_allclose(st["svecre"][:, l], new, atol=1e-10)
        np.testing.assert_allclose(st["hsvre"][:, l], Hd @ new, atol=1e-8)
        assert st["residual_states"][l] == pytest.approx(w[0], rel=1e-8)


def test_kscg_single_band_is_steepest_descent_on_rayleigh_quotient(orc):
    """With one band and one iteration the 2x2 step must lower the Rayleigh quotient to the
    lower eigenvalue of the 2x2 pencil [[eps, s],[s, pap]] (kscg_r_05 closed form)."""
    s = synth.make_system(8, 0.55, 2, 1, seed=10)
    st = H.init_state(s, orc)
    Hd = H.dense_h(s)
    psi0 = st["svecre"][:, 0].copy()
    kw = synth.abi_args(s, nsdmax=1)
    orc.call("kscg_r", **kw, **{k: st[k] for k in ("sval", "rspsep", "svecre", "ssvre",
                                                     "residual_states", "hsvre")})
    psi1 = st["svecre"][:, 0]
    rq0 = psi0 @ Hd @ psi0 * s.dv
    rq1 = psi1 @ Hd @ psi1 / (psi1 @ psi1)
    assert rq1 < rq0
    np.testing.assert_allclose(psi1 @ psi1 * s.dv, 1.0, atol=1e-12)
    np.testing.assert_allclose(st["hsvre"][:, 0], Hd @ psi1, atol=1e-9)


def test_reorderstates_stable_rank_sort(orc):
    s = synth.make_system(6, 0.6, 1, 5, seed=1)
    rng = np.random.default_rng(0)
    sval = np.array([0.3, -0.1, 0.3, -0.5, 0.2])
    psi = np.asfortranarray(rng.standard_normal((s.npts, 5)))
    spsi = np.asfortranarray(rng.standard_normal((s.npts, 5)))
    rsp = np.asfortranarray(rng.standard_normal((9, 1, 5)))
    res = np.arange(5.0)
    hs = np.zeros_like(psi, order="F")
    p0, s0, r0 = psi.copy(), spsi.copy(), rsp.copy()
    orc.call("reorderstates", **synth.abi_args(s, neigmx=5), ns1=1, sval=sval, residual_states=res,
             rspsep=rsp, svecre=psi, ssvre=spsi, hsvre=hs)
    order = [3, 1, 4, 0, 2]  # stable: the two 0.3 keep their order
    np.testing.assert_array_equal(sval, [-0.5, -0.1, 0.2, 0.3, 0.3])
    np.testing.assert_array_equal(res, np.array(order, dtype=float))
    np.testing.assert_array_equal(psi, p0[:, order])
    np.testing.assert_array_equal(spsi, s0[:, order])
    np.testing.assert_array_equal(rsp, r0[:, :, order])
    np.testing.assert_array_equal(hs, s0[:, order])


def test_charge_density_closed_form(orc):
    rng = np.random.default_rng(4)
    psi = np.asfortranarray(rng.standard_normal((5 * 6 * 7, 3)))
    f = np.array([2.0, 1.0, 0.5])
    rho = np.zeros(5 * 6 * 7)
    orc.call("scf_charge_smt_r", neigmx=3, nums=1, nspv=1, numk=1, ncpx=5, ncpy=6, ncpz=7,
             fnele=f, svecre=psi, rhosmt=rho)
    np.testing.assert_allclose(rho, (psi ** 2) @ f, rtol=1e-14)


def test_totalenergy_eig_closed_form(orc):
    sval = np.array([-1.0, 0.5, 2.0])
    f = np.array([1.0, 0.5, 0.0])
    out = orc.call("totalenergy_eig", nums=1, ncol=1, numk=1, neigmx=3, nwskk=np.array([1], dtype=np.int32),
                   nwskptot=2, tf=0.01, sval=sval, fnele=f)
    assert out["eneel"] == pytest.approx(-1.0 + 0.25)
    # tmpfac = nwskptot*nums/(2*nwskk) = 1: only f=0.5 contributes 0.5 ln 0.5 * 2
    assert out["eneth"] == pytest.approx(0.01 * (2 * 0.5 * math.log(0.5)))


@pytest.mark.parametrize("symmetric", [True, False])
def test_force_eig_matches_analytic_projector_gradient(orc, symmetric):
    """force_eig_r uses <p_j|d psi/dx> (FD derivative of psi, force8f.F90:119-131).  By
    integration by parts that is d<p_j|psi>/dR_a = -<dp_j/dr|psi>, computed here from the
    analytic projector gradient.  For symmetric D the result must be the physical force
    -dE_nl/dR_a; a non-symmetric D pins the index order D_ij of force8f.F90:157-162."""
    rng = np.random.default_rng(12)
    s = synth.make_system(24, 0.2, 2, 2, seed=12)
    dmat = rng.uniform(-0.5, 0.5, (9, 9))
    if symmetric:
        dmat = 0.5 * (dmat + dmat.T)
    s.dij = np.asfortranarray(np.repeat(dmat[:, :, None, None], s.natom, axis=3))
    c = s.grid_coords()
    X, Y, Z = np.meshgrid(c, c, c, indexing="ij")
    psi = np.exp(-((X - 0.3) ** 2 + (Y + 0.2) ** 2 + (Z - 0.1) ** 2) / (2 * 1.2 ** 2))
    psi = np.asfortranarray(psi.reshape(-1, 1, order="F") / math.sqrt((psi ** 2).sum() * s.dv))
    fnele = np.array([2.0])
    sval = np.array([0.7])
    fx, fy, fz = np.zeros(s.natom), np.zeros(s.natom), np.zeros(s.natom)
    orc.call("force_eig_r", **synth.abi_args(s, neigmx=1), svecre=psi, sval=sval, fnele=fnele,
             fatx=fx, faty=fy, fatz=fz)
    ref = np.zeros((s.natom, 3))
    L = s.n * s.h
    for a in range(s.natom):
        m = s.natinf[a]
        idx = s.lstvec2[:m, a] - 1
        pts = np.stack([X.ravel(order="F")[idx], Y.ravel(order="F")[idx], Z.ravel(order="F")[idx]])
        d = pts - s.atoms[a][:, None]
        d -= L * np.round(d / L)
        pv = synth.projector_values(*d)
        gp = synth.projector_grad(*d)
        pw = psi[idx, 0]
        cp = pv @ pw * s.dv                      # <p_i|psi>
        for ax in range(3):
            dcp = -(gp[ax] @ pw) * s.dv          # d<p_i|psi>/dR
            if symmetric:                        # -dE/dR, E = f c^T D c
                ref[a, ax] = -fnele[0] * (dcp @ dmat @ cp + cp @ dmat @ dcp)
            else:
                ref[a, ax] = -2.0 * fnele[0] * (cp @ dmat @ dcp)
    got = np.stack([fx, fy, fz], axis=1)
    np.testing.assert_allclose(got, ref, rtol=2e-3, atol=2e-3 * np.abs(ref).max())


# ------------------------------------------------------------------ differentiation pattern
def _adj_system(seed=51, n=12, h=0.45, natom=3, nb=6):
    s = synth.make_system(n, h, natom, nb, seed=seed)
    psi = synth.smooth_bands(synth.random_bands(s), s.n, s.h)
    f = np.linspace(1.0, 0.5, nb)
    return s, np.asfortranarray(psi), f


def _orc_band_energy(orc, s, psi, atoms, veff=None):
    pr = synth.build_projectors(s.n, s.h, atoms)
    kw = synth.abi_args(s)
    kw.update(pr)
    kw["num_list"] = pr["vnlocp"].shape[0]
    if veff is not None:
        kw["veff"] = veff
    e = np.zeros(s.nbands)
    orc.call("band_energy_r", **kw, svecre=psi, eband=e)
    return e


def test_band_energy_matches_dense(orc):
    s, psi, _ = _adj_system()
    e = _orc_band_energy(orc, s, psi, s.atoms)
    Hd = H.dense_h(s)
    ref = np.einsum("il,ij,jl->l", psi, Hd, psi) * s.dv
    np.testing.assert_allclose(e, ref, rtol=1e-12, atol=1e-12)


def test_energy_adjoint_matches_central_differences(orc):
    """dE/dR (projector part, veff fixed) against central differences of the oracle's own
    band energy with the projectors rebuilt at R +- 1e-4; dE/dveff along a random direction
    (E is linear in veff, so the difference quotient is exact up to rounding)."""
    s, psi, f = _adj_system()
    pr = synth.build_projectors(s.n, s.h, s.atoms, derivatives=True)
    kw = synth.abi_args(s)
    kw.update({k: v for k, v in pr.items() if k != "dvnlocp"})
    kw["num_list"] = pr["vnlocp"].shape[0]
    dedv = np.zeros(s.npts)
    dedr = np.zeros((s.natom, 3), order="F")
    orc.call("energy_adjoint_r", **kw, dvnlocp=pr["dvnlocp"], svecre=psi, fnele=f, dedv=dedv,
             dedr=dedr)
    np.testing.assert_allclose(dedv, (psi ** 2 @ f) * s.dv, rtol=1e-13, atol=1e-16)
    hd = 1e-4
    for a in range(s.natom):
        for d in range(3):
            Rp, Rm = s.atoms.copy(), s.atoms.copy()
            Rp[a, d] += hd
            Rm[a, d] -= hd
            fd = (f @ _orc_band_energy(orc, s, psi, Rp) - f @ _orc_band_energy(orc, s, psi, Rm)) / (2 * hd)
            assert abs(dedr[a, d] - fd) <= 1e-7 * max(1.0, abs(fd)), (a, d, dedr[a, d], fd)
    rng = np.random.default_rng(3)
    dv = rng.standard_normal(s.npts)
    v0 = s.veff.reshape(-1, order="F")
    ep = f @ _orc_band_energy(orc, s, psi, s.atoms, veff=np.asfortranarray(v0 + 1e-3 * dv))
    em = f @ _orc_band_energy(orc, s, psi, s.atoms, veff=np.asfortranarray(v0 - 1e-3 * dv))
    np.testing.assert_allclose((ep - em) / 2e-3, dedv @ dv, rtol=1e-9)


def test_energy_adjoint_agrees_with_force_eig_r(orc):
    """Translation invariance: <dp/dR|psi> = <p|grad psi>, so the adjoint's projector force
    equals force_eig_r's (which differentiates psi by finite differences, force8f.F90:468-)
    up to that derivative's discretisation error (smooth bands, symmetric D, S = 0)."""
    s, psi, f = _adj_system(n=16, h=0.4, natom=3, nb=6)
    pr = synth.build_projectors(s.n, s.h, s.atoms, derivatives=True)
    kw = synth.abi_args(s)
    kw.update({k: v for k, v in pr.items() if k != "dvnlocp"})
    kw["num_list"] = pr["vnlocp"].shape[0]
    dedv = np.zeros(s.npts)
    dedr = np.zeros((s.natom, 3), order="F")
    orc.call("energy_adjoint_r", **kw, dvnlocp=pr["dvnlocp"], svecre=psi, fnele=f, dedv=dedv,
             dedr=dedr)
    fx, fy, fz = np.zeros(s.natom), np.zeros(s.natom), np.zeros(s.natom)
    orc.call("force_eig_r", **kw, svecre=psi, sval=np.zeros(s.nbands), fnele=f, fatx=fx,
             faty=fy, fatz=fz)
    fref = np.stack([fx, fy, fz], axis=1)
    err = np.abs(-dedr - fref).max()
    print(f"adjoint vs force_eig_r: max diff {err:.3e}, max |F| {np.abs(fref).max():.3e}")
    assert err <= 2e-3 * np.abs(fref).max()


def test_band_sd_step_matches_dense_restatement(orc):
    """One batched-path band step on the CPU oracle against the same step written with the
    dense H of the 12^3 system and the dense 7-point preconditioner (1/2 I + 1/12 sum of the
    six periodic neighbours), eigh of the projected 2x2 problem, lower vector with a >= 0."""
    s = synth.make_system(12, 0.45, 3, 4, seed=61)
    psi0 = synth.smooth_bands(synth.random_bands(s), s.n, s.h)
    q, r = np.linalg.qr(psi0)
    psi0 = np.asfortranarray(q * np.sign(np.diag(r)) / np.sqrt(s.dv))
    psi = psi0.copy(order="F")
    sval = np.zeros(s.nbands)
    dij = np.asfortranarray(s.dij.reshape(9, 9, s.natom, order="F"))
    orc.call("band_sd_step", n=s.n, h=s.h, nbands=s.nbands, natom=s.natom, num_list=s.num_list,
             lstvec2=s.lstvec2, natinf=s.natinf, vnlocp=s.vnlocp, dij=dij,
             veff=np.asfortranarray(s.veff[..., 0]), svecre=psi, sval=sval)
    Hd = H.dense_h(s)
    n = s.n

    def K(v):
        g = v.reshape((n, n, n), order="F")
        out = 0.5 * g
        for ax in range(3):
            out = out + (np.roll(g, 1, ax) + np.roll(g, -1, ax)) / 12.0
        return out.reshape(-1, order="F")

    for l in range(s.nbands):
        x = psi0[:, l]
        e = x @ Hd @ x * s.dv
        p = K(-(Hd @ x - e * x))
        p = p - (x @ p * s.dv) * x
        p = p / np.sqrt(p @ p * s.dv)
        A = np.array([[e, x @ Hd @ p * s.dv], [p @ Hd @ x * s.dv, p @ Hd @ p * s.dv]])
        w, v = np.linalg.eigh(0.5 * (A + A.T))
        a, c = v[:, 0] * np.sign(v[0, 0])
        np.testing.assert_allclose(sval[l], w[0], rtol=1e-11)
        np.testing.assert_allclose(psi[:, l], a * x + c * p, atol=1e-10 * np.abs(x).max())
