"""Pins of the oracle's complex (Bloch, k-point) path (oracle/kukan_oracle_bloch.c) by closed
forms, dense operators built from the definitions and invariants (CPU).

- kslaplacian_c: a plane wave e^{i q.r} (q on the reciprocal grid) is an eigenvector of
  H_k = e^{-i k.r} (-1/2 lap_nf + V) e^{i k.r} with V = 0, eigenvalue
  sum_axes -1/2 [c0 + 2 sum_m c_m cos(m (k + q)_a h_a)] / h_a^2.  A wrong phase sign shows as
  k - q in place of k + q; at k = 0 the complex path equals the real one.
- nonlocaloperation_c_01/_02: the dense projector matrix P[r, 9a + i] = vnlocp(i, a) e^{i k.r}
  (r the projector point's unwrapped position) gives <p|u> = P^T u and the expansion
  conj(P) X c / dV; the resulting V_nl and S are Hermitian.
- The dense H_k and S_k assembled column by column from the oracle's own applications are
  Hermitian and equal the operator built from the definitions above; kscg_c (through
  scf_diag, nrc = 1) converges to eigh(H_k, S_k), rayleighritz_c reproduces the dense
  projected eigenproblem, orthogonalization_c gives Q^H S Q = I with the Gram-Schmidt Q.
- zheev (the oracle's Hermitian Jacobi) against numpy.linalg.eigh.
"""
import numpy as np
import pytest
import scipy.linalg

from rspacedist_amd import synth
from tests import helpers as H

K = (0.31, -0.22, 0.17)  # a generic k-point (1/bohr): no degeneracies from symmetry


def halo_c(u, nf, nsnd, nperi=3):
    return np.asfortranarray(H.halo(np.asfortranarray(u.real), nf, nsnd, nperi)
                             + 1j * H.halo(np.asfortranarray(u.imag), nf, nsnd, nperi))


def grid(n, h):
    j = np.arange(1, n + 1) * h - 0.5 * n * h - 0.5 * h
    return np.meshgrid(j, j, j, indexing="ij")


def lap_c(orc, u, n, h, k, nf=4, veff=None):
    """H_k u through orc_kslaplacian_c (one band)."""
    xm = 0.5 * n * h
    out = np.zeros((n ** 3, 1), dtype=np.complex128, order="F")
    v = halo_c(u.reshape((n, n, n), order="F"), nf, nf)
    orc.call("kslaplacian_c", ncpx=n, ncpy=n, ncpz=n, neigmx=1, ncol=1, nvef=1, nf=nf, dx=h, dy=h,
             dz=h, xmax=xm, ymax=xm, zmax=xm, skpx=k[0], skpy=k[1], skpz=k[2], vcm=v,
             veff=np.zeros((n, n, n), order="F") if veff is None else veff, l=1, avcm=out,
             workc=np.zeros((n, n, n), dtype=np.complex128, order="F"))
    return out[:, 0]


@pytest.mark.parametrize("nf", [1, 4, 8])
def test_kslaplacian_c_bloch_plane_wave(orc, nf):
    n, h = 12, 0.4
    L = n * h
    X, Y, Z = grid(n, h)
    a = H.fd_weights(nf)
    for q in [(1, 0, 0), (2, -1, 3), (-3, 2, 1)]:
        qv = 2 * np.pi * np.array(q) / L
        u = np.exp(1j * (qv[0] * X + qv[1] * Y + qv[2] * Z)).ravel(order="F")
        got = lap_c(orc, u, n, h, K, nf)
        lam = sum(-0.5 * (a[0] + 2 * sum(a[m] * np.cos(m * (K[d] + qv[d]) * h) for m in range(1, nf + 1)))
                  / h ** 2 for d in range(3))
        np.testing.assert_allclose(got, lam * u, rtol=0, atol=1e-11 * abs(lam))
        # the opposite phase sign gives lambda(k - q): different here
        lam_m = sum(-0.5 * (a[0] + 2 * sum(a[m] * np.cos(m * (K[d] - qv[d]) * h) for m in range(1, nf + 1)))
                    / h ** 2 for d in range(3))
        assert abs(lam - lam_m) > 1e-3


def test_kslaplacian_c_at_gamma_equals_real_path(orc):
    n, h, nf = 10, 0.45, 4
    rng = np.random.default_rng(3)
    u = rng.standard_normal(n ** 3)
    veff = np.asfortranarray(rng.standard_normal((n, n, n)))
    got = lap_c(orc, u.astype(np.complex128), n, h, (0.0, 0.0, 0.0), nf, veff)
    ref = np.zeros((n ** 3, 1), order="F")
    orc.call("kslaplacian_r", ncpx=n, ncpy=n, ncpz=n, neigmx=1, nf=nf, dx=h, dy=h, dz=h,
             vre=H.halo(u.reshape((n, n, n), order="F"), nf, nf), veff=veff, l=1, avre=ref)
    np.testing.assert_allclose(got.real, ref[:, 0], rtol=0, atol=1e-13 * np.abs(ref).max())
    assert np.abs(got.imag).max() == 0.0


def test_ksprecondition_c_symbol(orc):
    n, h = 10, 0.4
    X, Y, Z = grid(n, h)
    L = n * h
    qv = 2 * np.pi * np.array([1, 2, -1]) / L
    u = np.exp(1j * (qv[0] * X + qv[1] * Y + qv[2] * Z))
    out = np.zeros((n, n, n), dtype=np.complex128, order="F")
    orc.call("ksprecondition_c", ncpx=n, ncpy=n, ncpz=n, nf=4, ncol=1, vcm=halo_c(u, 4, 1), avcm=out)
    sym = 0.5 + sum(np.cos(qv[d] * h) for d in range(3)) / 6.0
    np.testing.assert_allclose(out, sym * u, rtol=0, atol=1e-14)


def phase_matrix(s, k):
    """P (N x 9 natom) complex: P[r, 9a + i] = vnlocp(j, i, a) exp(i k.x_j), x_j = (lst + nat) h - h/2."""
    P = np.zeros((s.npts, 9 * s.natom), dtype=np.complex128)
    for a in range(s.natom):
        m = s.natinf[a]
        x = (s.lstx[:m, a] + s.natx[a]) * s.h - 0.5 * s.h
        y = (s.lsty[:m, a] + s.naty[a]) * s.h - 0.5 * s.h
        z = (s.lstz[:m, a] + s.natz[a]) * s.h - 0.5 * s.h
        ph = np.exp(1j * (k[0] * x + k[1] * y + k[2] * z))
        idx = s.lstvec2[:m, a] - 1
        for i in range(9):
            np.add.at(P[:, 9 * a + i], idx, s.vnlocp[:m, i, a] * ph)
    return P


def dense_hk(s, k):
    """H_k = T_k + diag(V) + conj(P) D P^T / dV from the definitions."""
    n, h = s.n, s.h
    a = H.fd_weights(s.nf)

    def t1(kk):
        m = np.zeros((n, n), dtype=np.complex128)
        for i in range(n):
            m[i, i] += a[0]
            for j in range(1, s.nf + 1):
                m[i, (i + j) % n] += a[j] * np.exp(1j * kk * j * h)
                m[i, (i - j) % n] += a[j] * np.exp(-1j * kk * j * h)
        return -0.5 * m / h ** 2
    I = np.eye(n)
    T = (np.kron(I, np.kron(I, t1(k[0]))) + np.kron(I, np.kron(t1(k[1]), I))
         + np.kron(t1(k[2]), np.kron(I, I)))
    Hk = T + np.diag(s.veff[..., 0].ravel(order="F"))
    P = phase_matrix(s, k)
    Hk += np.conj(P) @ np.kron(np.eye(s.natom), s.dmat) @ P.T / s.dv
    return Hk


def dense_sk(s, k):
    P = phase_matrix(s, k)
    return np.eye(s.npts) + np.conj(P) @ np.kron(np.eye(s.natom), s.sss[:, :, 0]) @ P.T / s.dv


def test_nonlocal_c_against_dense_projectors(orc):
    s = synth.make_system(10, 0.5, 3, 2, seed=5, q_scale=2.0)
    kw = synth.bloch_args(s, K)
    rng = np.random.default_rng(6)
    u = np.asfortranarray(rng.standard_normal(s.npts) + 1j * rng.standard_normal(s.npts))
    cp = np.zeros((9, s.natom), dtype=np.complex128, order="F")
    orc.call("nonlocaloperation_c_01", **dict(kw, neigmx=1, l=1), vcm2=u, cpsep=cp,
             vcccm=np.zeros(s.num_list, dtype=np.complex128))
    P = phase_matrix(s, K)
    np.testing.assert_allclose(cp.ravel(order="F"), P.T @ u, rtol=0, atol=1e-13)
    for nhs, X in ((1, s.dij), (0, s.sss)):
        out = np.zeros(s.npts, dtype=np.complex128)
        orc.call("nonlocaloperation_c_02", **dict(kw, nhs=nhs, neigmx=1, l=1), cpsep=cp, dijsss=X,
                 avcm=out, vcccm=np.zeros(s.num_list, dtype=np.complex128),
                 avc=np.zeros(s.num_list, dtype=np.complex128))
        Xd = s.dmat if nhs == 1 else s.sss[:, :, 0]
        want = np.conj(P) @ (np.kron(np.eye(s.natom), Xd) @ (P.T @ u)) / s.dv
        np.testing.assert_allclose(out, want, rtol=0, atol=1e-11 * np.abs(want).max())


def _apply_dense(orc, s, k, which):
    """Columns of H_k (or S_k) from the oracle's applications to unit vectors."""
    kw = synth.bloch_args(s, k)
    cols = []
    for j in range(s.npts):
        e = np.zeros(s.npts, dtype=np.complex128)
        e[j] = 1.0
        cp = np.zeros((9, s.natom), dtype=np.complex128, order="F")
        orc.call("nonlocaloperation_c_01", **dict(kw, neigmx=1, l=1), vcm2=e, cpsep=cp,
                 vcccm=np.zeros(s.num_list, dtype=np.complex128))
        out = np.zeros(s.npts, dtype=np.complex128) if which == "h" else e.copy()
        X = s.dij if which == "h" else s.sss
        orc.call("nonlocaloperation_c_02", **dict(kw, nhs=1 if which == "h" else 0, neigmx=1, l=1),
                 cpsep=cp, dijsss=X, avcm=out, vcccm=np.zeros(s.num_list, dtype=np.complex128),
                 avc=np.zeros(s.num_list, dtype=np.complex128))
        if which == "h":
            out = out + lap_c(orc, e, s.n, s.h, k, s.nf, s.veff[..., 0])
        cols.append(out)
    return np.stack(cols, axis=1)


def test_dense_hk_from_oracle_is_hermitian_and_matches_definition(orc):
    s = synth.make_system(6, 0.55, 2, 2, seed=9, q_scale=1.0)
    Hk = _apply_dense(orc, s, K, "h")
    np.testing.assert_allclose(Hk, np.conj(Hk.T), rtol=0, atol=1e-10 * np.abs(Hk).max())
    np.testing.assert_allclose(Hk, dense_hk(s, K), rtol=0, atol=1e-10 * np.abs(Hk).max())
    Sk = _apply_dense(orc, s, K, "s")
    np.testing.assert_allclose(Sk, dense_sk(s, K), rtol=0, atol=1e-12)
    np.testing.assert_allclose(Sk, np.conj(Sk.T), rtol=0, atol=1e-12)


def _init_c(orc, s, k, m=None, seed=None):
    m = s.nbands if m is None else m
    psi = synth.random_bands_c(s, seed=seed, nbands=m)
    st = dict(sveccm=psi, ssvcm=np.zeros_like(psi), cspsep=np.zeros((9, s.natom, m), dtype=np.complex128, order="F"))
    orc.call("orthogonalization_c", **synth.bloch_args(s, k, neigmx=m), nflag=1, **st)
    st.update(hsvcm=np.zeros_like(psi), sval=np.zeros(m), residual_states=np.zeros(m))
    return st


@pytest.mark.parametrize("q", [0.0, 2.0])
def test_orthogonalization_c_is_gram_schmidt_in_s(orc, q):
    s = synth.make_system(8, 0.55, 2, 40, seed=6, q_scale=q)  # 40 bands: the mstep = 32 block path
    psi0 = synth.random_bands_c(s, smooth=False)
    st = dict(sveccm=psi0.copy(order="F"), ssvcm=np.zeros_like(psi0),
              cspsep=np.zeros((9, s.natom, 40), dtype=np.complex128, order="F"))
    orc.call("orthogonalization_c", **synth.bloch_args(s, K), nflag=1, **st)
    Sd = dense_sk(s, K)
    Q = st["sveccm"]
    np.testing.assert_allclose(np.conj(Q.T) @ Sd @ Q * s.dv, np.eye(40), rtol=0, atol=1e-12)
    np.testing.assert_allclose(st["ssvcm"], Sd @ Q, rtol=0, atol=1e-11)
    np.testing.assert_allclose(st["cspsep"].reshape(9 * s.natom, 40, order="F"), phase_matrix(s, K).T @ Q,
                               rtol=0, atol=1e-11)
    R = np.linalg.cholesky(np.conj(psi0.T) @ Sd @ psi0 * s.dv).conj().T  # G = R^H R
    np.testing.assert_allclose(Q, psi0 @ np.linalg.inv(R), rtol=0, atol=1e-9)


def test_zheev_matches_numpy(orc):
    rng = np.random.default_rng(2)
    A = rng.standard_normal((30, 30)) + 1j * rng.standard_normal((30, 30))
    A = A + A.conj().T
    w, Z = orc.zheev(A)
    np.testing.assert_allclose(w, np.linalg.eigvalsh(A), rtol=0, atol=1e-12)
    np.testing.assert_allclose(A @ Z, Z * w, rtol=0, atol=1e-11)
    np.testing.assert_allclose(np.conj(Z.T) @ Z, np.eye(30), rtol=0, atol=1e-13)
    big = np.argmax(np.abs(Z), axis=0)
    assert np.all(np.abs(Z[big, range(30)].imag) < 1e-15) and np.all(Z[big, range(30)].real > 0)


@pytest.mark.parametrize("q", [0.0, 2.0])
def test_rayleighritz_c_matches_dense_projection(orc, q):
    s = synth.make_system(8, 0.55, 2, 6, seed=7, q_scale=q)
    st = _init_c(orc, s, K)
    psi0 = st["sveccm"].copy()
    orc.call("rayleighritz_c", **synth.bloch_args(s, K), l_rrb=1,
             **{k: st[k] for k in ("cspsep", "sveccm", "ssvcm", "hsvcm", "sval")})
    Hd, Sd = dense_hk(s, K), dense_sk(s, K)
    w = scipy.linalg.eigh(np.conj(psi0.T) @ Hd @ psi0 * s.dv, np.conj(psi0.T) @ Sd @ psi0 * s.dv,
                          eigvals_only=True)
    np.testing.assert_allclose(st["sval"], w, rtol=0, atol=1e-11)
    Q = st["sveccm"]
    np.testing.assert_allclose(np.conj(Q.T) @ Sd @ Q * s.dv, np.eye(6), rtol=0, atol=1e-11)
    np.testing.assert_allclose(np.conj(Q.T) @ Hd @ Q * s.dv, np.diag(w), rtol=0, atol=1e-10)


@pytest.mark.parametrize("q", [0.0, 2.0])
def test_scf_diag_bloch_converges_to_dense_eigenvalues(orc, q):
    """kscg_c (band-sequential CG with the 2x2 Rayleigh-Ritz, through scf_diag nrc = 1 with a
    Rayleigh-Ritz step first) converges to the lowest eigenpairs of eigh(H_k, S_k).  (On an 8^3
    grid at h = 0.55 the residual left after a few steps sits at the (pi/h, pi/h, pi/h) mode,
    where the preconditioner's symbol 1/2 + (1/6) sum cos(kh) vanishes, so the reference's
    iteration stalls near 1e-8 Ha there; this grid converges.)"""
    s = synth.make_system(10, 0.5, 3, 6, seed=8, q_scale=q)
    st = _init_c(orc, s, K)
    kw = synth.bloch_args(s, K, nsdmax=4, epssd=1e-12)
    real_dummies = {k: np.zeros(1) for k in ("rspsep", "svecre", "hsvre", "ssvre")}
    for it in range(1, 61):
        kw["iscf"] = it
        orc.call("scf_diag", **kw, **st, **real_dummies)
    Hd, Sd = dense_hk(s, K), dense_sk(s, K)
    w = scipy.linalg.eigh(Hd, Sd, eigvals_only=True)[:6]
    np.testing.assert_allclose(st["sval"], w, rtol=0, atol=1e-8)
    Q = st["sveccm"]
    np.testing.assert_allclose(np.conj(Q.T) @ Sd @ Q * s.dv, np.eye(6), rtol=0, atol=1e-12)


def test_charge_c(orc):
    s = synth.make_system(6, 0.5, 1, 3, seed=1)
    psi = synth.random_bands_c(s, smooth=False)
    f = np.array([2.0, 1.0, 0.5])
    rho = np.zeros(s.npts)
    orc.call("scf_charge_smt_c", neigmx=3, nums=1, ncol=1, nspv=1, numk=1, ncpx=s.n, ncpy=s.n,
             ncpz=s.n, fnele=f, sveccm=psi, rhosmt=rho)
    np.testing.assert_allclose(rho, (np.abs(psi) ** 2) @ f, rtol=1e-14)


@pytest.mark.parametrize("q", [0.0, 2.0])
def test_scf_diag_bloch_with_diis_converges_to_dense_eigenvalues(orc, q):
    """CG for the first ten steps, RMM-DIIS (ksdiis_c, five sweeps per step) afterwards,
    Rayleigh-Ritz every step: the lowest eigenpairs of eigh(H_k, S_k).  (The DIIS of the
    reference, two vectors per band, converges linearly here: about a factor 3 per ten steps
    on the eigenvalue error, so the bar is 1e-7 after 80 steps.)"""
    s = synth.make_system(10, 0.5, 3, 6, seed=18, q_scale=q)
    st = _init_c(orc, s, K)
    kw = synth.bloch_args(s, K, nsdmax=4, epssd=1e-12, ncgscf=10, ndiismax=6)
    real_dummies = {k: np.zeros(1) for k in ("rspsep", "svecre", "hsvre", "ssvre")}
    for it in range(1, 81):
        kw["iscf"] = it
        orc.call("scf_diag", **kw, **st, **real_dummies)
    Hd, Sd = dense_hk(s, K), dense_sk(s, K)
    w = scipy.linalg.eigh(Hd, Sd, eigvals_only=True)[:6]
    np.testing.assert_allclose(st["sval"], w, rtol=0, atol=1e-7)
    Q = st["sveccm"]
    np.testing.assert_allclose(np.conj(Q.T) @ Sd @ Q * s.dv, np.eye(6), rtol=0, atol=1e-12)
