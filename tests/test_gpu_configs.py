"""The benchmark configurations at their full sizes (synth.CONFIGS; bands drawn and
orthonormalised on the device, rspacedist_amd.workload).

- 40^3 grid, 32 bands, 8 atoms (h = 0.25): H psi of every band, one band-sequential kscg_r
  sweep and the Hartree CG solve, each against the oracle on the same inputs.
- 96^3 grid, 512 bands, 64 atoms (h = 0.2), FP64: ten scf_diag calls, each one CG iteration
  over all bands with Gram-Schmidt and a 512 x 512 FP64 Rayleigh-Ritz step, then the Hartree
  solve to 1e-10: orthonormality, non-increasing band-energy sum, falling residuals, the
  Hartree residual recomputed on the host and the oracle's Hartree solution.
- 128^3 grid, 1024 bands, 128 atoms (h = 0.2): the split-FP16 TN and NN subspace products at
  2,097,152 x 1024 x 1024 against rocBLAS FP64, and one split-precision SCF step against one
  FP64 step from the same state.

Bounds as in the smaller tests: FP64 operator applications 1e-12 relative; iterative
routines as tests/test_gpu_parity.py; split products |C - C64| <= 2e-6 (|A|^T |B|)_ij, here
with the Cauchy-Schwarz bound ||a_i|| ||b_j|| >= (|A|^T |B|)_ij for TN (1.6x looser for
random operands) and the exact |X| |C| on a block of 65536 rows for NN.
"""
import numpy as np
import pytest

from rspacedist_amd import synth
from tests import helpers as H

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def kk():
    from rspacedist_amd import kukan
    kukan.init(0)
    kukan.set_param("kscg_mode", 0)
    kukan.set_param("precision", 0)
    yield kukan
    kukan.set_param("kscg_mode", 0)
    kukan.set_param("precision", 0)


# ------------------------------------------------------------------ 40^3 x 32 bands x 8 atoms

@pytest.fixture(scope="module")
def c0(orc):
    s = synth.make_config(0)
    st = H.init_state(s, orc, smooth=True)
    return s, st


def _hpsi(call, s, psi):
    """H psi of every band through the leaf entry points: halo fill, kslaplacian_r, then
    <p|psi> (nonlocaloperation_r_01) and its expansion (nonlocaloperation_r_02, nhs = 1)."""
    kw = synth.abi_args(s)
    n, nf, m = s.n, s.nf, psi.shape[1]
    out = np.zeros((s.npts, m), order="F")
    for l in range(m):
        v = np.zeros((n + 2 * nf,) * 3, order="F")
        v[nf:nf + n, nf:nf + n, nf:nf + n] = psi[:, l].reshape((n, n, n), order="F")
        call("overlap_finitedifference_r", nperi=3, ncpa=n, ncpb=n, ncpc=n, nf=nf, nf1=nf - 1,
             nsnd=nf, v_cc=v)
        a = np.zeros((s.npts, 1), order="F")
        call("kslaplacian_r", ncpx=n, ncpy=n, ncpz=n, neigmx=1, nf=nf, dx=s.h, dy=s.h, dz=s.h,
             vre=v, veff=s.veff, l=1, avre=a)
        rp = np.zeros((9, s.natom), order="F")
        call("nonlocaloperation_r_01", **kw, vre2=np.asfortranarray(psi[:, l]), rpsep=rp)
        call("nonlocaloperation_r_02", **dict(kw, nhs=1), rpsep=rp, dijsss=s.dij, avre=a,
             avr=np.zeros(s.num_list))
        out[:, l] = a[:, 0]
    return out


def test_c0_hpsi_all_bands(kk, orc, c0):
    s, st = c0
    g = _hpsi(kk.call, s, st["svecre"])
    o = _hpsi(orc.call, s, st["svecre"])
    np.testing.assert_allclose(g, o, rtol=0, atol=1e-12 * np.abs(o).max())
    # and the fused all-band path: dV <psi|H psi> per band
    eg, eo = np.zeros(s.nbands), np.zeros(s.nbands)
    kw = synth.abi_args(s)
    kk.call("band_energy_r", **kw, svecre=st["svecre"], eband=eg)
    orc.call("band_energy_r", **kw, svecre=st["svecre"], eband=eo)
    np.testing.assert_allclose(eg, eo, rtol=1e-12, atol=1e-12)
    np.testing.assert_allclose(eg, np.einsum("ij,ij->j", st["svecre"], o) * s.dv, rtol=1e-11)


def test_c0_kscg_r_sweep(kk, orc, c0):
    s, st0 = c0
    kk.set_param("kscg_mode", 0)
    arr = {k: st0[k].copy(order="F") for k in ("sval", "rspsep", "svecre", "ssvre", "residual_states", "hsvre")}
    res = []
    for f in (kk.call, orc.call):
        a = {k: v.copy(order="F") for k, v in arr.items()}
        r = f("kscg_r", **synth.abi_args(s), **a)
        res.append((a, r))
    (g, rg), (o, ro) = res
    np.testing.assert_allclose(g["sval"], o["sval"], rtol=1e-10, atol=1e-10)
    for k in ("svecre", "ssvre", "hsvre", "rspsep"):
        np.testing.assert_allclose(g[k], o[k], atol=1e-8 * np.abs(o[k]).max(), err_msg=k)
    np.testing.assert_allclose(g["residual_states"], o["residual_states"], rtol=1e-6, atol=1e-14)
    assert rg["ksitmax"] == ro["ksitmax"]


def test_c0_hartree(kk, orc, c0):
    s, st = c0
    rho = np.asfortranarray((st["svecre"] ** 2).sum(axis=1).reshape((s.n,) * 3, order="F"))
    z = np.zeros(1)
    kw = dict(nd=1, ncpx_d=s.n, ncpy_d=s.n, ncpz_d=s.n, nfh=4, ndisp=6, nperi=3, ncgres=801,
              ncgmin=1, ncgmax=800, xmax=s.xmax, ymax=s.xmax, zmax=s.xmax, epsvh=1e-10,
              rho_aug_dense=rho, boundx=z, boundy=z, boundz=z)
    vg, vo = np.zeros_like(rho), np.zeros_like(rho)
    kk.call("scf_hartree", **kw, vh_dense=vg)
    orc.call("scf_hartree", **kw, vh_dense=vo)
    np.testing.assert_allclose(vg, vo, rtol=0, atol=1e-8 * np.abs(vo).max())
    assert hartree_residual(rho, vg, s.h) <= 1.01e-10  # dp < 1e-10, summed in another order


def hartree_residual(rho, v, h, nfh=4):
    """sqrt(sum r^2 dV / Omega) of r = 2 pi (rho - mean rho) - (-1/2 lap_nfh) v on the periodic
    grid (scf_hartree8f.F90's stopping measure), with the FD weights of tools_definefdcoef."""
    a = H.fd_weights(nfh)
    lap = 3 * a[0] * v
    for ax in range(3):
        for m in range(1, nfh + 1):
            lap = lap + a[m] * (np.roll(v, m, axis=ax) + np.roll(v, -m, axis=ax))
    b = 2 * np.pi * (rho - rho.mean())
    r = b + 0.5 * lap / h ** 2
    return float(np.sqrt(np.mean(r * r)))


# ------------------------------------------------------------------ 96^3 x 512 bands x 64 atoms

def test_c1_fp64_ten_iterations(kk, orc):
    from rspacedist_amd import workload
    kk.set_param("kscg_mode", 1)
    kk.set_param("precision", 0)
    c = workload.Cell(cfg=1)
    try:
        c.init_bands(seed=4321)
        sums, resid = [], []
        for it in range(10):  # RR (512 x 512 FP64 eigh), one CG iteration, Gram-Schmidt
            c.diag(nsdmax=1)
            sv = c.st["sval"].numpy()
            sums.append(sv.sum())
            resid.append(c.st["residual_states"].numpy().sum())
            print(f"iteration {it + 1}: sum eps {sums[-1]:.10f} Ha, sum residual {resid[-1]:.3e}")
        assert all(b <= a + 1e-9 * abs(a) for a, b in zip(sums, sums[1:])), sums
        assert resid[-1] < 0.5 * resid[0]
        # orthonormality of the final bands: dV Psi^T Psi on the device in FP64
        m, npts = c.s.nbands, c.s.npts
        G = np.zeros((m, m), order="F")
        kk.subspace_gemm(1, m, m, npts, c.s.dv, c.st["svecre"], npts, c.st["svecre"], npts, 0.0, G, m)
        np.testing.assert_allclose(G, np.eye(m), rtol=0, atol=1e-12)
        # Hartree to 1e-10 for rho of the final bands
        its = c.charge_hartree(1e-10)
        rho = c.st["rho"].numpy().reshape((c.s.n,) * 3, order="F")
        vh = c.st["vh"].numpy().reshape((c.s.n,) * 3, order="F")
        res = hartree_residual(rho, vh, c.s.h)
        print(f"Hartree: {its} CG iterations, residual {res:.2e}")
        # the solver stops at its own dp < 1e-10; the host recomputation sums in another order
        assert res <= 1.01e-10
        vo = np.zeros_like(rho)
        z = np.zeros(1)
        orc.call("scf_hartree", nd=1, ncpx_d=c.s.n, ncpy_d=c.s.n, ncpz_d=c.s.n, nfh=4, ndisp=6,
                 nperi=3, ncgres=801, ncgmin=1, ncgmax=800, xmax=c.s.xmax, ymax=c.s.xmax,
                 zmax=c.s.xmax, epsvh=1e-10, rho_aug_dense=np.asfortranarray(rho), boundx=z,
                 boundy=z, boundz=z, vh_dense=vo)
        np.testing.assert_allclose(vh, vo, rtol=0, atol=1e-8 * np.abs(vo).max())
    finally:
        c.free()
        kk.set_param("kscg_mode", 0)


# ------------------------------------------------------------------ 128^3 x 1024 bands x 128 atoms

TOL = 2e-6


def test_c2_split_products_against_fp64(kk):
    R, M = 128 ** 3, 1024
    D = kk.DeviceArray
    A = D((R, M)).fill_normal(11)
    B = D((R, M)).fill_normal(12)
    try:
        def tn(X, Y, precision):
            kk.set_param("precision", precision)
            kk.set_param("mixed_min_rows", 65536)
            G = np.zeros((M, M), order="F")
            kk.subspace_gemm(1, M, M, R, 1.0, X, R, Y, R, 0.0, G, M)
            kk.set_param("precision", 0)
            return G
        g64 = tn(A, B, 0)
        gmx = tn(A, B, 1)
        na = np.sqrt(np.diag(tn(A, A, 0)))
        nb = np.sqrt(np.diag(tn(B, B, 0)))
        err = np.abs(gmx - g64) / (TOL * na[:, None] * nb[None, :])
        print(f"TN {R}x{M}x{M}: max err / bound {err.max():.3e}")
        assert err.max() <= 1.0
        # the same operand on both sides (one split serves both)
        gs = tn(A, A, 1)
        assert (np.abs(gs - tn(A, A, 0)) / (TOL * na[:, None] * na[None, :])).max() <= 1.0
        # NN: Y = A C with C of mixed column scales, beta = 0 and beta = 1
        rng = np.random.default_rng(13)
        Cm = np.asfortranarray(rng.standard_normal((M, M)) * np.exp2(rng.integers(-8, 8, (M, 1))))
        r0, r1 = 1_000_000, 1_065_536
        arows = workload_rows(A, R, r0, r1)
        ref_abs = np.abs(arows) @ np.abs(Cm)
        for beta in (0.0, 1.0):
            Y = {}
            for prec in (0, 1):
                B_copy = D((R, M))
                kk._ck(kk.lib().kukan_memcpy(B_copy.device_ptr, B.device_ptr, B.nbytes, 3), "memcpy")
                kk.set_param("precision", prec)
                kk.subspace_gemm(0, R, M, M, 1.0, A, R, Cm, M, beta, B_copy, R)
                kk.set_param("precision", 0)
                Y[prec] = workload_rows(B_copy, R, r0, r1)
                B_copy.free()
            bound = TOL * ref_abs + 1e-14 * np.abs(Y[0])
            e = np.abs(Y[1] - Y[0]) / bound
            print(f"NN {R}x{M}x{M} beta={beta}: max err / bound {e.max():.3e} (rows {r0}..{r1})")
            assert e.max() <= 1.0
    finally:
        A.free()
        B.free()
        kk.set_param("precision", 0)


def workload_rows(d, rows, r0, r1):
    from rspacedist_amd import workload
    return workload.column_rows(d, rows, r0, r1)


def test_c2_mixed_step_against_fp64_step(kk):
    """From one mid-SCF state (six FP64 steps after the seeded start, as the benchmark), one
    split-precision SCF step and one FP64 step: eigenvalues within 1e-6 Ha; both return bands
    orthonormal to FP64 rounding (the closing Cholesky-QR is FP64 in both)."""
    from rspacedist_amd import workload
    kk.set_param("kscg_mode", 1)
    c = workload.Cell(cfg=2)
    try:
        c.init_bands(seed=1234)
        for _ in range(6):
            c.scf_step()
        snap = c.snapshot()
        out = {}
        for prec in (0, 1):
            c.restore(snap)
            kk.set_param("precision", prec)
            c.scf_step()
            kk.set_param("precision", 0)
            out[prec] = c.st["sval"].numpy()
            # diagonal of dV Psi^T Psi in FP64
            psi = c.st["svecre"]
            m, npts = c.s.nbands, c.s.npts
            G = np.zeros((m, m), order="F")
            kk.subspace_gemm(1, m, m, npts, c.s.dv, psi, npts, psi, npts, 0.0, G, m)
            np.testing.assert_allclose(np.diag(G), 1.0, rtol=0, atol=1e-12)
            off = np.abs(G - np.diag(np.diag(G))).max()
            print(f"precision {prec}: max off-diagonal overlap {off:.2e}")
            assert off <= 1e-12
        d = np.abs(out[1] - out[0]).max()
        print(f"128^3 x 1024: max |eig mixed - eig fp64| after one step = {d:.2e} Ha")
        assert d <= 1e-6
        for v in snap.values():
            if hasattr(v, "free"):
                v.free()
    finally:
        c.free()
        kk.set_param("kscg_mode", 0)
        kk.set_param("precision", 0)
