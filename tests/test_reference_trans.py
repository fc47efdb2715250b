"""The oracle's potential-assembly transfers against the reference's own routines (CPU):
interpolation8f.F90, overlap_interpolation8f.F90, trans8f.F90 and scf_rhoaugdense8f.F90
compiled unchanged into oracle/_ref/libkukan_ref.so (one MPI rank).  Same summation order on
both sides, so the results agree to the last bits (1e-15 relative allowed for the compilers'
freedom in evaluating the multipole expression).  Skipped where the reference library was
not built."""
import numpy as np
import pytest

from tests.test_oracle_trans import halo_shape_c, halo_shape_d

ref = pytest.importorskip("oracle.ref")
pytestmark = pytest.mark.skipif(not ref.available(), reason="reference library not built")

CASES = [(nperi, nf, nmesh) for nperi in (0, 1, 2, 3) for nf, nmesh in ((2, 2), (4, 2), (3, 3))]


def both(orc, name, arrays, **kw):
    out = []
    for f in (orc.call, ref.call):
        a = {k: (v.copy(order="F") if isinstance(v, np.ndarray) else v) for k, v in arrays.items()}
        f(name, **kw, **a)
        out.append(a)
    return out


def close(o, r, k, rel=1e-15):
    scale = max(np.abs(r[k]).max(), 1e-300)
    err = np.abs(o[k] - r[k]).max() / scale
    assert err <= rel, f"{k}: relative difference {err:.2e}"


@pytest.mark.parametrize("nperi,nf,nmesh", CASES)
def test_interpolation_equals_reference(orc, nperi, nf, nmesh):
    n = (9, 8, 10)
    rng = np.random.default_rng(nperi * 10 + nf)
    vc = np.asfortranarray(rng.standard_normal(halo_shape_c(n, nf)))
    vd = np.asfortranarray(rng.standard_normal(halo_shape_d(n, nf, nmesh)))
    kw = {"ncpx": n[0], "ncpy": n[1], "ncpz": n[2], "nmesh": nmesh, "nperi": nperi, "nf": nf, "ndisp": 6}
    for is_ in (1, 2):
        o, r = both(orc, "interpolation", {"v_dd": vd, "v_cc": vc}, **kw, **{"is": is_})
        close(o, r, "v_dd")
        close(o, r, "v_cc")


@pytest.mark.parametrize("nperi", [0, 1, 2, 3])
@pytest.mark.parametrize("nflag", [1, 2])
def test_overlap_interpolation_equals_reference(orc, nperi, nflag):
    n, nf = (7, 9, 8), 3
    rng = np.random.default_rng(nperi + 5 * nflag)
    vc = np.asfortranarray(rng.standard_normal(halo_shape_c(n, nf)))
    o, r = both(orc, "overlap_interpolation", {"v_cc": vc}, ndisp=6, nperi=nperi, ncpa=n[0], ncpb=n[1],
                ncpc=n[2], nf=nf, nf1=nf - 1, nflag=nflag)
    np.testing.assert_array_equal(o["v_cc"], r["v_cc"])


@pytest.mark.parametrize("nperi,nf,nmesh", CASES)
def test_trans_routines_equal_reference(orc, nperi, nf, nmesh):
    n, h = (8, 8, 8), 0.45
    rng = np.random.default_rng(3 * nperi + nf + nmesh)
    nd = tuple(m * nmesh for m in n)
    g = dict(nmesh=nmesh, nperi=nperi, ndisp=6, nf=nf, ncpx=n[0], ncpy=n[1], ncpz=n[2])
    rho = np.asfortranarray(rng.random(n + (2,)))
    o, r = both(orc, "trans_c2d_smtcharge", {"rhosmt": rho, "rho_dense": np.zeros(nd + (2,), order="F")},
                nspv=2, **g)
    close(o, r, "rho_dense")
    dense = np.asfortranarray(rng.standard_normal(nd))
    o, r = both(orc, "trans_d2c_smtcharge", {"rho_aug_dense": dense, "rho_coarse": np.zeros(n, order="F")}, **g)
    close(o, r, "rho_coarse")
    vxd = np.asfortranarray(rng.standard_normal(nd + (2,)))
    o, r = both(orc, "trans_d2c_vxc", {"vx": np.zeros(n + (2,), order="F"), "vx_dense": vxd}, nspv=2, **g)
    close(o, r, "vx")
    natom = 3
    xm = 0.5 * n[0] * h
    cp = np.zeros((8, 2), order="F")
    cp[0] = [1.3, 2.1]
    atoms = rng.uniform(-xm * 0.6, xm * 0.6, (3, natom))
    o, r = both(orc, "trans_d2c_vh", {"vh_coarse": np.zeros(n, order="F"), "vh_dense": dense},
                natom=natom, num_spe=2, xmax=xm, ymax=xm, zmax=xm,
                indspe=np.array([1, 2, 1], dtype=np.int32),
                atmpole=np.asfortranarray(rng.standard_normal((10, natom))), cp=cp,
                atx=atoms[0].copy(), aty=atoms[1].copy(), atz=atoms[2].copy(), **g)
    close(o, r, "vh_coarse", 1e-14 if nperi < 3 else 1e-15)


def test_scf_rhoaugdense_equals_reference(orc):
    nd = (12, 10, 8)
    rng = np.random.default_rng(9)
    natom, nl = 5, 80
    lst = np.asfortranarray(rng.integers(1, np.prod(nd) + 1, size=(nl, natom)).astype(np.int32))
    arr = {"rho_dense": np.asfortranarray(rng.standard_normal(nd + (2,))),
           "rhoaug3d": np.asfortranarray(rng.standard_normal((nl, natom))),
           "rho_aug_dense": np.zeros(nd, order="F")}
    o, r = both(orc, "scf_rhoaugdense", arr, key_natpri_inps=1, key_pp_paw=1, natom=natom, num_spe=2,
                nspv=2, ncpx_d=nd[0], ncpy_d=nd[1], ncpz_d=nd[2], num_list_d=nl, num_ppcell_d=natom,
                indspe=np.array([1, 2, 1, 1, 2], dtype=np.int32),
                natprid=np.array([1, 1, 0, 1, 1], dtype=np.int32),
                napsd=np.arange(1, natom + 1, dtype=np.int32),
                natinfd=np.array([80, 40, 60, 75, 20], dtype=np.int32),
                ntyppp=np.array([1, 1], dtype=np.int32), lstvecd2=lst)
    np.testing.assert_array_equal(o["rho_aug_dense"], r["rho_aug_dense"])
