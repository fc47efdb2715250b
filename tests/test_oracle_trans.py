"""Pins of the oracle's potential-assembly transfers (oracle/kukan_oracle_trans.c), CPU.

- interpolation (is = 1): Lagrange interpolation of order 2 nf reproduces a polynomial of degree
  2 nf - 1 in each coordinate exactly (even and odd nmesh, where the dense points sit at
  different offsets); a wrong weight, offset or tap range fails it.
- interpolation (is = 2) with overlap_interpolation (nflag = 2) is the adjoint of the
  prolongation divided by nmesh^3: <R d, c> = <d, P c> / nmesh^3 for every nperi (periodic
  axes wrap, isolated axes pad with zeros); together with the polynomial pin this fixes the
  restriction's index arithmetic.
- overlap_interpolation (nflag = 1) against numpy's periodic extension, (nflag = 2) against the
  fold of each periodic halo onto the layers it wraps to.
- trans_d2c_vh's halo: the potential of point multipoles (charge, dipole, quadrupole) from
  their textbook closed forms, added once per isolated face a halo point lies in.
- scf_rhoaugdense against numpy's unbuffered scatter-add; the V_eff sum against its formula.
"""
import numpy as np
import pytest

NPERIS = [0, 1, 2, 3]


def coarse_coords(n, h, nf):
    j = np.arange(-(nf - 1), n + nf + 1)
    return j, j * h - 0.5 * n * h - 0.5 * h


def dense_coords(n, h, nf, nmesh):
    hd = h / nmesh
    nd = n * nmesh
    j = np.arange(-(nf * nmesh - 1), nd + nf * nmesh + 1)
    return j, j * hd - 0.5 * n * h - 0.5 * hd


def halo_shape_c(n, nf):
    return tuple(m + 2 * nf for m in n)


def halo_shape_d(n, nf, nmesh):
    return tuple(m * nmesh + 2 * nf * nmesh for m in n)


def interior_d(a, n, nf, nmesh):
    o = nf * nmesh
    return a[o:o + n[0] * nmesh, o:o + n[1] * nmesh, o:o + n[2] * nmesh]


def interior_c(a, n, nf):
    return a[nf:nf + n[0], nf:nf + n[1], nf:nf + n[2]]


@pytest.mark.parametrize("nf,nmesh", [(1, 2), (2, 2), (4, 2), (3, 3), (4, 3)])
def test_prolongation_reproduces_polynomials(orc, nf, nmesh):
    n, h = (7, 6, 5), 0.4
    rng = np.random.default_rng(nf * 10 + nmesh)
    deg = 2 * nf - 1
    cs = [rng.standard_normal(deg + 1) for _ in range(3)]
    f = lambda x, y, z: np.polyval(cs[0], x) * np.polyval(cs[1], y) + np.polyval(cs[2], z)
    xs = [coarse_coords(m, h, nf)[1] for m in n]
    X, Y, Z = np.meshgrid(*xs, indexing="ij")
    vc = np.asfortranarray(f(X, Y, Z))
    vd = np.zeros(halo_shape_d(n, nf, nmesh), order="F")
    orc.call("interpolation", **{"ncpx": n[0], "ncpy": n[1], "ncpz": n[2], "nmesh": nmesh, "v_dd": vd,
                                 "v_cc": vc, "nperi": 0, "nf": nf, "is": 1, "ndisp": 6})
    # the dense coordinates use h / nmesh about the same cell
    xd = [dense_coords(m, h, nf, nmesh)[1][nf * nmesh:nf * nmesh + m * nmesh] for m in n]
    Xd, Yd, Zd = np.meshgrid(*xd, indexing="ij")
    want = f(Xd, Yd, Zd)
    got = interior_d(vd, n, nf, nmesh)
    np.testing.assert_allclose(got, want, rtol=0, atol=1e-11 * np.abs(want).max())


def _P(orc, w, n, nf, nmesh, nperi):
    vc = np.zeros(halo_shape_c(n, nf), order="F")
    interior_c(vc, n, nf)[...] = w
    orc.call("overlap_interpolation", ndisp=6, nperi=nperi, ncpa=n[0], ncpb=n[1], ncpc=n[2], nf=nf,
             nf1=nf - 1, v_cc=vc, nflag=1)
    vd = np.zeros(halo_shape_d(n, nf, nmesh), order="F")
    orc.call("interpolation", **{"ncpx": n[0], "ncpy": n[1], "ncpz": n[2], "nmesh": nmesh, "v_dd": vd,
                                 "v_cc": vc, "nperi": nperi, "nf": nf, "is": 1, "ndisp": 6})
    return interior_d(vd, n, nf, nmesh).copy()


def _R(orc, d, n, nf, nmesh, nperi):
    vd = np.zeros(halo_shape_d(n, nf, nmesh), order="F")
    interior_d(vd, n, nf, nmesh)[...] = d
    vc = np.zeros(halo_shape_c(n, nf), order="F")
    orc.call("interpolation", **{"ncpx": n[0], "ncpy": n[1], "ncpz": n[2], "nmesh": nmesh, "v_dd": vd,
                                 "v_cc": vc, "nperi": nperi, "nf": nf, "is": 2, "ndisp": 6})
    orc.call("overlap_interpolation", ndisp=6, nperi=nperi, ncpa=n[0], ncpb=n[1], ncpc=n[2], nf=nf,
             nf1=nf - 1, v_cc=vc, nflag=2)
    return interior_c(vc, n, nf).copy()


@pytest.mark.parametrize("nperi", NPERIS)
@pytest.mark.parametrize("nf,nmesh", [(2, 2), (4, 2), (3, 3)])
def test_restriction_is_adjoint_of_prolongation(orc, nperi, nf, nmesh):
    n = (9, 8, 10)
    rng = np.random.default_rng(nperi + 7 * nf)
    w = rng.standard_normal(n)
    d = rng.standard_normal(tuple(m * nmesh for m in n))
    lhs = np.sum(_R(orc, d, n, nf, nmesh, nperi) * w)
    rhs = np.sum(d * _P(orc, w, n, nf, nmesh, nperi)) / nmesh ** 3
    assert abs(lhs - rhs) <= 1e-12 * np.sqrt(np.sum(d * d) * np.sum(w * w))


@pytest.mark.parametrize("nperi", NPERIS)
def test_overlap_interpolation_fill_and_fold(orc, nperi):
    n, nf = (7, 6, 8), 3
    rng = np.random.default_rng(nperi)
    w = rng.standard_normal(n)
    vc = np.zeros(halo_shape_c(n, nf), order="F")
    interior_c(vc, n, nf)[...] = w
    orc.call("overlap_interpolation", ndisp=6, nperi=nperi, ncpa=n[0], ncpb=n[1], ncpc=n[2], nf=nf,
             nf1=nf - 1, v_cc=vc, nflag=1)
    # periodic extension along the periodic axes, zeros along the isolated ones; the halo runs
    # from -(nf - 1) to n + nf
    want = np.zeros_like(vc)
    idx = []
    for a in range(3):
        j = np.arange(-(nf - 1), n[a] + nf + 1)
        if a < nperi:
            idx.append((j - 1) % n[a])
        else:
            idx.append(np.where((j >= 1) & (j <= n[a]), j - 1, -1))
    I, J, K = np.meshgrid(*idx, indexing="ij")
    ok = (I >= 0) & (J >= 0) & (K >= 0)
    want[ok] = w[I[ok], J[ok], K[ok]]
    np.testing.assert_array_equal(vc, want)
    # fold: each halo value lands on the interior point it wraps to (periodic axes only)
    h = rng.standard_normal(vc.shape)
    vc2 = np.asfortranarray(h.copy())
    orc.call("overlap_interpolation", ndisp=6, nperi=nperi, ncpa=n[0], ncpb=n[1], ncpc=n[2], nf=nf,
             nf1=nf - 1, v_cc=vc2, nflag=2)
    ok2 = np.ones(h.shape, bool)
    for a in range(3):
        if a >= nperi:
            j = np.arange(-(nf - 1), n[a] + nf + 1)
            sl = [None, None, None]
            sl[a] = slice(None)
            ok2 &= ((j >= 1) & (j <= n[a]))[tuple(sl)]
    idxw = []
    for a in range(3):
        j = np.arange(-(nf - 1), n[a] + nf + 1)
        idxw.append((j - 1) % n[a])
    Iw, Jw, Kw = np.meshgrid(*idxw, indexing="ij")
    acc = np.zeros(n)
    np.add.at(acc, (Iw[ok2], Jw[ok2], Kw[ok2]), h[ok2])
    np.testing.assert_allclose(interior_c(vc2, n, nf), acc, rtol=1e-14, atol=1e-14)


def _multipole_halo(n, h, nf, nmesh, nperi, atoms, cp1, mp):
    """The dense halo of trans_d2c_vh from the closed forms: the potential at r of a charge q,
    dipole p and traceless-form quadrupole Q placed at R (in the reference's convention, with
    s = R - r): q / |s| + p.s / |s|^3 + sum_ab Q_ab (3 s_a s_b - |s|^2 delta_ab) / |s|^5, each
    component of the multipole table divided by cp1 and multiplied back by it."""
    jd = [dense_coords(m, h, nf, nmesh) for m in n]
    X, Y, Z = np.meshgrid(*[c[1] for c in jd], indexing="ij")
    JX, JY, JZ = np.meshgrid(*[c[0] for c in jd], indexing="ij")
    nd = [m * nmesh for m in n]
    faces = [(JX < 1) | (JX > nd[0]), (JY < 1) | (JY > nd[1]), (JZ < 1) | (JZ > nd[2])]
    out = np.zeros(X.shape)
    for (ax, ay, az), m in zip(atoms, mp):
        s = np.stack([ax - X, ay - Y, az - Z])
        r = np.sqrt((s ** 2).sum(0))
        q = m[0]
        p = m[1:4]
        Q = np.array([[m[4], m[7], m[9]], [m[7], m[5], m[8]], [m[9], m[8], m[6]]])
        v = q / r + np.einsum("a,a...->...", p, s) / r ** 3
        v = v + sum(Q[a, b] * (3 * s[a] * s[b] - (r ** 2 if a == b else 0)) for a in range(3)
                    for b in range(3)) / r ** 5
        for a in range(3):
            if nperi <= a:
                out += np.where(faces[a], v, 0.0)
    return out


@pytest.mark.parametrize("nperi", [0, 1, 2])
def test_trans_d2c_vh_multipole_halo(orc, nperi):
    n, h, nf, nmesh = (8, 9, 7), 0.5, 3, 2
    rng = np.random.default_rng(11 + nperi)
    natom = 2
    atoms = [tuple(rng.uniform(-1.0, 1.0, 3)) for _ in range(natom)]
    mp = rng.standard_normal((natom, 10))
    cp = np.zeros((8, 2), order="F")
    cp[0] = [1.7, 0.9]
    indspe = np.array([1, 2], dtype=np.int32)
    xmax = [0.5 * m * h for m in n]
    vhd = np.zeros(tuple(m * nmesh for m in n), order="F")
    got = np.zeros(n, order="F")
    orc.call("trans_d2c_vh", natom=natom, num_spe=2, nmesh=nmesh, nperi=nperi, ndisp=6, nf=nf, ncpx=n[0],
             ncpy=n[1], ncpz=n[2], xmax=xmax[0], ymax=xmax[1], zmax=xmax[2], indspe=indspe,
             atmpole=np.asfortranarray(mp.T), cp=cp, atx=np.array([a[0] for a in atoms]),
             aty=np.array([a[1] for a in atoms]), atz=np.array([a[2] for a in atoms]),
             vh_coarse=got, vh_dense=vhd)
    # the same restriction applied to the closed-form halo (only cubic cells have one h here)
    assert len(set(xmax[a] / n[a] for a in range(3))) == 1
    halo = np.asfortranarray(_multipole_halo(n, h, nf, nmesh, nperi, atoms, cp[0], mp))
    vc = np.zeros(halo_shape_c(n, nf), order="F")
    orc.call("interpolation", **{"ncpx": n[0], "ncpy": n[1], "ncpz": n[2], "nmesh": nmesh, "v_dd": halo,
                                 "v_cc": vc, "nperi": nperi, "nf": nf, "is": 2, "ndisp": 6})
    orc.call("overlap_interpolation", ndisp=6, nperi=nperi, ncpa=n[0], ncpb=n[1], ncpc=n[2], nf=nf,
             nf1=nf - 1, v_cc=vc, nflag=2)
    want = interior_c(vc, n, nf)
    assert np.abs(want).max() > 1e-3  # the halo reaches the interior through the stencil
    np.testing.assert_allclose(got, want, rtol=0, atol=1e-12 * np.abs(want).max())


def test_scf_rhoaugdense_scatter(orc):
    nd = (10, 8, 6)
    rng = np.random.default_rng(3)
    natom, nl = 4, 50
    lst = np.asfortranarray(rng.integers(1, np.prod(nd) + 1, size=(nl, natom)).astype(np.int32))
    lst[5, 0] = lst[6, 0]  # a repeated point inside one list
    natinf = np.array([50, 30, 45, 10], dtype=np.int32)
    rho = np.asfortranarray(rng.standard_normal(nd + (2,)))
    aug = np.asfortranarray(rng.standard_normal((nl, natom)))
    ntyppp = np.array([1, 0], dtype=np.int32)
    indspe = np.array([1, 1, 2, 1], dtype=np.int32)
    natprid = np.array([1, 1, 1, 0], dtype=np.int32)
    napsd = np.array([1, 2, 3, 4], dtype=np.int32)
    out = np.zeros(nd, order="F")
    orc.call("scf_rhoaugdense", key_natpri_inps=1, key_pp_paw=1, natom=natom, num_spe=2, nspv=2,
             ncpx_d=nd[0], ncpy_d=nd[1], ncpz_d=nd[2], num_list_d=nl, num_ppcell_d=natom, indspe=indspe,
             natprid=natprid, napsd=napsd, natinfd=natinf, ntyppp=ntyppp, lstvecd2=lst, rho_dense=rho,
             rhoaug3d=aug, rho_aug_dense=out)
    want = (rho[..., 0] + rho[..., 1]).ravel(order="F")
    for a in range(natom):  # atoms 3 (not PAW) and 4 (natprid not inps) add nothing
        if ntyppp[indspe[a] - 1] == 1 and natprid[a] == 1:
            np.add.at(want, lst[:natinf[a], a] - 1, aug[:natinf[a], a])
    np.testing.assert_allclose(out.ravel(order="F"), want, rtol=1e-15, atol=1e-15)


@pytest.mark.parametrize("nspv", [1, 2, 4])
def test_scf_veff_sum(orc, nspv):
    n, h = (6, 5, 4), 0.3
    rng = np.random.default_rng(nspv)
    vloc, vh = rng.standard_normal(n), rng.standard_normal(n)
    vx = np.asfortranarray(rng.standard_normal(n + (nspv,)))
    bias = (0.01, -0.02, 0.03)
    xmax = [0.5 * m * h for m in n]
    veff = np.zeros(n + (nspv,), order="F")
    orc.call("scf_veff", nspv=nspv, ncpx=n[0], ncpy=n[1], ncpz=n[2], xmax=xmax[0], ymax=xmax[1],
             zmax=xmax[2], dx=h, dy=h, dz=h, biasx=bias[0], biasy=bias[1], biasz=bias[2],
             vloc_coarse=np.asfortranarray(vloc), vh_coarse=np.asfortranarray(vh), vx=vx, veff=veff)
    xs = [(np.arange(1, m + 1) - 0.5) * h - xm for m, xm in zip(n, xmax)]
    X, Y, Z = np.meshgrid(*xs, indexing="ij")
    vdia = vloc + vh + bias[0] * X + bias[1] * Y + bias[2] * Z
    for s in range(nspv):
        want = vx[..., s] + (vdia if (nspv < 3 or s < 2) else 0.0)
        np.testing.assert_allclose(veff[..., s], want, rtol=1e-14, atol=1e-14)
