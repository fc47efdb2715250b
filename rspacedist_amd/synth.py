"""Synthetic Kohn-Sham problems in the reference's array layouts.

The benchmark configurations are periodic cubic cells at the Gamma point with real
wavefunctions.  Everything a call of the hot path consumes is built here in exactly the
shapes the reference allocates (dim_global8f.F90:28-45, Fortran order):

* grid: ``2*nxmax`` points per axis, spacing ``h``, point ``j`` (1-based) at
  ``x_j = j*h - xmax - h/2`` (read_input_kukan8f.F90:373-378, kslaplacian8f.F90:583-586);
* ``veff(ncpx,ncpy,ncpz,nspv)``: V_loc = -2 sum_a exp(-|r-R_a|^2_mic / (2*0.8^2)) Ha;
* projectors: 9 per atom in the reference's real-harmonic order s, p_x, p_y, p_z,
  d_{x^2-y^2}, -d_{zx}, d_{3z^2-r^2}, -d_{yz}, d_{xy} (solid-harmonic forms of
  setup_initialwave8f.F90:83-91), radial part ``exp(-r^2/(2 sigma_p^2))`` with
  sigma_p = 0.5 bohr times the smooth cut-off ``(1-(r/r_c)^2)^3`` (r_c = 2.0 bohr), each
  normalised to unit L2 norm in the continuum; stored as ``vnlocp = p * dV``
  (pseudocalc8f.F90:3612-3629);
* index lists ``lstvec2(num_list,num_ppcell)`` (1-based, x fastest) and offsets
  ``lstx/lsty/lstz`` built in the order of pseudocalc8f.F90:954-1033 (z outer, x inner);
* ``dij(nprjmx,nprjmx,nums,natom)``: one seeded symmetric 9x9 with entries U[-0.5,0.5]
  replicated per atom; ``sss = 0`` (S = 1) unless a positive semi-definite q is asked for
  (``q_scale``, the S != 1 case of PAW data).

The smooth cut-off makes the projectors continuous functions of the atom position, so
energies can be differentiated by central differences.
"""
from __future__ import annotations

import dataclasses
import math

import numpy as np

KEY_NATPRI_IN = 2     # keys8f.F90:16
KEY_NATPRI_INPS = 1   # keys8f.F90:18
KEY_NATPRI_OUT = 0    # keys8f.F90:20
KEY_ORTHO_NOCMPT = 0  # keys8f.F90:89
KEY_ORTHO_CMPT = 1    # keys8f.F90:90
KEY_PP_PAW = 1        # keys8f.F90:12
KEY_SOC_CALC = 1      # keys8f.F90:74

SIGMA_VLOC = 0.8
DEPTH_VLOC = 2.0
SIGMA_P = 0.5
R_CUT = 2.0
NPRJ = 9

# Benchmark workloads: 40^3 CPU-sized case, 96^3 x 512 bands, 128^3 x 1024 bands (the
# split-precision step), 96^3 x 256 bands (forces and adjoints)
CONFIGS = {
    0: dict(n=40, h=0.25, nbands=32, natom=8),
    1: dict(n=96, h=0.20, nbands=512, natom=64),
    2: dict(n=128, h=0.20, nbands=1024, natom=128),
    3: dict(n=96, h=0.20, nbands=256, natom=32),
}


def _fcut(r):
    t = np.clip(1.0 - (r / R_CUT) ** 2, 0.0, None)
    return np.where(r < R_CUT, t ** 3, 0.0)


def _fcut_dr(r):
    t = np.clip(1.0 - (r / R_CUT) ** 2, 0.0, None)
    return np.where(r < R_CUT, 3.0 * t ** 2 * (-2.0 * r / R_CUT ** 2), 0.0)


def radial_norm(l: int) -> float:
    """N_l with N_l^2 * int_0^rc r^(2l+2) exp(-r^2/sigma^2) fcut(r)^2 dr = 1 (Simpson, 20001 pts)."""
    r = np.linspace(0.0, R_CUT, 20001)
    f = r ** (2 * l + 2) * np.exp(-(r ** 2) / SIGMA_P ** 2) * _fcut(r) ** 2
    hh = r[1] - r[0]
    integral = hh / 3.0 * (f[0] + f[-1] + 4.0 * f[1:-1:2].sum() + 2.0 * f[2:-1:2].sum())
    return 1.0 / math.sqrt(integral)


_NORM = [radial_norm(0), radial_norm(1), radial_norm(2)]
_C_S = math.sqrt(1.0 / (4.0 * math.pi))
_C_P = math.sqrt(3.0 / (4.0 * math.pi))
_C_D1 = math.sqrt(15.0 / (16.0 * math.pi))
_C_D2 = math.sqrt(15.0 / (4.0 * math.pi))
_C_D0 = math.sqrt(5.0 / (16.0 * math.pi))


def solid_harmonics(x, y, z):
    """The 9 real solid harmonics r^l Y_lm in the reference order, shape (9, ...)."""
    r2 = x * x + y * y + z * z
    return np.stack([
        _C_S * np.ones_like(x),
        _C_P * x, _C_P * y, _C_P * z,
        _C_D1 * (x * x - y * y), -_C_D2 * z * x, _C_D0 * (3.0 * z * z - r2),
        -_C_D2 * y * z, _C_D2 * x * y,
    ])


def solid_harmonics_grad(x, y, z):
    """d/dx, d/dy, d/dz of solid_harmonics, shape (3, 9, ...)."""
    o, zr = np.ones_like(x), np.zeros_like(x)
    gx = [zr, _C_P * o, zr, zr, _C_D1 * 2 * x, -_C_D2 * z, _C_D0 * (-2 * x), zr, _C_D2 * y]
    gy = [zr, zr, _C_P * o, zr, _C_D1 * (-2 * y), zr, _C_D0 * (-2 * y), -_C_D2 * z, _C_D2 * x]
    gz = [zr, zr, zr, _C_P * o, zr, -_C_D2 * x, _C_D0 * (4 * z), -_C_D2 * y, zr]
    return np.stack([np.stack(gx), np.stack(gy), np.stack(gz)])


_LOF = np.array([0, 1, 1, 1, 2, 2, 2, 2, 2])


def projector_values(dx, dy, dz):
    """p_i at displacement (point - atom); shape (9, npts)."""
    r = np.sqrt(dx * dx + dy * dy + dz * dz)
    g = np.exp(-(r ** 2) / (2 * SIGMA_P ** 2)) * _fcut(r)
    norm = np.array([_NORM[l] for l in _LOF])[:, None]
    return norm * solid_harmonics(dx, dy, dz) * g[None, :]


def projector_grad(dx, dy, dz):
    """d p_i / d(displacement) ; shape (3, 9, npts)."""
    r = np.sqrt(dx * dx + dy * dy + dz * dz)
    e = np.exp(-(r ** 2) / (2 * SIGMA_P ** 2))
    fc, fcd = _fcut(r), _fcut_dr(r)
    g = e * fc
    rs = np.where(r > 0, r, 1.0)
    dg_dr = e * (-r / SIGMA_P ** 2) * fc + e * fcd
    norm = np.array([_NORM[l] for l in _LOF])[:, None]
    S = solid_harmonics(dx, dy, dz)
    dS = solid_harmonics_grad(dx, dy, dz)
    out = np.empty((3,) + S.shape)
    for a, comp in enumerate((dx, dy, dz)):
        dgd = np.where(r > 0, dg_dr * comp / rs, 0.0)
        out[a] = norm * (dS[a] * g[None, :] + S * dgd[None, :])
    return out


@dataclasses.dataclass
class System:
    """One synthetic cell in reference layouts (all arrays Fortran-ordered)."""
    n: int
    h: float
    natom: int
    nbands: int
    seed: int
    nf: int = 4
    nums: int = 1
    nspv: int = 1
    numk: int = 1
    nperi: int = 3
    atoms: np.ndarray = None          # (natom, 3) positions, bohr
    veff: np.ndarray = None           # (n, n, n, nspv)
    vnlocp: np.ndarray = None         # (num_list, 9, natom)
    lstvec2: np.ndarray = None        # (num_list, natom) int32, 1-based
    lstx: np.ndarray = None
    lsty: np.ndarray = None
    lstz: np.ndarray = None
    natx: np.ndarray = None
    naty: np.ndarray = None
    natz: np.ndarray = None
    natinf: np.ndarray = None         # (natom,) int32
    dij: np.ndarray = None            # (9, 9, nums, natom)
    sss: np.ndarray = None            # (9, 9, num_spe)
    dmat: np.ndarray = None           # the 9x9 D

    @property
    def xmax(self):
        return 0.5 * self.n * self.h

    @property
    def npts(self):
        return self.n ** 3

    @property
    def dv(self):
        return self.h ** 3

    @property
    def num_list(self):
        return self.vnlocp.shape[0]

    @property
    def nprjmx(self):
        return NPRJ

    @property
    def num_ppcell(self):
        return self.natom

    def int_tables(self):
        """Per-atom integer tables of the reference (species, placement and ownership)."""
        na = self.natom
        return dict(
            nprj=np.array([NPRJ], dtype=np.int32),
            indspe=np.ones(na, dtype=np.int32),
            natpri=np.full(na, KEY_NATPRI_IN, dtype=np.int32),
            naps=np.arange(1, na + 1, dtype=np.int32),
            latom=np.arange(1, na + 1, dtype=np.int32),
            ntyppp=np.array([KEY_PP_PAW], dtype=np.int32),
            natsoc=np.zeros(na, dtype=np.int32),
        )

    def grid_coords(self):
        j = np.arange(1, self.n + 1, dtype=np.float64)
        return j * self.h - self.xmax - 0.5 * self.h


def _mic(d, L):
    return d - L * np.round(d / L)


def build_vloc(n, h, atoms):
    xmax = 0.5 * n * h
    L = n * h
    c = np.arange(1, n + 1) * h - xmax - 0.5 * h
    v = np.zeros((n, n, n), order="F")
    for a in atoms:
        dxx = _mic(c - a[0], L) ** 2
        dyy = _mic(c - a[1], L) ** 2
        dzz = _mic(c - a[2], L) ** 2
        v -= DEPTH_VLOC * np.exp(-(dxx[:, None, None] + dyy[None, :, None] + dzz[None, None, :])
                                 / (2 * SIGMA_VLOC ** 2))
    return v


def vloc_grad_density(n, h, atoms, rho):
    """dE/dR_a of E = dV * sum rho * V_loc(R): (natom, 3)."""
    xmax = 0.5 * n * h
    L = n * h
    c = np.arange(1, n + 1) * h - xmax - 0.5 * h
    out = np.zeros((len(atoms), 3))
    for ia, a in enumerate(atoms):
        dx = _mic(c - a[0], L)
        dy = _mic(c - a[1], L)
        dz = _mic(c - a[2], L)
        g = -DEPTH_VLOC * np.exp(-(dx[:, None, None] ** 2 + dy[None, :, None] ** 2 + dz[None, None, :] ** 2)
                                 / (2 * SIGMA_VLOC ** 2))
        # dV/dR = g * (r - R)/sigma^2
        w = rho * g / SIGMA_VLOC ** 2
        out[ia, 0] = np.sum(w * dx[:, None, None]) * h ** 3
        out[ia, 1] = np.sum(w * dy[None, :, None]) * h ** 3
        out[ia, 2] = np.sum(w * dz[None, None, :]) * h ** 3
    return out


def build_projectors(n, h, atoms, nperi=3, derivatives=False):
    """vnlocp, lstvec2, lstx/y/z, natx/y/z, natinf in the layout of pseudocalc8f.F90:954-1033.

    With ``derivatives``, also dvnlocp (num_list, 9, natom, 3) = d vnlocp / d R_x,y,z on the
    same lists (a projector centred at R is p(r - R), so d/dR = -grad p)."""
    xmax = 0.5 * n * h
    nat = len(atoms)
    npmax = int(math.ceil(R_CUT / h)) + 1
    offs = np.arange(-npmax + 1, npmax + 1)
    oz, oy, ox = np.meshgrid(offs, offs, offs, indexing="ij")  # z outer, x inner
    ox, oy, oz = ox.ravel(), oy.ravel(), oz.ravel()
    lists = []
    for a in atoms:
        nat_i = [int(round((a[k] + 0.5 * h) / h)) for k in range(3)]
        tmp = [a[k] - (nat_i[k] * h - 0.5 * h) for k in range(3)]
        dx = ox * h - tmp[0]
        dy = oy * h - tmp[1]
        dz = oz * h - tmp[2]
        r = np.sqrt(dx * dx + dy * dy + dz * dz)
        sel = r < R_CUT
        nx = n // 2
        i = ox[sel] + nat_i[0] + nx
        j = oy[sel] + nat_i[1] + nx
        k = oz[sel] + nat_i[2] + nx
        i = (i - 1) % n + 1
        j = (j - 1) % n + 1
        k = (k - 1) % n + 1
        lin = (k - 1) * n * n + (j - 1) * n + i
        pv = projector_values(dx[sel], dy[sel], dz[sel]) * h ** 3
        lists.append((lin.astype(np.int32), ox[sel], oy[sel], oz[sel], pv, nat_i,
                      np.stack([dx[sel], dy[sel], dz[sel]])))
    num_list = max(len(t[0]) for t in lists)
    vnlocp = np.zeros((num_list, NPRJ, nat), order="F")
    lstvec2 = np.zeros((num_list, nat), dtype=np.int32, order="F")
    lstx = np.zeros((num_list, nat), dtype=np.int32, order="F")
    lsty = np.zeros_like(lstx)
    lstz = np.zeros_like(lstx)
    natinf = np.zeros(nat, dtype=np.int32)
    natxyz = np.zeros((3, nat), dtype=np.int32)
    dvnlocp = np.zeros((num_list, NPRJ, nat, 3), order="F") if derivatives else None
    for ia, (lin, ix, iy, iz, pv, nat_i, disp) in enumerate(lists):
        m = len(lin)
        natinf[ia] = m
        lstvec2[:m, ia] = lin
        lstx[:m, ia], lsty[:m, ia], lstz[:m, ia] = ix, iy, iz
        vnlocp[:m, :, ia] = pv.T
        natxyz[:, ia] = nat_i
        if derivatives:
            g = projector_grad(disp[0], disp[1], disp[2]) * h ** 3  # (3, 9, m), d/d(r - R)
            for d in range(3):
                dvnlocp[:m, :, ia, d] = -g[d].T
    out = dict(vnlocp=vnlocp, lstvec2=lstvec2, lstx=lstx, lsty=lsty, lstz=lstz,
               natx=natxyz[0].copy(), naty=natxyz[1].copy(), natz=natxyz[2].copy(), natinf=natinf)
    if derivatives:
        out["dvnlocp"] = dvnlocp
    return out


def seeded_dmat(rng):
    a = rng.uniform(-0.5, 0.5, size=(NPRJ, NPRJ))
    return np.triu(a) + np.triu(a, 1).T


def seeded_qmat(seed, scale):
    """A seeded symmetric positive semi-definite 9x9 augmentation matrix q (sss of one
    species): scale * R R^T / 9 with R uniform in [-0.5, 0.5].  With q >= 0 the overlap
    S = 1 + sum_a sum_ij |p_i> q_ij <p_j| is positive definite, as PAW overlaps are."""
    r = np.random.default_rng(seed + 104729).uniform(-0.5, 0.5, size=(NPRJ, NPRJ))
    return scale * (r @ r.T) / NPRJ


def make_system(n, h, natom, nbands, seed, nf=4, atoms=None, dmat=None, q_scale=0.0):
    """Build a System; ``atoms``/``dmat`` override the seeded draws (used for finite differences).
    ``q_scale`` > 0 gives a nonzero overlap matrix sss (seeded_qmat), i.e. S != 1."""
    rng = np.random.default_rng(seed)
    L = n * h
    xmax = 0.5 * L
    at_draw = rng.uniform(-xmax, xmax, size=(natom, 3))
    d_draw = seeded_dmat(rng)
    atoms = at_draw if atoms is None else np.asarray(atoms, dtype=np.float64)
    dmat = d_draw if dmat is None else dmat
    s = System(n=n, h=h, natom=natom, nbands=nbands, seed=seed, nf=nf)
    s.atoms = atoms
    s.dmat = dmat
    s.veff = np.asfortranarray(build_vloc(n, h, atoms)[..., None])
    pr = build_projectors(n, h, atoms)
    for k, v in pr.items():
        setattr(s, k, v)
    s.dij = np.asfortranarray(np.repeat(dmat[:, :, None, None], natom, axis=3).reshape(NPRJ, NPRJ, 1, natom))
    s.sss = np.zeros((NPRJ, NPRJ, 1), order="F")
    if q_scale:
        s.sss[:, :, 0] = seeded_qmat(seed, q_scale)
    return s


def make_config(index, **over):
    c = dict(CONFIGS[index])
    c.update(over)
    return make_system(c["n"], c["h"], c["natom"], c["nbands"], seed=index)


def random_bands(sys: System, seed=None, nbands=None):
    """Seeded N(0,1) bands, shape (n^3, M) Fortran order (not yet orthonormal)."""
    rng = np.random.default_rng(sys.seed + 7919 if seed is None else seed)
    m = sys.nbands if nbands is None else nbands
    return np.asfortranarray(rng.standard_normal((sys.npts, m)))


def abi_args(s: System, **extra):
    """Keyword arguments shared by the entry points for system ``s`` (defaults of
    parameters.inp: nf=4, nfh=4, nsdmax=4, northo=1, ...); ``extra`` adds or overrides."""
    kw = dict(
        natom=s.natom, num_spe=1, neigmx=s.nbands, nums=s.nums, nspv=s.nspv, numk=s.numk,
        nperi=s.nperi, northo=1, ns=1, nk=1, nf=s.nf, nkscg=1, nsdmax=4, nprecon_cg=1,
        nprjmx=NPRJ, num_list=s.num_list, num_ppcell=s.natom, ndisp=6,
        ncpx=s.n, ncpy=s.n, ncpz=s.n,
        key_ortho_nocmpt_innerproduct=KEY_ORTHO_NOCMPT, key_ortho_cmpt_innerproduct=KEY_ORTHO_CMPT,
        key_natpri_in=KEY_NATPRI_IN, key_natpri_inps=KEY_NATPRI_INPS,
        key_natpri_out=KEY_NATPRI_OUT, key_soc_calc=KEY_SOC_CALC, key_pp_paw=KEY_PP_PAW,
        dx=s.h, dy=s.h, dz=s.h, xmax=s.xmax, ymax=s.xmax, zmax=s.xmax, epssd=1.0e-6,
        ndiismax=4, nprecon_diis=1, eps_eig_diis=1.0e-14, alambda_diis=0.5, ratio_diis=0.3,
        alambda_min=0.1, alambda_max=1.0, l_hsv=0,
        nrc=0, nso=0, ncol=1, iscf=1, ncgscf=200, nretcg=2000, nchange=1, nrrz=1,
        npxmax=s.n // 2, npymax=s.n // 2, npzmax=s.n // 2,
        dij=s.dij, veff=s.veff, vnlocp=s.vnlocp, sss=s.sss, lstvec2=s.lstvec2, natinf=s.natinf,
        lstx=s.lstx, lsty=s.lsty, lstz=s.lstz, natx=s.natx, naty=s.naty, natz=s.natz,
        skpxx=np.zeros(1), skpyy=np.zeros(1), skpzz=np.zeros(1), dijsoc=np.zeros(1),
        svalsoc=np.zeros(1),
    )
    kw.update(s.int_tables())
    kw.update(extra)
    return kw


def smooth_bands(psi, n, h, width=None):
    """Low-pass filter (Gaussian in Fourier space, width 2h by default) applied to each band of
    psi (n^3, M).  Seeded N(0,1) bands carry large content near the Nyquist mode, where the
    reference's preconditioner (symbol 1/2 + 1/6 sum cos kh) vanishes; the filtered bands
    converge like the smooth atomic guesses the reference starts from."""
    w = 2.0 * h if width is None else width
    k = 2 * np.pi * np.fft.fftfreq(n, d=h)
    g = np.exp(-0.5 * w * w * (k[:, None, None] ** 2 + k[None, :, None] ** 2 + k[None, None, :] ** 2))
    out = np.empty_like(psi, order="F")
    for j in range(psi.shape[1]):
        f = np.fft.fftn(psi[:, j].reshape((n, n, n), order="F"))
        out[:, j] = np.real(np.fft.ifftn(f * g)).ravel(order="F")
    return out


def random_bands_c(sys: System, seed=None, nbands=None, smooth=True):
    """Seeded complex N(0,1) bands (re and im independent), shape (n^3, M) Fortran order,
    low-pass filtered like smooth_bands unless smooth is False (not yet orthonormal)."""
    rng = np.random.default_rng(sys.seed + 7927 if seed is None else seed)
    m = sys.nbands if nbands is None else nbands
    re = rng.standard_normal((sys.npts, m))
    im = rng.standard_normal((sys.npts, m))
    if smooth:
        re = smooth_bands(np.asfortranarray(re), sys.n, sys.h)
        im = smooth_bands(np.asfortranarray(im), sys.n, sys.h)
    return np.asfortranarray(re + 1j * im)


def bloch_args(s: System, k, **extra):
    """abi_args for the complex (Bloch) path at one k-point k = (kx, ky, kz) (1/bohr): the
    scalar skpx/y/z of the band routines and the per-k-point arrays skpxx/yy/zz (numk = 1)."""
    kx, ky, kz = (float(v) for v in k)
    return abi_args(s, nrc=1, skpx=kx, skpy=ky, skpz=kz, skpxx=np.array([kx]),
                    skpyy=np.array([ky]), skpzz=np.array([kz]), ns1=1, nso=0, ncol=1,
                    natsoc=np.zeros(s.natom, dtype=np.int32), dijsoc=np.zeros(1), **extra)
