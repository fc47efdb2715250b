"""Device-resident benchmark workloads: one synthetic cell (synth.CONFIGS) with every array in
HBM, and the SCF step the benchmark times.

One SCF step (the reference's kukan8f.F90:536,605 loop body with nrrz = 1, nsdmax = 4):

  rho = sum_n psi_n^2 (scf_charge_smt_r) -> Hartree CG to 1e-10, warm-started
  (scf_hartree) -> scf_diag: Rayleigh-Ritz, block-CG iterations, Cholesky-QR, reorder.

Bands start as seeded N(0,1) values drawn on the device (kukan_fill_normal) and are
orthonormalised by orthogonalization_r, so no grid-sized array is built on the host.
"""
from __future__ import annotations

import numpy as np

from . import kukan as kk
from . import synth


class Cell:
    """A synthetic cell with its static tables (dev) and SCF state (st) in HBM."""

    def __init__(self, cfg=None, n=None, h=None, natom=None, nbands=None, seed=None, q_scale=0.0):
        if cfg is not None:
            c = synth.CONFIGS[cfg]
            n, h, natom, nbands = c["n"], c["h"], c["natom"], c["nbands"]
            seed = cfg if seed is None else seed
        self.s = s = synth.make_system(n, h, natom, nbands, seed=seed, q_scale=q_scale)
        npts, m = s.npts, s.nbands
        D = kk.DeviceArray
        self.dev = dict(veff=D.from_numpy(s.veff), vnlocp=D.from_numpy(s.vnlocp),
                        lstvec2=D.from_numpy(s.lstvec2), dij=D.from_numpy(s.dij))
        self.st = dict(svecre=D((npts, m)), hsvre=D((npts, m)), rspsep=D((9, s.natom, m)),
                       sval=D((m,)), residual_states=D((m,)), rho=D((npts,)), vh=D((npts,)),
                       fnele=D.from_numpy(np.ones(m)))
        self.spsi = self.st["svecre"] if q_scale == 0.0 else D((npts, m))
        self.iscf = 0

    def kw(self, **extra):
        return synth.abi_args(self.s, **self.dev, **extra)

    def init_bands(self, seed):
        """Seeded N(0,1) bands, S-orthonormalised (with S psi and <p|psi>); V_H = 0."""
        st = self.st
        st["svecre"].fill_normal(seed)
        kk.call("orthogonalization_r", **self.kw(), nflag=1, svecre=st["svecre"], ssvre=self.spsi,
                rspsep=st["rspsep"])
        zero = np.zeros(self.s.npts)
        kk._ck(kk.lib().kukan_memcpy(st["vh"].device_ptr, zero.ctypes.data, zero.nbytes, 1), "memcpy")
        kk.synchronize()
        self.iscf = 0

    def charge_hartree(self, epsvh=1e-10):
        """rho = sum_n f_n psi_n^2, then the Hartree CG (warm-started); returns its iterations."""
        s, st = self.s, self.st
        z = np.zeros(1)
        kk.call("scf_charge_smt_r", neigmx=s.nbands, nums=1, nspv=1, numk=1, ncpx=s.n, ncpy=s.n,
                ncpz=s.n, fnele=st["fnele"], svecre=st["svecre"], rhosmt=st["rho"])
        kk.call("scf_hartree", nd=1, ncpx_d=s.n, ncpy_d=s.n, ncpz_d=s.n, nfh=4, ndisp=6, nperi=3,
                ncgres=801, ncgmin=1, ncgmax=800, xmax=s.xmax, ymax=s.xmax, zmax=s.xmax,
                epsvh=epsvh, rho_aug_dense=st["rho"], boundx=z, boundy=z, boundz=z,
                vh_dense=st["vh"])
        return int(kk.get_param("hartree_iterations"))

    def diag(self, nsdmax=4, **extra):
        """One scf_diag call (RR before, band update, Cholesky-QR, reorder)."""
        st = self.st
        self.iscf += 1
        kw = self.kw(iscf=self.iscf, nsdmax=nsdmax, epssd=extra.pop("epssd", 1e-12), **extra)
        return kk.call("scf_diag", **kw, svecre=st["svecre"], ssvre=self.spsi, hsvre=st["hsvre"],
                       rspsep=st["rspsep"], sval=st["sval"], residual_states=st["residual_states"])

    def scf_step(self, nsdmax=4):
        """Charge, Hartree, scf_diag; returns the Hartree iteration count."""
        it = self.charge_hartree()
        self.diag(nsdmax=nsdmax)
        return it

    def forces(self):
        """force_eig_r (FP64) of the current bands: (natom, 3) Ha/bohr."""
        s, st = self.s, self.st
        fx, fy, fz = np.zeros(s.natom), np.zeros(s.natom), np.zeros(s.natom)
        kk.call("force_eig_r", **self.kw(), svecre=st["svecre"], sval=st["sval"],
                fnele=st["fnele"], fatx=fx, faty=fy, fatz=fz)
        return np.stack([fx, fy, fz], axis=1)

    def total_energy(self):
        """E = sum_n f_n eps_n + E_H of the current bands: the band energy of totalenergy_eig
        (totalenergy8f.F90:151-199) and E_H = 1/2 dV sum rho V_H with V_H solved (FP64) for
        the current rho.  Returns (E, E_band, E_H)."""
        s, st = self.s, self.st
        self.charge_hartree()
        m = s.nbands
        sval = st["sval"].numpy()
        f = np.ones(m)
        r = kk.call("totalenergy_eig", nums=1, ncol=1, numk=1, neigmx=m,
                    nwskk=np.array([1], dtype=np.int32), nwskptot=1, tf=0.0, sval=sval, fnele=f)
        rho, vh = st["rho"].numpy(), st["vh"].numpy()
        eh = 0.5 * s.dv * float(np.dot(rho, vh))
        return r["eneel"] + eh, r["eneel"], eh

    def snapshot(self):
        """Device copies of the state one SCF step reads."""
        keys = ("svecre", "rspsep", "sval", "residual_states", "vh")
        snap = {}
        for k in keys:
            d = kk.DeviceArray(self.st[k].shape)
            kk._ck(kk.lib().kukan_memcpy(d.device_ptr, self.st[k].device_ptr, self.st[k].nbytes, 3),
                   "memcpy")
            snap[k] = d
        snap["_iscf"] = self.iscf
        kk.synchronize()
        return snap

    def restore(self, snap):
        for k, d in snap.items():
            if k == "_iscf":
                self.iscf = d
                continue
            kk._ck(kk.lib().kukan_memcpy(self.st[k].device_ptr, d.device_ptr, d.nbytes, 3), "memcpy")
        kk.synchronize()

    def free(self):
        for d in list(self.dev.values()) + list(self.st.values()):
            d.free()
        if self.spsi is not self.st["svecre"]:
            self.spsi.free()


def column_rows(d, rows, r0, r1):
    """Rows [r0, r1) of every column of a device (rows x cols) Fortran array, to the host."""
    cols = d.shape[1]
    out = np.empty((r1 - r0, cols), order="F")
    L = kk.lib()
    for j in range(cols):
        kk._ck(L.kukan_memcpy(out[:, j].ctypes.data, d.device_ptr + 8 * (j * rows + r0),
                              8 * (r1 - r0), 2), "memcpy")
    kk.synchronize()
    return out
