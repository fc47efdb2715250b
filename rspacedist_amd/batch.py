"""Batching pattern: many independent periodic systems of uneven sizes on one GPU.

The workload of the batching configuration: 48 systems with grid edges drawn uniformly from
{32, 48, 64, 80}, spacing 0.25 bohr, 16..128 bands and 2..24 atoms each (seeded), the
synthetic V_loc / projector recipe of synth.py.  One step advances every system by a batched
H psi, one preconditioned band update and a Hartree CG solve (include/kukan_hip.h,
kukan_batch_*).  ``Batch.step(batched=True)`` runs the persistent, graph-replayed path;
``batched=False`` runs the same kernels system by system for comparison.
"""
from __future__ import annotations

import ctypes

import numpy as np

from . import kukan, synth

SIZES = (32, 48, 64, 80)
H = 0.25


def batch_configs(nsys: int = 48, seed: int = 4):
    """Seeded (n, nbands, natom, seed) per system: sizes in SIZES, bands 16..128, atoms 2..24."""
    rng = np.random.default_rng(seed)
    n = rng.choice(SIZES, nsys)
    m = rng.integers(16, 129, nsys)
    a = rng.integers(2, 25, nsys)
    return [dict(n=int(n[i]), nbands=int(m[i]), natom=int(a[i]), seed=seed * 1000 + i)
            for i in range(nsys)]


def make_bands(s: synth.System) -> "kukan.DeviceArray":
    """Seeded N(0,1) bands orthonormalised on the GPU (orthogonalization_r)."""
    D = kukan.DeviceArray((s.npts, s.nbands))
    D.fill_normal(s.seed + 17)
    kukan.call("orthogonalization_r", **synth.abi_args(s), nflag=1, svecre=D, ssvre=D,
               rspsep=np.zeros((9, s.natom, s.nbands), order="F"))
    return D


class Batch:
    def __init__(self):
        self.lib = kukan.lib()
        h = ctypes.c_void_p()
        kukan._ck(self.lib.kukan_batch_create(ctypes.byref(h)), "kukan_batch_create")
        self.handle = h
        self.systems = []

    def add(self, s: synth.System, psi) -> int:
        """Add system s with orthonormal bands psi (numpy (npts, nbands) or DeviceArray)."""
        from . import abi
        idx = ctypes.c_int(-1)
        dij = np.asfortranarray(s.dij.reshape(9, 9, s.natom, order="F"))
        veff = np.asfortranarray(s.veff[..., 0])
        p = lambda x, code, nm: abi._ptr(x, code, nm)
        kukan._ck(self.lib.kukan_batch_add_system(
            self.handle, s.n, s.h, s.nbands, s.natom, s.num_list, p(s.lstvec2, "I", "lstvec2"),
            p(s.natinf, "I", "natinf"), p(s.vnlocp, "D", "vnlocp"), p(dij, "D", "dij"),
            p(veff, "D", "veff"), p(psi, "D", "svecre"), ctypes.byref(idx)),
            "kukan_batch_add_system")
        self.systems.append(s)
        return idx.value

    def step(self, batched: bool = True, ncgmax: int = 800, epsvh: float = 1e-10) -> None:
        kukan._ck(self.lib.kukan_batch_step(self.handle, 1 if batched else 0, ncgmax, epsvh),
                  "kukan_batch_step")

    def get(self, i: int):
        """(psi, sval, vh, hartree iterations) of system i."""
        s = self.systems[i]
        psi = np.zeros((s.npts, s.nbands), order="F")
        sval = np.zeros(s.nbands)
        vh = np.zeros(s.npts)
        it = ctypes.c_int(0)
        kukan._ck(self.lib.kukan_batch_get(self.handle, i, psi.ctypes.data, sval.ctypes.data,
                                           vh.ctypes.data, ctypes.byref(it)), "kukan_batch_get")
        return psi, sval, vh, it.value

    def close(self) -> None:
        if self.handle:
            kukan._ck(self.lib.kukan_batch_destroy(self.handle), "kukan_batch_destroy")
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
