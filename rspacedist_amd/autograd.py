"""Differentiation pattern: a torch.autograd.Function over libkukan_hip.so.

    E(V_loc, R) = sum_l f_l dV <psi_l| T + V_loc + V_H + V_nl(R) |psi_l>  +  E_H

with the bands psi held fixed (orthonormal, resident in HBM), V_nl(R) = sum_a |p_a> D <p_a|
built from projectors centred at the atom positions R, and V_H, E_H = dV/2 sum rho V_H from
the Hartree solve of rho = sum_l f_l psi_l^2 (constants here, since psi is fixed).

Forward: ``kukan_band_energy_r`` (e_l = dV <psi_l|H|psi_l> on the GPU).  Backward:
``kukan_energy_adjoint_r``, the hand-written adjoint: dE/dV_eff = dV sum_l f_l psi_l^2 (the
density) and the projector part of dE/dR.  The local part of dE/dR follows from autograd
through V_loc(R) (``vloc_torch``), so ``R.grad`` is the total -F.

Tensors are CPU float64; the GPU is driven only through the C-ABI library (torch's own HIP
runtime is not used).  Reference forces for comparison: force_eig_r (force8f.F90:9-195).
"""
from __future__ import annotations

import numpy as np
import torch

from . import kukan, synth


class BandState:
    """Fixed bands and the system constants of one differentiable energy."""

    def __init__(self, s: synth.System, psi, f: np.ndarray | None = None):
        """psi: (npts, nbands) numpy array (copied to HBM) or a resident kukan.DeviceArray."""
        self.s = s
        self.n, self.h, self.natom, self.nbands = s.n, s.h, s.natom, s.nbands
        self.f = np.ones(s.nbands) if f is None else np.asarray(f, dtype=np.float64)
        self.psi = psi if hasattr(psi, "device_ptr") else \
            kukan.DeviceArray.from_numpy(np.asfortranarray(psi))
        self.fdev = kukan.DeviceArray.from_numpy(self.f)
        self.vh = np.zeros((s.n, s.n, s.n), order="F")
        self.e_h = 0.0

    def kwargs(self, proj: dict) -> dict:
        kw = synth.abi_args(self.s)
        kw.update({k: v for k, v in proj.items() if k != "dvnlocp"})
        kw["num_list"] = proj["vnlocp"].shape[0]
        for k in ("veff", "svecre", "fnele"):
            kw.pop(k, None)
        return kw

    def hartree(self) -> None:
        """V_H and E_H of rho = sum_l f_l psi_l^2 (periodic cell, CG to 1e-10)."""
        n = self.n
        rho = np.zeros(n ** 3)
        kukan.call("scf_charge_smt_r", neigmx=self.nbands, nums=1, nspv=1, numk=1, ncpx=n,
                   ncpy=n, ncpz=n, fnele=self.fdev, svecre=self.psi, rhosmt=rho)
        z = np.zeros(1)
        vh = np.zeros(n ** 3)
        kukan.call("scf_hartree", nd=1, ncpx_d=n, ncpy_d=n, ncpz_d=n, nfh=4, ndisp=6, nperi=3,
                   ncgres=801, ncgmin=1, ncgmax=800, xmax=self.s.xmax, ymax=self.s.xmax,
                   zmax=self.s.xmax, epsvh=1e-10, rho_aug_dense=rho, boundx=z, boundy=z, boundz=z,
                   vh_dense=vh)
        self.vh = vh.reshape((n, n, n), order="F")
        self.e_h = 0.5 * self.h ** 3 * float(np.dot(rho, vh))


def vloc_torch(R: torch.Tensor, n: int, h: float) -> torch.Tensor:
    """V_loc(r) = -DEPTH sum_a exp(-|r - R_a|^2_mic / (2 sigma^2)) on the cell-centred grid
    (same recipe as synth.build_vloc), differentiable in R.  Shape (n, n, n), x fastest in
    Fortran order, i.e. index [ix, iy, iz]."""
    L = n * h
    c = torch.arange(1, n + 1, dtype=torch.float64) * h - 0.5 * L - 0.5 * h
    v = torch.zeros((n, n, n), dtype=torch.float64)
    two_s2 = 2.0 * synth.SIGMA_VLOC ** 2
    for a in range(R.shape[0]):
        d = [c - R[a, k] for k in range(3)]
        d = [x - L * torch.round(x.detach() / L) for x in d]  # minimum image (shift is constant)
        ex = torch.exp(-d[0] ** 2 / two_s2)
        ey = torch.exp(-d[1] ** 2 / two_s2)
        ez = torch.exp(-d[2] ** 2 / two_s2)
        v = v - synth.DEPTH_VLOC * ex[:, None, None] * ey[None, :, None] * ez[None, None, :]
    return v


class BandEnergy(torch.autograd.Function):
    """E_band(veff, R) = sum_l f_l dV <psi_l|T + veff + V_nl(R)|psi_l> on the GPU."""

    @staticmethod
    def forward(ctx, veff: torch.Tensor, R: torch.Tensor, state: BandState):
        proj = synth.build_projectors(state.n, state.h, R.detach().numpy(), derivatives=True)
        kw = state.kwargs(proj)
        e = np.zeros(state.nbands)
        v = np.asfortranarray(veff.detach().numpy())
        kukan.call("band_energy_r", **kw, veff=v, svecre=state.psi, eband=e)
        ctx.state, ctx.proj, ctx.shape = state, proj, veff.shape
        return torch.tensor(float(np.dot(state.f, e)), dtype=torch.float64)

    @staticmethod
    def backward(ctx, g):
        state, proj = ctx.state, ctx.proj
        kw = state.kwargs(proj)
        dedv = np.zeros(state.n ** 3)
        dedr = np.zeros((state.natom, 3), order="F")
        kukan.call("energy_adjoint_r", **kw, dvnlocp=proj["dvnlocp"], svecre=state.psi,
                   fnele=state.fdev, dedv=dedv, dedr=dedr)
        gv = torch.from_numpy(dedv.reshape(ctx.shape, order="F").copy()) * g
        gr = torch.from_numpy(np.ascontiguousarray(dedr)) * g
        return gv, gr, None


def total_energy(R: torch.Tensor, state: BandState) -> torch.Tensor:
    """E(R) with V_loc(R) built in torch (local forces by autograd) and V_H, E_H fixed."""
    veff = vloc_torch(R, state.n, state.h) + torch.from_numpy(state.vh)
    return BandEnergy.apply(veff, R, state) + state.e_h


def forces(R: np.ndarray, state: BandState) -> tuple[float, np.ndarray]:
    """(E, F = -dE/dR) at positions R (natom, 3)."""
    Rt = torch.tensor(R, dtype=torch.float64, requires_grad=True)
    E = total_energy(Rt, state)
    E.backward()
    return float(E), -Rt.grad.numpy().copy()


def energy_at(R: np.ndarray, state: BandState) -> float:
    with torch.no_grad():
        return float(total_energy(torch.tensor(R, dtype=torch.float64), state))


__all__ = ["BandState", "BandEnergy", "vloc_torch", "total_energy", "forces", "energy_at"]
