"""Differentiation pattern on the GPU: kukan_band_energy_r / kukan_energy_adjoint_r against
the oracle, and the torch.autograd harness (rspacedist_amd/autograd.py) against central
finite differences of the total energy and against force_eig_r."""
import time

import numpy as np
import pytest

from rspacedist_amd import synth

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def kk():
    from rspacedist_amd import kukan
    kukan.init(0)
    return kukan


def _system(n=12, h=0.45, natom=3, nb=6, seed=51):
    s = synth.make_system(n, h, natom, nb, seed=seed)
    psi = synth.smooth_bands(synth.random_bands(s), s.n, s.h)
    return s, np.asfortranarray(psi), np.linspace(1.0, 0.5, nb)


def _kw(s, pr):
    kw = synth.abi_args(s)
    kw.update({k: v for k, v in pr.items() if k != "dvnlocp"})
    kw["num_list"] = pr["vnlocp"].shape[0]
    return kw


def test_band_energy_and_adjoint_match_oracle(kk, orc):
    s, psi, f = _system()
    pr = synth.build_projectors(s.n, s.h, s.atoms, derivatives=True)
    kw = _kw(s, pr)
    out = []
    for lib in (kk, orc):
        e = np.zeros(s.nbands)
        dedv = np.zeros(s.npts)
        dedr = np.zeros((s.natom, 3), order="F")
        lib.call("band_energy_r", **kw, svecre=psi, eband=e)
        lib.call("energy_adjoint_r", **kw, dvnlocp=pr["dvnlocp"], svecre=psi, fnele=f,
                 dedv=dedv, dedr=dedr)
        out.append((e, dedv, dedr))
    (eg, vg, rg), (eo, vo, ro) = out
    np.testing.assert_allclose(eg, eo, rtol=1e-12, atol=1e-13)
    np.testing.assert_allclose(vg, vo, rtol=1e-13, atol=1e-18)
    np.testing.assert_allclose(rg, ro, rtol=1e-11, atol=1e-14)


def test_autograd_forces_match_central_differences(kk):
    """R.grad from torch autograd (projector part from the GPU adjoint, local part through
    V_loc(R)) against (E(R + h e) - E(R - h e)) / 2h with h = 1e-4 bohr, all 3 natom
    coordinates; and dE/dV_loc against a difference quotient along a random direction."""
    import torch
    from rspacedist_amd import autograd as AG
    s, psi, f = _system()
    st = AG.BandState(s, psi, f)
    st.hartree()
    E, F = AG.forces(s.atoms, st)
    hd = 1e-4
    for a in range(s.natom):
        for d in range(3):
            Rp, Rm = s.atoms.copy(), s.atoms.copy()
            Rp[a, d] += hd
            Rm[a, d] -= hd
            fd = -(AG.energy_at(Rp, st) - AG.energy_at(Rm, st)) / (2 * hd)
            assert abs(F[a, d] - fd) <= 1e-6 * max(1.0, abs(fd)), (a, d, F[a, d], fd)
    veff = torch.tensor(np.asarray(s.veff[..., 0]), dtype=torch.float64, requires_grad=True)
    R = torch.tensor(s.atoms)
    Eb = AG.BandEnergy.apply(veff, R, st)
    Eb.backward()
    rng = np.random.default_rng(4)
    dv = torch.tensor(rng.standard_normal(veff.shape))
    with torch.no_grad():
        ep = AG.BandEnergy.apply(veff + 1e-3 * dv, R, st)
        em = AG.BandEnergy.apply(veff - 1e-3 * dv, R, st)
    np.testing.assert_allclose(float((ep - em) / 2e-3), float((veff.grad * dv).sum()), rtol=1e-9)


def test_adjoint_projector_force_matches_force_eig_r(kk):
    """-dE_nl/dR from the adjoint equals the reference's force_eig_r (FD derivative of psi)
    up to that derivative's discretisation error."""
    s, psi, f = _system(n=16, h=0.4)
    pr = synth.build_projectors(s.n, s.h, s.atoms, derivatives=True)
    kw = _kw(s, pr)
    dedv = np.zeros(s.npts)
    dedr = np.zeros((s.natom, 3), order="F")
    kk.call("energy_adjoint_r", **kw, dvnlocp=pr["dvnlocp"], svecre=psi, fnele=f, dedv=dedv,
            dedr=dedr)
    fx, fy, fz = np.zeros(s.natom), np.zeros(s.natom), np.zeros(s.natom)
    kk.call("force_eig_r", **kw, svecre=psi, sval=np.zeros(s.nbands), fnele=f, fatx=fx, faty=fy,
            fatz=fz)
    fref = np.stack([fx, fy, fz], axis=1)
    assert np.abs(-dedr - fref).max() <= 2e-3 * np.abs(fref).max()


def test_differentiation_config_96cube_256bands(kk):
    """The 96^3 x 256-band x 32-atom configuration: forward + adjoint on the GPU, forces
    against central differences (h = 1e-4) on three coordinates, timings printed."""
    from rspacedist_amd import autograd as AG
    c = synth.CONFIGS[3]
    s = synth.make_system(c["n"], c["h"], c["natom"], c["nbands"], seed=3)
    D = kk.DeviceArray((s.npts, s.nbands))
    D.fill_normal(33)
    kw = synth.abi_args(s)
    kk.call("orthogonalization_r", **kw, nflag=1, svecre=D, ssvre=D,
            rspsep=np.zeros((9, s.natom, s.nbands), order="F"))
    st = AG.BandState(s, D)
    st.hartree()
    t0 = time.perf_counter()
    E, F = AG.forces(s.atoms, st)
    t1 = time.perf_counter()
    print(f"96^3 x 256 bands x 32 atoms: E = {E:.10f} Ha, forward+backward {t1 - t0:.2f} s wall "
          f"(host projector tables included)")
    hd = 1e-4
    for a, d in ((0, 0), (7, 1), (31, 2)):
        Rp, Rm = s.atoms.copy(), s.atoms.copy()
        Rp[a, d] += hd
        Rm[a, d] -= hd
        fd = -(AG.energy_at(Rp, st) - AG.energy_at(Rm, st)) / (2 * hd)
        print(f"  F[{a},{d}] adjoint {F[a, d]:+.10e}  central difference {fd:+.10e}")
        assert abs(F[a, d] - fd) <= 1e-5 * max(1.0, abs(fd))
