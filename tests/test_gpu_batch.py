"""Batching pattern on the GPU: the persistent, graph-replayed batched step against the same
kernels launched system by system (bit-identical), and each system against the oracle's
one-system restatement (orc_band_sd_step, the orthonormalisation orc_orthogonalization_r, then
orc_scf_hartree); the bands leave each step orthonormal to FP64 rounding."""
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


SMALL = [dict(n=16, h=0.35, natom=2, nbands=5, seed=901), dict(n=24, h=0.3, natom=3, nbands=3, seed=902),
         dict(n=16, h=0.4, natom=3, nbands=7, seed=903), dict(n=20, h=0.35, natom=2, nbands=4, seed=904)]


def _systems(cfgs):
    out = []
    for c in cfgs:
        s = synth.make_system(c["n"], c["h"], c["natom"], c["nbands"], seed=c["seed"])
        psi = synth.smooth_bands(synth.random_bands(s), s.n, s.h)
        q, r = np.linalg.qr(psi)
        psi = np.asfortranarray(q * np.sign(np.diag(r)) / np.sqrt(s.dv))
        out.append((s, psi))
    return out


def _batch(systems):
    from rspacedist_amd.batch import Batch
    b = Batch()
    for s, psi in systems:
        b.add(s, psi)
    return b


def _oracle_step(orc, s, psi0):
    psi = psi0.copy(order="F")
    sval = np.zeros(s.nbands)
    dij = np.asfortranarray(s.dij.reshape(9, 9, s.natom, order="F"))
    orc.call("band_sd_step", n=s.n, h=s.h, nbands=s.nbands, natom=s.natom, num_list=s.num_list,
             lstvec2=s.lstvec2, natinf=s.natinf, vnlocp=s.vnlocp, dij=dij,
             veff=np.asfortranarray(s.veff[..., 0]), svecre=psi, sval=sval)
    # the batched step orthonormalises the updated bands (Cholesky-QR; the reference's
    # Gram-Schmidt gives the same Q)
    orc.call("orthogonalization_r", **synth.abi_args(s), nflag=1, svecre=psi, ssvre=psi.copy(order="F"),
             rspsep=np.zeros((9, s.natom, s.nbands), order="F"))
    rho = (psi ** 2).sum(axis=1)
    z = np.zeros(1)
    vo = np.zeros(s.npts)
    xm = s.xmax
    orc.call("scf_hartree", nd=1, ncpx_d=s.n, ncpy_d=s.n, ncpz_d=s.n, nfh=4, ndisp=6, nperi=3,
             ncgres=801, ncgmin=1, ncgmax=800, xmax=xm, ymax=xm, zmax=xm, epsvh=1e-10,
             rho_aug_dense=rho, boundx=z, boundy=z, boundz=z, vh_dense=vo)
    return psi, sval, vo


def test_batched_and_loop_match_oracle(kk, orc):
    """One step from the start, each system against the oracle, in both modes."""
    systems = _systems(SMALL)
    for batched in (False, True):
        b = _batch(systems)
        b.step(batched=batched)
        for i, (s, psi0) in enumerate(systems):
            psi, sval, vo = _oracle_step(orc, s, psi0)
            pg, sg, vg, ig = b.get(i)
            print(f"batched={batched} system {i}: n={s.n} hartree its {ig}, "
                  f"|dsval| {np.abs(sg - sval).max():.2e} |dvh| {np.abs(vg - vo).max():.2e}")
            np.testing.assert_allclose(sg, sval, rtol=1e-11, atol=1e-12)
            np.testing.assert_allclose(pg, psi, rtol=0, atol=1e-10 * np.abs(psi).max())
            gram = s.dv * (pg.T @ pg)
            assert np.abs(gram - np.eye(s.nbands)).max() < 1e-12
            np.testing.assert_allclose(vg, vo, rtol=0, atol=1e-8 * np.abs(vo).max())
        b.close()


def test_batched_equals_loop_bitwise(kk):
    systems = _systems(SMALL)
    bb, bl = _batch(systems), _batch(systems)
    for step in range(2):
        bb.step(batched=True)
        bl.step(batched=False)
        for i in range(len(systems)):
            pb, sb, vb, ib = bb.get(i)
            pl, sl, vl, il = bl.get(i)
            assert ib == il, (step, i, ib, il)
            np.testing.assert_array_equal(pb, pl)
            np.testing.assert_array_equal(sb, sl)
            np.testing.assert_array_equal(vb, vl)
    print(f"tuned tile P={kk.get_param('batch_tile_p'):.0f}, grid={kk.get_param('batch_grid'):.0f}")


def test_batching_config_48_systems(kk):
    """The 48-system configuration (sizes in {32,48,64,80}^3, 16..128 bands, 2..24 atoms):
    batched (persistent stencil launches + graph replays) vs the per-system launch loop, same
    results bit for bit, bands orthonormal after five steps; throughput printed."""
    from rspacedist_amd import batch as BT
    cfgs = BT.batch_configs(48, seed=4)
    systems = []
    for c in cfgs:
        s = synth.make_system(c["n"], BT.H, c["natom"], c["nbands"], seed=c["seed"])
        systems.append((s, BT.make_bands(s)))
    bb, bl = BT.Batch(), BT.Batch()
    for s, D in systems:
        bb.add(s, D)
        bl.add(s, D)
    for s, D in systems:
        D.free()
    bb.step(batched=True)   # first step: autotune + graph capture
    bl.step(batched=False)
    kk.synchronize()
    reps = 4  # five steps in all
    t0 = time.perf_counter()
    for _ in range(reps):
        bb.step(batched=True)
    t1 = time.perf_counter()
    for _ in range(reps):
        bl.step(batched=False)
    t2 = time.perf_counter()
    tb, tl = (t1 - t0) / reps, (t2 - t1) / reps
    pts = sum(s.npts * s.nbands for s, _ in systems)
    print(f"48 systems, {sum(s.nbands for s, _ in systems)} bands, {pts / 1e6:.1f} M point-bands: "
          f"batched {tb * 1e3:.1f} ms/step ({48 / tb:.0f} systems/s), per-system loop "
          f"{tl * 1e3:.1f} ms/step ({48 / tl:.0f} systems/s), speedup {tl / tb:.2f}x; "
          f"tile P={kk.get_param('batch_tile_p'):.0f} grid={kk.get_param('batch_grid'):.0f}")
    for i in (0, 17, 47):
        pb, sb, vb, ib = bb.get(i)
        pl, sl, vl, il = bl.get(i)
        np.testing.assert_array_equal(sb, sl)
        np.testing.assert_array_equal(vb, vl)
        assert ib == il
        s = systems[i][0]
        gram = s.dv * (pb.T @ pb)
        assert np.abs(gram - np.eye(s.nbands)).max() < 1e-11, i
    bb.close()
    bl.close()
