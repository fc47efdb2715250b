"""Hartree CG as device-side iteration blocks replayed from a captured HIP graph.

Every replay must reproduce the eagerly launched blocks bit for bit (same kernels, same
order), also after other eager work on the stream, after the graph is rebuilt for another
grid, and with an iteration cap that stops inside a block (scf_hartree8f.F90 stopping rule:
ncgmin, ncgmax, dp <= epsvh)."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def kk():
    from rspacedist_amd import kukan
    kukan.init(0)
    kukan.set_param("hartree_persistent", 0)  # the iteration blocks (the persistent CG has its own tests)
    yield kukan
    kukan.set_param("hartree_graph", 1)
    kukan.set_param("hartree_block", 16)
    kukan.set_param("hartree_persistent", 0)


def _solve(kk, n, h, rho, vh0, graph, ncgmax=800, eps=1e-10, nperi=3):
    kk.set_param("hartree_graph", 1 if graph else 0)
    xmax = n * h / 2
    z = np.zeros(1)
    vh = vh0.copy(order="F")
    kk.call("scf_hartree", nd=1, ncpx_d=n, ncpy_d=n, ncpz_d=n, nfh=4, ndisp=6, nperi=nperi,
            ncgres=801, ncgmin=1, ncgmax=ncgmax, xmax=xmax, ymax=xmax, zmax=xmax, epsvh=eps,
            rho_aug_dense=rho, boundx=z, boundy=z, boundz=z, vh_dense=vh)
    return vh, int(kk.get_param("hartree_iterations"))


def test_graph_replays_match_eager(kk, orc):
    rng = np.random.default_rng(11)
    for n, h in ((24, 0.3), (32, 0.25), (24, 0.3)):   # rebuild for a new grid, then back
        rho = np.asfortranarray(rng.random((n, n, n)))
        vh0 = np.zeros((n, n, n), order="F")
        ref, it_ref = _solve(kk, n, h, rho, vh0, graph=False)
        for rep in range(3):
            got, it = _solve(kk, n, h, rho, vh0, graph=True)
            np.testing.assert_array_equal(got, ref)
            assert it == it_ref
            # eager work between replays on the same stream
            _solve(kk, n, h, rho, vh0, graph=False)
        # warm start from a partially converged potential: fewer iterations, same answer
        half, _ = _solve(kk, n, h, rho, vh0, graph=True, ncgmax=7)
        got, it = _solve(kk, n, h, rho, half, graph=True)
        ref2, it2 = _solve(kk, n, h, rho, half, graph=False)
        np.testing.assert_array_equal(got, ref2)
        assert it == it2
    # and the device-side loop solves what the oracle's host-driven loop solves
    n, h = 24, 0.3
    rho = np.asfortranarray(np.random.default_rng(12).random((n, n, n)))
    z = np.zeros(1)
    vo = np.zeros((n, n, n), order="F")
    xmax = n * h / 2
    orc.call("scf_hartree", nd=1, ncpx_d=n, ncpy_d=n, ncpz_d=n, nfh=4, ndisp=6, nperi=3,
             ncgres=801, ncgmin=1, ncgmax=800, xmax=xmax, ymax=xmax, zmax=xmax, epsvh=1e-10,
             rho_aug_dense=rho, boundx=z, boundy=z, boundz=z, vh_dense=vo)
    vg, _ = _solve(kk, n, h, rho, np.zeros((n, n, n), order="F"), graph=True)
    np.testing.assert_allclose(vg, vo, rtol=0, atol=1e-8 * np.abs(vo).max())


def test_iteration_cap_inside_a_block(kk):
    n, h = 24, 0.3
    rho = np.asfortranarray(np.random.default_rng(13).random((n, n, n)))
    vh0 = np.zeros((n, n, n), order="F")
    for cap in (1, 5, 17, 33):
        g, itg = _solve(kk, n, h, rho, vh0, graph=True, ncgmax=cap)
        e, ite = _solve(kk, n, h, rho, vh0, graph=False, ncgmax=cap)
        assert itg == ite == cap
        np.testing.assert_array_equal(g, e)
