"""kukan_scf_hartree_noalloc (the Hartree solve on the library's fixed per-GPU workspace, for
callers that cannot allocate) against kukan_scf_hartree on the same device inputs: two calls
in a row, the second on a smaller grid (the workspace keeps the first call's contents), a
Dirichlet (isolated) case, and the size limit.  The no-alloc entry point always runs the
Hartree CG as iteration blocks; with the allocating entry point's opt-in persistent CG off
(hartree_persistent 0, the default) both run the same kernels in the same order, so the
potentials and iteration counts agree exactly."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def kk():
    from rspacedist_amd import kukan
    kukan.init(0)
    kukan.set_param("hartree_persistent", 0)
    yield kukan
    kukan.set_param("hartree_persistent", 0)


def _solve(kk, name, n, rho, vh0, nperi=3, nd=1, bounds=None, epsvh=1e-10):
    D = kk.DeviceArray
    r = D.from_numpy(rho)
    v = D.from_numpy(vh0)
    z = np.zeros(1)
    b = [D.from_numpy(x) for x in bounds] if bounds else [D.from_numpy(z)] * 3
    h = 0.3
    xm = n * h / 2
    kk.call(name, nd=nd, ncpx_d=n, ncpy_d=n, ncpz_d=n, nfh=4, ndisp=6, nperi=nperi, ncgres=801,
            ncgmin=1, ncgmax=2000, xmax=xm, ymax=xm, zmax=xm, epsvh=epsvh, rho_aug_dense=r,
            boundx=b[0], boundy=b[1], boundz=b[2], vh_dense=v)
    its = int(kk.get_param("hartree_iterations"))
    out = v.numpy()
    for x in [r, v] + (b if bounds else [b[0]]):
        x.free()
    return out, its


@pytest.mark.parametrize("first,second", [(64, 40), (96, 24)])
def test_noalloc_equals_allocating_entry_point_two_calls(kk, first, second):
    rng = np.random.default_rng(first)
    for n in (first, second):  # second call: smaller grid on the workspace the first one left
        rho = np.asfortranarray(rng.random((n, n, n)))
        vh0 = np.asfortranarray(0.01 * rng.standard_normal((n, n, n)))  # warm start
        ref, it_ref = _solve(kk, "scf_hartree", n, rho, vh0)
        got, it_got = _solve(kk, "scf_hartree_noalloc", n, rho, vh0)
        print(f"{n}^3: {it_got} iterations (allocating entry point {it_ref})")
        assert it_got == it_ref
        np.testing.assert_array_equal(got, ref)


def test_noalloc_isolated_cell(kk):
    n, nfh = 20, 4
    rng = np.random.default_rng(5)
    rho = np.asfortranarray(rng.random((n, n, n)))
    bounds = [np.asfortranarray(rng.random((2 * nfh, n, n))), np.asfortranarray(rng.random((n, 2 * nfh, n))),
              np.asfortranarray(rng.random((n, n, 2 * nfh)))]
    vh0 = np.zeros((n, n, n), order="F")
    ref, it_ref = _solve(kk, "scf_hartree", n, rho, vh0, nperi=0, nd=0, bounds=bounds, epsvh=1e-12)
    got, it_got = _solve(kk, "scf_hartree_noalloc", n, rho, vh0, nperi=0, nd=0, bounds=bounds, epsvh=1e-12)
    assert it_got == it_ref
    np.testing.assert_array_equal(got, ref)


def test_noalloc_rejects_oversized_grid_and_host_arrays(kk):
    from rspacedist_amd import kukan
    n = 104  # 104^3 = 1,124,864 points > 1,048,576
    D = kk.DeviceArray
    r = D((n ** 3,))
    v = D((n ** 3,))
    z = np.zeros(1)
    with pytest.raises(kukan.KukanError, match="exceeds the workspace limit"):
        kk.call("scf_hartree_noalloc", nd=1, ncpx_d=n, ncpy_d=n, ncpz_d=n, nfh=4, ndisp=6, nperi=3,
                ncgres=801, ncgmin=1, ncgmax=10, xmax=1.0, ymax=1.0, zmax=1.0, epsvh=1e-10,
                rho_aug_dense=r, boundx=z, boundy=z, boundz=z, vh_dense=v)
    r.free()
    v.free()
    with pytest.raises(kukan.KukanError, match="must be device pointers"):
        kk.call("scf_hartree_noalloc", nd=1, ncpx_d=8, ncpy_d=8, ncpz_d=8, nfh=4, ndisp=6, nperi=3,
                ncgres=801, ncgmin=1, ncgmax=10, xmax=1.0, ymax=1.0, zmax=1.0, epsvh=1e-10,
                rho_aug_dense=np.zeros(512), boundx=z, boundy=z, boundz=z, vh_dense=np.zeros(512))
