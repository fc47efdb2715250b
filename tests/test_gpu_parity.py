"""Parity of libkukan_hip.so (MI355X) with the FP64 oracle, entry point by entry point.

Both libraries are called with the same reference-named arguments on the same seeded
inputs.  Tolerances: bit-exact for index work (halo wrap, reordering), 1e-12..1e-13
relative for single FP64 operator applications (summation order differs), and looser
bounds for iterative routines whose rounding differences are amplified by the iteration
(stated per test).
"""
import math

import numpy as np
import pytest
import scipy.linalg

from rspacedist_amd import synth
from tests import helpers as H

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def kk():
    from rspacedist_amd import kukan
    kukan.init(0)
    kukan.set_param("kscg_mode", 0)
    return kukan


def both(kk, orc, name, arrays, **kw):
    """Run `name` through both libraries on copies of `arrays`; return (gpu, oracle) dicts."""
    out = []
    for f in (kk.call, orc.call):
        a = {k: (v.copy(order="F") if isinstance(v, np.ndarray) else v) for k, v in arrays.items()}
        r = f(name, **kw, **a)
        a["_ret"] = r
        out.append(a)
    return out


def test_halo_fill_bit_exact(kk, orc):
    rng = np.random.default_rng(0)
    for nperi in (0, 1, 2, 3):
        for nsnd in (1, 4):
            v = np.zeros((9 + 8, 8 + 8, 7 + 8), order="F")
            v[4:13, 4:12, 4:11] = rng.standard_normal((9, 8, 7))
            g, o = both(kk, orc, "overlap_finitedifference_r", {"v_cc": v}, nperi=nperi, ncpa=9,
                        ncpb=8, ncpc=7, nf=4, nf1=3, nsnd=nsnd)
            np.testing.assert_array_equal(g["v_cc"], o["v_cc"])


@pytest.mark.parametrize("nf", [1, 4, 8])
def test_kslaplacian_leaf(kk, orc, nf):
    rng = np.random.default_rng(nf)
    n1, n2, n3 = 20, 17, 13
    psi = np.asfortranarray(rng.standard_normal((n1, n2, n3)))
    vre = H.halo(psi, nf, nf)
    veff = np.asfortranarray(rng.standard_normal((n1, n2, n3)))
    avre = np.asfortranarray(rng.standard_normal((n1, n2, n3, 3)))
    g, o = both(kk, orc, "kslaplacian_r", {"vre": vre, "veff": veff, "avre": avre}, ncpx=n1,
                ncpy=n2, ncpz=n3, neigmx=3, nf=nf, dx=0.3, dy=0.35, dz=0.4, l=2)
    np.testing.assert_allclose(g["avre"], o["avre"], rtol=1e-12, atol=1e-12 * np.abs(o["avre"]).max())
    np.testing.assert_array_equal(g["avre"][..., 0], avre[..., 0])


def test_ksprecondition_leaf(kk, orc):
    rng = np.random.default_rng(3)
    psi = np.asfortranarray(rng.standard_normal((12, 10, 9, 2)))
    vre = np.stack([H.halo(psi[..., i], 4, 1) for i in range(2)], axis=-1)
    vre = np.asfortranarray(vre)
    g, o = both(kk, orc, "ksprecondition", {"vre": vre, "avre": np.zeros((12, 10, 9, 2), order="F")},
                ncpx=12, ncpy=10, ncpz=9, nf=4, ncol=2)
    np.testing.assert_allclose(g["avre"], o["avre"], rtol=1e-14, atol=1e-14)


def _nl_system(sym=False):
    s = synth.make_system(16, 0.35, 5, 4, seed=21)
    rng = np.random.default_rng(5)
    d = rng.uniform(-0.5, 0.5, (9, 9))
    if sym:
        d = 0.5 * (d + d.T)
    s.dij = np.asfortranarray(np.repeat(d[:, :, None, None], s.natom, axis=3))
    s.sss = np.asfortranarray(rng.uniform(-0.1, 0.1, (9, 9, 1)))
    return s, rng


def test_nonlocal_r01(kk, orc):
    s, rng = _nl_system()
    psi = rng.standard_normal(s.npts)
    rp = np.asfortranarray(rng.standard_normal((9, s.natom)))
    g, o = both(kk, orc, "nonlocaloperation_r_01", {"vre2": psi, "rpsep": rp}, **synth.abi_args(s))
    np.testing.assert_allclose(g["rpsep"], o["rpsep"], rtol=1e-12, atol=1e-13)


def test_projector_plan_follows_lstvec2_in_place(kk, orc):
    """The library reuses the projector plan of the previous call when the atom tables and
    lstvec2 (address and contents) are the same; changing lstvec2 in place at the same device
    address, or reordering it, must give the new result."""
    import ctypes
    s, rng = _nl_system()
    avre0 = rng.standard_normal(s.npts)
    rp = np.asfortranarray(rng.standard_normal((9, s.natom)))
    args = synth.abi_args(s)
    lst = kk.DeviceArray.from_numpy(np.asfortranarray(args["lstvec2"]))
    args = dict(args, lstvec2=lst, nhs=1)
    h = np.asfortranarray(args["natinf"]).copy()
    for trial in range(3):
        host = np.asfortranarray(synth.abi_args(s)["lstvec2"]).copy(order="F")
        if trial == 1:  # the same points in another order (same checksum class of entries)
            for na in range(s.natom):
                host[: h[na], na] = host[: h[na], na][::-1]
        if trial == 2:  # shifted points
            host[: h[0], 0] = (host[: h[0], 0] % s.npts) + 1
        kk.lib().kukan_memcpy(ctypes.c_void_p(lst.device_ptr), host.ctypes.data, host.nbytes, 1)
        kk.synchronize()
        assert np.array_equal(lst.numpy(), host)
        g = {"rpsep": rp.copy(order="F"), "dijsss": s.dij, "avre": avre0.copy(), "avr": np.zeros(s.num_list)}
        kk.call("nonlocaloperation_r_02", **args, **g)
        o = {"rpsep": rp.copy(order="F"), "dijsss": s.dij, "avre": avre0.copy(), "avr": np.zeros(s.num_list)}
        orc.call("nonlocaloperation_r_02", **dict(args, lstvec2=host), **o)
        np.testing.assert_allclose(g["avre"], o["avre"], rtol=1e-12, atol=1e-12, err_msg=str(trial))
    lst.free()


@pytest.mark.parametrize("nhs", [1, 0])
def test_nonlocal_r02(kk, orc, nhs):
    s, rng = _nl_system()
    rp = np.asfortranarray(rng.standard_normal((9, s.natom)))
    avre = rng.standard_normal(s.npts)
    X = s.dij if nhs == 1 else s.sss
    g, o = both(kk, orc, "nonlocaloperation_r_02", {"rpsep": rp, "dijsss": X, "avre": avre,
                "avr": np.zeros(s.num_list)}, **dict(synth.abi_args(s), nhs=nhs))
    np.testing.assert_allclose(g["avre"], o["avre"], rtol=1e-12, atol=1e-12)


def test_hartree_periodic_and_correction(kk, orc):
    n, h = 24, 0.3
    xmax = n * h / 2
    rho = np.asfortranarray(np.random.default_rng(7).random((n, n, n)))
    z = np.zeros(1)
    arr = {"rho_aug_dense": rho, "boundx": z, "boundy": z, "boundz": z,
           "vh_dense": np.zeros((n, n, n), order="F")}
    kw = dict(nd=1, ncpx_d=n, ncpy_d=n, ncpz_d=n, nfh=4, ndisp=6, nperi=3, ncgres=801, ncgmin=1,
              ncgmax=800, xmax=xmax, ymax=xmax, zmax=xmax, epsvh=1e-11)
    g, o = both(kk, orc, "scf_hartree", arr, **kw)
    # the same CG recurrence from the same start: iterates agree to rounding amplification
    np.testing.assert_allclose(g["vh_dense"], o["vh_dense"], rtol=0, atol=1e-8 * np.abs(o["vh_dense"]).max())
    kc = dict(ncpx_d=n, ncpy_d=n, ncpz_d=n, xmax=xmax, ymax=xmax, zmax=xmax, ddx=h, ddy=h, ddz=h)
    g2, o2 = both(kk, orc, "tools_correctpotential", {"vh_dense": g["vh_dense"]}, **kc)
    np.testing.assert_allclose(g2["vh_dense"], o2["vh_dense"], atol=1e-13 * np.abs(o2["vh_dense"]).max())
    assert abs(g2["vh_dense"].mean()) < 1e-12


def test_hartree_isolated(kk, orc):
    n, h, nfh = 10, 0.4, 4
    xmax = n * h / 2
    rng = np.random.default_rng(8)
    arr = {"rho_aug_dense": np.asfortranarray(rng.random((n, n, n))),
           "boundx": np.asfortranarray(rng.random((2 * nfh, n, n))),
           "boundy": np.asfortranarray(rng.random((n, 2 * nfh, n))),
           "boundz": np.asfortranarray(rng.random((n, n, 2 * nfh))),
           "vh_dense": np.zeros((n, n, n), order="F")}
    kw = dict(nd=0, ncpx_d=n, ncpy_d=n, ncpz_d=n, nfh=nfh, ndisp=6, nperi=0, ncgres=801, ncgmin=1,
              ncgmax=2000, xmax=xmax, ymax=xmax, zmax=xmax, epsvh=1e-12)
    g, o = both(kk, orc, "scf_hartree", arr, **kw)
    np.testing.assert_allclose(g["vh_dense"], o["vh_dense"], atol=1e-8 * np.abs(o["vh_dense"]).max())


def _state(s, orc, m=None, smooth=True):
    return H.init_state(s, orc, nbands=m, smooth=smooth)


def test_orthogonalization_r_block_path(kk, orc):
    s = synth.make_system(12, 0.45, 3, 40, seed=31)
    psi0 = synth.random_bands(s)
    arr = {"svecre": psi0, "ssvre": np.zeros_like(psi0, order="F"),
           "rspsep": np.zeros((9, s.natom, 40), order="F")}
    g, o = both(kk, orc, "orthogonalization_r", arr, **dict(synth.abi_args(s), nflag=1))
    for k in ("svecre", "ssvre", "rspsep"):
        np.testing.assert_allclose(g[k], o[k], atol=1e-10 * np.abs(o[k]).max())
    q = g["svecre"]
    np.testing.assert_allclose(q.T @ q * s.dv, np.eye(40), atol=1e-13)


def test_orthogonalization_r_01_single_band(kk, orc):
    s = synth.make_system(12, 0.45, 3, 8, seed=32)
    st = _state(s, orc)
    rng = np.random.default_rng(1)
    st["svecre"][:, 5] += 0.3 * rng.standard_normal(s.npts)
    arr = {k: st[k] for k in ("svecre", "ssvre", "rspsep")}
    g, o = both(kk, orc, "orthogonalization_r_01", arr, **dict(synth.abi_args(s), l_inp=6, nflag=1))
    for k in arr:
        np.testing.assert_allclose(g[k], o[k], atol=1e-12 * np.abs(o[k]).max())


def test_rayleighritz(kk, orc):
    s = synth.make_system(12, 0.45, 3, 10, seed=33)
    st = _state(s, orc)
    arr = {k: st[k] for k in ("rspsep", "svecre", "ssvre", "hsvre", "sval")}
    g, o = both(kk, orc, "rayleighritz_r", arr, **dict(synth.abi_args(s), l_rrb=1))
    np.testing.assert_allclose(g["sval"], o["sval"], rtol=1e-12, atol=1e-12)
    for k in ("svecre", "ssvre", "rspsep", "hsvre"):
        np.testing.assert_allclose(g[k], o[k], atol=1e-9 * np.abs(o[k]).max())


def test_reorderstates_exact(kk, orc):
    s = synth.make_system(8, 0.5, 2, 7, seed=34)
    rng = np.random.default_rng(2)
    arr = {"sval": np.array([0.3, -0.1, 0.3, -0.5, 0.2, 0.9, -0.1]),
           "residual_states": np.arange(7.0),
           "rspsep": np.asfortranarray(rng.standard_normal((9, s.natom, 7))),
           "svecre": np.asfortranarray(rng.standard_normal((s.npts, 7))),
           "ssvre": np.asfortranarray(rng.standard_normal((s.npts, 7))),
           "hsvre": np.zeros((s.npts, 7), order="F")}
    g, o = both(kk, orc, "reorderstates", arr, **dict(synth.abi_args(s, neigmx=7), ns1=1))
    for k in arr:
        np.testing.assert_array_equal(g[k], o[k])


def test_kscg_r_band_sequential_matches_oracle(kk, orc):
    """kscg_mode=0 is the reference algorithm band by band; iterates agree with the oracle
    up to rounding amplified over nsdmax iterations per band."""
    kk.set_param("kscg_mode", 0)
    s = synth.make_system(12, 0.45, 3, 8, seed=35)
    st = _state(s, orc)
    arr = {k: st[k] for k in ("sval", "rspsep", "svecre", "ssvre", "residual_states", "hsvre")}
    g, o = both(kk, orc, "kscg_r", arr, **synth.abi_args(s))
    np.testing.assert_allclose(g["sval"], o["sval"], rtol=1e-10, atol=1e-10)
    for k in ("svecre", "ssvre", "hsvre", "rspsep"):
        np.testing.assert_allclose(g[k], o[k], atol=1e-8 * np.abs(o[k]).max())
    np.testing.assert_allclose(g["residual_states"], o["residual_states"], rtol=1e-6, atol=1e-14)
    assert g["_ret"]["ksitmax"] == o["_ret"]["ksitmax"]


def test_ksdiis_r_matches_oracle(kk, orc):
    s = synth.make_system(12, 0.45, 3, 8, seed=36)
    st = _state(s, orc)
    arr = {k: st[k] for k in ("rspsep", "sval", "svecre", "ssvre", "residual_states", "hsvre")}
    for l_hsv in (0, 1):
        g, o = both(kk, orc, "ksdiis_r", arr, **synth.abi_args(s, l_hsv=l_hsv))
        np.testing.assert_allclose(g["sval"], o["sval"], rtol=1e-11, atol=1e-11)
        for k in ("svecre", "ssvre", "hsvre", "rspsep"):
            np.testing.assert_allclose(g[k], o[k], atol=1e-9 * np.abs(o[k]).max())
        np.testing.assert_allclose(g["residual_states"], o["residual_states"], rtol=1e-8)


def test_scf_diag_sequential_matches_oracle_over_steps(kk, orc):
    kk.set_param("kscg_mode", 0)
    s = synth.make_system(12, 0.45, 3, 8, seed=37)
    st0 = _state(s, orc)
    res = {}
    for name, f in (("gpu", kk.call), ("orc", orc.call)):
        st = {k: v.copy(order="F") for k, v in st0.items()}
        kw = synth.abi_args(s)
        for it in range(1, 4):
            kw["iscf"] = it
            f("scf_diag", **kw, **st)
        res[name] = st
    np.testing.assert_allclose(res["gpu"]["sval"], res["orc"]["sval"], rtol=1e-9, atol=1e-10)


@pytest.mark.parametrize("proj", [1, 0])
def test_scf_diag_block_converges_to_dense_eigenvalues(kk, orc, proj):
    """Block CG with K r projected on bands 1..b (as kscg_r) or on all bands."""
    kk.set_param("kscg_mode", 1)
    kk.set_param("bcg_proj", proj)
    try:
        s = synth.make_system(10, 0.5, 3, 8, seed=8)
        st = _state(s, orc)
        kw = synth.abi_args(s, nsdmax=4, epssd=1e-12)
        for it in range(1, 41):
            kw["iscf"] = it
            kk.call("scf_diag", **kw, **st)
        w = scipy.linalg.eigh(H.dense_h(s), eigvals_only=True)[:8]
        np.testing.assert_allclose(st["sval"], w, atol=1e-8)
        q = st["svecre"]
        np.testing.assert_allclose(q.T @ q * s.dv, np.eye(8), atol=1e-12)
    finally:
        kk.set_param("bcg_proj", 1)
        kk.set_param("kscg_mode", 0)


def test_block_cg_combination_stencil_matches_materialised_direction(kk, orc):
    """The block CG forms p = fac P + mdf psi inside the H p stencil (with <psi|H p> partials)
    instead of storing p; the same steps with p stored agree to rounding."""
    s = synth.make_system(16, 0.4, 3, 12, seed=81)
    base = _state(s, orc)
    kw = synth.abi_args(s, nsdmax=4, epssd=1e-12)
    out = []
    kk.set_param("kscg_mode", 1)
    try:
        for comb in (1, 0):
            kk.set_param("bcg_comb", comb)
            st = {k: (v.copy(order="F") if isinstance(v, np.ndarray) else v) for k, v in base.items()}
            for it in range(1, 4):
                kw["iscf"] = it
                kk.call("scf_diag", **kw, **st)
            out.append(st)
    finally:
        kk.set_param("kscg_mode", 0)
        kk.set_param("bcg_comb", 1)
    a, b = out
    np.testing.assert_allclose(a["sval"], b["sval"], rtol=0, atol=1e-11)
    np.testing.assert_allclose(a["svecre"], b["svecre"], rtol=0, atol=1e-9 * np.abs(b["svecre"]).max())


@pytest.mark.parametrize("comb", [1, 0])
def test_projector_space_partials_match_grid_partials(kk, orc, comb):
    """<p|H p> and <psi|H p> in the block CG take their projector part from projector space
    (sum coef <proj|.>, no second pass over the bands); the grid-pass partials give the same
    steps to rounding."""
    s = synth.make_system(16, 0.4, 3, 12, seed=83)
    base = _state(s, orc)
    kw = synth.abi_args(s, nsdmax=4, epssd=1e-12)
    out = []
    kk.set_param("kscg_mode", 1)
    kk.set_param("bcg_comb", comb)
    try:
        for pd in (1, 0):
            kk.set_param("nl_proj_dots", pd)
            st = {k: (v.copy(order="F") if isinstance(v, np.ndarray) else v) for k, v in base.items()}
            for it in range(1, 4):
                kw["iscf"] = it
                kk.call("scf_diag", **kw, **st)
            out.append(st)
    finally:
        kk.set_param("kscg_mode", 0)
        kk.set_param("bcg_comb", 1)
        kk.set_param("nl_proj_dots", 1)
    a, b = out
    np.testing.assert_allclose(a["sval"], b["sval"], rtol=0, atol=1e-11)
    np.testing.assert_allclose(a["svecre"], b["svecre"], rtol=0, atol=1e-9 * np.abs(b["svecre"]).max())


@pytest.mark.parametrize("asym", [0, 1])
def test_symmetric_combination_dots(kk, orc, asym):
    """With dij symmetric, H is symmetric and the block CG forms dV <psi|H p> as
    dV <H psi|fac P> + mdf dV <psi|H psi>, so the H p stencil does not read psi; the steps agree
    with the direct sums to rounding.  An asymmetric dij keeps the direct sums (bitwise equal)."""
    s = synth.make_system(16, 0.4, 3, 12, seed=85)
    if asym:
        s.dij = np.asfortranarray(s.dij.copy())
        s.dij[0, 1] += 0.05
    base = _state(s, orc)
    kw = synth.abi_args(s, nsdmax=4, epssd=1e-12)
    out = []
    kk.set_param("kscg_mode", 1)
    try:
        for sym in (1, 0):
            kk.set_param("comb_sym", sym)
            st = {k: (v.copy(order="F") if isinstance(v, np.ndarray) else v) for k, v in base.items()}
            for it in range(1, 4):
                kw["iscf"] = it
                kk.call("scf_diag", **kw, **st)
            out.append(st)
    finally:
        kk.set_param("kscg_mode", 0)
        kk.set_param("comb_sym", 1)
    a, b = out
    if asym:
        np.testing.assert_array_equal(a["sval"], b["sval"])
        np.testing.assert_array_equal(a["svecre"], b["svecre"])
    else:
        np.testing.assert_allclose(a["sval"], b["sval"], rtol=0, atol=1e-11)
        np.testing.assert_allclose(a["svecre"], b["svecre"], rtol=0, atol=1e-9 * np.abs(b["svecre"]).max())


@pytest.mark.parametrize("n,epssd,upd_p", [(16, 1e-12, 1), (16, 3.0, 1), (32, 1e-12, 1), (32, 1e-12, 2)])
def test_fused_band_update_and_preconditioner(kk, orc, n, epssd, upd_p):
    """With H symmetric the block CG's band update and the next residual preconditioner run as
    one stencil pass, the new bands alternating between two buffers; the steps agree with the
    separate passes to rounding.  epssd = 3 makes every band converge early, so the loop also
    ends after a fused update (the bands are copied back into the caller's arrays).  At 32^3
    one and two points per thread give different tilings (two y-tiles or one, 16 or 8
    work-groups per band) from the separate preconditioner's, so the partial sums' layout of
    the fused pass is checked against its own tiling."""
    s = synth.make_system(n, 0.4, 3, 12, seed=87)
    base = _state(s, orc)
    kw = synth.abi_args(s, nsdmax=4, epssd=epssd)
    out = []
    kk.set_param("kscg_mode", 1)
    kk.set_param("upd_p", upd_p)
    try:
        for fu in (1, 0):
            kk.set_param("bcg_fuse_update", fu)
            st = {k: (v.copy(order="F") if isinstance(v, np.ndarray) else v) for k, v in base.items()}
            for it in range(1, 4):
                kw["iscf"] = it
                kk.call("scf_diag", **kw, **st)
            out.append(st)
    finally:
        kk.set_param("kscg_mode", 0)
        kk.set_param("bcg_fuse_update", 1)
        kk.set_param("upd_p", 1)
    a, b = out
    np.testing.assert_allclose(a["sval"], b["sval"], rtol=0, atol=1e-11)
    for k, v in b.items():
        if isinstance(v, np.ndarray) and v.dtype == np.float64 and v.size > 1:
            np.testing.assert_allclose(a[k], v, rtol=0, atol=1e-9 * max(1.0, np.abs(v).max()), err_msg=k)


def test_charge_force_energy(kk, orc):
    s, rng = _nl_system(sym=True)
    s.sss[:] = 0.02
    m = 4
    psi = synth.smooth_bands(synth.random_bands(s, nbands=m), s.n, s.h)
    f = np.array([2.0, 2.0, 1.0, 0.5])
    sv = np.array([-0.4, -0.1, 0.2, 0.5])
    g, o = both(kk, orc, "scf_charge_smt_r", {"fnele": f, "svecre": psi, "rhosmt": np.zeros(s.npts)},
                **synth.abi_args(s, neigmx=m))
    np.testing.assert_allclose(g["rhosmt"], o["rhosmt"], rtol=1e-13, atol=1e-15)
    arr = {"svecre": psi, "sval": sv, "fnele": f, "fatx": np.zeros(s.natom),
           "faty": np.zeros(s.natom), "fatz": np.zeros(s.natom)}
    g, o = both(kk, orc, "force_eig_r", arr, **synth.abi_args(s, neigmx=m))
    for k in ("fatx", "faty", "fatz"):
        np.testing.assert_allclose(g[k], o[k], rtol=1e-10, atol=1e-12 * np.abs(o[k]).max())
    ek = dict(nums=1, ncol=1, numk=1, neigmx=m, nwskk=np.array([1], dtype=np.int32), nwskptot=2, tf=0.01)
    g = kk.call("totalenergy_eig", sval=sv, fnele=f / 2, **ek)
    o = orc.call("totalenergy_eig", sval=sv, fnele=f / 2, **ek)
    assert g["eneel"] == o["eneel"] and g["eneth"] == o["eneth"]


def test_device_pointers_and_binding(kk, orc):
    """Entry points use device arrays in place and keep bound host arrays resident."""
    s = synth.make_system(12, 0.45, 3, 6, seed=38)
    st = _state(s, orc)
    ref = {k: v.copy(order="F") for k, v in st.items()}
    orc.call("rayleighritz_r", **synth.abi_args(s), l_rrb=1,
             **{k: ref[k] for k in ("rspsep", "svecre", "ssvre", "hsvre", "sval")})
    dev = {k: kk.DeviceArray.from_numpy(v) for k, v in st.items()}
    kw = synth.abi_args(s, l_rrb=1)
    for k in ("veff", "vnlocp", "dij", "lstvec2"):
        kw[k] = kk.DeviceArray.from_numpy(kw[k])
    kk.call("rayleighritz_r", **kw, **{k: dev[k] for k in ("rspsep", "svecre", "ssvre", "hsvre", "sval")})
    np.testing.assert_allclose(dev["sval"].numpy(), ref["sval"], rtol=1e-12, atol=1e-12)
    np.testing.assert_allclose(dev["svecre"].numpy(), ref["svecre"], atol=1e-9 * np.abs(ref["svecre"]).max())
    # bound host array: results stay on the device until sync_to_host
    psi = st["svecre"].copy(order="F")
    L = kk.lib()
    assert L.kukan_bind(psi.ctypes.data, psi.nbytes) == 0
    before = psi.copy()
    kk.call("orthogonalization_r", **synth.abi_args(s), nflag=1, svecre=psi,
            ssvre=np.zeros_like(psi, order="F"), rspsep=np.zeros((9, s.natom, 6), order="F"))
    np.testing.assert_array_equal(psi, before)
    assert L.kukan_sync_to_host(psi.ctypes.data) == 0
    np.testing.assert_allclose(psi.T @ psi * s.dv, np.eye(6), atol=1e-12)
    assert L.kukan_unbind(psi.ctypes.data) == 0


def test_fill_normal_statistics(kk):
    d = kk.DeviceArray((1 << 20,)).fill_normal(7)
    x = d.numpy()
    assert abs(x.mean()) < 5e-3 and abs(x.std() - 1.0) < 5e-3
    np.testing.assert_array_equal(x, kk.DeviceArray((1 << 20,)).fill_normal(7).numpy())


def test_fortran_shim_driver(kk, orc, tmp_path):
    """The Fortran drop-in modules, called as the reference host calls them, give the
    oracle's results (kslaplacian_r with halo exchange, ksprecondition, scf_hartree and
    tools_correctpotential on a 12^3 periodic cell)."""
    import os
    import subprocess
    exe = os.path.join(os.path.dirname(kk.LIB_PATH), "fortran", "shim_driver")
    out = str(tmp_path / "shim.bin")
    r = subprocess.run([exe, out], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    n, nf, h = 12, 4, 0.4
    got = np.fromfile(out, dtype=np.float64)
    avre_g, kout_g, vh_g = got[:2 * n ** 3], got[2 * n ** 3:3 * n ** 3], got[3 * n ** 3:]
    i = np.arange(1, n + 1) * h
    X, Y, Z = np.meshgrid(i, i, i, indexing="ij")
    L = n * h
    psi = np.sin(2 * np.pi * X / L) * np.cos(4 * np.pi * Y / L) + 0.3 * np.cos(2 * np.pi * Z / L)
    ii = np.arange(1, n + 1)
    IX, IY, IZ = np.meshgrid(ii, ii, ii, indexing="ij")
    veff = np.asfortranarray(-1.0 + 0.1 * IX - 0.05 * IY + 0.02 * IZ)
    rho = np.asfortranarray(1.0 + 0.5 * np.sin(2 * np.pi * X / L) * np.sin(2 * np.pi * Y / L)
                            + 0.25 * np.cos(2 * np.pi * Z / L))
    avre = np.ones((n, n, n, 2), order="F")
    orc.call("kslaplacian_r", ncpx=n, ncpy=n, ncpz=n, neigmx=2, nf=nf, dx=h, dy=h, dz=h,
             vre=H.halo(np.asfortranarray(psi), nf, nf), veff=veff, l=2, avre=avre)
    np.testing.assert_allclose(avre_g, avre.ravel(order="F"), rtol=1e-12, atol=1e-12)
    kout = np.zeros((n, n, n), order="F")
    orc.call("ksprecondition", ncpx=n, ncpy=n, ncpz=n, nf=nf, ncol=1,
             vre=H.halo(np.asfortranarray(psi), nf, 1), avre=kout)
    np.testing.assert_allclose(kout_g, kout.ravel(order="F"), rtol=1e-14, atol=1e-14)
    vh = np.zeros((n, n, n), order="F")
    z = np.zeros(1)
    xmax = n * h / 2
    orc.call("scf_hartree", nd=1, ncpx_d=n, ncpy_d=n, ncpz_d=n, nfh=4, ndisp=6, nperi=3, ncgres=801,
             ncgmin=1, ncgmax=800, xmax=xmax, ymax=xmax, zmax=xmax, epsvh=1e-11, rho_aug_dense=rho,
             boundx=z, boundy=z, boundz=z, vh_dense=vh)
    orc.call("tools_correctpotential", ncpx_d=n, ncpy_d=n, ncpz_d=n, xmax=xmax, ymax=xmax, zmax=xmax,
             ddx=h, ddy=h, ddz=h, vh_dense=vh)
    np.testing.assert_allclose(vh_g, vh.ravel(order="F"), atol=1e-8 * np.abs(vh).max())


def test_error_codes_carry_reference_messages(kk):
    from rspacedist_amd import kukan
    with pytest.raises(kukan.KukanError, match="nhs=0 and ns/=1"):
        s, rng = _nl_system()
        kk.call("nonlocaloperation_r_02", **dict(synth.abi_args(s), nhs=0, ns=2),
                rpsep=np.zeros((9, s.natom), order="F"), dijsss=s.sss, avre=np.zeros(s.npts),
                avr=np.zeros(s.num_list))
    with pytest.raises(kukan.KukanError, match="nfh ne 4, 5 or 6"):
        z = np.zeros(1)
        kk.call("scf_hartree", nd=1, ncpx_d=8, ncpy_d=8, ncpz_d=8, nfh=3, ndisp=6, nperi=3,
                ncgres=801, ncgmin=1, ncgmax=10, xmax=1.0, ymax=1.0, zmax=1.0, epsvh=1e-10,
                rho_aug_dense=np.zeros(512), boundx=z, boundy=z, boundz=z, vh_dense=np.zeros(512))


@pytest.mark.parametrize("nperi", [3, 0])
def test_hpsi_tiled_path_ragged_grid(kk, orc, nperi):
    """H psi through the tiled stencil on a non-cubic grid whose y extent is not a multiple
    of the tile height (200 x 23 x 9), with arbitrary projector lists, via rayleighritz_r."""
    nx, ny, nz, m, natom = 200, 23, 9, 5, 3
    rng = np.random.default_rng(40)
    n = nx * ny * nz
    num_list = 60
    lst = np.zeros((num_list, natom), dtype=np.int32, order="F")
    for a in range(natom):
        lst[:, a] = rng.choice(n, size=num_list, replace=False) + 1
    kw = dict(natom=natom, num_spe=1, num_ppcell=natom, nprjmx=9, neigmx=m, numk=1, nums=1, nspv=1,
              num_list=num_list, ncpx=nx, ncpy=ny, ncpz=nz, nperi=nperi, nk=1, ns=1, nf=4, ndisp=6,
              key_natpri_in=2, key_natpri_inps=1, dx=0.3, dy=0.25, dz=0.35, l_rrb=1,
              nprj=np.array([9], dtype=np.int32), natinf=np.full(natom, num_list, dtype=np.int32),
              lstvec2=lst, indspe=np.ones(natom, dtype=np.int32),
              natpri=np.full(natom, 2, dtype=np.int32), naps=np.arange(1, natom + 1, dtype=np.int32),
              ntyppp=np.array([1], dtype=np.int32),
              dij=np.asfortranarray(rng.uniform(-0.5, 0.5, (9, 9, 1, natom))),
              veff=np.asfortranarray(rng.standard_normal((nx, ny, nz, 1))),
              vnlocp=np.asfortranarray(rng.standard_normal((num_list, 9, natom)) * 0.01))
    psi = np.asfortranarray(rng.standard_normal((n, m)))
    psi, _ = np.linalg.qr(psi)
    psi = np.asfortranarray(psi / math.sqrt(0.3 * 0.25 * 0.35))
    arr = {"svecre": psi, "ssvre": psi.copy(order="F"), "hsvre": np.zeros_like(psi, order="F"),
           "rspsep": np.asfortranarray(rng.standard_normal((9, natom, m))), "sval": np.zeros(m)}
    g, o = both(kk, orc, "rayleighritz_r", arr, **kw)
    np.testing.assert_allclose(g["sval"], o["sval"], rtol=1e-11, atol=1e-10)
    np.testing.assert_allclose(g["svecre"], o["svecre"], atol=1e-8 * np.abs(o["svecre"]).max())
