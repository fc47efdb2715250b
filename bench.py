"""SCF-step benchmark of the kukan Kohn-Sham hot path on one MI355X.

Workload (default): periodic 128^3 cell, h = 0.2 bohr, 1024 bands, 128 atoms with 9
projectors each (the synthetic recipe of rspacedist_amd/synth.py).  One SCF step is

  rho = sum_n psi_n^2 -> Hartree CG to 1e-10 (warm-started) -> scf_diag: Rayleigh-Ritz,
  4 block-CG iterations, Cholesky-QR (FP64), reorder

with every array resident in HBM before the timed region (rspacedist_amd/workload.py).
Seeded bands are orthonormalised and advanced by --warmup FP64 steps (untimed; a mid-SCF
state instead of the random start).  From that same state:

- one step in each precision: the one-step eigenvalue and total-energy differences;
- --steps timed steps in FP64 and in split precision (--precision both, the default): the
  step times (`value` is the split-precision one), and the eigenvalue, force and total-energy
  differences after those steps.

Beside the headline it reports the other workloads of the suite at their sizes: 40^3 x 32
bands (H psi + a band-sequential CG sweep + Hartree, GPU and CPU oracle), 96^3 x 512 bands in
FP64 (ten RR + CG + Gram-Schmidt iterations and the Hartree solve), the 96^3 x 256-band
energy with its adjoint (forward + backward), the 48-system batch (batched vs per-system loop),
and the projector pipeline at the headline cell (list vectors, vnlocp and forces for 128 atoms).
Prints one JSON line.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

PEAK_HBM_GBS = 8000.0
PEAK_FP64_TFS = 78.6
PEAK_F16_TFS = 2500.0  # dense MFMA fp16


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def run_mode(cell, kk, precision, steps):
    """Time `steps` SCF steps from the current state in the given precision."""
    kk.set_param("kscg_mode", 1)
    kk.set_param("precision", 1 if precision == "mixed" else 0)
    kk.synchronize()
    kk.timing_reset()
    kk.set_param("timing", 1)
    hart = []
    t0 = time.perf_counter()
    for _ in range(steps):
        hart.append(cell.scf_step())
    kk.synchronize()
    dt = (time.perf_counter() - t0) / steps
    kk.set_param("timing", 0)
    tim = kk.timing()
    kk.set_param("precision", 0)
    return dt, tim, hart, cell.st["sval"].numpy()


# timing tag -> kernel whose calibrated HBM bytes per launch the committed counter profile holds
TRAFFIC_KERNEL = {"gemm_mixed_nn": "k_gemm_nn<3>", "gemm_mixed_tn": "k_gemm_tnf<3>",
                  "hpsi_stencil": "k_stencil_lds<4, 2, 3>", "hp_comb_stencil": "k_stencil_lds<4, 1, 5>",
                  "precond_resid": "k_stencil_lds<1, 2, 1>", "update_precond": "k_stencil_lds<1, 1, 6>",
                  "band_update_dots": "k_pair_update_dots", "p_update_dots": "k_pupdate_dots",
                  "dots": "k_dot_part", "gemm_scale": "k_colmax", "nl_expand": "k_expand_rows<16>",
                  "nl_gather": "k_gather<9, 4>"}


def measured_traffic(tag):
    """HBM bytes per launch from the committed rocprofv3 FETCH_SIZE/WRITE_SIZE profile (or None)."""
    path = os.path.join(ROOT, "profiles", "scf_step_128cube_1024bands_hbm_traffic.json")
    k = TRAFFIC_KERNEL.get(tag)
    if not k or not os.path.exists(path):
        return None
    with open(path) as fh:
        ent = json.load(fh)["kernels"].get(k)
    return ent["hbm_bytes_per_launch"] if ent else None


MFMA_TAGS = {"gemm_f64": PEAK_FP64_TFS, "syrk_f64": PEAK_FP64_TFS, "trsm_f64": PEAK_FP64_TFS,
             "syevd": PEAK_FP64_TFS, "potrf": PEAK_FP64_TFS, "trtri": PEAK_FP64_TFS,
             "gemm_mixed_tn": PEAK_F16_TFS, "gemm_mixed_nn": PEAK_F16_TFS}


def roofline_entry(tim, tag):
    """Achieved vs peak for one timing tag from its algorithmic work (HBM bytes, or flops of
    the full product for the subspace GEMMs: 2 N M^2 whatever triangle or split is computed).
    For a split-precision product, mfma_pass_frac is the MFMA pipe's share of peak on the
    f16 flops it actually issued (three passes per product over the tiles computed)."""
    ms, cnt, work = tim.get(tag, (0.0, 0, 0.0))
    if cnt == 0 or ms == 0:
        return None
    per_launch_s = ms / cnt / 1e3
    if tag not in MFMA_TAGS:
        ach = work / cnt / per_launch_s / 1e9
        return {"bound": "hbm", "achieved": round(ach, 1), "peak": PEAK_HBM_GBS, "unit": "GB/s",
                "frac": round(ach / PEAK_HBM_GBS, 4), "traffic": measured_traffic(tag), "kernel": tag,
                "avg_launch_ms": round(ms / cnt, 4), "launches": cnt}
    peak = MFMA_TAGS[tag]
    ach = work / cnt / per_launch_s / 1e12
    out = {"bound": "mfma", "achieved": round(ach, 2), "peak": peak, "unit": "TFLOP/s",
           "frac": round(ach / peak, 4), "traffic": measured_traffic(tag), "kernel": tag,
           "avg_launch_ms": round(ms / cnt, 4), "launches": cnt}
    if tag + ".mfma" in tim:
        issued = tim[tag + ".mfma"][2]
        out["mfma_pass_frac"] = round(issued / (ms / 1e3) / 1e12 / peak, 4)
    return out


def host_info():
    model = ""
    try:
        with open("/proc/cpuinfo") as fh:
            for line in fh:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    try:
        aff = len(os.sched_getaffinity(0))
    except AttributeError:
        aff = os.cpu_count()
    return {"cpu_model": model, "nproc": aff, "cpu_count": os.cpu_count()}


def oracle_threads():
    # the host's thread budget for this process (the GPU boxes set OMP_NUM_THREADS to the CPU
    # share of one GPU); otherwise every core this process may run on
    os.environ.setdefault("OMP_NUM_THREADS", str(host_info()["nproc"] or 1))
    from oracle import oracle
    return oracle.lib().orc_num_threads()


def config0(kk, synth):
    """40^3 x 32 bands x 8 atoms: H psi of all bands (dV <psi|H psi>), one band-sequential
    kscg_r sweep (nsdmax = 4) and the Hartree CG solve of rho = sum psi^2, on the GPU and on the
    oracle with the same inputs (seconds per workload, and the eigenvalue agreement)."""
    from oracle import oracle
    s = synth.make_config(0)
    psi = synth.smooth_bands(synth.random_bands(s), s.n, s.h)
    m = s.nbands
    base = dict(svecre=psi, ssvre=np.zeros_like(psi, order="F"), rspsep=np.zeros((9, s.natom, m), order="F"),
                hsvre=np.zeros_like(psi, order="F"), sval=np.zeros(m), residual_states=np.zeros(m))
    oracle.call("orthogonalization_r", **synth.abi_args(s), nflag=1, svecre=base["svecre"],
                ssvre=base["ssvre"], rspsep=base["rspsep"])
    out = {}
    kk.set_param("kscg_mode", 0)
    for who, call in (("gpu", kk.call), ("cpu", oracle.call)):
        st = {k: v.copy(order="F") for k, v in base.items()}
        eb = np.zeros(m)
        rho = np.zeros(s.npts)
        vh = np.zeros(s.npts)
        z = np.zeros(1)
        if who == "gpu":  # the GPU side: all arrays resident, as in the headline
            st = {k: kk.DeviceArray.from_numpy(v) for k, v in st.items()}
            eb, rho, vh = (kk.DeviceArray.from_numpy(x) for x in (eb, rho, vh))
            kk.synchronize()
        t0 = time.perf_counter()
        call("band_energy_r", **synth.abi_args(s), svecre=st["svecre"], eband=eb)
        r = call("kscg_r", **synth.abi_args(s), **st)
        call("scf_charge_smt_r", neigmx=m, nums=1, nspv=1, numk=1, ncpx=s.n, ncpy=s.n, ncpz=s.n,
             fnele=np.ones(m), svecre=st["svecre"], rhosmt=rho)
        call("scf_hartree", nd=1, ncpx_d=s.n, ncpy_d=s.n, ncpz_d=s.n, nfh=4, ndisp=6, nperi=3,
             ncgres=801, ncgmin=1, ncgmax=800, xmax=s.xmax, ymax=s.xmax, zmax=s.xmax, epsvh=1e-10,
             rho_aug_dense=rho, boundx=z, boundy=z, boundz=z, vh_dense=vh)
        if who == "gpu":
            kk.synchronize()
        out[who] = time.perf_counter() - t0
        out[who + "_sval"] = st["sval"].numpy() if who == "gpu" else st["sval"]
        out[who + "_ksitmax"] = r["ksitmax"]
    return {"workload": "40^3 h=0.25, 32 bands, 8 atoms: H psi + band-sequential CG sweep (nsdmax 4) + Hartree CG to 1e-10",
            "gpu_s": round(out["gpu"], 4), "cpu_s": round(out["cpu"], 3),
            "eig_diff_gpu_vs_oracle_Ha": float(np.abs(out["gpu_sval"] - out["cpu_sval"]).max())}


def config1(kk):
    """96^3 x 512 bands x 64 atoms in FP64 on the GPU: ten scf_diag calls, each a 512 x 512
    Rayleigh-Ritz, one block-CG iteration and Gram-Schmidt (Cholesky-QR), then the Hartree CG of
    rho to 1e-10 (seconds for the whole workload)."""
    from rspacedist_amd import workload
    kk.set_param("kscg_mode", 1)
    kk.set_param("precision", 0)
    c = workload.Cell(cfg=1)
    try:
        c.init_bands(seed=4321)
        c.diag(nsdmax=1)  # first call: library buffers, rocSOLVER workspaces
        c.init_bands(seed=4321)
        kk.timing_reset()
        kk.set_param("timing", 1)
        t0 = time.perf_counter()
        for _ in range(10):
            c.diag(nsdmax=1)
        its = c.charge_hartree(1e-10)
        kk.synchronize()
        dt = time.perf_counter() - t0
        kk.set_param("timing", 0)
        tim = kk.timing()
        res = float(c.st["residual_states"].numpy().sum())
        return {"workload": "96^3 h=0.2, 512 bands, 64 atoms, FP64: 10 x (RR 512x512 + 1 CG iteration + Cholesky-QR) + Hartree CG to 1e-10",
                "gpu_s": round(dt, 4), "hartree_iterations": its, "sum_residual_final": res,
                "phases_ms": {k[6:]: round(v[0], 2) for k, v in tim.items() if k.startswith("phase_")}}
    finally:
        c.free()


def config3(kk):
    """96^3 x 256 bands x 32 atoms: E = sum <psi|H|psi> + E_H forward and the adjoint (dE/dV_loc
    and dE/dR for all 96 coordinates) through the torch.autograd.Function."""
    import torch
    from rspacedist_amd import autograd as AG, synth
    c = synth.CONFIGS[3]
    s = synth.make_system(c["n"], c["h"], c["natom"], c["nbands"], seed=3)
    D = kk.DeviceArray((s.npts, s.nbands)).fill_normal(33)
    kk.call("orthogonalization_r", **synth.abi_args(s), nflag=1, svecre=D, ssvre=D,
            rspsep=np.zeros((9, s.natom, s.nbands), order="F"))
    st = AG.BandState(s, D)
    st.hartree()
    R = torch.tensor(s.atoms, dtype=torch.float64, requires_grad=True)
    AG.total_energy(R, st).backward()  # first call: projector plans and buffers
    R.grad = None
    kk.synchronize()
    kk.timing_reset()
    kk.set_param("timing", 1)
    t0 = time.perf_counter()
    E = AG.total_energy(R, st)
    t1 = time.perf_counter()
    E.backward()
    kk.synchronize()
    t2 = time.perf_counter()
    kk.set_param("timing", 0)
    tim = kk.timing()
    D.free()
    kern = {k: v for k, v in tim.items() if not k.startswith("phase_") and v[1] > 0}
    return {"workload": "96^3 h=0.2, 256 bands, 32 atoms: E = sum <psi|H|psi> + E_H and dE/dV_loc, dE/dR by torch.autograd",
            "forward_s": round(t1 - t0, 4), "backward_s": round(t2 - t1, 4), "energy_Ha": float(E),
            "kernels": {k: roofline_entry(tim, k) for k in kern if kern[k][2] > 0}}


def config4(kk):
    """48 systems ({32,48,64,80}^3, 16..128 bands, 2..24 atoms): one batched step (persistent
    stencil launches + graph replays) against the per-system launch loop."""
    from rspacedist_amd import batch as BT, synth
    cfgs = BT.batch_configs(48, seed=4)
    bb, bl = BT.Batch(), BT.Batch()
    pts = bands = 0
    for cf in cfgs:
        s = synth.make_system(cf["n"], BT.H, cf["natom"], cf["nbands"], seed=cf["seed"])
        D = BT.make_bands(s)
        bb.add(s, D)
        bl.add(s, D)
        D.free()
        pts += s.npts * s.nbands
        bands += s.nbands
    bb.step(batched=True)  # first step: autotune + graph capture
    bl.step(batched=False)
    kk.synchronize()
    reps = 3
    kk.timing_reset()
    kk.set_param("timing", 1)
    t0 = time.perf_counter()
    for _ in range(reps):
        bb.step(batched=True)
    kk.synchronize()
    t1 = time.perf_counter()
    kk.set_param("timing", 0)
    tim = kk.timing()
    for _ in range(reps):
        bl.step(batched=False)
    kk.synchronize()
    t2 = time.perf_counter()
    tb, tl = (t1 - t0) / reps, (t2 - t1) / reps
    bb.close()
    bl.close()
    # lower bound on the step's HBM fraction: only H psi's algorithmic bytes (psi read and
    # H psi written once, 16 bytes per point-band) over the whole batched step time
    hbm = 16.0 * pts / tb / 1e9
    return {"workload": f"48 systems {{32,48,64,80}}^3 h=0.25, {bands} bands, {pts / 1e6:.0f} M point-bands: H psi + band update + FP64 Cholesky-QR orthonormalisation + Hartree CG per system",
            "batched_ms_per_step": round(tb * 1e3, 2), "loop_ms_per_step": round(tl * 1e3, 2),
            "systems_per_s": round(48 / tb, 1), "speedup_vs_loop": round(tl / tb, 2),
            "hbm_frac_lower_bound": round(hbm / PEAK_HBM_GBS, 4),
            "tile_p": kk.get_param("batch_tile_p"), "grid": kk.get_param("batch_grid"),
            "device_ms_per_step": {t: round(v[0] / reps, 2) for t, v in tim.items() if v[1] > 0}}


def config_pp(kk, cpu=True):
    """The projector pipeline at the headline cell: 128^3 h=0.2, 128 atoms of two species (9 and
    18 projectors), nfdg 4, npmesh 2, nmesh 2, two spins: list vectors (pseudocalc_01b), vnlocp
    (pseudocalc_07) and force() with every array in HBM.  The CPU oracle is timed on a sample
    (the lists and vnlocp of 4 atoms, the forces of 2 atoms with one spin) and scaled by atoms
    and spins; it is a reference figure, not the target."""
    from rspacedist_amd import pp_synth as PS
    d, x = PS.cell(seed=128, nperi=3, nxmax=64, nymax=64, nzmax=64, h=0.2, nmesh=2, nfdg=4, npmesh=2,
                   natom=128, nprj=(9, 18), rc=(1.0, 0.8), nqmx=24)
    npx = int(np.ceil(d["radial"].max() * d["psctoff"] / d["dx"])) + d["nfdg"] + 1
    d.update(npxmax=npx, npymax=npx, npzmax=npx, num_list=(2 * npx) ** 3, num_list_d=(2 * npx * 2) ** 3)
    dev = lambda a: kk.DeviceArray.from_numpy(a)
    L = {k: dev(v) for k, v in PS.list_outputs(d).items()}
    w = dev(np.zeros(2 * d["nfdg"] * x["npmesh"]))
    kk.call("pseudocalc_01a", nfdg=d["nfdg"], npmesh=x["npmesh"], wxyz=w)
    args01b = {**d, **L}
    n = d["ncpx"] * d["ncpy"] * d["ncpz"]
    nd, na, nspv = n * 8, d["natom"], 2
    host_lists = None

    def run(timed):
        nonlocal host_lists
        t = {}
        kk.synchronize()
        t0 = time.perf_counter()
        kk.call("pseudocalc_01b", **args01b)
        kk.synchronize()
        t["lists_s"] = time.perf_counter() - t0
        if host_lists is None:
            host_lists = {k: v.numpy() for k, v in L.items()}
        a = PS.pp07_args(d, x, host_lists, w)
        a.update({k: L[k] for k in ("natpri", "naps", "natinf", "natx", "naty", "natz", "lstx", "lsty", "lstz")})
        a.update(vnlocp=vnl, vtmp_pp=dummy, vtmp_pp1=dummy, vtmp_pp2=dummy, vnlspd=dummy)
        t0 = time.perf_counter()
        kk.call("pseudocalc_07", **a)
        kk.synchronize()
        t["vnlocp_s"] = time.perf_counter() - t0
        t0 = time.perf_counter()
        kk.call("force", **fargs)
        kk.synchronize()
        t["force_s"] = time.perf_counter() - t0
        return t

    vnl = kk.DeviceArray((d["num_list"], x["nprjmx"], d["num_ppcell"]))
    dummy = np.zeros(1)
    # force() inputs: the big fields seeded on the device; the small tables from the host
    fargs = PS.force_args(dict(d, natom=1, num_atcell=1, num_ppcell_d=1),
                          x, {k: np.zeros(1, dtype=np.int32) for k in ("natpri", "natpri_inf", "natprid", "napsd",
                                                                       "natinfd", "natinfd_vloc", "lstvecd2",
                                                                       "lstdx", "lstdy", "lstdz", "ndatx", "ndaty",
                                                                       "ndatz")}, nspv=nspv, new_pw=3, new_rs=1)
    fargs.update({k: d[k] for k in ("natom", "num_atcell", "num_ppcell_d", "indspe", "atx", "aty", "atz")})
    fargs.update({k: L[k] for k in ("natpri", "natpri_inf", "natprid", "napsd", "natinfd", "natinfd_vloc",
                                    "lstvecd2", "lstdx", "lstdy", "lstdz", "ndatx", "ndaty", "ndatz")})
    nld = d["num_list_d"] * d["num_ppcell_d"]
    big = {}
    for i, (k, shape) in enumerate([(f"dvlocd{c}_scw", (n, na)) for c in "xyz"] +
                                   [(f"dvlocd{c}_hdp", (nld,)) for c in "xyz"] +
                                   [(f"drhoaug3dd{c}", (nld,)) for c in "xyz"] +
                                   [("rho_coarse", (n,)), ("rho_aug_dense", (nd,)), ("vloc_dense", (nd,)),
                                    ("vh_dense", (nd,)), ("vx_dense", (nd, nspv))]):
        big[k] = kk.DeviceArray(shape).fill_normal(900 + i)
    fargs.update(big)
    sh = (d["nradmx"], fargs["npoint"], nspv, na)
    fargs.update(rhotrur=np.zeros(sh, order="F"), rhosmtr=np.zeros(sh, order="F"),
                 rhoaugr=np.zeros((d["nradmx"], fargs["npoint"], na), order="F"))
    for c in ("fatx", "faty", "fatz"):
        fargs[c] = dev(np.zeros(na))
    run(False)  # first call: buffers
    kk.timing_reset()
    kk.set_param("timing", 1)
    t = run(True)
    kk.set_param("timing", 0)
    tim = kk.timing()
    out = {"workload": f"128^3 h=0.2, {na} atoms (9 and 18 projectors), nfdg 4, npmesh 2, nmesh 2, {nspv} spins: "
                       "pseudocalc_01b lists + pseudocalc_07 vnlocp + force()",
           "natinf_mean": float(np.mean(host_lists["natinf"])), "num_list": d["num_list"],
           **{k: round(v, 4) for k, v in t.items()},
           "device_ms": {k: round(v[0], 3) for k, v in tim.items() if v[1] > 0 and not k.startswith("phase_")},
           "force_local": roofline_entry(tim, "force_local")}
    for v in list(big.values()) + [vnl, w] + list(L.values()):
        v.free()
    if cpu:
        from oracle import oracle as orc
        oracle_threads()
        s4 = dict(d, natom=4, num_atcell=4, num_ppcell=4, num_ppcell_d=4, indspe=d["indspe"][:4].copy(),
                  atx=d["atx"][:4].copy(), aty=d["aty"][:4].copy(), atz=d["atz"][:4].copy())
        o = {**s4, **PS.list_outputs(s4)}
        t0 = time.perf_counter()
        orc.call("pseudocalc_01b", **o)
        t1 = time.perf_counter()
        wh = np.zeros(2 * d["nfdg"] * x["npmesh"])
        orc.call("pseudocalc_01a", nfdg=d["nfdg"], npmesh=x["npmesh"], wxyz=wh)
        a = PS.pp07_args(s4, x, o, wh)
        t2 = time.perf_counter()
        orc.call("pseudocalc_07", **a)
        t3 = time.perf_counter()
        s2 = dict(s4, natom=2, num_atcell=2, num_ppcell=2, num_ppcell_d=2, indspe=d["indspe"][:2].copy(),
                  atx=d["atx"][:2].copy(), aty=d["aty"][:2].copy(), atz=d["atz"][:2].copy())
        o2 = {**s2, **PS.list_outputs(s2)}
        orc.call("pseudocalc_01b", **o2)
        fa = PS.force_args(s2, x, o2, nspv=1, new_pw=3, new_rs=1)
        t4 = time.perf_counter()
        orc.call("force", **fa)
        t5 = time.perf_counter()
        cpu_s = (t1 - t0 + t3 - t2) * na / 4 + (t5 - t4) * na / 2 * nspv
        out["cpu_oracle_s"] = round(cpu_s, 2)
        out["cpu_sample"] = (f"oracle (threads {oracle_threads()}) lists + vnlocp of 4 atoms ({t1 - t0 + t3 - t2:.2f} s) "
                             f"x {na // 4}; force of 2 atoms, 1 spin ({t5 - t4:.2f} s) x {na // 2 * nspv}")
    return out


def cpu_baseline(cfg):
    """The FP64 oracle (a C/OpenMP restatement of the reference's band-sequential path) on
    this host, as an extrapolation to the full step (a full step at 128^3 x 1024 bands is about
    ten minutes per CG sweep on the CPU).  At a fixed atom density the band solve costs a*M*N
    (stencil, projectors, dot products) + b*M^2*N (Gram-Schmidt against lower bands, the RR
    products and rotation) + the M^3 dense eigensolve, so it is measured on a grid of 1/8 the
    volume (same spacing, 1/8 of the atoms) at 32, 64 and 128 bands, fitted as T = a M + b M^2,
    scaled by the volume ratio and evaluated at the full band count; the oracle's eigensolver is
    timed at half the band count and scaled by 8.  The Hartree solve is timed on the full grid
    for a fixed number of CG iterations and scaled to the GPU's iteration count by the caller."""
    from oracle import oracle
    from rspacedist_amd import synth
    threads = oracle_threads()
    c = synth.CONFIGS[cfg]
    M = c["nbands"]
    n_s = max(8, c["n"] // 2)
    vol = (c["n"] / n_s) ** 3
    natom_s = max(1, int(round(c["natom"] / vol)))
    times = {}
    for m in sorted({min(M, k) for k in (32, 64, 128)}):
        s = synth.make_system(n_s, c["h"], natom_s, m, seed=cfg)
        psi = synth.random_bands(s, nbands=m)
        st = dict(svecre=psi, ssvre=np.zeros_like(psi, order="F"),
                  rspsep=np.zeros((9, s.natom, m), order="F"), hsvre=np.zeros_like(psi, order="F"),
                  sval=np.zeros(m), residual_states=np.zeros(m))
        kw = synth.abi_args(s, neigmx=m)
        oracle.call("orthogonalization_r", **kw, nflag=1, svecre=st["svecre"], ssvre=st["ssvre"],
                    rspsep=st["rspsep"])
        t0 = time.perf_counter()
        oracle.call("scf_diag", **kw, **st)
        times[m] = time.perf_counter() - t0
    ms_ = np.array(sorted(times), dtype=float)
    tt = np.array([times[int(m)] for m in ms_])
    if len(ms_) >= 2:
        (a, b), *_ = np.linalg.lstsq(np.stack([ms_, ms_ ** 2], axis=1), tt, rcond=None)
    else:
        a, b = tt[0] / ms_[0], 0.0
    a, b = max(float(a), 0.0), max(float(b), 0.0)
    mh = max(2, M // 2)
    A = np.random.default_rng(1).standard_normal((mh, mh))
    t0 = time.perf_counter()
    oracle.dsyev(A + A.T)
    eig_s = 8.0 * (time.perf_counter() - t0)
    diag = vol * (a * M + b * M * M) + eig_s
    n = c["n"]
    xmax = 0.5 * n * c["h"]
    rho = np.asfortranarray(np.random.default_rng(0).random((n, n, n)))
    vh = np.zeros_like(rho)
    z = np.zeros(1)
    t0 = time.perf_counter()
    oracle.call("scf_hartree", nd=1, ncpx_d=n, ncpy_d=n, ncpz_d=n, nfh=4, ndisp=6, nperi=3,
                ncgres=801, ncgmin=10, ncgmax=10, xmax=xmax, ymax=xmax, zmax=xmax, epsvh=0.0,
                rho_aug_dense=rho, boundx=z, boundy=z, boundz=z, vh_dense=vh)
    hart_per_it = (time.perf_counter() - t0) / 10
    return dict(threads=threads, diag_s=diag, samples={str(k): round(v, 3) for k, v in times.items()},
                hartree_per_iter_s=hart_per_it, a=a, b=b, n_s=n_s, natom_s=natom_s, vol=vol,
                eig_s=eig_s, mh=mh)


def cpu_config1_sweep():
    """One scf_diag call with one CG iteration (RR 512 x 512 + band-sequential CG with its
    Gram-Schmidt) of the oracle at 96^3 x 512 bands x 64 atoms; x10 gives the ten-iteration
    workload.  Minutes of CPU time: only with --cpu-config1."""
    from oracle import oracle
    from rspacedist_amd import synth
    threads = oracle_threads()
    s = synth.make_config(1)
    m = s.nbands
    psi = synth.random_bands(s)
    st = dict(svecre=psi, ssvre=np.zeros_like(psi, order="F"), rspsep=np.zeros((9, s.natom, m), order="F"),
              hsvre=np.zeros_like(psi, order="F"), sval=np.zeros(m), residual_states=np.zeros(m))
    kw = synth.abi_args(s, nsdmax=1)
    oracle.call("orthogonalization_r", **kw, nflag=1, svecre=st["svecre"], ssvre=st["ssvre"],
                rspsep=st["rspsep"])
    t0 = time.perf_counter()
    oracle.call("scf_diag", **kw, **st)
    return {"threads": threads, "one_iteration_s": round(time.perf_counter() - t0, 2)}


def product_only(kk, args):
    """One split-FP16 subspace product alone at the headline shape (2,097,152 x 1024 x 1024;
    NN: Y = X C, TN: C = X^T X), --warmup untimed launches then --steps timed ones, with the
    operands resident in HBM.  For profiling the product without the rest of the step."""
    R, M = 128 ** 3, 1024
    X = kk.DeviceArray((R, M)).fill_normal(3)
    C = kk.DeviceArray((M, M)).fill_normal(4)
    Y = kk.DeviceArray((R, M)) if args.product == "nn" else None
    kk.set_param("precision", 1)
    kk.set_param("mixed_min_rows", 0)
    tag = "gemm_mixed_" + args.product

    def launch():
        if args.product == "nn":
            kk.subspace_gemm(0, R, M, M, 1.0, X, R, C, M, 0.0, Y, R)
        else:
            kk.subspace_gemm(1, M, M, R, 1.0, X, R, X, R, 0.0, C, M)

    for _ in range(args.warmup):
        launch()
    kk.synchronize()
    kk.timing_reset()
    kk.set_param("timing", 1)
    for _ in range(args.steps):
        launch()
    kk.synchronize()
    kk.set_param("timing", 0)
    tim = kk.timing()
    ent = roofline_entry(tim, tag)
    print(json.dumps({"metric": f"split-FP16 {args.product.upper()} product, {R} x {M} x {M}",
                      "value": round(ent["avg_launch_ms"], 4), "unit": "ms/product", "n_gpus": 1,
                      "steps": args.steps, "warmup": args.warmup, "higher_is_better": False,
                      "dtype": "f64 via split-f16", "data": "synthetic", "roofline": ent}), flush=True)
    for a in (X, C, Y):
        if a is not None:
            a.free()


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=6)
    ap.add_argument("--config", type=int, default=2)
    ap.add_argument("--precision", choices=["mixed", "fp64", "both"], default="both")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--param", action="append", default=[], metavar="KEY=VALUE",
                    help="set a library parameter (kukan_set_param) before the run")
    ap.add_argument("--no-suite", action="store_true",
                    help="skip the 40^3, 96^3 FP64, adjoint and batch workloads")
    ap.add_argument("--product", choices=["nn", "tn"],
                    help="time one split-FP16 product alone at the headline shape, nothing else")
    ap.add_argument("--cpu-config1", action="store_true",
                    help="also time one oracle iteration at 96^3 x 512 bands (minutes)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one MI355X only (this build has no multi-GPU path); got --gpus {args.gpus}",
              file=sys.stderr)
        sys.exit(2)
    from rspacedist_amd import kukan as kk, synth, workload
    kk.init(0)
    for kv in args.param:
        k, v = kv.split("=", 1)
        kk.set_param(k, float(v))
    if args.product:
        product_only(kk, args)
        return
    log(f"[bench] config {args.config}: building synthetic system")
    cell = workload.Cell(cfg=args.config)
    s = cell.s
    log(f"[bench] n={s.n}^3 bands={s.nbands} atoms={s.natom} num_list={s.num_list}")
    modes = ["fp64", "mixed"] if args.precision == "both" else [args.precision]
    cell.init_bands(seed=1234)
    kk.set_param("kscg_mode", 1)
    kk.set_param("precision", 0)
    log(f"[bench] warmup: {args.warmup} FP64 steps")
    cold = None
    for it in range(1, args.warmup + 1):
        if it == 1:  # the first Hartree solve starts from V_H = 0
            kk.timing_reset()
            kk.set_param("timing", 1)
        h = cell.scf_step()
        if it == 1:
            kk.set_param("timing", 0)
            cold = {"iterations": h, "ms": round(kk.timing().get("phase_hartree", (0.0,))[0], 2)}
    kk.synchronize()
    snap = cell.snapshot()
    # one step in each precision from the same state
    one = {}
    for mode in modes:
        cell.restore(snap)
        kk.set_param("precision", 1 if mode == "mixed" else 0)
        cell.scf_step()
        kk.set_param("precision", 0)
        one[mode] = dict(sval=cell.st["sval"].numpy(), energy=cell.total_energy()[0])
    res = {}
    for mode in modes:
        cell.restore(snap)
        log(f"[bench] {mode}: timed {args.steps} steps")
        dt, tim, hart, sval = run_mode(cell, kk, mode, args.steps)
        res[mode] = dict(dt=dt, tim=tim, hart=hart, sval=sval,
                         force=cell.forces() if args.precision == "both" else None,
                         energy=cell.total_energy())
        log(f"[bench] {mode}: {dt:.4f} s/step, hartree its {hart}")
        for k, v in sorted(tim.items(), key=lambda kv: -kv[1][0]):
            if not k.startswith("phase_") and v[1] > 0:
                log(f"    {k:16s} {v[0] / args.steps:9.2f} ms/step  {v[1] // args.steps:6d} launches/step")
        log("    phases: " + ", ".join(f"{k[6:]} {v[0] / args.steps:.1f} ms" for k, v in tim.items()
                                       if k.startswith("phase_")))
    for k, d in snap.items():
        if k != "_iscf":
            d.free()
    head = "mixed" if "mixed" in res else "fp64"
    tim = res[head]["tim"]
    errs = {}
    if "mixed" in res and "fp64" in res:
        errs = {
            "eig_err_mixed_vs_fp64_Ha": float(np.abs(res["mixed"]["sval"] - res["fp64"]["sval"]).max()),
            "force_err_mixed_vs_fp64_Ha_per_bohr": float(np.abs(res["mixed"]["force"] - res["fp64"]["force"]).max()),
            "total_energy_err_mixed_vs_fp64_Ha": abs(res["mixed"]["energy"][0] - res["fp64"]["energy"][0]),
            "one_step_eig_err_Ha": float(np.abs(one["mixed"]["sval"] - one["fp64"]["sval"]).max()),
            "one_step_total_energy_err_Ha": abs(one["mixed"]["energy"] - one["fp64"]["energy"]),
        }
    total_energy = {m: {"E_Ha": r["energy"][0], "E_band_Ha": r["energy"][1], "E_H_Ha": r["energy"][2]}
                    for m, r in res.items()}
    phases = {k[6:]: round(v[0] / args.steps, 2) for k, v in tim.items() if k.startswith("phase_")}
    kern = {k: v for k, v in tim.items() if not k.startswith("phase_") and v[1] > 0}
    dominant = max(kern.items(), key=lambda kv: kv[1][0])[0] if kern else None
    roof = roofline_entry(tim, dominant) if dominant else None
    extra = {t: roofline_entry(tim, t) for t in kern if t != dominant}
    cell.free()
    suite = {}
    if not args.no_suite:
        for name, fn in (("c0_40cube_32bands", lambda: config0(kk, synth)),
                         ("c1_96cube_512bands_fp64", lambda: config1(kk)),
                         ("c3_96cube_256bands_adjoint", lambda: config3(kk)),
                         ("c4_48_systems_batch", lambda: config4(kk)),
                         ("pp_128cube_128atoms", lambda: config_pp(kk, cpu=not args.no_cpu_baseline))):
            log(f"[bench] {name}")
            t0 = time.perf_counter()
            suite[name] = fn()
            log(f"[bench] {name}: {json.dumps(suite[name])[:400]} ({time.perf_counter() - t0:.1f} s)")
    cpu = None
    if not args.no_cpu_baseline:
        log("[bench] CPU baseline (oracle on host cores)")
        cb = cpu_baseline(args.config)
        hits = float(np.mean(res[head]["hart"])) if res[head]["hart"] else 1.0
        val = cb["diag_s"] + cb["hartree_per_iter_s"] * hits
        cpu = {"value": round(val, 2), "unit": "s/SCF-step", "cores": cb["threads"], "kind": "port",
               "extrapolated": True,
               "sample": (f"oracle scf_diag (band-sequential CG, FP64) on a {cb['n_s']}^3 grid "
                          f"({cb['natom_s']} atoms, 1/{cb['vol']:.0f} of the volume) with "
                          f"{', '.join(cb['samples'])} bands ({cb['samples']} s), fitted a*M+b*M^2, "
                          f"scaled by {cb['vol']:.0f} and evaluated at M={s.nbands}; plus the oracle "
                          f"eigensolver at M={cb['mh']} x 8 ({cb['eig_s']:.2f} s); plus 10 oracle "
                          f"Hartree CG iterations on the full grid ({cb['hartree_per_iter_s']:.3f} s/it) "
                          f"x {hits:.0f} GPU iterations"),
               "host": host_info()}
        if "c0_40cube_32bands" in suite:
            cpu["measured_40cube_32bands_s"] = suite["c0_40cube_32bands"]["cpu_s"]
        if args.cpu_config1:
            log("[bench] CPU: one oracle iteration at 96^3 x 512 bands")
            cpu["config1"] = cpu_config1_sweep()
    line = {
        "metric": "SCF-step s at 128^3x1024 bands: mixed vs FP64, eigval err <=1e-6 Ha",
        "value": round(res[head]["dt"], 4),
        "unit": "s/SCF-step",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(res[head]["dt"] * 1e3, 2),
        "higher_is_better": False,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64" if head == "fp64" else "f64+split-f16",
        "data": "synthetic",
        "config": {"workload": f"SCF step, {s.n}^3 grid h={s.h}, {s.nbands} bands, {s.natom} atoms x 9 projectors, "
                               "Hartree CG to 1e-10 + RR + 4 block-CG + Cholesky-QR (FP64) + reorder",
                   "precision": head},
        "fp64_s_per_step": round(res["fp64"]["dt"], 4) if "fp64" in res else None,
        "mixed_s_per_step": round(res["mixed"]["dt"], 4) if "mixed" in res else None,
        **errs,
        "total_energy": total_energy,
        "phases_ms_per_step": phases,
        "hartree_cold_first_step": cold,
        "roofline": roof,
        "kernels": extra,
        "suite": suite,
        "cpu_baseline": cpu,
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
