"""SCF-step benchmark of the kukan Kohn-Sham hot path on one MI355X.

Workload (default): periodic 128^3 cell, h = 0.2 bohr, 1024 bands, 128 atoms with 9
projectors each (the synthetic recipe of rspacedist_amd/synth.py).  One SCF step is

  rho = sum_n psi_n^2 -> Hartree CG to 1e-10 (warm-started) -> scf_diag: Rayleigh-Ritz,
  4 block-CG iterations, Cholesky-QR, reorder

with every array resident in HBM before the timed region.  Seeded bands are
orthonormalised and advanced by --warmup FP64 steps (untimed; the default 6 leaves a mid-SCF
state, largest band residual about 0.4, instead of the random start); from that same state
the step is timed --steps times in FP64 and in split precision (--precision both, the
default), and the largest eigenvalue difference between the two runs is reported.  `value`
is the split-precision time per step.  Prints one JSON line.
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


def setup(cfg, kk, synth):
    c = synth.CONFIGS[cfg]
    s = synth.make_system(c["n"], c["h"], c["natom"], c["nbands"], seed=cfg)
    n, m = s.npts, s.nbands
    D = kk.DeviceArray
    dev = dict(veff=D.from_numpy(s.veff), vnlocp=D.from_numpy(s.vnlocp),
               lstvec2=D.from_numpy(s.lstvec2), dij=D.from_numpy(s.dij))
    st = dict(svecre=D((n, m)), hsvre=D((n, m)), rspsep=D((9, s.natom, m)),
              sval=D((m,)), residual_states=D((m,)), rho=D((n,)), vh=D((n,)),
              fnele=D.from_numpy(np.ones(m)))
    return s, dev, st


def init_state(s, dev, st, kk, synth, seed):
    st["svecre"].fill_normal(seed)
    kw = synth.abi_args(s, **dev)
    kk.call("orthogonalization_r", **kw, nflag=1, svecre=st["svecre"], ssvre=st["svecre"],
            rspsep=st["rspsep"])
    zero = np.zeros(s.npts)
    kk._ck(kk.lib().kukan_memcpy(st["vh"].device_ptr, zero.ctypes.data, zero.nbytes, 1), "memcpy")
    kk.synchronize()


def scf_step(s, dev, st, kk, synth, iscf, nsdmax=4):
    z = np.zeros(1)
    kk.call("scf_charge_smt_r", neigmx=s.nbands, nums=1, nspv=1, numk=1, ncpx=s.n, ncpy=s.n,
            ncpz=s.n, fnele=st["fnele"], svecre=st["svecre"], rhosmt=st["rho"])
    kk.call("scf_hartree", nd=1, ncpx_d=s.n, ncpy_d=s.n, ncpz_d=s.n, nfh=4, ndisp=6, nperi=3,
            ncgres=801, ncgmin=1, ncgmax=800, xmax=s.xmax, ymax=s.xmax, zmax=s.xmax, epsvh=1e-10,
            rho_aug_dense=st["rho"], boundx=z, boundy=z, boundz=z, vh_dense=st["vh"])
    hart_it = int(kk.get_param("hartree_iterations"))
    kw = synth.abi_args(s, iscf=iscf, nsdmax=nsdmax, epssd=1e-12, **dev)
    r = kk.call("scf_diag", **kw, svecre=st["svecre"], ssvre=st["svecre"], hsvre=st["hsvre"],
                rspsep=st["rspsep"], sval=st["sval"], residual_states=st["residual_states"])
    return hart_it, r


def snapshot(s, st, kk):
    """Device copies of the state one SCF step reads (bands, <p|psi>, eigenvalues, V_H)."""
    keys = ("svecre", "rspsep", "sval", "residual_states", "vh")
    snap = {}
    for k in keys:
        d = kk.DeviceArray(st[k].shape)
        kk._ck(kk.lib().kukan_memcpy(d.device_ptr, st[k].device_ptr, st[k].nbytes, 3), "memcpy")
        snap[k] = d
    kk.synchronize()
    return snap


def restore(snap, st, kk):
    for k, d in snap.items():
        kk._ck(kk.lib().kukan_memcpy(st[k].device_ptr, d.device_ptr, d.nbytes, 3), "memcpy")
    kk.synchronize()


def run_mode(s, dev, st, kk, synth, precision, steps, it0):
    """Time `steps` SCF steps from the current state in the given precision."""
    kk.set_param("kscg_mode", 1)
    kk.set_param("precision", 1 if precision == "mixed" else 0)
    kk.synchronize()
    kk.timing_reset()
    kk.set_param("timing", 1)
    hart = []
    it = it0
    t0 = time.perf_counter()
    for _ in range(steps):
        it += 1
        h, _ = scf_step(s, dev, st, kk, synth, it)
        hart.append(h)
    kk.synchronize()
    dt = (time.perf_counter() - t0) / steps
    kk.set_param("timing", 0)
    tim = kk.timing()
    kk.set_param("precision", 0)
    return dt, tim, hart, st["sval"].numpy()


def forces(s, dev, st, kk, synth):
    """force_eig_r (FP64) of the current bands: (natom, 3) Ha/bohr."""
    kw = synth.abi_args(s, **dev)
    fx, fy, fz = np.zeros(s.natom), np.zeros(s.natom), np.zeros(s.natom)
    kk.call("force_eig_r", **kw, svecre=st["svecre"], sval=st["sval"], fnele=st["fnele"], fatx=fx,
            faty=fy, fatz=fz)
    return np.stack([fx, fy, fz], axis=1)


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


def roofline_entry(tim, tag, bound):
    ms, cnt, work = tim.get(tag, (0.0, 0, 0.0))
    if cnt == 0 or ms == 0:
        return None
    per_launch_s = ms / cnt / 1e3
    if bound == "hbm":
        ach = work / cnt / per_launch_s / 1e9
        return {"bound": "hbm", "achieved": round(ach, 1), "peak": PEAK_HBM_GBS, "unit": "GB/s",
                "frac": round(ach / PEAK_HBM_GBS, 4), "traffic": measured_traffic(tag), "kernel": tag,
                "avg_launch_ms": round(ms / cnt, 4), "launches": cnt}
    ach = work / cnt / per_launch_s / 1e12
    peak = PEAK_F16_TFS if tag.startswith("gemm_mixed") else PEAK_FP64_TFS
    return {"bound": "mfma", "achieved": round(ach, 2), "peak": peak, "unit": "TFLOP/s",
            "frac": round(ach / peak, 4), "traffic": measured_traffic(tag), "kernel": tag,
            "avg_launch_ms": round(ms / cnt, 4), "launches": cnt}


def cpu_baseline(cfg, budget_s=30.0):
    """The FP64 oracle (a C/OpenMP restatement of the reference's band-sequential path) on
    this host.  At a fixed atom density the band solve costs a*M*N (stencil, projectors, dot
    products) + b*M^2*N (Gram-Schmidt against lower bands, the RR products and rotation) + the
    M^3 dense eigensolve, so it is measured on a grid of 1/8 the volume (same spacing, 1/8 of
    the atoms) at 32, 64 and 128 bands, fitted as T = a M + b M^2, scaled by the volume ratio
    and evaluated at the full band count; the oracle's eigensolver is timed at half the band
    count and scaled by 8.  The Hartree solve is timed on the full grid for a fixed number of
    CG iterations and scaled to the GPU's iteration count by the caller."""
    threads = min(16, os.cpu_count() or 1)
    os.environ.setdefault("OMP_NUM_THREADS", str(threads))
    from oracle import oracle
    from rspacedist_amd import synth
    lib = oracle.lib()
    threads = lib.orc_num_threads()
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
    # T(M) = a M + b M^2 by least squares over the samples (no constant term)
    ms_ = np.array(sorted(times), dtype=float)
    tt = np.array([times[int(m)] for m in ms_])
    if len(ms_) >= 2:
        (a, b), *_ = np.linalg.lstsq(np.stack([ms_, ms_ ** 2], axis=1), tt, rcond=None)
    else:
        a, b = tt[0] / ms_[0], 0.0
    a, b = max(float(a), 0.0), max(float(b), 0.0)
    # the RR eigensolve at the full band count: the oracle's tred2/tql2 at M/2, times 8
    mh = max(2, M // 2)
    A = np.random.default_rng(1).standard_normal((mh, mh))
    t0 = time.perf_counter()
    oracle.dsyev(A + A.T)
    eig_s = 8.0 * (time.perf_counter() - t0)
    diag = vol * (a * M + b * M * M) + eig_s
    # Hartree: fixed 10 CG iterations on the full grid
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
    out = dict(threads=threads, diag_s=diag, samples={str(k): round(v, 3) for k, v in times.items()},
               hartree_per_iter_s=hart_per_it, a=a, b=b, n_s=n_s, natom_s=natom_s, vol=vol,
               eig_s=eig_s, mh=mh)
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=6)
    ap.add_argument("--config", type=int, default=2)
    ap.add_argument("--precision", choices=["mixed", "fp64", "both"], default="both")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one MI355X only (this build has no multi-GPU path); got --gpus {args.gpus}",
              file=sys.stderr)
        sys.exit(2)
    from rspacedist_amd import kukan as kk, synth
    kk.init(0)
    log(f"[bench] config {args.config}: building synthetic system")
    s, dev, st = setup(args.config, kk, synth)
    log(f"[bench] n={s.n}^3 bands={s.nbands} atoms={s.natom} num_list={s.num_list}")
    modes = ["fp64", "mixed"] if args.precision == "both" else [args.precision]
    # shared start: seeded bands, orthonormalised, then `warmup` FP64 steps (untimed); every
    # precision is timed from that same state, so their eigenvalues can be compared
    init_state(s, dev, st, kk, synth, seed=1234)
    kk.set_param("kscg_mode", 1)
    kk.set_param("precision", 0)
    log(f"[bench] warmup: {args.warmup} FP64 steps")
    cold = None
    for it in range(1, args.warmup + 1):
        if it == 1:  # the first Hartree solve starts from V_H = 0
            kk.timing_reset()
            kk.set_param("timing", 1)
        h, _ = scf_step(s, dev, st, kk, synth, it)
        if it == 1:
            kk.set_param("timing", 0)
            cold = {"iterations": h, "ms": round(kk.timing().get("phase_hartree", (0.0,))[0], 2)}
    kk.synchronize()
    snap = snapshot(s, st, kk)  # (also the known-traffic copy kernel counter runs calibrate on)
    res = {}
    for i, mode in enumerate(modes):
        if i > 0:
            restore(snap, st, kk)
        log(f"[bench] {mode}: timed {args.steps} steps")
        dt, tim, hart, sval = run_mode(s, dev, st, kk, synth, mode, args.steps, args.warmup)
        res[mode] = dict(dt=dt, tim=tim, hart=hart, sval=sval,
                         force=forces(s, dev, st, kk, synth) if args.precision == "both" else None)
        log(f"[bench] {mode}: {dt:.4f} s/step, hartree its {hart}")
        for k, v in sorted(tim.items(), key=lambda kv: -kv[1][0]):
            if not k.startswith("phase_"):
                log(f"    {k:16s} {v[0] / args.steps:9.2f} ms/step  {v[1] // args.steps:6d} launches/step")
        log("    phases: " + ", ".join(f"{k[6:]} {v[0] / args.steps:.1f} ms" for k, v in tim.items()
                                       if k.startswith("phase_")))
    if snap:
        for d in snap.values():
            d.free()
    head = "mixed" if "mixed" in res else "fp64"
    tim = res[head]["tim"]
    eig_err = force_err = None
    if "mixed" in res and "fp64" in res:
        eig_err = float(np.abs(res["mixed"]["sval"] - res["fp64"]["sval"]).max())
        force_err = float(np.abs(res["mixed"]["force"] - res["fp64"]["force"]).max())
    phases = {k[6:]: round(v[0] / args.steps, 2) for k, v in tim.items() if k.startswith("phase_")}
    kern = {k: v for k, v in tim.items() if not k.startswith("phase_")}
    dominant = max(kern.items(), key=lambda kv: kv[1][0])[0] if kern else None
    bound = {"gemm_f64": "mfma", "gemm_mixed_tn": "mfma", "gemm_mixed_nn": "mfma", "trsm_f64": "mfma", "syevd": "mfma",
             "potrf": "mfma"}.get(dominant, "hbm")
    roof = roofline_entry(tim, dominant, bound) if dominant else None
    extra = {t: roofline_entry(tim, t, "hbm") for t in ("band_update_dots", "p_update_dots", "hpsi_stencil",
                                                        "hp_comb_stencil", "precond_resid", "update_precond", "dots", "axpby", "nl_gather", "nl_expand",
                                                        "gemm_scale")
             if t in tim and t != dominant}
    for t in ("gemm_f64", "gemm_mixed_tn", "gemm_mixed_nn"):
        if t in tim and t != dominant:
            extra[t] = roofline_entry(tim, t, "mfma")
    cpu = None
    if not args.no_cpu_baseline:
        log("[bench] CPU baseline (oracle on host cores)")
        cb = cpu_baseline(args.config)
        hits = float(np.mean(res[head]["hart"])) if res[head]["hart"] else 1.0
        val = cb["diag_s"] + cb["hartree_per_iter_s"] * hits
        cpu = {"value": round(val, 2), "unit": "s/SCF-step", "cores": cb["threads"], "kind": "port",
               "sample": (f"oracle scf_diag (band-sequential CG, FP64) on a {cb['n_s']}^3 grid "
                          f"({cb['natom_s']} atoms, 1/{cb['vol']:.0f} of the volume) with "
                          f"{', '.join(cb['samples'])} bands ({cb['samples']} s), fitted a*M+b*M^2, "
                          f"scaled by {cb['vol']:.0f} and evaluated at M={s.nbands}; plus the oracle "
                          f"eigensolver at M={cb['mh']} x 8 ({cb['eig_s']:.2f} s); plus 10 oracle "
                          f"Hartree CG iterations on the full grid ({cb['hartree_per_iter_s']:.3f} s/it) "
                          f"x {hits:.0f} GPU iterations")}
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
                               "Hartree CG to 1e-10 + RR + 4 block-CG + Cholesky-QR + reorder",
                   "precision": head},
        "fp64_s_per_step": round(res["fp64"]["dt"], 4) if "fp64" in res else None,
        "mixed_s_per_step": round(res["mixed"]["dt"], 4) if "mixed" in res else None,
        "eig_err_mixed_vs_fp64_Ha": eig_err,
        "force_err_mixed_vs_fp64_Ha_per_bohr": force_err,
        "phases_ms_per_step": phases,
        "hartree_cold_first_step": cold,
        "roofline": roof,
        "kernels": extra,
        "cpu_baseline": cpu,
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
