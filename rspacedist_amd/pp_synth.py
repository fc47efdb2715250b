"""Synthetic inputs for the projector-table, list-vector and force routines (pseudocalc8f.F90,
force8f.F90): a small cell with a few atoms of two or three species, radial grids, Bessel
coefficient tables and the dense-grid fields force() reads.  Shared by the oracle, reference and
GPU tests; arrays are Fortran-ordered as the routines take them."""
from __future__ import annotations

import numpy as np

KEY_IN, KEY_INPS, KEY_OUT = 1, 2, 0
NPRJ_CASES = (1, 2, 4, 5, 7, 8, 9, 10, 12, 13, 14, 15, 17, 18)


def f(a, dtype=None):
    return np.asfortranarray(a if dtype is None else np.asarray(a, dtype=dtype))


def cell(seed=0, nperi=3, nxmax=6, nymax=5, nzmax=5, h=0.45, nmesh=2, nfdg=3, npmesh=2, natom=3,
         nprj=(9, 4), rc=(1.05, 0.85), psctoff=1.3, npx=None, nqmx=24, nradmx=12, edge=False):
    """The list-vector inputs of pseudocalc_01b plus the species tables pseudocalc_07 and force
    read.  ``edge`` puts the first atom on a cell face (its sphere wraps, or is cut when the
    axis is isolated)."""
    rng = np.random.default_rng(seed)
    num_spe = len(nprj)
    dx = dy = dz = h
    xmax, ymax, zmax = nxmax * dx, nymax * dy, nzmax * dz
    spe = rng.integers(1, num_spe + 1, natom).astype(np.int32)
    spe[:num_spe] = np.arange(1, num_spe + 1)
    pos = np.stack([rng.uniform(-m, m, natom) for m in (xmax, ymax, zmax)])
    if edge:
        pos[:, 0] = [-xmax + 0.3 * dx, ymax - 0.2 * dy, 0.1]
    nradct = np.full(num_spe, nradmx - 2, dtype=np.int32)
    radial = np.zeros((nradmx, num_spe))
    for s in range(num_spe):
        radial[:, s] = rc[s] * (np.arange(1, nradmx + 1) / nradct[s]) ** 1.5
    rmax = max(rc) * psctoff
    if npx is None:
        npx = int(np.ceil(rmax / h)) + 1
    d = dict(nperi=nperi, natom=natom, num_spe=num_spe, nradmx=nradmx, nmesh=nmesh, nfdg=nfdg, ndisp=6,
             num_atcell=natom, num_ppcell=natom, num_ppcell_d=natom, num_list=(2 * npx) ** 3,
             num_list_d=(2 * npx * nmesh) ** 3, nxmax=nxmax, nymax=nymax, nzmax=nzmax, ncpx=2 * nxmax,
             ncpy=2 * nymax, ncpz=2 * nzmax, npxmax=npx, npymax=npx, npzmax=npx, key_natpri_in=KEY_IN,
             key_natpri_inps=KEY_INPS, key_natpri_out=KEY_OUT, psctoff=psctoff, xmax=xmax, ymax=ymax,
             zmax=zmax, dx=dx, dy=dy, dz=dz, indspe=f(spe), nradct=f(nradct), radial=f(radial),
             atx=f(pos[0]), aty=f(pos[1]), atz=f(pos[2]))
    # species tables
    coef = np.zeros((nqmx + 1, 8, num_spe))
    for s in range(num_spe):
        iq = np.arange(nqmx + 1)
        for c in range(8):
            coef[:, c, s] = rng.standard_normal(nqmx + 1) / (1.0 + 0.3 * iq) ** 2
    cp = np.zeros((8, num_spe))
    cp[0] = rng.uniform(1.0, 4.0, num_spe)
    cp[1] = rng.uniform(0.8, 2.0, num_spe)
    cp[2] = rng.uniform(2.0, 5.0, num_spe)
    cp[3] = rng.uniform(0.3, 0.7, num_spe)
    cp[4] = 1.0 - cp[3]
    extra = dict(npmesh=npmesh, nprj=f(np.asarray(nprj, dtype=np.int32)), nprjmx=max(nprj), nqmx=nqmx,
                 nqct=f(np.full(num_spe, nqmx - 3, dtype=np.int32)),
                 nqctpcc=f(np.full(num_spe, nqmx - 5, dtype=np.int32)), coef=f(coef), cp=f(cp),
                 psftrad=0.9, eps=1.0e-8, filpp=0.1)
    return d, extra


def list_outputs(d, fill=-7):
    """Output arrays of pseudocalc_01b, prefilled so untouched entries are visible."""
    na, nl, nld = d["natom"], d["num_list"], d["num_list_d"]
    npc, npd = d["num_ppcell"], d["num_ppcell_d"]
    o = {k: f(np.full(na, fill, dtype=np.int32)) for k in
         ("natx", "naty", "natz", "ndatx", "ndaty", "ndatz", "natinf", "natinfd", "natinfd_vloc", "natpri",
          "natprid", "natpri_inf", "naps", "napsd")}
    for k in ("lstx", "lsty", "lstz", "lstvec2"):
        o[k] = f(np.full((nl, npc), fill, dtype=np.int32))
    for k in ("lstdx", "lstdy", "lstdz", "lstvecd2"):
        o[k] = f(np.full((nld, npd), fill, dtype=np.int32))
    return o


def spd_dims(d, x):
    """nxspd.. and the work arrays of pseudocalc_07 (pseudocalc8f.F90:227-233)."""
    npm, nfdg = x["npmesh"], d["nfdg"]
    nxs = npm * (d["npxmax"] + nfdg) - (npm + 1) // 2
    nys = npm * (d["npymax"] + nfdg) - (npm + 1) // 2
    nzs = npm * (d["npzmax"] + nfdg) - (npm + 1) // 2
    w = dict(nxspd=nxs, nyspd=nys, nzspd=nzs,
             vtmp_pp=f(np.zeros(8 * d["npxmax"] * d["npymax"] * d["npzmax"])),
             vtmp_pp1=f(np.zeros(8 * nxs * nys * d["npzmax"])),
             vtmp_pp2=f(np.zeros(8 * nxs * d["npymax"] * d["npzmax"])),
             vnlspd=f(np.zeros(8 * nxs * nys * nzs * x["nprjmx"])))
    return w


def pp07_args(d, x, lists, wxyz):
    """Keyword arguments of pseudocalc_07 from the cell, its species tables and its lists."""
    a = {k: d[k] for k in ("natom", "num_spe", "nfdg", "num_list", "num_ppcell", "nradmx", "nxmax", "nymax",
                           "nzmax", "npxmax", "npymax", "npzmax", "key_natpri_in", "key_natpri_inps", "psctoff",
                           "dx", "dy", "dz", "indspe", "nradct", "radial", "atx", "aty", "atz")}
    a.update({k: x[k] for k in ("npmesh", "nprjmx", "nqmx", "eps", "psftrad", "filpp", "nprj", "nqct", "coef")})
    a.update({k: lists[k] for k in ("natpri", "naps", "natinf", "natx", "naty", "natz", "lstx", "lsty", "lstz")})
    a.update(spd_dims(d, x))
    a["wxyz"] = wxyz
    a["nfiltyp"] = 1
    a["vnlocp"] = f(np.full((d["num_list"], x["nprjmx"], d["num_ppcell"]), 7.0))
    return a


def force_args(d, x, lists, seed=1, nspv=2, jel=False, bias=(0.0, 0.0, 0.0), new_pw=2, new_rs=1):
    """Keyword arguments of force() with seeded fields on the coarse and dense grids."""
    rng = np.random.default_rng(seed)
    nm = d["nmesh"]
    n = d["ncpx"] * d["ncpy"] * d["ncpz"]
    nd = n * nm ** 3
    na, nld, npd = d["natom"], d["num_list_d"], d["num_ppcell_d"]
    npoint = 6
    point = rng.standard_normal((npoint, 3))
    point /= np.linalg.norm(point, axis=1)[:, None]
    dradial = np.gradient(d["radial"], axis=0)
    a = {k: d[k] for k in ("natom", "num_spe", "num_atcell", "num_ppcell_d", "num_list_d", "nradmx", "nmesh",
                           "nperi", "ncpx", "ncpy", "ncpz", "key_natpri_in", "key_natpri_inps", "indspe",
                           "nradct", "psctoff", "radial", "xmax", "ymax", "zmax", "atx", "aty", "atz")}
    a.update({k: lists[k] for k in ("natpri", "natpri_inf", "natprid", "napsd", "natinfd", "natinfd_vloc",
                                    "lstvecd2", "lstdx", "lstdy", "lstdz", "ndatx", "ndaty", "ndatz")})
    a.update(nqmx=x["nqmx"], npoint=npoint, nspv=nspv, nint1dmax=40, nzmax=d["nzmax"],
             jelcalc=1 if jel else 0, ncpx_d=d["ncpx"] * nm, ncpy_d=d["ncpy"] * nm, ncpz_d=d["ncpz"] * nm,
             new_pwx=new_pw, new_pwy=new_pw, new_pwz=new_pw, new_rsx=new_rs, new_rsy=new_rs, new_rsz=new_rs,
             key_jel_calc=1, nqctpcc=x["nqctpcc"], veta=0.9, psftrad=x["psftrad"], dradial=f(dradial),
             point=f(point), wt=f(np.full(npoint, 4 * np.pi / npoint)), biasx=bias[0], biasy=bias[1],
             biasz=bias[2], cp=x["cp"], coef=x["coef"], chrjel=0.3, strjel=-0.5 * d["zmax"],
             endjel=0.4 * d["zmax"])
    for c in "xyz":
        a[f"dvlocd{c}_scw"] = f(rng.standard_normal((n, na)))
        a[f"dvlocd{c}_hdp"] = f(rng.standard_normal((nld, npd)))
        a[f"drhoaug3dd{c}"] = f(rng.standard_normal((nld, npd)))
    sh = (d["nradmx"], npoint, nspv, d["num_atcell"])
    a.update(rhotrur=f(rng.random(sh)), rhosmtr=f(rng.random(sh)),
             rhoaugr=f(rng.random((d["nradmx"], npoint, d["num_atcell"]))), rho_coarse=f(rng.random(n)),
             rho_aug_dense=f(rng.random(nd)), vloc_dense=f(rng.standard_normal(nd)),
             vh_dense=f(rng.standard_normal(nd)), vx_dense=f(rng.standard_normal((nd, nspv))))
    a.update(fatx=f(rng.standard_normal(na)), faty=f(rng.standard_normal(na)), fatz=f(rng.standard_normal(na)))
    return a
