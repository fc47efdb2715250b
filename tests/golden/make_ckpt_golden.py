"""Writes tests/golden/ckpt/: checkpoint files written by the reference's own write_wf and
write_rhovht (rw_wfpot8f.F90, built into oracle/_ref/libkukan_ref.so by `make -C oracle ref`)
from seeded data, for the tests that read them where the reference is absent.  The data come
from `case_data` below, which the tests call to regenerate them.  Run from the repository root:

    python tests/golden/make_ckpt_golden.py
"""
from __future__ import annotations

import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
OUT = os.path.join(ROOT, "tests", "golden", "ckpt")
N = (6, 5, 4)

# case -> (nrc, nums, ncol, neigmx, numk, ksym) of the file
WF_CASES = {
    "wf_real_2spin": (0, 2, 1, 4, 1, 1),
    "wf_cplx_halfk": (1, 1, 1, 3, 2, 2),
    "wf_real_1spin": (0, 1, 1, 4, 1, 1),
    "wf_cplx_2spin": (1, 2, 1, 3, 1, 1),
}
RHO = dict(nspv=2, nradmx=4, npoint=3, num_atcell=2, natom=3, nmesh=2)


def case_data(name):
    """The seeded arrays written into case `name`."""
    rng = np.random.default_rng(sum(map(ord, name)))
    n = int(np.prod(N))
    if name in WF_CASES:
        nrc, nums, ncol, neig, numk, _ = WF_CASES[name]
        sh = (n, neig, nums, numk)
        v = rng.standard_normal(sh) + (1j * rng.standard_normal(sh) if nrc else 0)
        return dict(svecre=np.asfortranarray(v if nrc == 0 else np.zeros(1)),
                    sveccm=np.asfortranarray(v if nrc else np.zeros(1, dtype=np.complex128)),
                    sval=np.asfortranarray(rng.standard_normal((neig, nums + 1 - ncol, numk))))
    sh = (RHO["nradmx"], RHO["npoint"], RHO["nspv"], RHO["num_atcell"])
    return dict(rhosmt_i=np.asfortranarray(rng.random((n, RHO["nspv"]))),
                rhotrur_i=np.asfortranarray(rng.random(sh)), rhosmtr_i=np.asfortranarray(rng.random(sh)),
                rhosmt_o=np.asfortranarray(rng.random((n, RHO["nspv"]))),
                rhotrur_o=np.asfortranarray(rng.random(sh)), rhosmtr_o=np.asfortranarray(rng.random(sh)),
                vh_dense=np.asfortranarray(rng.standard_normal(n * RHO["nmesh"] ** 3)),
                natpri=np.array([2, 0, 2], dtype=np.int32))


def main():
    sys.path.insert(0, ROOT)
    from oracle import ref
    os.makedirs(OUT, exist_ok=True)
    for name, (nrc, nums, ncol, neig, numk, ksym) in WF_CASES.items():
        d = os.path.join(OUT, name)
        os.makedirs(d, exist_ok=True)
        ref.call("write_wf", nr1In=0, nr2In=0, chdir=d, nrc=nrc, nums=nums, ncol=ncol, neigmx=neig,
                 numk=numk, ncpx=N[0], ncpy=N[1], ncpz=N[2], nrxIn=0, nryIn=0, nrzIn=0, ksym=ksym,
                 **case_data(name))
    d = os.path.join(OUT, "rhovht")
    os.makedirs(d, exist_ok=True)
    a = case_data("rhovht")
    ref.call("write_rhovht", nrIn=0, chdir=d, nrho=2, nele=0, npre=0, nevhist=3, nspv=RHO["nspv"], ncpx=N[0],
             ncpy=N[1], ncpz=N[2], nmesh=RHO["nmesh"], nradmx=RHO["nradmx"], npoint=RHO["npoint"],
             num_atcell=RHO["num_atcell"], nrxIn=0, nryIn=0, nrzIn=0, natom=RHO["natom"], key_natpri_in=2, **a)
    for dp, _, fs in os.walk(OUT):
        for f in sorted(fs):
            print(os.path.relpath(os.path.join(dp, f), ROOT), os.path.getsize(os.path.join(dp, f)))


if __name__ == "__main__":
    main()
