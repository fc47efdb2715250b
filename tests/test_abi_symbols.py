"""CPU checks of the C-ABI library: it loads and exports every symbol include/kukan_hip.h
declares (no compute calls: there is no GPU here)."""
import ctypes
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def header_symbols():
    src = open(os.path.join(ROOT, "include", "kukan_hip.h")).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"\b(kukan_\w+)\s*\(", src)))


def test_header_declares_reference_routines():
    syms = header_symbols()
    for r in ("kslaplacian_r", "nonlocaloperation_r_01", "nonlocaloperation_r_02", "ksprecondition",
              "kscg_r", "ksdiis_r", "orthogonalization_r", "orthogonalization_r_01",
              "rayleighritz_r", "reorderstates", "scf_diag", "scf_hartree",
              "tools_correctpotential", "scf_charge_smt_r", "force_eig_r", "totalenergy_eig",
              "overlap_finitedifference_r", "overlap_fdcheck_r", "tools_definefdcoef"):
        assert "kukan_" + r in syms


def test_library_exports_every_declared_symbol():
    path = os.path.join(ROOT, "rspacedist_amd", "libkukan_hip.so")
    if not os.path.exists(path):
        pytest.skip("libkukan_hip.so not built")
    lib = ctypes.CDLL(path)
    missing = [s for s in header_symbols() if not hasattr(lib, s)]
    assert not missing, missing


def test_signature_table_matches_header():
    from rspacedist_amd import abi
    src = open(os.path.join(ROOT, "include", "kukan_hip.h")).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    for name, args in abi.SIGNATURES.items():
        m = re.search(r"int kukan_%s\((.*?)\);" % name, src, flags=re.S)
        assert m, name
        params = [p.strip() for p in m.group(1).split(",")]
        names = [re.findall(r"(\w+)$", p)[0] for p in params]
        assert names == [a for a, _ in args], name


def test_no_oracle_in_product_path():
    pkg = os.path.join(ROOT, "rspacedist_amd")
    for dp, _, fs in os.walk(pkg):
        for f in fs:
            if f.endswith((".py", ".hip", ".hpp", ".F90", ".cpp")):
                txt = open(os.path.join(dp, f), errors="ignore").read()
                for pat in ("import oracle", "from oracle", "liboracle", "kukan_oracle"):
                    assert pat not in txt, (f, pat)


def test_subspace_gemm_wrapper_marshals_arguments():
    """The ctypes wrapper reaches the library; without a device the call reports the HIP error
    instead of computing anything on the host."""
    import numpy as np
    from rspacedist_amd import kukan
    A = np.asfortranarray(np.arange(8.0).reshape(4, 2))
    C = np.zeros((2, 2), order="F")
    try:
        kukan.subspace_gemm(1, 2, 2, 4, 1.0, A, 4, A, 4, 0.0, C, 2)
    except kukan.KukanError as e:
        assert "device" in str(e) or "HIP" in str(e)
        assert not C.any()
    else:  # a GPU is present
        np.testing.assert_allclose(C, A.T @ A)


def _shim_interfaces():
    """bind(C) interfaces of the Fortran drop-in: C symbol -> [(argument, kind)], kind one of
    'i'/'d' (integer/real by value), 'I'/'D'/'Z'/'S' (integer/real/complex/character array),
    'oi'/'od' (integer/real by reference)."""
    src = open(os.path.join(ROOT, "rspacedist_amd", "fortran", "kukan_hip_shim.F90")).read()
    src = re.sub(r"&\s*\n\s*", "", src)  # join continuation lines
    out = {}
    for m in re.finditer(r"function\s+c_\w+\(([^)]*)\)\s*bind\(C,\s*name='(kukan_\w+)'\)(.*?)end function", src,
                         flags=re.S | re.I):
        names = [a.strip() for a in m.group(1).split(",") if a.strip()]
        decl = {}
        for line in m.group(3).splitlines():
            line = line.strip()
            if "::" not in line:
                continue
            typ, vars_ = line.split("::", 1)
            val = "value" in typ
            for v in re.split(r",\s*(?![^(]*\))", vars_):
                v = v.strip()
                arr = "(" in v
                nm = v.split("(")[0].strip()
                if typ.lower().startswith("integer"):
                    k = "I" if arr else ("i" if val else "oi")
                elif typ.lower().startswith("real"):
                    k = "D" if arr else ("d" if val else "od")
                elif typ.lower().startswith("complex"):
                    k = "Z"
                else:
                    k = "S"
                decl[nm] = k
        out[m.group(2)] = [(a, decl.get(a)) for a in names]
    return out


def test_fortran_shim_interfaces_match_signature_table():
    """Every bind(C) interface of the Fortran drop-in passes its arguments in the order and
    with the kinds of the C-ABI (rspacedist_amd/abi.py, checked against the header above)."""
    from rspacedist_amd import abi
    ifs = _shim_interfaces()
    assert len(ifs) > 20
    for sym, args in ifs.items():
        name = sym[len("kukan_"):]
        if name not in abi.SIGNATURES:
            continue  # library utilities (init, set_param, bind, ...)
        want = [t for _, t in abi.SIGNATURES[name]]
        got = [k for _, k in args]
        assert got == want, (sym, [(a, k, w) for (a, k), w in zip(args, want) if k != w])
    for name in ("pseudocalc_01a", "pseudocalc_01b", "pseudocalc_07", "force", "scf_diag", "kscg_c"):
        assert "kukan_" + name in ifs
