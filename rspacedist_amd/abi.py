"""Argument tables of the C-ABI entry points (include/kukan_hip.h) and a ctypes binder.

Each entry is the reference Fortran routine's argument list in order, with a type code:
``i`` int32 scalar, ``d`` double scalar, ``I`` int32 array, ``D`` double array,
``Z`` complex*16 array (complex128, optional where the real path passes a dummy),
``oi``/``od`` int/double output scalar (passed by pointer).  ``bind(lib, prefix)`` attaches ctypes prototypes for every routine
present in ``lib`` under ``prefix + name``; ``call`` marshals keyword arguments:
numpy arrays (host memory, Fortran order) and torch tensors (``data_ptr()``, host or
device) are passed as pointers, scalars by value.
"""
from __future__ import annotations

import ctypes

import numpy as np

_I, _D = "i", "d"

SIGNATURES = {
    "tools_definefdcoef": [("nf", "i"), ("acfd", "D")],
    "overlap_finitedifference_r": [("nperi", "i"), ("ncpa", "i"), ("ncpb", "i"), ("ncpc", "i"),
                                   ("nf", "i"), ("nf1", "i"), ("nsnd", "i"), ("v_cc", "D")],
    "kslaplacian_r": [("ncpx", "i"), ("ncpy", "i"), ("ncpz", "i"), ("neigmx", "i"), ("nf", "i"),
                      ("dx", "d"), ("dy", "d"), ("dz", "d"), ("vre", "D"), ("veff", "D"),
                      ("l", "i"), ("avre", "D")],
    "ksprecondition": [("ncpx", "i"), ("ncpy", "i"), ("ncpz", "i"), ("nf", "i"), ("ncol", "i"),
                       ("vre", "D"), ("avre", "D")],
    "nonlocaloperation_r_01": [("natom", "i"), ("nprjmx", "i"), ("num_spe", "i"),
                               ("num_ppcell", "i"), ("num_list", "i"), ("ncpx", "i"),
                               ("ncpy", "i"), ("ncpz", "i"), ("key_natpri_in", "i"),
                               ("key_natpri_inps", "i"), ("nprj", "I"), ("indspe", "I"),
                               ("natinf", "I"), ("natpri", "I"), ("naps", "I"), ("lstvec2", "I"),
                               ("vnlocp", "D"), ("vre2", "D"), ("rpsep", "D")],
    "nonlocaloperation_r_02": [("nhs", "i"), ("natom", "i"), ("num_spe", "i"), ("nums", "i"),
                               ("nprjmx", "i"), ("ns", "i"), ("num_list", "i"),
                               ("num_ppcell", "i"), ("ncpx", "i"), ("ncpy", "i"), ("ncpz", "i"),
                               ("key_natpri_in", "i"), ("key_natpri_inps", "i"), ("dx", "d"),
                               ("dy", "d"), ("dz", "d"), ("nprj", "I"), ("indspe", "I"),
                               ("natinf", "I"), ("lstvec2", "I"), ("natpri", "I"), ("naps", "I"),
                               ("vnlocp", "D"), ("rpsep", "D"), ("dijsss", "D"), ("avre", "D"),
                               ("avr", "D")],
    "kscg_r": [("natom", "i"), ("num_spe", "i"), ("neigmx", "i"), ("nums", "i"), ("nspv", "i"),
               ("numk", "i"), ("nperi", "i"), ("northo", "i"), ("ns", "i"), ("nk", "i"),
               ("nf", "i"), ("nkscg", "i"), ("nsdmax", "i"), ("nprecon_cg", "i"),
               ("nprjmx", "i"), ("num_list", "i"), ("num_ppcell", "i"), ("ndisp", "i"),
               ("ncpx", "i"), ("ncpy", "i"), ("ncpz", "i"),
               ("key_ortho_nocmpt_innerproduct", "i"), ("key_ortho_cmpt_innerproduct", "i"),
               ("key_natpri_in", "i"), ("key_natpri_inps", "i"), ("key_natpri_out", "i"),
               ("dx", "d"), ("dy", "d"), ("dz", "d"), ("epssd", "d"), ("indspe", "I"),
               ("natpri", "I"), ("naps", "I"), ("nprj", "I"), ("natinf", "I"), ("lstvec2", "I"),
               ("latom", "I"), ("ntyppp", "I"), ("dij", "D"), ("veff", "D"), ("vnlocp", "D"),
               ("sss", "D"), ("ksconv", "oi"), ("ksitmax", "oi"), ("sval", "D"), ("rspsep", "D"),
               ("svecre", "D"), ("ssvre", "D"), ("residual_states", "D"), ("hsvre", "D")],
    "ksdiis_r": [("l_hsv", "i"), ("natom", "i"), ("num_spe", "i"), ("neigmx", "i"), ("nums", "i"),
                 ("nspv", "i"), ("numk", "i"), ("nperi", "i"), ("ndiismax", "i"),
                 ("nprecon_diis", "i"), ("nk", "i"), ("ns", "i"), ("nf", "i"), ("nprjmx", "i"),
                 ("num_list", "i"), ("num_ppcell", "i"), ("ndisp", "i"), ("northo", "i"),
                 ("ncpx", "i"), ("ncpy", "i"), ("ncpz", "i"),
                 ("key_ortho_cmpt_innerproduct", "i"), ("key_natpri_in", "i"),
                 ("key_natpri_inps", "i"), ("eps_eig_diis", "d"), ("alambda_diis", "d"),
                 ("ratio_diis", "d"), ("alambda_min", "d"), ("alambda_max", "d"), ("epssd", "d"),
                 ("xmax", "d"), ("ymax", "d"), ("zmax", "d"), ("indspe", "I"), ("natpri", "I"),
                 ("naps", "I"), ("nprj", "I"), ("natinf", "I"), ("lstvec2", "I"), ("latom", "I"),
                 ("ntyppp", "I"), ("dij", "D"), ("veff", "D"), ("vnlocp", "D"), ("sss", "D"),
                 ("ksconv", "oi"), ("ksitmax", "oi"), ("rspsep", "D"), ("sval", "D"),
                 ("svecre", "D"), ("ssvre", "D"), ("residual_states", "D"), ("hsvre", "D")],
    "orthogonalization_r": [("natom", "i"), ("neigmx", "i"), ("nprjmx", "i"), ("nums", "i"),
                            ("numk", "i"), ("num_ppcell", "i"), ("num_spe", "i"), ("northo", "i"),
                            ("num_list", "i"), ("nflag", "i"), ("ndisp", "i"), ("ncpx", "i"),
                            ("ncpy", "i"), ("ncpz", "i"), ("key_ortho_cmpt_innerproduct", "i"),
                            ("key_natpri_in", "i"), ("key_natpri_inps", "i"), ("dx", "d"),
                            ("dy", "d"), ("dz", "d"), ("nprj", "I"), ("indspe", "I"),
                            ("natpri", "I"), ("naps", "I"), ("natinf", "I"), ("lstvec2", "I"),
                            ("latom", "I"), ("ntyppp", "I"), ("vnlocp", "D"), ("sss", "D"),
                            ("svecre", "D"), ("ssvre", "D"), ("rspsep", "D")],
    "orthogonalization_r_01": [("natom", "i"), ("neigmx", "i"), ("nprjmx", "i"), ("nums", "i"),
                               ("numk", "i"), ("nk", "i"), ("ns", "i"), ("l_inp", "i"),
                               ("num_list", "i"), ("num_ppcell", "i"), ("num_spe", "i"),
                               ("nflag", "i"), ("ndisp", "i"), ("ncpx", "i"), ("ncpy", "i"),
                               ("ncpz", "i"), ("key_ortho_cmpt_innerproduct", "i"),
                               ("key_natpri_in", "i"), ("key_natpri_inps", "i"), ("dx", "d"),
                               ("dy", "d"), ("dz", "d"), ("nprj", "I"), ("indspe", "I"),
                               ("natpri", "I"), ("naps", "I"), ("natinf", "I"), ("lstvec2", "I"),
                               ("latom", "I"), ("vnlocp", "D"), ("sss", "D"), ("svecre", "D"),
                               ("ssvre", "D"), ("rspsep", "D")],
    "rayleighritz_r": [("natom", "i"), ("num_ppcell", "i"), ("num_spe", "i"), ("nprjmx", "i"),
                       ("neigmx", "i"), ("numk", "i"), ("nums", "i"), ("nspv", "i"),
                       ("num_list", "i"), ("ncpx", "i"), ("ncpy", "i"), ("ncpz", "i"),
                       ("nperi", "i"), ("nk", "i"), ("ns", "i"), ("nf", "i"), ("ndisp", "i"),
                       ("key_natpri_in", "i"), ("key_natpri_inps", "i"), ("dx", "d"), ("dy", "d"),
                       ("dz", "d"), ("nprj", "I"), ("natinf", "I"), ("lstvec2", "I"),
                       ("indspe", "I"), ("natpri", "I"), ("naps", "I"), ("ntyppp", "I"),
                       ("dij", "D"), ("veff", "D"), ("vnlocp", "D"), ("l_rrb", "i"),
                       ("rspsep", "D"), ("svecre", "D"), ("ssvre", "D"), ("hsvre", "D"),
                       ("sval", "D")],
    "reorderstates": [("nrc", "i"), ("num_ppcell", "i"), ("nprjmx", "i"), ("neigmx", "i"),
                      ("nums", "i"), ("ncol", "i"), ("numk", "i"), ("ncpx", "i"), ("ncpy", "i"),
                      ("ncpz", "i"), ("ns1", "i"), ("nk", "i"), ("sval", "D"),
                      ("residual_states", "D"), ("rspsep", "D"), ("cspsep", "Z"),
                      ("svecre", "D"), ("sveccm", "Z"), ("ssvre", "D"), ("ssvcm", "Z"),
                      ("hsvre", "D"), ("hsvcm", "Z")],
    "scf_diag": [("nrc", "i"), ("nso", "i"), ("nprjmx", "i"), ("num_ppcell", "i"),
                 ("num_spe", "i"), ("iscf", "i"), ("ncgscf", "i"), ("nretcg", "i"),
                 ("natom", "i"), ("neigmx", "i"), ("nums", "i"), ("ncol", "i"), ("nspv", "i"),
                 ("numk", "i"), ("nperi", "i"), ("northo", "i"), ("nf", "i"), ("ndisp", "i"),
                 ("num_list", "i"), ("ncpx", "i"), ("ncpy", "i"), ("ncpz", "i"), ("npxmax", "i"),
                 ("npymax", "i"), ("npzmax", "i"), ("nchange", "i"), ("nrrz", "i"),
                 ("key_ortho_nocmpt_innerproduct", "i"), ("key_ortho_cmpt_innerproduct", "i"),
                 ("key_natpri_in", "i"), ("key_natpri_inps", "i"), ("key_natpri_out", "i"),
                 ("key_soc_calc", "i"), ("key_pp_paw", "i"), ("skpxx", "D"), ("skpyy", "D"),
                 ("skpzz", "D"), ("dx", "d"), ("dy", "d"), ("dz", "d"), ("xmax", "d"),
                 ("ymax", "d"), ("zmax", "d"), ("dij", "D"), ("dijsoc", "D"), ("veff", "D"),
                 ("vnlocp", "D"), ("sss", "D"), ("nkscg", "i"), ("nsdmax", "i"),
                 ("nprecon_cg", "i"), ("epssd", "d"), ("ndiismax", "i"), ("nprecon_diis", "i"),
                 ("eps_eig_diis", "d"), ("alambda_diis", "d"), ("ratio_diis", "d"),
                 ("alambda_min", "d"), ("alambda_max", "d"), ("indspe", "I"), ("natpri", "I"),
                 ("naps", "I"), ("nprj", "I"), ("natinf", "I"), ("lstvec2", "I"), ("latom", "I"),
                 ("natsoc", "I"), ("ntyppp", "I"), ("lstx", "I"), ("lsty", "I"), ("lstz", "I"),
                 ("natx", "I"), ("naty", "I"), ("natz", "I"), ("rspsep", "D"), ("cspsep", "Z"),
                 ("svecre", "D"), ("sveccm", "Z"), ("hsvre", "D"), ("hsvcm", "Z"),
                 ("ssvre", "D"), ("ssvcm", "Z"), ("sval", "D"), ("svalsoc", "D"),
                 ("cspsepsoc", "Z"), ("svecsoc", "Z"), ("residual_states", "D"),
                 ("ksconv", "oi"), ("ksitmax", "oi")],
    "scf_hartree": [("nd", "i"), ("ncpx_d", "i"), ("ncpy_d", "i"), ("ncpz_d", "i"), ("nfh", "i"),
                    ("ndisp", "i"), ("nperi", "i"), ("ncgres", "i"), ("ncgmin", "i"),
                    ("ncgmax", "i"), ("xmax", "d"), ("ymax", "d"), ("zmax", "d"), ("epsvh", "d"),
                    ("rho_aug_dense", "D"), ("boundx", "D"), ("boundy", "D"), ("boundz", "D"),
                    ("vh_dense", "D")],
    # the allocation-free variant: same arguments, device pointers only
    "scf_hartree_noalloc": [("nd", "i"), ("ncpx_d", "i"), ("ncpy_d", "i"), ("ncpz_d", "i"),
                            ("nfh", "i"), ("ndisp", "i"), ("nperi", "i"), ("ncgres", "i"),
                            ("ncgmin", "i"), ("ncgmax", "i"), ("xmax", "d"), ("ymax", "d"),
                            ("zmax", "d"), ("epsvh", "d"), ("rho_aug_dense", "D"), ("boundx", "D"),
                            ("boundy", "D"), ("boundz", "D"), ("vh_dense", "D")],
    "tools_correctpotential": [("ncpx_d", "i"), ("ncpy_d", "i"), ("ncpz_d", "i"), ("xmax", "d"),
                               ("ymax", "d"), ("zmax", "d"), ("ddx", "d"), ("ddy", "d"),
                               ("ddz", "d"), ("vh_dense", "D")],
    "scf_charge_smt_r": [("neigmx", "i"), ("nums", "i"), ("nspv", "i"), ("numk", "i"),
                         ("ncpx", "i"), ("ncpy", "i"), ("ncpz", "i"), ("fnele", "D"),
                         ("svecre", "D"), ("rhosmt", "D")],
    "force_eig_r": [("nperi", "i"), ("numk", "i"), ("nums", "i"), ("neigmx", "i"), ("nf", "i"),
                    ("ncpx", "i"), ("ncpy", "i"), ("ncpz", "i"), ("num_list", "i"),
                    ("natom", "i"), ("num_spe", "i"), ("num_ppcell", "i"), ("nprjmx", "i"),
                    ("key_natpri_in", "i"), ("key_natpri_inps", "i"), ("lstvec2", "I"),
                    ("latom", "I"), ("lstx", "I"), ("lsty", "I"), ("lstz", "I"), ("nprj", "I"),
                    ("natpri", "I"), ("naps", "I"), ("indspe", "I"), ("natinf", "I"),
                    ("xmax", "d"), ("ymax", "d"), ("zmax", "d"), ("sss", "D"), ("vnlocp", "D"),
                    ("dij", "D"), ("svecre", "D"), ("sval", "D"), ("fnele", "D"), ("fatx", "D"),
                    ("faty", "D"), ("fatz", "D")],
    "band_energy_r": [("nperi", "i"), ("neigmx", "i"), ("nf", "i"), ("ncpx", "i"), ("ncpy", "i"),
                      ("ncpz", "i"), ("num_list", "i"), ("natom", "i"), ("num_spe", "i"),
                      ("num_ppcell", "i"), ("nprjmx", "i"), ("key_natpri_in", "i"),
                      ("key_natpri_inps", "i"), ("lstvec2", "I"), ("nprj", "I"), ("natpri", "I"),
                      ("naps", "I"), ("indspe", "I"), ("natinf", "I"), ("xmax", "d"),
                      ("ymax", "d"), ("zmax", "d"), ("vnlocp", "D"), ("dij", "D"), ("veff", "D"),
                      ("svecre", "D"), ("eband", "D")],
    "energy_adjoint_r": [("nperi", "i"), ("neigmx", "i"), ("nf", "i"), ("ncpx", "i"),
                         ("ncpy", "i"), ("ncpz", "i"), ("num_list", "i"), ("natom", "i"),
                         ("num_spe", "i"), ("num_ppcell", "i"), ("nprjmx", "i"),
                         ("key_natpri_in", "i"), ("key_natpri_inps", "i"), ("lstvec2", "I"),
                         ("nprj", "I"), ("natpri", "I"), ("naps", "I"), ("indspe", "I"),
                         ("natinf", "I"), ("xmax", "d"), ("ymax", "d"), ("zmax", "d"),
                         ("vnlocp", "D"), ("dvnlocp", "D"), ("dij", "D"), ("svecre", "D"),
                         ("fnele", "D"), ("dedv", "D"), ("dedr", "D")],
    "totalenergy_eig": [("nums", "i"), ("ncol", "i"), ("numk", "i"), ("neigmx", "i"),
                        ("nwskk", "I"), ("nwskptot", "i"), ("tf", "d"), ("sval", "D"),
                        ("fnele", "D"), ("eneel", "od"), ("eneth", "od")],
}

# complex (Bloch) path: nrc = 1, collinear spin, no spin-orbit (the reference argument
# lists; Z = complex*16)
BLOCH_SIGNATURES = {
    "overlap_finitedifference_c": [("nperi", "i"), ("ncpa", "i"), ("ncpb", "i"), ("ncpc", "i"),
                                   ("nf", "i"), ("nf1", "i"), ("nsnd", "i"), ("ncol", "i"),
                                   ("v_cc", "Z")],
    "kslaplacian_c": [("ncpx", "i"), ("ncpy", "i"), ("ncpz", "i"), ("neigmx", "i"), ("ncol", "i"),
                      ("nvef", "i"), ("nf", "i"), ("dx", "d"), ("dy", "d"), ("dz", "d"),
                      ("xmax", "d"), ("ymax", "d"), ("zmax", "d"), ("skpx", "d"), ("skpy", "d"),
                      ("skpz", "d"), ("vcm", "Z"), ("veff", "D"), ("l", "i"), ("avcm", "Z"),
                      ("workc", "Z")],
    "ksprecondition_c": [("ncpx", "i"), ("ncpy", "i"), ("ncpz", "i"), ("nf", "i"), ("ncol", "i"),
                         ("vcm", "Z"), ("avcm", "Z")],
    "nonlocaloperation_c_01": [("natom", "i"), ("nprjmx", "i"), ("num_spe", "i"),
                               ("num_ppcell", "i"), ("num_list", "i"), ("neigmx", "i"),
                               ("ncol", "i"), ("l", "i"), ("ncpx", "i"), ("ncpy", "i"),
                               ("ncpz", "i"), ("key_natpri_in", "i"), ("key_natpri_inps", "i"),
                               ("dx", "d"), ("dy", "d"), ("dz", "d"), ("skpx", "d"), ("skpy", "d"),
                               ("skpz", "d"), ("nprj", "I"), ("indspe", "I"), ("natinf", "I"),
                               ("natpri", "I"), ("naps", "I"), ("lstvec2", "I"), ("lstx", "I"),
                               ("lsty", "I"), ("lstz", "I"), ("natx", "I"), ("naty", "I"),
                               ("natz", "I"), ("vnlocp", "D"), ("vcm2", "Z"), ("cpsep", "Z"),
                               ("vcccm", "Z")],
    "nonlocaloperation_c_02": [("nhs", "i"), ("nso", "i"), ("natom", "i"), ("num_spe", "i"),
                               ("neigmx", "i"), ("nums", "i"), ("nprjmx", "i"), ("ncol", "i"),
                               ("ns1", "i"), ("l", "i"), ("num_list", "i"), ("num_ppcell", "i"),
                               ("ncpx", "i"), ("ncpy", "i"), ("ncpz", "i"), ("npxmax", "i"),
                               ("npymax", "i"), ("npzmax", "i"), ("key_natpri_in", "i"),
                               ("key_natpri_inps", "i"), ("key_soc_calc", "i"), ("dx", "d"),
                               ("dy", "d"), ("dz", "d"), ("skpx", "d"), ("skpy", "d"),
                               ("skpz", "d"), ("nprj", "I"), ("indspe", "I"), ("natinf", "I"),
                               ("lstvec2", "I"), ("natpri", "I"), ("naps", "I"), ("natsoc", "I"),
                               ("lstx", "I"), ("lsty", "I"), ("lstz", "I"), ("natx", "I"),
                               ("naty", "I"), ("natz", "I"), ("vnlocp", "D"), ("cpsep", "Z"),
                               ("dijsss", "D"), ("dijsoc", "D"), ("avcm", "Z"), ("vcccm", "Z"),
                               ("avc", "Z")],
    "orthogonalization_c_01": [("natom", "i"), ("neigmx", "i"), ("nprjmx", "i"), ("nums", "i"),
                               ("ncol", "i"), ("numk", "i"), ("nk", "i"), ("ns1", "i"),
                               ("l_inp", "i"), ("num_list", "i"), ("num_ppcell", "i"),
                               ("num_spe", "i"), ("nflag", "i"), ("ndisp", "i"), ("ncpx", "i"),
                               ("ncpy", "i"), ("ncpz", "i"), ("npxmax", "i"), ("npymax", "i"),
                               ("npzmax", "i"), ("key_ortho_cmpt_innerproduct", "i"),
                               ("key_natpri_in", "i"), ("key_natpri_inps", "i"), ("dx", "d"),
                               ("dy", "d"), ("dz", "d"), ("skpx", "d"), ("skpy", "d"),
                               ("skpz", "d"), ("nprj", "I"), ("indspe", "I"), ("natpri", "I"),
                               ("naps", "I"), ("natinf", "I"), ("lstvec2", "I"), ("latom", "I"),
                               ("lstx", "I"), ("lsty", "I"), ("lstz", "I"), ("natx", "I"),
                               ("naty", "I"), ("natz", "I"), ("vnlocp", "D"), ("sss", "D"),
                               ("sveccm", "Z"), ("ssvcm", "Z"), ("cspsep", "Z")],
    "orthogonalization_c": [("natom", "i"), ("neigmx", "i"), ("nprjmx", "i"), ("nums", "i"),
                            ("ncol", "i"), ("numk", "i"), ("num_ppcell", "i"), ("num_spe", "i"),
                            ("northo", "i"), ("num_list", "i"), ("nflag", "i"), ("ndisp", "i"),
                            ("ncpx", "i"), ("ncpy", "i"), ("ncpz", "i"), ("npxmax", "i"),
                            ("npymax", "i"), ("npzmax", "i"), ("key_ortho_cmpt_innerproduct", "i"),
                            ("key_natpri_in", "i"), ("key_natpri_inps", "i"), ("dx", "d"),
                            ("dy", "d"), ("dz", "d"), ("nprj", "I"), ("indspe", "I"),
                            ("natpri", "I"), ("naps", "I"), ("natinf", "I"), ("lstvec2", "I"),
                            ("latom", "I"), ("ntyppp", "I"), ("lstx", "I"), ("lsty", "I"),
                            ("lstz", "I"), ("natx", "I"), ("naty", "I"), ("natz", "I"),
                            ("skpxx", "D"), ("skpyy", "D"), ("skpzz", "D"), ("vnlocp", "D"),
                            ("sss", "D"), ("sveccm", "Z"), ("ssvcm", "Z"), ("cspsep", "Z")],
    "kscg_c": [("nso", "i"), ("natom", "i"), ("num_spe", "i"), ("neigmx", "i"), ("nums", "i"),
               ("ncol", "i"), ("nspv", "i"), ("numk", "i"), ("nperi", "i"), ("northo", "i"),
               ("ns1", "i"), ("nk", "i"), ("nf", "i"), ("nkscg", "i"), ("nsdmax", "i"),
               ("nprecon_cg", "i"), ("nprjmx", "i"), ("num_list", "i"), ("num_ppcell", "i"),
               ("ndisp", "i"), ("ncpx", "i"), ("ncpy", "i"), ("ncpz", "i"), ("npxmax", "i"),
               ("npymax", "i"), ("npzmax", "i"), ("key_ortho_nocmpt_innerproduct", "i"),
               ("key_ortho_cmpt_innerproduct", "i"), ("key_natpri_in", "i"),
               ("key_natpri_inps", "i"), ("key_natpri_out", "i"), ("key_soc_calc", "i"),
               ("dx", "d"), ("dy", "d"), ("dz", "d"), ("skpx", "d"), ("skpy", "d"), ("skpz", "d"),
               ("epssd", "d"), ("indspe", "I"), ("natpri", "I"), ("naps", "I"), ("nprj", "I"),
               ("natinf", "I"), ("lstvec2", "I"), ("latom", "I"), ("natsoc", "I"), ("ntyppp", "I"),
               ("lstx", "I"), ("lsty", "I"), ("lstz", "I"), ("natx", "I"), ("naty", "I"),
               ("natz", "I"), ("dij", "D"), ("dijsoc", "D"), ("veff", "D"), ("vnlocp", "D"),
               ("sss", "D"), ("ksconv", "oi"), ("ksitmax", "oi"), ("sval", "D"), ("cspsep", "Z"),
               ("sveccm", "Z"), ("ssvcm", "Z"), ("residual_states", "D"), ("hsvcm", "Z")],
    "ksdiis_c": [("l_hsv", "i"), ("nso", "i"), ("natom", "i"), ("num_spe", "i"), ("neigmx", "i"),
                 ("nums", "i"), ("ncol", "i"), ("nspv", "i"), ("numk", "i"), ("nperi", "i"),
                 ("ndiismax", "i"), ("nprecon_diis", "i"), ("nk", "i"), ("ns1", "i"), ("nf", "i"),
                 ("nprjmx", "i"), ("num_list", "i"), ("num_ppcell", "i"), ("ndisp", "i"),
                 ("northo", "i"), ("ncpx", "i"), ("ncpy", "i"), ("ncpz", "i"), ("npxmax", "i"),
                 ("npymax", "i"), ("npzmax", "i"), ("key_ortho_cmpt_innerproduct", "i"),
                 ("key_natpri_in", "i"), ("key_natpri_inps", "i"), ("key_soc_calc", "i"),
                 ("eps_eig_diis", "d"), ("alambda_diis", "d"), ("ratio_diis", "d"),
                 ("alambda_min", "d"), ("alambda_max", "d"), ("epssd", "d"), ("xmax", "d"),
                 ("ymax", "d"), ("zmax", "d"), ("skpx", "d"), ("skpy", "d"), ("skpz", "d"),
                 ("indspe", "I"), ("natpri", "I"), ("naps", "I"), ("nprj", "I"), ("natinf", "I"),
                 ("lstvec2", "I"), ("latom", "I"), ("natsoc", "I"), ("ntyppp", "I"), ("lstx", "I"),
                 ("lsty", "I"), ("lstz", "I"), ("natx", "I"), ("naty", "I"), ("natz", "I"),
                 ("dij", "D"), ("dijsoc", "D"), ("veff", "D"), ("vnlocp", "D"), ("sss", "D"),
                 ("ksconv", "oi"), ("ksitmax", "oi"), ("cspsep", "Z"), ("sval", "D"),
                 ("sveccm", "Z"), ("ssvcm", "Z"), ("residual_states", "D"), ("hsvcm", "Z")],
    "rayleighritz_c": [("nso", "i"), ("natom", "i"), ("num_ppcell", "i"), ("num_spe", "i"),
                       ("nprjmx", "i"), ("neigmx", "i"), ("numk", "i"), ("nums", "i"),
                       ("ncol", "i"), ("nspv", "i"), ("num_list", "i"), ("ncpx", "i"),
                       ("ncpy", "i"), ("ncpz", "i"), ("npxmax", "i"), ("npymax", "i"),
                       ("npzmax", "i"), ("nperi", "i"), ("nk", "i"), ("ns1", "i"), ("nf", "i"),
                       ("ndisp", "i"), ("key_natpri_in", "i"), ("key_natpri_inps", "i"),
                       ("key_soc_calc", "i"), ("dx", "d"), ("dy", "d"), ("dz", "d"), ("xmax", "d"),
                       ("ymax", "d"), ("zmax", "d"), ("skpxx", "D"), ("skpyy", "D"),
                       ("skpzz", "D"), ("nprj", "I"), ("natinf", "I"), ("lstvec2", "I"),
                       ("indspe", "I"), ("natpri", "I"), ("naps", "I"), ("natsoc", "I"),
                       ("ntyppp", "I"), ("lstx", "I"), ("lsty", "I"), ("lstz", "I"), ("natx", "I"),
                       ("naty", "I"), ("natz", "I"), ("dij", "D"), ("dijsoc", "D"), ("veff", "D"),
                       ("vnlocp", "D"), ("l_rrb", "i"), ("cspsep", "Z"), ("sveccm", "Z"),
                       ("ssvcm", "Z"), ("hsvcm", "Z"), ("sval", "D")],
    "scf_charge_smt_c": [("neigmx", "i"), ("nums", "i"), ("ncol", "i"), ("nspv", "i"),
                         ("numk", "i"), ("ncpx", "i"), ("ncpy", "i"), ("ncpz", "i"),
                         ("fnele", "D"), ("sveccm", "Z"), ("rhosmt", "D")],
}
SIGNATURES.update(BLOCH_SIGNATURES)

# potential assembly: the coarse <-> dense grid transfers and the augmented dense density
# (reference argument lists), and the V_eff sum of scf_potentials
TRANS_SIGNATURES = {
    "interpolation": [("ncpx", "i"), ("ncpy", "i"), ("ncpz", "i"), ("nmesh", "i"), ("v_dd", "D"),
                      ("v_cc", "D"), ("nperi", "i"), ("nf", "i"), ("is", "i"), ("ndisp", "i")],
    "overlap_interpolation": [("ndisp", "i"), ("nperi", "i"), ("ncpa", "i"), ("ncpb", "i"),
                              ("ncpc", "i"), ("nf", "i"), ("nf1", "i"), ("v_cc", "D"), ("nflag", "i")],
    "trans_c2d_smtcharge": [("nmesh", "i"), ("nspv", "i"), ("nperi", "i"), ("ndisp", "i"), ("nf", "i"),
                            ("ncpx", "i"), ("ncpy", "i"), ("ncpz", "i"), ("rhosmt", "D"),
                            ("rho_dense", "D")],
    "trans_d2c_smtcharge": [("nmesh", "i"), ("nperi", "i"), ("ndisp", "i"), ("nf", "i"), ("ncpx", "i"),
                            ("ncpy", "i"), ("ncpz", "i"), ("rho_aug_dense", "D"), ("rho_coarse", "D")],
    "trans_d2c_vh": [("natom", "i"), ("num_spe", "i"), ("nmesh", "i"), ("nperi", "i"), ("ndisp", "i"),
                     ("nf", "i"), ("ncpx", "i"), ("ncpy", "i"), ("ncpz", "i"), ("xmax", "d"),
                     ("ymax", "d"), ("zmax", "d"), ("indspe", "I"), ("atmpole", "D"), ("cp", "D"),
                     ("atx", "D"), ("aty", "D"), ("atz", "D"), ("vh_coarse", "D"), ("vh_dense", "D")],
    "trans_d2c_vxc": [("nmesh", "i"), ("nspv", "i"), ("nperi", "i"), ("ndisp", "i"), ("nf", "i"),
                      ("ncpx", "i"), ("ncpy", "i"), ("ncpz", "i"), ("vx", "D"), ("vx_dense", "D")],
    "scf_rhoaugdense": [("key_natpri_inps", "i"), ("key_pp_paw", "i"), ("natom", "i"), ("num_spe", "i"),
                        ("nspv", "i"), ("ncpx_d", "i"), ("ncpy_d", "i"), ("ncpz_d", "i"),
                        ("num_list_d", "i"), ("num_ppcell_d", "i"), ("indspe", "I"), ("natprid", "I"),
                        ("napsd", "I"), ("natinfd", "I"), ("ntyppp", "I"), ("lstvecd2", "I"),
                        ("rho_dense", "D"), ("rhoaug3d", "D"), ("rho_aug_dense", "D")],
    # no reference routine of its own: the last loop of scf_potentials (scf_potentials8f.F90:255-289)
    "scf_veff": [("nspv", "i"), ("ncpx", "i"), ("ncpy", "i"), ("ncpz", "i"), ("xmax", "d"),
                 ("ymax", "d"), ("zmax", "d"), ("dx", "d"), ("dy", "d"), ("dz", "d"), ("biasx", "d"),
                 ("biasy", "d"), ("biasz", "d"), ("vloc_coarse", "D"), ("vh_coarse", "D"), ("vx", "D"),
                 ("veff", "D")],
    # no reference routine of its own: scf_potentials' grid chain from the smooth density to
    # V_eff (scf_potentials8f.F90:115-289 without the PAW one-centre and XC terms)
    "scf_potentials_smooth": [
        ("key_natpri_inps", "i"), ("key_pp_paw", "i"), ("nmesh", "i"), ("ncpx", "i"), ("ncpy", "i"),
        ("ncpz", "i"), ("natom", "i"), ("num_spe", "i"), ("num_list_d", "i"), ("num_ppcell_d", "i"),
        ("nspv", "i"), ("ndisp", "i"), ("nperi", "i"), ("nf", "i"), ("nfh", "i"), ("ncgres", "i"),
        ("ncgmin", "i"), ("ncgmax", "i"), ("indspe", "I"), ("natprid", "I"), ("napsd", "I"),
        ("natinfd", "I"), ("ntyppp", "I"), ("lstvecd2", "I"), ("epsvh", "d"), ("xmax", "d"),
        ("ymax", "d"), ("zmax", "d"), ("dx", "d"), ("dy", "d"), ("dz", "d"), ("biasx", "d"),
        ("biasy", "d"), ("biasz", "d"), ("cp", "D"), ("atx", "D"), ("aty", "D"), ("atz", "D"),
        ("atmpole", "D"), ("rhosmt", "D"), ("rhoaug3d", "D"), ("vloc_coarse", "D"), ("vx", "D"),
        ("boundx", "D"), ("boundy", "D"), ("boundz", "D"), ("rho_aug_dense", "D"),
        ("vh_dense", "D"), ("vh_coarse", "D"), ("veff", "D")],
}
SIGNATURES.update(TRANS_SIGNATURES)

# checkpoint files wfc.* / rho.* / r_o.* / vht.* (rw_wfpot8f.F90; S = character, passed as a
# NUL-terminated string)
CKPT_SIGNATURES = {
    "write_wf": [("nr1In", "i"), ("nr2In", "i"), ("chdir", "S"), ("nrc", "i"), ("nums", "i"),
                 ("ncol", "i"), ("neigmx", "i"), ("numk", "i"), ("ncpx", "i"), ("ncpy", "i"),
                 ("ncpz", "i"), ("nrxIn", "i"), ("nryIn", "i"), ("nrzIn", "i"), ("ksym", "i"),
                 ("svecre", "D"), ("sveccm", "Z"), ("sval", "D")],
    "read_wf": [("nr1In", "i"), ("nr2In", "i"), ("chdir", "S"), ("nrc", "i"), ("nums", "i"),
                ("ncol", "i"), ("neigmx", "i"), ("numk", "i"), ("ncpx", "i"), ("ncpy", "i"),
                ("ncpz", "i"), ("ksym", "i"), ("key_ksym_inv2", "i"), ("svecre", "D"),
                ("sveccm", "Z"), ("sval", "D")],
    "read_rho": [("nrIn", "i"), ("chdir", "S"), ("nevhist", "i"), ("nspv", "i"), ("ncpx", "i"),
                 ("ncpy", "i"), ("ncpz", "i"), ("nradmx", "i"), ("npoint", "i"),
                 ("num_atcell", "i"), ("lbrydn", "oi"), ("na_atcell", "oi"), ("natcell_inf", "I"),
                 ("rhosmt", "D"), ("rhotrur", "D"), ("rhosmtr", "D")],
    "read_vht": [("nrIn", "i"), ("chdir", "S"), ("nevhist", "i"), ("ncpx", "i"), ("ncpy", "i"),
                 ("ncpz", "i"), ("nmesh", "i"), ("vh_dense", "D")],
    "write_rhovht": [("nrIn", "i"), ("chdir", "S"), ("nrho", "i"), ("nele", "i"), ("npre", "i"),
                     ("nevhist", "i"), ("nspv", "i"), ("ncpx", "i"), ("ncpy", "i"), ("ncpz", "i"),
                     ("nmesh", "i"), ("nradmx", "i"), ("npoint", "i"), ("num_atcell", "i"),
                     ("nrxIn", "i"), ("nryIn", "i"), ("nrzIn", "i"), ("natom", "i"),
                     ("key_natpri_in", "i"), ("natpri", "I"), ("rhosmt_i", "D"), ("rhotrur_i", "D"),
                     ("rhosmtr_i", "D"), ("rhosmt_o", "D"), ("rhotrur_o", "D"), ("rhosmtr_o", "D"),
                     ("vh_dense", "D")],
    "filenumbers_rhovht": [("chdir", "S"), ("nrrho", "oi"), ("nrpot", "oi")],
}
SIGNATURES.update(CKPT_SIGNATURES)

# projector tables, list vectors and forces (pseudocalc8f.F90, force8f.F90)
PP_SIGNATURES = {
    "pseudocalc_01a": [("nfdg", "i"), ("npmesh", "i"), ("wxyz", "D")],
    "pseudocalc_01b": [
        ("nperi", "i"), ("natom", "i"), ("num_spe", "i"), ("nradmx", "i"), ("nmesh", "i"), ("nfdg", "i"),
        ("ndisp", "i"), ("num_atcell", "i"), ("num_ppcell", "i"), ("num_ppcell_d", "i"), ("num_list", "i"),
        ("num_list_d", "i"), ("nxmax", "i"), ("nymax", "i"), ("nzmax", "i"), ("ncpx", "i"), ("ncpy", "i"),
        ("ncpz", "i"), ("npxmax", "i"), ("npymax", "i"), ("npzmax", "i"), ("key_natpri_in", "i"),
        ("key_natpri_inps", "i"), ("key_natpri_out", "i"), ("psctoff", "d"), ("xmax", "d"), ("ymax", "d"),
        ("zmax", "d"), ("dx", "d"), ("dy", "d"), ("dz", "d"), ("indspe", "I"), ("nradct", "I"),
        ("radial", "D"), ("atx", "D"), ("aty", "D"), ("atz", "D"), ("natx", "I"), ("naty", "I"),
        ("natz", "I"), ("ndatx", "I"), ("ndaty", "I"), ("ndatz", "I"), ("lstx", "I"), ("lsty", "I"),
        ("lstz", "I"), ("lstdx", "I"), ("lstdy", "I"), ("lstdz", "I"), ("natinf", "I"), ("natinfd", "I"),
        ("natinfd_vloc", "I"), ("lstvec2", "I"), ("lstvecd2", "I"), ("natpri", "I"), ("natprid", "I"),
        ("natpri_inf", "I"), ("naps", "I"), ("napsd", "I")],
    "pseudocalc_07": [
        ("natom", "i"), ("num_spe", "i"), ("npmesh", "i"), ("nprjmx", "i"), ("nfdg", "i"), ("nfiltyp", "i"),
        ("num_list", "i"), ("num_ppcell", "i"), ("nradmx", "i"), ("nqmx", "i"), ("nxmax", "i"),
        ("nymax", "i"), ("nzmax", "i"), ("npxmax", "i"), ("npymax", "i"), ("npzmax", "i"), ("nxspd", "i"),
        ("nyspd", "i"), ("nzspd", "i"), ("key_natpri_in", "i"), ("key_natpri_inps", "i"), ("eps", "d"),
        ("psftrad", "d"), ("psctoff", "d"), ("filpp", "d"), ("dx", "d"), ("dy", "d"), ("dz", "d"),
        ("indspe", "I"), ("nprj", "I"), ("nradct", "I"), ("natpri", "I"), ("naps", "I"), ("natinf", "I"),
        ("nqct", "I"), ("natx", "I"), ("naty", "I"), ("natz", "I"), ("lstx", "I"), ("lsty", "I"),
        ("lstz", "I"), ("wxyz", "D"), ("radial", "D"), ("coef", "D"), ("atx", "D"), ("aty", "D"),
        ("atz", "D"), ("vnlocp", "D"), ("vtmp_pp", "D"), ("vtmp_pp1", "D"), ("vtmp_pp2", "D"),
        ("vnlspd", "D")],
    "force": [
        ("natom", "i"), ("num_spe", "i"), ("num_atcell", "i"), ("num_ppcell_d", "i"), ("num_list_d", "i"),
        ("nqmx", "i"), ("nradmx", "i"), ("npoint", "i"), ("nmesh", "i"), ("nspv", "i"), ("nperi", "i"),
        ("nint1dmax", "i"), ("nzmax", "i"), ("jelcalc", "i"), ("ncpx", "i"), ("ncpy", "i"), ("ncpz", "i"),
        ("ncpx_d", "i"), ("ncpy_d", "i"), ("ncpz_d", "i"), ("new_pwx", "i"), ("new_pwy", "i"),
        ("new_pwz", "i"), ("new_rsx", "i"), ("new_rsy", "i"), ("new_rsz", "i"), ("key_jel_calc", "i"),
        ("key_natpri_in", "i"), ("key_natpri_inps", "i"), ("indspe", "I"), ("natpri", "I"),
        ("natpri_inf", "I"), ("nradct", "I"), ("natprid", "I"), ("napsd", "I"), ("natinfd", "I"),
        ("natinfd_vloc", "I"), ("lstvecd2", "I"), ("nqctpcc", "I"), ("lstdx", "I"), ("lstdy", "I"),
        ("lstdz", "I"), ("veta", "d"), ("psctoff", "d"), ("psftrad", "d"), ("radial", "D"),
        ("dradial", "D"), ("point", "D"), ("wt", "D"), ("xmax", "d"), ("ymax", "d"), ("zmax", "d"),
        ("biasx", "d"), ("biasy", "d"), ("biasz", "d"), ("cp", "D"), ("coef", "D"), ("atx", "D"),
        ("aty", "D"), ("atz", "D"), ("dvlocdx_scw", "D"), ("dvlocdy_scw", "D"), ("dvlocdz_scw", "D"),
        ("dvlocdx_hdp", "D"), ("dvlocdy_hdp", "D"), ("dvlocdz_hdp", "D"), ("rhotrur", "D"),
        ("rhosmtr", "D"), ("rhoaugr", "D"), ("drhoaug3ddx", "D"), ("drhoaug3ddy", "D"),
        ("drhoaug3ddz", "D"), ("rho_coarse", "D"), ("rho_aug_dense", "D"), ("vloc_dense", "D"),
        ("vh_dense", "D"), ("vx_dense", "D"), ("chrjel", "d"), ("strjel", "d"), ("endjel", "d"),
        ("ndatx", "I"), ("ndaty", "I"), ("ndatz", "I"), ("fatx", "D"), ("faty", "D"), ("fatz", "D"),
        ("fjel", "od")],
}
SIGNATURES.update(PP_SIGNATURES)

_CT = {"i": ctypes.c_int, "d": ctypes.c_double, "I": ctypes.c_void_p, "D": ctypes.c_void_p,
       "Z": ctypes.c_void_p, "oi": ctypes.POINTER(ctypes.c_int),
       "od": ctypes.POINTER(ctypes.c_double), "S": ctypes.c_char_p}


def bind(lib, prefix, table=None):
    """Attach prototypes for every routine of SIGNATURES (or ``table``) found in lib."""
    for name, args in (table or SIGNATURES).items():
        fn = getattr(lib, prefix + name, None)
        if fn is None:
            continue
        fn.restype = ctypes.c_int
        fn.argtypes = [_CT[t] for _, t in args]
    return lib


def _ptr(x, code, name):
    if x is None:
        return None
    if isinstance(x, np.ndarray):
        want = {"I": np.int32, "D": np.float64, "Z": np.complex128}.get(code)
        if want is not None and x.dtype != want:
            raise TypeError(f"argument {name}: dtype {x.dtype}, expected {want}")
        if not (x.flags.f_contiguous or x.flags.c_contiguous):
            raise TypeError(f"argument {name}: array must be contiguous")
        if x.ndim > 1 and not x.flags.f_contiguous:
            raise TypeError(f"argument {name}: multi-dimensional array must be Fortran-ordered")
        return x.ctypes.data
    if hasattr(x, "device_ptr"):  # kukan.DeviceArray
        return x.device_ptr
    if hasattr(x, "data_ptr"):  # torch tensor
        return x.data_ptr()
    if isinstance(x, int):
        return x
    raise TypeError(f"argument {name}: cannot pass {type(x)}")


class Outs(dict):
    """Output scalars of a call (ksconv, ksitmax, eneel, ...)."""


def call(lib, prefix, name, err=None, _table=None, **kw):
    """Call prefix+name with keyword arguments named as in the reference.  Scalar outputs
    (``oi``/``od``) may be given as initial values; they are returned in an Outs dict
    together with the return code under key ``rc``."""
    args = (_table or SIGNATURES)[name]
    fn = getattr(lib, prefix + name)
    cargs, outs = [], Outs()
    for an, t in args:
        if t in ("i", "d"):
            if an not in kw:
                raise TypeError(f"{name}: missing scalar {an}")
            cargs.append(kw[an])
        elif t == "S":  # a character argument: NUL-terminated bytes
            v = kw.get(an, "")
            cargs.append(v.encode() if isinstance(v, str) else bytes(v))
        elif t in ("oi", "od"):
            box = (ctypes.c_int if t == "oi" else ctypes.c_double)(kw.get(an, 0))
            outs[an] = box
            cargs.append(ctypes.byref(box))
        else:
            if an not in kw and t != "Z":
                raise TypeError(f"{name}: missing array {an}")
            cargs.append(_ptr(kw.get(an), t, an))
    rc = fn(*cargs)
    res = Outs({k: v.value for k, v in outs.items()})
    res["rc"] = rc
    if rc != 0 and err is not None:
        raise RuntimeError(f"{prefix}{name} failed ({rc}): {err()}")
    return res
