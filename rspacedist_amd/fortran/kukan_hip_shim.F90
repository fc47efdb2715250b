! kukan_hip_shim.F90 -- Fortran drop-in for the RSPACE kukan v8f Kohn-Sham hot path on MI355X.
!
! The modules below are named after the reference modules with a _hip suffix (so that both
! can be linked into one program)
! and carry the reference routine names with the reference argument lists
! (RSPACEes.v08.01/kukan/*.F90, cited per routine).  Each routine forwards
! its arrays by address and its scalars by value to libkukan_hip.so through ISO_C_BINDING
! (include/kukan_hip.h) and maps a nonzero return code to the reference's abort path with
! the library's message (which repeats the reference's stopp text).
!
! Build with -DKUKAN_USE_STOPP to abort through the host program's mod_stopp::stopp;
! otherwise the shim prints the message and stops with ERROR STOP.

module kukan_hip_c
  use iso_c_binding
  implicit none
  interface
    integer(c_int) function kukan_init(device) bind(C, name='kukan_init')
      import :: c_int
      integer(c_int), value :: device
    end function
    integer(c_int) function kukan_finalize() bind(C, name='kukan_finalize')
      import :: c_int
    end function
    type(c_ptr) function kukan_last_error() bind(C, name='kukan_last_error')
      import :: c_ptr
    end function
    integer(c_int) function kukan_set_param(key, val) bind(C, name='kukan_set_param')
      import :: c_int, c_char, c_double
      character(kind=c_char), intent(in) :: key(*)
      real(c_double), value :: val
    end function
    integer(c_int) function kukan_bind(p, nbytes) bind(C, name='kukan_bind')
      import :: c_int, c_ptr, c_size_t
      type(c_ptr), value :: p
      integer(c_size_t), value :: nbytes
    end function
    integer(c_int) function kukan_unbind(p) bind(C, name='kukan_unbind')
      import :: c_int, c_ptr
      type(c_ptr), value :: p
    end function
    integer(c_int) function kukan_sync_to_host(p) bind(C, name='kukan_sync_to_host')
      import :: c_int, c_ptr
      type(c_ptr), value :: p
    end function
    integer(c_int) function kukan_sync_to_device(p) bind(C, name='kukan_sync_to_device')
      import :: c_int, c_ptr
      type(c_ptr), value :: p
    end function
    integer(c_int) function c_definefdcoef(nf, acfd) bind(C, name='kukan_tools_definefdcoef')
      import :: c_int, c_double
      integer(c_int), value :: nf
      real(c_double) :: acfd(*)
    end function
    integer(c_int) function c_overlap_fd(nperi, na, nb, nc, nf, nf1, nsnd, v) &
        bind(C, name='kukan_overlap_finitedifference_r')
      import :: c_int, c_double
      integer(c_int), value :: nperi, na, nb, nc, nf, nf1, nsnd
      real(c_double) :: v(*)
    end function
    integer(c_int) function c_kslaplacian_r(ncpx, ncpy, ncpz, neigmx, nf, dx, dy, dz, vre, veff, l, avre) &
        bind(C, name='kukan_kslaplacian_r')
      import :: c_int, c_double
      integer(c_int), value :: ncpx, ncpy, ncpz, neigmx, nf, l
      real(c_double), value :: dx, dy, dz
      real(c_double) :: vre(*), veff(*), avre(*)
    end function
    integer(c_int) function c_ksprecondition(ncpx, ncpy, ncpz, nf, ncol, vre, avre) &
        bind(C, name='kukan_ksprecondition')
      import :: c_int, c_double
      integer(c_int), value :: ncpx, ncpy, ncpz, nf, ncol
      real(c_double) :: vre(*), avre(*)
    end function
    integer(c_int) function c_nonlocal_r_01(natom, nprjmx, num_spe, num_ppcell, num_list, ncpx, ncpy, &
        ncpz, kin, kinps, nprj, indspe, natinf, natpri, naps, lstvec2, vnlocp, vre2, rpsep) &
        bind(C, name='kukan_nonlocaloperation_r_01')
      import :: c_int, c_double
      integer(c_int), value :: natom, nprjmx, num_spe, num_ppcell, num_list, ncpx, ncpy, ncpz, kin, kinps
      integer(c_int) :: nprj(*), indspe(*), natinf(*), natpri(*), naps(*), lstvec2(*)
      real(c_double) :: vnlocp(*), vre2(*), rpsep(*)
    end function
    integer(c_int) function c_nonlocal_r_02(nhs, natom, num_spe, nums, nprjmx, ns, num_list, num_ppcell, &
        ncpx, ncpy, ncpz, kin, kinps, dx, dy, dz, nprj, indspe, natinf, lstvec2, natpri, naps, &
        vnlocp, rpsep, dijsss, avre, avr) bind(C, name='kukan_nonlocaloperation_r_02')
      import :: c_int, c_double
      integer(c_int), value :: nhs, natom, num_spe, nums, nprjmx, ns, num_list, num_ppcell, ncpx, ncpy, ncpz
      integer(c_int), value :: kin, kinps
      real(c_double), value :: dx, dy, dz
      integer(c_int) :: nprj(*), indspe(*), natinf(*), lstvec2(*), natpri(*), naps(*)
      real(c_double) :: vnlocp(*), rpsep(*), dijsss(*), avre(*), avr(*)
    end function
    integer(c_int) function c_kscg_r(natom, num_spe, neigmx, nums, nspv, numk, nperi, northo, ns, nk, nf, &
        nkscg, nsdmax, nprecon_cg, nprjmx, num_list, num_ppcell, ndisp, ncpx, ncpy, ncpz, knocmpt, kcmpt, &
        kin, kinps, kout, dx, dy, dz, epssd, indspe, natpri, naps, nprj, natinf, lstvec2, latom, ntyppp, &
        dij, veff, vnlocp, sss, ksconv, ksitmax, sval, rspsep, svecre, ssvre, residual_states, hsvre) &
        bind(C, name='kukan_kscg_r')
      import :: c_int, c_double
      integer(c_int), value :: natom, num_spe, neigmx, nums, nspv, numk, nperi, northo, ns, nk, nf
      integer(c_int), value :: nkscg, nsdmax, nprecon_cg, nprjmx, num_list, num_ppcell, ndisp
      integer(c_int), value :: ncpx, ncpy, ncpz, knocmpt, kcmpt, kin, kinps, kout
      real(c_double), value :: dx, dy, dz, epssd
      integer(c_int) :: indspe(*), natpri(*), naps(*), nprj(*), natinf(*), lstvec2(*), latom(*), ntyppp(*)
      real(c_double) :: dij(*), veff(*), vnlocp(*), sss(*)
      integer(c_int) :: ksconv, ksitmax
      real(c_double) :: sval(*), rspsep(*), svecre(*), ssvre(*), residual_states(*), hsvre(*)
    end function
    integer(c_int) function c_ksdiis_r(l_hsv, natom, num_spe, neigmx, nums, nspv, numk, nperi, ndiismax, &
        nprecon_diis, nk, ns, nf, nprjmx, num_list, num_ppcell, ndisp, northo, ncpx, ncpy, ncpz, kcmpt, &
        kin, kinps, eps_eig_diis, alambda_diis, ratio_diis, alambda_min, alambda_max, epssd, xmax, ymax, &
        zmax, indspe, natpri, naps, nprj, natinf, lstvec2, latom, ntyppp, dij, veff, vnlocp, sss, ksconv, &
        ksitmax, rspsep, sval, svecre, ssvre, residual_states, hsvre) bind(C, name='kukan_ksdiis_r')
      import :: c_int, c_double
      integer(c_int), value :: l_hsv, natom, num_spe, neigmx, nums, nspv, numk, nperi, ndiismax
      integer(c_int), value :: nprecon_diis, nk, ns, nf, nprjmx, num_list, num_ppcell, ndisp, northo
      integer(c_int), value :: ncpx, ncpy, ncpz, kcmpt, kin, kinps
      real(c_double), value :: eps_eig_diis, alambda_diis, ratio_diis, alambda_min, alambda_max, epssd
      real(c_double), value :: xmax, ymax, zmax
      integer(c_int) :: indspe(*), natpri(*), naps(*), nprj(*), natinf(*), lstvec2(*), latom(*), ntyppp(*)
      real(c_double) :: dij(*), veff(*), vnlocp(*), sss(*)
      integer(c_int) :: ksconv, ksitmax
      real(c_double) :: rspsep(*), sval(*), svecre(*), ssvre(*), residual_states(*), hsvre(*)
    end function
    integer(c_int) function c_ortho_r(natom, neigmx, nprjmx, nums, numk, num_ppcell, num_spe, northo, &
        num_list, nflag, ndisp, ncpx, ncpy, ncpz, kcmpt, kin, kinps, dx, dy, dz, nprj, indspe, natpri, &
        naps, natinf, lstvec2, latom, ntyppp, vnlocp, sss, svecre, ssvre, rspsep) &
        bind(C, name='kukan_orthogonalization_r')
      import :: c_int, c_double
      integer(c_int), value :: natom, neigmx, nprjmx, nums, numk, num_ppcell, num_spe, northo, num_list
      integer(c_int), value :: nflag, ndisp, ncpx, ncpy, ncpz, kcmpt, kin, kinps
      real(c_double), value :: dx, dy, dz
      integer(c_int) :: nprj(*), indspe(*), natpri(*), naps(*), natinf(*), lstvec2(*), latom(*), ntyppp(*)
      real(c_double) :: vnlocp(*), sss(*), svecre(*), ssvre(*), rspsep(*)
    end function
    integer(c_int) function c_ortho_r_01(natom, neigmx, nprjmx, nums, numk, nk, ns, l_inp, num_list, &
        num_ppcell, num_spe, nflag, ndisp, ncpx, ncpy, ncpz, kcmpt, kin, kinps, dx, dy, dz, nprj, indspe, &
        natpri, naps, natinf, lstvec2, latom, vnlocp, sss, svecre, ssvre, rspsep) &
        bind(C, name='kukan_orthogonalization_r_01')
      import :: c_int, c_double
      integer(c_int), value :: natom, neigmx, nprjmx, nums, numk, nk, ns, l_inp, num_list, num_ppcell
      integer(c_int), value :: num_spe, nflag, ndisp, ncpx, ncpy, ncpz, kcmpt, kin, kinps
      real(c_double), value :: dx, dy, dz
      integer(c_int) :: nprj(*), indspe(*), natpri(*), naps(*), natinf(*), lstvec2(*), latom(*)
      real(c_double) :: vnlocp(*), sss(*), svecre(*), ssvre(*), rspsep(*)
    end function
    integer(c_int) function c_rayleighritz_r(natom, num_ppcell, num_spe, nprjmx, neigmx, numk, nums, nspv, &
        num_list, ncpx, ncpy, ncpz, nperi, nk, ns, nf, ndisp, kin, kinps, dx, dy, dz, nprj, natinf, &
        lstvec2, indspe, natpri, naps, ntyppp, dij, veff, vnlocp, l_rrb, rspsep, svecre, ssvre, hsvre, &
        sval) bind(C, name='kukan_rayleighritz_r')
      import :: c_int, c_double
      integer(c_int), value :: natom, num_ppcell, num_spe, nprjmx, neigmx, numk, nums, nspv, num_list
      integer(c_int), value :: ncpx, ncpy, ncpz, nperi, nk, ns, nf, ndisp, kin, kinps, l_rrb
      real(c_double), value :: dx, dy, dz
      integer(c_int) :: nprj(*), natinf(*), lstvec2(*), indspe(*), natpri(*), naps(*), ntyppp(*)
      real(c_double) :: dij(*), veff(*), vnlocp(*), rspsep(*), svecre(*), ssvre(*), hsvre(*), sval(*)
    end function
    integer(c_int) function c_reorderstates(nrc, num_ppcell, nprjmx, neigmx, nums, ncol, numk, ncpx, ncpy, &
        ncpz, ns1, nk, sval, residual_states, rspsep, cspsep, svecre, sveccm, ssvre, ssvcm, hsvre, hsvcm) &
        bind(C, name='kukan_reorderstates')
      import :: c_int, c_double, c_double_complex
      integer(c_int), value :: nrc, num_ppcell, nprjmx, neigmx, nums, ncol, numk, ncpx, ncpy, ncpz, ns1, nk
      real(c_double) :: sval(*), residual_states(*), rspsep(*), svecre(*), ssvre(*), hsvre(*)
      complex(c_double_complex) :: cspsep(*), sveccm(*), ssvcm(*), hsvcm(*)
    end function
    integer(c_int) function c_scf_diag(nrc, nso, nprjmx, num_ppcell, num_spe, iscf, ncgscf, nretcg, natom, &
        neigmx, nums, ncol, nspv, numk, nperi, northo, nf, ndisp, num_list, ncpx, ncpy, ncpz, npxmax, &
        npymax, npzmax, nchange, nrrz, knocmpt, kcmpt, kin, kinps, kout, ksoc, kpaw, skpxx, skpyy, skpzz, &
        dx, dy, dz, xmax, ymax, zmax, dij, dijsoc, veff, vnlocp, sss, nkscg, nsdmax, nprecon_cg, epssd, &
        ndiismax, nprecon_diis, eps_eig_diis, alambda_diis, ratio_diis, alambda_min, alambda_max, indspe, &
        natpri, naps, nprj, natinf, lstvec2, latom, natsoc, ntyppp, lstx, lsty, lstz, natx, naty, natz, &
        rspsep, cspsep, svecre, sveccm, hsvre, hsvcm, ssvre, ssvcm, sval, svalsoc, cspsepsoc, svecsoc, &
        residual_states, ksconv, ksitmax) bind(C, name='kukan_scf_diag')
      import :: c_int, c_double, c_double_complex
      integer(c_int), value :: nrc, nso, nprjmx, num_ppcell, num_spe, iscf, ncgscf, nretcg, natom, neigmx
      integer(c_int), value :: nums, ncol, nspv, numk, nperi, northo, nf, ndisp, num_list, ncpx, ncpy, ncpz
      integer(c_int), value :: npxmax, npymax, npzmax, nchange, nrrz, knocmpt, kcmpt, kin, kinps, kout
      integer(c_int), value :: ksoc, kpaw, nkscg, nsdmax, nprecon_cg, ndiismax, nprecon_diis
      real(c_double), value :: dx, dy, dz, xmax, ymax, zmax, epssd, eps_eig_diis, alambda_diis
      real(c_double), value :: ratio_diis, alambda_min, alambda_max
      real(c_double) :: skpxx(*), skpyy(*), skpzz(*), dij(*), dijsoc(*), veff(*), vnlocp(*), sss(*)
      integer(c_int) :: indspe(*), natpri(*), naps(*), nprj(*), natinf(*), lstvec2(*), latom(*), natsoc(*)
      integer(c_int) :: ntyppp(*), lstx(*), lsty(*), lstz(*), natx(*), naty(*), natz(*)
      real(c_double) :: rspsep(*), svecre(*), hsvre(*), ssvre(*), sval(*), svalsoc(*), residual_states(*)
      complex(c_double_complex) :: cspsep(*), sveccm(*), hsvcm(*), ssvcm(*), cspsepsoc(*), svecsoc(*)
      integer(c_int) :: ksconv, ksitmax
    end function
    integer(c_int) function c_scf_hartree(nd, nx, ny, nz, nfh, ndisp, nperi, ncgres, ncgmin, ncgmax, xmax, &
        ymax, zmax, epsvh, rho, bx, by, bz, vh) bind(C, name='kukan_scf_hartree')
      import :: c_int, c_double
      integer(c_int), value :: nd, nx, ny, nz, nfh, ndisp, nperi, ncgres, ncgmin, ncgmax
      real(c_double), value :: xmax, ymax, zmax, epsvh
      real(c_double) :: rho(*), bx(*), by(*), bz(*), vh(*)
    end function
    integer(c_int) function c_correctpotential(nx, ny, nz, xmax, ymax, zmax, ddx, ddy, ddz, vh) &
        bind(C, name='kukan_tools_correctpotential')
      import :: c_int, c_double
      integer(c_int), value :: nx, ny, nz
      real(c_double), value :: xmax, ymax, zmax, ddx, ddy, ddz
      real(c_double) :: vh(*)
    end function
    integer(c_int) function c_charge_smt_r(neigmx, nums, nspv, numk, ncpx, ncpy, ncpz, fnele, svecre, rho) &
        bind(C, name='kukan_scf_charge_smt_r')
      import :: c_int, c_double
      integer(c_int), value :: neigmx, nums, nspv, numk, ncpx, ncpy, ncpz
      real(c_double) :: fnele(*), svecre(*), rho(*)
    end function
    integer(c_int) function c_kscg_c( &
    nso, natom, num_spe, neigmx, nums, ncol, nspv, numk, nperi, northo, ns1, nk, nf, nkscg, nsdmax, &
        nprecon_cg, nprjmx, num_list, num_ppcell, ndisp, ncpx, ncpy, ncpz, npxmax, npymax, npzmax, &
        key_ortho_nocmpt_innerproduct, key_ortho_cmpt_innerproduct, key_natpri_in, key_natpri_inps, &
        key_natpri_out, key_soc_calc, dx, dy, dz, skpx, skpy, skpz, epssd, indspe, natpri, naps, nprj, &
        natinf, lstvec2, latom, natsoc, ntyppp, lstx, lsty, lstz, natx, naty, natz, dij, dijsoc, veff, &
        vnlocp, sss, ksconv, ksitmax, sval, cspsep, sveccm, ssvcm, residual_states, hsvcm) &
        bind(C, name='kukan_kscg_c')
      import :: c_int, c_double, c_double_complex
      integer(c_int), value :: nso, natom, num_spe, neigmx, nums, ncol, nspv, numk, nperi, northo, ns1
      integer(c_int), value :: nk, nf, nkscg, nsdmax, nprecon_cg, nprjmx, num_list, num_ppcell, ndisp, ncpx
      integer(c_int), value :: ncpy, ncpz, npxmax, npymax, npzmax, key_ortho_nocmpt_innerproduct, key_ortho_cmpt_innerproduct
      integer(c_int), value :: key_natpri_in, key_natpri_inps, key_natpri_out, key_soc_calc
      real(c_double), value :: dx, dy, dz, skpx, skpy, skpz, epssd
      integer(c_int) :: indspe(*), natpri(*), naps(*), nprj(*), natinf(*), lstvec2(*), latom(*)
      integer(c_int) :: natsoc(*), ntyppp(*), lstx(*), lsty(*), lstz(*), natx(*), naty(*), natz(*)
      real(c_double) :: dij(*), dijsoc(*), veff(*), vnlocp(*), sss(*), sval(*), residual_states(*)
      complex(c_double_complex) :: cspsep(*), sveccm(*), ssvcm(*), hsvcm(*)
      integer(c_int) :: ksconv, ksitmax
    end function
    integer(c_int) function c_ksdiis_c( &
    l_hsv, nso, natom, num_spe, neigmx, nums, ncol, nspv, numk, nperi, ndiismax, nprecon_diis, nk, ns1, &
        nf, nprjmx, num_list, num_ppcell, ndisp, northo, ncpx, ncpy, ncpz, npxmax, npymax, npzmax, &
        key_ortho_cmpt_innerproduct, key_natpri_in, key_natpri_inps, key_soc_calc, eps_eig_diis, &
        alambda_diis, ratio_diis, alambda_min, alambda_max, epssd, xmax, ymax, zmax, skpx, skpy, skpz, &
        indspe, natpri, naps, nprj, natinf, lstvec2, latom, natsoc, ntyppp, lstx, lsty, lstz, natx, &
        naty, natz, dij, dijsoc, veff, vnlocp, sss, ksconv, ksitmax, cspsep, sval, sveccm, ssvcm, &
        residual_states, hsvcm) &
        bind(C, name='kukan_ksdiis_c')
      import :: c_int, c_double, c_double_complex
      integer(c_int), value :: l_hsv, nso, natom, num_spe, neigmx, nums, ncol, nspv, numk, nperi, ndiismax
      integer(c_int), value :: nprecon_diis, nk, ns1, nf, nprjmx, num_list, num_ppcell, ndisp, northo, ncpx
      integer(c_int), value :: ncpy, ncpz, npxmax, npymax, npzmax, key_ortho_cmpt_innerproduct, key_natpri_in
      integer(c_int), value :: key_natpri_inps, key_soc_calc
      real(c_double), value :: eps_eig_diis, alambda_diis, ratio_diis, alambda_min, alambda_max, epssd
      real(c_double), value :: xmax, ymax, zmax, skpx, skpy, skpz
      integer(c_int) :: indspe(*), natpri(*), naps(*), nprj(*), natinf(*), lstvec2(*), latom(*)
      integer(c_int) :: natsoc(*), ntyppp(*), lstx(*), lsty(*), lstz(*), natx(*), naty(*), natz(*)
      real(c_double) :: dij(*), dijsoc(*), veff(*), vnlocp(*), sss(*), sval(*), residual_states(*)
      complex(c_double_complex) :: cspsep(*), sveccm(*), ssvcm(*), hsvcm(*)
      integer(c_int) :: ksconv, ksitmax
    end function
    integer(c_int) function c_orthogonalization_c( &
    natom, neigmx, nprjmx, nums, ncol, numk, num_ppcell, num_spe, northo, num_list, nflag, ndisp, ncpx, &
        ncpy, ncpz, npxmax, npymax, npzmax, key_ortho_cmpt_innerproduct, key_natpri_in, &
        key_natpri_inps, dx, dy, dz, nprj, indspe, natpri, naps, natinf, lstvec2, latom, ntyppp, lstx, &
        lsty, lstz, natx, naty, natz, skpxx, skpyy, skpzz, vnlocp, sss, sveccm, ssvcm, cspsep) &
        bind(C, name='kukan_orthogonalization_c')
      import :: c_int, c_double, c_double_complex
      integer(c_int), value :: natom, neigmx, nprjmx, nums, ncol, numk, num_ppcell, num_spe, northo, num_list
      integer(c_int), value :: nflag, ndisp, ncpx, ncpy, ncpz, npxmax, npymax, npzmax, key_ortho_cmpt_innerproduct
      integer(c_int), value :: key_natpri_in, key_natpri_inps
      real(c_double), value :: dx, dy, dz
      integer(c_int) :: nprj(*), indspe(*), natpri(*), naps(*), natinf(*), lstvec2(*), latom(*)
      integer(c_int) :: ntyppp(*), lstx(*), lsty(*), lstz(*), natx(*), naty(*), natz(*)
      real(c_double) :: skpxx(*), skpyy(*), skpzz(*), vnlocp(*), sss(*)
      complex(c_double_complex) :: sveccm(*), ssvcm(*), cspsep(*)
    end function
    integer(c_int) function c_orthogonalization_c_01( &
    natom, neigmx, nprjmx, nums, ncol, numk, nk, ns1, l_inp, num_list, num_ppcell, num_spe, nflag, &
        ndisp, ncpx, ncpy, ncpz, npxmax, npymax, npzmax, key_ortho_cmpt_innerproduct, key_natpri_in, &
        key_natpri_inps, dx, dy, dz, skpx, skpy, skpz, nprj, indspe, natpri, naps, natinf, lstvec2, &
        latom, lstx, lsty, lstz, natx, naty, natz, vnlocp, sss, sveccm, ssvcm, cspsep) &
        bind(C, name='kukan_orthogonalization_c_01')
      import :: c_int, c_double, c_double_complex
      integer(c_int), value :: natom, neigmx, nprjmx, nums, ncol, numk, nk, ns1, l_inp, num_list, num_ppcell
      integer(c_int), value :: num_spe, nflag, ndisp, ncpx, ncpy, ncpz, npxmax, npymax, npzmax, key_ortho_cmpt_innerproduct
      integer(c_int), value :: key_natpri_in, key_natpri_inps
      real(c_double), value :: dx, dy, dz, skpx, skpy, skpz
      integer(c_int) :: nprj(*), indspe(*), natpri(*), naps(*), natinf(*), lstvec2(*), latom(*)
      integer(c_int) :: lstx(*), lsty(*), lstz(*), natx(*), naty(*), natz(*)
      real(c_double) :: vnlocp(*), sss(*)
      complex(c_double_complex) :: sveccm(*), ssvcm(*), cspsep(*)
    end function
    integer(c_int) function c_rayleighritz_c( &
    nso, natom, num_ppcell, num_spe, nprjmx, neigmx, numk, nums, ncol, nspv, num_list, ncpx, ncpy, &
        ncpz, npxmax, npymax, npzmax, nperi, nk, ns1, nf, ndisp, key_natpri_in, key_natpri_inps, &
        key_soc_calc, dx, dy, dz, xmax, ymax, zmax, skpxx, skpyy, skpzz, nprj, natinf, lstvec2, indspe, &
        natpri, naps, natsoc, ntyppp, lstx, lsty, lstz, natx, naty, natz, dij, dijsoc, veff, vnlocp, &
        l_rrb, cspsep, sveccm, ssvcm, hsvcm, sval) &
        bind(C, name='kukan_rayleighritz_c')
      import :: c_int, c_double, c_double_complex
      integer(c_int), value :: nso, natom, num_ppcell, num_spe, nprjmx, neigmx, numk, nums, ncol, nspv
      integer(c_int), value :: num_list, ncpx, ncpy, ncpz, npxmax, npymax, npzmax, nperi, nk, ns1, nf, ndisp
      integer(c_int), value :: key_natpri_in, key_natpri_inps, key_soc_calc, l_rrb
      real(c_double), value :: dx, dy, dz, xmax, ymax, zmax
      integer(c_int) :: nprj(*), natinf(*), lstvec2(*), indspe(*), natpri(*), naps(*), natsoc(*)
      integer(c_int) :: ntyppp(*), lstx(*), lsty(*), lstz(*), natx(*), naty(*), natz(*)
      real(c_double) :: skpxx(*), skpyy(*), skpzz(*), dij(*), dijsoc(*), veff(*), vnlocp(*), sval(*)
      complex(c_double_complex) :: cspsep(*), sveccm(*), ssvcm(*), hsvcm(*)
    end function
    integer(c_int) function c_scf_charge_smt_c( &
    neigmx, nums, ncol, nspv, numk, ncpx, ncpy, ncpz, fnele, sveccm, rhosmt) &
        bind(C, name='kukan_scf_charge_smt_c')
      import :: c_int, c_double, c_double_complex
      integer(c_int), value :: neigmx, nums, ncol, nspv, numk, ncpx, ncpy, ncpz
      real(c_double) :: fnele(*), rhosmt(*)
      complex(c_double_complex) :: sveccm(*)
    end function
    integer(c_int) function c_pseudocalc_01a(nfdg, npmesh, wxyz) &
        bind(C, name='kukan_pseudocalc_01a')
      import :: c_int, c_double
      integer(c_int), value :: nfdg, npmesh
      real(c_double) :: wxyz(*)
    end function
    integer(c_int) function c_pseudocalc_01b(nperi, natom, num_spe, nradmx, nmesh, nfdg, ndisp, &
        num_atcell, num_ppcell, num_ppcell_d, num_list, num_list_d, nxmax, nymax, nzmax, ncpx, ncpy, &
        ncpz, npxmax, npymax, npzmax, key_natpri_in, key_natpri_inps, key_natpri_out, psctoff, xmax, &
        ymax, zmax, dx, dy, dz, indspe, nradct, radial, atx, aty, atz, natx, naty, natz, ndatx, &
        ndaty, ndatz, lstx, lsty, lstz, lstdx, lstdy, lstdz, natinf, natinfd, natinfd_vloc, lstvec2, &
        lstvecd2, natpri, natprid, natpri_inf, naps, napsd) &
        bind(C, name='kukan_pseudocalc_01b')
      import :: c_int, c_double
      integer(c_int), value :: nperi, natom, num_spe, nradmx, nmesh, nfdg, ndisp, num_atcell, &
        num_ppcell, num_ppcell_d, num_list, num_list_d, nxmax, nymax, nzmax, ncpx, ncpy, ncpz, &
        npxmax, npymax, npzmax, key_natpri_in, key_natpri_inps, key_natpri_out
      real(c_double), value :: psctoff, xmax, ymax, zmax, dx, dy, dz
      integer(c_int) :: indspe(*), nradct(*), natx(*), naty(*), natz(*), ndatx(*), ndaty(*), &
        ndatz(*), lstx(*), lsty(*), lstz(*), lstdx(*), lstdy(*), lstdz(*), natinf(*), natinfd(*), &
        natinfd_vloc(*), lstvec2(*), lstvecd2(*), natpri(*), natprid(*), natpri_inf(*), naps(*), &
        napsd(*)
      real(c_double) :: radial(*), atx(*), aty(*), atz(*)
    end function
    integer(c_int) function c_pseudocalc_07(natom, num_spe, npmesh, nprjmx, nfdg, nfiltyp, num_list, &
        num_ppcell, nradmx, nqmx, nxmax, nymax, nzmax, npxmax, npymax, npzmax, nxspd, nyspd, nzspd, &
        key_natpri_in, key_natpri_inps, eps, psftrad, psctoff, filpp, dx, dy, dz, indspe, nprj, &
        nradct, natpri, naps, natinf, nqct, natx, naty, natz, lstx, lsty, lstz, wxyz, radial, coef, &
        atx, aty, atz, vnlocp, vtmp_pp, vtmp_pp1, vtmp_pp2, vnlspd) &
        bind(C, name='kukan_pseudocalc_07')
      import :: c_int, c_double
      integer(c_int), value :: natom, num_spe, npmesh, nprjmx, nfdg, nfiltyp, num_list, num_ppcell, &
        nradmx, nqmx, nxmax, nymax, nzmax, npxmax, npymax, npzmax, nxspd, nyspd, nzspd, &
        key_natpri_in, key_natpri_inps
      real(c_double), value :: eps, psftrad, psctoff, filpp, dx, dy, dz
      integer(c_int) :: indspe(*), nprj(*), nradct(*), natpri(*), naps(*), natinf(*), nqct(*), &
        natx(*), naty(*), natz(*), lstx(*), lsty(*), lstz(*)
      real(c_double) :: wxyz(*), radial(*), coef(*), atx(*), aty(*), atz(*), vnlocp(*), vtmp_pp(*), &
        vtmp_pp1(*), vtmp_pp2(*), vnlspd(*)
    end function
    integer(c_int) function c_force(natom, num_spe, num_atcell, num_ppcell_d, num_list_d, nqmx, &
        nradmx, npoint, nmesh, nspv, nperi, nint1dmax, nzmax, jelcalc, ncpx, ncpy, ncpz, ncpx_d, &
        ncpy_d, ncpz_d, new_pwx, new_pwy, new_pwz, new_rsx, new_rsy, new_rsz, key_jel_calc, &
        key_natpri_in, key_natpri_inps, indspe, natpri, natpri_inf, nradct, natprid, napsd, natinfd, &
        natinfd_vloc, lstvecd2, nqctpcc, lstdx, lstdy, lstdz, veta, psctoff, psftrad, radial, &
        dradial, point, wt, xmax, ymax, zmax, biasx, biasy, biasz, cp, coef, atx, aty, atz, &
        dvlocdx_scw, dvlocdy_scw, dvlocdz_scw, dvlocdx_hdp, dvlocdy_hdp, dvlocdz_hdp, rhotrur, &
        rhosmtr, rhoaugr, drhoaug3ddx, drhoaug3ddy, drhoaug3ddz, rho_coarse, rho_aug_dense, &
        vloc_dense, vh_dense, vx_dense, chrjel, strjel, endjel, ndatx, ndaty, ndatz, fatx, faty, &
        fatz, fjel) &
        bind(C, name='kukan_force')
      import :: c_int, c_double
      integer(c_int), value :: natom, num_spe, num_atcell, num_ppcell_d, num_list_d, nqmx, nradmx, &
        npoint, nmesh, nspv, nperi, nint1dmax, nzmax, jelcalc, ncpx, ncpy, ncpz, ncpx_d, ncpy_d, &
        ncpz_d, new_pwx, new_pwy, new_pwz, new_rsx, new_rsy, new_rsz, key_jel_calc, key_natpri_in, &
        key_natpri_inps
      real(c_double), value :: veta, psctoff, psftrad, xmax, ymax, zmax, biasx, biasy, biasz, chrjel, &
        strjel, endjel
      integer(c_int) :: indspe(*), natpri(*), natpri_inf(*), nradct(*), natprid(*), napsd(*), &
        natinfd(*), natinfd_vloc(*), lstvecd2(*), nqctpcc(*), lstdx(*), lstdy(*), lstdz(*), ndatx(*), &
        ndaty(*), ndatz(*)
      real(c_double) :: radial(*), dradial(*), point(*), wt(*), cp(*), coef(*), atx(*), aty(*), &
        atz(*), dvlocdx_scw(*), dvlocdy_scw(*), dvlocdz_scw(*), dvlocdx_hdp(*), dvlocdy_hdp(*), &
        dvlocdz_hdp(*), rhotrur(*), rhosmtr(*), rhoaugr(*), drhoaug3ddx(*), drhoaug3ddy(*), &
        drhoaug3ddz(*), rho_coarse(*), rho_aug_dense(*), vloc_dense(*), vh_dense(*), vx_dense(*), &
        fatx(*), faty(*), fatz(*)
      real(c_double) :: fjel
    end function
  end interface

contains

  subroutine kukan_check(rc)
#ifdef KUKAN_USE_STOPP
    use mod_stopp, only: stopp
#endif
    integer(c_int), intent(in) :: rc
    character(len=512) :: msg
    character(kind=c_char), pointer :: cs(:)
    integer :: i
    if (rc == 0) return
    call c_f_pointer(kukan_last_error(), cs, [512])
    msg = ' '
    do i = 1, 512
      if (cs(i) == c_null_char) exit
      msg(i:i) = cs(i)
    end do
#ifdef KUKAN_USE_STOPP
    call stopp(trim(msg))
#else
    write(0, '(a)') trim(msg)
    error stop 1
#endif
  end subroutine kukan_check

  subroutine kukan_setup(device)
    integer, intent(in) :: device
    call kukan_check(kukan_init(int(device, c_int)))
  end subroutine kukan_setup

  subroutine kukan_option(key, val)
    character(len=*), intent(in) :: key
    real(8), intent(in) :: val
    call kukan_check(kukan_set_param(trim(key)//c_null_char, real(val, c_double)))
  end subroutine kukan_option

  integer(c_int) function l2i(l)
    logical, intent(in) :: l
    l2i = 0
    if (l) l2i = 1
  end function l2i

end module kukan_hip_c


! ------------------------------------------------------------- tools8f.F90:144-237, 354-378
module mod_kukan_tools
  use kukan_hip_c
  implicit none
contains
  subroutine tools_definefdcoef(nf, acfd0, acfd1, acfd2, acfd3, acfd4, acfd5, acfd6, acfd7, acfd8)
    integer nf
    real*8 acfd0, acfd1, acfd2, acfd3, acfd4, acfd5, acfd6, acfd7, acfd8
    real(c_double) :: a(0:8)
    call kukan_check(c_definefdcoef(int(nf, c_int), a))
    acfd0 = a(0); acfd1 = a(1); acfd2 = a(2); acfd3 = a(3); acfd4 = a(4)
    acfd5 = a(5); acfd6 = a(6); acfd7 = a(7); acfd8 = a(8)
  end subroutine tools_definefdcoef

  subroutine tools_correctpotential(ncpx_d, ncpy_d, ncpz_d, xmax, ymax, zmax, ddx, ddy, ddz, vh_dense)
    integer, intent(in) :: ncpx_d, ncpy_d, ncpz_d
    real*8, intent(in) :: xmax, ymax, zmax, ddx, ddy, ddz
    real*8, intent(inout) :: vh_dense(ncpx_d, ncpy_d, ncpz_d)
    call kukan_check(c_correctpotential(ncpx_d, ncpy_d, ncpz_d, xmax, ymax, zmax, ddx, ddy, ddz, vh_dense))
  end subroutine tools_correctpotential
end module mod_kukan_tools


! ------------------------------------------------------------- overlap_finitedifference8f.F90
module mod_overlap_finitedifference_hip
  use kukan_hip_c
  implicit none
contains
  subroutine overlap_finitedifference_init(ncpa, ncpb, ncpc, nf, ncol, nrc)
    integer ncpa, ncpb, ncpc, nf, ncol, nrc
  end subroutine
  subroutine overlap_finitedifference_final
  end subroutine
  ! :36 -- pack + local periodic wrap; the wrap completes here on one GPU
  subroutine overlap_finitedifference_r(nperi, ncpa, ncpb, ncpc, nf, nf1, nsnd, v_cc)
    integer nperi, ncpa, ncpb, ncpc, nf, nf1, nsnd
    real*8 v_cc(-nf1:ncpa+nf, -nf1:ncpb+nf, -nf1:ncpc+nf)
    call kukan_check(c_overlap_fd(nperi, ncpa, ncpb, ncpc, nf, nf1, nsnd, v_cc))
  end subroutine
  ! :648 -- nothing left to wait for
  subroutine overlap_fdcheck_r(nperi, ncpa, ncpb, ncpc, nf, nf1, nsnd, v_cc)
    integer nperi, ncpa, ncpb, ncpc, nf, nf1, nsnd
    real*8 v_cc(-nf1:ncpa+nf, -nf1:ncpb+nf, -nf1:ncpc+nf)
  end subroutine
end module mod_overlap_finitedifference_hip


! ------------------------------------------------------------- kslaplacian8f.F90:9-557
module mod_kslaplacian_hip
  use kukan_hip_c
  implicit none
contains
  subroutine kslaplacian_initialize(nf)
    integer nf
  end subroutine
  subroutine kslaplacian_finalize
  end subroutine
  subroutine kslaplacian_r(ncpx, ncpy, ncpz, neigmx, nf, dx, dy, dz, vre, veff, l, avre)
    integer, intent(in) :: ncpx, ncpy, ncpz, neigmx, l, nf
    real*8, intent(in) :: dx, dy, dz
    real*8, intent(in) :: vre(-(nf-1):ncpx+nf, -(nf-1):ncpy+nf, -(nf-1):ncpz+nf)
    real*8, intent(in) :: veff(ncpx, ncpy, ncpz)
    real*8, intent(inout) :: avre(ncpx, ncpy, ncpz, neigmx)
    call kukan_check(c_kslaplacian_r(ncpx, ncpy, ncpz, neigmx, nf, dx, dy, dz, vre, veff, l, avre))
  end subroutine
end module mod_kslaplacian_hip


! ------------------------------------------------------------- ksprecondition8f.F90:8-39
module mod_ksprecondition_hip
  use kukan_hip_c
  implicit none
contains
  subroutine ksprecondition(ncpx, ncpy, ncpz, nf, ncol, vre, avre)
    integer, intent(in) :: ncpx, ncpy, ncpz, nf, ncol
    real*8, intent(in) :: vre(-(nf-1):ncpx+nf, -(nf-1):ncpy+nf, -(nf-1):ncpz+nf, ncol)
    real*8, intent(out) :: avre(ncpx, ncpy, ncpz, ncol)
    call kukan_check(c_ksprecondition(ncpx, ncpy, ncpz, nf, ncol, vre, avre))
  end subroutine
end module mod_ksprecondition_hip


! ------------------------------------------------------------- nonlocaloperation8f.F90:9-235
module mod_nonlocaloperation_hip
  use kukan_hip_c
  implicit none
contains
  subroutine nonlocaloperation_r_01(natom, nprjmx, num_spe, num_ppcell, num_list, ncpx, ncpy, ncpz, &
                                    key_natpri_in, key_natpri_inps, nprj, indspe, natinf, natpri, &
                                    naps, lstvec2, vnlocp, vre2, rpsep)
    integer, intent(in) :: natom, nprjmx, num_spe, num_ppcell, num_list, ncpx, ncpy, ncpz
    integer, intent(in) :: key_natpri_in, key_natpri_inps
    integer, intent(in) :: nprj(num_spe), indspe(natom), natinf(natom), lstvec2(num_list, num_ppcell)
    integer, intent(in) :: natpri(natom), naps(natom)
    real*8, intent(in) :: vnlocp(num_list, nprjmx, num_ppcell), vre2(ncpx, ncpy, ncpz)
    real*8, intent(inout) :: rpsep(nprjmx, num_ppcell)
    call kukan_check(c_nonlocal_r_01(natom, nprjmx, num_spe, num_ppcell, num_list, ncpx, ncpy, ncpz, &
                     key_natpri_in, key_natpri_inps, nprj, indspe, natinf, natpri, naps, lstvec2, &
                     vnlocp, vre2, rpsep))
  end subroutine

  subroutine nonlocaloperation_r_02(nhs, natom, num_spe, nums, nprjmx, ns, num_list, num_ppcell, ncpx, &
                                    ncpy, ncpz, key_natpri_in, key_natpri_inps, dx, dy, dz, nprj, &
                                    indspe, natinf, lstvec2, natpri, naps, vnlocp, rpsep, dijsss, avre, avr)
    integer, intent(in) :: nhs, natom, num_spe, nums, nprjmx, ns, num_list, num_ppcell, ncpx, ncpy, ncpz
    integer, intent(in) :: key_natpri_in, key_natpri_inps
    real*8, intent(in) :: dx, dy, dz
    integer, intent(in) :: nprj(num_spe), indspe(natom), natinf(natom), lstvec2(num_list, num_ppcell)
    integer, intent(in) :: natpri(natom), naps(natom)
    real*8, intent(in) :: vnlocp(num_list, nprjmx, num_ppcell), rpsep(nprjmx, num_ppcell)
    real*8, intent(in) :: dijsss(nprjmx, nprjmx, nums*nhs-nhs+1, natom*nhs+num_spe*(1-nhs))
    real*8, intent(inout) :: avre(ncpx, ncpy, ncpz)
    real*8, intent(inout) :: avr(num_list)
    call kukan_check(c_nonlocal_r_02(nhs, natom, num_spe, nums, nprjmx, ns, num_list, num_ppcell, ncpx, &
                     ncpy, ncpz, key_natpri_in, key_natpri_inps, dx, dy, dz, nprj, indspe, natinf, &
                     lstvec2, natpri, naps, vnlocp, rpsep, dijsss, avre, avr))
  end subroutine
end module mod_nonlocaloperation_hip


! ------------------------------------------------------------- kscg8f.F90:7-489
module mod_kscg_hip
  use kukan_hip_c
  implicit none
contains
  subroutine kscg_r(natom, num_spe, neigmx, nums, nspv, numk, nperi, northo, ns, nk, nf, nkscg, nsdmax, &
                    nprecon_cg, nprjmx, num_list, num_ppcell, ndisp, ncpx, ncpy, ncpz, &
                    key_ortho_nocmpt_innerproduct, key_ortho_cmpt_innerproduct, key_natpri_in, &
                    key_natpri_inps, key_natpri_out, dx, dy, dz, epssd, indspe, natpri, naps, nprj, &
                    natinf, lstvec2, latom, ntyppp, dij, veff, vnlocp, sss, ksconv, ksitmax, sval, &
                    rspsep, svecre, ssvre, residual_states, hsvre)
    integer, intent(in) :: natom, num_spe, neigmx, nums, nspv, numk, nperi, northo, ns, nk, nf, nkscg, nsdmax
    integer, intent(in) :: nprecon_cg, nprjmx, num_list, num_ppcell, ndisp, ncpx, ncpy, ncpz
    integer, intent(in) :: key_ortho_nocmpt_innerproduct, key_ortho_cmpt_innerproduct
    integer, intent(in) :: key_natpri_in, key_natpri_inps, key_natpri_out
    real*8, intent(in) :: dx, dy, dz, epssd
    integer, intent(in) :: indspe(natom), natpri(natom), naps(natom), nprj(num_spe)
    integer, intent(in) :: natinf(natom), lstvec2(num_list, num_ppcell), latom(natom), ntyppp(num_spe)
    real*8, intent(in) :: dij(nprjmx, nprjmx, nums, natom), veff(ncpx, ncpy, ncpz, nspv)
    real*8, intent(in) :: vnlocp(num_list, nprjmx, num_ppcell), sss(nprjmx, nprjmx, num_spe)
    integer, intent(inout) :: ksitmax
    logical, intent(inout) :: ksconv
    real*8, intent(inout) :: sval(neigmx, nums, numk), rspsep(nprjmx, num_ppcell, neigmx, nums, numk)
    real*8, intent(inout) :: svecre(ncpx, ncpy, ncpz, neigmx, nums, numk)
    real*8, intent(inout) :: ssvre(ncpx, ncpy, ncpz, neigmx, nums, numk)
    real*8, intent(out) :: residual_states(neigmx), hsvre(ncpx, ncpy, ncpz, neigmx)
    integer(c_int) :: ic, it
    ic = l2i(ksconv); it = ksitmax
    call kukan_check(c_kscg_r(natom, num_spe, neigmx, nums, nspv, numk, nperi, northo, ns, nk, nf, nkscg, &
         nsdmax, nprecon_cg, nprjmx, num_list, num_ppcell, ndisp, ncpx, ncpy, ncpz, &
         key_ortho_nocmpt_innerproduct, key_ortho_cmpt_innerproduct, key_natpri_in, key_natpri_inps, &
         key_natpri_out, dx, dy, dz, epssd, indspe, natpri, naps, nprj, natinf, lstvec2, latom, ntyppp, &
         dij, veff, vnlocp, sss, ic, it, sval, rspsep, svecre, ssvre, residual_states, hsvre))
    ksconv = ic /= 0; ksitmax = it
  end subroutine
  ! kscg8f.F90:492-1032 (complex path, ncol = 1, nso = 0)
  subroutine kscg_c( &
      nso, natom, num_spe, neigmx, nums, ncol, nspv, numk, nperi, northo, ns1, nk, nf, nkscg, nsdmax, &
          nprecon_cg, nprjmx, num_list, num_ppcell, ndisp, ncpx, ncpy, ncpz, npxmax, npymax, npzmax, &
          key_ortho_nocmpt_innerproduct, key_ortho_cmpt_innerproduct, key_natpri_in, key_natpri_inps, &
          key_natpri_out, key_soc_calc, dx, dy, dz, skpx, skpy, skpz, epssd, indspe, natpri, naps, &
          nprj, natinf, lstvec2, latom, natsoc, ntyppp, lstx, lsty, lstz, natx, naty, natz, dij, &
          dijsoc, veff, vnlocp, sss, ksconv, ksitmax, sval, cspsep, sveccm, ssvcm, residual_states, &
          hsvcm)
    integer, intent(in) :: nso, natom, num_spe, neigmx, nums, ncol, nspv, numk, nperi, northo, ns1
    integer, intent(in) :: nk, nf, nkscg, nsdmax, nprecon_cg, nprjmx, num_list, num_ppcell, ndisp, ncpx
    integer, intent(in) :: ncpy, ncpz, npxmax, npymax, npzmax, key_ortho_nocmpt_innerproduct, key_ortho_cmpt_innerproduct
    integer, intent(in) :: key_natpri_in, key_natpri_inps, key_natpri_out, key_soc_calc
    real*8, intent(in) :: dx, dy, dz, skpx, skpy, skpz, epssd
    integer, intent(in) :: indspe(*), natpri(*), naps(*), nprj(*), natinf(*), lstvec2(*), latom(*)
    integer, intent(in) :: natsoc(*), ntyppp(*), lstx(*), lsty(*), lstz(*), natx(*), naty(*), natz(*)
    real*8, intent(in) :: dij(*), dijsoc(*), veff(*), vnlocp(*), sss(*)
    real*8, intent(inout) :: sval(*), residual_states(*)
    complex*16, intent(inout) :: cspsep(*), sveccm(*), ssvcm(*), hsvcm(*)
    integer, intent(inout) :: ksitmax
    logical, intent(inout) :: ksconv
    integer(c_int) :: ic, it
    ic = l2i(ksconv); it = ksitmax
    call kukan_check(c_kscg_c(nso, natom, num_spe, neigmx, nums, ncol, nspv, numk, nperi, northo, ns1, nk, nf, nkscg, &
             nsdmax, nprecon_cg, nprjmx, num_list, num_ppcell, ndisp, ncpx, ncpy, ncpz, npxmax, npymax, &
             npzmax, key_ortho_nocmpt_innerproduct, key_ortho_cmpt_innerproduct, key_natpri_in, &
             key_natpri_inps, key_natpri_out, key_soc_calc, dx, dy, dz, skpx, skpy, skpz, epssd, &
             indspe, natpri, naps, nprj, natinf, lstvec2, latom, natsoc, ntyppp, lstx, lsty, lstz, &
             natx, naty, natz, dij, dijsoc, veff, vnlocp, sss, ic, it, sval, cspsep, sveccm, ssvcm, &
             residual_states, hsvcm))
    ksconv = ic /= 0; ksitmax = it
  end subroutine
end module mod_kscg_hip


! ------------------------------------------------------------- ksdiis8f.F90:8-712
module mod_ksdiis_hip
  use kukan_hip_c
  implicit none
contains
  subroutine ksdiis_r(l_hsv, natom, num_spe, neigmx, nums, nspv, numk, nperi, ndiismax, nprecon_diis, &
                      nk, ns, nf, nprjmx, num_list, num_ppcell, ndisp, northo, ncpx, ncpy, ncpz, &
                      key_ortho_cmpt_innerproduct, key_natpri_in, key_natpri_inps, eps_eig_diis, &
                      alambda_diis, ratio_diis, alambda_min, alambda_max, epssd, xmax, ymax, zmax, &
                      indspe, natpri, naps, nprj, natinf, lstvec2, latom, ntyppp, dij, veff, vnlocp, &
                      sss, ksconv, ksitmax, rspsep, sval, svecre, ssvre, residual_states, hsvre)
    logical, intent(in) :: l_hsv
    integer, intent(in) :: natom, num_spe, neigmx, nums, nspv, numk, nperi, ndiismax
    integer, intent(in) :: nprecon_diis, nk, ns, nf, nprjmx, num_list, num_ppcell, ndisp, northo
    integer, intent(in) :: ncpx, ncpy, ncpz, key_ortho_cmpt_innerproduct, key_natpri_in, key_natpri_inps
    real*8, intent(in) :: eps_eig_diis, alambda_diis, ratio_diis, alambda_min, alambda_max, epssd
    real*8, intent(in) :: xmax, ymax, zmax
    integer, intent(in) :: indspe(natom), natpri(natom), naps(natom), nprj(num_spe)
    integer, intent(in) :: natinf(natom), lstvec2(num_list, num_ppcell), latom(natom), ntyppp(num_spe)
    real*8, intent(in) :: dij(nprjmx, nprjmx, nums, natom), veff(ncpx, ncpy, ncpz, nspv)
    real*8, intent(in) :: vnlocp(num_list, nprjmx, num_ppcell), sss(nprjmx, nprjmx, num_spe)
    logical, intent(inout) :: ksconv
    integer, intent(inout) :: ksitmax
    real*8, intent(inout) :: rspsep(nprjmx, num_ppcell, neigmx, nums, numk), sval(neigmx, nums, numk)
    real*8, intent(inout) :: svecre(ncpx, ncpy, ncpz, neigmx, nums, numk)
    real*8, intent(inout) :: ssvre(ncpx, ncpy, ncpz, neigmx, nums, numk)
    real*8, intent(inout) :: residual_states(neigmx)
    real*8, intent(out) :: hsvre(ncpx, ncpy, ncpz, neigmx)
    integer(c_int) :: ic, it
    ic = l2i(ksconv); it = ksitmax
    call kukan_check(c_ksdiis_r(l2i(l_hsv), natom, num_spe, neigmx, nums, nspv, numk, nperi, ndiismax, &
         nprecon_diis, nk, ns, nf, nprjmx, num_list, num_ppcell, ndisp, northo, ncpx, ncpy, ncpz, &
         key_ortho_cmpt_innerproduct, key_natpri_in, key_natpri_inps, eps_eig_diis, alambda_diis, &
         ratio_diis, alambda_min, alambda_max, epssd, xmax, ymax, zmax, indspe, natpri, naps, nprj, &
         natinf, lstvec2, latom, ntyppp, dij, veff, vnlocp, sss, ic, it, rspsep, sval, svecre, ssvre, &
         residual_states, hsvre))
    ksconv = ic /= 0; ksitmax = it
  end subroutine
  ! ksdiis8f.F90:715-1509 (complex path, ncol = 1, nso = 0)
  subroutine ksdiis_c( &
      l_hsv, nso, natom, num_spe, neigmx, nums, ncol, nspv, numk, nperi, ndiismax, nprecon_diis, nk, &
          ns1, nf, nprjmx, num_list, num_ppcell, ndisp, northo, ncpx, ncpy, ncpz, npxmax, npymax, &
          npzmax, key_ortho_cmpt_innerproduct, key_natpri_in, key_natpri_inps, key_soc_calc, &
          eps_eig_diis, alambda_diis, ratio_diis, alambda_min, alambda_max, epssd, xmax, ymax, zmax, &
          skpx, skpy, skpz, indspe, natpri, naps, nprj, natinf, lstvec2, latom, natsoc, ntyppp, lstx, &
          lsty, lstz, natx, naty, natz, dij, dijsoc, veff, vnlocp, sss, ksconv, ksitmax, cspsep, sval, &
          sveccm, ssvcm, residual_states, hsvcm)
    integer, intent(in) :: nso, natom, num_spe, neigmx, nums, ncol, nspv, numk, nperi, ndiismax, nprecon_diis
    integer, intent(in) :: nk, ns1, nf, nprjmx, num_list, num_ppcell, ndisp, northo, ncpx, ncpy, ncpz
    integer, intent(in) :: npxmax, npymax, npzmax, key_ortho_cmpt_innerproduct, key_natpri_in, key_natpri_inps
    integer, intent(in) :: key_soc_calc
    logical, intent(in) :: l_hsv
    real*8, intent(in) :: eps_eig_diis, alambda_diis, ratio_diis, alambda_min, alambda_max, epssd
    real*8, intent(in) :: xmax, ymax, zmax, skpx, skpy, skpz
    integer, intent(in) :: indspe(*), natpri(*), naps(*), nprj(*), natinf(*), lstvec2(*), latom(*)
    integer, intent(in) :: natsoc(*), ntyppp(*), lstx(*), lsty(*), lstz(*), natx(*), naty(*), natz(*)
    real*8, intent(in) :: dij(*), dijsoc(*), veff(*), vnlocp(*), sss(*)
    real*8, intent(inout) :: sval(*), residual_states(*)
    complex*16, intent(inout) :: cspsep(*), sveccm(*), ssvcm(*), hsvcm(*)
    integer, intent(inout) :: ksitmax
    logical, intent(inout) :: ksconv
    integer(c_int) :: ic, it
    ic = l2i(ksconv); it = ksitmax
    call kukan_check(c_ksdiis_c(l2i(l_hsv), nso, natom, num_spe, neigmx, nums, ncol, nspv, numk, nperi, ndiismax, &
             nprecon_diis, nk, ns1, nf, nprjmx, num_list, num_ppcell, ndisp, northo, ncpx, ncpy, ncpz, &
             npxmax, npymax, npzmax, key_ortho_cmpt_innerproduct, key_natpri_in, key_natpri_inps, &
             key_soc_calc, eps_eig_diis, alambda_diis, ratio_diis, alambda_min, alambda_max, epssd, &
             xmax, ymax, zmax, skpx, skpy, skpz, indspe, natpri, naps, nprj, natinf, lstvec2, latom, &
             natsoc, ntyppp, lstx, lsty, lstz, natx, naty, natz, dij, dijsoc, veff, vnlocp, sss, ic, &
             it, cspsep, sval, sveccm, ssvcm, residual_states, hsvcm))
    ksconv = ic /= 0; ksitmax = it
  end subroutine
end module mod_ksdiis_hip


! ------------------------------------------------------------- orthogonalization8f.F90:8-376
module mod_orthogonalization_hip
  use kukan_hip_c
  implicit none
contains
  subroutine orthogonalization_r(natom, neigmx, nprjmx, nums, numk, num_ppcell, num_spe, northo, &
                                 num_list, nflag, ndisp, ncpx, ncpy, ncpz, key_ortho_cmpt_innerproduct, &
                                 key_natpri_in, key_natpri_inps, dx, dy, dz, nprj, indspe, natpri, naps, &
                                 natinf, lstvec2, latom, ntyppp, vnlocp, sss, svecre, ssvre, rspsep)
    integer, intent(in) :: natom, neigmx, nprjmx, nums, numk, num_ppcell, num_spe, northo, num_list
    integer, intent(in) :: nflag, ndisp, ncpx, ncpy, ncpz, key_ortho_cmpt_innerproduct
    integer, intent(in) :: key_natpri_in, key_natpri_inps
    real*8, intent(in) :: dx, dy, dz
    integer, intent(in) :: indspe(natom), natpri(natom), naps(natom), nprj(num_spe)
    integer, intent(in) :: natinf(natom), lstvec2(num_list, num_ppcell), latom(natom), ntyppp(num_spe)
    real*8, intent(in) :: vnlocp(num_list, nprjmx, num_ppcell), sss(nprjmx, nprjmx, num_spe)
    real*8, intent(inout) :: svecre(ncpx, ncpy, ncpz, neigmx, nums, numk)
    real*8, intent(inout) :: ssvre(ncpx, ncpy, ncpz, neigmx, nums, numk)
    real*8, intent(inout) :: rspsep(nprjmx, num_ppcell, neigmx, nums, numk)
    call kukan_check(c_ortho_r(natom, neigmx, nprjmx, nums, numk, num_ppcell, num_spe, northo, num_list, &
         nflag, ndisp, ncpx, ncpy, ncpz, key_ortho_cmpt_innerproduct, key_natpri_in, key_natpri_inps, &
         dx, dy, dz, nprj, indspe, natpri, naps, natinf, lstvec2, latom, ntyppp, vnlocp, sss, svecre, &
         ssvre, rspsep))
  end subroutine

  subroutine orthogonalization_r_01(natom, neigmx, nprjmx, nums, numk, nk, ns, l_inp, num_list, &
                                    num_ppcell, num_spe, nflag, ndisp, ncpx, ncpy, ncpz, &
                                    key_ortho_cmpt_innerproduct, key_natpri_in, key_natpri_inps, dx, dy, &
                                    dz, nprj, indspe, natpri, naps, natinf, lstvec2, latom, vnlocp, sss, &
                                    svecre, ssvre, rspsep)
    integer, intent(in) :: natom, neigmx, nprjmx, nums, numk, nk, ns, l_inp, num_list, num_ppcell
    integer, intent(in) :: num_spe, nflag, ndisp, ncpx, ncpy, ncpz, key_ortho_cmpt_innerproduct
    integer, intent(in) :: key_natpri_in, key_natpri_inps
    integer, intent(in) :: indspe(natom), natpri(natom), naps(natom), nprj(num_spe)
    integer, intent(in) :: natinf(natom), lstvec2(num_list, num_ppcell), latom(natom)
    real*8, intent(in) :: dx, dy, dz
    real*8, intent(in) :: vnlocp(num_list, nprjmx, num_ppcell), sss(nprjmx, nprjmx, num_spe)
    real*8, intent(inout) :: svecre(ncpx, ncpy, ncpz, neigmx, nums, numk)
    real*8, intent(inout) :: ssvre(ncpx, ncpy, ncpz, neigmx, nums, numk)
    real*8, intent(inout) :: rspsep(nprjmx, num_ppcell, neigmx, nums, numk)
    call kukan_check(c_ortho_r_01(natom, neigmx, nprjmx, nums, numk, nk, ns, l_inp, num_list, num_ppcell, &
         num_spe, nflag, ndisp, ncpx, ncpy, ncpz, key_ortho_cmpt_innerproduct, key_natpri_in, &
         key_natpri_inps, dx, dy, dz, nprj, indspe, natpri, naps, natinf, lstvec2, latom, vnlocp, sss, &
         svecre, ssvre, rspsep))
  end subroutine
  ! orthogonalization8f.F90:47-90, 379-724 (complex path, ncol = 1)
  subroutine orthogonalization_c( &
      natom, neigmx, nprjmx, nums, ncol, numk, num_ppcell, num_spe, northo, num_list, nflag, ndisp, &
          ncpx, ncpy, ncpz, npxmax, npymax, npzmax, key_ortho_cmpt_innerproduct, key_natpri_in, &
          key_natpri_inps, dx, dy, dz, nprj, indspe, natpri, naps, natinf, lstvec2, latom, ntyppp, &
          lstx, lsty, lstz, natx, naty, natz, skpxx, skpyy, skpzz, vnlocp, sss, sveccm, ssvcm, cspsep)
    integer, intent(in) :: natom, neigmx, nprjmx, nums, ncol, numk, num_ppcell, num_spe, northo, num_list
    integer, intent(in) :: nflag, ndisp, ncpx, ncpy, ncpz, npxmax, npymax, npzmax, key_ortho_cmpt_innerproduct
    integer, intent(in) :: key_natpri_in, key_natpri_inps
    real*8, intent(in) :: dx, dy, dz
    integer, intent(in) :: nprj(*), indspe(*), natpri(*), naps(*), natinf(*), lstvec2(*), latom(*)
    integer, intent(in) :: ntyppp(*), lstx(*), lsty(*), lstz(*), natx(*), naty(*), natz(*)
    real*8, intent(in) :: skpxx(*), skpyy(*), skpzz(*), vnlocp(*), sss(*)
    complex*16, intent(inout) :: sveccm(*), ssvcm(*), cspsep(*)
    call kukan_check(c_orthogonalization_c(natom, neigmx, nprjmx, nums, ncol, numk, num_ppcell, num_spe, northo, num_list, nflag, ndisp, &
             ncpx, ncpy, ncpz, npxmax, npymax, npzmax, key_ortho_cmpt_innerproduct, key_natpri_in, &
             key_natpri_inps, dx, dy, dz, nprj, indspe, natpri, naps, natinf, lstvec2, latom, ntyppp, &
             lstx, lsty, lstz, natx, naty, natz, skpxx, skpyy, skpzz, vnlocp, sss, sveccm, ssvcm, cspsep))
  end subroutine
  subroutine orthogonalization_c_01( &
      natom, neigmx, nprjmx, nums, ncol, numk, nk, ns1, l_inp, num_list, num_ppcell, num_spe, nflag, &
          ndisp, ncpx, ncpy, ncpz, npxmax, npymax, npzmax, key_ortho_cmpt_innerproduct, key_natpri_in, &
          key_natpri_inps, dx, dy, dz, skpx, skpy, skpz, nprj, indspe, natpri, naps, natinf, lstvec2, &
          latom, lstx, lsty, lstz, natx, naty, natz, vnlocp, sss, sveccm, ssvcm, cspsep)
    integer, intent(in) :: natom, neigmx, nprjmx, nums, ncol, numk, nk, ns1, l_inp, num_list, num_ppcell
    integer, intent(in) :: num_spe, nflag, ndisp, ncpx, ncpy, ncpz, npxmax, npymax, npzmax, key_ortho_cmpt_innerproduct
    integer, intent(in) :: key_natpri_in, key_natpri_inps
    real*8, intent(in) :: dx, dy, dz, skpx, skpy, skpz
    integer, intent(in) :: nprj(*), indspe(*), natpri(*), naps(*), natinf(*), lstvec2(*), latom(*)
    integer, intent(in) :: lstx(*), lsty(*), lstz(*), natx(*), naty(*), natz(*)
    real*8, intent(in) :: vnlocp(*), sss(*)
    complex*16, intent(inout) :: sveccm(*), ssvcm(*), cspsep(*)
    call kukan_check(c_orthogonalization_c_01(natom, neigmx, nprjmx, nums, ncol, numk, nk, ns1, l_inp, num_list, num_ppcell, num_spe, nflag, &
             ndisp, ncpx, ncpy, ncpz, npxmax, npymax, npzmax, key_ortho_cmpt_innerproduct, &
             key_natpri_in, key_natpri_inps, dx, dy, dz, skpx, skpy, skpz, nprj, indspe, natpri, naps, &
             natinf, lstvec2, latom, lstx, lsty, lstz, natx, naty, natz, vnlocp, sss, sveccm, ssvcm, &
             cspsep))
  end subroutine
end module mod_orthogonalization_hip


! ------------------------------------------------------------- rayleigh-ritz8f.F90:9-125
module mod_rayleighritz_hip
  use kukan_hip_c
  implicit none
contains
  subroutine rayleighritz_blacs_ini(neigmx, nrc)
    integer, intent(in) :: neigmx, nrc
  end subroutine
  subroutine rayleighritz_blacs_fin
  end subroutine
  subroutine rayleighritz_r(natom, num_ppcell, num_spe, nprjmx, neigmx, numk, nums, nspv, num_list, ncpx, &
                            ncpy, ncpz, nperi, nk, ns, nf, ndisp, key_natpri_in, key_natpri_inps, dx, dy, &
                            dz, nprj, natinf, lstvec2, indspe, natpri, naps, ntyppp, dij, veff, vnlocp, &
                            l_rrb, rspsep, svecre, ssvre, hsvre, sval)
    integer, intent(in) :: natom, num_ppcell, num_spe, nprjmx, neigmx, numk, nums, nspv
    integer, intent(in) :: num_list, ncpx, ncpy, ncpz, nperi, nk, ns, nf, ndisp
    integer, intent(in) :: key_natpri_in, key_natpri_inps
    real*8, intent(in) :: dx, dy, dz
    integer, intent(in) :: nprj(num_spe), natinf(natom), lstvec2(num_list, num_ppcell)
    integer, intent(in) :: indspe(natom), natpri(natom), naps(natom), ntyppp(num_spe)
    real*8, intent(in) :: dij(nprjmx, nprjmx, nums, natom), veff(ncpx, ncpy, ncpz, nspv)
    real*8, intent(in) :: vnlocp(num_list, nprjmx, num_ppcell)
    logical, intent(in) :: l_rrb
    real*8, intent(inout) :: svecre(ncpx, ncpy, ncpz, neigmx, nums, numk)
    real*8, intent(inout) :: rspsep(nprjmx, num_ppcell, neigmx, nums, numk)
    real*8, intent(inout) :: ssvre(ncpx, ncpy, ncpz, neigmx, nums, numk)
    real*8, intent(inout) :: hsvre(ncpx, ncpy, ncpz, neigmx)
    real*8, intent(out) :: sval(neigmx, nums, numk)
    call kukan_check(c_rayleighritz_r(natom, num_ppcell, num_spe, nprjmx, neigmx, numk, nums, nspv, &
         num_list, ncpx, ncpy, ncpz, nperi, nk, ns, nf, ndisp, key_natpri_in, key_natpri_inps, dx, dy, &
         dz, nprj, natinf, lstvec2, indspe, natpri, naps, ntyppp, dij, veff, vnlocp, l2i(l_rrb), rspsep, &
         svecre, ssvre, hsvre, sval))
  end subroutine
  ! rayleigh-ritz8f.F90:128-242 (complex path, ncol = 1, nso = 0)
  subroutine rayleighritz_c( &
      nso, natom, num_ppcell, num_spe, nprjmx, neigmx, numk, nums, ncol, nspv, num_list, ncpx, ncpy, &
          ncpz, npxmax, npymax, npzmax, nperi, nk, ns1, nf, ndisp, key_natpri_in, key_natpri_inps, &
          key_soc_calc, dx, dy, dz, xmax, ymax, zmax, skpxx, skpyy, skpzz, nprj, natinf, lstvec2, &
          indspe, natpri, naps, natsoc, ntyppp, lstx, lsty, lstz, natx, naty, natz, dij, dijsoc, veff, &
          vnlocp, l_rrb, cspsep, sveccm, ssvcm, hsvcm, sval)
    integer, intent(in) :: nso, natom, num_ppcell, num_spe, nprjmx, neigmx, numk, nums, ncol, nspv
    integer, intent(in) :: num_list, ncpx, ncpy, ncpz, npxmax, npymax, npzmax, nperi, nk, ns1, nf, ndisp
    integer, intent(in) :: key_natpri_in, key_natpri_inps, key_soc_calc
    logical, intent(in) :: l_rrb
    real*8, intent(in) :: dx, dy, dz, xmax, ymax, zmax
    integer, intent(in) :: nprj(*), natinf(*), lstvec2(*), indspe(*), natpri(*), naps(*), natsoc(*)
    integer, intent(in) :: ntyppp(*), lstx(*), lsty(*), lstz(*), natx(*), naty(*), natz(*)
    real*8, intent(in) :: skpxx(*), skpyy(*), skpzz(*), dij(*), dijsoc(*), veff(*), vnlocp(*)
    real*8, intent(inout) :: sval(*)
    complex*16, intent(inout) :: cspsep(*), sveccm(*), ssvcm(*), hsvcm(*)
    call kukan_check(c_rayleighritz_c(nso, natom, num_ppcell, num_spe, nprjmx, neigmx, numk, nums, ncol, nspv, num_list, ncpx, ncpy, &
             ncpz, npxmax, npymax, npzmax, nperi, nk, ns1, nf, ndisp, key_natpri_in, key_natpri_inps, &
             key_soc_calc, dx, dy, dz, xmax, ymax, zmax, skpxx, skpyy, skpzz, nprj, natinf, lstvec2, &
             indspe, natpri, naps, natsoc, ntyppp, lstx, lsty, lstz, natx, naty, natz, dij, dijsoc, &
             veff, vnlocp, l2i(l_rrb), cspsep, sveccm, ssvcm, hsvcm, sval))
  end subroutine
end module mod_rayleighritz_hip


! ------------------------------------------------------------- reorderstates8f.F90:8-124
module mod_reorderstates_hip
  use kukan_hip_c
  implicit none
contains
  subroutine reorderstates(nrc, num_ppcell, nprjmx, nei000gmx, nums, ncol, numk, ncpx, ncpy, ncpz, ns1, nk, &
                           sval, residual_states, rspsep, cspsep, svecre, sveccm, ssvre, ssvcm, hsvre, hsvcm)
    integer, intent(in) :: nrc, num_ppcell, nprjmx, nei000gmx, nums, ncol, numk, ncpx, ncpy, ncpz, ns1, nk
    real*8, intent(inout) :: sval(nei000gmx, nums+1-ncol, numk), residual_states(nei000gmx, nums+1-ncol, numk)
    real*8, intent(inout) :: rspsep(nprjmx*(1-nrc)+nrc, num_ppcell*(1-nrc)+nrc, nei000gmx*(1-nrc)+nrc, &
                                    nums*(1-nrc)+nrc, numk*(1-nrc)+nrc)
    complex*16, intent(inout) :: cspsep(*), sveccm(*), ssvcm(*), hsvcm(*)
    real*8, intent(inout) :: svecre(ncpx*(1-nrc)+nrc, ncpy*(1-nrc)+nrc, ncpz*(1-nrc)+nrc, &
                                    nei000gmx*(1-nrc)+nrc, nums*(1-nrc)+nrc, numk*(1-nrc)+nrc)
    real*8, intent(inout) :: ssvre(ncpx*(1-nrc)+nrc, ncpy*(1-nrc)+nrc, ncpz*(1-nrc)+nrc, &
                                   nei000gmx*(1-nrc)+nrc, nums*(1-nrc)+nrc, numk*(1-nrc)+nrc)
    real*8, intent(inout) :: hsvre(ncpx*(1-nrc)+nrc, ncpy*(1-nrc)+nrc, ncpz*(1-nrc)+nrc, nei000gmx*(1-nrc)+nrc)
    call kukan_check(c_reorderstates(nrc, num_ppcell, nprjmx, nei000gmx, nums, ncol, numk, ncpx, ncpy, ncpz, &
         ns1, nk, sval, residual_states, rspsep, cspsep, svecre, sveccm, ssvre, ssvcm, hsvre, hsvcm))
  end subroutine
end module mod_reorderstates_hip


! ------------------------------------------------------------- scf_diag8f.F90:6-316
module mod_scf_diag_hip
  use kukan_hip_c
  implicit none
contains
  subroutine scf_diag(nrc, nso, nprjmx, num_ppcell, num_spe, iscf, ncgscf, nretcg, natom, neigmx, nums, &
                      ncol, nspv, numk, nperi, northo, nf, ndisp, num_list, ncpx, ncpy, ncpz, npxmax, &
                      npymax, npzmax, nchange, nrrz, key_ortho_nocmpt_innerproduct, &
                      key_ortho_cmpt_innerproduct, key_natpri_in, key_natpri_inps, key_natpri_out, &
                      key_soc_calc, key_pp_paw, skpxx, skpyy, skpzz, dx, dy, dz, xmax, ymax, zmax, dij, &
                      dijsoc, veff, vnlocp, sss, nkscg, nsdmax, nprecon_cg, epssd, ndiismax, nprecon_diis, &
                      eps_eig_diis, alambda_diis, ratio_diis, alambda_min, alambda_max, indspe, natpri, &
                      naps, nprj, natinf, lstvec2, latom, natsoc, ntyppp, lstx, lsty, lstz, natx, naty, &
                      natz, rspsep, cspsep, svecre, sveccm, hsvre, hsvcm, ssvre, ssvcm, sval, svalsoc, &
                      cspsepsoc, svecsoc, residual_states, ksconv, ksitmax)
    integer, intent(in) :: nrc, nso, nprjmx, num_ppcell, num_spe, iscf, ncgscf, nretcg
    integer, intent(in) :: natom, neigmx, nums, ncol, nspv, numk, nperi, northo, nf, ndisp, num_list
    integer, intent(in) :: ncpx, ncpy, ncpz, npxmax, npymax, npzmax, nchange, nrrz
    integer, intent(in) :: key_ortho_nocmpt_innerproduct, key_ortho_cmpt_innerproduct
    integer, intent(in) :: key_natpri_in, key_natpri_inps, key_natpri_out, key_soc_calc, key_pp_paw
    real*8, intent(in) :: skpxx(numk), skpyy(numk), skpzz(numk), dx, dy, dz, xmax, ymax, zmax
    real*8, intent(in) :: dij(nprjmx, nprjmx, nums*ncol, natom)
    real*8, intent(in) :: dijsoc(*)
    real*8, intent(in) :: veff(ncpx, ncpy, ncpz, nspv)
    real*8, intent(in) :: vnlocp(num_list, nprjmx, num_ppcell), sss(nprjmx, nprjmx, num_spe)
    integer, intent(in) :: nkscg, nsdmax, nprecon_cg
    real*8, intent(in) :: epssd
    integer, intent(in) :: ndiismax, nprecon_diis
    real*8, intent(in) :: eps_eig_diis, alambda_diis, ratio_diis, alambda_min, alambda_max
    integer, intent(in) :: indspe(natom), natpri(natom), naps(natom), nprj(num_spe)
    integer, intent(in) :: natinf(natom), lstvec2(num_list, num_ppcell), latom(natom), natsoc(natom)
    integer, intent(in) :: ntyppp(num_spe)
    integer, intent(in) :: lstx(num_list, num_ppcell), lsty(num_list, num_ppcell), lstz(num_list, num_ppcell)
    integer, intent(in) :: natx(natom), naty(natom), natz(natom)
    real*8, intent(inout) :: rspsep(nprjmx*(1-nrc)+nrc, num_ppcell*(1-nrc)+nrc, neigmx*(1-nrc)+nrc, &
                                    nums*(1-nrc)+nrc, numk*(1-nrc)+nrc)
    complex*16, intent(inout) :: cspsep(*), sveccm(*), hsvcm(*), ssvcm(*)
    real*8, intent(inout) :: svecre(ncpx*(1-nrc)+nrc, ncpy*(1-nrc)+nrc, ncpz*(1-nrc)+nrc, &
                                    neigmx*(1-nrc)+nrc, nums*(1-nrc)+nrc, numk*(1-nrc)+nrc)
    real*8, intent(inout) :: hsvre(ncpx*(1-nrc)+nrc, ncpy*(1-nrc)+nrc, ncpz*(1-nrc)+nrc, neigmx*(1-nrc)+nrc)
    real*8, intent(inout) :: ssvre(ncpx*(1-nrc)+nrc, ncpy*(1-nrc)+nrc, ncpz*(1-nrc)+nrc, &
                                   neigmx*(1-nrc)+nrc, nums*(1-nrc)+nrc, numk*(1-nrc)+nrc)
    real*8, intent(out) :: sval(neigmx, nums+1-ncol, numk)
    real*8, intent(out) :: svalsoc(*)
    complex*16, intent(out) :: cspsepsoc(*), svecsoc(*)
    real*8, intent(out) :: residual_states(neigmx, nums+1-ncol, numk)
    logical, intent(out) :: ksconv
    integer, intent(out) :: ksitmax
    integer(c_int) :: ic, it
    ic = 1; it = 0
    call kukan_check(c_scf_diag(nrc, nso, nprjmx, num_ppcell, num_spe, iscf, ncgscf, nretcg, natom, neigmx, &
         nums, ncol, nspv, numk, nperi, northo, nf, ndisp, num_list, ncpx, ncpy, ncpz, npxmax, npymax, &
         npzmax, nchange, nrrz, key_ortho_nocmpt_innerproduct, key_ortho_cmpt_innerproduct, key_natpri_in, &
         key_natpri_inps, key_natpri_out, key_soc_calc, key_pp_paw, skpxx, skpyy, skpzz, dx, dy, dz, xmax, &
         ymax, zmax, dij, dijsoc, veff, vnlocp, sss, nkscg, nsdmax, nprecon_cg, epssd, ndiismax, &
         nprecon_diis, eps_eig_diis, alambda_diis, ratio_diis, alambda_min, alambda_max, indspe, natpri, &
         naps, nprj, natinf, lstvec2, latom, natsoc, ntyppp, lstx, lsty, lstz, natx, naty, natz, rspsep, &
         cspsep, svecre, sveccm, hsvre, hsvcm, ssvre, ssvcm, sval, svalsoc, cspsepsoc, svecsoc, &
         residual_states, ic, it))
    ksconv = ic /= 0; ksitmax = it
  end subroutine
end module mod_scf_diag_hip


! ------------------------------------------------------------- scf_hartree8f.F90:11-611
module mod_scf_hartree_hip
  use kukan_hip_c
  implicit none
contains
  subroutine scf_hartree(nd, ncpx_d, ncpy_d, ncpz_d, nfh, ndisp, nperi, ncgres, ncgmin, ncgmax, xmax, ymax, &
                         zmax, epsvh, rho_aug_dense, boundx, boundy, boundz, vh_dense)
    integer, intent(in) :: nd, ncpx_d, ncpy_d, ncpz_d, nfh, ndisp, nperi, ncgres, ncgmin, ncgmax
    real*8, intent(in) :: xmax, ymax, zmax, epsvh
    real*8, intent(in) :: rho_aug_dense(ncpx_d, ncpy_d, ncpz_d)
    real*8, intent(in) :: boundx(-(nfh-1)*(1-nd)+nd:nfh*(1-nd)+nd, ncpy_d*(1-nd)+nd, ncpz_d*(1-nd)+nd)
    real*8, intent(in) :: boundy(ncpx_d*(1-nd)+nd, -(nfh-1)*(1-nd)+nd:nfh*(1-nd)+nd, ncpz_d*(1-nd)+nd)
    real*8, intent(in) :: boundz(ncpx_d*(1-nd)+nd, ncpy_d*(1-nd)+nd, -(nfh-1)*(1-nd)+nd:nfh*(1-nd)+nd)
    real*8, intent(inout) :: vh_dense(ncpx_d, ncpy_d, ncpz_d)
    call kukan_check(c_scf_hartree(nd, ncpx_d, ncpy_d, ncpz_d, nfh, ndisp, nperi, ncgres, ncgmin, ncgmax, &
         xmax, ymax, zmax, epsvh, rho_aug_dense, boundx, boundy, boundz, vh_dense))
  end subroutine
end module mod_scf_hartree_hip


! ------------------------------------------------------------- scf_charge8f.F90:290-321
module mod_kukan_charge
  use kukan_hip_c
  implicit none
contains
  subroutine scf_charge_smt_r(neigmx, nums, nspv, numk, ncpx, ncpy, ncpz, fnele, svecre, rhosmt)
    integer, intent(in) :: neigmx, nums, nspv, numk, ncpx, ncpy, ncpz
    real*8, intent(in) :: fnele(neigmx, nums, numk)
    real*8, intent(in) :: svecre(ncpx, ncpy, ncpz, neigmx, nums, numk)
    real*8, intent(out) :: rhosmt(ncpx, ncpy, ncpz, nspv)
    call kukan_check(c_charge_smt_r(neigmx, nums, nspv, numk, ncpx, ncpy, ncpz, fnele, svecre, rhosmt))
  end subroutine
  ! scf_charge8f.F90:324-381
  subroutine scf_charge_smt_c( &
      neigmx, nums, ncol, nspv, numk, ncpx, ncpy, ncpz, fnele, sveccm, rhosmt)
    integer, intent(in) :: neigmx, nums, ncol, nspv, numk, ncpx, ncpy, ncpz
    real*8, intent(in) :: fnele(*)
    real*8, intent(inout) :: rhosmt(*)
    complex*16, intent(in) :: sveccm(*)
    call kukan_check(c_scf_charge_smt_c(neigmx, nums, ncol, nspv, numk, ncpx, ncpy, ncpz, fnele, sveccm, rhosmt))
  end subroutine
end module mod_kukan_charge

! -----------------------------------------------------------------------------
! mod_pseudocalc (pseudocalc8f.F90): the weights, list vectors and projector table
module mod_pseudocalc_hip
  use kukan_hip_c
  implicit none
contains
  ! pseudocalc8f.F90:343-629
  subroutine pseudocalc_01a(nfdg, npmesh, wxyz)
    integer, intent(in) :: nfdg, npmesh
    real*8 :: wxyz(*)
    call kukan_check(c_pseudocalc_01a(nfdg, npmesh, wxyz))
  end subroutine

  ! pseudocalc8f.F90:633-1159
  subroutine pseudocalc_01b(nperi, natom, num_spe, nradmx, nmesh, nfdg, ndisp, num_atcell, &
      num_ppcell, num_ppcell_d, num_list, num_list_d, nxmax, nymax, nzmax, ncpx, ncpy, ncpz, npxmax, &
      npymax, npzmax, key_natpri_in, key_natpri_inps, key_natpri_out, psctoff, xmax, ymax, zmax, dx, &
      dy, dz, indspe, nradct, radial, atx, aty, atz, natx, naty, natz, ndatx, ndaty, ndatz, lstx, &
      lsty, lstz, lstdx, lstdy, lstdz, natinf, natinfd, natinfd_vloc, lstvec2, lstvecd2, natpri, &
      natprid, natpri_inf, naps, napsd)
    integer, intent(in) :: nperi, natom, num_spe, nradmx, nmesh, nfdg, ndisp, num_atcell, num_ppcell, &
      num_ppcell_d, num_list, num_list_d, nxmax, nymax, nzmax, ncpx, ncpy, ncpz, npxmax, npymax, &
      npzmax, key_natpri_in, key_natpri_inps, key_natpri_out
    real*8, intent(in) :: psctoff, xmax, ymax, zmax, dx, dy, dz
    integer :: indspe(*), nradct(*), natx(*), naty(*), natz(*), ndatx(*), ndaty(*), ndatz(*), &
      lstx(*), lsty(*), lstz(*), lstdx(*), lstdy(*), lstdz(*), natinf(*), natinfd(*), &
      natinfd_vloc(*), lstvec2(*), lstvecd2(*), natpri(*), natprid(*), natpri_inf(*), naps(*), &
      napsd(*)
    real*8 :: radial(*), atx(*), aty(*), atz(*)
    call kukan_check(c_pseudocalc_01b(nperi, natom, num_spe, nradmx, nmesh, nfdg, ndisp, num_atcell, &
      num_ppcell, num_ppcell_d, num_list, num_list_d, nxmax, nymax, nzmax, ncpx, ncpy, ncpz, npxmax, &
      npymax, npzmax, key_natpri_in, key_natpri_inps, key_natpri_out, psctoff, xmax, ymax, zmax, dx, &
      dy, dz, indspe, nradct, radial, atx, aty, atz, natx, naty, natz, ndatx, ndaty, ndatz, lstx, &
      lsty, lstz, lstdx, lstdy, lstdz, natinf, natinfd, natinfd_vloc, lstvec2, lstvecd2, natpri, &
      natprid, natpri_inf, naps, napsd))
  end subroutine

  ! pseudocalc8f.F90:2515-3702
  subroutine pseudocalc_07(natom, num_spe, npmesh, nprjmx, nfdg, nfiltyp, num_list, num_ppcell, &
      nradmx, nqmx, nxmax, nymax, nzmax, npxmax, npymax, npzmax, nxspd, nyspd, nzspd, key_natpri_in, &
      key_natpri_inps, eps, psftrad, psctoff, filpp, dx, dy, dz, indspe, nprj, nradct, natpri, naps, &
      natinf, nqct, natx, naty, natz, lstx, lsty, lstz, wxyz, radial, coef, atx, aty, atz, vnlocp, &
      vtmp_pp, vtmp_pp1, vtmp_pp2, vnlspd)
    integer, intent(in) :: natom, num_spe, npmesh, nprjmx, nfdg, nfiltyp, num_list, num_ppcell, &
      nradmx, nqmx, nxmax, nymax, nzmax, npxmax, npymax, npzmax, nxspd, nyspd, nzspd, key_natpri_in, &
      key_natpri_inps
    real*8, intent(in) :: eps, psftrad, psctoff, filpp, dx, dy, dz
    integer :: indspe(*), nprj(*), nradct(*), natpri(*), naps(*), natinf(*), nqct(*), natx(*), &
      naty(*), natz(*), lstx(*), lsty(*), lstz(*)
    real*8 :: wxyz(*), radial(*), coef(*), atx(*), aty(*), atz(*), vnlocp(*), vtmp_pp(*), &
      vtmp_pp1(*), vtmp_pp2(*), vnlspd(*)
    call kukan_check(c_pseudocalc_07(natom, num_spe, npmesh, nprjmx, nfdg, nfiltyp, num_list, &
      num_ppcell, nradmx, nqmx, nxmax, nymax, nzmax, npxmax, npymax, npzmax, nxspd, nyspd, nzspd, &
      key_natpri_in, key_natpri_inps, eps, psftrad, psctoff, filpp, dx, dy, dz, indspe, nprj, nradct, &
      natpri, naps, natinf, nqct, natx, naty, natz, lstx, lsty, lstz, wxyz, radial, coef, atx, aty, &
      atz, vnlocp, vtmp_pp, vtmp_pp1, vtmp_pp2, vnlspd))
  end subroutine
end module mod_pseudocalc_hip

! -----------------------------------------------------------------------------
! mod_force (force8f.F90): the forces beyond the band-structure term
module mod_force_hip
  use kukan_hip_c
  implicit none
contains
  ! force8f.F90:1617-2190
  subroutine force(natom, num_spe, num_atcell, num_ppcell_d, num_list_d, nqmx, nradmx, npoint, nmesh, &
      nspv, nperi, nint1dmax, nzmax, jelcalc, ncpx, ncpy, ncpz, ncpx_d, ncpy_d, ncpz_d, new_pwx, &
      new_pwy, new_pwz, new_rsx, new_rsy, new_rsz, key_jel_calc, key_natpri_in, key_natpri_inps, &
      indspe, natpri, natpri_inf, nradct, natprid, napsd, natinfd, natinfd_vloc, lstvecd2, nqctpcc, &
      lstdx, lstdy, lstdz, veta, psctoff, psftrad, radial, dradial, point, wt, xmax, ymax, zmax, &
      biasx, biasy, biasz, cp, coef, atx, aty, atz, dvlocdx_scw, dvlocdy_scw, dvlocdz_scw, &
      dvlocdx_hdp, dvlocdy_hdp, dvlocdz_hdp, rhotrur, rhosmtr, rhoaugr, drhoaug3ddx, drhoaug3ddy, &
      drhoaug3ddz, rho_coarse, rho_aug_dense, vloc_dense, vh_dense, vx_dense, chrjel, strjel, endjel, &
      ndatx, ndaty, ndatz, fatx, faty, fatz, fjel)
    integer, intent(in) :: natom, num_spe, num_atcell, num_ppcell_d, num_list_d, nqmx, nradmx, &
      npoint, nmesh, nspv, nperi, nint1dmax, nzmax, jelcalc, ncpx, ncpy, ncpz, ncpx_d, ncpy_d, &
      ncpz_d, new_pwx, new_pwy, new_pwz, new_rsx, new_rsy, new_rsz, key_jel_calc, key_natpri_in, &
      key_natpri_inps
    real*8, intent(in) :: veta, psctoff, psftrad, xmax, ymax, zmax, biasx, biasy, biasz, chrjel, &
      strjel, endjel
    integer :: indspe(*), natpri(*), natpri_inf(*), nradct(*), natprid(*), napsd(*), natinfd(*), &
      natinfd_vloc(*), lstvecd2(*), nqctpcc(*), lstdx(*), lstdy(*), lstdz(*), ndatx(*), ndaty(*), &
      ndatz(*)
    real*8 :: radial(*), dradial(*), point(*), wt(*), cp(*), coef(*), atx(*), aty(*), atz(*), &
      dvlocdx_scw(*), dvlocdy_scw(*), dvlocdz_scw(*), dvlocdx_hdp(*), dvlocdy_hdp(*), dvlocdz_hdp(*), &
      rhotrur(*), rhosmtr(*), rhoaugr(*), drhoaug3ddx(*), drhoaug3ddy(*), drhoaug3ddz(*), &
      rho_coarse(*), rho_aug_dense(*), vloc_dense(*), vh_dense(*), vx_dense(*), fatx(*), faty(*), &
      fatz(*)
    real*8, intent(out) :: fjel
    call kukan_check(c_force(natom, num_spe, num_atcell, num_ppcell_d, num_list_d, nqmx, nradmx, &
      npoint, nmesh, nspv, nperi, nint1dmax, nzmax, jelcalc, ncpx, ncpy, ncpz, ncpx_d, ncpy_d, &
      ncpz_d, new_pwx, new_pwy, new_pwz, new_rsx, new_rsy, new_rsz, key_jel_calc, key_natpri_in, &
      key_natpri_inps, indspe, natpri, natpri_inf, nradct, natprid, napsd, natinfd, natinfd_vloc, &
      lstvecd2, nqctpcc, lstdx, lstdy, lstdz, veta, psctoff, psftrad, radial, dradial, point, wt, &
      xmax, ymax, zmax, biasx, biasy, biasz, cp, coef, atx, aty, atz, dvlocdx_scw, dvlocdy_scw, &
      dvlocdz_scw, dvlocdx_hdp, dvlocdy_hdp, dvlocdz_hdp, rhotrur, rhosmtr, rhoaugr, drhoaug3ddx, &
      drhoaug3ddy, drhoaug3ddz, rho_coarse, rho_aug_dense, vloc_dense, vh_dense, vx_dense, chrjel, &
      strjel, endjel, ndatx, ndaty, ndatz, fatx, faty, fatz, fjel))
  end subroutine
end module mod_force_hip
