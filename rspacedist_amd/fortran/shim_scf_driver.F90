! shim_scf_driver.F90 -- the coarse drop-in routines called from Fortran exactly as the
! reference host calls them (kukan8f.F90 / scf_diag8f.F90 argument lists), on a system read
! from disk: the real path (orthogonalization_r, scf_diag with nrc = 0 over CG and DIIS steps,
! kscg_r, rayleighritz_r, reorderstates, ksdiis_r, scf_charge_smt_r) and the complex path
! (the same with the _c routines and nrc = 1).  Every logical in and out of the shim
! (ksconv, l_rrb, l_hsv) is exercised.
!
! usage: shim_scf_driver <namelist file> <array file> <output file>
! The array file is a stream of the system's tables (the order below); the output file
! receives, as raw doubles, the eigenvalues, residuals, ksconv and ksitmax after every call
! and the final bands and densities.
program shim_scf_driver
  use mod_orthogonalization_hip, only: orthogonalization_r, orthogonalization_c
  use mod_scf_diag_hip, only: scf_diag
  use mod_kscg_hip, only: kscg_r, kscg_c
  use mod_ksdiis_hip, only: ksdiis_r, ksdiis_c
  use mod_rayleighritz_hip, only: rayleighritz_r, rayleighritz_c
  use mod_reorderstates_hip, only: reorderstates
  use mod_kukan_charge, only: scf_charge_smt_r, scf_charge_smt_c
  use kukan_hip_c, only: kukan_setup
  implicit none
  integer :: n, nbands, natom, num_list, nprjmx, num_spe, nf, nperi, nsteps
  integer :: ncgscf, nrrz, nchange, ndiismax, nsdmax, northo, nkscg, nprecon_cg, nprecon_diis
  integer :: kin, kinps, kout, knocmpt, kcmpt
  real*8 :: h, kx, ky, kz, epssd, eps_eig_diis, alambda_diis, ratio_diis, alambda_min, alambda_max
  namelist /sys/ n, nbands, natom, num_list, nprjmx, num_spe, nf, nperi, nsteps, ncgscf, nrrz, nchange, &
                 ndiismax, nsdmax, northo, nkscg, nprecon_cg, nprecon_diis, kin, kinps, kout, knocmpt, &
                 kcmpt, h, kx, ky, kz, epssd, eps_eig_diis, alambda_diis, ratio_diis, alambda_min, alambda_max
  integer, allocatable :: nprj(:), indspe(:), natinf(:), natpri(:), naps(:), latom(:), ntyppp(:), natsoc(:)
  integer, allocatable :: lstvec2(:,:), lstx(:,:), lsty(:,:), lstz(:,:), natx(:), naty(:), natz(:)
  real*8, allocatable :: vnlocp(:,:,:), dij(:,:,:,:), sss(:,:,:), veff(:,:,:,:)
  real*8, allocatable :: svecre(:,:,:,:,:,:), ssvre(:,:,:,:,:,:), hsvre(:,:,:,:), rspsep(:,:,:,:,:)
  real*8, allocatable :: sval(:,:,:), resid(:,:,:), rho(:,:,:,:)
  complex*16, allocatable :: sveccm(:,:,:,:,:,:), ssvcm(:,:,:,:,:,:), hsvcm(:,:,:,:,:), cspsep(:,:,:,:,:)
  real*8 :: skpxx(1), skpyy(1), skpzz(1), dijsoc(1), rdum(1), svalsoc(1)
  complex*16 :: cdum(1)
  real*8 :: xmax
  logical :: ksconv
  integer :: ksitmax, iscf, u, uo
  character(len=512) :: fnml, farr, fout

  call get_command_argument(1, fnml)
  call get_command_argument(2, farr)
  call get_command_argument(3, fout)
  open(newunit=u, file=trim(fnml), status='old')
  read(u, nml=sys)
  close(u)
  xmax = 0.5d0 * n * h
  skpxx(1) = kx; skpyy(1) = ky; skpzz(1) = kz
  dijsoc = 0.0d0
  allocate(nprj(num_spe), indspe(natom), natinf(natom), natpri(natom), naps(natom), latom(natom))
  allocate(ntyppp(num_spe), natsoc(natom), natx(natom), naty(natom), natz(natom))
  allocate(lstvec2(num_list, natom), lstx(num_list, natom), lsty(num_list, natom), lstz(num_list, natom))
  allocate(vnlocp(num_list, nprjmx, natom), dij(nprjmx, nprjmx, 1, natom), sss(nprjmx, nprjmx, num_spe))
  allocate(veff(n, n, n, 1))
  allocate(svecre(n, n, n, nbands, 1, 1), ssvre(n, n, n, nbands, 1, 1), hsvre(n, n, n, nbands))
  allocate(rspsep(nprjmx, natom, nbands, 1, 1), sval(nbands, 1, 1), resid(nbands, 1, 1), rho(n, n, n, 1))
  allocate(sveccm(n, n, n, nbands, 1, 1), ssvcm(n, n, n, nbands, 1, 1), hsvcm(n, n, n, nbands, 1))
  allocate(cspsep(nprjmx, natom, nbands, 1, 1))
  natsoc = 0
  open(newunit=u, file=trim(farr), access='stream', form='unformatted', status='old')
  read(u) nprj, indspe, natinf, natpri, naps, latom, ntyppp, lstvec2, lstx, lsty, lstz, natx, naty, natz
  read(u) vnlocp, dij, sss, veff, svecre, sveccm
  close(u)

  call kukan_setup(0)
  open(newunit=uo, file=trim(fout), access='stream', form='unformatted', status='replace')
  ssvre = 0.0d0; hsvre = 0.0d0; rspsep = 0.0d0; sval = 0.0d0; resid = 0.0d0

  ! ------------------------------------------------------------ real path (nrc = 0)
  call orthogonalization_r(natom, nbands, nprjmx, 1, 1, natom, num_spe, northo, num_list, kcmpt, 6, &
                           n, n, n, kcmpt, kin, kinps, h, h, h, nprj, indspe, natpri, naps, natinf, &
                           lstvec2, latom, ntyppp, vnlocp, sss, svecre, ssvre, rspsep)
  do iscf = 1, nsteps
    call scf_diag(0, 0, nprjmx, natom, num_spe, iscf, ncgscf, 2000, natom, nbands, 1, 1, 1, 1, nperi, &
                  northo, nf, 6, num_list, n, n, n, n/2, n/2, n/2, nchange, nrrz, knocmpt, kcmpt, kin, &
                  kinps, kout, 1, 1, skpxx, skpyy, skpzz, h, h, h, xmax, xmax, xmax, dij, dijsoc, veff, &
                  vnlocp, sss, nkscg, nsdmax, nprecon_cg, epssd, ndiismax, nprecon_diis, eps_eig_diis, &
                  alambda_diis, ratio_diis, alambda_min, alambda_max, indspe, natpri, naps, nprj, natinf, &
                  lstvec2, latom, natsoc, ntyppp, lstx, lsty, lstz, natx, naty, natz, rspsep, cdum, &
                  svecre, cdum, hsvre, cdum, ssvre, cdum, sval, svalsoc, cdum, cdum, resid, ksconv, ksitmax)
    call record()
  end do
  ksconv = .true.; ksitmax = 0
  call kscg_r(natom, num_spe, nbands, 1, 1, 1, nperi, northo, 1, 1, nf, nkscg, nsdmax, nprecon_cg, &
              nprjmx, num_list, natom, 6, n, n, n, knocmpt, kcmpt, kin, kinps, kout, h, h, h, epssd, &
              indspe, natpri, naps, nprj, natinf, lstvec2, latom, ntyppp, dij, veff, vnlocp, sss, &
              ksconv, ksitmax, sval, rspsep, svecre, ssvre, resid, hsvre)
  call record()
  call rayleighritz_r(natom, natom, num_spe, nprjmx, nbands, 1, 1, 1, num_list, n, n, n, nperi, 1, 1, &
                      nf, 6, kin, kinps, h, h, h, nprj, natinf, lstvec2, indspe, natpri, naps, ntyppp, &
                      dij, veff, vnlocp, .true., rspsep, svecre, ssvre, hsvre, sval)
  call record()
  call reorderstates(0, natom, nprjmx, nbands, 1, 1, 1, n, n, n, 1, 1, sval, resid, rspsep, cdum, &
                     svecre, cdum, ssvre, cdum, hsvre, cdum)
  call record()
  ksconv = .true.; ksitmax = 0
  call ksdiis_r(.true., natom, num_spe, nbands, 1, 1, 1, nperi, ndiismax, nprecon_diis, 1, 1, nf, &
                nprjmx, num_list, natom, 6, northo, n, n, n, kcmpt, kin, kinps, eps_eig_diis, &
                alambda_diis, ratio_diis, alambda_min, alambda_max, epssd, xmax, xmax, xmax, indspe, &
                natpri, naps, nprj, natinf, lstvec2, latom, ntyppp, dij, veff, vnlocp, sss, ksconv, &
                ksitmax, rspsep, sval, svecre, ssvre, resid, hsvre)
  call record()
  call scf_charge_smt_r(nbands, 1, 1, 1, n, n, n, spread(1.0d0, 1, nbands), svecre, rho)
  write(uo) svecre, rho

  ! ------------------------------------------------------------ complex path (nrc = 1)
  ssvcm = (0.0d0, 0.0d0); hsvcm = (0.0d0, 0.0d0); cspsep = (0.0d0, 0.0d0); sval = 0.0d0; resid = 0.0d0
  call orthogonalization_c(natom, nbands, nprjmx, 1, 1, 1, natom, num_spe, northo, num_list, kcmpt, 6, &
                           n, n, n, n/2, n/2, n/2, kcmpt, kin, kinps, h, h, h, nprj, indspe, natpri, &
                           naps, natinf, lstvec2, latom, ntyppp, lstx, lsty, lstz, natx, naty, natz, &
                           skpxx, skpyy, skpzz, vnlocp, sss, sveccm, ssvcm, cspsep)
  do iscf = 1, nsteps
    call scf_diag(1, 0, nprjmx, natom, num_spe, iscf, ncgscf, 2000, natom, nbands, 1, 1, 1, 1, nperi, &
                  northo, nf, 6, num_list, n, n, n, n/2, n/2, n/2, nchange, nrrz, knocmpt, kcmpt, kin, &
                  kinps, kout, 1, 1, skpxx, skpyy, skpzz, h, h, h, xmax, xmax, xmax, dij, dijsoc, veff, &
                  vnlocp, sss, nkscg, nsdmax, nprecon_cg, epssd, ndiismax, nprecon_diis, eps_eig_diis, &
                  alambda_diis, ratio_diis, alambda_min, alambda_max, indspe, natpri, naps, nprj, natinf, &
                  lstvec2, latom, natsoc, ntyppp, lstx, lsty, lstz, natx, naty, natz, rdum, cspsep, &
                  rdum, sveccm, rdum, hsvcm, rdum, ssvcm, sval, svalsoc, cdum, cdum, resid, ksconv, ksitmax)
    call record()
  end do
  ksconv = .true.; ksitmax = 0
  call kscg_c(0, natom, num_spe, nbands, 1, 1, 1, 1, nperi, northo, 1, 1, nf, nkscg, nsdmax, nprecon_cg, &
              nprjmx, num_list, natom, 6, n, n, n, n/2, n/2, n/2, knocmpt, kcmpt, kin, kinps, kout, 1, &
              h, h, h, kx, ky, kz, epssd, indspe, natpri, naps, nprj, natinf, lstvec2, latom, natsoc, &
              ntyppp, lstx, lsty, lstz, natx, naty, natz, dij, dijsoc, veff, vnlocp, sss, ksconv, &
              ksitmax, sval, cspsep, sveccm, ssvcm, resid, hsvcm)
  call record()
  call rayleighritz_c(0, natom, natom, num_spe, nprjmx, nbands, 1, 1, 1, 1, num_list, n, n, n, n/2, &
                      n/2, n/2, nperi, 1, 1, nf, 6, kin, kinps, 1, h, h, h, xmax, xmax, xmax, skpxx, &
                      skpyy, skpzz, nprj, natinf, lstvec2, indspe, natpri, naps, natsoc, ntyppp, lstx, &
                      lsty, lstz, natx, naty, natz, dij, dijsoc, veff, vnlocp, .true., cspsep, sveccm, &
                      ssvcm, hsvcm, sval)
  call record()
  call reorderstates(1, natom, nprjmx, nbands, 1, 1, 1, n, n, n, 1, 1, sval, resid, rdum, cspsep, &
                     rdum, sveccm, rdum, ssvcm, rdum, hsvcm)
  call record()
  ksconv = .true.; ksitmax = 0
  call ksdiis_c(.true., 0, natom, num_spe, nbands, 1, 1, 1, 1, nperi, ndiismax, nprecon_diis, 1, 1, nf, &
                nprjmx, num_list, natom, 6, northo, n, n, n, n/2, n/2, n/2, kcmpt, kin, kinps, 1, &
                eps_eig_diis, alambda_diis, ratio_diis, alambda_min, alambda_max, epssd, xmax, xmax, &
                xmax, kx, ky, kz, indspe, natpri, naps, nprj, natinf, lstvec2, latom, natsoc, ntyppp, &
                lstx, lsty, lstz, natx, naty, natz, dij, dijsoc, veff, vnlocp, sss, ksconv, ksitmax, &
                cspsep, sval, sveccm, ssvcm, resid, hsvcm)
  call record()
  call scf_charge_smt_c(nbands, 1, 1, 1, 1, n, n, n, spread(1.0d0, 1, nbands), sveccm, rho)
  write(uo) sveccm, rho
  close(uo)
  write(*, '(a)') 'shim_scf_driver: ok'

contains

  subroutine record()
    real*8 :: flags(2)
    flags(1) = merge(1.0d0, 0.0d0, ksconv)
    flags(2) = dble(ksitmax)
    write(uo) sval, resid, flags
  end subroutine record

end program shim_scf_driver
