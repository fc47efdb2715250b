! shim_driver.F90 -- calls the drop-in modules exactly as the reference host code does
! (kukan8f.F90 / scf_potentials8f.F90 call sites) on a small periodic cell and writes the
! results as raw doubles for comparison with the oracle.
program shim_driver
  use mod_overlap_finitedifference_hip, only: overlap_finitedifference_r, overlap_fdcheck_r
  use mod_kslaplacian_hip, only: kslaplacian_r
  use mod_ksprecondition_hip, only: ksprecondition
  use mod_scf_hartree_hip, only: scf_hartree
  use mod_kukan_tools, only: tools_correctpotential
  use kukan_hip_c, only: kukan_setup
  implicit none
  integer, parameter :: n = 12, nf = 4
  real*8, parameter :: h = 0.4d0
  real*8 :: xmax, pi, x, y, z
  real*8, allocatable :: vre(:,:,:), veff(:,:,:), avre(:,:,:,:), rho(:,:,:), vh(:,:,:)
  real*8, allocatable :: pre(:,:,:,:), kout(:,:,:,:)
  real*8 :: bnd(1,1,1)
  integer :: ix, iy, iz, u
  character(len=256) :: out
  call get_command_argument(1, out)
  call kukan_setup(0)
  pi = dacos(-1.0d0)
  xmax = 0.5d0 * n * h
  allocate(vre(-(nf-1):n+nf, -(nf-1):n+nf, -(nf-1):n+nf), veff(n,n,n), avre(n,n,n,2))
  allocate(rho(n,n,n), vh(n,n,n), pre(-(nf-1):n+nf, -(nf-1):n+nf, -(nf-1):n+nf, 1), kout(n,n,n,1))
  vre = 0.0d0
  pre = 0.0d0
  do iz = 1, n
  do iy = 1, n
  do ix = 1, n
    x = ix*h; y = iy*h; z = iz*h
    vre(ix,iy,iz) = sin(2*pi*x/(n*h)) * cos(4*pi*y/(n*h)) + 0.3d0*cos(2*pi*z/(n*h))
    pre(ix,iy,iz,1) = vre(ix,iy,iz)
    veff(ix,iy,iz) = -1.0d0 + 0.1d0*ix - 0.05d0*iy + 0.02d0*iz
    rho(ix,iy,iz) = 1.0d0 + 0.5d0*sin(2*pi*x/(n*h))*sin(2*pi*y/(n*h)) + 0.25d0*cos(2*pi*z/(n*h))
  end do
  end do
  end do
  avre = 1.0d0
  call overlap_finitedifference_r(3, n, n, n, nf, nf-1, nf, vre)
  call overlap_fdcheck_r(3, n, n, n, nf, nf-1, nf, vre)
  call kslaplacian_r(n, n, n, 2, nf, h, h, h, vre, veff, 2, avre)
  call overlap_finitedifference_r(3, n, n, n, nf, nf-1, 1, pre)
  call ksprecondition(n, n, n, nf, 1, pre, kout)
  vh = 0.0d0
  call scf_hartree(1, n, n, n, 4, 6, 3, 801, 1, 800, xmax, xmax, xmax, 1.0d-11, rho, bnd, bnd, bnd, vh)
  call tools_correctpotential(n, n, n, xmax, xmax, xmax, h, h, h, vh)
  open(newunit=u, file=trim(out), access='stream', form='unformatted', status='replace')
  write(u) avre, kout, vh
  close(u)
  write(*,'(a)') 'shim_driver: ok'
end program shim_driver
