! polar_driver -- analyze-style driver of one polarization solve.
!
! Usage: polar_driver <natom> <box> <nfft> [seed] [precision 0|1] [nsolute] [blob radius]
! Builds the seeded water (+ solute) system, rebuilds the neighbor list,
! solves for the induced dipoles and prints one line
!   EP=<kcal/mol> NITER=<n> RESNRM=<x> MURMS_D=<Debye> NPAIRS=<list pairs>
program polar_driver
  use iso_c_binding
  use thp_bind
  use thp_polar
  implicit none
  integer :: natom, nfft, prec, nsol, nargs
  integer(c_int64_t) :: seed
  real(c_double) :: box, radius
  real(c_double), allocatable, target :: x(:), y(:), z(:), rpole(:,:), pol(:), thl(:), dpu(:,:)
  integer(c_int), allocatable, target :: ik(:,:)
  real(8), allocatable :: dep(:,:)
  integer(c_int) :: nscaled
  character(len=64) :: arg
  real(c_double), parameter :: debye = 4.80321d0
  real(c_double) :: murms

  nargs = command_argument_count()
  if (nargs < 3) then
    write (*, '(a)') ' usage: polar_driver natom box nfft [seed] [precision] [nsolute] [radius]'
    stop 1
  end if
  call get_command_argument(1, arg); read (arg, *) natom
  call get_command_argument(2, arg); read (arg, *) box
  call get_command_argument(3, arg); read (arg, *) nfft
  seed = 1; prec = 0; nsol = 0; radius = 0.0d0
  if (nargs >= 4) then; call get_command_argument(4, arg); read (arg, *) seed; end if
  if (nargs >= 5) then; call get_command_argument(5, arg); read (arg, *) prec; end if
  if (nargs >= 6) then; call get_command_argument(6, arg); read (arg, *) nsol; end if
  if (nargs >= 7) then; call get_command_argument(7, arg); read (arg, *) radius; end if

  allocate (x(natom), y(natom), z(natom), rpole(13, natom), pol(natom), thl(natom))
  allocate (ik(2, 3 * natom), dpu(3, 3 * natom))
  if (thp_make_system(int(natom, c_int), box, int(nsol, c_int), radius, seed, x, y, z, rpole, pol, &
                      thl, nscaled, ik, dpu) /= THP_OK) error stop 'thp_make_system failed'

  call polar_init(0)
  call polar_setup(natom, [box, box, box], x, y, z, rpole, pol, thl, int(nscaled), ik, dpu, &
                   0.4d0, 7.0d0, min(2.0d0, 0.5d0 * box - 7.05d0), [nfft, nfft, nfft], 5, prec)
  call nblist(0)
  poleps = 1.0d-8
  call newinduce_pme2
  murms = sqrt(sum(uind**2) / (3.0d0 * natom)) * debye
  write (*, '(a,es24.16,a,i0,a,es12.4,a,es14.6,a,i0)') ' EP=', ep, ' NITER=', niter_last, &
    ' RESNRM=', resnrm_last, ' MURMS_D=', murms, ' NPAIRS=', nblist_npairs()
  ! the polarization gradient of that solve (global multipoles: forces at
  ! fixed orientations)
  allocate (dep(3, natom))
  call polar_gradient(dep)
  write (*, '(a,es24.16)') ' GRADRMS=', sqrt(sum(dep**2) / (3.0d0 * natom))
  call polar_finalize
end program polar_driver
