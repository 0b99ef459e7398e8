! polar_driver -- analyze-style driver of the polarization path.
!
! Usage: polar_driver <natom> <box> <nfft> [seed] [precision 0|1] [nsolute] [blob radius]
! Builds the seeded water (+ solute) system, rebuilds the neighbor list,
! solves for the induced dipoles and prints one line
!   EP=<kcal/mol> NITER=<n> RESNRM=<x> MURMS_D=<Debye> NPAIRS=<list pairs>
!
! Usage: polar_driver --shims <input> <output>
! Reads a system (and optional local frames and probe vectors) from a
! binary stream file, calls every routine of the reference call surface in
! module thp_polar in the order a Tinker-HP step does (nblist, efld0_direct,
! efld0_recip, tmatxb_pme with and without the diagonal, tmatxbrecip,
! grid_uind_site, fphi_uind_site, newinduce_pme2, polar_gradient) and writes
! every result to the output stream file, for comparison with the oracle.
! Layout of the input (little-endian, no record markers):
!   int32  n, nscaled, nfft(3), precision, has_frames
!   real64 box(3), aewald, cut, buffer, poleps
!   real64 x(n), y(n), z(n), rpole(13,n), polarity(n), thole(n)
!   int32  ik(2,nscaled) (1-based); real64 dpu(3,nscaled)
!   if has_frames: int32 axistype(n), zaxis(n), xaxis(n), yaxis(n)
!                  (1-based, 0 = none); real64 pole(13,n)
!   real64 mu(3,2,n), fuind(3,n), fuinp(3,n)
! Layout of the output:
!   int64  npairs; int32 niter
!   real64 ef(3,2,n), cphi(10,n), efi(3,2,n), efi_nodiag(3,2,n),
!          dipfield(3,2,n), qgrid(2,nfft1,nfft2,nfft3), fdip_phi1(3,n),
!          fdip_phi2(3,n), uind(3,n), uinp(3,n), ep, dep(3,n), torque(3,n)
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
  real(c_double), allocatable, target :: dep(:,:)
  integer(c_int) :: nscaled
  character(len=512) :: arg
  real(c_double), parameter :: debye = 4.80321d0
  real(c_double) :: murms

  nargs = command_argument_count()
  if (nargs >= 3) then
    call get_command_argument(1, arg)
    if (trim(arg) == '--shims') then
      call shims()
      stop
    end if
  end if
  if (nargs < 3) then
    write (*, '(a)') ' usage: polar_driver natom box nfft [seed] [precision] [nsolute] [radius]'
    write (*, '(a)') '        polar_driver --shims input output'
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

contains

  subroutine shims()
    character(len=512) :: fin, fout
    integer :: u, n, ns, nf(3), precision, has_frames
    real(c_double) :: bx(3), aew, cutoff, buf, eps
    real(c_double), allocatable, target :: px(:), py(:), pz(:), rp(:,:), pl(:), th(:), dp(:,:), pole(:,:)
    real(c_double), allocatable :: mu(:,:,:), fuind(:,:), fuinp(:,:)
    real(c_double), allocatable :: ef(:,:,:), cphi(:,:), efi(:,:,:), efi0(:,:,:), dipf(:,:,:), murec(:,:,:)
    real(c_double), allocatable :: dipfb(:,:,:), qgrid(:), phi1(:,:), phi2(:,:)
    real(c_double), allocatable, target :: grad(:,:), torq(:,:)
    integer(c_int), allocatable, target :: pik(:,:)
    integer, allocatable :: at(:), za(:), xa(:), ya(:)
    integer(c_int) :: hdr(7)

    call get_command_argument(2, fin)
    call get_command_argument(3, fout)
    open (newunit=u, file=trim(fin), access='stream', form='unformatted', status='old')
    read (u) hdr
    n = hdr(1); ns = hdr(2); nf = hdr(3:5); precision = hdr(6); has_frames = hdr(7)
    read (u) bx, aew, cutoff, buf, eps
    allocate (px(n), py(n), pz(n), rp(13, n), pl(n), th(n), pik(2, max(ns, 1)), dp(3, max(ns, 1)))
    read (u) px, py, pz, rp, pl, th
    if (ns > 0) then
      read (u) pik(:, 1:ns)
      read (u) dp(:, 1:ns)
    end if
    pik = pik - 1  ! 1-based pairs in the file, 0-based at the ABI
    if (has_frames /= 0) then
      allocate (at(n), za(n), xa(n), ya(n), pole(13, n))
      read (u) at, za, xa, ya
      read (u) pole
    end if
    allocate (mu(3, 2, n), fuind(3, n), fuinp(3, n))
    read (u) mu, fuind, fuinp
    close (u)

    call polar_init(0)
    call polar_setup(n, bx, px, py, pz, rp, pl, th, ns, pik, dp, aew, cutoff, buf, nf, 5, precision)
    if (has_frames /= 0) call polar_set_frames(at, za, xa, ya, pole)
    poleps = eps
    allocate (ef(3, 2, n), cphi(10, n), efi(3, 2, n), efi0(3, 2, n), dipf(3, 2, n), murec(3, 2, 1))
    allocate (dipfb(3, 2, 1), qgrid(2 * nf(1) * nf(2) * nf(3)), phi1(3, n), phi2(3, n))
    allocate (grad(3, n), torq(3, n))
    call nblist(0)
    call efld0_direct(2, ef)
    call efld0_recip(cphi)
    call tmatxb_pme(2, .true., mu, efi)
    call tmatxb_pme(2, .false., mu, efi0)
    call tmatxbrecip(mu, murec, 2, dipf, dipfb)
    call grid_uind_site(fuind, fuinp, qgrid)
    call fphi_uind_site(qgrid, phi1, phi2)
    call newinduce_pme2
    call polar_gradient(grad, torq)

    open (newunit=u, file=trim(fout), access='stream', form='unformatted', status='replace')
    write (u) nblist_npairs(), int(niter_last, c_int)
    write (u) ef, cphi, efi, efi0, dipf, qgrid, phi1, phi2, uind, uinp, ep, grad, torq
    close (u)
    write (*, '(a,es24.16,a,i0)') ' SHIMS EP=', ep, ' NITER=', niter_last
    call polar_finalize
  end subroutine

end program polar_driver
