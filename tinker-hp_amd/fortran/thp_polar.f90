! Fortran host of the AMOEBA induced-dipole solve: the reference call
! surface over libthp.
!
! Module state plays the role of the reference modules (atoms, mpole,
! polar, pme, neigh, uprior): the system is registered once with
! polar_setup and the routines below take the reference arguments:
!
!   nblist(istep) / mlistcell .............. v1.2/source/nblist.f:17-247
!   efld0_direct(nrhs, ef) .................. v1.2/source/newinduce_pme.f:948
!   efld0_recip(cphi) ....................... v1.2/source/newinduce_pme.f:1671
!   tmatxb_pme(nrhs, dodiag, mu, efi) ....... v1.2/source/newinduce_pme.f:1392
!   tmatxbrecip(mu, murec, nrhs, dipfield, dipfieldbis)
!                                           v1.2/source/newinduce_pme.f:1919
!   grid_uind_site / fphi_uind_site ......... v1.2/source/pmestuff.f:280, :746
!   newinduce_pme2 .......................... v1.2/source/newinduce_pme2.f:16
!
! Arrays keep the reference layouts: ef/mu/efi(3,nrhs,n), cphi(10,n),
! rpole(13,n).  A failing library call is fatal, as in the reference.
module thp_polar
  use iso_c_binding
  use thp_bind
  implicit none

  type(c_ptr), save :: ctx = c_null_ptr, sys = c_null_ptr
  integer, save :: n = 0
  ! solver keywords (POLAR-EPS, POLAR-ITER, POLAR-GUESS; prmkey.f:395-455)
  real(c_double), save :: poleps = 1.0d-8
  integer, save :: politer = 500
  integer, save :: polgsf = 0
  ! results of the last solve (uind/uinp of module polar, ep of energi)
  real(c_double), allocatable, target, save :: uind(:,:), uinp(:,:)
  real(c_double), save :: ep = 0.0d0
  integer, save :: niter_last = 0
  real(c_double), save :: resnrm_last = 0.0d0

contains

  subroutine fatal_status(rc, what)
    integer(c_int), intent(in) :: rc
    character(*), intent(in) :: what
    if (rc /= THP_OK) then
      write (*, '(a,a,a,i0,a,a)') ' THP  --  ', what, ' failed, status ', rc, ': ', trim(thp_error_text(ctx))
      error stop 1
    end if
  end subroutine

  subroutine polar_init(device)
    integer, intent(in) :: device
    call fatal_status(thp_init(ctx, int(device, c_int)), 'thp_init')
  end subroutine

  subroutine polar_setup(natom, box, x, y, z, rpole, polarity, thole, nscaled, ik, dpu, &
                         aewald, cut, buffer, nfft, order, precision)
    integer, intent(in) :: natom, nscaled, nfft(3), order, precision
    real(c_double), intent(in) :: box(3), aewald, cut, buffer
    real(c_double), intent(in), target :: x(natom), y(natom), z(natom), rpole(13, natom)
    real(c_double), intent(in), target :: polarity(natom), thole(natom)
    integer(c_int), intent(in), target :: ik(2, *)
    real(c_double), intent(in), target :: dpu(3, *)
    type(thp_system_desc) :: d
    if (c_associated(sys)) call fatal_status(thp_system_destroy(sys), 'thp_system_destroy')
    n = natom
    d%n = natom
    d%box = box
    d%x = c_loc(x)
    d%y = c_loc(y)
    d%z = c_loc(z)
    d%rpole = c_loc(rpole)
    d%polarity = c_loc(polarity)
    d%thole = c_loc(thole)
    d%nscaled = nscaled
    d%scaled_ik = c_loc(ik(1, 1))
    d%scaled_dpu = c_loc(dpu(1, 1))
    d%aewald = aewald
    d%cut = cut
    d%buffer = buffer
    d%nfft = nfft
    d%order = order
    d%precision = precision
    call fatal_status(thp_system_create(ctx, d, sys), 'thp_system_create')
    if (allocated(uind)) deallocate (uind, uinp)
    allocate (uind(3, n), uinp(3, n))
    uind = 0.0d0
    uinp = 0.0d0
  end subroutine

  subroutine polar_update_coords(x, y, z)
    real(c_double), intent(in) :: x(n), y(n), z(n)
    call fatal_status(thp_system_update_coords(sys, x, y, z), 'thp_system_update_coords')
  end subroutine

  ! nblist(istep): rebuild the Verlet list (cut + buffer, k > i pairs)
  subroutine nblist(istep)
    integer, intent(in) :: istep
    if (istep < 0) return
    call mlistcell
  end subroutine

  subroutine mlistcell
    call fatal_status(thp_nblist(sys, 1_c_int), 'thp_nblist')
  end subroutine

  ! number of pairs of the list (nelst summed over atoms)
  integer(c_int64_t) function nblist_npairs()
    call fatal_status(thp_nblist_get(sys, nblist_npairs, c_null_ptr, 0_c_int64_t), 'thp_nblist_get')
  end function

  subroutine efld0_direct(nrhs, ef)
    integer, intent(in) :: nrhs
    real(c_double), intent(out) :: ef(3, nrhs, n)
    if (nrhs /= 2) error stop 'efld0_direct: nrhs must be 2'
    call fatal_status(thp_efld0_direct(sys, ef), 'thp_efld0_direct')
  end subroutine

  subroutine efld0_recip(cphi)
    real(c_double), intent(out) :: cphi(10, n)
    call fatal_status(thp_efld0_recip(sys, cphi), 'thp_efld0_recip')
  end subroutine

  subroutine tmatxb_pme(nrhs, dodiag, mu, efi)
    integer, intent(in) :: nrhs
    logical, intent(in) :: dodiag
    real(c_double), intent(in) :: mu(3, nrhs, n)
    real(c_double), intent(out) :: efi(3, nrhs, n)
    integer(c_int) :: dd
    if (nrhs /= 2) error stop 'tmatxb_pme: nrhs must be 2'
    dd = 0
    if (dodiag) dd = 1
    call fatal_status(thp_tmatxb(sys, mu, efi, dd), 'thp_tmatxb')
  end subroutine

  ! one rank: murec and dipfieldbis (the other ranks' share) stay unused
  subroutine tmatxbrecip(mu, murec, nrhs, dipfield, dipfieldbis)
    integer, intent(in) :: nrhs
    real(c_double), intent(in) :: mu(3, nrhs, n), murec(3, nrhs, *)
    real(c_double), intent(out) :: dipfield(3, nrhs, n), dipfieldbis(3, nrhs, *)
    if (nrhs /= 2) error stop 'tmatxbrecip: nrhs must be 2'
    call fatal_status(thp_tmatxb_recip(sys, mu, dipfield), 'thp_tmatxb_recip')
  end subroutine

  ! all sites at once: fractional dipoles fuind/fuinp(3,n) onto the grid;
  ! qgrid(2,nfft1,nfft2,nfft3) receives the spread grid
  subroutine grid_uind_site(fuind, fuinp, qgrid)
    real(c_double), intent(in) :: fuind(3, n), fuinp(3, n)
    real(c_double), intent(out), target :: qgrid(*)
    call fatal_status(thp_grid_uind(sys, fuind, fuinp, c_loc(qgrid)), 'thp_grid_uind')
  end subroutine

  ! first derivatives of the induced-dipole potential at every site
  ! (fdip_phi1/fdip_phi2(2:4) of the reference) from a given grid
  subroutine fphi_uind_site(qgrid, fdip_phi1, fdip_phi2)
    real(c_double), intent(in), target :: qgrid(*)
    real(c_double), intent(out) :: fdip_phi1(3, n), fdip_phi2(3, n)
    call fatal_status(thp_fphi_uind(sys, c_loc(qgrid), fdip_phi1, fdip_phi2), 'thp_fphi_uind')
  end subroutine

  ! induced dipoles (uind/uinp) and polarization energy ep of the current
  ! coordinates; polgsf = 1 starts from the previous uind/uinp
  subroutine newinduce_pme2
    type(thp_solve_opts) :: o
    type(thp_solve_stats) :: st
    o%poleps = poleps
    o%politer = politer
    o%guess = 0
    if (polgsf == 1) o%guess = 1
    call fatal_status(thp_induce(sys, o, uind, uinp, ep, st), 'thp_induce')
    niter_last = st%niter
    resnrm_last = st%resnrm
    if (st%status /= THP_OK) write (*, '(a,i0)') ' NEWINDUCE_PME2  --  Warning, PCG not converged, status ', st%status
  end subroutine

  ! global-frame multipoles rpole(13,n) of this step, after rotpole
  subroutine polar_update_multipoles(rpole)
    real(c_double), intent(in) :: rpole(13, n)
    call fatal_status(thp_system_update_multipoles(sys, rpole), 'thp_system_update_multipoles')
  end subroutine

  ! the POLAR-ALG keyword (prmkey.f:396-398): 1 PCG, 2 Jacobi/DIIS
  subroutine polar_set_solver(polalg)
    integer, intent(in) :: polalg
    call fatal_status(thp_set_solver(sys, int(polalg, c_int)), 'thp_set_solver')
  end subroutine

  ! the polarization gradient of the last solve, dep(3,n) in kcal/mol/A
  ! (epolar1c's dep, epolar1.f:57-376).  With local frames (polar_set_frames)
  ! the multipole torques are carried to the frame atoms on the device, as
  ! torque.f does; with global multipoles dep holds the forces at fixed
  ! orientations and torque(3,n), if present, the torques on the sites
  ! (kcal/mol), for the host to distribute.
  subroutine polar_gradient(dep, torque)
    real(c_double), intent(out), target, contiguous :: dep(:,:)
    real(c_double), intent(out), target, contiguous, optional :: torque(:,:)
    if (size(dep, 1) /= 3 .or. size(dep, 2) < n) error stop 'polar_gradient: dep must be (3, n)'
    if (present(torque)) then
      if (size(torque, 1) /= 3 .or. size(torque, 2) < n) error stop 'polar_gradient: torque must be (3, n)'
      call fatal_status(thp_epolar1(sys, dep, c_loc(torque)), 'thp_epolar1')
    else
      call fatal_status(thp_epolar1(sys, dep, c_null_ptr), 'thp_epolar1')
    end if
  end subroutine

  ! local-frame multipoles (kmpole's ipole/pole and the axis atoms) that the
  ! device rotates to the global frame now and after every coordinate update
  ! (rotpole.f:24-339 with chkpole.f:20-98, each step as epolar1.f:113-116
  ! does): axistype 0 None, 1 Z-Only, 2 Z-then-X, 3 Bisector, 4 Z-Bisect,
  ! 5 3-Fold; frame atoms 1-based (0 where the type takes none), pole(13,n)
  ! in the local frame
  subroutine polar_set_frames(axistype, zaxis, xaxis, yaxis, pole)
    integer, intent(in) :: axistype(n), zaxis(n), xaxis(n), yaxis(n)
    real(c_double), intent(in) :: pole(13, n)
    integer(c_int), allocatable :: at(:), zi(:), xi(:), yi(:)
    allocate (at(n), zi(n), xi(n), yi(n))
    at = int(axistype, c_int)
    zi = int(zaxis - 1, c_int)
    xi = int(xaxis - 1, c_int)
    yi = int(yaxis - 1, c_int)
    call fatal_status(thp_system_set_frames(sys, at, zi, xi, yi, pole), 'thp_system_set_frames')
  end subroutine

  ! use_polarshortreal with ewaldshortcut (cutoffs.f:59): the short-range
  ! solves of RESPA1's inner steps; 0 restores the full solve
  subroutine polar_set_shortreal(shortcut)
    real(8), intent(in) :: shortcut
    call fatal_status(thp_set_shortreal(sys, real(shortcut, c_double)), 'thp_set_shortreal')
  end subroutine

  ! the truncated CG options of POLAR-ALG 3 (tcgorder, tcgprec, tcgguess,
  ! tcgpeek, tcgomega; defaults 2, .true., .false., .true., 1.0 in
  ! initprm.f:447-455)
  subroutine polar_set_tcg(tcgorder, tcgprec, tcgguess, tcgpeek, tcgomega)
    integer, intent(in) :: tcgorder
    logical, intent(in) :: tcgprec, tcgguess, tcgpeek
    real(8), intent(in) :: tcgomega
    call fatal_status(thp_set_tcg(sys, int(tcgorder, c_int), merge(1_c_int, 0_c_int, tcgprec), &
                                  merge(1_c_int, 0_c_int, tcgguess), merge(1_c_int, 0_c_int, tcgpeek), &
                                  real(tcgomega, c_double)), 'thp_set_tcg')
  end subroutine

  ! the POLAR-PREDICT keyword (mdinit.f:185-190): 'NONE', 'GEAR', 'ASPC'
  ! or 'LSQR'; the stored solves are cleared
  subroutine polar_set_predictor(polpred)
    character(len=*), intent(in) :: polpred
    integer(c_int) :: m
    select case (polpred)
    case ('NONE')
      m = 0
    case ('GEAR')
      m = 1
    case ('ASPC')
      m = 2
    case ('LSQR')
      m = 3
    case default
      error stop 'polar_set_predictor: unknown predictor'
    end select
    call fatal_status(thp_set_predictor(sys, m), 'thp_set_predictor')
  end subroutine

  subroutine polar_finalize
    if (c_associated(sys)) call fatal_status(thp_system_destroy(sys), 'thp_system_destroy')
    if (c_associated(ctx)) call fatal_status(thp_finalize(ctx), 'thp_finalize')
    sys = c_null_ptr
    ctx = c_null_ptr
  end subroutine

end module thp_polar
