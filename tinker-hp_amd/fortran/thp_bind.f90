! ISO_C_BINDING interfaces to libthp (include/thp.h).
!
! The derived types mirror the C structs field by field (bind(C) layout);
! arrays are passed as plain addresses, 0-based atom indices at the ABI,
! every function returns an integer status (0 = ok).
module thp_bind
  use iso_c_binding
  implicit none

  integer(c_int), parameter :: THP_OK = 0
  integer(c_int), parameter :: THP_PREC_FP64 = 0, THP_PREC_MIXED = 1

  type, bind(C) :: thp_system_desc
    integer(c_int) :: n
    real(c_double) :: box(3)
    type(c_ptr) :: x, y, z, rpole, polarity, thole
    integer(c_int) :: nscaled
    type(c_ptr) :: scaled_ik, scaled_dpu
    real(c_double) :: aewald, cut, buffer
    integer(c_int) :: nfft(3), order, precision
  end type

  type, bind(C) :: thp_solve_opts
    real(c_double) :: poleps
    integer(c_int) :: politer, guess
  end type

  type, bind(C) :: thp_solve_stats
    integer(c_int) :: niter, status
    real(c_double) :: resnrm, ep
    integer(c_int) :: n_indefinite
    real(c_double) :: t_field_ms, t_solve_ms
  end type

  interface
    integer(c_int) function thp_init(ctx, device) bind(C, name="thp_init")
      import :: c_int, c_ptr
      type(c_ptr), intent(out) :: ctx
      integer(c_int), value :: device
    end function
    integer(c_int) function thp_finalize(ctx) bind(C, name="thp_finalize")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
    end function
    type(c_ptr) function thp_last_error(ctx) bind(C, name="thp_last_error")
      import :: c_ptr
      type(c_ptr), value :: ctx
    end function
    integer(c_int) function thp_system_create(ctx, desc, sys) bind(C, name="thp_system_create")
      import :: c_int, c_ptr, thp_system_desc
      type(c_ptr), value :: ctx
      type(thp_system_desc), intent(in) :: desc
      type(c_ptr), intent(out) :: sys
    end function
    integer(c_int) function thp_system_destroy(sys) bind(C, name="thp_system_destroy")
      import :: c_int, c_ptr
      type(c_ptr), value :: sys
    end function
    integer(c_int) function thp_system_update_coords(sys, x, y, z) bind(C, name="thp_system_update_coords")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: sys
      real(c_double), intent(in) :: x(*), y(*), z(*)
    end function
    integer(c_int) function thp_nblist(sys, force) bind(C, name="thp_nblist")
      import :: c_int, c_ptr
      type(c_ptr), value :: sys
      integer(c_int), value :: force
    end function
    integer(c_int) function thp_nblist_get(sys, npairs, ik, cap) bind(C, name="thp_nblist_get")
      import :: c_int, c_ptr, c_int64_t
      type(c_ptr), value :: sys
      integer(c_int64_t), intent(out) :: npairs
      type(c_ptr), value :: ik
      integer(c_int64_t), value :: cap
    end function
    integer(c_int) function thp_efld0_direct(sys, ef) bind(C, name="thp_efld0_direct")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: sys
      real(c_double), intent(out) :: ef(*)
    end function
    integer(c_int) function thp_efld0_recip(sys, cphi) bind(C, name="thp_efld0_recip")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: sys
      real(c_double), intent(out) :: cphi(*)
    end function
    integer(c_int) function thp_epolar1(sys, dep, torque) bind(C, name="thp_epolar1")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: sys
      real(c_double), intent(out) :: dep(*)
      type(c_ptr), value :: torque
    end function
    integer(c_int) function thp_tmatxb(sys, mu, efi, dodiag) bind(C, name="thp_tmatxb")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: sys
      real(c_double), intent(in) :: mu(*)
      real(c_double), intent(out) :: efi(*)
      integer(c_int), value :: dodiag
    end function
    integer(c_int) function thp_tmatxb_recip(sys, mu, dipfield) bind(C, name="thp_tmatxb_recip")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: sys
      real(c_double), intent(in) :: mu(*)
      real(c_double), intent(out) :: dipfield(*)
    end function
    integer(c_int) function thp_grid_uind(sys, fuind, fuinp, qgrid) bind(C, name="thp_grid_uind")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: sys
      real(c_double), intent(in) :: fuind(*), fuinp(*)
      type(c_ptr), value :: qgrid
    end function
    integer(c_int) function thp_fphi_uind(sys, qgrid, phi1, phi2) bind(C, name="thp_fphi_uind")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: sys
      type(c_ptr), value :: qgrid
      real(c_double), intent(out) :: phi1(*), phi2(*)
    end function
    integer(c_int) function thp_pme_convolve(sys) bind(C, name="thp_pme_convolve")
      import :: c_int, c_ptr
      type(c_ptr), value :: sys
    end function
    integer(c_int) function thp_induce(sys, opts, uind, uinp, ep, stats) bind(C, name="thp_induce")
      import :: c_int, c_ptr, c_double, thp_solve_opts, thp_solve_stats
      type(c_ptr), value :: sys
      type(thp_solve_opts), intent(in) :: opts
      real(c_double), intent(inout) :: uind(*), uinp(*)
      real(c_double), intent(out) :: ep
      type(thp_solve_stats), intent(out) :: stats
    end function
    ! the same solve with the dipoles in the layout of the reference's
    ! mu(3,nrhs,npolebloc) work array (newinduce_pme2.f:71)
    integer(c_int) function thp_induce_mu(sys, opts, mu, ep, stats) bind(C, name="thp_induce_mu")
      import :: c_int, c_ptr, c_double, thp_solve_opts, thp_solve_stats
      type(c_ptr), value :: sys
      type(thp_solve_opts), intent(in) :: opts
      real(c_double), intent(inout) :: mu(*)
      real(c_double), intent(out) :: ep
      type(thp_solve_stats), intent(out) :: stats
    end function
    integer(c_int) function thp_system_update_multipoles(sys, rpole) bind(C, name="thp_system_update_multipoles")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: sys
      real(c_double), intent(in) :: rpole(*)
    end function
    integer(c_int) function thp_set_solver(sys, polalg) bind(C, name="thp_set_solver")
      import :: c_int, c_ptr
      type(c_ptr), value :: sys
      integer(c_int), value :: polalg
    end function
    integer(c_int) function thp_set_tcg(sys, order, precond, guess, peek, omega) bind(C, name="thp_set_tcg")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: sys
      integer(c_int), value :: order, precond, guess, peek
      real(c_double), value :: omega
    end function
    integer(c_int) function thp_set_shortreal(sys, shortcut) bind(C, name="thp_set_shortreal")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: sys
      real(c_double), value :: shortcut
    end function
    integer(c_int) function thp_set_predictor(sys, method) bind(C, name="thp_set_predictor")
      import :: c_int, c_ptr
      type(c_ptr), value :: sys
      integer(c_int), value :: method
    end function
    integer(c_int) function thp_system_set_frames(sys, axistype, zaxis, xaxis, yaxis, pole) &
        bind(C, name="thp_system_set_frames")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: sys
      integer(c_int), intent(in) :: axistype(*), zaxis(*), xaxis(*), yaxis(*)
      real(c_double), intent(in) :: pole(*)
    end function
    ! device-pointer entry points ordered on the caller's stream (arrays are
    ! device addresses, e.g. from !$acc host_data use_device; stream is a
    ! hipStream_t, c_null_ptr for a synchronous call)
    integer(c_int) function thp_efld0_direct_dev(sys, ef, stream) bind(C, name="thp_efld0_direct_dev")
      import :: c_int, c_ptr
      type(c_ptr), value :: sys, ef, stream
    end function
    integer(c_int) function thp_efld0_recip_dev(sys, cphi, stream) bind(C, name="thp_efld0_recip_dev")
      import :: c_int, c_ptr
      type(c_ptr), value :: sys, cphi, stream
    end function
    integer(c_int) function thp_tmatxb_dev_async(sys, mu, efi, dodiag, stream) bind(C, name="thp_tmatxb_dev_async")
      import :: c_int, c_ptr
      type(c_ptr), value :: sys, mu, efi, stream
      integer(c_int), value :: dodiag
    end function
    integer(c_int) function thp_tmatxb_recip_dev(sys, mu, dipfield, stream) bind(C, name="thp_tmatxb_recip_dev")
      import :: c_int, c_ptr
      type(c_ptr), value :: sys, mu, dipfield, stream
    end function
    integer(c_int) function thp_grid_uind_dev(sys, fuind, fuinp, qgrid, stream) bind(C, name="thp_grid_uind_dev")
      import :: c_int, c_ptr
      type(c_ptr), value :: sys, fuind, fuinp, qgrid, stream
    end function
    integer(c_int) function thp_fphi_uind_dev(sys, qgrid, phi1, phi2, stream) bind(C, name="thp_fphi_uind_dev")
      import :: c_int, c_ptr
      type(c_ptr), value :: sys, qgrid, phi1, phi2, stream
    end function
    ! the gradient without device allocation (scratch in a file-scope
    ! device array of libthp; at most THP_NOALLOC_MAX_ATOMS = 131072 atoms)
    integer(c_int) function thp_epolar1_noalloc(sys, dep, torque, stream) bind(C, name="thp_epolar1_noalloc")
      import :: c_int, c_ptr
      type(c_ptr), value :: sys, dep, torque, stream
    end function
    integer(c_int) function thp_make_system(n, box, nsol, radius, seed, x, y, z, rpole, polarity, &
                                            thole, nscaled, ik, dpu) bind(C, name="thp_make_system")
      import :: c_int, c_double, c_int64_t
      integer(c_int), value :: n, nsol
      real(c_double), value :: box, radius
      integer(c_int64_t), value :: seed
      real(c_double), intent(out) :: x(*), y(*), z(*), rpole(*), polarity(*), thole(*)
      integer(c_int), intent(out) :: nscaled
      integer(c_int), intent(out) :: ik(*)
      real(c_double), intent(out) :: dpu(*)
    end function
  end interface

contains

  ! C string of thp_last_error as a Fortran string
  function thp_error_text(ctx) result(msg)
    type(c_ptr), intent(in) :: ctx
    character(len=512) :: msg
    character(kind=c_char), pointer :: p(:)
    integer :: i
    msg = ''
    call c_f_pointer(thp_last_error(ctx), p, [512])
    do i = 1, 512
      if (p(i) == c_null_char) exit
      msg(i:i) = p(i)
    end do
  end function

end module thp_bind
