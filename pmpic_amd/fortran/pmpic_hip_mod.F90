!> iso_c_binding interface to libpmpic_hip.so (include/pmpic_hip.h).
!!
!! One interface per C entry point; arguments are plain C types and contiguous arrays passed by
!! address, so the reference's Fortran state arrays (parcel_type SoA, prognostic data(z,y,x))
!! can be handed over without copies.  pmpic_check turns a non-zero status into the same
!! `error stop` the reference raises for that condition.
module pmpic_hip_mod
  use, intrinsic :: iso_c_binding
  implicit none

  private
  public :: pmpic_config_t, pmpic_check, pmpic_message
  public :: pmpic_create, pmpic_destroy, pmpic_set_parcels, pmpic_get_parcels, pmpic_get_state, &
       pmpic_set_time, pmpic_vort2vel, pmpic_vorticity_tendency, pmpic_rk4_integrator, &
       pmpic_parcel_splitting, pmpic_parcel_mixing, pmpic_step, pmpic_par2grid, pmpic_grid2par, &
       pmpic_get_grid, pmpic_conservation_sums, pmpic_last_counts, pmpic_writebinary_grids

  !> pmpic_config of include/pmpic_hip.h
  type, bind(C) :: pmpic_config_t
    integer(c_int32_t) :: nx, ny, nz
    real(c_double) :: xmin, xmax, ymin, ymax, zmin, zmax
    integer(c_int64_t) :: max_parcels
    real(c_double) :: dtm, dtmax, timestep_prefactor
    integer(c_int32_t) :: precision, splitting_enabled, mixing_enabled, device
  end type pmpic_config_t

  interface
    integer(c_int) function pmpic_create(cfg, ctx) bind(C, name="pmpic_create")
      import :: c_int, c_ptr, pmpic_config_t
      type(pmpic_config_t), intent(in) :: cfg
      type(c_ptr), intent(out) :: ctx
    end function
    integer(c_int) function pmpic_destroy(ctx) bind(C, name="pmpic_destroy")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
    end function
    type(c_ptr) function pmpic_status_message(status) bind(C, name="pmpic_status_message")
      import :: c_int, c_ptr
      integer(c_int), value :: status
    end function
    integer(c_int) function pmpic_set_parcels(ctx, n, fields) bind(C, name="pmpic_set_parcels")
      import :: c_int, c_ptr, c_int64_t
      type(c_ptr), value :: ctx
      integer(c_int64_t), value :: n
      type(c_ptr), intent(in) :: fields(17)
    end function
    integer(c_int) function pmpic_get_parcels(ctx, n, fields) bind(C, name="pmpic_get_parcels")
      import :: c_int, c_ptr, c_int64_t
      type(c_ptr), value :: ctx
      integer(c_int64_t), intent(out) :: n
      type(c_ptr), intent(in) :: fields(17)
    end function
    integer(c_int) function pmpic_get_state(ctx, n, time, dtm) bind(C, name="pmpic_get_state")
      import :: c_int, c_ptr, c_int64_t, c_double
      type(c_ptr), value :: ctx
      integer(c_int64_t), intent(out) :: n
      real(c_double), intent(out) :: time, dtm
    end function
    integer(c_int) function pmpic_set_time(ctx, time, dtm) bind(C, name="pmpic_set_time")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      real(c_double), value :: time, dtm
    end function
    integer(c_int) function pmpic_vort2vel(ctx) bind(C, name="pmpic_vort2vel")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
    end function
    integer(c_int) function pmpic_vorticity_tendency(ctx) bind(C, name="pmpic_vorticity_tendency")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
    end function
    integer(c_int) function pmpic_rk4_integrator(ctx) bind(C, name="pmpic_rk4_integrator")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
    end function
    integer(c_int) function pmpic_parcel_splitting(ctx) bind(C, name="pmpic_parcel_splitting")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
    end function
    integer(c_int) function pmpic_parcel_mixing(ctx) bind(C, name="pmpic_parcel_mixing")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
    end function
    integer(c_int) function pmpic_step(ctx, nsteps) bind(C, name="pmpic_step")
      import :: c_int, c_ptr, c_int32_t
      type(c_ptr), value :: ctx
      integer(c_int32_t), value :: nsteps
    end function
    integer(c_int) function pmpic_par2grid(ctx, var, grid) bind(C, name="pmpic_par2grid")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      real(c_double), intent(in) :: var(*)
      real(c_double), intent(out) :: grid(*)
    end function
    integer(c_int) function pmpic_grid2par(ctx, grid, var) bind(C, name="pmpic_grid2par")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      real(c_double), intent(in) :: grid(*)
      real(c_double), intent(out) :: var(*)
    end function
    integer(c_int) function pmpic_get_grid(ctx, id, grid) bind(C, name="pmpic_get_grid")
      import :: c_int, c_ptr, c_double, c_int32_t
      type(c_ptr), value :: ctx
      integer(c_int32_t), value :: id
      real(c_double), intent(out) :: grid(*)
    end function
    integer(c_int) function pmpic_writebinary_grids(ctx) bind(C, name="pmpic_writebinary_grids")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
    end function
    integer(c_int) function pmpic_conservation_sums(ctx, out6) bind(C, name="pmpic_conservation_sums")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      real(c_double), intent(out) :: out6(6)
    end function
    integer(c_int) function pmpic_last_counts(ctx, out3) bind(C, name="pmpic_last_counts")
      import :: c_int, c_ptr, c_int64_t
      type(c_ptr), value :: ctx
      integer(c_int64_t), intent(out) :: out3(3)
    end function
  end interface

contains

  !> text of a status code
  function pmpic_message(status) result(msg)
    integer(c_int), intent(in) :: status
    character(len=:), allocatable :: msg
    type(c_ptr) :: p
    character(kind=c_char), pointer :: chars(:)
    integer :: n
    p = pmpic_status_message(status)
    call c_f_pointer(p, chars, [256])
    n = 0
    do while (n < 256)
      if (chars(n + 1) == c_null_char) exit
      n = n + 1
    end do
    allocate(character(len=n) :: msg)
    msg = transfer(chars(1:n), msg)
  end function pmpic_message

  !> the reference fails with `error stop`; so does the shim on any non-zero status
  subroutine pmpic_check(status, where)
    integer(c_int), intent(in) :: status
    character(len=*), intent(in) :: where
    if (status /= 0) then
      print *, trim(where), ": ", pmpic_message(status)
      error stop 1
    end if
  end subroutine pmpic_check

end module pmpic_hip_mod
