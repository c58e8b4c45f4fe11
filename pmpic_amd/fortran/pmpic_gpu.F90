!> The GPU context of the PMPIC hot path inside a MONC-structured model: this module owns the
!! libpmpic_hip.so context and keeps the model state's host parcel arrays (current_state%parcels,
!! model_core/src/parcels/parcels.F90) consistent with the parcels the GPU steps.
!!
!! * pmpic_gpu_initialise creates the context from the configuration the reference reads
!!   (setup_grid.F90:44-104: nx, ny, nz, xmin .. zmax; parcelsetup_base.F90:49: max_parcels;
!!   rk4_integrator / vorticity_tendency: dtm, dtmax, timestep_prefactor; the parcel_splitting
!!   and parcel_mixing components' *_enabled switches) and uploads the host parcels, which a
!!   parcel set-up component (readfromfile_parcelsetup, plume_parcelsetup, ...) has filled.
!! * The five component callbacks (vort2vel, vorticity_tendency, rk4_integrator,
!!   parcel_splitting, parcel_mixing) forward to the library; after the ones that change the
!!   parcel count or the timestep, numparcels_local and dtm are refreshed from the library's
!!   host copies (no device transfer).
!! * The host arrays then hold stale values: pmpic_gpu_to_host copies the parcels back (grown
!!   to the count when needed) and must run before anything reads current_state%parcels, as
!!   writebinary does (writebinary.F90:149-209); pmpic_gpu_to_device sends host arrays that
!!   something else changed (a restart read after the context exists).
!! * pmpic_gpu_grids_to_host fills the gridded fields writebinary dumps (u, v, w, p, q, r, b,
!!   hg, hgliq, vol; writebinary.F90:211-299) after the library's grid computation of
!!   write_grids_to_file (writebinary.F90:217-233), each (nz+1, ny+4, nx+4) with the halos.
!! * pmpic_gpu_finalise copies the parcels back and destroys the context.
!!
!! Errors from the library stop the run with its status message (pmpic_check).
module pmpic_gpu_mod
  use, intrinsic :: iso_c_binding
  use datadefn_mod, only : DEFAULT_PRECISION, PARCEL_INTEGER
  use state_mod, only : model_state_type
  use optionsdatabase_mod, only : options_get_real, options_get_integer, options_get_logical, &
       options_has_key
  use pmpic_hip_mod
  implicit none

  private
  public :: pmpic_gpu_initialise, pmpic_gpu_finalise, pmpic_gpu_to_host, pmpic_gpu_to_device, &
       pmpic_gpu_grids_to_host, &
       pmpic_gpu_vort2vel, pmpic_gpu_vorticity_tendency, pmpic_gpu_rk4_integrator, &
       pmpic_gpu_parcel_splitting, pmpic_gpu_parcel_mixing, pmpic_gpu_context

  type(c_ptr), save :: ctx = c_null_ptr
  logical, save :: host_stale = .false.  ! the GPU holds parcels newer than current_state%parcels

contains

  !> the library context (null before pmpic_gpu_initialise), for other library calls
  type(c_ptr) function pmpic_gpu_context()
    pmpic_gpu_context = ctx
  end function pmpic_gpu_context

  logical function enabled(current_state, key, default)
    type(model_state_type), intent(in) :: current_state
    character(len=*), intent(in) :: key
    logical, intent(in) :: default
    enabled = default
    if (options_has_key(current_state%options_database, key)) &
         enabled = options_get_logical(current_state%options_database, key)
  end function enabled

  subroutine pmpic_gpu_initialise(current_state)
    type(model_state_type), target, intent(inout) :: current_state
    type(pmpic_config_t) :: cfg
    if (c_associated(ctx)) return
    associate (db => current_state%options_database)
      cfg%nx = options_get_integer(db, "nx")
      cfg%ny = options_get_integer(db, "ny")
      cfg%nz = options_get_integer(db, "nz")
      cfg%xmin = options_get_real(db, "xmin")
      cfg%xmax = options_get_real(db, "xmax")
      cfg%ymin = options_get_real(db, "ymin")
      cfg%ymax = options_get_real(db, "ymax")
      cfg%zmin = options_get_real(db, "zmin")
      cfg%zmax = options_get_real(db, "zmax")
      cfg%dtm = options_get_real(db, "dtm")
      cfg%dtmax = options_get_real(db, "dtmax")
      cfg%timestep_prefactor = options_get_real(db, "timestep_prefactor")
      cfg%precision = 0
      if (options_has_key(db, "precision")) cfg%precision = options_get_integer(db, "precision")
    end associate
    cfg%max_parcels = current_state%parcels%maxparcels_local
    cfg%splitting_enabled = merge(1, 0, enabled(current_state, "parcel_splitting_enabled", .true.))
    cfg%mixing_enabled = merge(1, 0, enabled(current_state, "parcel_mixing_enabled", .true.))
    cfg%device = 0
    call pmpic_check(pmpic_create(cfg, ctx), "pmpic_create")
    if (current_state%dtm == 0.0_DEFAULT_PRECISION) current_state%dtm = cfg%dtm
    call pmpic_gpu_to_device(current_state)
  end subroutine pmpic_gpu_initialise

  !> the host parcels (numparcels_local of them), time and dtm to the GPU
  subroutine pmpic_gpu_to_device(current_state)
    type(model_state_type), target, intent(inout) :: current_state
    type(c_ptr) :: ptrs(17)
    call field_pointers(current_state, ptrs)
    call pmpic_check(pmpic_set_parcels(ctx, int(current_state%parcels%numparcels_local, &
         c_int64_t), ptrs), "pmpic_set_parcels")
    call pmpic_check(pmpic_set_time(ctx, current_state%time, current_state%dtm), "pmpic_set_time")
    host_stale = .false.
  end subroutine pmpic_gpu_to_device

  !> the GPU's parcels into the host arrays (grown when the count exceeds them)
  subroutine pmpic_gpu_to_host(current_state)
    type(model_state_type), target, intent(inout) :: current_state
    type(c_ptr) :: ptrs(17)
    integer(c_int64_t) :: n
    real(c_double) :: t, dtm
    if (.not. c_associated(ctx) .or. .not. host_stale) return
    call pmpic_check(pmpic_get_state(ctx, n, t, dtm), "pmpic_get_state")
    if (n > size(current_state%parcels%x, kind=c_int64_t)) call grow(current_state, n)
    call field_pointers(current_state, ptrs)
    call pmpic_check(pmpic_get_parcels(ctx, n, ptrs), "pmpic_get_parcels")
    current_state%parcels%numparcels_local = n
    host_stale = .false.
  end subroutine pmpic_gpu_to_host

  subroutine pmpic_gpu_grids_to_host(current_state)
    type(model_state_type), target, intent(inout) :: current_state
    integer :: nx, ny, nz
    if (.not. c_associated(ctx)) return
    nx = options_get_integer(current_state%options_database, "nx")
    ny = options_get_integer(current_state%options_database, "ny")
    nz = options_get_integer(current_state%options_database, "nz")
    call pmpic_check(pmpic_writebinary_grids(ctx), "pmpic_writebinary_grids")
    call get1(current_state%u%data, 0);  call get1(current_state%v%data, 1)
    call get1(current_state%w%data, 2);  call get1(current_state%p%data, 3)
    call get1(current_state%q%data, 4);  call get1(current_state%r%data, 5)
    call get1(current_state%b%data, 6);  call get1(current_state%hg%data, 7)
    call get1(current_state%hgliq%data, 8); call get1(current_state%vol%data, 9)
  contains
    subroutine get1(a, id)
      real(kind=DEFAULT_PRECISION), allocatable, intent(inout) :: a(:,:,:)
      integer, intent(in) :: id
      if (.not. allocated(a)) allocate(a(nz + 1, ny + 4, nx + 4))
      call pmpic_check(pmpic_get_grid(ctx, int(id, c_int32_t), a), "pmpic_get_grid")
    end subroutine get1
  end subroutine pmpic_gpu_grids_to_host

  subroutine pmpic_gpu_finalise(current_state)
    type(model_state_type), target, intent(inout) :: current_state
    if (.not. c_associated(ctx)) return
    call pmpic_gpu_to_host(current_state)
    call pmpic_check(pmpic_destroy(ctx), "pmpic_destroy")
    ctx = c_null_ptr
  end subroutine pmpic_gpu_finalise

  subroutine pmpic_gpu_vort2vel(current_state)
    type(model_state_type), target, intent(inout) :: current_state
    if (.not. c_associated(ctx)) call pmpic_gpu_initialise(current_state)
    call pmpic_check(pmpic_vort2vel(ctx), "vort2vel")
  end subroutine pmpic_gpu_vort2vel

  subroutine pmpic_gpu_vorticity_tendency(current_state)
    type(model_state_type), target, intent(inout) :: current_state
    if (.not. c_associated(ctx)) call pmpic_gpu_initialise(current_state)
    call pmpic_check(pmpic_vorticity_tendency(ctx), "vorticity_tendency")
  end subroutine pmpic_gpu_vorticity_tendency

  subroutine pmpic_gpu_rk4_integrator(current_state)
    type(model_state_type), target, intent(inout) :: current_state
    if (.not. c_associated(ctx)) call pmpic_gpu_initialise(current_state)
    call pmpic_check(pmpic_rk4_integrator(ctx), "rk4_integrator")
    call after_parcel_update(current_state)
  end subroutine pmpic_gpu_rk4_integrator

  subroutine pmpic_gpu_parcel_splitting(current_state)
    type(model_state_type), target, intent(inout) :: current_state
    if (.not. c_associated(ctx)) call pmpic_gpu_initialise(current_state)
    call pmpic_check(pmpic_parcel_splitting(ctx), "parcel_splitting")
    call after_parcel_update(current_state)
  end subroutine pmpic_gpu_parcel_splitting

  subroutine pmpic_gpu_parcel_mixing(current_state)
    type(model_state_type), target, intent(inout) :: current_state
    if (.not. c_associated(ctx)) call pmpic_gpu_initialise(current_state)
    call pmpic_check(pmpic_parcel_mixing(ctx), "parcel_mixing")
    call after_parcel_update(current_state)
  end subroutine pmpic_gpu_parcel_mixing

  !> the count and timestep the model core reads between callbacks (host copies kept by the
  !> library); the host parcel arrays are stale from here on
  subroutine after_parcel_update(current_state)
    type(model_state_type), target, intent(inout) :: current_state
    integer(c_int64_t) :: n
    real(c_double) :: t, dtm
    call pmpic_check(pmpic_get_state(ctx, n, t, dtm), "pmpic_get_state")
    current_state%parcels%numparcels_local = n
    current_state%dtm = dtm
    host_stale = .true.
  end subroutine after_parcel_update

  subroutine field_pointers(current_state, ptrs)
    type(model_state_type), target, intent(inout) :: current_state
    type(c_ptr), intent(out) :: ptrs(17)
    associate (p => current_state%parcels)
      ptrs(1) = c_loc(p%x(1));       ptrs(2) = c_loc(p%y(1));       ptrs(3) = c_loc(p%z(1))
      ptrs(4) = c_loc(p%p(1));       ptrs(5) = c_loc(p%q(1));       ptrs(6) = c_loc(p%r(1))
      ptrs(7) = c_loc(p%dxdt(1));    ptrs(8) = c_loc(p%dydt(1));    ptrs(9) = c_loc(p%dzdt(1))
      ptrs(10) = c_loc(p%dpdt(1));   ptrs(11) = c_loc(p%dqdt(1));   ptrs(12) = c_loc(p%drdt(1))
      ptrs(13) = c_loc(p%h(1));      ptrs(14) = c_loc(p%b(1));      ptrs(15) = c_loc(p%vol(1))
      ptrs(16) = c_loc(p%stretch(1)); ptrs(17) = c_loc(p%tag(1))
    end associate
  end subroutine field_pointers

  !> reallocate the 17 host arrays to hold n parcels, keeping the first numparcels_local
  subroutine grow(current_state, n)
    type(model_state_type), target, intent(inout) :: current_state
    integer(c_int64_t), intent(in) :: n
    integer(c_int64_t) :: m
    m = current_state%parcels%numparcels_local
    associate (p => current_state%parcels)
      call grow1(p%x, n, m);    call grow1(p%y, n, m);    call grow1(p%z, n, m)
      call grow1(p%p, n, m);    call grow1(p%q, n, m);    call grow1(p%r, n, m)
      call grow1(p%dxdt, n, m); call grow1(p%dydt, n, m); call grow1(p%dzdt, n, m)
      call grow1(p%dpdt, n, m); call grow1(p%dqdt, n, m); call grow1(p%drdt, n, m)
      call grow1(p%h, n, m);    call grow1(p%b, n, m);    call grow1(p%vol, n, m)
      call grow1(p%stretch, n, m); call grow1(p%tag, n, m)
    end associate
    current_state%parcels%maxparcels_local = n
  end subroutine grow

  subroutine grow1(a, n, keep)
    real(kind=DEFAULT_PRECISION), allocatable, intent(inout) :: a(:)
    integer(c_int64_t), intent(in) :: n, keep
    real(kind=DEFAULT_PRECISION), allocatable :: b(:)
    allocate(b(n))
    b = 0.0_DEFAULT_PRECISION
    if (allocated(a) .and. keep > 0) b(1:keep) = a(1:keep)
    call move_alloc(b, a)
  end subroutine grow1

end module pmpic_gpu_mod
