!> MONC-style component shims of the PMPIC hot path.
!!
!! The reference registers each component through `<name>_get_descriptor()`, a
!! component_descriptor_type with name, version and initialisation / timestep / finalisation
!! procedure pointers taking the whole model state (model_core/src/components/
!! monc_component.F90:51-89; e.g. components/vort2vel/src/vort2vel.F90:37-43).  The registry
!! runs the integrator group's timestep callbacks once per RK stage
!! (model_core/src/components/timestepper.F90:89-106) in the order of global_config:43.
!! Here the same five components keep those names and that calling convention; their callbacks
!! forward to libpmpic_hip.so, where the state lives on the GPU.  pmpic_state_type carries what
!! the path needs from model_state_type (state.F90:38-67): the options, the time and timestep and
!! the parcel count, plus the device context.
module pmpic_components_mod
  use, intrinsic :: iso_c_binding
  use pmpic_hip_mod
  implicit none

  private
  public :: pmpic_state_type, pmpic_component_descriptor_type, vort2vel_get_descriptor, &
       vorticity_tendency_get_descriptor, rk4_integrator_get_descriptor, &
       parcel_splitting_get_descriptor, parcel_mixing_get_descriptor, pmpic_context_init, &
       pmpic_context_finalise, pmpic_integrator_group

  type :: pmpic_state_type
    type(pmpic_config_t) :: cfg
    type(c_ptr) :: ctx = c_null_ptr
    real(c_double) :: time = 0.0_c_double, dtm = 0.0_c_double
    integer(c_int64_t) :: numparcels = 0
    integer :: rksteps = 4
  end type pmpic_state_type

  abstract interface
    subroutine component_callback(current_state)
      import :: pmpic_state_type
      type(pmpic_state_type), target, intent(inout) :: current_state
    end subroutine component_callback
  end interface

  type :: pmpic_component_descriptor_type
    procedure(component_callback), nopass, pointer :: initialisation => null()
    procedure(component_callback), nopass, pointer :: timestep => null()
    procedure(component_callback), nopass, pointer :: finalisation => null()
    character(len=64) :: name = ""
    real(c_double) :: version = 0.1_c_double
  end type pmpic_component_descriptor_type

contains

  type(pmpic_component_descriptor_type) function vort2vel_get_descriptor()
    vort2vel_get_descriptor%name = "vort2vel"
    vort2vel_get_descriptor%timestep => vort2vel_timestep
  end function

  type(pmpic_component_descriptor_type) function vorticity_tendency_get_descriptor()
    vorticity_tendency_get_descriptor%name = "vorticity_tendency"
    vorticity_tendency_get_descriptor%timestep => vorticity_tendency_timestep
  end function

  type(pmpic_component_descriptor_type) function rk4_integrator_get_descriptor()
    rk4_integrator_get_descriptor%name = "rk4_integrator"
    rk4_integrator_get_descriptor%timestep => rk4_integrator_timestep
  end function

  type(pmpic_component_descriptor_type) function parcel_splitting_get_descriptor()
    parcel_splitting_get_descriptor%name = "parcel_splitting"
    parcel_splitting_get_descriptor%timestep => parcel_splitting_timestep
  end function

  type(pmpic_component_descriptor_type) function parcel_mixing_get_descriptor()
    parcel_mixing_get_descriptor%name = "parcel_mixing"
    parcel_mixing_get_descriptor%timestep => parcel_mixing_timestep
  end function

  !> setup_grid / parcelsetup_base / rk4_integrator / vort2vel initialisation in one call
  subroutine pmpic_context_init(state)
    type(pmpic_state_type), target, intent(inout) :: state
    call pmpic_check(pmpic_create(state%cfg, state%ctx), "pmpic_create")
    state%dtm = state%cfg%dtm
  end subroutine

  subroutine pmpic_context_finalise(state)
    type(pmpic_state_type), target, intent(inout) :: state
    call pmpic_check(pmpic_destroy(state%ctx), "pmpic_destroy")
    state%ctx = c_null_ptr
  end subroutine

  subroutine sync_state(state)
    type(pmpic_state_type), target, intent(inout) :: state
    real(c_double) :: t
    call pmpic_check(pmpic_get_state(state%ctx, state%numparcels, t, state%dtm), "pmpic_get_state")
  end subroutine

  subroutine vort2vel_timestep(current_state)
    type(pmpic_state_type), target, intent(inout) :: current_state
    call pmpic_check(pmpic_vort2vel(current_state%ctx), "vort2vel")
  end subroutine

  subroutine vorticity_tendency_timestep(current_state)
    type(pmpic_state_type), target, intent(inout) :: current_state
    call pmpic_check(pmpic_vorticity_tendency(current_state%ctx), "vorticity_tendency")
  end subroutine

  subroutine rk4_integrator_timestep(current_state)
    type(pmpic_state_type), target, intent(inout) :: current_state
    call pmpic_check(pmpic_rk4_integrator(current_state%ctx), "rk4_integrator")
    call sync_state(current_state)
  end subroutine

  subroutine parcel_splitting_timestep(current_state)
    type(pmpic_state_type), target, intent(inout) :: current_state
    call pmpic_check(pmpic_parcel_splitting(current_state%ctx), "parcel_splitting")
    call sync_state(current_state)
  end subroutine

  subroutine parcel_mixing_timestep(current_state)
    type(pmpic_state_type), target, intent(inout) :: current_state
    call pmpic_check(pmpic_parcel_mixing(current_state%ctx), "parcel_mixing")
    call sync_state(current_state)
  end subroutine

  !> one timestep: the rk_integrated group run rksteps times (timestepper.F90:89-106), then
  !> time += dtm (monc.F90:223-227)
  subroutine pmpic_integrator_group(state, group)
    type(pmpic_state_type), target, intent(inout) :: state
    type(pmpic_component_descriptor_type), intent(in) :: group(:)
    integer :: s, c
    do s = 1, state%rksteps
      do c = 1, size(group)
        if (associated(group(c)%timestep)) call group(c)%timestep(state)
      end do
    end do
    state%time = state%time + state%dtm
    call pmpic_check(pmpic_set_time(state%ctx, state%time, state%dtm), "pmpic_set_time")
  end subroutine

end module pmpic_components_mod
