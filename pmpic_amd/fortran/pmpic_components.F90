!> MONC component descriptors of the PMPIC hot path, calling the GPU through pmpic_gpu_mod.
!!
!! The reference registers each component through `<name>_get_descriptor()`, a
!! component_descriptor_type with name, version and initialisation / timestep / finalisation
!! hooks that take the whole model state, `type(model_state_type), target, intent(inout) ::
!! current_state` (model_core/src/components/monc_component.F90:51-89; e.g.
!! components/vort2vel/src/vort2vel.F90:37-43).  The registry runs the integrator group's
!! timestep hooks once per RK stage, current_state%rksteps times per timestep
!! (model_core/src/components/timestepper.F90:89-106), in the order of global_config:43.
!! These descriptors keep the components' names and that signature; their hooks are the ones a
!! maintainer puts in the reference's own component files (see the integration notes), so the
!! model core drives the GPU path unchanged.
module pmpic_components_mod
  use datadefn_mod, only : DEFAULT_PRECISION
  use state_mod, only : model_state_type
  use monc_component_mod, only : component_descriptor_type
  use pmpic_gpu_mod
  implicit none

  private
  public :: vort2vel_get_descriptor, vorticity_tendency_get_descriptor, &
       rk4_integrator_get_descriptor, parcel_splitting_get_descriptor, &
       parcel_mixing_get_descriptor

contains

  type(component_descriptor_type) function vort2vel_get_descriptor()
    vort2vel_get_descriptor%name = "vort2vel"
    vort2vel_get_descriptor%version = 0.1_DEFAULT_PRECISION
    vort2vel_get_descriptor%initialisation => vort2vel_initialisation
    vort2vel_get_descriptor%timestep => vort2vel_timestep
    vort2vel_get_descriptor%finalisation => vort2vel_finalisation
  end function vort2vel_get_descriptor

  type(component_descriptor_type) function vorticity_tendency_get_descriptor()
    vorticity_tendency_get_descriptor%name = "vorticity_tendency"
    vorticity_tendency_get_descriptor%version = 0.1_DEFAULT_PRECISION
    vorticity_tendency_get_descriptor%timestep => vorticity_tendency_timestep
  end function vorticity_tendency_get_descriptor

  type(component_descriptor_type) function rk4_integrator_get_descriptor()
    rk4_integrator_get_descriptor%name = "rk4_integrator"
    rk4_integrator_get_descriptor%version = 0.1_DEFAULT_PRECISION
    rk4_integrator_get_descriptor%initialisation => rk4_integrator_initialisation
    rk4_integrator_get_descriptor%timestep => rk4_integrator_timestep
  end function rk4_integrator_get_descriptor

  type(component_descriptor_type) function parcel_splitting_get_descriptor()
    parcel_splitting_get_descriptor%name = "parcel_splitting"
    parcel_splitting_get_descriptor%version = 0.1_DEFAULT_PRECISION
    parcel_splitting_get_descriptor%timestep => parcel_splitting_timestep
  end function parcel_splitting_get_descriptor

  type(component_descriptor_type) function parcel_mixing_get_descriptor()
    parcel_mixing_get_descriptor%name = "parcel_mixing"
    parcel_mixing_get_descriptor%version = 0.1_DEFAULT_PRECISION
    parcel_mixing_get_descriptor%timestep => parcel_mixing_timestep
  end function parcel_mixing_get_descriptor

  !> after the parcel set-up components (global_config order): the context and the upload
  subroutine vort2vel_initialisation(current_state)
    type(model_state_type), target, intent(inout) :: current_state
    call pmpic_gpu_initialise(current_state)
  end subroutine vort2vel_initialisation

  subroutine vort2vel_finalisation(current_state)
    type(model_state_type), target, intent(inout) :: current_state
    call pmpic_gpu_finalise(current_state)
  end subroutine vort2vel_finalisation

  !> the four RK stages the integrator group runs per timestep
  subroutine rk4_integrator_initialisation(current_state)
    type(model_state_type), target, intent(inout) :: current_state
    current_state%rksteps = 4
  end subroutine rk4_integrator_initialisation

  subroutine vort2vel_timestep(current_state)
    type(model_state_type), target, intent(inout) :: current_state
    call pmpic_gpu_vort2vel(current_state)
  end subroutine vort2vel_timestep

  subroutine vorticity_tendency_timestep(current_state)
    type(model_state_type), target, intent(inout) :: current_state
    call pmpic_gpu_vorticity_tendency(current_state)
  end subroutine vorticity_tendency_timestep

  subroutine rk4_integrator_timestep(current_state)
    type(model_state_type), target, intent(inout) :: current_state
    call pmpic_gpu_rk4_integrator(current_state)
  end subroutine rk4_integrator_timestep

  subroutine parcel_splitting_timestep(current_state)
    type(model_state_type), target, intent(inout) :: current_state
    call pmpic_gpu_parcel_splitting(current_state)
  end subroutine parcel_splitting_timestep

  subroutine parcel_mixing_timestep(current_state)
    type(model_state_type), target, intent(inout) :: current_state
    call pmpic_gpu_parcel_mixing(current_state)
  end subroutine parcel_mixing_timestep

end module pmpic_components_mod
