!> Build stand-ins for the reference's own modules, so that the component shims and the context
!! module compile and run here without the reference tree.  Each keeps the module name, type
!! name and the members the shims use of the reference module it stands in for:
!!
!!   datadefn_mod        model_core/src/datadefn.F90 (DEFAULT_PRECISION, PARCEL_INTEGER)
!!   collections_mod     the hashmap_type the options live in (a flat key/value list here)
!!   optionsdatabase_mod options_get_real / _integer / _logical (and options_add to fill it)
!!   grids_mod           model_core/src/grid/grids.F90:14,55-63 (Z/Y/X_INDEX, global_grid_type)
!!   parcel_mod          model_core/src/parcels/parcels.F90 (parcel_type: counts and the 17
!!                       parcel arrays)
!!   prognostics_mod     prognostic_field_type (its data array, (z, y, x) with the halos)
!!   state_mod           model_core/src/state.F90:38-67 (model_state_type: options, grid, time,
!!                       dtm, rksteps, iterations, timestep, the gridded fields u .. vol,
!!                       parcels)
!!   monc_component_mod  model_core/src/components/monc_component.F90:51-89
!!                       (component_descriptor_type and its callback signature)
!!
!! In the reference tree these modules exist already and this file is not compiled; the shims
!! use only the names listed here, so they build against either.
module datadefn_mod
  use, intrinsic :: iso_fortran_env, only : int64
  implicit none
  integer, public, parameter :: DOUBLE_PRECISION = selected_real_kind(15, 307)
  integer, public, parameter :: DEFAULT_PRECISION = DOUBLE_PRECISION
  integer, public, parameter :: LONG_INTEGER = int64
  integer, public, parameter :: PARCEL_INTEGER = LONG_INTEGER
end module datadefn_mod

module collections_mod
  implicit none
  private
  type, public :: hashmap_type
    character(len=64), allocatable :: keys(:)
    character(len=64), allocatable :: values(:)
    integer :: n = 0
  end type hashmap_type
end module collections_mod

module optionsdatabase_mod
  use datadefn_mod, only : DEFAULT_PRECISION
  use collections_mod, only : hashmap_type
  implicit none
  private
  public :: options_add, options_has_key, options_get_real, options_get_integer, options_get_logical

contains

  !> key = value (text, as a config.mcf line holds it); a later value replaces an earlier one
  subroutine options_add(db, key, value)
    type(hashmap_type), intent(inout) :: db
    character(len=*), intent(in) :: key, value
    character(len=64), allocatable :: k2(:), v2(:)
    integer :: i
    do i = 1, db%n
      if (trim(db%keys(i)) == key) then
        db%values(i) = value
        return
      end if
    end do
    if (.not. allocated(db%keys)) then
      allocate(db%keys(16), db%values(16))
    else if (db%n == size(db%keys)) then
      allocate(k2(2 * db%n), v2(2 * db%n))
      k2(1:db%n) = db%keys(1:db%n)
      v2(1:db%n) = db%values(1:db%n)
      call move_alloc(k2, db%keys)
      call move_alloc(v2, db%values)
    end if
    db%n = db%n + 1
    db%keys(db%n) = key
    db%values(db%n) = value
  end subroutine options_add

  logical function options_has_key(db, key)
    type(hashmap_type), intent(in) :: db
    character(len=*), intent(in) :: key
    integer :: i
    options_has_key = .false.
    do i = 1, db%n
      if (trim(db%keys(i)) == key) options_has_key = .true.
    end do
  end function options_has_key

  function lookup(db, key) result(v)
    type(hashmap_type), intent(in) :: db
    character(len=*), intent(in) :: key
    character(len=64) :: v
    integer :: i
    do i = 1, db%n
      if (trim(db%keys(i)) == key) then
        v = db%values(i)
        return
      end if
    end do
    print *, "option not found: ", key
    error stop 1
  end function lookup

  real(kind=DEFAULT_PRECISION) function options_get_real(db, key)
    type(hashmap_type), intent(in) :: db
    character(len=*), intent(in) :: key
    read(lookup(db, key), *) options_get_real
  end function options_get_real

  integer function options_get_integer(db, key)
    type(hashmap_type), intent(in) :: db
    character(len=*), intent(in) :: key
    read(lookup(db, key), *) options_get_integer
  end function options_get_integer

  logical function options_get_logical(db, key)
    type(hashmap_type), intent(in) :: db
    character(len=*), intent(in) :: key
    read(lookup(db, key), *) options_get_logical
  end function options_get_logical
end module optionsdatabase_mod

module grids_mod
  use datadefn_mod, only : DEFAULT_PRECISION
  implicit none
  integer, parameter, public :: Z_INDEX = 1, Y_INDEX = 2, X_INDEX = 3
  type, public :: global_grid_type
    real(kind=DEFAULT_PRECISION), dimension(3) :: bottom, top, resolution
    integer, dimension(3) :: size
  end type global_grid_type
end module grids_mod

module parcel_mod
  use datadefn_mod, only : DEFAULT_PRECISION, PARCEL_INTEGER
  implicit none
  private
  type, public :: parcel_type
    integer(kind=PARCEL_INTEGER) :: numparcels_global = 0, numparcels_local = 0
    integer(kind=PARCEL_INTEGER) :: maxparcels_global = 0, maxparcels_local = 0
    integer :: n_properties = 17
    real(kind=DEFAULT_PRECISION), allocatable, dimension(:) :: x, y, z, p, q, r, dxdt, dydt, &
         dzdt, dpdt, dqdt, drdt, h, b, vol, stretch, tag
  end type parcel_type
end module parcel_mod

module prognostics_mod
  use datadefn_mod, only : DEFAULT_PRECISION
  implicit none
  private
  type, public :: prognostic_field_type
    real(kind=DEFAULT_PRECISION), dimension(:,:,:), allocatable :: data  ! (z, y, x) with halos
  end type prognostic_field_type
end module prognostics_mod

module state_mod
  use datadefn_mod, only : DEFAULT_PRECISION
  use collections_mod, only : hashmap_type
  use grids_mod, only : global_grid_type
  use parcel_mod, only : parcel_type
  use prognostics_mod, only : prognostic_field_type
  implicit none
  private
  type, public :: model_state_type
    logical :: continue_timestep = .true., initialised = .false.
    type(hashmap_type) :: options_database
    type(global_grid_type) :: global_grid
    real(kind=DEFAULT_PRECISION) :: time = 0.0_DEFAULT_PRECISION, dtm = 0.0_DEFAULT_PRECISION, &
         dtmax = 0.0_DEFAULT_PRECISION
    integer :: rksteps = 0, iterations = 0, timestep = 1
    type(prognostic_field_type) :: u, v, w, p, q, r, b, hg, hgliq, vol
    type(parcel_type) :: parcels
  end type model_state_type
end module state_mod

module monc_component_mod
  use datadefn_mod, only : DEFAULT_PRECISION
  use state_mod, only : model_state_type
  implicit none
  private
  abstract interface
    subroutine component_callback(current_state)
      import :: model_state_type
      type(model_state_type), target, intent(inout) :: current_state
    end subroutine component_callback
  end interface
  type, public :: component_descriptor_type
    character(len=64) :: name = ""
    real(kind=DEFAULT_PRECISION) :: version = 0.1_DEFAULT_PRECISION
    procedure(component_callback), nopass, pointer :: initialisation => null(), &
         timestep => null(), finalisation => null()
  end type component_descriptor_type
end module monc_component_mod
