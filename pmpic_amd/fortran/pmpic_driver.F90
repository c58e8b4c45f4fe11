!> Fortran host driver: reads a parcel dump in the writebinary / readfromfile_parcelsetup
!! stream format (components/writebinary/src/writebinary.F90:165-205: time, six bounds, int64
!! count, then the 17 parcel arrays), advances it nsteps timesteps through the component
!! callbacks of pmpic_components_mod, and writes the result in the same format.
!!
!! usage: pmpic_driver in.dat out.dat nsteps nx ny nz xmax ymax zmax max_parcels dtm dtmax
!!        timestep_prefactor            (xmin = ymin = zmin = 0)
program pmpic_driver
  use, intrinsic :: iso_c_binding
  use pmpic_hip_mod
  use pmpic_components_mod
  implicit none

  type(pmpic_state_type), target :: state
  type(pmpic_component_descriptor_type) :: group(5)
  character(len=512) :: fin, fout, arg
  integer :: nsteps, step, f
  integer(c_int64_t) :: n, nread
  real(c_double) :: time, bounds(6)
  real(c_double), allocatable, target :: fields(:, :)
  type(c_ptr) :: ptrs(17)

  call get_command_argument(1, fin)
  call get_command_argument(2, fout)
  call get_command_argument(3, arg); read(arg, *) nsteps
  call get_command_argument(4, arg); read(arg, *) state%cfg%nx
  call get_command_argument(5, arg); read(arg, *) state%cfg%ny
  call get_command_argument(6, arg); read(arg, *) state%cfg%nz
  call get_command_argument(7, arg); read(arg, *) state%cfg%xmax
  call get_command_argument(8, arg); read(arg, *) state%cfg%ymax
  call get_command_argument(9, arg); read(arg, *) state%cfg%zmax
  call get_command_argument(10, arg); read(arg, *) state%cfg%max_parcels
  call get_command_argument(11, arg); read(arg, *) state%cfg%dtm
  call get_command_argument(12, arg); read(arg, *) state%cfg%dtmax
  call get_command_argument(13, arg); read(arg, *) state%cfg%timestep_prefactor
  state%cfg%xmin = 0.0_c_double
  state%cfg%ymin = 0.0_c_double
  state%cfg%zmin = 0.0_c_double
  state%cfg%precision = 0
  state%cfg%splitting_enabled = 1
  state%cfg%mixing_enabled = 1
  state%cfg%device = 0

  ! readfromfile_parcelsetup.F90:217-316 layout
  open(unit=10, file=trim(fin), access="stream", form="unformatted", status="old")
  read(10) time
  read(10) bounds
  read(10) nread
  allocate(fields(max(nread, 1_c_int64_t), 17))
  do f = 1, 17
    read(10) fields(1:nread, f)
  end do
  close(10)

  call pmpic_context_init(state)
  do f = 1, 17
    ptrs(f) = c_loc(fields(1, f))
  end do
  call pmpic_check(pmpic_set_parcels(state%ctx, nread, ptrs), "pmpic_set_parcels")
  state%time = time
  call pmpic_check(pmpic_set_time(state%ctx, time, state%cfg%dtm), "pmpic_set_time")

  ! the integrator group of global_config:43 (euler_integrator disabled)
  group(1) = vort2vel_get_descriptor()
  group(2) = vorticity_tendency_get_descriptor()
  group(3) = rk4_integrator_get_descriptor()
  group(4) = parcel_splitting_get_descriptor()
  group(5) = parcel_mixing_get_descriptor()
  do step = 1, nsteps
    call pmpic_integrator_group(state, group)
    print "(a,i6,a,i12,a,es14.6)", "step ", step, "  parcels ", state%numparcels, "  dt ", state%dtm
  end do

  call pmpic_check(pmpic_get_state(state%ctx, n, time, state%dtm), "pmpic_get_state")
  deallocate(fields)
  allocate(fields(max(n, 1_c_int64_t), 17))
  do f = 1, 17
    ptrs(f) = c_loc(fields(1, f))
  end do
  call pmpic_check(pmpic_get_parcels(state%ctx, n, ptrs), "pmpic_get_parcels")
  call pmpic_context_finalise(state)

  ! writebinary.F90:170-205 layout
  open(unit=11, file=trim(fout), access="stream", form="unformatted", status="replace")
  write(11) state%time
  write(11) bounds
  write(11) n
  do f = 1, 17
    write(11) fields(1:n, f)
  end do
  close(11)
end program pmpic_driver
