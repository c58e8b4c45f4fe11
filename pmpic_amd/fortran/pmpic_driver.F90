!> A model-core stand-in driving the PMPIC components as the reference's model core does: the
!! configuration into the options database (setup_grid.F90:44-104 names), the parcels of a dump
!! in the writebinary / readfromfile_parcelsetup stream format (writebinary.F90:165-205: time,
!! six bounds, int64 count, then the 17 parcel arrays) into current_state%parcels, the
!! components' initialisation hooks in registration order, nsteps timesteps of the integrator
!! group (rksteps RK stages of every component's timestep hook, timestepper.F90:89-106, then
!! time = time + dtm, monc.F90:223-227), the host-mirror sync writebinary needs, the dump in the
!! same format (and writebinary's ten grids, raw, in out.dat.grids), and the finalisation hooks.
!!
!! usage: pmpic_driver in.dat out.dat nsteps nx ny nz xmax ymax zmax max_parcels dtm dtmax
!!        timestep_prefactor            (xmin = ymin = zmin = 0)
program pmpic_driver
  use, intrinsic :: iso_c_binding
  use datadefn_mod, only : DEFAULT_PRECISION, PARCEL_INTEGER
  use state_mod, only : model_state_type
  use optionsdatabase_mod, only : options_add
  use monc_component_mod, only : component_descriptor_type
  use pmpic_components_mod
  use pmpic_gpu_mod, only : pmpic_gpu_to_host, pmpic_gpu_grids_to_host
  implicit none

  type(model_state_type), target :: current_state
  type(component_descriptor_type) :: group(5)
  character(len=512) :: fin, fout, arg
  character(len=24), parameter :: keys(10) = [character(len=24) :: "nx", "ny", "nz", "xmax", &
       "ymax", "zmax", "max_parcels", "dtm", "dtmax", "timestep_prefactor"]
  integer :: nsteps, step, s, c, k
  integer(PARCEL_INTEGER) :: n, cap
  real(c_double) :: bounds(6)

  call get_command_argument(1, fin)
  call get_command_argument(2, fout)
  call get_command_argument(3, arg); read(arg, *) nsteps
  do k = 1, size(keys)
    call get_command_argument(3 + k, arg)
    call options_add(current_state%options_database, trim(keys(k)), trim(arg))
  end do
  call options_add(current_state%options_database, "xmin", "0.0")
  call options_add(current_state%options_database, "ymin", "0.0")
  call options_add(current_state%options_database, "zmin", "0.0")
  call options_add(current_state%options_database, "parcel_splitting_enabled", ".true.")
  call options_add(current_state%options_database, "parcel_mixing_enabled", ".true.")
  call get_command_argument(10, arg); read(arg, *) cap

  ! readfromfile_parcelsetup's layout, into arrays of max_parcels (parcelsetup_base.F90:63-81)
  open(unit=10, file=trim(fin), access="stream", form="unformatted", status="old")
  read(10) current_state%time
  read(10) bounds
  read(10) n
  associate (p => current_state%parcels)
    p%maxparcels_local = cap
    p%numparcels_local = n
    allocate(p%x(cap), p%y(cap), p%z(cap), p%p(cap), p%q(cap), p%r(cap), p%dxdt(cap), &
         p%dydt(cap), p%dzdt(cap), p%dpdt(cap), p%dqdt(cap), p%drdt(cap), p%h(cap), p%b(cap), &
         p%vol(cap), p%stretch(cap), p%tag(cap))
    read(10) p%x(1:n), p%y(1:n), p%z(1:n), p%p(1:n), p%q(1:n), p%r(1:n), p%dxdt(1:n), &
         p%dydt(1:n), p%dzdt(1:n), p%dpdt(1:n), p%dqdt(1:n), p%drdt(1:n), p%h(1:n), p%b(1:n), &
         p%vol(1:n), p%stretch(1:n), p%tag(1:n)
  end associate
  close(10)

  ! the integrator group of global_config:43 (euler_integrator disabled)
  group(1) = vort2vel_get_descriptor()
  group(2) = vorticity_tendency_get_descriptor()
  group(3) = rk4_integrator_get_descriptor()
  group(4) = parcel_splitting_get_descriptor()
  group(5) = parcel_mixing_get_descriptor()
  ! initialisation hooks in registration order: rk4_integrator's, then vort2vel's (the context)
  call group(3)%initialisation(current_state)
  call group(1)%initialisation(current_state)

  do step = 1, nsteps
    do s = 1, current_state%rksteps
      do c = 1, size(group)
        if (associated(group(c)%timestep)) call group(c)%timestep(current_state)
      end do
    end do
    current_state%time = current_state%time + current_state%dtm
    current_state%iterations = current_state%iterations + 1
    print "(a,i6,a,i12,a,es14.6)", "step ", step, "  parcels ", &
         current_state%parcels%numparcels_local, "  dt ", current_state%dtm
  end do

  ! writebinary's timestep hook syncs the host mirror before it reads the parcels
  call pmpic_gpu_to_host(current_state)
  n = current_state%parcels%numparcels_local
  open(unit=11, file=trim(fout), access="stream", form="unformatted", status="replace")
  write(11) current_state%time
  write(11) bounds
  write(11) n
  associate (p => current_state%parcels)
    write(11) p%x(1:n), p%y(1:n), p%z(1:n), p%p(1:n), p%q(1:n), p%r(1:n), p%dxdt(1:n), &
         p%dydt(1:n), p%dzdt(1:n), p%dpdt(1:n), p%dqdt(1:n), p%drdt(1:n), p%h(1:n), p%b(1:n), &
         p%vol(1:n), p%stretch(1:n), p%tag(1:n)
  end associate
  close(11)

  ! and the grids it dumps, written raw (u .. vol, each (nz+1, ny+4, nx+4)) beside the parcels
  call pmpic_gpu_grids_to_host(current_state)
  open(unit=12, file=trim(fout)//".grids", access="stream", form="unformatted", status="replace")
  write(12) current_state%u%data, current_state%v%data, current_state%w%data, &
       current_state%p%data, current_state%q%data, current_state%r%data, current_state%b%data, &
       current_state%hg%data, current_state%hgliq%data, current_state%vol%data
  close(12)

  do c = 1, size(group)
    if (associated(group(c)%finalisation)) call group(c)%finalisation(current_state)
  end do
end program pmpic_driver
