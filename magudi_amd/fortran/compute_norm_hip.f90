! compute_norm_hip.f90 -- magudi's compute_norm (bin/compute_norm.f90) on the GPU library:
! the diagonal of the optimiser's control-space metric for the multiple-shooting drivers.
! Reads magudi.inp (or --input <file>) with its time_splitting options and writes
!   <output_prefix>.norm_<actuator>.dat  the SBP norm H at the actuator patch points times
!        the RK4 quadrature weight and dt of every stage (collectThermalActuatorNorm,
!        src/ThermalActuatorImpl.f90:710-763), latest stage first, in the layout of the
!        control forcing file;
!   <output_prefix>.norm_ic.q  state_controllability times H in every conserved variable
!        (the metric of every segment's initial condition);
!   <output_prefix>.layout.txt (or --layout-output <file>)  the slots of the control vector:
!        "actuator <name> <size>" per ACTUATOR patch and "ic <k> <size>" for k = 1 ..
!        number_of_segments - 1, sizes in real64 values;
! and with --save_metrics the grid's <output_prefix>.Jacobian.f and <output_prefix>.metrics.f.
program compute_norm_hip

  use, intrinsic :: iso_c_binding
  use magudi_hip
  use magudi_hip_input
  use magudi_hip_setup
  implicit none

  type(t_Region) :: region
  type(t_Solver) :: solver
  type(t_RunSetup) :: run
  type(t_PatchRow), allocatable :: rows(:)
  character(len = OPTION_LENGTH) :: argument, inputFilename, layoutFilename, filename, prefix,  &
       message
  logical :: lookForInput, lookForLayout, layoutFlag, saveMetrics, exists
  integer :: k, i, j, kk, d, u, nSplit, nts, startTimestep, nd, nv, np, p, lo(3), hi(3), istat
  integer :: nStages, s, stage
  real(c_double) :: controllability, dt, aux(4), w(4)
  real(c_double), allocatable :: H(:), hp(:), normData(:,:), state(:,:), grid(:,:)
  integer(c_long_long) :: fileBytes, icSize

  inputFilename = "magudi.inp"
  lookForInput = .false.; lookForLayout = .false.; layoutFlag = .false.; saveMetrics = .false.
  do k = 1, command_argument_count()
     call get_command_argument(k, argument)
     select case (trim(adjustl(argument)))
     case ("--input")
        lookForInput = .true.
     case ("--layout-output")
        lookForLayout = .true.
     case ("--save_metrics")
        saveMetrics = .true.
     case default
        if (lookForInput) then
           inputFilename = adjustl(argument)
           inquire(file = trim(inputFilename), exist = exists)
           if (.not. exists) call inputError("input file " // trim(inputFilename) //           &
                " does not exist!")
           lookForInput = .false.
        else if (lookForLayout) then
           layoutFilename = adjustl(argument)
           lookForLayout = .false.
           layoutFlag = .true.
        else
           call inputError("option " // trim(argument) // " unknown!")
        end if
     end select
  end do

  call parseInputFile(inputFilename)
  if (.not. hasOption("enable_adjoint_solver")) call inputError(                               &
       "magudi.inp must contain 'enable_adjoint_solver' for compute_norm.")
  prefix = getOption("output_prefix", "magudi")
  if (.not. layoutFlag) layoutFilename = trim(prefix) // ".layout.txt"
  write(*, '(2A)') "Input file: ", trim(inputFilename)
  write(*, '(2A)') "Layout file: ", trim(layoutFilename)
  call getRequiredOption("time_splitting/number_of_segments", nSplit)
  call getRequiredOption("time_splitting/segment_length", nts)
  call getRequiredOption("time_splitting/start_timestep", startTimestep)
  controllability = getOption("time_splitting/state_controllability", 1.0_c_double)
  if (nSplit < 1 .or. nts < 1 .or. startTimestep < 0 .or. .not. controllability > 0.0_c_double) &
       call inputError("time_splitting: number_of_segments and segment_length must be "        &
       // "positive, start_timestep not negative, state_controllability positive")

  call setupFromInput(region, solver, run, loadInitialCondition = .false.,                    &
       nControlStages = 0)
  if (solver%nTimesteps /= nts * nSplit) then
     write(message, '(A,I0,A,I0,A,I0,A)') "magudi.inp number_of_timesteps = ",                 &
          solver%nTimesteps, " but compute_norm requires Nts * Nsplit = ", nts, " * ", nSplit,   &
          ". Set number_of_timesteps accordingly before running compute_norm."
     call inputError(message)
  end if
  nd = count(run%gridSize > 1)
  nv = region%nUnknowns
  allocate(H(region%nGridPoints))
  call check(mgd_region_get_grid(region%handle, 2_c_int, H), "mgd_region_get_grid")
  dt = solver%timeStepSize

  ! the actuator norm: H at the patch points (the patch box, i fastest) times norm(i) dt of
  ! every RK4 stage (src/RK4IntegratorImpl.f90: 1/6, 1/3, 1/3, 1/6)
  if (run%nActuatorPoints > 0) then
     call readBoundaryConditions(getOption("boundary_condition_file", ""), rows)
     do i = 1, size(rows)
        if (trim(rows(i)%name) == trim(run%actuatorName)) exit
     end do
     do d = 1, 3
        lo(d) = rows(i)%extent(2 * d - 1)
        hi(d) = rows(i)%extent(2 * d)
        if (lo(d) < 0) lo(d) = lo(d) + run%gridSize(d) + 1
        if (hi(d) < 0) hi(d) = hi(d) + run%gridSize(d) + 1
     end do
     np = run%nActuatorPoints
     allocate(hp(np))
     p = 0
     do kk = lo(3), hi(3)
        do j = lo(2), hi(2)
           do k = lo(1), hi(1)
              p = p + 1
              hp(p) = H(k + run%gridSize(1) * ((j - 1) + run%gridSize(2) * (kk - 1)))
           end do
        end do
     end do
     w = [1.0_c_double / 6.0_c_double, 1.0_c_double / 3.0_c_double,                          &
          1.0_c_double / 3.0_c_double, 1.0_c_double / 6.0_c_double]
     nStages = 4 * solver%nTimesteps
     allocate(normData(np, nStages))
     do s = 1, nStages
        stage = mod(s - 1, 4) + 1
        normData(:, s) = hp * (w(stage) * dt)
     end do
     filename = trim(prefix) // ".norm_" // trim(run%actuatorName) // ".dat"
     call check(mgd_control_write(cstr(filename), int(nStages, c_int), int(np, c_int),        &
          normData), "mgd_control_write " // trim(filename))
     write(*, '(2A)') "Wrote control space norm to ", trim(filename)
  end if

  ! the initial conditions' norm; the header holds the initial condition's step and time
  allocate(state(region%nGridPoints, nv))
  do u = 1, nv
     state(:, u) = controllability * H
  end do
  aux = 0.0_c_double
  filename = getOption("initial_condition_file", "")
  if (len_trim(filename) > 0) then
     inquire(file = trim(filename), exist = exists)
     if (exists) then
        allocate(grid(region%nGridPoints, nv))
        call check(mgd_plot3d_read_solution(cstr(filename), 0_c_int, int(nv, c_int), grid, aux), &
             "mgd_plot3d_read_solution " // trim(filename))
        deallocate(grid)
        aux(2:3) = 0.0_c_double
     end if
  end if
  filename = trim(prefix) // ".norm_ic.q"
  call check(mgd_plot3d_write_solution(cstr(filename), int(run%gridSize, c_int),              &
       int(nv, c_int), state, aux), "mgd_plot3d_write_solution " // trim(filename))
  write(*, '(2A)') "Wrote shared IC norm slab to ", trim(filename)

  if (saveMetrics) then
     filename = trim(prefix) // ".Jacobian.f"
     call check(mgd_region_get_grid(region%handle, 1_c_int, H), "mgd_region_get_grid")
     call check(mgd_plot3d_write_function(cstr(filename), int(run%gridSize, c_int), 1_c_int, H), &
          "mgd_plot3d_write_function " // trim(filename))
     allocate(grid(region%nGridPoints, nd * nd))
     call check(mgd_region_get_grid(region%handle, 0_c_int, grid), "mgd_region_get_grid")
     filename = trim(prefix) // ".metrics.f"
     call check(mgd_plot3d_write_function(cstr(filename), int(run%gridSize, c_int),           &
          int(nd * nd, c_int), grid), "mgd_plot3d_write_function " // trim(filename))
  end if

  ! the layout (kind, identifier, size in real64 values)
  open(newunit = u, file = trim(layoutFilename), action = 'write', status = 'replace',        &
       iostat = istat)
  if (istat /= 0) call inputError("cannot write " // trim(layoutFilename))
  write(u, '(A)') "# kind identifier size_in_real64"
  if (run%nActuatorPoints > 0) then
     filename = trim(prefix) // ".norm_" // trim(run%actuatorName) // ".dat"
     inquire(file = trim(filename), size = fileBytes)
     write(u, '(A,1X,A,1X,I0)') "actuator", trim(run%actuatorName), fileBytes / 8
  end if
  icSize = int(region%nGridPoints, c_long_long) * nv
  do k = 1, nSplit - 1
     write(u, '(A,1X,I0,1X,I0)') "ic", k, icSize
  end do
  close(u)
  call region%cleanup()

end program compute_norm_hip
