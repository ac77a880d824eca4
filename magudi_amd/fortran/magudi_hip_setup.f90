! magudi_hip_setup.f90 -- the set-up half of magudi's drivers on the GPU library: the
! Region%setup / Solver%setup of bin/Forward.f90 and bin/Adjoint.f90 from magudi.inp, bc.dat
! and the PLOT3D grid, initial condition, target state and mollifier files
! (src/RegionImpl.f90:1225-1339, src/SolverImpl.f90:465-560), with the option names and
! defaults of the reference (src/SolverOptionsImpl.f90, src/SimulationFlagsImpl.f90,
! src/GridImpl.f90:195-215, src/StateImpl.f90:130-150, the patch set-ups).
module magudi_hip_setup

  use, intrinsic :: iso_c_binding
  use magudi_hip
  use magudi_hip_input
  implicit none
  private

  ! what the drivers need beyond the region and the solver
  type, public :: t_RunSetup
     character(len = OPTION_LENGTH) :: outputPrefix = "magudi"
     character(len = OPTION_LENGTH) :: actuatorName = ""
     integer :: nActuatorPoints = 0
     integer :: gridSize(3) = 1                      ! the PLOT3D grid's size
     logical :: enableController = .false., enableFunctional = .false.
     logical :: controllerSwitch = .false.
     real(c_double), allocatable :: control(:,:)     ! (nActuatorPoints, 4 nTimesteps)
  end type t_RunSetup

  public :: setupFromInput, cstr, writeScalar

contains

  function cstr(s) result(c)
    ! a C string of a Fortran string
    character(len = *), intent(in) :: s
    character(kind = c_char, len = len_trim(s) + 1) :: c
    c = trim(s) // c_null_char
  end function cstr

  subroutine writeScalar(filename, value)
    ! the drivers' result files: one value in (1X,SP,ES23.15E3)
    character(len = *), intent(in) :: filename
    real(c_double), intent(in) :: value
    integer :: u, istat
    open(newunit = u, file = trim(filename), action = 'write', status = 'replace', iostat = istat)
    if (istat /= 0) call inputError("cannot write " // trim(filename))
    write(u, '(1X,SP,ES23.15E3)') value
    close(u)
  end subroutine writeScalar

  function patchCode(typeName) result(code)
    character(len = *), intent(in) :: typeName
    integer :: code
    select case (trim(typeName))
    case ("SAT_FAR_FIELD")
       code = MGD_PATCH_FAR_FIELD
    case ("SPONGE")
       code = MGD_PATCH_SPONGE
    case ("ACTUATOR")
       code = MGD_PATCH_ACTUATOR
    case ("COST_TARGET")
       code = MGD_PATCH_COST_TARGET
    case ("KOLMOGOROV_FORCING")
       code = MGD_PATCH_KOLMOGOROV
    case ("SAT_SLIP_WALL")
       code = MGD_PATCH_IMPENETRABLE_WALL
    case ("SAT_ISOTHERMAL_WALL")
       code = MGD_PATCH_ISOTHERMAL_WALL
    case default
       code = -1
       call inputError("patch type " // trim(typeName) // " is not built")
    end select
  end function patchCode

  function patchPoints(extent, n) result(np)
    ! number of points of a bc.dat box (negative indices count from the end)
    integer, intent(in) :: extent(6), n(3)
    integer :: np, d, lo, hi
    np = 1
    do d = 1, 3
       lo = extent(2 * d - 1)
       hi = extent(2 * d)
       if (lo < 0) lo = lo + n(d) + 1
       if (hi < 0) hi = hi + n(d) + 1
       np = np * (hi - lo + 1)
    end do
  end function patchPoints

  subroutine readFunction(filename, n, nScalars, f)
    character(len = *), intent(in) :: filename
    integer, intent(in) :: n, nScalars
    real(c_double), intent(out) :: f(n, nScalars)
    call check(mgd_plot3d_read_function(cstr(filename), 0_c_int, int(nScalars, c_int), f),    &
         "mgd_plot3d_read_function " // trim(filename))
  end subroutine readFunction

  subroutine setupFromInput(region, solver, run, nInstances, loadInitialCondition,            &
       nControlStages)
    ! nInstances: independent copies of the state (the multiple-shooting segments);
    ! loadInitialCondition = .false. leaves the conserved state to the caller;
    ! nControlStages: stages of the control forcing file (default 4 number_of_timesteps; 0:
    ! the file is not read)
    type(t_Region), intent(inout) :: region
    type(t_Solver), intent(inout) :: solver
    type(t_RunSetup), intent(inout) :: run
    integer, intent(in), optional :: nInstances
    logical, intent(in), optional :: loadInitialCondition
    integer, intent(in), optional :: nControlStages
    type(mgd_region_desc) :: desc
    type(t_PatchRow), allocatable :: rows(:)
    character(len = OPTION_LENGTH) :: filename, key, val
    integer(c_int) :: nGrids, fileType, nScalars, hasIblank, nativeEndian, sizes(3)
    integer(c_int), allocatable :: iblank(:)
    real(c_double), allocatable :: xyz(:,:), coordinates(:,:), target(:,:), f(:,:), state(:,:)
    real(c_double) :: aux(4), gamma, re, pr, params(4), loc(3), limits(4)
    integer :: nd, nGridPoints, d, i, nv, nSources, code, nStages
    logical :: doLoad

    run%outputPrefix = getOption("output_prefix", "magudi")

    ! grid (the reference reads a PLOT3D whole-format grid; one block is built here)
    call getRequiredOption("grid_file", filename)
    call check(mgd_plot3d_detect(cstr(filename), 0_c_int, nGrids, fileType, nScalars,        &
         hasIblank, nativeEndian, sizes, 1_c_int), "mgd_plot3d_detect " // trim(filename))
    if (nGrids /= 1 .or. fileType /= 0) call inputError(trim(filename) // ": one grid block expected")
    nd = 3
    if (sizes(3) == 1) nd = 2
    if (sizes(3) == 1 .and. sizes(2) == 1) nd = 1
    nGridPoints = product(sizes)
    allocate(xyz(nGridPoints, 3), iblank(nGridPoints))
    call check(mgd_plot3d_read_grid(cstr(filename), 0_c_int, xyz, iblank), "mgd_plot3d_read_grid")
    allocate(coordinates(nGridPoints, nd))
    coordinates = xyz(:, 1:nd)

    ! solver options and simulation flags
    desc%nd = nd
    desc%n = sizes
    do d = 1, nd
       write(key, '(A,I1,A)') "grid001/dir", d, "/"
       val = getOption(trim(key) // "periodicity_type", "")
       desc%periodicity(d) = MGD_PERIODIC_NONE
       if (trim(val) == "PLANE") desc%periodicity(d) = MGD_PERIODIC_PLANE
       if (trim(val) == "OVERLAP") desc%periodicity(d) = MGD_PERIODIC_OVERLAP
       if (desc%periodicity(d) == MGD_PERIODIC_PLANE)                                         &
            call getRequiredOption(trim(key) // "periodic_length", desc%periodic_length(d))
    end do
    val = getOption("defaults/discretization_scheme", "SBP 4-8")
    if (trim(val) /= "SBP 4-8") call inputError("discretization scheme " // trim(val) //       &
         " is not built (SBP 4-8 is)")
    gamma = getOption("ratio_of_specific_heats", 1.4_c_double)
    desc%gamma = gamma
    desc%viscous = merge(1, 0, getOption("include_viscous_terms", .false.))
    if (desc%viscous == 1) then
       re = max(0.0_c_double, getOption("Reynolds_number", 0.0_c_double))
       pr = max(0.0_c_double, getOption("Prandtl_number", 0.72_c_double))
       desc%reynolds = re
       desc%prandtl = pr
       if (re > 0.0_c_double .and. pr > 0.0_c_double) then
          desc%power_law = getOption("viscosity_power_law_exponent", 0.666_c_double)
          desc%bulk_ratio = getOption("bulk_viscosity_ratio", 0.6_c_double)
       end if
    end if
    desc%dissipation = merge(1, 0, getOption("add_dissipation", .false.))
    if (desc%dissipation == 1) call getRequiredOption("dissipation_amount", desc%dissipation_amount)
    desc%composite_dissipation = merge(1, 0, getOption("composite_dissipation", .true.))
    if (getOption("use_constant_CFL_mode", .true.))                                           &
         call inputError("use_constant_CFL_mode = true: the runs need a fixed time_step_size")
    if (.not. getOption("repeat_first_derivative", .true.))                                   &
         call inputError("repeat_first_derivative = false is not built")
    desc%n_instances = 1
    if (present(nInstances)) desc%n_instances = max(1, nInstances)
    run%gridSize = sizes
    call region%setup(desc, coordinates)

    ! boundary conditions (bc.dat rows of grid 1) with their patch options
    call getRequiredOption("boundary_condition_file", filename)
    call readBoundaryConditions(filename, rows)
    run%enableController = getOption("enable_controller", .false.)
    run%enableFunctional = getOption("enable_functional", .false.)
    do i = 1, size(rows)
       if (rows(i)%gridIndex /= 1) cycle
       code = patchCode(rows(i)%type)
       key = "patches/" // trim(rows(i)%name) // "/"
       params = 0.0_c_double
       select case (code)
       case (MGD_PATCH_FAR_FIELD)
          params(1) = getOption(trim(key) // "inviscid_penalty_amount",                       &
               getOption("defaults/inviscid_penalty_amount", 1.0_c_double))
          params(2) = getOption(trim(key) // "viscous_penalty_amount",                        &
               getOption("defaults/viscous_penalty_amount", 1.0_c_double))
       case (MGD_PATCH_SPONGE)
          params(1) = getOption(trim(key) // "sponge_amount",                                 &
               getOption("defaults/sponge_amount", 1.0_c_double))
          params(2) = real(getOption(trim(key) // "sponge_exponent",                          &
               getOption("defaults/sponge_exponent", 2)), c_double)
       case (MGD_PATCH_KOLMOGOROV)
          call getRequiredOption(trim(key) // "wavenumber", d)
          params(1) = real(d, c_double)
          call getRequiredOption(trim(key) // "amplitude", params(2))
       case (MGD_PATCH_IMPENETRABLE_WALL)
          params(1) = getOption(trim(key) // "inviscid_penalty_amount",                       &
               getOption("defaults/inviscid_penalty_amount", 1.0_c_double))
       case (MGD_PATCH_ISOTHERMAL_WALL)
          params(1) = getOption(trim(key) // "inviscid_penalty_amount",                       &
               getOption("defaults/inviscid_penalty_amount", 1.0_c_double))
          params(2) = getOption(trim(key) // "viscous_penalty_amount",                        &
               getOption("defaults/viscous_penalty_amount", 1.0_c_double))
          params(3) = getOption(trim(key) // "temperature", 1.0_c_double / (gamma - 1.0_c_double))
       case (MGD_PATCH_ACTUATOR)
          if (len_trim(run%actuatorName) > 0) call inputError("one ACTUATOR patch is built")
          run%actuatorName = rows(i)%name
          run%nActuatorPoints = patchPoints(rows(i)%extent, desc%n)
       end select
       call region%addPatch(code, rows(i)%normalDirection, rows(i)%extent, params,            &
            patchPoints(rows(i)%extent, desc%n))
    end do

    ! target state (target_state_file, else the quiescent state), mollifiers, mean pressure
    nv = nd + 2
    allocate(target(nGridPoints, nv), state(nGridPoints, 5), f(nGridPoints, 1))
    filename = getOption("target_state_file", "")
    if (len_trim(filename) > 0) then
       call check(mgd_plot3d_read_solution(cstr(filename), 0_c_int, int(nv, c_int), target, aux), &
            "mgd_plot3d_read_solution " // trim(filename))
    else
       ! makeQuiescent: rho = 1, u = 0, p = 1 / gamma
       target = 0.0_c_double
       target(:, 1) = 1.0_c_double
       target(:, nv) = 1.0_c_double / gamma / (gamma - 1.0_c_double)
    end if
    call check(mgd_region_set_field(region%handle, MGD_FIELD_TARGET_STATE, target), "target state")
    if (run%enableController) then
       filename = getOption("control_mollifier_file", "")
       f = 1.0_c_double
       if (len_trim(filename) > 0) call readFunction(filename, nGridPoints, 1, f)
       call check(mgd_region_set_field(region%handle, MGD_FIELD_CONTROL_MOLLIFIER, f),        &
            "control mollifier")
    end if
    if (run%enableFunctional) then
       filename = getOption("target_mollifier_file", "")
       f = 1.0_c_double
       if (len_trim(filename) > 0) call readFunction(filename, nGridPoints, 1, f)
       call check(mgd_region_set_field(region%handle, MGD_FIELD_TARGET_MOLLIFIER, f),         &
            "target mollifier")
       filename = getOption("mean_pressure_file", "")
       if (len_trim(filename) > 0) then
          call readFunction(filename, nGridPoints, 1, f)
          call check(mgd_region_set_field(region%handle, MGD_FIELD_MEAN_PRESSURE, f), "mean pressure")
       end if
    end if

    ! acoustic sources
    nSources = min(getOption("number_of_acoustic_sources", 0), 99)
    do i = 1, nSources
       write(key, '(A,I2.2,A)') "acoustic_source", i, "/"
       loc(1) = getOption(trim(key) // "x", 0.0_c_double)
       loc(2) = getOption(trim(key) // "y", 0.0_c_double)
       loc(3) = getOption(trim(key) // "z", 0.0_c_double)
       call check(mgd_region_add_source(region%handle, loc,                                    &
            getOption(trim(key) // "amplitude", 1.0_c_double),                                  &
            getOption(trim(key) // "frequency", 1.0_c_double),                                  &
            getOption(trim(key) // "radius", 1.0_c_double),                                     &
            getOption(trim(key) // "phase", 0.0_c_double)), "mgd_region_add_source")
    end do
    call region%finalize()

    ! solution limits
    region%enableSolutionLimits = getOption("enable_solution_limits", .false.)
    if (region%enableSolutionLimits) then
       call getRequiredOption("minimum_density", limits(1))
       call getRequiredOption("maximum_density", limits(2))
       call getRequiredOption("minimum_temperature", limits(3))
       call getRequiredOption("maximum_temperature", limits(4))
       call check(mgd_region_set_solution_limits(region%handle, 1_c_int, limits),             &
            "mgd_region_set_solution_limits")
    end if

    ! initial condition (initial_condition_file, else the target state)
    doLoad = .true.
    if (present(loadInitialCondition)) doLoad = loadInitialCondition
    if (doLoad) then
       filename = getOption("initial_condition_file", "")
       if (len_trim(filename) > 0) then
          call check(mgd_plot3d_read_solution(cstr(filename), 0_c_int, int(nv, c_int), state,   &
               aux), "mgd_plot3d_read_solution " // trim(filename))
          call region%setState(MGD_CONSERVED, 1, state(:, 1:nv))
       else
          call region%setState(MGD_CONSERVED, 1, target)
       end if
    end if

    ! time stepping
    call getRequiredOption("time_step_size", solver%timeStepSize)
    solver%nTimesteps = getOption("number_of_timesteps", 1000)
    solver%saveInterval = getOption("save_interval", -1)
    if (solver%saveInterval <= 0) solver%saveInterval = solver%nTimesteps

    ! actuator forcing of the forward run (controller_switch): the control forcing file
    ! (not read with nControlStages = 0)
    run%controllerSwitch = run%enableController .and. getOption("controller_switch", .false.)
    nStages = 4 * solver%nTimesteps
    if (present(nControlStages)) nStages = nControlStages
    if (run%controllerSwitch .and. run%nActuatorPoints > 0 .and. nStages > 0) then
       filename = trim(getOption("control_forcing_directory", "")) // trim(run%outputPrefix) //  &
            ".control_forcing_" // trim(run%actuatorName) // ".dat"
       filename = getOption("control_forcing_filename", trim(filename))
       allocate(run%control(run%nActuatorPoints, nStages))
       call check(mgd_control_read(cstr(filename), int(nStages, c_int),                         &
            int(run%nActuatorPoints, c_int), run%control), "mgd_control_read " // trim(filename))
    end if
  end subroutine setupFromInput

end module magudi_hip_setup
