! msadjoint_hip.f90 -- magudi's multiple-shooting adjoint driver (bin/msadjoint.f90) on the
! GPU library: reads magudi.inp (or --input <file>) with its time_splitting options, the
! files it names and the per-segment initial conditions <output_prefix>-<k>.ic.q, runs the
! segments forward (as msforward does; the snapshots stay in HBM) and their adjoints in one
! batched pass from the matching terminal conditions
! penalty_scale (end_k - ic_(k+1)), and writes the total gradient inner product (control
! forcing and initial-condition parts) to --output <file> (default
! <output_prefix>.adjoint_run.txt), the per-segment parts to --segments-output <file>
! (default <output_prefix>.sub_adjoint_run.txt), the initial-condition gradients to
! <output_prefix>-<k>.ic.adjoint.q, the terminal conditions to
! <output_prefix>-<k>-<start_timestep + (k+1) segment_length>.adjoint.q and the control
! gradient of every stage of every segment, latest first, to
! <output_prefix>.gradient_<patch>.dat (gradient_filename overrides).
program msadjoint_hip

  use, intrinsic :: iso_c_binding
  use magudi_hip
  use magudi_hip_input
  use magudi_hip_setup
  use magudi_hip_shooting
  implicit none

  type(t_Region) :: region
  type(t_Solver) :: solver
  type(t_RunSetup) :: run
  type(t_Shooting) :: ms
  character(len = OPTION_LENGTH) :: argument, inputFilename, outputFilename, subOutputFilename, &
       filename
  logical :: lookForInput, lookForOutput, lookForSubOutput, outputFlag, subOutputFlag, crashed
  integer :: k, u, istat, last, ts
  real(c_double) :: scale, total
  real(c_double), allocatable, target :: gradient(:,:)
  real(c_double), allocatable :: ctrlIP(:), icIP(:), ip(:)
  type(c_ptr) :: gradientPtr

  inputFilename = "magudi.inp"
  lookForInput = .false.; lookForOutput = .false.; lookForSubOutput = .false.
  outputFlag = .false.; subOutputFlag = .false.
  do k = 1, command_argument_count()
     call get_command_argument(k, argument)
     select case (trim(adjustl(argument)))
     case ("--input")
        lookForInput = .true.
     case ("--output")
        lookForOutput = .true.
     case ("--segments-output")
        lookForSubOutput = .true.
     case default
        if (lookForInput) then
           inputFilename = adjustl(argument)
           lookForInput = .false.
        else if (lookForOutput) then
           outputFilename = adjustl(argument)
           lookForOutput = .false.
           outputFlag = .true.
        else if (lookForSubOutput) then
           subOutputFilename = adjustl(argument)
           lookForSubOutput = .false.
           subOutputFlag = .true.
        else
           call inputError("option " // trim(argument) // " unknown!")
        end if
     end select
  end do

  call parseInputFile(inputFilename)
  if (.not. hasOption("enable_adjoint_solver")) call inputError(                               &
       "magudi.inp must contain 'enable_adjoint_solver' to use msadjoint.")
  if (.not. hasOption("initial_condition_file")) call inputError(                              &
       "magudi.inp must contain 'initial_condition_file' (placeholder ok) for msadjoint.")
  if (.not. hasOption("adjoint_nonzero_initial_condition")) call inputError(                   &
       "magudi.inp must contain 'adjoint_nonzero_initial_condition' for msadjoint.")
  run%outputPrefix = getOption("output_prefix", "magudi")
  solver%saveInterval = getOption("save_interval", -1)
  if (solver%saveInterval <= 0) solver%saveInterval = huge(0)
  call readShootingOptions(ms, solver)
  call setupFromInput(region, solver, run, nInstances = ms%nSplit,                            &
       loadInitialCondition = .false., nControlStages = 4 * ms%nts * ms%nSplit)
  ms%saveInterval = max(1, min(ms%nts, solver%saveInterval))
  if (.not. outputFlag) outputFilename = trim(run%outputPrefix) // ".adjoint_run.txt"
  if (.not. subOutputFlag) subOutputFilename = trim(run%outputPrefix) // ".sub_adjoint_run.txt"
  write(*, '(2A)') "Input file: ", trim(inputFilename)
  write(*, '(2A)') "Output file (total inner product): ", trim(outputFilename)
  write(*, '(2A)') "Per-segment output file: ", trim(subOutputFilename)

  call loadSegments(region, run, ms)
  call msForward(region, solver, run, ms)
  if (ms%crashed) call inputError("msadjoint: the forward segments met the solution limits")

  ! matching adjoint terminal conditions (bin/msadjoint.f90:377-400): segment k < last starts
  ! back from penalty_scale (end_k - ic_(k+1)), the last from zero
  scale = ms%penaltyWeight / penaltyGradientFactor(ms, ms%L2sqSum)
  last = ms%nSplit - 1
  do k = 0, last - 1
     call check(mgd_state_lincomb(region%handle, MGD_ADJOINT_VARS, int(k, c_int),              &
          MGD_FINAL_STATE, int(k, c_int), MGD_INITIAL_STATE, int(k + 1, c_int), scale,         &
          -1_c_int, 0_c_int), "mgd_state_lincomb")
     ts = ms%startTimestep + (k + 1) * ms%nts
     call writeState(region, run, MGD_ADJOINT_VARS, k + 1,                                     &
          segmentFile(run%outputPrefix, k, ts, ".adjoint.q"), ts,                              &
          ms%t0(k) + ms%nts * solver%timeStepSize)
  end do
  call check(mgd_state_lincomb(region%handle, MGD_ADJOINT_VARS, int(last, c_int),              &
       MGD_FINAL_STATE, int(last, c_int), MGD_FINAL_STATE, int(last, c_int), 0.0_c_double,     &
       -1_c_int, 0_c_int), "mgd_state_lincomb")

  ! every segment's adjoint in one batched pass: control gradient and sensitivity per segment
  allocate(ctrlIP(0:last), icIP(0:last), ip(0:last))
  gradientPtr = c_null_ptr
  if (run%nActuatorPoints > 0) then
     allocate(gradient(run%nActuatorPoints, 4 * ms%nts * ms%nSplit))
     gradientPtr = c_loc(gradient)
  end if
  crashed = checkRun(mgd_run_adjoint_checkpointed(region%handle, 1_c_int, gradientPtr, ctrlIP), &
       "mgd_run_adjoint_checkpointed")

  ! initial-condition gradients (bin/msadjoint.f90:340-352): lambda_k + penalty_scale
  ! (ic_k - end_(k-1)) for k >= 1, the adjoint at the start of segment 0 for k = 0
  do k = 1, last
     call check(mgd_state_lincomb(region%handle, MGD_SCRATCH, int(k, c_int), MGD_INITIAL_STATE, &
          int(k, c_int), MGD_FINAL_STATE, int(k - 1, c_int), scale, MGD_ADJOINT_VARS,          &
          int(k, c_int)), "mgd_state_lincomb")
  end do
  call check(mgd_inner_product(region%handle, MGD_SCRATCH, MGD_SCRATCH, ip), "mgd_inner_product")
  icIP = 0.0_c_double
  do k = 1, last
     icIP(k) = ip(k)
     call writeState(region, run, MGD_SCRATCH, k + 1,                                          &
          segmentFile(run%outputPrefix, k, -1, ".ic.adjoint.q"),                               &
          ms%startTimestep + k * ms%nts, ms%t0(k))
  end do
  call writeState(region, run, MGD_ADJOINT_VARS, 1,                                            &
       segmentFile(run%outputPrefix, 0, -1, ".ic.adjoint.q"), ms%startTimestep, ms%t0(0))

  total = 0.0_c_double
  do k = 0, last
     total = total + ctrlIP(k) + icIP(k)
  end do
  write(*, '(A,1X,SP,ES23.15E3)') "msadjoint: total gradient inner product = ", total
  call writeScalar(outputFilename, total)
  open(newunit = u, file = trim(subOutputFilename), action = 'write', status = 'replace',     &
       iostat = istat)
  if (istat /= 0) call inputError("cannot write " // trim(subOutputFilename))
  write(u, '(A)') "# segment  control_forcing_inner_product   ic_inner_product"
  do k = 0, last
     write(u, '(I8,2(1X,SP,ES23.15E3))') k, ctrlIP(k), icIP(k)
  end do
  close(u)
  if (run%nActuatorPoints > 0) then
     filename = trim(getOption("gradient_directory", "")) // trim(run%outputPrefix) //          &
          ".gradient_" // trim(run%actuatorName) // ".dat"
     filename = getOption("gradient_filename", trim(filename))
     call check(mgd_control_write(cstr(filename), int(4 * ms%nts * ms%nSplit, c_int),          &
          int(run%nActuatorPoints, c_int), gradient), "mgd_control_write " // trim(filename))
  end if
  call region%cleanup()

end program msadjoint_hip
