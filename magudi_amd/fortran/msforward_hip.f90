! msforward_hip.f90 -- magudi's multiple-shooting forward driver (bin/msforward.f90) on the
! GPU library: reads magudi.inp (or --input <file>) with its time_splitting options, the
! files it names and the per-segment initial conditions <output_prefix>-<k>.ic.q, runs the
! segments as the instances of one region, and writes the total cost (time integral +
! matching penalty) to --output <file> (default <output_prefix>.forward_run.txt), the
! per-segment costs and mismatches to --segments-output <file> (default
! <output_prefix>.sub_J.txt) and each segment's end state to
! <output_prefix>-<k>-<start_timestep + (k+1) segment_length>.q, as the reference does.
program msforward_hip

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
  character(len = OPTION_LENGTH) :: argument, inputFilename, outputFilename, subOutputFilename
  logical :: lookForInput, lookForOutput, lookForSubOutput, outputFlag, subOutputFlag
  integer :: k, u, istat

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
  if (.not. hasOption("initial_condition_file")) call inputError(                              &
       "magudi.inp must contain 'initial_condition_file' (placeholder ok) for msforward.")
  run%outputPrefix = getOption("output_prefix", "magudi")
  solver%saveInterval = getOption("save_interval", -1)
  if (solver%saveInterval <= 0) solver%saveInterval = huge(0)
  call readShootingOptions(ms, solver)
  call setupFromInput(region, solver, run, nInstances = ms%nSplit,                            &
       loadInitialCondition = .false., nControlStages = 4 * ms%nts * ms%nSplit)
  ms%saveInterval = max(1, min(ms%nts, solver%saveInterval))
  if (.not. outputFlag) outputFilename = trim(run%outputPrefix) // ".forward_run.txt"
  if (.not. subOutputFlag) subOutputFilename = trim(run%outputPrefix) // ".sub_J.txt"
  write(*, '(2A)') "Input file: ", trim(inputFilename)
  write(*, '(2A)') "Output file (total J): ", trim(outputFilename)
  write(*, '(2A)') "Per-segment output file: ", trim(subOutputFilename)

  call loadSegments(region, run, ms)
  call msForward(region, solver, run, ms)
  ! the end-of-segment snapshots (the reference's showProgress files that msadjoint reads)
  do k = 0, ms%nSplit - 1
     call writeState(region, run, MGD_FINAL_STATE, k + 1,                                      &
          segmentFile(run%outputPrefix, k, ms%startTimestep + (k + 1) * ms%nts, ".q"),         &
          ms%startTimestep + (k + 1) * ms%nts, ms%t0(k) + ms%nts * solver%timeStepSize)
  end do

  if (ms%crashed) then
     write(*, '(A)') "msforward: hard crash detected; writing HUGE to outputFilename."
  else
     write(*, '(A,1X,SP,ES23.15E3)') "msforward: time-integral J  = ", ms%timeIntegral
     write(*, '(A,1X,SP,ES23.15E3)') "msforward: matching penalty = ", ms%penalty
     write(*, '(A,1X,SP,ES23.15E3)') "msforward: total J          = ", ms%J
  end if
  call writeScalar(outputFilename, ms%J)
  open(newunit = u, file = trim(subOutputFilename), action = 'write', status = 'replace',     &
       iostat = istat)
  if (istat /= 0) call inputError("cannot write " // trim(subOutputFilename))
  write(u, '(A)') "# segment  time_integral_J            L2sq_mismatch"
  do k = 0, ms%nSplit - 1
     write(u, '(I8,2(1X,SP,ES23.15E3))') k, ms%segmentCost(k), ms%segmentL2sq(k)
  end do
  close(u)
  call region%cleanup()

end program msforward_hip
