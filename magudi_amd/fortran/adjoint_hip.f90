! adjoint_hip.f90 -- magudi's adjoint driver (bin/Adjoint.f90) on the GPU library: reads
! magudi.inp (or --input <file>) and the files it names, runs the forward run (its cost
! functional to <output_prefix>.forward_run.txt) with snapshots in HBM, then Solver%runAdjoint
! with the uniform checkpointer's window recompute; writes the cost sensitivity to --output
! <file> (default <output_prefix>.adjoint_run.txt) in (1X,SP,ES23.15E3) and the control
! gradient of every stage, latest first, to <output_prefix>.gradient_<patch>.dat
! (gradient_filename overrides).  --whole-run uses the library's whole-run entry points.
program adjoint_hip

  use, intrinsic :: iso_c_binding
  use magudi_hip
  use magudi_hip_input
  use magudi_hip_setup
  implicit none

  type(t_Region) :: region
  type(t_Solver) :: solver
  type(t_RunSetup) :: run
  character(len = OPTION_LENGTH) :: argument, inputFilename, outputFilename, filename
  logical :: lookForInput, lookForOutput, outputFlag
  integer :: k
  real(c_double) :: costFunctional, costSensitivity
  real(c_double), allocatable :: gradient(:,:)

  inputFilename = "magudi.inp"
  lookForInput = .false.; lookForOutput = .false.; outputFlag = .false.
  do k = 1, command_argument_count()
     call get_command_argument(k, argument)
     select case (trim(adjustl(argument)))
     case ("--input")
        lookForInput = .true.
     case ("--output")
        lookForOutput = .true.
     case ("--whole-run")
        solver%useWholeRun = .true.
     case default
        if (lookForInput) then
           inputFilename = adjustl(argument)
           lookForInput = .false.
        else if (lookForOutput) then
           outputFilename = adjustl(argument)
           lookForOutput = .false.
           outputFlag = .true.
        else
           call inputError("option " // trim(argument) // " unknown!")
        end if
     end select
  end do

  call parseInputFile(inputFilename)
  call setupFromInput(region, solver, run)
  if (.not. outputFlag) outputFilename = trim(run%outputPrefix) // ".adjoint_run.txt"
  write(*, '(2A)') "Input file: ", trim(inputFilename)
  write(*, '(2A)') "Output file: ", trim(outputFilename)

  allocate(gradient(max(1, run%nActuatorPoints), 4 * solver%nTimesteps))
  if (allocated(run%control)) then
     costSensitivity = solver%runAdjoint(region, costFunctional, run%control, gradient)
  else
     costSensitivity = solver%runAdjoint(region, costFunctional, gradient = gradient)
  end if
  call writeScalar(trim(run%outputPrefix) // ".forward_run.txt", costFunctional)
  call writeScalar(outputFilename, costSensitivity)
  if (run%nActuatorPoints > 0) then
     filename = trim(getOption("gradient_directory", "")) // trim(run%outputPrefix) //          &
          ".gradient_" // trim(run%actuatorName) // ".dat"
     filename = getOption("gradient_filename", trim(filename))
     call check(mgd_control_write(cstr(filename), int(4 * solver%nTimesteps, c_int),           &
          int(run%nActuatorPoints, c_int), gradient), "mgd_control_write " // trim(filename))
  end if
  if (solver%solutionCrashed) write(*, '(A)') "Solution limits reached: cost HUGE"
  write(*, '(A,1X,SP,ES23.15E3)') "Forward run: cost functional =", costFunctional
  write(*, '(A,1X,SP,ES23.15E3)') "Adjoint run: cost sensitivity =", costSensitivity
  call region%cleanup()

end program adjoint_hip
