! forward_hip.f90 -- magudi's forward driver (bin/Forward.f90) on the GPU library: reads
! magudi.inp (or --input <file>), bc.dat and the PLOT3D grid / initial condition / target
! state / mollifier files it names, runs Solver%runForward and writes the cost functional to
! --output <file> (default <output_prefix>.forward_run.txt) in (1X,SP,ES23.15E3).  With
! controller_switch = true the actuator forcing comes from
! <output_prefix>.control_forcing_<patch>.dat.  --whole-run uses the library's whole-run
! entry point instead of the stage loop.
program forward_hip

  use, intrinsic :: iso_c_binding
  use magudi_hip
  use magudi_hip_input
  use magudi_hip_setup
  implicit none

  type(t_Region) :: region
  type(t_Solver) :: solver
  type(t_RunSetup) :: run
  character(len = OPTION_LENGTH) :: argument, inputFilename, outputFilename
  logical :: lookForInput, lookForOutput, outputFlag
  integer :: k
  real(c_double) :: costFunctional

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
  if (.not. outputFlag) outputFilename = trim(run%outputPrefix) // ".forward_run.txt"
  write(*, '(2A)') "Input file: ", trim(inputFilename)
  write(*, '(2A)') "Output file: ", trim(outputFilename)

  if (allocated(run%control)) then
     costFunctional = solver%runForward(region, run%control)
  else
     costFunctional = solver%runForward(region)
  end if
  if (solver%solutionCrashed) write(*, '(A)') "Solution limits reached: cost functional HUGE"
  if (run%enableFunctional) call writeScalar(outputFilename, costFunctional)
  write(*, '(A,1X,SP,ES23.15E3)') "Forward run: cost functional =", costFunctional
  call region%cleanup()

end program forward_hip
