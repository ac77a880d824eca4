! magudi_hip_shooting.f90 -- the multiple-shooting part of the msforward / msadjoint drivers
! (bin/msforward.f90:242-365, bin/msadjoint.f90:287-384) on the GPU library.  The segments
! are the instances of one region: each starts from its own initial condition
! <prefix>-<k>.ic.q at the time in that file's header, and one batched forward pass with
! HBM snapshots runs them all (the reference runs them one after another, each with its
! own .q snapshots); the matching mismatches and inner products run on the device.
! Without the state mollifier (time_splitting/state_mollifier/enabled, not built).
module magudi_hip_shooting

  use, intrinsic :: iso_c_binding
  use magudi_hip
  use magudi_hip_input
  use magudi_hip_setup
  implicit none
  private

  ! time_splitting options and the per-segment results of a forward pass
  type, public :: t_Shooting
     integer :: nSplit = 1, nts = 1, startTimestep = 0, saveInterval = 1
     real(c_double) :: penaltyWeight = 0.0_c_double
     character(len = OPTION_LENGTH) :: penaltyType = "quadratic"
     real(c_double) :: huberThreshold = 1.0e-12_c_double
     real(c_double), allocatable :: t0(:), segmentCost(:), segmentL2sq(:)
     real(c_double) :: L2sqSum = 0.0_c_double, timeIntegral = 0.0_c_double
     real(c_double) :: penalty = 0.0_c_double, J = 0.0_c_double
     logical :: crashed = .false.
  end type t_Shooting

  public :: readShootingOptions, penaltyNorm, penaltyGradientFactor, loadSegments, msForward,  &
       writeState, segmentFile

contains

  subroutine readShootingOptions(ms, solver)
    ! the time_splitting options of magudi.inp (bin/msforward.f90:143-166)
    type(t_Shooting), intent(inout) :: ms
    type(t_Solver), intent(in) :: solver
    call getRequiredOption("time_splitting/number_of_segments", ms%nSplit)
    call getRequiredOption("time_splitting/segment_length", ms%nts)
    call getRequiredOption("time_splitting/start_timestep", ms%startTimestep)
    call getRequiredOption("time_splitting/penalty_weight", ms%penaltyWeight)
    if (ms%nSplit < 1 .or. ms%nts < 1 .or. ms%startTimestep < 0)                              &
         call inputError("time_splitting: number_of_segments and segment_length must be "     &
         // "positive, start_timestep not negative")
    ms%penaltyType = getOption("time_splitting/penalty_type", "quadratic")
    ms%huberThreshold = getOption("time_splitting/huber/threshold", 1.0e-12_c_double)
    select case (trim(ms%penaltyType))
    case ("quadratic", "huber")
    case default
       call inputError("time_splitting/penalty_type '" // trim(ms%penaltyType) //               &
            "': expected 'quadratic' or 'huber'.")
    end select
    if (trim(ms%penaltyType) == "huber" .and. .not. ms%huberThreshold > 0.0_c_double)          &
         call inputError("time_splitting/huber/threshold must be positive")
    if (getOption("time_splitting/state_mollifier/enabled", .false.))                          &
         call inputError("time_splitting/state_mollifier/enabled = true is not built")
    ms%saveInterval = max(1, min(ms%nts, solver%saveInterval))
    allocate(ms%t0(0:ms%nSplit - 1), ms%segmentCost(0:ms%nSplit - 1),                           &
         ms%segmentL2sq(0:ms%nSplit - 1))
  end subroutine readShootingOptions

  function penaltyNorm(ms, L2sq) result(v)
    ! MSPenalty%norm (src/MSPenaltyImpl.f90): quadratic 0.5 L2sq; Huber with threshold h:
    ! sqrt(L2sq) - h/2 above h, 0.5 L2sq / h below
    type(t_Shooting), intent(in) :: ms
    real(c_double), intent(in) :: L2sq
    real(c_double) :: v, n
    if (trim(ms%penaltyType) == "quadratic") then
       v = 0.5_c_double * L2sq
    else
       n = sqrt(L2sq)
       if (n >= ms%huberThreshold) then
          v = n - 0.5_c_double * ms%huberThreshold
       else
          v = 0.5_c_double * L2sq / ms%huberThreshold
       end if
    end if
  end function penaltyNorm

  function penaltyGradientFactor(ms, L2sq) result(v)
    ! MSPenalty%gradientFactor (src/MSPenaltyImpl.f90): 1 (quadratic), max(sqrt(L2sq), h)
    type(t_Shooting), intent(in) :: ms
    real(c_double), intent(in) :: L2sq
    real(c_double) :: v
    if (trim(ms%penaltyType) == "quadratic") then
       v = 1.0_c_double
    else
       v = max(sqrt(L2sq), ms%huberThreshold)
    end if
  end function penaltyGradientFactor

  function segmentFile(prefix, k, timestep, suffix) result(filename)
    ! <prefix>-<k><suffix>, or <prefix>-<k>-<timestep, 8 digits><suffix> (timestep >= 0)
    character(len = *), intent(in) :: prefix, suffix
    integer, intent(in) :: k, timestep
    character(len = OPTION_LENGTH) :: filename
    if (timestep < 0) then
       write(filename, '(2A,I0,A)') trim(prefix), "-", k, suffix
    else
       write(filename, '(2A,I0,A,I8.8,A)') trim(prefix), "-", k, "-", timestep, suffix
    end if
  end function segmentFile

  subroutine loadSegments(region, run, ms)
    ! the per-segment initial conditions <prefix>-<k>.ic.q into the instances' initial
    ! states, their times (the files' headers) as the instances' time offsets, and the
    ! conserved states set to them
    type(t_Region), intent(inout) :: region
    type(t_RunSetup), intent(in) :: run
    type(t_Shooting), intent(inout) :: ms
    real(c_double), allocatable :: state(:,:)
    real(c_double) :: aux(4)
    character(len = OPTION_LENGTH) :: filename
    logical :: exists
    integer :: k
    allocate(state(region%nGridPoints, region%nUnknowns))
    do k = 0, ms%nSplit - 1
       filename = segmentFile(run%outputPrefix, k, -1, ".ic.q")
       inquire(file = trim(filename), exist = exists)
       if (.not. exists) call inputError("Per-segment IC file " // trim(filename) //          &
            " does not exist!")
       call check(mgd_plot3d_read_solution(cstr(filename), 0_c_int,                           &
            int(region%nUnknowns, c_int), state, aux), "mgd_plot3d_read_solution " // trim(filename))
       call region%setState(MGD_INITIAL_STATE, k + 1, state)
       ms%t0(k) = aux(4)
    end do
    call check(mgd_region_set_time_offsets(region%handle, ms%t0), "mgd_region_set_time_offsets")
    call check(mgd_state_copy(region%handle, MGD_INITIAL_STATE, MGD_CONSERVED), "mgd_state_copy")
  end subroutine loadSegments

  subroutine msForward(region, solver, run, ms)
    ! the segment loop of msforward (bin/msforward.f90:242-365): every segment's run with
    ! its slice of the control (controlTimestepOffset = k nts), the mismatches
    ! ic_k - end_(k-1) and their SBP-norm squares, J = sum of the segment costs +
    ! penalty_weight norm(sum of the squares); HUGE when a segment met the solution limits
    type(t_Region), intent(inout) :: region
    type(t_Solver), intent(in) :: solver
    type(t_RunSetup), target, intent(in) :: run
    type(t_Shooting), intent(inout) :: ms
    type(c_ptr) :: control
    real(c_double), allocatable :: ip(:)
    integer :: k
    control = c_null_ptr
    if (allocated(run%control)) control = c_loc(run%control)
    ms%crashed = checkRun(mgd_run_forward_checkpointed(region%handle, 0.0_c_double,            &
         solver%timeStepSize, int(ms%nts, c_int), int(ms%saveInterval, c_int), control,        &
         ms%segmentCost), "mgd_run_forward_checkpointed")
    do k = 1, ms%nSplit - 1
       call check(mgd_state_lincomb(region%handle, MGD_SCRATCH, int(k, c_int), MGD_INITIAL_STATE, &
            int(k, c_int), MGD_FINAL_STATE, int(k - 1, c_int), 1.0_c_double, -1_c_int, 0_c_int), &
            "mgd_state_lincomb")
    end do
    call check(mgd_state_lincomb(region%handle, MGD_SCRATCH, 0_c_int, MGD_INITIAL_STATE, 0_c_int, &
         MGD_INITIAL_STATE, 0_c_int, 0.0_c_double, -1_c_int, 0_c_int), "mgd_state_lincomb")
    allocate(ip(0:ms%nSplit - 1))
    call check(mgd_inner_product(region%handle, MGD_SCRATCH, MGD_SCRATCH, ip), "mgd_inner_product")
    ms%segmentL2sq = 0.0_c_double
    ms%L2sqSum = 0.0_c_double
    do k = 1, ms%nSplit - 1
       ms%segmentL2sq(k) = ip(k)
       ms%L2sqSum = ms%L2sqSum + ip(k)
    end do
    ms%timeIntegral = 0.0_c_double
    do k = 0, ms%nSplit - 1
       ms%timeIntegral = ms%timeIntegral + ms%segmentCost(k)
    end do
    ms%penalty = ms%penaltyWeight * penaltyNorm(ms, ms%L2sqSum)
    ms%J = ms%timeIntegral + ms%penalty
    if (ms%crashed) then
       ms%J = huge(0.0_c_double)
       ms%timeIntegral = huge(0.0_c_double)
       ms%penalty = 0.0_c_double
    end if
  end subroutine msForward

  subroutine writeState(region, run, which, instance, filename, timestep, time)
    ! one instance of a state array as a PLOT3D solution file (header: timestep, time)
    type(t_Region), intent(inout) :: region
    type(t_RunSetup), intent(in) :: run
    integer(c_int), intent(in) :: which
    integer, intent(in) :: instance, timestep
    character(len = *), intent(in) :: filename
    real(c_double), intent(in) :: time
    real(c_double), allocatable :: state(:,:)
    real(c_double) :: aux(4)
    allocate(state(region%nGridPoints, region%nUnknowns))
    call region%getState(which, instance, state)
    aux = [real(timestep, c_double), 0.0_c_double, 0.0_c_double, time]
    call check(mgd_plot3d_write_solution(cstr(filename), int(run%gridSize, c_int),            &
         int(region%nUnknowns, c_int), state, aux), "mgd_plot3d_write_solution " // trim(filename))
  end subroutine writeState

end module magudi_hip_shooting
