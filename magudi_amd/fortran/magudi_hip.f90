! magudi_hip.f90 -- Fortran 2008 host interface to the MI355X RHS library.
!
! ISO_C_BINDING declarations of include/magudi_hip.h, plus thin Fortran types that
! mirror magudi's t_Region / t_RK4Integrator / t_Solver (include/Region.f90,
! include/TimeIntegrator.f90, include/Solver.f90) so a magudi driver (bin/Forward.f90,
! bin/Adjoint.f90) hands its hot path to the GPU:
!   Region%computeRhs(mode)        -> mgd_compute_rhs
!   RK4Integrator%substepForward / substepAdjoint -> mgd_rk4_substep
!   ReverseMigrator%migrateTo      -> mgd_checkpoint_migrate
!   Solver%runForward / runAdjoint -> the reference's time loops over those stage calls, or
!                                     the whole-run entry points (mgd_run_forward,
!                                     mgd_run_forward_checkpointed / mgd_run_adjoint_checkpointed)
! Arrays keep magudi's layout x(nGridPoints, nComponents).
module magudi_hip

  use, intrinsic :: iso_c_binding
  implicit none
  private

  integer(c_int), parameter, public :: MGD_FORWARD = 1, MGD_ADJOINT = -1
  integer(c_int), parameter, public :: MGD_PATCH_FAR_FIELD = 1, MGD_PATCH_SPONGE = 2,         &
       MGD_PATCH_ACTUATOR = 3, MGD_PATCH_COST_TARGET = 4, MGD_PATCH_KOLMOGOROV = 5
  integer(c_int), parameter, public :: MGD_CONSERVED = 0, MGD_ADJOINT_VARS = 1, MGD_RHS = 2,   &
       MGD_INITIAL_STATE = 4, MGD_FINAL_STATE = 5, MGD_SCRATCH = 6
  integer(c_int), parameter, public :: MGD_FIELD_TARGET_STATE = 0,                            &
       MGD_FIELD_CONTROL_MOLLIFIER = 1, MGD_FIELD_TARGET_MOLLIFIER = 2
  integer(c_int), parameter, public :: MGD_OPT_FORCE_PLAIN = 1, MGD_OPT_TILE_SHAPE = 2,         &
       MGD_OPT_RESIDENT_STATES = 3, MGD_OPT_WINDOW_CACHE = 4, MGD_OPT_Z_CHUNK = 5,               &
       MGD_OPT_MARCH_TILE = 6, MGD_OPT_STEP_STORE = 7
  integer(c_int), parameter, public :: MGD_PATCH_IMPENETRABLE_WALL = 6,                        &
       MGD_PATCH_ISOTHERMAL_WALL = 7
  integer(c_int), parameter, public :: MGD_PERIODIC_NONE = 0, MGD_PERIODIC_PLANE = 1,            &
       MGD_PERIODIC_OVERLAP = 2
  integer(c_int), parameter, public :: MGD_FIELD_MEAN_PRESSURE = 3
  integer(c_int), parameter, public :: MGD_SOLUTION_LIMIT = 1

  type, bind(C), public :: mgd_region_desc
     integer(c_int) :: nd = 2
     integer(c_int) :: n(3) = 1
     integer(c_int) :: periodicity(3) = 0
     real(c_double) :: periodic_length(3) = 0.0_c_double
     integer(c_int) :: scheme = 0
     integer(c_int) :: viscous = 0
     integer(c_int) :: dissipation = 0
     integer(c_int) :: composite_dissipation = 0
     real(c_double) :: gamma = 1.4_c_double
     real(c_double) :: reynolds = 0.0_c_double
     real(c_double) :: prandtl = 0.72_c_double
     real(c_double) :: power_law = 0.0_c_double
     real(c_double) :: bulk_ratio = 0.0_c_double
     real(c_double) :: dissipation_amount = 0.0_c_double
     integer(c_int) :: n_instances = 1
  end type mgd_region_desc

  type, bind(C), public :: mgd_patch_desc
     integer(c_int) :: type = 0
     integer(c_int) :: normal_dir = 0
     integer(c_int) :: extent(6) = 1
     real(c_double) :: p(4) = 0.0_c_double
  end type mgd_patch_desc

  interface
     function mgd_last_error() bind(C, name="mgd_last_error") result(msg)
       import :: c_ptr
       type(c_ptr) :: msg
     end function mgd_last_error
     function mgd_region_create(desc, coords, out) bind(C, name="mgd_region_create") result(rc)
       import :: mgd_region_desc, c_double, c_ptr, c_int
       type(mgd_region_desc), intent(in) :: desc
       real(c_double), intent(in) :: coords(*)
       type(c_ptr), intent(out) :: out
       integer(c_int) :: rc
     end function mgd_region_create
     function mgd_region_destroy(r) bind(C, name="mgd_region_destroy") result(rc)
       import :: c_ptr, c_int
       type(c_ptr), value :: r
       integer(c_int) :: rc
     end function mgd_region_destroy
     function mgd_region_add_patch(r, pd) bind(C, name="mgd_region_add_patch") result(rc)
       import :: c_ptr, c_int, mgd_patch_desc
       type(c_ptr), value :: r
       type(mgd_patch_desc), intent(in) :: pd
       integer(c_int) :: rc
     end function mgd_region_add_patch
     function mgd_region_set_field(r, which, v) bind(C, name="mgd_region_set_field") result(rc)
       import :: c_ptr, c_int, c_double
       type(c_ptr), value :: r
       integer(c_int), value :: which
       real(c_double), intent(in) :: v(*)
       integer(c_int) :: rc
     end function mgd_region_set_field
     function mgd_region_add_source(r, loc, amplitude, frequency, radius, phase)               &
          bind(C, name="mgd_region_add_source") result(rc)
       import :: c_ptr, c_int, c_double
       type(c_ptr), value :: r
       real(c_double), intent(in) :: loc(3)
       real(c_double), value :: amplitude, frequency, radius, phase
       integer(c_int) :: rc
     end function mgd_region_add_source
     function mgd_region_set_option(r, option, val) bind(C, name="mgd_region_set_option")  &
          result(rc)
       import :: c_ptr, c_int
       type(c_ptr), value :: r
       integer(c_int), value :: option, val
       integer(c_int) :: rc
     end function mgd_region_set_option
     ! regions in one launch per RK stage; controls / gradients / sensitivities / Js are
     ! c_loc of arrays of c_ptr (one per member) or c_null_ptr
     function mgd_group_create(regions, n, group) bind(C, name="mgd_group_create") result(rc)
       import :: c_ptr, c_int
       type(c_ptr), intent(in) :: regions(*)
       integer(c_int), value :: n
       type(c_ptr), intent(out) :: group
       integer(c_int) :: rc
     end function mgd_group_create
     function mgd_group_destroy(group) bind(C, name="mgd_group_destroy") result(rc)
       import :: c_ptr, c_int
       type(c_ptr), value :: group
       integer(c_int) :: rc
     end function mgd_group_destroy
     function mgd_group_run_adjoint(group, t0, dt, nsteps, save_interval, controls,         &
          gradients, sensitivities, js) bind(C, name="mgd_group_run_adjoint") result(rc)
       import :: c_ptr, c_int, c_double
       type(c_ptr), value :: group
       real(c_double), value :: t0
       real(c_double), intent(in) :: dt(*)
       integer(c_int), value :: nsteps, save_interval
       type(c_ptr), value :: controls, gradients, sensitivities, js
       integer(c_int) :: rc
     end function mgd_group_run_adjoint
     function mgd_region_finalize(r) bind(C, name="mgd_region_finalize") result(rc)
       import :: c_ptr, c_int
       type(c_ptr), value :: r
       integer(c_int) :: rc
     end function mgd_region_finalize
     function mgd_state_put(r, which, inst, v) bind(C, name="mgd_state_put") result(rc)
       import :: c_ptr, c_int, c_double
       type(c_ptr), value :: r
       integer(c_int), value :: which, inst
       real(c_double), intent(in) :: v(*)
       integer(c_int) :: rc
     end function mgd_state_put
     function mgd_state_get(r, which, inst, v) bind(C, name="mgd_state_get") result(rc)
       import :: c_ptr, c_int, c_double
       type(c_ptr), value :: r
       integer(c_int), value :: which, inst
       real(c_double), intent(out) :: v(*)
       integer(c_int) :: rc
     end function mgd_state_get
     function mgd_compute_rhs(r, mode, time, control, factor) bind(C, name="mgd_compute_rhs") &
          result(rc)
       import :: c_ptr, c_int, c_double
       type(c_ptr), value :: r
       integer(c_int), value :: mode
       real(c_double), value :: time, factor
       type(c_ptr), value :: control
       integer(c_int) :: rc
     end function mgd_compute_rhs
     function mgd_run_forward(r, t0, dt, nsteps, control, J, use_graph)                        &
          bind(C, name="mgd_run_forward") result(rc)
       import :: c_ptr, c_int, c_double
       type(c_ptr), value :: r
       real(c_double), value :: t0, dt
       integer(c_int), value :: nsteps, use_graph
       type(c_ptr), value :: control
       real(c_double), intent(out) :: J(*)
       integer(c_int) :: rc
     end function mgd_run_forward
     function mgd_run_adjoint(r, t0, dt, nsteps, save_interval, control, gradient,            &
          sensitivity, J, use_graph) bind(C, name="mgd_run_adjoint") result(rc)
       import :: c_ptr, c_int, c_double
       type(c_ptr), value :: r
       real(c_double), value :: t0, dt
       integer(c_int), value :: nsteps, save_interval, use_graph
       type(c_ptr), value :: control, gradient
       real(c_double), intent(out) :: sensitivity(*), J(*)
       integer(c_int) :: rc
     end function mgd_run_adjoint
     function mgd_run_forward_checkpointed(r, t0, dt, nsteps, save_interval, control, J)      &
          bind(C, name="mgd_run_forward_checkpointed") result(rc)
       import :: c_ptr, c_int, c_double
       type(c_ptr), value :: r
       real(c_double), value :: t0, dt
       integer(c_int), value :: nsteps, save_interval
       type(c_ptr), value :: control
       real(c_double), intent(out) :: J(*)
       integer(c_int) :: rc
     end function mgd_run_forward_checkpointed
     function mgd_run_adjoint_checkpointed(r, terminal, gradient, sensitivity)                &
          bind(C, name="mgd_run_adjoint_checkpointed") result(rc)
       import :: c_ptr, c_int, c_double
       type(c_ptr), value :: r
       integer(c_int), value :: terminal
       type(c_ptr), value :: gradient
       real(c_double), intent(out) :: sensitivity(*)
       integer(c_int) :: rc
     end function mgd_run_adjoint_checkpointed
     function mgd_state_copy(r, src, dst) bind(C, name="mgd_state_copy") result(rc)
       import :: c_ptr, c_int
       type(c_ptr), value :: r
       integer(c_int), value :: src, dst
       integer(c_int) :: rc
     end function mgd_state_copy
     function mgd_state_lincomb(r, dst, dst_inst, a, a_inst, b, b_inst, s, c, c_inst)         &
          bind(C, name="mgd_state_lincomb") result(rc)
       import :: c_ptr, c_int, c_double
       type(c_ptr), value :: r
       integer(c_int), value :: dst, dst_inst, a, a_inst, b, b_inst, c, c_inst
       real(c_double), value :: s
       integer(c_int) :: rc
     end function mgd_state_lincomb
     function mgd_inner_product(r, a, b, out) bind(C, name="mgd_inner_product") result(rc)
       import :: c_ptr, c_int, c_double
       type(c_ptr), value :: r
       integer(c_int), value :: a, b
       real(c_double), intent(out) :: out(*)
       integer(c_int) :: rc
     end function mgd_inner_product
     function mgd_region_set_time_offsets(r, offsets)                                          &
          bind(C, name="mgd_region_set_time_offsets") result(rc)
       import :: c_ptr, c_int, c_double
       type(c_ptr), value :: r
       real(c_double), intent(in) :: offsets(*)
       integer(c_int) :: rc
     end function mgd_region_set_time_offsets
     ! stage-level entry points (include/TimeIntegrator.f90:64-83,
     ! include/ReverseMigrator.f90:92-106)
     function mgd_rk4_substep(r, mode, stage, time, dt, control) bind(C, name="mgd_rk4_substep") &
          result(rc)
       import :: c_ptr, c_int, c_double
       type(c_ptr), value :: r
       integer(c_int), value :: mode, stage
       real(c_double), value :: time, dt
       type(c_ptr), value :: control
       integer(c_int) :: rc
     end function mgd_rk4_substep
     function mgd_checkpoint_migrate(r, timestep, stage) bind(C, name="mgd_checkpoint_migrate") &
          result(rc)
       import :: c_ptr, c_int
       type(c_ptr), value :: r
       integer(c_int), value :: timestep, stage
       integer(c_int) :: rc
     end function mgd_checkpoint_migrate
     function mgd_control_gradient(r, gradient, sensitivity)                                  &
          bind(C, name="mgd_control_gradient") result(rc)
       import :: c_ptr, c_int
       type(c_ptr), value :: r, gradient, sensitivity
       integer(c_int) :: rc
     end function mgd_control_gradient
     function mgd_adjoint_initial_condition(r, keep)                                          &
          bind(C, name="mgd_adjoint_initial_condition") result(rc)
       import :: c_ptr, c_int
       type(c_ptr), value :: r
       integer(c_int), value :: keep
       integer(c_int) :: rc
     end function mgd_adjoint_initial_condition
     function mgd_update_adjoint_forcing(r, zero) bind(C, name="mgd_update_adjoint_forcing")  &
          result(rc)
       import :: c_ptr, c_int
       type(c_ptr), value :: r
       integer(c_int), value :: zero
       integer(c_int) :: rc
     end function mgd_update_adjoint_forcing
     function mgd_functional(r, out) bind(C, name="mgd_functional") result(rc)
       import :: c_ptr, c_int, c_double
       type(c_ptr), value :: r
       real(c_double), intent(out) :: out(*)
       integer(c_int) :: rc
     end function mgd_functional
     function mgd_region_set_solution_limits(r, enable, range)                                &
          bind(C, name="mgd_region_set_solution_limits") result(rc)
       import :: c_ptr, c_int, c_double
       type(c_ptr), value :: r
       integer(c_int), value :: enable
       real(c_double), intent(in) :: range(4)
       integer(c_int) :: rc
     end function mgd_region_set_solution_limits
     function mgd_check_solution_limits(r, crashed) bind(C, name="mgd_check_solution_limits") &
          result(rc)
       import :: c_ptr, c_int
       type(c_ptr), value :: r
       integer(c_int), intent(out) :: crashed(*)
       integer(c_int) :: rc
     end function mgd_check_solution_limits
     function mgd_time_step_size(r, cfl, dt) bind(C, name="mgd_time_step_size") result(rc)
       import :: c_ptr, c_int, c_double
       type(c_ptr), value :: r
       real(c_double), value :: cfl
       real(c_double), intent(out) :: dt(*)
       integer(c_int) :: rc
     end function mgd_time_step_size
     function mgd_sync(r) bind(C, name="mgd_sync") result(rc)
       import :: c_ptr, c_int
       type(c_ptr), value :: r
       integer(c_int) :: rc
     end function mgd_sync
     ! PLOT3D and control files (src/PLOT3DFormat.c, src/ActuatorPatchImpl.f90:226-458)
     function mgd_plot3d_detect(path, include_function, n_grids, file_type, n_scalars,         &
          has_iblank, native_endian, sizes, max_grids) bind(C, name="mgd_plot3d_detect")      &
          result(rc)
       import :: c_char, c_int
       character(kind = c_char), intent(in) :: path(*)
       integer(c_int), value :: include_function, max_grids
       integer(c_int), intent(out) :: n_grids, file_type, n_scalars, has_iblank, native_endian
       integer(c_int), intent(out) :: sizes(*)
       integer(c_int) :: rc
     end function mgd_plot3d_detect
     function mgd_plot3d_read_grid(path, block, coords, iblank)                               &
          bind(C, name="mgd_plot3d_read_grid") result(rc)
       import :: c_char, c_int, c_double
       character(kind = c_char), intent(in) :: path(*)
       integer(c_int), value :: block
       real(c_double), intent(out) :: coords(*)
       integer(c_int), intent(out) :: iblank(*)
       integer(c_int) :: rc
     end function mgd_plot3d_read_grid
     function mgd_plot3d_read_solution(path, block, nv, state, aux)                           &
          bind(C, name="mgd_plot3d_read_solution") result(rc)
       import :: c_char, c_int, c_double
       character(kind = c_char), intent(in) :: path(*)
       integer(c_int), value :: block, nv
       real(c_double), intent(out) :: state(*), aux(4)
       integer(c_int) :: rc
     end function mgd_plot3d_read_solution
     function mgd_plot3d_read_function(path, block, n_scalars, data)                          &
          bind(C, name="mgd_plot3d_read_function") result(rc)
       import :: c_char, c_int, c_double
       character(kind = c_char), intent(in) :: path(*)
       integer(c_int), value :: block, n_scalars
       real(c_double), intent(out) :: data(*)
       integer(c_int) :: rc
     end function mgd_plot3d_read_function
     function mgd_plot3d_write_solution(path, n, nv, state, aux)                              &
          bind(C, name="mgd_plot3d_write_solution") result(rc)
       import :: c_char, c_int, c_double
       character(kind = c_char), intent(in) :: path(*)
       integer(c_int), intent(in) :: n(3)
       integer(c_int), value :: nv
       real(c_double), intent(in) :: state(*), aux(4)
       integer(c_int) :: rc
     end function mgd_plot3d_write_solution
     function mgd_control_write(path, nstages, npoints, forward_order)                        &
          bind(C, name="mgd_control_write") result(rc)
       import :: c_char, c_int, c_double
       character(kind = c_char), intent(in) :: path(*)
       integer(c_int), value :: nstages, npoints
       real(c_double), intent(in) :: forward_order(*)
       integer(c_int) :: rc
     end function mgd_control_write
     function mgd_control_read(path, nstages, npoints, forward_order)                         &
          bind(C, name="mgd_control_read") result(rc)
       import :: c_char, c_int, c_double
       character(kind = c_char), intent(in) :: path(*)
       integer(c_int), value :: nstages, npoints
       real(c_double), intent(out) :: forward_order(*)
       integer(c_int) :: rc
     end function mgd_control_read
     function mgd_plot3d_write_function(path, n, n_scalars, data)                             &
          bind(C, name="mgd_plot3d_write_function") result(rc)
       import :: c_char, c_int, c_double
       character(kind = c_char), intent(in) :: path(*)
       integer(c_int), intent(in) :: n(3)
       integer(c_int), value :: n_scalars
       real(c_double), intent(in) :: data(*)
       integer(c_int) :: rc
     end function mgd_plot3d_write_function
     ! Grid%update's arrays (0 metrics, 1 jacobian, 2 norm, 3 arc lengths)
     function mgd_region_get_grid(r, which, host) bind(C, name="mgd_region_get_grid") result(rc)
       import :: c_ptr, c_int, c_double
       type(c_ptr), value :: r
       integer(c_int), value :: which
       real(c_double), intent(out) :: host(*)
       integer(c_int) :: rc
     end function mgd_region_get_grid
  end interface

  public :: mgd_run_forward_checkpointed, mgd_run_adjoint_checkpointed, mgd_state_lincomb,      &
       mgd_inner_product, mgd_region_set_time_offsets, mgd_state_copy
  public :: mgd_rk4_substep, mgd_checkpoint_migrate, mgd_control_gradient,                      &
       mgd_adjoint_initial_condition, mgd_update_adjoint_forcing, mgd_functional,               &
       mgd_region_set_solution_limits, mgd_check_solution_limits, mgd_time_step_size, mgd_sync
  public :: mgd_plot3d_detect, mgd_plot3d_read_grid, mgd_plot3d_read_solution,                  &
       mgd_plot3d_read_function, mgd_plot3d_write_solution, mgd_control_write, mgd_control_read, &
       mgd_plot3d_write_function, mgd_region_get_grid
  public :: mgd_region_set_option, mgd_group_create, mgd_group_destroy, mgd_group_run_adjoint
  public :: mgd_last_error, mgd_region_create, mgd_region_destroy, mgd_region_add_patch,      &
       mgd_region_set_field, mgd_region_add_source, mgd_region_finalize, mgd_state_put,         &
       mgd_state_get, mgd_compute_rhs, mgd_run_forward, mgd_run_adjoint

  ! t_Region over the GPU library (include/Region.f90:29-66): the grid, patches and the
  ! HBM-resident states of nInstances independent copies
  type, public :: t_Region
     type(c_ptr) :: handle = c_null_ptr
     integer :: nDimensions = 0, nUnknowns = 0, nGridPoints = 0, nInstances = 1
     integer :: nActuatorPoints = 0
     logical :: enableSolutionLimits = .false.
   contains
     procedure, pass :: setup => regionSetup
     procedure, pass :: addPatch => regionAddPatch
     procedure, pass :: finalize => regionFinalize
     procedure, pass :: computeRhs => regionComputeRhs
     procedure, pass :: setState => regionSetState
     procedure, pass :: getState => regionGetState
     procedure, pass :: getTimeStepSize => regionGetTimeStepSize
     procedure, pass :: cleanup => regionCleanup
  end type t_Region

  ! t_RK4Integrator (include/RK4Integrator.f90, src/RK4IntegratorImpl.f90:65-270): stages on
  ! the region's device state, RK buffers held by the library
  type, public :: t_RK4Integrator
     integer :: nStages = 4
     real(c_double) :: norm(4) = [1.0_c_double / 6.0_c_double, 1.0_c_double / 3.0_c_double,  &
          1.0_c_double / 3.0_c_double, 1.0_c_double / 6.0_c_double]
   contains
     procedure, pass :: substepForward => rk4SubstepForward
     procedure, pass :: substepAdjoint => rk4SubstepAdjoint
  end type t_RK4Integrator

  ! t_ReverseMigrator (include/ReverseMigrator.f90:21, 92-106) over the HBM snapshots of the
  ! last checkpointed forward run (the uniform checkpointer)
  type, public :: t_ReverseMigrator
   contains
     procedure, pass :: migrateTo => migratorMigrateTo
  end type t_ReverseMigrator

  ! t_Solver (include/Solver.f90:13-31).  runForward / runAdjoint follow the reference's time
  ! loops (src/SolverImpl.f90:672-912, 914-1245) over the stage calls; with useWholeRun the
  ! library's whole-run entry points do the same work in one call each.
  type, public :: t_Solver
     real(c_double) :: timeStepSize = 0.0_c_double
     integer :: nTimesteps = 0, saveInterval = 0
     logical :: useWholeRun = .false.
     logical :: solutionCrashed = .false.
   contains
     procedure, pass :: runForward => solverRunForward
     procedure, pass :: runAdjoint => solverRunAdjoint
  end type t_Solver

  public :: check, checkRun

contains

  subroutine check(rc, what)
    ! gracefulExit-style failure (src/ErrorHandlerImpl.f90:162-204) on a non-zero status
    integer(c_int), intent(in) :: rc
    character(len = *), intent(in) :: what
    character(kind = c_char), pointer :: msg(:)
    integer :: i
    if (rc == 0) return
    call c_f_pointer(mgd_last_error(), msg, [512])
    write(*, '(3A,I0)', advance = 'no') "magudi_hip: ", what, " failed with status ", rc
    write(*, '(A)', advance = 'no') ": "
    do i = 1, 512
       if (msg(i) == c_null_char) exit
       write(*, '(A)', advance = 'no') msg(i)
    end do
    write(*, *)
    stop 1
  end subroutine check

  function checkRun(rc, what) result(crashed)
    ! a run's status: MGD_SOLUTION_LIMIT is the reference's solution crash (cost HUGE),
    ! anything else non-zero fails as check does
    integer(c_int), intent(in) :: rc
    character(len = *), intent(in) :: what
    logical :: crashed
    crashed = rc == MGD_SOLUTION_LIMIT
    if (.not. crashed) call check(rc, what)
  end function checkRun

  subroutine regionSetup(this, desc, coordinates)
    class(t_Region) :: this
    type(mgd_region_desc), intent(in) :: desc
    real(c_double), intent(in) :: coordinates(:,:)
    this%nDimensions = desc%nd
    this%nUnknowns = desc%nd + 2
    this%nGridPoints = size(coordinates, 1)
    this%nInstances = max(1, desc%n_instances)
    call check(mgd_region_create(desc, coordinates, this%handle), "mgd_region_create")
  end subroutine regionSetup

  subroutine regionAddPatch(this, patchType, normalDirection, extent, params, nPatchPoints)
    ! one bc.dat row with its patch options; nPatchPoints counts an actuator's points
    class(t_Region) :: this
    integer, intent(in) :: patchType, normalDirection, extent(6)
    real(c_double), intent(in) :: params(:)
    integer, intent(in), optional :: nPatchPoints
    type(mgd_patch_desc) :: pd
    if (size(params) > size(pd%p)) then
       write(*, '(A,I0,A)') "magudi_hip: a patch takes at most ", size(pd%p), " parameters"
       stop 1
    end if
    pd%type = patchType
    pd%normal_dir = normalDirection
    pd%extent = extent
    pd%p(1:size(params)) = params
    call check(mgd_region_add_patch(this%handle, pd), "mgd_region_add_patch")
    if (patchType == MGD_PATCH_ACTUATOR .and. present(nPatchPoints))                          &
         this%nActuatorPoints = nPatchPoints
  end subroutine regionAddPatch

  subroutine regionFinalize(this)
    class(t_Region) :: this
    call check(mgd_region_finalize(this%handle), "mgd_region_finalize")
  end subroutine regionFinalize

  subroutine regionComputeRhs(this, mode, time, control, forcingFactor)
    ! Region%computeRhs (src/RegionImpl.f90:1877-2027) into the library's rhs buffer;
    ! control: the actuator forcing per instance and actuator point (forward);
    ! forcingFactor: State%adjointForcingFactor of the cost-target forcing (adjoint)
    class(t_Region) :: this
    integer, intent(in) :: mode
    real(c_double), intent(in) :: time
    real(c_double), intent(in), target, optional :: control(:)
    real(c_double), intent(in), optional :: forcingFactor
    type(c_ptr) :: cp
    real(c_double) :: factor
    cp = c_null_ptr
    if (present(control)) then
       if (size(control) > 0) cp = c_loc(control)
    end if
    factor = 1.0_c_double
    if (present(forcingFactor)) factor = forcingFactor
    call check(mgd_compute_rhs(this%handle, int(mode, c_int), time, cp, factor),             &
         "mgd_compute_rhs")
  end subroutine regionComputeRhs

  subroutine regionSetState(this, which, instance, x)
    ! Region%loadData for an in-memory state: x(nGridPoints, nUnknowns) of one instance
    class(t_Region) :: this
    integer, intent(in) :: which, instance
    real(c_double), intent(in) :: x(:,:)
    call check(mgd_state_put(this%handle, int(which, c_int), int(instance - 1, c_int), x),    &
         "mgd_state_put")
  end subroutine regionSetState

  subroutine regionGetState(this, which, instance, x)
    class(t_Region) :: this
    integer, intent(in) :: which, instance
    real(c_double), intent(out) :: x(:,:)
    call check(mgd_state_get(this%handle, int(which, c_int), int(instance - 1, c_int), x),    &
         "mgd_state_get")
  end subroutine regionGetState

  function regionGetTimeStepSize(this, fixedTimeStepSize, cfl) result(timeStepSize)
    ! Region%getTimeStepSize (src/RegionImpl.f90:1842-1875): the fixed time_step_size, or
    ! with cfl (use_constant_CFL_mode) the smallest computeTimeStepSize of the instances
    class(t_Region) :: this
    real(c_double), intent(in) :: fixedTimeStepSize
    real(c_double), intent(in), optional :: cfl
    real(c_double) :: timeStepSize
    real(c_double), allocatable :: dts(:)
    timeStepSize = fixedTimeStepSize
    if (.not. present(cfl)) return
    allocate(dts(this%nInstances))
    call check(mgd_time_step_size(this%handle, cfl, dts), "mgd_time_step_size")
    timeStepSize = minval(dts)
  end function regionGetTimeStepSize

  subroutine regionCleanup(this)
    class(t_Region) :: this
    integer(c_int) :: rc
    if (c_associated(this%handle)) rc = mgd_region_destroy(this%handle)
    this%handle = c_null_ptr
  end subroutine regionCleanup

  subroutine rk4SubstepForward(this, region, time, timeStepSize, timestep, stage, control)
    ! stage `stage` of the forward step at State%time `time` (t, t + dt/2, t + dt/2, t + dt);
    ! control: this stage's actuator forcing per instance and point (or absent)
    class(t_RK4Integrator) :: this
    class(t_Region) :: region
    real(c_double), intent(in) :: time, timeStepSize
    integer, intent(in) :: timestep, stage
    real(c_double), intent(in), target, optional :: control(:)
    type(c_ptr) :: cp
    cp = c_null_ptr
    if (present(control)) then
       if (size(control) > 0) cp = c_loc(control)
    end if
    call check(mgd_rk4_substep(region%handle, MGD_FORWARD, int(stage, c_int), time,          &
         timeStepSize, cp), "mgd_rk4_substep")
  end subroutine rk4SubstepForward

  subroutine rk4SubstepAdjoint(this, region, time, timeStepSize, timestep, stage)
    ! stage `stage` (4..1) of the reversed step about the forward stage state the
    ! migrator put in the conserved variables
    class(t_RK4Integrator) :: this
    class(t_Region) :: region
    real(c_double), intent(in) :: time, timeStepSize
    integer, intent(in) :: timestep, stage
    call check(mgd_rk4_substep(region%handle, MGD_ADJOINT, int(stage, c_int), time,          &
         timeStepSize, c_null_ptr), "mgd_rk4_substep")
  end subroutine rk4SubstepAdjoint

  subroutine migratorMigrateTo(this, region, timestep, stage)
    class(t_ReverseMigrator) :: this
    class(t_Region) :: region
    integer, intent(in) :: timestep, stage
    call check(mgd_checkpoint_migrate(region%handle, int(timestep, c_int), int(stage, c_int)), &
         "mgd_checkpoint_migrate")
  end subroutine migratorMigrateTo

  function solverRunForward(this, region, control) result(costFunctional)
    ! Solver%runForward (src/SolverImpl.f90:672-912) from the region's conserved state of
    ! instance 1; control(nActuatorPoints, 4 nTimesteps): the actuator forcing of every
    ! stage in forward order (controller_switch), or absent.  The cost is HUGE when a
    ! solution limit is reached.
    class(t_Solver) :: this
    class(t_Region) :: region
    real(c_double), intent(in), target, optional :: control(:,:)
    real(c_double) :: costFunctional
    type(t_RK4Integrator) :: timeIntegrator
    real(c_double) :: J(1), Jinst(1), time
    integer(c_int) :: crashed(1)
    integer :: timestep, i, sidx
    type(c_ptr) :: cp
    this%solutionCrashed = .false.
    if (this%useWholeRun) then
       cp = c_null_ptr
       if (present(control)) cp = c_loc(control)
       this%solutionCrashed = checkRun(mgd_run_forward(region%handle, 0.0_c_double,          &
            this%timeStepSize, int(this%nTimesteps, c_int), cp, J, 0_c_int), "mgd_run_forward")
       costFunctional = J(1)
       return
    end if
    time = 0.0_c_double
    costFunctional = 0.0_c_double
    do timestep = 1, this%nTimesteps
       do i = 1, timeIntegrator%nStages
          if (region%enableSolutionLimits) then
             if (checkRun(mgd_check_solution_limits(region%handle, crashed),                   &
                  "mgd_check_solution_limits")) then
                this%solutionCrashed = .true.
                costFunctional = huge(0.0_c_double)
                return
             end if
          end if
          ! State%time as substepForwardRK4 advances it (src/RK4IntegratorImpl.f90:115-156)
          if (i == 2 .or. i == 4) time = time + this%timeStepSize / 2.0_c_double
          sidx = 4 * (timestep - 1) + i
          if (present(control)) then
             call timeIntegrator%substepForward(region, time, this%timeStepSize, timestep, i,  &
                  control(:, sidx))
          else
             call timeIntegrator%substepForward(region, time, this%timeStepSize, timestep, i)
          end if
          call check(mgd_functional(region%handle, Jinst), "mgd_functional")
          costFunctional = costFunctional + timeIntegrator%norm(i) * this%timeStepSize * Jinst(1)
       end do
    end do
  end function solverRunForward

  function solverRunAdjoint(this, region, costFunctional, control, gradient) result(costSensitivity)
    ! Solver%runAdjoint (src/SolverImpl.f90:914-1245): the forward run with a snapshot every
    ! saveInterval steps in HBM, then the discrete adjoint back to t0 with the uniform
    ! checkpointer's window recompute.  gradient(nActuatorPoints, 4 nTimesteps): the control
    ! gradient of every stage in forward order (the .gradient_<patch>.dat file holds it in
    ! reverse order); costFunctional: J of the forward run.
    class(t_Solver) :: this
    class(t_Region) :: region
    real(c_double), intent(out) :: costFunctional
    real(c_double), intent(in), target, optional :: control(:,:)
    real(c_double), intent(out), target, optional :: gradient(:,:)
    real(c_double) :: costSensitivity
    type(t_RK4Integrator) :: timeIntegrator
    type(t_ReverseMigrator) :: reverseMigrator
    real(c_double) :: J(1)
    real(c_double), target :: S(1)
    integer :: timestep, i, na
    type(c_ptr) :: cp, gp
    cp = c_null_ptr
    if (present(control)) cp = c_loc(control)
    this%solutionCrashed = checkRun(mgd_run_forward_checkpointed(region%handle, 0.0_c_double, &
         this%timeStepSize, int(this%nTimesteps, c_int), int(this%saveInterval, c_int), cp, J), &
         "mgd_run_forward_checkpointed")
    costFunctional = J(1)
    costSensitivity = 0.0_c_double
    if (this%solutionCrashed) then
       costSensitivity = huge(0.0_c_double)
       return
    end if
    na = region%nActuatorPoints
    if (this%useWholeRun) then
       gp = c_null_ptr
       if (present(gradient)) gp = c_loc(gradient)
       call check(mgd_run_adjoint_checkpointed(region%handle, 0_c_int, gp, S),                 &
            "mgd_run_adjoint_checkpointed")
       costSensitivity = S(1)
       return
    end if
    ! loadInitialCondition(ADJOINT), src/SolverImpl.f90:368-428
    call check(mgd_adjoint_initial_condition(region%handle, 0_c_int), "mgd_adjoint_initial_condition")
    do timestep = this%nTimesteps - 1, 0, -1
       do i = timeIntegrator%nStages, 1, -1
          if (i == 1) then
             call reverseMigrator%migrateTo(region, timestep, timeIntegrator%nStages)
          else
             call reverseMigrator%migrateTo(region, timestep + 1, i - 1)
          end if
          ! controller%updateGradient and computeSensitivity
          gp = c_null_ptr
          if (present(gradient) .and. na > 0) gp = c_loc(gradient(1, 4 * timestep + i))
          call check(mgd_control_gradient(region%handle, gp, c_loc(S)), "mgd_control_gradient")
          costSensitivity = costSensitivity + timeIntegrator%norm(i) * this%timeStepSize * S(1)
          ! no adjoint forcing at the final stage (src/SolverImpl.f90:1177-1187)
          call check(mgd_update_adjoint_forcing(region%handle,                                 &
               merge(1_c_int, 0_c_int, timestep == 0 .and. i == 1)), "mgd_update_adjoint_forcing")
          call timeIntegrator%substepAdjoint(region, 0.0_c_double, this%timeStepSize, timestep, i)
       end do
    end do
  end function solverRunAdjoint

end module magudi_hip
