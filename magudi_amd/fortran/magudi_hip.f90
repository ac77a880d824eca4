! magudi_hip.f90 -- Fortran 2008 host interface to the MI355X RHS library.
!
! ISO_C_BINDING declarations of include/magudi_hip.h, plus thin Fortran types that
! mirror magudi's t_Region / t_Solver (include/Region.f90, include/Solver.f90) so a
! magudi driver (bin/Forward.f90, bin/Adjoint.f90) can hand its hot path to the GPU:
!   Region%computeRhs(mode)     -> mgd_compute_rhs
!   Solver%runForward           -> mgd_run_forward
!   Solver%runAdjoint           -> mgd_run_adjoint (HBM checkpoint + recompute)
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
       MGD_OPT_RESIDENT_STATES = 3

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
  end interface

  public :: mgd_run_forward_checkpointed, mgd_run_adjoint_checkpointed, mgd_state_lincomb,      &
       mgd_inner_product, mgd_region_set_time_offsets
  public :: mgd_region_set_option, mgd_group_create, mgd_group_destroy, mgd_group_run_adjoint
  public :: mgd_last_error, mgd_region_create, mgd_region_destroy, mgd_region_add_patch,      &
       mgd_region_set_field, mgd_region_add_source, mgd_region_finalize, mgd_state_put,         &
       mgd_state_get, mgd_compute_rhs, mgd_run_forward, mgd_run_adjoint

  ! t_Region over the GPU library (include/Region.f90:29-66)
  type, public :: t_Region
     type(c_ptr) :: handle = c_null_ptr
     integer :: nDimensions = 0, nUnknowns = 0, nGridPoints = 0
   contains
     procedure, pass :: setup => regionSetup
     procedure, pass :: addPatch => regionAddPatch
     procedure, pass :: finalize => regionFinalize
     procedure, pass :: computeRhs => regionComputeRhs
     procedure, pass :: cleanup => regionCleanup
  end type t_Region

  ! t_Solver (include/Solver.f90:13-31): time marching on the GPU
  type, public :: t_Solver
     real(c_double) :: timeStepSize = 0.0_c_double
     integer :: nTimesteps = 0, saveInterval = 0
   contains
     procedure, pass :: runForward => solverRunForward
     procedure, pass :: runAdjoint => solverRunAdjoint
  end type t_Solver

  public :: check

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

  subroutine regionSetup(this, desc, coordinates)
    class(t_Region) :: this
    type(mgd_region_desc), intent(in) :: desc
    real(c_double), intent(in) :: coordinates(:,:)
    this%nDimensions = desc%nd
    this%nUnknowns = desc%nd + 2
    this%nGridPoints = size(coordinates, 1)
    call check(mgd_region_create(desc, coordinates, this%handle), "mgd_region_create")
  end subroutine regionSetup

  subroutine regionAddPatch(this, patchType, normalDirection, extent, params)
    class(t_Region) :: this
    integer, intent(in) :: patchType, normalDirection, extent(6)
    real(c_double), intent(in) :: params(:)
    type(mgd_patch_desc) :: pd
    pd%type = patchType
    pd%normal_dir = normalDirection
    pd%extent = extent
    pd%p(1:size(params)) = params
    call check(mgd_region_add_patch(this%handle, pd), "mgd_region_add_patch")
  end subroutine regionAddPatch

  subroutine regionFinalize(this)
    class(t_Region) :: this
    call check(mgd_region_finalize(this%handle), "mgd_region_finalize")
  end subroutine regionFinalize

  subroutine regionComputeRhs(this, mode, time)
    ! Region%computeRhs (src/RegionImpl.f90:1877-2027) into the library's rhs buffer
    class(t_Region) :: this
    integer, intent(in) :: mode
    real(c_double), intent(in) :: time
    call check(mgd_compute_rhs(this%handle, int(mode, c_int), time, c_null_ptr, 1.0_c_double), &
         "mgd_compute_rhs")
  end subroutine regionComputeRhs

  subroutine regionCleanup(this)
    class(t_Region) :: this
    integer(c_int) :: rc
    if (c_associated(this%handle)) rc = mgd_region_destroy(this%handle)
    this%handle = c_null_ptr
  end subroutine regionCleanup

  function solverRunForward(this, region) result(costFunctional)
    ! Solver%runForward (src/SolverImpl.f90:672-912) from the region's conserved state
    class(t_Solver) :: this
    class(t_Region) :: region
    real(c_double) :: costFunctional
    real(c_double) :: J(1)
    call check(mgd_run_forward(region%handle, 0.0_c_double, this%timeStepSize,               &
         int(this%nTimesteps, c_int), c_null_ptr, J, 0_c_int), "mgd_run_forward")
    costFunctional = J(1)
  end function solverRunForward

  function solverRunAdjoint(this, region, costFunctional) result(costSensitivity)
    ! Solver%runAdjoint (src/SolverImpl.f90:914-1245) with HBM checkpointing
    class(t_Solver) :: this
    class(t_Region) :: region
    real(c_double), intent(out) :: costFunctional
    real(c_double) :: costSensitivity
    real(c_double) :: J(1), S(1)
    call check(mgd_run_adjoint(region%handle, 0.0_c_double, this%timeStepSize,               &
         int(this%nTimesteps, c_int), int(this%saveInterval, c_int), c_null_ptr, c_null_ptr,   &
         S, J, 0_c_int), "mgd_run_adjoint")
    costFunctional = J(1)
    costSensitivity = S(1)
  end function solverRunAdjoint

end module magudi_hip
