/*
 * magudi_hip.h -- C ABI of the MI355X implementation of magudi's right-hand-side
 * hot path (forward RHS, discrete-adjoint RHS, RK4, HBM checkpointing).
 *
 * Plain C: pointers, sizes and an int status.  0 = success, < 0 = error
 * (mgd_last_error() gives the message), MGD_SOLUTION_LIMIT (> 0): a run or check found
 * a density or temperature outside the limits set with mgd_region_set_solution_limits.
 * Host arrays use magudi's layout x(nGridPoints, nComponents): component-major,
 * grid index i fastest.  Device buffers are owned by the region and stay
 * resident in HBM across RK4 stages; host copies happen only in put/get calls.
 *
 * Each entry point names the magudi interface it replaces (paths relative to
 * the magudi source tree).  Fortran hosts bind these with ISO_C_BINDING.
 * A region is not thread-safe; use one per host thread.
 */
#ifndef MAGUDI_HIP_H
#define MAGUDI_HIP_H

#ifdef __cplusplus
extern "C" {
#endif

#define MGD_OK 0
#define MGD_ERR_ARG (-1)
#define MGD_ERR_HIP (-2)
#define MGD_ERR_UNSUPPORTED (-3)
#define MGD_ERR_STATE (-4)
#define MGD_SOLUTION_LIMIT 1

/* Region%computeRhs modes (include/Region.f90:3-12) */
#define MGD_FORWARD 1
#define MGD_ADJOINT (-1)

/* Patch types: bc.dat "Type" column (src/PatchFactoryImpl.f90:41-82) */
#define MGD_PATCH_FAR_FIELD 1
#define MGD_PATCH_SPONGE 2
#define MGD_PATCH_ACTUATOR 3
#define MGD_PATCH_COST_TARGET 4
#define MGD_PATCH_KOLMOGOROV 5
#define MGD_PATCH_IMPENETRABLE_WALL 6   /* SAT_SLIP_WALL (src/ImpenetrableWallImpl.f90) */
#define MGD_PATCH_ISOTHERMAL_WALL 7     /* SAT_ISOTHERMAL_WALL (src/IsothermalWallImpl.f90) */

/* Grid periodicity (Grid_enum NONE / PLANE / OVERLAP) */
#define MGD_PERIODIC_NONE 0
#define MGD_PERIODIC_PLANE 1
#define MGD_PERIODIC_OVERLAP 2

/* State arrays for put/get */
#define MGD_CONSERVED 0
#define MGD_ADJOINT_VARS 1
#define MGD_RHS 2
#define MGD_TARGET 3
#define MGD_INITIAL_STATE 4     /* device-resident copy of an initial condition */
#define MGD_FINAL_STATE 5       /* forward state at the end of the last checkpointed run */
#define MGD_SCRATCH 6           /* work state for host drivers (matching conditions) */

/* Fields for mgd_region_set_field */
#define MGD_FIELD_TARGET_STATE 0
#define MGD_FIELD_CONTROL_MOLLIFIER 1
#define MGD_FIELD_TARGET_MOLLIFIER 2
#define MGD_FIELD_MEAN_PRESSURE 3       /* mean_pressure_file (N); default: from the target state */

/* Stencil operators for mgd_stencil_apply (t_Grid members, include/Grid.f90) */
#define MGD_OP_FIRST_DERIVATIVE 0
#define MGD_OP_ADJOINT_FIRST_DERIVATIVE 1
#define MGD_OP_DISSIPATION 2
#define MGD_OP_DISSIPATION_TRANSPOSE 3
#define MGD_OP_SECOND_DERIVATIVE 4      /* t_Grid%secondDerivative (src/StencilOperatorImpl.f90:1745-1840) */

/* Solver options and simulation flags a region is built from
 * (src/SolverOptionsImpl.f90:3-140, src/SimulationFlagsImpl.f90:3-70, grid options
 * of src/GridImpl.f90:142-291). */
typedef struct {
  int nd;                       /* 1, 2 or 3 */
  int n[3];                     /* global grid size */
  int periodicity[3];           /* MGD_PERIODIC_* */
  double periodic_length[3];
  int scheme;                   /* 0: "SBP 4-8" (defaults/discretization_scheme) */
  int viscous;                  /* include_viscous_terms */
  int dissipation;              /* add_dissipation */
  int composite_dissipation;    /* composite_dissipation */
  double gamma;                 /* ratio_of_specific_heats */
  double reynolds;              /* Reynolds_number (acoustic) */
  double prandtl;               /* Prandtl_number */
  double power_law;             /* viscosity_power_law_exponent */
  double bulk_ratio;            /* bulk_viscosity_ratio */
  double dissipation_amount;    /* dissipation_amount */
  int n_instances;              /* independent states sharing grid and patches (batching) */
} mgd_region_desc;

/* One bc.dat row plus its patch options.  p[]: far field (inviscid_penalty_amount,
 * viscous_penalty_amount); sponge (sponge_amount, sponge_exponent); Kolmogorov
 * (wavenumber, amplitude); slip wall (inviscid_penalty_amount); isothermal wall
 * (inviscid_penalty_amount, viscous_penalty_amount, wall temperature). */
typedef struct {
  int type;
  int normal_dir;
  int extent[6];                /* iMin iMax jMin jMax kMin kMax, 1-based, < 0 from end */
  double p[4];
} mgd_patch_desc;

typedef struct mgd_region mgd_region;

/* Last error message of the calling thread. */
const char* mgd_last_error(void);

/* Number of HIP devices visible (0 when no GPU). Calls no kernel. */
int mgd_device_count(int* count);

/* Region%setup + Grid%setupSpatialDiscretization + Grid%update
 * (src/RegionImpl.f90:1225-1339, src/GridImpl.f90:487-1065).  coords: nd x N
 * host array of a rectilinear grid.  Replaces t_Region%setup. */
int mgd_region_create(const mgd_region_desc* desc, const double* coords, mgd_region** out);
int mgd_region_destroy(mgd_region* r);

/* t_PatchFactory%connect + t_Patch%setup for one bc.dat row (src/PatchFactoryImpl.f90). */
int mgd_region_add_patch(mgd_region* r, const mgd_patch_desc* pd);

/* Target state (N x nv) and control / target mollifiers (N): Solver%setup inputs
 * (src/SolverImpl.f90:505-560). */
int mgd_region_set_field(mgd_region* r, int which, const double* host);

/* t_AcousticSource%setup (src/AcousticSourceImpl.f90:3-32) */
int mgd_region_add_source(mgd_region* r, const double* location, double amplitude,
                          double frequency, double radius, double phase);

/* Remainder of Solver%setup: sponge strengths (computeSpongeStrengths,
 * src/PatchFactoryImpl.f90:161-374), far-field penalties and A+ blocks
 * (src/FarFieldPatchImpl.f90:3-74), mean pressure (src/AcousticNoiseImpl.f90:3-71). */
int mgd_region_finalize(mgd_region* r);

/* Solution limits (enable_solution_limits with minimum_density, maximum_density,
 * minimum_temperature, maximum_temperature; src/SolverOptionsImpl.f90:72-77).  With
 * enable = 1, every forward stage of a run checks its input state first, as
 * checkSolutionLimits does (src/SolverImpl.f90:189-304, 814-819): an instance whose
 * density or temperature reaches a bound (f <= min or f >= max) or is not a number gets
 * J = HUGE (DBL_MAX), in adjoint runs also sensitivity = HUGE, and the run returns
 * MGD_SOLUTION_LIMIT.  The kernels run to the end (the flagged results are discarded
 * instead of stopping at the stage).  range = {rho_min, rho_max, T_min, T_max}.
 * Not applied to region groups.  mgd_check_solution_limits: the same check of the
 * current conserved state for a host Solver loop; crashed[i] per instance. */
int mgd_region_set_solution_limits(mgd_region* r, int enable, const double* range);
int mgd_check_solution_limits(mgd_region* r, int* crashed);

/* Region%getTimeStepSize in use_constant_CFL_mode (src/RegionImpl.f90:1842-1875, through
 * State%computeTimeStepSize, src/StateImpl.f90:605-670) and State%computeCfl
 * (src/StateImpl.f90:539-603): per instance, computeTimeStepSize / computeCfl
 * (src/CNSHelperImpl.f90:842-982) of the current conserved state -- the advective limit
 * and, in viscous regions, the diffusive one.  dt / cfl: n_instances values (the
 * reference's getTimeStepSize is their minimum). */
int mgd_time_step_size(mgd_region* r, double cfl, double* dt);
int mgd_cfl(mgd_region* r, double dt, double* cfl);

/* Per-instance time offsets (n_instances values, NULL to clear): instance i runs at
 * t + offsets[i] wherever a time enters (acoustic source phase, stage times of a run).
 * Lets independent time segments (multiple shooting, bin/msforward.f90:242-365) share
 * one batched region. */
int mgd_region_set_time_offsets(mgd_region* r, const double* offsets);

/* Copies a state array of one instance (any MGD_CONSERVED ... MGD_SCRATCH id) between
 * host and HBM.  Replaces Region%loadData / saveData for in-memory states. */
int mgd_state_put(mgd_region* r, int which, int instance, const double* host);
int mgd_state_get(mgd_region* r, int which, int instance, double* host);

/* Device-to-device copy of a state array of all instances (e.g. MGD_INITIAL_STATE to
 * MGD_CONSERVED to restart from a resident initial condition).  Asynchronous. */
int mgd_state_copy(mgd_region* r, int src, int dst);

/* dst[dst_inst] = c[c_inst] + s * (a[a_inst] - b[b_inst]) on one instance of state
 * arrays (c < 0: no c term).  The matching-condition updates of the multiple-shooting
 * drivers: mismatch ic_k - end_(k-1) (bin/msforward.f90:300-318), IC gradient
 * lambda_(k+1) + s (ic_(k+1) - end_k) and adjoint terminal s (end_k - ic_(k+1))
 * (bin/msadjoint.f90:340-400). */
int mgd_state_lincomb(mgd_region* r, int dst, int dst_inst, int a, int a_inst, int b, int b_inst,
                      double s, int c, int c_inst);

/* Grid%computeInnerProduct (src/GridImpl.f90:1067-1170) per instance: out[inst] = sum over
 * components and grid points of a * norm * b. */
int mgd_inner_product(mgd_region* r, int a, int b, double* out);

/* mgd_inner_product without any allocation, for callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops); it has no counterpart in
 * the reference.  out_device: device array of n_instances doubles.  The per-block partial
 * sums go to a fixed 2 MiB device array of the library (one per GPU), so the region may
 * hold at most 2^26 grid points over all instances (MGD_ERR_ARG beyond).  Asynchronous on
 * the region's stream (mgd_sync waits); calls that overlap on one GPU share the array
 * and race. */
int mgd_inner_product_noalloc(mgd_region* r, int a, int b, double* out_device);

/* Region%computeRhs(mode) (src/RegionImpl.f90:1877-2027) for all instances into the
 * region's rhs buffer.  time: State%time for acoustic sources.  control: actuator
 * forcing per instance per actuator patch point (host, may be NULL).  adjoint
 * forcing of the cost target is taken from the last mgd_update_adjoint_forcing,
 * scaled by forcing_factor (State%adjointForcingFactor). */
int mgd_compute_rhs(mgd_region* r, int mode, double time, const double* control,
                    double forcing_factor);

/* Functional%updateAdjointForcing (src/FunctionalImpl.f90:37-91) on the conserved state;
 * zero=1 clears it (the final adjoint step). */
int mgd_update_adjoint_forcing(mgd_region* r, int zero);

/* AcousticNoise%compute (src/AcousticNoiseImpl.f90:123-206) per instance. */
int mgd_functional(mgd_region* r, double* out);

/* Solver%runForward with RK4 (src/SolverImpl.f90:672-912, src/RK4IntegratorImpl.f90:65-162)
 * from the current conserved state.  control: per instance, per stage (nsteps*4,
 * forward order), per actuator point, or NULL.  J: per instance cost functional.
 * use_graph: replay one captured RK4 step per time step. */
int mgd_run_forward(mgd_region* r, double t0, double dt, int nsteps, const double* control,
                    double* J, int use_graph);

/* Solver%runAdjoint with its checkpoints held in HBM
 * (src/SolverImpl.f90:914-1245, src/UniformCheckpointerImpl.f90:78-208,
 * src/RK4IntegratorImpl.f90:164-270) starting from the current conserved state as
 * the initial condition: runs the forward pass (keeping its checkpoints as
 * mgd_run_forward_checkpointed does), then marches the adjoint back, recomputing the
 * stage states not held.
 * gradient: per instance, per stage (forward order), per actuator point (may be NULL);
 * sensitivity and J: per instance.  The adjoint state at t0 stays in the region
 * (MGD_ADJOINT_VARS). */
int mgd_run_adjoint(mgd_region* r, double t0, double dt, int nsteps, int save_interval,
                    const double* control, double* gradient, double* sensitivity, double* J,
                    int use_graph);

/* Cost on the final state, for runs whose adjoint starts from a nonzero terminal state:
 * J per instance of the current conserved state, and its gradient dJ/dq written to
 * MGD_ADJOINT_VARS as the terminal adjoint state for mgd_run_adjoint_checkpointed(r, 1, ..).
 * It stands in for the user-supplied <prefix>.adjoint.q of adjoint_nonzero_initial_condition
 * (src/SolverImpl.f90:368-428, 1096-1102); magudi's own functionals are time integrals.
 * MGD_TERMINAL_KINETIC_ENERGY: J = sum_p H_p |rho u|^2 / (2 rho) with the SBP norm H. */
#define MGD_TERMINAL_KINETIC_ENERGY 0
int mgd_terminal_cost(mgd_region* r, int kind, double* J);

/* The two halves of mgd_run_adjoint, for drivers that set the adjoint terminal condition
 * from the forward result (multiple shooting, bin/msadjoint.f90:287-384).
 * mgd_run_forward_checkpointed: the forward pass of Solver%runAdjoint from the current
 * conserved state, keeping its checkpoints in HBM (the .q files of showProgress,
 * src/SolverImpl.f90:672-912: the state of every step and stage states of the last
 * steps, snapshots every save_interval steps and windows, or every stage state, per
 * MGD_OPT_STEP_STORE / MGD_OPT_WINDOW_CACHE / MGD_OPT_RESIDENT_STATES) and the final
 * state (MGD_FINAL_STATE).
 * mgd_run_adjoint_checkpointed: the adjoint march back over those checkpoints.
 * terminal = 0 starts from zero, terminal = 1 from the current MGD_ADJOINT_VARS
 * (adjoint_nonzero_initial_condition, loadInitialCondition ADJOINT,
 * src/SolverImpl.f90:368-428); the terminal cost forcing of the final state is added in
 * both cases. */
int mgd_run_forward_checkpointed(mgd_region* r, double t0, double dt, int nsteps,
                                 int save_interval, const double* control, double* J);
int mgd_run_adjoint_checkpointed(mgd_region* r, int terminal, double* gradient,
                                 double* sensitivity);

/* Stage-level entry points, for a host that keeps its own Solver time loop (the reference's
 * runForward / runAdjoint, src/SolverImpl.f90:672-1245).
 *
 * mgd_rk4_substep: t_TimeIntegrator%subStepForward / subStepAdjoint
 * (include/TimeIntegrator.f90:19-21, 64-83; RK4: src/RK4IntegratorImpl.f90:65-270) on every
 * instance: one RK4 stage of the conserved state (MGD_FORWARD, stages 1..4 in order) or of
 * the adjoint state (MGD_ADJOINT, stages 4..1), the RK buffers held by the region.
 * time: State%time of the stage (acoustic source phase; forward only: t, t + dt/2,
 * t + dt/2, t + dt); control: the stage's actuator forcing per instance and actuator point
 * (forward, or NULL).  An adjoint stage linearises about the forward stage state in
 * MGD_CONSERVED (mgd_checkpoint_migrate puts it there), applies
 * State%adjointForcingFactor itself (2, 1, 1/2, 1 for stages 4..1) to the cost-target
 * forcing of the last mgd_update_adjoint_forcing.  Asynchronous on the region's stream.
 *
 * mgd_checkpoint_migrate: t_ReverseMigrator%migrateTo (include/ReverseMigrator.f90:21,
 * 92-106; UniformCheckpointer, src/UniformCheckpointerImpl.f90:78-208): MGD_CONSERVED :=
 * the forward state after stage `stage` of time step `timestep` (1-based; (0, 4) is the
 * initial condition) of the last mgd_run_forward_checkpointed, from the run's HBM storage
 * (MGD_OPT_STEP_STORE, MGD_OPT_RESIDENT_STATES): recomputing the step's inner stage
 * states from the stored step state, or the window of save_interval steps from its
 * snapshot, when the state is not held.
 *
 * mgd_control_gradient: ThermalActuator%updateGradient and computeSensitivity
 * (src/ThermalActuatorImpl.f90:83-159, 383-443) at the current adjoint state: gradient per
 * instance and actuator point (lambda_E times the control mollifier), sensitivity per
 * instance (its H-norm square); either may be NULL.
 *
 * mgd_adjoint_initial_condition: loadInitialCondition(ADJOINT) (src/SolverImpl.f90:368-428)
 * after mgd_run_forward_checkpointed: the adjoint state := 0 (keep = 0) or kept (keep = 1,
 * adjoint_nonzero_initial_condition), plus the cost-target forcing of the final state
 * with factor -dt/6. */
int mgd_rk4_substep(mgd_region* r, int mode, int stage, double time, double dt,
                    const double* control);
int mgd_checkpoint_migrate(mgd_region* r, int timestep, int stage);
int mgd_control_gradient(mgd_region* r, double* gradient, double* sensitivity);
int mgd_adjoint_initial_condition(mgd_region* r, int keep);

/* t_StencilOperator%apply (include/StencilOperator.f90:19-27,
 * src/StencilOperatorImpl.f90:2374-2410) of one of the region's operators along
 * direction dir (0-based) to ncomp host fields of size N, in place. */
int mgd_stencil_apply(mgd_region* r, int op, int dir, double* host, int ncomp);

/* Grid arrays after Grid%update as magudi stores them (for checks):
 * 0 metrics (N x nd^2), 1 jacobian (N), 2 norm (N), 3 arcLengths (N x nd). */
int mgd_region_get_grid(mgd_region* r, int which, double* host);

/* Time (ms) of the last run's kernels on the region stream, split by kernel family:
 * out[0] total, out[1] forward RHS, out[2] adjoint RHS; counts in launches[]. */
int mgd_last_timings(mgd_region* r, double* out, int* launches);

/* Region options.  MGD_OPT_FORCE_PLAIN = 1 runs the generic per-operator kernels even
 * where the fused 2-D stage kernels apply (comparison and testing).
 * MGD_OPT_TILE_SHAPE selects the tile of the fused 2-D kernels: -1 (default) autotune at
 * the first run, 0: 32 x 16, 1: 16 x 32, 2: 16 x 16, 3: 24 x 16 points with one thread
 * per point, 4: 16 x 16 and 5: 16 x 12 points with two threads per point.  No result
 * depends on the tile shape.
 * MGD_OPT_RESIDENT_STATES (fused 2-D path): 1 keeps every RK stage state of a
 * checkpointed forward run in device memory ((4 nsteps + 1) states, e.g. 19.8 GB for
 * 401^2 x 960 steps) so that the adjoint pass recomputes nothing; 0: the step store
 * (MGD_OPT_STEP_STORE) or snapshots every save_interval steps and window recompute
 * (UniformCheckpointer, src/UniformCheckpointerImpl.f90:78-208), each recompute stage
 * in one launch with an adjoint stage; -1 (default): resident where the states fit in the
 * device memory left free at the run's preparation less a reserve of max(8 GiB, 1/32 of
 * the device), except for the members of a group launch (windows).  Identical results.
 * MGD_OPT_WINDOW_CACHE (generic-kernel path, e.g. 3-D regions): the checkpointed forward
 * pass also keeps the stage states of its last n windows (4 save_interval states each) and
 * the adjoint passes read them instead of recomputing those windows; identical results.
 * -1 (default): as many windows as fit in the device memory left free at the run's
 * preparation less a reserve of max(8 GiB, 1/32 of the device), at most 8; 0: none;
 * n > 0: at most n.
 * MGD_OPT_Z_CHUNK: planes per work unit of the 3-D z marches (periodic 3-D regions): 0
 * (default) autotune at the first run -- one forward RK4 step and one adjoint rhs timed
 * for each of 32, 64 and 128 --, or one of those fixed.  MGD_OPT_MARCH_TILE: the width of
 * their tiles, 16 (16 x 16 points, 256 threads) or 32 (32 x 16, 512 threads; x sizes that
 * are multiples of 32), 0 (default) autotuned with the z chunk; the forward march and the
 * adjoint marches are tuned apart.  No result depends on either.
 * MGD_OPT_STEP_STORE (generic-kernel path; fused 2-D path where the stage states are not
 * resident, except for the members of a group launch): the checkpointed forward pass keeps the state
 * at the start of every step (nsteps states) and, in the memory left (MGD_OPT_WINDOW_CACHE
 * n >= 0: at most n save_interval steps), the three inner stage states of its last steps;
 * the adjoint pass recomputes the inner stage states of the other steps from the stored
 * state, three RK stages per step instead of 4 save_interval - 1 per window from a
 * snapshot, with no window buffer.  -1 (default): where nsteps + 3 states (fused path:
 * nsteps + 6, two recompute buffers) fit in the memory left free less the reserve and MGD_OPT_WINDOW_CACHE is at its default; 0: off
 * (snapshots and windows); 1: on where the states of every step can be allocated.
 * Identical results. */
#define MGD_OPT_FORCE_PLAIN 1
#define MGD_OPT_TILE_SHAPE 2
#define MGD_OPT_RESIDENT_STATES 3
#define MGD_OPT_WINDOW_CACHE 4
#define MGD_OPT_Z_CHUNK 5
#define MGD_OPT_MARCH_TILE 6
#define MGD_OPT_STEP_STORE 7
int mgd_region_set_option(mgd_region* r, int option, int value);

/* Storage plan of the last checkpointed forward run: plan[0] = 1 step store, plan[1] steps
 * whose inner stage states are held, from step plan[2] on; 0 snapshots and windows,
 * plan[1] windows held, from step plan[2] on (the generic kernels); 2 the fused 2-D path
 * with every stage state resident (plan[1] = 4 nsteps + 1) or with snapshots and its own
 * window recompute (plan[1] = 0).  No counterpart
 * in the reference (whose UniformCheckpointer writes .q files every save_interval
 * steps). */
int mgd_region_get_checkpoint_plan(mgd_region* r, int* plan);

/* Tile shape the fused 2-D kernels use now (-1: the region runs the generic kernels). */
int mgd_region_get_tile_shape(mgd_region* r, int* shape);
/* Work-unit shape the 3-D marches use now (MGD_OPT_MARCH_TILE, MGD_OPT_Z_CHUNK; after the
 * autotuner, if it ran): shape[0..1] tile width and z chunk of the forward march and the
 * stress pass, shape[2..3] those of the adjoint marches. */
int mgd_region_get_march_shape(mgd_region* r, int* shape);

/* Times one kernel on the region stream with HIP events after one warm-up repetition.
 * which: MGD_KERNEL_FWD_STAGE (fused 2-D forward stage kernel, `reps` RK4 steps of 4
 * stage launches), MGD_KERNEL_ADJ_STAGE (fused 2-D adjoint stage kernel, `reps` adjoint
 * RK4 steps of 4 launches), MGD_KERNEL_COPY (`reps` state-sized copies, the bandwidth
 * reference), MGD_KERNEL_RHS_FORWARD / MGD_KERNEL_RHS_ADJOINT (`reps` evaluations of the
 * rhs by the generic per-operator kernels, any region).  ms_per_launch: average per
 * launch (per rhs evaluation for the last two); bytes: algorithmic HBM bytes of one
 * launch (evaluation), averaged over the 4 stages.  Benchmark support; it overwrites the
 * region's RK, adjoint and rhs buffers (the conserved state is kept).
 * MGD_KERNEL_FWD_MARCH3D: the 3-D z-march forward kernel alone (periodic 3-D regions whose
 * rhs it serves), `reps` RK4 steps of 4 out-of-place stage launches; bytes: its stage
 * state, stress / heat-flux and RK operands, averaged over the 4 stages. */
#define MGD_KERNEL_FWD_STAGE 0
#define MGD_KERNEL_ADJ_STAGE 1
#define MGD_KERNEL_COPY 2
#define MGD_KERNEL_RHS_FORWARD 3
#define MGD_KERNEL_RHS_ADJOINT 4
#define MGD_KERNEL_FWD_MARCH3D 5
int mgd_time_kernel(mgd_region* r, int which, int reps, double* ms_per_launch, double* bytes);

/* Per-launch timing of the 3-D z-march forward kernel inside the runs: with on = 1 every
 * later launch of it by the region is bracketed by two HIP events on the region stream
 * (the record restarts); on = 0 stops recording.  No counterpart in the reference: the
 * measurement of the kernel inside a timed run (which = MGD_KERNEL_FWD_MARCH3D only). */
int mgd_kernel_timer(mgd_region* r, int which, int on);
/* Launches recorded, how many of them also wrote a copy of the stage state (the window
 * recompute and cached windows), and their summed duration in ms (waits for the last). */
int mgd_kernel_timer_read(mgd_region* r, long* launches, long* with_copy, double* ms);

/* Region groups: regions of different sizes run in lock step with one kernel launch per
 * RK stage for all of them (the uneven batches of the multiple-shooting and ensemble
 * workflows; no counterpart in the reference, which runs them as separate jobs).
 * Members: finalized 2-D regions on the fused path, each fitting 16 x 16 tiles; any
 * number of instances each.  mgd_group_run_adjoint is mgd_run_adjoint of every member
 * (same nsteps and save_interval, own dt[k] and controls[k]); gradients / sensitivities /
 * Js hold one output array per member (or null).  With MGD_OPT_RESIDENT_STATES = 1 on
 * every member each keeps all its stage states resident (no recompute; an error where
 * they cannot be allocated); otherwise (the option on some members only is an error)
 * every member keeps snapshots and the windows are recomputed on a second stream, which
 * costs little where the members are small.  Results equal the members' own runs.
 * Members keep their streams and tile shapes after the run. */
typedef struct mgd_group mgd_group;
int mgd_group_create(mgd_region* const* regions, int n, mgd_group** group);
int mgd_group_destroy(mgd_group* group);
int mgd_group_run_adjoint(mgd_group* group, double t0, const double* dt, int nsteps,
                          int save_interval, const double* const* controls,
                          double* const* gradients, double* const* sensitivities,
                          double* const* Js);

/* PLOT3D whole-format files, single or multi block, either byte order
 * (src/PLOT3DFormat.c:88-294, src/PLOT3DHelperImpl.f90:554-1286).  Host only.
 * mgd_plot3d_detect: file_type 0 grid, 1 solution, 2 function (include_function enables
 * function files); sizes receives nx, ny, nz of up to max_grids blocks.
 * Grids hold x, y, z (N each, z = 0 in 2-D) and iblank (1 when absent); solutions a
 * 4-double auxiliary record (timestep, -, -, time) and 5 slots, an nv-component state
 * using slots 1 .. nv-1 and 5; function files n_scalars components.  Arrays are
 * component-major with the grid index fastest, as everywhere in this interface. */
/* Record-marker width of the PLOT3D writers: 4 (default) or 8 bytes (the reference's
 * USE_EXTENDED_PLOT3D build, include/config.h.in:34-40).  Readers detect the width. */
int mgd_plot3d_set_record_markers(int bytes);
int mgd_plot3d_detect(const char* path, int include_function, int* n_grids, int* file_type,
                      int* n_scalars, int* has_iblank, int* native_endian, int* sizes,
                      int max_grids);
int mgd_plot3d_read_grid(const char* path, int block, double* coords, int* iblank);
int mgd_plot3d_read_solution(const char* path, int block, int nv, double* state, double* aux);
int mgd_plot3d_read_function(const char* path, int block, int n_scalars, double* data);
int mgd_plot3d_write_grid(const char* path, int nd, const int* n, const double* coords,
                          const int* iblank);
int mgd_plot3d_write_solution(const char* path, const int* n, int nv, const double* state,
                              const double* aux);
int mgd_plot3d_write_function(const char* path, const int* n, int n_scalars, const double* data);

/* Actuator control-forcing and gradient files (src/ActuatorPatchImpl.f90:226-458): raw
 * native doubles, one record of npoints values per RK stage, latest stage first.  The
 * arrays here are in forward stage order (as mgd_run_adjoint returns the gradient).
 * mgd_control_zaxpy: z = a x + y over whole files, y may be NULL (bin/ZAXPY.f90,
 * src/ControlSpaceAdvancerImpl.f90:3-191). */
int mgd_control_write(const char* path, int nstages, int npoints, const double* forward_order);
int mgd_control_read(const char* path, int nstages, int npoints, double* forward_order);
int mgd_control_zaxpy(const char* z, double a, const char* x, const char* y);

/* Waits for all work of the region. */
int mgd_sync(mgd_region* r);

#ifdef __cplusplus
}
#endif
#endif
