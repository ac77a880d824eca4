/*
 * magudi_oracle.h -- CPU restatement of magudi's right-hand-side hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker for the HIP
 * implementation in magudi_amd/: only tests/, __graft_entry__.smoke() and
 * bench.py's cpu_baseline leg may load it.  The product never links it.
 *
 * It restates, in plain C99 with the reference's operation order, the Fortran
 * routines listed beside each declaration (paths relative to the magudi
 * source tree).  Arrays use the reference layout x(nGridPoints, nComponents):
 * x[l*N + p] with p = i + nx*(j + ny*k), all 0-based here.
 */
#ifndef MAGUDI_ORACLE_H
#define MAGUDI_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

#define ORC_MAXB 20   /* max boundary width/depth of any operator (adjoint 4-8: 16x12) */

/* t_StencilOperator (include/StencilOperator.f90:7-29) */
typedef struct {
  int sym;          /* 0 SYMMETRIC, 1 SKEW_SYMMETRIC (src/StencilOperatorImpl.f90:8-11) */
  int iw;           /* interiorWidth */
  int bw;           /* boundaryWidth (columns) */
  int bd;           /* boundaryDepth (rows) */
  double interior[2 * ORC_MAXB + 1];   /* rhsInterior(-iw/2:iw/2) at [k + ORC_MAXB] */
  double b1[ORC_MAXB * ORC_MAXB];      /* rhsBoundary1(c,m) at [(m-1)*ORC_MAXB + (c-1)] */
  double b2[ORC_MAXB * ORC_MAXB];      /* rhsBoundary2 (right-boundary mirror) */
  double norm[ORC_MAXB];               /* normBoundary(1:bd') */
  int nnorm;
} orc_op;

enum { ORC_FIRST = 0, ORC_SECOND = 1, ORC_DISS = 2, ORC_DISS_T = 3, ORC_COMPOSITE = 4 };
enum { ORC_SBP48 = 0, ORC_SBP36 = 1, ORC_SBP24 = 2 };
enum { ORC_NONE = 0, ORC_PLANE = 1, ORC_OVERLAP = 2 };
enum { ORC_FORWARD = 1, ORC_ADJOINT = -1 };
enum { ORC_FAR_FIELD = 1, ORC_SPONGE = 2, ORC_ACTUATOR = 3, ORC_COST_TARGET = 4,
       ORC_KOLMOGOROV = 5, ORC_IMPENETRABLE_WALL = 6, ORC_ISOTHERMAL_WALL = 7 };

int  orc_op_setup(orc_op* op, int scheme, int kind);          /* setupOperator */
void orc_op_adjoint(const orc_op* op, orc_op* adj);           /* getAdjointOperator */
/* applyOperator (src/StencilOperatorImpl.f90:2374-2410, 35-457): in-place along dir (0-based). */
void orc_op_apply(const orc_op* op, int dir, const int n[3], int periodicity,
                  double* x, int ncomp);
void orc_op_apply_norm(const orc_op* op, int dir, const int n[3], int periodic,
                       double* x, int ncomp, int inverse);

/* Problem description shared by the oracle and the tests. */
typedef struct {
  int nd;
  int n[3];
  int periodicity[3];
  double periodic_length[3];
  int scheme;
  int viscous;
  int dissipation;
  int composite_dissipation;
  double gamma, reynolds, prandtl, power_law, bulk_ratio, dissipation_amount;
} orc_desc;

typedef struct {
  int type;
  int normal_dir;
  int extent[6];    /* as bc.dat: 1-based, negative counts from the end */
  double p[4];
} orc_patch_desc;

typedef struct orc_region orc_region;

orc_region* orc_region_create(const orc_desc* d, const double* coords);
void orc_region_destroy(orc_region* r);
int  orc_region_add_patch(orc_region* r, const orc_patch_desc* pd);
/* which: 0 target state (N x nv), 1 control mollifier (N), 2 target mollifier (N) */
int  orc_region_set_field(orc_region* r, int which, const double* v);
int  orc_region_add_source(orc_region* r, const double* loc, double amplitude,
                           double frequency, double radius, double phase);
int  orc_region_finalize(orc_region* r);   /* sponge strengths, mean pressure, far-field sigma */
int  orc_region_npoints(const orc_region* r);
/* grid arrays after Grid%update: which 0 metrics(N x nd^2), 1 jacobian(N), 2 norm(N), 3 arcLengths(N x nd) */
int  orc_region_get_grid(const orc_region* r, int which, double* out);

/* Region%computeRhs (src/RegionImpl.f90:1877-2027) for one state.
 * mode +1: rhs from q at time t.  mode -1: adjoint rhs from (q, lambda); for the
 * adjoint the cost-target forcing uses `adjoint_forcing` (patch-point array,
 * nv per point, may be NULL) scaled by forcing_factor. control_forcing: actuator
 * forcing for the energy equation per actuator patch point (may be NULL). */
int orc_compute_rhs(orc_region* r, int mode, const double* q, const double* lambda,
                    double time, const double* control_forcing,
                    const double* adjoint_forcing, double forcing_factor, double* rhs);

/* AcousticNoise J_inst (src/AcousticNoiseImpl.f90:123-206) */
double orc_functional(orc_region* r, const double* q);
/* computeTimeStepSize / computeCfl (src/CNSHelperImpl.f90:842-982) of state q */
double orc_time_step_size(orc_region* r, const double* q, double cfl);
double orc_cfl(orc_region* r, const double* q, double dt);
/* AcousticNoise adjoint forcing (src/AcousticNoiseImpl.f90:208-280), patch-point array nv x np */
int orc_adjoint_forcing(orc_region* r, const double* q, double* forcing);
int orc_cost_target_npoints(const orc_region* r);
int orc_actuator_npoints(const orc_region* r);
/* gather lambda_E * control mollifier on the actuator patch (ThermalActuatorImpl.f90:383-443) */
int orc_gradient(orc_region* r, const double* lambda, double* g);
double orc_sensitivity(orc_region* r, const double* lambda);   /* :83-159 */

/* Solver%runForward (src/SolverImpl.f90:672-912) with RK4. control: per stage
 * (nsteps*4) arrays of actuator patch forcing in forward order, or NULL.
 * Returns J.  q is in/out. */
double orc_run_forward(orc_region* r, double* q, double t0, double dt, int nsteps,
                       const double* control);
/* Solver%runAdjoint with uniform checkpointing (src/SolverImpl.f90:914-1245,
 * src/UniformCheckpointerImpl.f90:78-208).  q0: initial forward state.
 * gradient (out, may be NULL): per stage (nsteps*4) actuator arrays in forward order.
 * lambda0 (out, may be NULL): adjoint state at t0.  Returns the sensitivity. */
double orc_run_adjoint(orc_region* r, const double* q0, double t0, double dt, int nsteps,
                       int save_interval, const double* control, double* gradient,
                       double* lambda0, double* J_out);
/* The same with a given adjoint terminal state lambdaT (NULL: zero) and the final forward
 * state returned in qT_out (may be NULL). */
double orc_run_adjoint_ic(orc_region* r, const double* q0, double t0, double dt, int nsteps,
                          int save_interval, const double* control, const double* lambdaT,
                          double* gradient, double* lambda0, double* J_out, double* qT_out);

/* CNSHelper pointwise routines (src/CNSHelperImpl.f90) exposed for unit tests. */
/* OpenMP threads of the oracle's per-point and per-pencil loops (1 when built without
 * OpenMP).  Results do not depend on the count: every value is computed as in a serial
 * run and the quadrature sums stay serial. */
int orc_threads(void);
void orc_set_threads(int n);

void orc_inviscid_jacobian(int nd, const double* q, const double* metrics, double gamma,
                           double* A);     /* nv x nv column-major */
void orc_incoming_jacobian(int nd, const double* q, const double* metrics, double gamma,
                           int incoming_direction, double* A);
void orc_inviscid_flux(int nd, const double* q, const double* metrics, double gamma,
                       double* F);       /* contravariant flux along metrics, nv */

#ifdef __cplusplus
}
#endif
#endif
