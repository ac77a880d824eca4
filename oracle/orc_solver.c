/*
 * orc_solver.c -- RK4 time marching and the discrete-adjoint driver of the
 * oracle (TEST INFRASTRUCTURE ONLY).  Restates:
 *   RK4Integrator substepForward/substepAdjoint  src/RK4IntegratorImpl.f90:65-270
 *   Solver%runForward                            src/SolverImpl.f90:672-912
 *   Solver%runAdjoint + loadInitialCondition     src/SolverImpl.f90:306-441, 914-1245
 *   UniformCheckpointer%migrateTo                src/UniformCheckpointerImpl.f90:78-208
 * Snapshots that the reference writes to .q files every save_interval steps are
 * kept in memory here; they hold exactly the bits the files would.
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>
#include "orc_internal.h"

static const double rk_norm[4] = {1.0 / 6.0, 1.0 / 3.0, 1.0 / 3.0, 1.0 / 6.0};

typedef struct {
  orc_region* r;
  double *buf1, *buf2, *rhs;
  long M;   /* N * nv */
} rk4;

static void rk4_init(rk4* k, orc_region* r) {
  k->r = r;
  k->M = r->N * r->nv;
  k->buf1 = calloc(k->M, sizeof(double));
  k->buf2 = calloc(k->M, sizeof(double));
  k->rhs = calloc(k->M, sizeof(double));
}

static void rk4_free(rk4* k) { free(k->buf1); free(k->buf2); free(k->rhs); }

/* substepForwardRK4; *time is the state time, advanced in stages 2 and 4. */
static void substep_forward(rk4* k, double* q, double* time, double dt, int stage,
                            const double* ctrl) {
  const long M = k->M;
  double* rhs = k->rhs;
  switch (stage) {
    case 1:
      memcpy(k->buf1, q, sizeof(double) * M);
      orc_compute_rhs(k->r, ORC_FORWARD, q, NULL, *time, ctrl, NULL, 0.0, rhs);
      #pragma omp parallel for schedule(static)
      for (long i = 0; i < M; ++i) {
        k->buf2[i] = q[i] + dt * rhs[i] / 6.0;
        q[i] = k->buf1[i] + dt * rhs[i] / 2.0;
      }
      break;
    case 2:
      *time = *time + dt / 2.0;
      orc_compute_rhs(k->r, ORC_FORWARD, q, NULL, *time, ctrl, NULL, 0.0, rhs);
      #pragma omp parallel for schedule(static)
      for (long i = 0; i < M; ++i) {
        k->buf2[i] = k->buf2[i] + dt * rhs[i] / 3.0;
        q[i] = k->buf1[i] + dt * rhs[i] / 2.0;
      }
      break;
    case 3:
      orc_compute_rhs(k->r, ORC_FORWARD, q, NULL, *time, ctrl, NULL, 0.0, rhs);
      #pragma omp parallel for schedule(static)
      for (long i = 0; i < M; ++i) {
        k->buf2[i] = k->buf2[i] + dt * rhs[i] / 3.0;
        q[i] = k->buf1[i] + dt * rhs[i];
      }
      break;
    case 4:
      *time = *time + dt / 2.0;
      orc_compute_rhs(k->r, ORC_FORWARD, q, NULL, *time, ctrl, NULL, 0.0, rhs);
      #pragma omp parallel for schedule(static)
      for (long i = 0; i < M; ++i) q[i] = k->buf2[i] + dt * rhs[i] / 6.0;
      break;
  }
}

/* substepAdjointRK4: q is the forward stage state, lam the adjoint state. */
static void substep_adjoint(rk4* k, const double* q, double* lam, double dt, int stage,
                            const double* aforcing) {
  const long M = k->M;
  double* rhs = k->rhs;
  switch (stage) {
    case 4:
      memcpy(k->buf1, lam, sizeof(double) * M);
      orc_compute_rhs(k->r, ORC_ADJOINT, q, lam, 0.0, NULL, aforcing, 2.0, rhs);
      #pragma omp parallel for schedule(static)
      for (long i = 0; i < M; ++i) {
        k->buf2[i] = lam[i] - dt * rhs[i] / 6.0;
        lam[i] = k->buf1[i] - dt * rhs[i] / 2.0;
      }
      break;
    case 3:
      orc_compute_rhs(k->r, ORC_ADJOINT, q, lam, 0.0, NULL, aforcing, 1.0, rhs);
      #pragma omp parallel for schedule(static)
      for (long i = 0; i < M; ++i) {
        k->buf2[i] = k->buf2[i] - dt * rhs[i] / 3.0;
        lam[i] = k->buf1[i] - dt * rhs[i] / 2.0;
      }
      break;
    case 2:
      orc_compute_rhs(k->r, ORC_ADJOINT, q, lam, 0.0, NULL, aforcing, 0.5, rhs);
      #pragma omp parallel for schedule(static)
      for (long i = 0; i < M; ++i) {
        k->buf2[i] = k->buf2[i] - dt * rhs[i] / 3.0;
        lam[i] = k->buf1[i] - dt * rhs[i];
      }
      break;
    case 1:
      orc_compute_rhs(k->r, ORC_ADJOINT, q, lam, 0.0, NULL, aforcing, 1.0, rhs);
      #pragma omp parallel for schedule(static)
      for (long i = 0; i < M; ++i) lam[i] = k->buf2[i] - dt * rhs[i] / 6.0;
      break;
  }
}

static const double* ctrl_at(const orc_region* r, const double* control, int step0, int stage) {
  /* forcing of forward step `step0` (0-based) and `stage` (1..4) */
  if (!control) return NULL;
  return control + ((long)step0 * 4 + (stage - 1)) * orc_actuator_npoints(r);
}

double orc_run_forward(orc_region* r, double* q, double t0, double dt, int nsteps,
                       const double* control) {
  rk4 k;
  rk4_init(&k, r);
  double J = 0.0, time = t0;
  const int has_ct = orc_cost_target_npoints(r) > 0;
  for (int s = 0; s < nsteps; ++s)
    for (int i = 1; i <= 4; ++i) {
      substep_forward(&k, q, &time, dt, i, ctrl_at(r, control, s, i));
      if (has_ct) J = J + rk_norm[i - 1] * dt * orc_functional(r, q);
    }
  rk4_free(&k);
  return J;
}

double orc_run_adjoint(orc_region* r, const double* q0, double t0, double dt, int nsteps,
                       int save_interval, const double* control, double* gradient,
                       double* lambda0, double* J_out) {
  return orc_run_adjoint_ic(r, q0, t0, dt, nsteps, save_interval, control, NULL, gradient,
                            lambda0, J_out, NULL);
}

/* Solver%runAdjoint with adjoint_nonzero_initial_condition (src/SolverImpl.f90:1096-1102,
 * loadInitialCondition ADJOINT :368-428): lambda(T) = lambdaT (NULL: zero) plus the
 * terminal cost forcing; qT_out (may be NULL) receives the final forward state. */
double orc_run_adjoint_ic(orc_region* r, const double* q0, double t0, double dt, int nsteps,
                          int save_interval, const double* control, const double* lambdaT,
                          double* gradient, double* lambda0, double* J_out, double* qT_out) {
  const long M = r->N * r->nv;
  const int K = save_interval;
  const int nsnap = nsteps / K + 2;
  /* forward run writing snapshots at t0 and every save_interval steps (showProgress) */
  double* snaps = malloc(sizeof(double) * M * nsnap);
  double* snap_time = calloc(nsnap, sizeof(double));
  double* q = malloc(sizeof(double) * M);
  memcpy(q, q0, sizeof(double) * M);
  memcpy(snaps, q, sizeof(double) * M);
  snap_time[0] = t0;
  rk4 k;
  rk4_init(&k, r);
  double J = 0.0, time = t0;
  const int has_ct = orc_cost_target_npoints(r) > 0;
  for (int s = 0; s < nsteps; ++s) {
    for (int i = 1; i <= 4; ++i) {
      substep_forward(&k, q, &time, dt, i, ctrl_at(r, control, s, i));
      if (has_ct) J = J + rk_norm[i - 1] * dt * orc_functional(r, q);
    }
    if ((s + 1) % K == 0 || s + 1 == nsteps) {
      const int slot = (s + 1) % K == 0 ? (s + 1) / K : nsnap - 1;
      memcpy(snaps + (long)slot * M, q, sizeof(double) * M);
      snap_time[slot] = time;
    }
  }
  if (J_out) *J_out = J;
  if (qT_out) memcpy(qT_out, q, sizeof(double) * M);
  /* adjoint terminal condition (loadInitialCondition ADJOINT) on the final state */
  const int nct = orc_cost_target_npoints(r);
  const int nv = r->nv;
  double* aforce = calloc((size_t)(nct > 0 ? nct : 1) * nv, sizeof(double));
  double* lam = calloc(M, sizeof(double));
  if (lambdaT) memcpy(lam, lambdaT, sizeof(double) * M);
  if (nct > 0) {
    orc_adjoint_forcing(r, q, aforce);
    double* rhs = calloc(M, sizeof(double));
    /* rhs = 0; CostTarget%updateRhs(ADJOINT) with factor -dt/6; lambda += rhs */
    orc_patch* P = NULL;
    for (int i = 0; i < r->npatches; ++i) if (r->patches[i].pd.type == ORC_COST_TARGET) P = &r->patches[i];
    const double factor = -dt / 6.0;
    for (int l = 0; l < nv; ++l)
      for (int c = 0; c < P->np; ++c) {
        const int a = c % P->ls[0], b = (c / P->ls[0]) % P->ls[1], e = c / (P->ls[0] * P->ls[1]);
        const long p = (P->lo[0] + a) + (long)r->d.n[0] * ((P->lo[1] + b) + (long)r->d.n[1] * (P->lo[2] + e));
        rhs[l * r->N + p] = rhs[l * r->N + p] + factor * aforce[(long)l * P->np + c];
      }
    #pragma omp parallel for schedule(static)
    for (long i = 0; i < M; ++i) lam[i] = lam[i] + rhs[i];
    free(rhs);
  }
  /* uniform checkpointing window: stage states of K steps */
  double* window = malloc(sizeof(double) * M * (4 * K));
  int loaded = -1;   /* step index (0-based time level) of the loaded snapshot */
  rk4 kw;
  rk4_init(&kw, r);
  const int nact = orc_actuator_npoints(r);
  double sens = 0.0;
  double* qs = malloc(sizeof(double) * M);
  for (int t = nsteps - 1; t >= 0; --t) {   /* reverses forward step t (t -> t+1) */
    for (int i = 4; i >= 1; --i) {
      /* migrateTo(timestep, stage): state after stage (i-1) of forward step t, or level t */
      int ts = i == 1 ? t : t + 1;      /* 1-based forward step count convention */
      int st = i == 1 ? 4 : i - 1;
      const int end = nsteps;
      int need = loaded == -1 || ts < loaded || ts > loaded + K ||
                 (ts == loaded && st < 4) || (ts == loaded + K && st == 4);
      if (need) {
        int lt;
        if (ts % K == 0 && st == 4) lt = ts;
        else if (ts % K == 0) lt = ts - K;
        else lt = ts - ts % K;
        /* load the snapshot for time level lt */
        const int slot = lt == end && end % K != 0 ? nsnap - 1 : lt / K;
        memcpy(qs, snaps + (long)slot * M, sizeof(double) * M);
        double tw = snap_time[slot];
        loaded = lt;
        memcpy(window, qs, sizeof(double) * M);
        if (loaded != end) {
          int idx = 1;
          for (int s2 = loaded + 1; s2 <= loaded + K; ++s2) {
            int stop = 0;
            for (int st2 = 1; st2 <= 4; ++st2) {
              substep_forward(&kw, qs, &tw, dt, st2, ctrl_at(r, control, s2 - 1, st2));
              if (s2 == loaded + K && st2 == 4) { stop = 1; break; }
              memcpy(window + (long)idx * M, qs, sizeof(double) * M);
              idx++;
            }
            if (stop) break;
            if (s2 == end) break;
          }
        }
      }
      const int widx = (ts - 1 - loaded) * 4 + st;   /* 0-based buffer index */
      const double* qcur = window + (long)widx * M;
      if (gradient && nact > 0)
        orc_gradient(r, lam, gradient + ((long)t * 4 + (i - 1)) * nact);
      if (nact > 0) sens = sens + rk_norm[i - 1] * dt * orc_sensitivity(r, lam);
      const int final_step = (t == 0 && i == 1);
      if (nct > 0) {
        if (final_step) memset(aforce, 0, sizeof(double) * nct * nv);
        else orc_adjoint_forcing(r, qcur, aforce);
      }
      substep_adjoint(&k, qcur, lam, dt, i, nct > 0 ? aforce : NULL);
    }
  }
  if (lambda0) memcpy(lambda0, lam, sizeof(double) * M);
  free(qs); free(window); free(lam); free(aforce); free(snaps); free(snap_time); free(q);
  rk4_free(&k);
  rk4_free(&kw);
  return sens;
}
