/*
 * orc_region.c -- grid, state, RHS, patches and functionals of the oracle
 * (TEST INFRASTRUCTURE ONLY).  Restates for a single block on one rank,
 * non-curvilinear (curvilinear_domain = false) grids:
 *   Grid%computeCoordinateDerivatives / Grid%update  src/GridImpl.f90:621-1065
 *   Grid%computeGradient                             src/GridImpl.f90:1172-1421
 *   State%update                                     src/StateImpl.f90:466-537
 *   computeRhsForward / computeRhsAdjoint / addDissipation /
 *   addFarFieldAdjointPenalty                        src/RhsHelperImpl.f90:10-596
 *   Region%computeRhs                                src/RegionImpl.f90:1877-2027
 *   FarField / Sponge / Actuator / CostTarget / Kolmogorov patch updates
 *   computeSpongeStrengths, computeQuadratureOnPatches src/PatchFactoryImpl.f90:161-444
 *   AcousticSource, AcousticNoise, ThermalActuator gradient/sensitivity.
 */
#include <float.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>
#include "orc_internal.h"

static double* zalloc(long n) { return (double*)calloc((size_t)(n > 0 ? n : 1), sizeof(double)); }

static int is_periodic(const orc_region* r, int d) { return r->d.periodicity[d] != ORC_NONE; }

/* index helpers */
static inline void unflat(const orc_region* r, long p, int* i, int* j, int* k) {
  const int* n = r->d.n;
  *i = (int)(p % n[0]);
  *j = (int)((p / n[0]) % n[1]);
  *k = (int)(p / ((long)n[0] * n[1]));
}

/* Grid%computeCoordinateDerivatives (src/GridImpl.f90:621-744) */
static void coordinate_derivatives(orc_region* r, int dir, double* out /* N x nd */) {
  const int nd = r->d.nd;
  const long N = r->N;
  const int* n = r->d.n;
  memcpy(out, r->coords, sizeof(double) * N * nd);
  if (r->d.periodicity[dir] != ORC_PLANE) {
    orc_op_apply(&r->first[dir], dir, n, r->d.periodicity[dir], out, nd);
    return;
  }
  /* PLANE periodicity: ghost coordinates shifted by the periodic length, then the
   * interior stencil at every point. */
  const orc_op* op = &r->first[dir];
  const int hw = op->iw / 2;
  const int len = n[dir];
  const long stride = dir == 0 ? 1 : (dir == 1 ? n[0] : (long)n[0] * n[1]);
  const double L = r->d.periodic_length[dir];
  double* g = malloc(sizeof(double) * (len + 2 * hw));
  for (long p = 0; p < N; ++p) {
    int ijk[3];
    unflat(r, p, &ijk[0], &ijk[1], &ijk[2]);
    if (ijk[dir] != 0) continue;   /* one pencil per start point */
    for (int l = 0; l < nd; ++l) {
      const double* x = r->coords + l * N + p;
      for (int i = -hw; i < len + hw; ++i) {
        int w = i < 0 ? i + len : (i >= len ? i - len : i);
        double v = x[w * stride];
        if (l == dir) {
          if (i < 0) v = v - L;
          else if (i >= len) v = v + L;
        }
        g[i + hw] = v;
      }
      for (int i = 0; i < len; ++i) {
        double s = 0.0;
        for (int k = 1; k <= hw; ++k)
          s += op->interior[k + ORC_MAXB] * (g[i + k + hw] - g[i - k + hw]);
        out[l * N + p + i * stride] = s;
      }
    }
  }
  free(g);
}

static void grid_update(orc_region* r) {
  const int nd = r->d.nd;
  const long N = r->N;
  const int* n = r->d.n;
  double* jmi = zalloc(N * nd * nd);
  double* cd = zalloc(N * nd);
  for (int j = 0; j < nd; ++j) {
    coordinate_derivatives(r, j, cd);
    for (int l = 0; l < nd; ++l) memcpy(jmi + (j * nd + l) * N, cd + l * N, sizeof(double) * N);
  }
  double* m = r->metrics;
  memset(m, 0, sizeof(double) * N * nd * nd);
  if (nd == 1) {
    for (long p = 0; p < N; ++p) {
      r->jacobian[p] = jmi[p];
      m[p] = 1.0;
      r->arc[p] = fabs(m[p]);
    }
  } else if (nd == 2) {
    for (long p = 0; p < N; ++p) {
      r->jacobian[p] = jmi[0 * N + p] * jmi[3 * N + p];
      m[0 * N + p] = jmi[3 * N + p];
      m[3 * N + p] = jmi[0 * N + p];
      r->arc[0 * N + p] = fabs(m[0 * N + p]);
      r->arc[1 * N + p] = fabs(m[3 * N + p]);
    }
  } else {
    int anyplane = 0;
    for (int d = 0; d < 3; ++d) anyplane |= r->d.periodicity[d] == ORC_PLANE;
    for (long p = 0; p < N; ++p)
      r->jacobian[p] = jmi[0 * N + p] * jmi[4 * N + p] * jmi[8 * N + p];
    if (anyplane) {
      for (long p = 0; p < N; ++p) {
        m[0 * N + p] = jmi[4 * N + p] * jmi[8 * N + p];
        m[4 * N + p] = jmi[0 * N + p] * jmi[8 * N + p];
        m[8 * N + p] = jmi[0 * N + p] * jmi[4 * N + p];
      }
    } else {
      double* F = zalloc(N);
      for (long p = 0; p < N; ++p) F[p] = jmi[4 * N + p] * r->coords[2 * N + p];
      orc_op_apply(&r->first[2], 2, n, r->d.periodicity[2], F, 1);
      memcpy(m + 0 * N, F, sizeof(double) * N);
      for (long p = 0; p < N; ++p) F[p] = jmi[8 * N + p] * r->coords[0 * N + p];
      orc_op_apply(&r->first[0], 0, n, r->d.periodicity[0], F, 1);
      memcpy(m + 4 * N, F, sizeof(double) * N);
      for (long p = 0; p < N; ++p) F[p] = jmi[0 * N + p] * r->coords[1 * N + p];
      orc_op_apply(&r->first[1], 1, n, r->d.periodicity[1], F, 1);
      memcpy(m + 8 * N, F, sizeof(double) * N);
      free(F);
    }
    for (long p = 0; p < N; ++p) {
      r->arc[0 * N + p] = fabs(m[0 * N + p]);
      r->arc[1 * N + p] = fabs(m[4 * N + p]);
      r->arc[2 * N + p] = fabs(m[8 * N + p]);
    }
  }
  for (long p = 0; p < N; ++p) r->norm[p] = 1.0;
  for (int d = 0; d < nd; ++d)
    orc_op_apply_norm(&r->first[d], d, n, is_periodic(r, d), r->norm, 1, 0);
  for (long p = 0; p < N; ++p) r->norm[p] = r->norm[p] * r->jacobian[p];
  for (long p = 0; p < N; ++p) r->jacobian[p] = 1.0 / r->jacobian[p];
  free(jmi);
  free(cd);
}

orc_region* orc_region_create(const orc_desc* d, const double* coords) {
  orc_region* r = calloc(1, sizeof *r);
  r->d = *d;
  for (int i = d->nd; i < 3; ++i) { r->d.n[i] = 1; r->d.periodicity[i] = ORC_NONE; }
  r->nv = d->nd + 2;
  r->N = (long)r->d.n[0] * r->d.n[1] * r->d.n[2];
  const long N = r->N;
  const int nd = d->nd, nv = r->nv;
  r->coords = zalloc(N * nd);
  memcpy(r->coords, coords, sizeof(double) * N * nd);
  for (int i = 0; i < nd; ++i) {
    if (orc_op_setup(&r->first[i], d->scheme, ORC_FIRST) ||
        orc_op_setup(&r->diss[i], d->scheme, d->composite_dissipation ? ORC_COMPOSITE : ORC_DISS) ||
        orc_op_setup(&r->dissT[i], d->scheme, ORC_DISS_T)) {
      orc_region_destroy(r);
      return NULL;
    }
    orc_op_adjoint(&r->first[i], &r->adjfirst[i]);
  }
  r->metrics = zalloc(N * nd * nd);
  r->jacobian = zalloc(N);
  r->norm = zalloc(N);
  r->arc = zalloc(N * nd);
  r->target = zalloc(N * nv);
  r->ctrl_moll = zalloc(N);
  r->tgt_moll = zalloc(N);
  r->mean_p = zalloc(N);
  r->sv = zalloc(N); r->vel = zalloc(N * nd); r->p = zalloc(N); r->T = zalloc(N);
  r->mu = zalloc(N); r->lam = zalloc(N); r->kap = zalloc(N);
  r->tau = zalloc(N * nd * nd); r->hf = zalloc(N * nd);
  r->w1 = zalloc(N * nv * nd); r->w2 = zalloc(N * nv * nd); r->w3 = zalloc(N * nv);
  if (d->viscous) {
    r->re_inv = d->reynolds > 0.0 ? 1.0 / d->reynolds : 0.0;
    r->pr_inv = d->prandtl > 0.0 ? 1.0 / d->prandtl : 0.0;
  }
  grid_update(r);
  return r;
}

void orc_region_destroy(orc_region* r) {
  if (!r) return;
  double* arrs[] = {r->coords, r->metrics, r->jacobian, r->norm, r->arc, r->target,
                    r->ctrl_moll, r->tgt_moll, r->mean_p, r->sv, r->vel, r->p, r->T,
                    r->mu, r->lam, r->kap, r->tau, r->hf, r->w1, r->w2, r->w3};
  for (size_t i = 0; i < sizeof arrs / sizeof arrs[0]; ++i) free(arrs[i]);
  for (int i = 0; i < r->npatches; ++i) {
    free(r->patches[i].sponge);
    free(r->patches[i].kforce);
    free(r->patches[i].tvflux);
  }
  free(r);
}

int orc_region_npoints(const orc_region* r) { return (int)r->N; }

int orc_region_get_grid(const orc_region* r, int which, double* out) {
  const long N = r->N;
  const int nd = r->d.nd;
  switch (which) {
    case 0: memcpy(out, r->metrics, sizeof(double) * N * nd * nd); return 0;
    case 1: memcpy(out, r->jacobian, sizeof(double) * N); return 0;
    case 2: memcpy(out, r->norm, sizeof(double) * N); return 0;
    case 3: memcpy(out, r->arc, sizeof(double) * N * nd); return 0;
  }
  return -1;
}

int orc_region_add_patch(orc_region* r, const orc_patch_desc* pd) {
  if (r->npatches >= ORC_MAX_PATCHES) return -1;
  orc_patch* P = &r->patches[r->npatches];
  memset(P, 0, sizeof *P);
  P->pd = *pd;
  int np = 1;
  for (int i = 0; i < 3; ++i) {
    int a = pd->extent[2 * i], b = pd->extent[2 * i + 1];
    if (i >= r->d.nd) { a = 1; b = 1; }
    if (a < 0) a = a + r->d.n[i] + 1;
    if (b < 0) b = b + r->d.n[i] + 1;
    if (a < 1 || b > r->d.n[i] || b < a) return -2;
    P->lo[i] = a - 1;
    P->hi[i] = b - 1;
    P->ls[i] = b - a + 1;
    np *= P->ls[i];
  }
  P->np = np;
  r->npatches++;
  return 0;
}

int orc_region_set_field(orc_region* r, int which, const double* v) {
  const long N = r->N;
  switch (which) {
    case 0: memcpy(r->target, v, sizeof(double) * N * r->nv); r->has_target = 1; return 0;
    case 1: memcpy(r->ctrl_moll, v, sizeof(double) * N); r->has_ctrl = 1; return 0;
    case 2: memcpy(r->tgt_moll, v, sizeof(double) * N); r->has_tgt = 1; return 0;
  }
  return -1;
}

int orc_region_add_source(orc_region* r, const double* loc, double amplitude, double frequency,
                          double radius, double phase) {
  /* setupAcousticSource (src/AcousticSourceImpl.f90:3-32) */
  if (r->nsources >= ORC_MAX_SOURCES) return -1;
  orc_source* s = &r->sources[r->nsources++];
  const double pi = 4.0 * atan(1.0);
  for (int i = 0; i < 3; ++i) s->loc[i] = i < r->d.nd ? loc[i] : 0.0;
  s->amplitude = amplitude;
  s->omega = 2.0 * pi * frequency;
  s->gfac = 9.0 / (2.0 * (radius * radius));
  s->phase = phase;
  return 0;
}

static long patch_grid_index(const orc_region* r, const orc_patch* P, int pi) {
  const int a = pi % P->ls[0];
  const int b = (pi / P->ls[0]) % P->ls[1];
  const int c = pi / (P->ls[0] * P->ls[1]);
  return (P->lo[0] + a) + (long)r->d.n[0] * ((P->lo[1] + b) + (long)r->d.n[1] * (P->lo[2] + c));
}

/* State%update: dependent and transport variables plus, for viscous flows, the stress
 * tensor and heat flux from repeated first derivatives. */
static void update_state(orc_region* r, const double* q) {
  const int nd = r->d.nd;
  const long N = r->N;
  const double g = r->d.gamma;
  #pragma omp parallel for schedule(static)
  for (long p = 0; p < N; ++p) {
    const double sv = 1.0 / q[p];
    r->sv[p] = sv;
    double ke = 0.0;
    for (int i = 0; i < nd; ++i) {
      const double u = sv * q[(i + 1) * N + p];
      r->vel[i * N + p] = u;
      ke += u * u;
    }
    r->p[p] = (g - 1.0) * (q[(nd + 1) * N + p] - 0.5 * q[p] * ke);
    r->T[p] = g * r->p[p] / (g - 1.0) * sv;
  }
  if (!r->d.viscous) return;
  #pragma omp parallel for schedule(static)
  for (long p = 0; p < N; ++p) {
    if (r->d.power_law <= 0.0) {
      r->mu[p] = r->re_inv;
      r->lam[p] = (r->d.bulk_ratio - 2.0 / 3.0) * r->re_inv;
      r->kap[p] = r->re_inv * r->pr_inv;
    } else {
      r->mu[p] = pow((g - 1.0) * r->T[p], r->d.power_law) * r->re_inv;
      r->lam[p] = (r->d.bulk_ratio - 2.0 / 3.0) * r->mu[p];
      r->kap[p] = r->mu[p] * r->pr_inv;
    }
  }
  /* velocity gradient: tau(:, i*nd + j) = du_i/dx_j  (computeGradientOfVector) */
  double* tmp = zalloc(N * nd);
  for (int j = 0; j < nd; ++j) {
    memcpy(tmp, r->vel, sizeof(double) * N * nd);
    orc_op_apply(&r->first[j], j, r->d.n, r->d.periodicity[j], tmp, nd);
    for (int i = 0; i < nd; ++i)
      #pragma omp parallel for schedule(static)
      for (long p = 0; p < N; ++p)
        r->tau[(i * nd + j) * N + p] =
            r->jacobian[p] * r->metrics[(j * nd + j) * N + p] * tmp[i * N + p];
  }
  /* computeStressTensor, in-place form (src/CNSHelperImpl.f90:406-455) */
  #pragma omp parallel for schedule(static)
  for (long p = 0; p < N; ++p) {
    double* t = r->tau;
    const double mu = r->mu[p], la = r->lam[p];
    if (nd == 1) {
      t[p] = (2.0 * mu + la) * t[p];
    } else if (nd == 2) {
      const double dv = la * (t[0 * N + p] + t[3 * N + p]);
      t[0 * N + p] = 2.0 * mu * t[0 * N + p] + dv;
      t[1 * N + p] = mu * (t[1 * N + p] + t[2 * N + p]);
      t[2 * N + p] = t[1 * N + p];
      t[3 * N + p] = 2.0 * mu * t[3 * N + p] + dv;
    } else {
      const double dv = la * (t[0 * N + p] + t[4 * N + p] + t[8 * N + p]);
      t[0 * N + p] = 2.0 * mu * t[0 * N + p] + dv;
      t[1 * N + p] = mu * (t[1 * N + p] + t[3 * N + p]);
      t[2 * N + p] = mu * (t[2 * N + p] + t[6 * N + p]);
      t[3 * N + p] = t[1 * N + p];
      t[4 * N + p] = 2.0 * mu * t[4 * N + p] + dv;
      t[5 * N + p] = mu * (t[5 * N + p] + t[7 * N + p]);
      t[6 * N + p] = t[2 * N + p];
      t[7 * N + p] = t[5 * N + p];
      t[8 * N + p] = 2.0 * mu * t[8 * N + p] + dv;
    }
  }
  /* heat flux = -kappa grad T (computeGradientOfScalar) */
  for (int j = 0; j < nd; ++j) {
    memcpy(tmp, r->T, sizeof(double) * N);
    orc_op_apply(&r->first[j], j, r->d.n, r->d.periodicity[j], tmp, 1);
    #pragma omp parallel for schedule(static)
    for (long p = 0; p < N; ++p) {
      const double gT = r->jacobian[p] * r->metrics[(j * nd + j) * N + p] * tmp[p];
      r->hf[j * N + p] = -r->kap[p] * gT;
    }
  }
  free(tmp);
}

/* computeCartesianViscousFluxes (src/CNSHelperImpl.f90:621-689) at one point, into
 * F[(d*nv + l)] for direction d. */
static void viscous_flux_point(const orc_region* r, long p, double* F) {
  const int nd = r->d.nd, nv = r->nv;
  const long N = r->N;
  const double* t = r->tau;
  for (int d = 0; d < nd; ++d) {
    F[d * nv + 0] = 0.0;
    for (int i = 0; i < nd; ++i) F[d * nv + 1 + i] = t[(i * nd + d) * N + p];
    double e = 0.0;
    for (int i = 0; i < nd; ++i) e += r->vel[i * N + p] * t[(i * nd + d) * N + p];
    F[d * nv + nd + 1] = e - r->hf[d * N + p];
  }
}

static void add_dissipation(orc_region* r, int mode, const double* x, double* rhs) {
  /* addDissipation (src/RhsHelperImpl.f90:10-87) */
  const int nd = r->d.nd, nv = r->nv;
  const long N = r->N;
  const double amount = mode == ORC_FORWARD ? r->d.dissipation_amount : -r->d.dissipation_amount;
  double* t = r->w3;
  for (int d = 0; d < nd; ++d) {
    memcpy(t, x, sizeof(double) * N * nv);
    orc_op_apply(&r->diss[d], d, r->d.n, r->d.periodicity[d], t, nv);
    if (!r->d.composite_dissipation) {
      for (int j = 0; j < nv; ++j)
        #pragma omp parallel for schedule(static)
        for (long p = 0; p < N; ++p) t[j * N + p] = -r->arc[d * N + p] * t[j * N + p];
      orc_op_apply(&r->dissT[d], d, r->d.n, r->d.periodicity[d], t, nv);
      orc_op_apply_norm(&r->first[d], d, r->d.n, is_periodic(r, d), t, nv, 1);
    }
    #pragma omp parallel for schedule(static)
    for (long k = 0; k < N * nv; ++k) rhs[k] = rhs[k] + amount * t[k];
  }
}

static void rhs_forward(orc_region* r, const double* q, double* rhs) {
  /* computeRhsForward (src/RhsHelperImpl.f90:254-354), non-curvilinear */
  const int nd = r->d.nd, nv = r->nv;
  const long N = r->N;
  double* F = r->w1;   /* transformed fluxes (N x nv x nd) */
  #pragma omp parallel for schedule(static)
  for (long k = 0; k < N * nv; ++k) rhs[k] = 0.0;
  #pragma omp parallel for schedule(static)
  for (long p = 0; p < N; ++p) {
    const double pr = r->p[p];
    double f[3][5];
    for (int d = 0; d < nd; ++d) {
      const double ud = r->vel[d * N + p];
      f[d][0] = q[(d + 1) * N + p];
      for (int i = 0; i < nd; ++i) {
        if (i == d) f[d][i + 1] = q[(d + 1) * N + p] * ud + pr;
        else if (i < d) f[d][i + 1] = q[(i + 1) * N + p] * ud;   /* = F(:, d+1, i) */
        else f[d][i + 1] = q[(d + 1) * N + p] * r->vel[i * N + p];
      }
      f[d][nd + 1] = ud * (q[(nd + 1) * N + p] + pr);
    }
    if (r->d.viscous) {
      double fv[15];
      viscous_flux_point(r, p, fv);
      for (int d = 0; d < nd; ++d)
        for (int l = 0; l < nv; ++l) f[d][l] = f[d][l] - fv[d * nv + l];
    }
    for (int d = 0; d < nd; ++d) {
      const double m = r->metrics[(d * nd + d) * N + p];
      for (int l = 0; l < nv; ++l) F[(d * nv + l) * N + p] = m * f[d][l];
    }
  }
  for (int d = 0; d < nd; ++d)
    orc_op_apply(&r->first[d], d, r->d.n, r->d.periodicity[d], F + (long)d * nv * N, nv);
  #pragma omp parallel for schedule(static)
  for (long k = 0; k < N * nv; ++k) {
    double s = F[k];
    for (int d = 1; d < nd; ++d) s = s + F[(long)d * nv * N + k];
    rhs[k] = rhs[k] - s;
  }
  if (r->d.dissipation) add_dissipation(r, ORC_FORWARD, q, rhs);
}

/* primitive -> conservative chain rule shared by the viscous adjoint terms
 * (src/RhsHelperImpl.f90:555-570 and :227-236); t2 is N x (nv-1), sign +1 adds. */
static void viscous_chain(orc_region* r, const double* q, double* t2, double sign,
                          double* rhs) {
  const int nd = r->d.nd;
  const long N = r->N;
  const double g = r->d.gamma;
  #pragma omp parallel for schedule(static)
  for (long p = 0; p < N; ++p) t2[nd * N + p] = g * r->sv[p] * t2[nd * N + p];
  for (int i = 0; i < nd; ++i)
    #pragma omp parallel for schedule(static)
    for (long p = 0; p < N; ++p)
      t2[i * N + p] = r->sv[p] * t2[i * N + p] - r->vel[i * N + p] * t2[nd * N + p];
  for (int l = 1; l <= nd + 1; ++l)
    #pragma omp parallel for schedule(static)
    for (long p = 0; p < N; ++p)
      rhs[l * N + p] = sign > 0 ? rhs[l * N + p] + t2[(l - 1) * N + p]
                                : rhs[l * N + p] - t2[(l - 1) * N + p];
  #pragma omp parallel for schedule(static)
  for (long p = 0; p < N; ++p) {
    double s = r->vel[p] * t2[p];
    for (int i = 1; i < nd; ++i) s = s + r->vel[i * N + p] * t2[i * N + p];
    const double a = r->sv[p] * q[(nd + 1) * N + p] * t2[nd * N + p];
    if (sign > 0) rhs[p] = rhs[p] - a - s;
    else rhs[p] = rhs[p] + a + s;
  }
}

static void rhs_adjoint(orc_region* r, const double* q, const double* lam, double* rhs) {
  /* computeRhsAdjoint (src/RhsHelperImpl.f90:356-596) */
  const int nd = r->d.nd, nv = r->nv;
  const long N = r->N;
  const double g = r->d.gamma;
  double* t1 = r->w1;   /* N x nv x nd */
  #pragma omp parallel for schedule(static)
  for (long k = 0; k < N * nv; ++k) rhs[k] = 0.0;
  for (int d = 0; d < nd; ++d) {
    memcpy(t1 + (long)d * nv * N, lam, sizeof(double) * N * nv);
    orc_op_apply(&r->adjfirst[d], d, r->d.n, r->d.periodicity[d], t1 + (long)d * nv * N, nv);
  }
  #pragma omp parallel for schedule(static)
  for (long p = 0; p < N; ++p) {
    double qp[5], u[3], tau[9], hf[3], A[25], B[25];
    for (int l = 0; l < nv; ++l) qp[l] = q[l * N + p];
    for (int i = 0; i < nd; ++i) u[i] = r->vel[i * N + p];
    if (r->d.viscous) {
      for (int k = 0; k < nd * nd; ++k) tau[k] = r->tau[k * N + p];
      for (int i = 0; i < nd; ++i) hf[i] = r->hf[i * N + p];
    }
    for (int d = 0; d < nd; ++d) {
      double m[3];
      for (int i = 0; i < nd; ++i) m[i] = r->metrics[(d * nd + i) * N + p];
      orc_inviscid_jacobian_full(nd, qp, m, g, r->sv[p], u, r->T[p], A);
      if (r->d.viscous) {
        orc_first_partial_viscous(nd, qp, m, tau, hf, r->d.power_law, g, r->sv[p], u, r->T[p], B);
        for (int k = 0; k < nv * nv; ++k) A[k] = A[k] - B[k];
      }
      for (int i = 0; i < nv; ++i) {
        double s = 0.0;
        for (int k = 0; k < nv; ++k) s += A[i * nv + k] * t1[((long)d * nv + k) * N + p];
        rhs[i * N + p] = rhs[i * N + p] + s;
      }
    }
  }
  if (r->d.viscous) {
    /* second-partial viscous terms (src/RhsHelperImpl.f90:509-570) */
    const int nw = nv - 1;
    #pragma omp parallel for schedule(static)
    for (long p = 0; p < N; ++p) {
      double u[3], diff[3][4], B[16];
      for (int i = 0; i < nd; ++i) u[i] = r->vel[i * N + p];
      memset(diff, 0, sizeof diff);
      for (int j = 0; j < nd; ++j) {
        double m2[3];
        for (int i = 0; i < nd; ++i) m2[i] = r->metrics[(j * nd + i) * N + p];
        for (int i = 0; i < nd; ++i) {
          double m1[3];
          for (int k = 0; k < nd; ++k) m1[k] = r->metrics[(i * nd + k) * N + p];
          orc_second_partial_viscous(nd, u, r->mu[p], r->lam[p], r->kap[p], r->jacobian[p],
                                     m1, m2, B);
          for (int a = 0; a < nw; ++a) {
            double s = 0.0;
            for (int k = 0; k < nw; ++k) s += B[a * nw + k] * t1[((long)i * nv + 1 + k) * N + p];
            diff[j][a] = diff[j][a] + s;
          }
        }
      }
      for (int j = 0; j < nd; ++j)
        for (int a = 0; a < nw; ++a) t1[((long)j * nv + 1 + a) * N + p] = diff[j][a];
    }
    for (int j = 0; j < nd; ++j)
      orc_op_apply(&r->adjfirst[j], j, r->d.n, r->d.periodicity[j], t1 + ((long)j * nv + 1) * N,
                   nw);
    double* t2 = r->w2;
    for (int a = 0; a < nw; ++a)
      #pragma omp parallel for schedule(static)
      for (long p = 0; p < N; ++p) {
        double s = t1[(1 + a) * N + p];
        for (int j = 1; j < nd; ++j) s = s + t1[((long)j * nv + 1 + a) * N + p];
        t2[a * N + p] = s;
      }
    viscous_chain(r, q, t2, -1.0, rhs);
  }
  if (r->d.dissipation) add_dissipation(r, ORC_ADJOINT, lam, rhs);
  if (r->d.viscous) {
    /* addFarFieldAdjointPenalty (src/RhsHelperImpl.f90:89-250) */
    int any = 0;
    for (int i = 0; i < r->npatches; ++i) any |= r->patches[i].pd.type == ORC_FAR_FIELD;
    if (any) {
      const int nw = nv - 1;
      double* tt = r->w1;
      memset(tt, 0, sizeof(double) * N * nw * nd);
      for (int ip = 0; ip < r->npatches; ++ip) {
        const orc_patch* P = &r->patches[ip];
        if (P->pd.type != ORC_FAR_FIELD) continue;
        const int dir = abs(P->pd.normal_dir) - 1;
        for (int pi = 0; pi < P->np; ++pi) {
          const long p = patch_grid_index(r, P, pi);
          double u[3], m1[3], B[16];
          for (int i = 0; i < nd; ++i) u[i] = r->vel[i * N + p];
          for (int i = 0; i < nd; ++i) m1[i] = r->metrics[(dir * nd + i) * N + p];
          for (int l = 0; l < nd; ++l) {
            double m2[3];
            for (int i = 0; i < nd; ++i) m2[i] = r->metrics[(l * nd + i) * N + p];
            orc_second_partial_viscous(nd, u, r->mu[p], r->lam[p], r->kap[p], r->jacobian[p],
                                       m1, m2, B);
            for (int a = 0; a < nw; ++a) {
              double s = 0.0;
              for (int k = 0; k < nw; ++k) s += B[a * nw + k] * lam[(1 + k) * N + p];
              tt[((long)l * nw + a) * N + p] = tt[((long)l * nw + a) * N + p] - P->sigma_visc * s;
            }
          }
        }
      }
      for (int l = 0; l < nd; ++l)
        orc_op_apply(&r->adjfirst[l], l, r->d.n, r->d.periodicity[l], tt + (long)l * nw * N, nw);
      double* t2 = r->w2;
      for (int a = 0; a < nw; ++a)
        #pragma omp parallel for schedule(static)
        for (long p = 0; p < N; ++p) {
          double s = tt[a * N + p];
          for (int l = 1; l < nd; ++l) s = s + tt[((long)l * nw + a) * N + p];
          t2[a * N + p] = s;
        }
      viscous_chain(r, q, t2, +1.0, rhs);
    }
  }
}

static double quadrature(const orc_region* r, int type, const double* integrand_full);

/* normalizeKolmogorovForcing (src/RegionImpl.f90:910-999, called at :1477): the forces of
 * every Kolmogorov patch, summed per grid point (disperseAdd), and the count of patches
 * covering each point are integrated over the union of those patches
 * (computeQuadratureOnPatches); each patch's force then loses globalForcing /
 * globalVolume, so that the forcing has zero quadrature mean over the patches. */
static void normalize_kolmogorov(orc_region* r) {
  const long N = r->N;
  int any = 0;
  for (int ip = 0; ip < r->npatches; ++ip) any |= r->patches[ip].pd.type == ORC_KOLMOGOROV;
  if (!any) return;
  double* gf = zalloc(N);
  double* gv = zalloc(N);
  for (int ip = 0; ip < r->npatches; ++ip) {
    const orc_patch* P = &r->patches[ip];
    if (P->pd.type != ORC_KOLMOGOROV) continue;
    for (int k = 0; k < P->np; ++k) {
      const long p = patch_grid_index(r, P, k);
      gf[p] = gf[p] + P->kforce[k];
      gv[p] = gv[p] + 1.0;
    }
  }
  const double volume = quadrature(r, ORC_KOLMOGOROV, gv);
  const double forcing = quadrature(r, ORC_KOLMOGOROV, gf);
  for (int ip = 0; ip < r->npatches; ++ip) {
    orc_patch* P = &r->patches[ip];
    if (P->pd.type != ORC_KOLMOGOROV) continue;
    for (int k = 0; k < P->np; ++k) P->kforce[k] = P->kforce[k] - forcing / volume;
  }
  free(gf);
  free(gv);
}

int orc_region_finalize(orc_region* r) {
  const int nd = r->d.nd, nv = r->nv;
  const long N = r->N;
  const double g = r->d.gamma;
  /* mean pressure from the target state (src/AcousticNoiseImpl.f90:56-71) */
  #pragma omp parallel for schedule(static)
  for (long p = 0; p < N; ++p) {
    double s = 0.0;
    for (int i = 0; i < nd; ++i) s += r->target[(i + 1) * N + p] * r->target[(i + 1) * N + p];
    r->mean_p[p] = (g - 1.0) * (r->target[(nd + 1) * N + p] - 0.5 * s / r->target[p]);
  }
  /* sponge strengths (computeSpongeStrengths, src/PatchFactoryImpl.f90:161-374) */
  double* cd = zalloc(N * nd);
  for (int dir = 0; dir < nd; ++dir) {
    int any = 0;
    for (int i = 0; i < r->npatches; ++i)
      any |= r->patches[i].pd.type == ORC_SPONGE && abs(r->patches[i].pd.normal_dir) == dir + 1;
    if (!any) continue;
    coordinate_derivatives(r, dir, cd);
    const int len = r->d.n[dir];
    double* arcl = zalloc(N);
    #pragma omp parallel for schedule(static)
    for (long p = 0; p < N; ++p) {
      double s = 0.0;
      for (int l = 0; l < nd; ++l) s += cd[l * N + p] * cd[l * N + p];
      arcl[p] = sqrt(s);
    }
    const long stride = dir == 0 ? 1 : (dir == 1 ? r->d.n[0] : (long)r->d.n[0] * r->d.n[1]);
    double* cli = zalloc(len);
    for (int ip = 0; ip < r->npatches; ++ip) {
      orc_patch* P = &r->patches[ip];
      if (P->pd.type != ORC_SPONGE || abs(P->pd.normal_dir) != dir + 1) continue;
      if (!P->sponge) P->sponge = zalloc(P->np);
      const int mn = P->lo[dir], mx = P->hi[dir];   /* 0-based iMin-1, iMax-1 */
      for (int pi = 0; pi < P->np; ++pi) {
        const long p = patch_grid_index(r, P, pi);
        int ijk[3];
        unflat(r, p, &ijk[0], &ijk[1], &ijk[2]);
        const long p0 = p - (long)ijk[dir] * stride;
        for (int i = 0; i < len; ++i) cli[i] = arcl[p0 + i * stride];
        const int i = ijk[dir];
        double a = 0.0, b = 0.0;
        if (P->pd.normal_dir > 0) {
          for (int k = mn; k <= i - 1; ++k) a += cli[k];
          for (int k = mn; k <= mx - 1; ++k) b += cli[k];
        } else {
          for (int k = i + 1; k <= mx; ++k) a += cli[k];
          for (int k = mn + 1; k <= mx; ++k) b += cli[k];
        }
        P->sponge[pi] = a / b;
      }
      const double amount = P->pd.p[0];
      const double expo = P->pd.p[1];
      for (int pi = 0; pi < P->np; ++pi)
        P->sponge[pi] = amount * pow(1.0 - P->sponge[pi], expo);
    }
    free(cli);
    free(arcl);
  }
  free(cd);
  /* far-field penalty parameters (src/FarFieldPatchImpl.f90:3-74) and target viscous
   * fluxes (src/PatchFactoryImpl.f90:543-568) */
  int anyff = 0;
  for (int ip = 0; ip < r->npatches; ++ip) {
    orc_patch* P = &r->patches[ip];
    if (P->pd.type == ORC_FAR_FIELD) {
      anyff = 1;
      const int dir = abs(P->pd.normal_dir) - 1;
      const double h0 = r->first[dir].norm[0];
      P->sigma_inv = copysign(P->pd.p[0], (double)P->pd.normal_dir) / h0;
      P->sigma_visc = r->d.viscous ? copysign(P->pd.p[1], (double)P->pd.normal_dir) / h0 : 0.0;
    } else if (P->pd.type == ORC_IMPENETRABLE_WALL || P->pd.type == ORC_ISOTHERMAL_WALL) {
      /* setupImpenetrableWall (src/ImpenetrableWallImpl.f90:3-40): sign(amount, normal) / H00;
       * setupIsothermalWall (src/IsothermalWallImpl.f90:3-80): viscous amount / H00 / Re */
      const int dir = abs(P->pd.normal_dir) - 1;
      const double h0 = r->first[dir].norm[0];
      P->sigma_inv = copysign(P->pd.p[0], (double)P->pd.normal_dir) / h0;
      P->sigma_visc = 0.0;
      if (P->pd.type == ORC_ISOTHERMAL_WALL && r->d.viscous)
        P->sigma_visc = P->pd.p[1] / h0 * r->re_inv;
    } else if (P->pd.type == ORC_KOLMOGOROV) {
      const double pi = 4.0 * atan(1.0);
      const int wn = P->pd.p[0] > 0 ? (int)P->pd.p[0] : 0;
      P->kforce = zalloc(P->np);
      #pragma omp parallel for schedule(static)
      for (int k = 0; k < P->np; ++k) {
        const long p = patch_grid_index(r, P, k);
        P->kforce[k] = P->pd.p[1] * sin(2.0 * pi * wn * r->coords[1 * N + p]);
      }
    }
  }
  normalize_kolmogorov(r);
  if (anyff && r->d.viscous) {
    update_state(r, r->target);
    for (int ip = 0; ip < r->npatches; ++ip) {
      orc_patch* P = &r->patches[ip];
      if (P->pd.type != ORC_FAR_FIELD) continue;
      P->tvflux = zalloc((long)P->np * nv * nd);
      #pragma omp parallel for schedule(static)
      for (int k = 0; k < P->np; ++k) {
        double fv[15];
        viscous_flux_point(r, patch_grid_index(r, P, k), fv);
        for (int d = 0; d < nd; ++d)
          for (int l = 0; l < nv; ++l) P->tvflux[((long)d * nv + l) * P->np + k] = fv[d * nv + l];
      }
    }
  }
  return 0;
}

static int patch_index_of_type(const orc_region* r, int type) {
  for (int i = 0; i < r->npatches; ++i) if (r->patches[i].pd.type == type) return i;
  return -1;
}

int orc_cost_target_npoints(const orc_region* r) {
  int i = patch_index_of_type(r, ORC_COST_TARGET);
  return i < 0 ? 0 : r->patches[i].np;
}

int orc_actuator_npoints(const orc_region* r) {
  int i = patch_index_of_type(r, ORC_ACTUATOR);
  return i < 0 ? 0 : r->patches[i].np;
}

int orc_compute_rhs(orc_region* r, int mode, const double* q, const double* lambda,
                    double time, const double* control_forcing, const double* adjoint_forcing,
                    double forcing_factor, double* rhs) {
  const int nd = r->d.nd, nv = r->nv;
  const long N = r->N;
  update_state(r, q);
  if (mode == ORC_FORWARD) rhs_forward(r, q, rhs);
  else rhs_adjoint(r, q, lambda, rhs);
  for (int j = 0; j < nv; ++j)
    #pragma omp parallel for schedule(static)
    for (long p = 0; p < N; ++p) rhs[j * N + p] = rhs[j * N + p] * r->jacobian[p];
  for (int ip = 0; ip < r->npatches; ++ip) {
    const orc_patch* P = &r->patches[ip];
    switch (P->pd.type) {
      case ORC_FAR_FIELD: {
        /* addFarFieldPenalty (src/FarFieldPatchImpl.f90:93-286) */
        const int dir = abs(P->pd.normal_dir) - 1;
        #pragma omp parallel for schedule(static)
        for (int k = 0; k < P->np; ++k) {
          const long p = patch_grid_index(r, P, k);
          double qt[5], m[3], A[25];
          for (int l = 0; l < nv; ++l) qt[l] = r->target[l * N + p];
          for (int i = 0; i < nd; ++i) m[i] = r->metrics[(dir * nd + i) * N + p];
          orc_incoming_jacobian(nd, qt, m, r->d.gamma, P->pd.normal_dir, A);
          const double sj = P->sigma_inv * r->jacobian[p];
          if (mode == ORC_FORWARD) {
            for (int i = 0; i < nv; ++i) {
              double s = 0.0;
              for (int c = 0; c < nv; ++c) s += A[c * nv + i] * (q[c * N + p] - qt[c]);
              rhs[i * N + p] = rhs[i * N + p] - sj * s;
            }
            if (r->d.viscous) {
              double fv[15];
              viscous_flux_point(r, p, fv);
              const double vj = P->sigma_visc * r->jacobian[p];
              for (int l = 0; l < nv; ++l) {
                double s = 0.0;
                for (int d = 0; d < nd; ++d)
                  s += (fv[d * nv + l] - P->tvflux[((long)d * nv + l) * P->np + k]) * m[d];
                rhs[l * N + p] = rhs[l * N + p] + vj * s;
              }
            }
          } else {
            for (int i = 0; i < nv; ++i) {
              double s = 0.0;
              for (int c = 0; c < nv; ++c) s += A[i * nv + c] * lambda[c * N + p];
              rhs[i * N + p] = rhs[i * N + p] + sj * s;
            }
            if (r->d.viscous) {
              double qp[5], u[3], tau[9], hf[3], B[25];
              for (int l = 0; l < nv; ++l) qp[l] = q[l * N + p];
              for (int i = 0; i < nd; ++i) u[i] = r->vel[i * N + p];
              for (int c = 0; c < nd * nd; ++c) tau[c] = r->tau[c * N + p];
              for (int i = 0; i < nd; ++i) hf[i] = r->hf[i * N + p];
              orc_first_partial_viscous(nd, qp, m, tau, hf, r->d.power_law, r->d.gamma, r->sv[p],
                                        u, r->T[p], B);
              const double vj = P->sigma_visc * r->jacobian[p];
              for (int i = 0; i < nv; ++i) {
                double s = 0.0;
                for (int c = 0; c < nv; ++c) s += B[i * nv + c] * lambda[c * N + p];
                rhs[i * N + p] = rhs[i * N + p] - vj * s;
              }
            }
          }
        }
        break;
      }
      case ORC_IMPENETRABLE_WALL:
      case ORC_ISOTHERMAL_WALL: {
        /* addImpenetrableWallPenalty (src/ImpenetrableWallImpl.f90:60-207) */
        const int dir = abs(P->pd.normal_dir) - 1;
        const double g = r->d.gamma;
        #pragma omp parallel for schedule(static)
        for (int k = 0; k < P->np; ++k) {
          const long p = patch_grid_index(r, P, k);
          double qp[5], m[3];
          for (int l = 0; l < nv; ++l) qp[l] = q[l * N + p];
          for (int i = 0; i < nd; ++i) m[i] = r->metrics[(dir * nd + i) * N + p];
          double un = 0.0;
          for (int i = 0; i < nd; ++i) un += qp[i + 1] * m[i];
          const double sj = P->sigma_inv * r->jacobian[p];
          if (mode == ORC_FORWARD) {
            double pen[5];
            pen[0] = un;
            for (int i = 0; i < nd; ++i) pen[i + 1] = un * r->vel[i * N + p];
            pen[nd + 1] = un * r->sv[p] * (qp[nd + 1] + r->p[p]);
            for (int l = 0; l < nv; ++l) rhs[l * N + p] = rhs[l * N + p] - sj * pen[l];
          } else {
            /* d(penalty)/dq: inviscid flux Jacobian along the normal metrics minus the
             * pressure term of the momentum rows */
            double A[25], dp[5], ke = 0.0;
            orc_inviscid_jacobian(nd, qp, m, g, A);
            for (int i = 0; i < nd; ++i) ke += r->vel[i * N + p] * r->vel[i * N + p];
            dp[0] = 0.5 * ke;
            for (int i = 0; i < nd; ++i) dp[i + 1] = -r->vel[i * N + p];
            dp[nd + 1] = 1.0;
            for (int l = 0; l < nv; ++l) dp[l] = dp[l] * (g - 1.0);
            for (int i = 0; i < nd; ++i)
              for (int c = 0; c < nv; ++c) A[c * nv + (i + 1)] = A[c * nv + (i + 1)] - m[i] * dp[c];
            for (int i = 0; i < nv; ++i) {
              double s = 0.0;
              for (int c = 0; c < nv; ++c) s += A[i * nv + c] * lambda[c * N + p];
              rhs[i * N + p] = rhs[i * N + p] + sj * s;
            }
          }
          if (P->pd.type == ORC_ISOTHERMAL_WALL && r->d.viscous) {
            /* addIsothermalWallPenalty (src/IsothermalWallImpl.f90:101-330) */
            const double tw = P->pd.p[2];
            const double jac = r->jacobian[p];
            if (mode == ORC_FORWARD) {
              double pen[5];
              pen[0] = 0.0;
              for (int l = 1; l < nv; ++l) pen[l] = qp[l];
              pen[nd + 1] = pen[nd + 1] - qp[0] * tw / g;
              for (int l = 0; l < nv; ++l) pen[l] = jac * pen[l];
              for (int l = 0; l < nv; ++l) rhs[l * N + p] = rhs[l * N + p] - P->sigma_visc * pen[l];
            } else {
              double pen[5];
              pen[0] = -lambda[(nd + 1) * N + p] * tw / g;
              for (int l = 1; l < nv; ++l) pen[l] = lambda[l * N + p];
              for (int l = 0; l < nv; ++l) pen[l] = jac * pen[l];
              for (int l = 0; l < nv; ++l) rhs[l * N + p] = rhs[l * N + p] + P->sigma_visc * pen[l];
            }
          }
        }
        break;
      }
      case ORC_SPONGE: {
        /* addDamping (src/SpongePatchImpl.f90:65-166) */
        for (int l = 0; l < nv; ++l)
          #pragma omp parallel for schedule(static)
          for (int k = 0; k < P->np; ++k) {
            const long p = patch_grid_index(r, P, k);
            if (mode == ORC_FORWARD)
              rhs[l * N + p] = rhs[l * N + p] - P->sponge[k] * (q[l * N + p] - r->target[l * N + p]);
            else
              rhs[l * N + p] = rhs[l * N + p] + P->sponge[k] * lambda[l * N + p];
          }
        break;
      }
      case ORC_ACTUATOR: {
        /* updateActuatorPatch (src/ActuatorPatchImpl.f90:108-181): energy only */
        if (mode != ORC_FORWARD || !control_forcing) break;
        #pragma omp parallel for schedule(static)
        for (int k = 0; k < P->np; ++k) {
          const long p = patch_grid_index(r, P, k);
          rhs[(nd + 1) * N + p] = rhs[(nd + 1) * N + p] + r->ctrl_moll[p] * control_forcing[k];
        }
        break;
      }
      case ORC_COST_TARGET: {
        /* addAdjointForcing (src/CostTargetPatchImpl.f90:72-136) */
        if (mode == ORC_FORWARD || !adjoint_forcing) break;
        for (int l = 0; l < nv; ++l)
          #pragma omp parallel for schedule(static)
          for (int k = 0; k < P->np; ++k) {
            const long p = patch_grid_index(r, P, k);
            rhs[l * N + p] = rhs[l * N + p] + forcing_factor * adjoint_forcing[(long)l * P->np + k];
          }
        break;
      }
      case ORC_KOLMOGOROV: {
        #pragma omp parallel for schedule(static)
        for (int k = 0; k < P->np; ++k) {
          const long p = patch_grid_index(r, P, k);
          if (mode == ORC_FORWARD)
            rhs[1 * N + p] = rhs[1 * N + p] + q[p] * P->kforce[k];
          else
            rhs[p] = rhs[p] - lambda[1 * N + p] * P->kforce[k];
        }
        break;
      }
    }
  }
  if (mode == ORC_FORWARD) {
    /* addAcousticSource (src/AcousticSourceImpl.f90:34-64) */
    for (int s = 0; s < r->nsources; ++s) {
      const orc_source* S = &r->sources[s];
      const double a = S->amplitude * cos(S->omega * time + S->phase);
      #pragma omp parallel for schedule(static)
      for (long p = 0; p < N; ++p) {
        double rr = 0.0;
        for (int i = 0; i < nd; ++i) {
          const double dx = r->coords[i * N + p] - S->loc[i];
          rr += dx * dx;
        }
        rhs[(nd + 1) * N + p] = rhs[(nd + 1) * N + p] + a * exp(-S->gfac * rr);
      }
    }
  }
  return 0;
}

static double quadrature(const orc_region* r, int type, const double* integrand_full) {
  /* computeQuadratureOnPatches: mask of all patches of `type`, then sum(mask*norm*f) */
  const long N = r->N;
  double* mask = zalloc(N);
  for (int ip = 0; ip < r->npatches; ++ip) {
    const orc_patch* P = &r->patches[ip];
    if (P->pd.type != type) continue;
    #pragma omp parallel for schedule(static)
    for (int k = 0; k < P->np; ++k) mask[patch_grid_index(r, P, k)] = 1.0;
  }
  double s = 0.0;
  for (long p = 0; p < N; ++p) s += mask[p] * r->norm[p] * integrand_full[p];
  free(mask);
  return s;
}

/* Local wave speeds of computeTimeStepSize / computeCfl (src/CNSHelperImpl.f90:842-982) at
 * point p after update_state: advection J (c |m| + sum_j |u . m_j|) with m_j the j-th row
 * of the metrics; diffusion J^2 sum(metrics^2) max(2 mu, kappa) (its first loop over the
 * rows is overwritten in the reference and has no effect). */
static double advective_speed(const orc_region* r, long p) {
  const int nd = r->d.nd;
  const long N = r->N;
  const double c = sqrt((r->d.gamma - 1.0) * r->T[p]);
  double w = 0.0;
  for (int j = 0; j < nd; ++j) {
    double s = 0.0;
    for (int k = 0; k < nd; ++k) {
      const double m = r->metrics[(j * nd + k) * N + p];
      s += m * m;
    }
    w = w + s;
  }
  w = c * sqrt(w);
  for (int j = 0; j < nd; ++j) {
    double d = 0.0;
    for (int k = 0; k < nd; ++k) d += r->vel[k * N + p] * r->metrics[(j * nd + k) * N + p];
    w = w + fabs(d);
  }
  return r->jacobian[p] * w;
}

static double diffusive_speed(const orc_region* r, long p) {
  const int nd = r->d.nd;
  const long N = r->N;
  double s = 0.0;
  for (int k = 0; k < nd * nd; ++k) {
    const double m = r->metrics[k * N + p];
    s += m * m;
  }
  const double jac = r->jacobian[p];
  const double a = 2.0 * r->mu[p], b = r->kap[p];
  return jac * jac * s * (a > b ? a : b);
}

/* computeTimeStepSize (src/CNSHelperImpl.f90:913-982): min over points of cfl / speed. */
double orc_time_step_size(orc_region* r, const double* q, double cfl) {
  update_state(r, q);
  double dt = DBL_MAX;
  for (long p = 0; p < r->N; ++p) {
    const double v = cfl / advective_speed(r, p);
    dt = v < dt ? v : dt;
  }
  if (r->d.viscous)
    for (long p = 0; p < r->N; ++p) {
      const double v = cfl / diffusive_speed(r, p);
      dt = v < dt ? v : dt;
    }
  return dt;
}

/* computeCfl (src/CNSHelperImpl.f90:842-911): max over points of speed * dt. */
double orc_cfl(orc_region* r, const double* q, double dt) {
  update_state(r, q);
  double cfl = 0.0;
  for (long p = 0; p < r->N; ++p) {
    const double v = advective_speed(r, p) * dt;
    cfl = v > cfl ? v : cfl;
  }
  if (r->d.viscous)
    for (long p = 0; p < r->N; ++p) {
      const double v = diffusive_speed(r, p) * dt;
      cfl = v > cfl ? v : cfl;
    }
  return cfl;
}

double orc_functional(orc_region* r, const double* q) {
  update_state(r, q);
  const long N = r->N;
  double* f = zalloc(N);
  #pragma omp parallel for schedule(static)
  for (long p = 0; p < N; ++p) {
    const double F = r->p[p] - r->mean_p[p];
    f[p] = F * F * r->tgt_moll[p];
  }
  double J = quadrature(r, ORC_COST_TARGET, f);
  free(f);
  return J;
}

int orc_adjoint_forcing(orc_region* r, const double* q, double* forcing) {
  const int ip = patch_index_of_type(r, ORC_COST_TARGET);
  if (ip < 0) return -1;
  update_state(r, q);
  const int nd = r->d.nd;
  const long N = r->N;
  const orc_patch* P = &r->patches[ip];
  #pragma omp parallel for schedule(static)
  for (int k = 0; k < P->np; ++k) {
    const long p = patch_grid_index(r, P, k);
    const double F = -2.0 * r->tgt_moll[p] * 1.0 * (r->d.gamma - 1.0) * (r->p[p] - r->mean_p[p]);
    forcing[(long)(nd + 1) * P->np + k] = F;
    double ke = 0.0;
    for (int i = 0; i < nd; ++i) {
      forcing[(long)(i + 1) * P->np + k] = -r->vel[i * N + p] * F;
      ke += r->vel[i * N + p] * r->vel[i * N + p];
    }
    forcing[k] = 0.5 * ke * F;
  }
  return 0;
}

int orc_gradient(orc_region* r, const double* lambda, double* g) {
  const int ip = patch_index_of_type(r, ORC_ACTUATOR);
  if (ip < 0) return -1;
  const orc_patch* P = &r->patches[ip];
  const int nd = r->d.nd;
  #pragma omp parallel for schedule(static)
  for (int k = 0; k < P->np; ++k) {
    const long p = patch_grid_index(r, P, k);
    g[k] = lambda[(nd + 1) * r->N + p] * (r->ctrl_moll[p] * 1.0);
  }
  return 0;
}

double orc_sensitivity(orc_region* r, const double* lambda) {
  const long N = r->N;
  const int nd = r->d.nd;
  double* f = zalloc(N);
  #pragma omp parallel for schedule(static)
  for (long p = 0; p < N; ++p) {
    const double F = lambda[(nd + 1) * N + p] * r->ctrl_moll[p] * 1.0;
    f[p] = F * F;
  }
  double s = quadrature(r, ORC_ACTUATOR, f);
  free(f);
  return s;
}
