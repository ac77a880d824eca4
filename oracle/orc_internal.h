/* orc_internal.h -- private declarations of the oracle (TEST INFRASTRUCTURE ONLY). */
#ifndef ORC_INTERNAL_H
#define ORC_INTERNAL_H
#include "magudi_oracle.h"

void orc_inviscid_jacobian_full(int nd, const double* q, const double* m, double g, double sv,
                                const double* u, double T, double* A);
void orc_first_partial_viscous(int nd, const double* q, const double* m, const double* tau,
                               const double* hf, double pl, double g, double sv,
                               const double* u, double T, double* A);
void orc_second_partial_viscous(int nd, const double* u, double mu, double lam, double kap,
                                double jac, const double* m1, const double* m2, double* B);

#define ORC_MAX_PATCHES 32
#define ORC_MAX_SOURCES 8

typedef struct {
  orc_patch_desc pd;
  int lo[3], hi[3];        /* 0-based inclusive */
  int ls[3];
  int np;
  double* sponge;          /* sponge strength per patch point */
  double* kforce;          /* Kolmogorov force per unit mass per patch point */
  double* tvflux;          /* far field: target viscous fluxes (np x nv x nd) */
  double sigma_inv, sigma_visc;
} orc_patch;

typedef struct {
  double loc[3], amplitude, omega, gfac, phase;
} orc_source;

struct orc_region {
  orc_desc d;
  int nv;
  long N;
  double* coords;
  orc_op first[3], adjfirst[3], diss[3], dissT[3];
  double *metrics, *jacobian, *norm, *arc;
  double *target, *ctrl_moll, *tgt_moll, *mean_p;
  int has_target, has_ctrl, has_tgt;
  orc_patch patches[ORC_MAX_PATCHES];
  int npatches;
  orc_source sources[ORC_MAX_SOURCES];
  int nsources;
  double re_inv, pr_inv;
  /* work arrays */
  double *sv, *vel, *p, *T, *mu, *lam, *kap, *tau, *hf, *w1, *w2, *w3;
};

#endif
