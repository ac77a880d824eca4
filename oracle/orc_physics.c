/*
 * orc_physics.c -- pointwise compressible-flow routines of the oracle
 * (TEST INFRASTRUCTURE ONLY).  Restates src/CNSHelperImpl.f90 of magudi:
 *   computeJacobianOfInviscidFlux2D/3D          :1098-1444
 *   computeIncomingJacobianOfInviscidFlux2D/3D  :1669-2342 (base matrix only)
 *   computeFirstPartialViscousJacobian2D/3D     :2412-2600
 *   computeSecondPartialViscousJacobian2D/3D    :2636-2756
 * Matrices are column-major nv x nv: A[(j-1)*nv + (i-1)] = A(i,j).
 */
#include <math.h>
#include <string.h>
#include "orc_internal.h"

#define M(A, n, i, j) ((A)[((j) - 1) * (n) + ((i) - 1)])

void orc_inviscid_jacobian_full(int nd, const double* q, const double* m, double g,
                                double sv, const double* u, double T, double* A) {
  if (nd == 2) {
    const double U = m[0] * u[0] + m[1] * u[1];
    const double phi2 = 0.5 * (g - 1.0) * (u[0] * u[0] + u[1] * u[1]);
    M(A, 4, 1, 1) = 0.0;
    M(A, 4, 2, 1) = phi2 * m[0] - U * u[0];
    M(A, 4, 3, 1) = phi2 * m[1] - U * u[1];
    M(A, 4, 4, 1) = U * ((g - 2.0) / (g - 1.0) * phi2 - T);
    M(A, 4, 1, 2) = m[0];
    M(A, 4, 2, 2) = U - (g - 2.0) * u[0] * m[0];
    M(A, 4, 3, 2) = u[1] * m[0] - (g - 1.0) * u[0] * m[1];
    M(A, 4, 4, 2) = (T + phi2 / (g - 1.0)) * m[0] - (g - 1.0) * U * u[0];
    M(A, 4, 1, 3) = m[1];
    M(A, 4, 2, 3) = u[0] * m[1] - (g - 1.0) * u[1] * m[0];
    M(A, 4, 3, 3) = U - (g - 2.0) * u[1] * m[1];
    M(A, 4, 4, 3) = (T + phi2 / (g - 1.0)) * m[1] - (g - 1.0) * U * u[1];
    M(A, 4, 1, 4) = 0.0;
    M(A, 4, 2, 4) = (g - 1.0) * m[0];
    M(A, 4, 3, 4) = (g - 1.0) * m[1];
    M(A, 4, 4, 4) = g * U;
  } else if (nd == 3) {
    const double U = m[0] * u[0] + m[1] * u[1] + m[2] * u[2];
    const double phi2 = 0.5 * (g - 1.0) * (u[0] * u[0] + u[1] * u[1] + u[2] * u[2]);
    M(A, 5, 1, 1) = 0.0;
    M(A, 5, 2, 1) = phi2 * m[0] - U * u[0];
    M(A, 5, 3, 1) = phi2 * m[1] - U * u[1];
    M(A, 5, 4, 1) = phi2 * m[2] - U * u[2];
    M(A, 5, 5, 1) = U * ((g - 2.0) / (g - 1.0) * phi2 - T);
    M(A, 5, 1, 2) = m[0];
    M(A, 5, 2, 2) = U - (g - 2.0) * u[0] * m[0];
    M(A, 5, 3, 2) = u[1] * m[0] - (g - 1.0) * u[0] * m[1];
    M(A, 5, 4, 2) = u[2] * m[0] - (g - 1.0) * u[0] * m[2];
    M(A, 5, 5, 2) = (T + phi2 / (g - 1.0)) * m[0] - (g - 1.0) * U * u[0];
    M(A, 5, 1, 3) = m[1];
    M(A, 5, 2, 3) = u[0] * m[1] - (g - 1.0) * u[1] * m[0];
    M(A, 5, 3, 3) = U - (g - 2.0) * u[1] * m[1];
    M(A, 5, 4, 3) = u[2] * m[1] - (g - 1.0) * u[1] * m[2];
    M(A, 5, 5, 3) = (T + phi2 / (g - 1.0)) * m[1] - (g - 1.0) * U * u[1];
    M(A, 5, 1, 4) = m[2];
    M(A, 5, 2, 4) = u[0] * m[2] - (g - 1.0) * u[2] * m[0];
    M(A, 5, 3, 4) = u[1] * m[2] - (g - 1.0) * u[2] * m[1];
    M(A, 5, 4, 4) = U - (g - 2.0) * u[2] * m[2];
    M(A, 5, 5, 4) = (T + phi2 / (g - 1.0)) * m[2] - (g - 1.0) * U * u[2];
    M(A, 5, 1, 5) = 0.0;
    M(A, 5, 2, 5) = (g - 1.0) * m[0];
    M(A, 5, 3, 5) = (g - 1.0) * m[1];
    M(A, 5, 4, 5) = (g - 1.0) * m[2];
    M(A, 5, 5, 5) = g * U;
  } else {
    /* 1D (computeJacobianOfInviscidFlux1D, :984-1096) */
    const double U = m[0] * u[0];
    const double phi2 = 0.5 * (g - 1.0) * u[0] * u[0];
    M(A, 3, 1, 1) = 0.0;
    M(A, 3, 2, 1) = phi2 * m[0] - U * u[0];
    M(A, 3, 3, 1) = U * ((g - 2.0) / (g - 1.0) * phi2 - T);
    M(A, 3, 1, 2) = m[0];
    M(A, 3, 2, 2) = U - (g - 2.0) * u[0] * m[0];
    M(A, 3, 3, 2) = (T + phi2 / (g - 1.0)) * m[0] - (g - 1.0) * U * u[0];
    M(A, 3, 1, 3) = 0.0;
    M(A, 3, 2, 3) = (g - 1.0) * m[0];
    M(A, 3, 3, 3) = g * U;
  }
  (void)q; (void)sv;
}

static void prim(int nd, const double* q, double g, double* sv, double* u, double* T) {
  /* default branches of the Jacobian routines: specific volume, velocity, temperature */
  *sv = 1.0 / q[0];
  double ke = 0.0;
  for (int i = 0; i < nd; ++i) u[i] = *sv * q[i + 1];
  for (int i = 0; i < nd; ++i) ke += u[i] * u[i];
  *T = g * (*sv * q[nd + 1] - 0.5 * ke);
}

void orc_inviscid_jacobian(int nd, const double* q, const double* metrics, double gamma,
                           double* A) {
  double sv, u[3], T;
  prim(nd, q, gamma, &sv, u, &T);
  orc_inviscid_jacobian_full(nd, q, metrics, gamma, sv, u, T, A);
}

void orc_inviscid_flux(int nd, const double* q, const double* m, double g, double* F) {
  /* contravariant Cartesian inviscid flux m . F(q) (computeCartesianInviscidFluxes :563-619 +
   * transformFluxes :772-840 for a general metric row) -- used by unit tests only. */
  double sv, u[3], T;
  prim(nd, q, g, &sv, u, &T);
  double ke = 0.0;
  for (int i = 0; i < nd; ++i) ke += u[i] * u[i];
  const double p = (g - 1.0) * (q[nd + 1] - 0.5 * q[0] * ke);
  double U = 0.0;
  for (int i = 0; i < nd; ++i) U += m[i] * u[i];
  F[0] = q[0] * U;
  for (int i = 0; i < nd; ++i) F[i + 1] = q[i + 1] * U + m[i] * p;
  F[nd + 1] = (q[nd + 1] + p) * U;
  (void)T;
}

void orc_incoming_jacobian(int nd, const double* q, const double* metrics, double g, int dirn,
                           double* A) {
  double sv, u[3], T;
  prim(nd, q, g, &sv, u, &T);
  const int n = nd + 2;
  double R[25], L[25], lam[5], nm[3];
  double arc = 0.0;
  for (int i = 0; i < nd; ++i) arc += metrics[i] * metrics[i];
  arc = sqrt(arc);
  for (int i = 0; i < nd; ++i) nm[i] = metrics[i] / arc;
  double U = 0.0, ke = 0.0;
  for (int i = 0; i < nd; ++i) U += nm[i] * u[i];
  for (int i = 0; i < nd; ++i) ke += u[i] * u[i];
  const double c = sqrt((g - 1.0) * T);
  const double phi2 = 0.5 * (g - 1.0) * ke;
  const double rho = q[0];
  if (nd == 2) {
    lam[0] = U; lam[1] = U; lam[2] = U + c; lam[3] = U - c;
    for (int i = 0; i < 4; ++i) lam[i] = arc * lam[i];
    for (int i = 0; i < 4; ++i) if ((double)dirn * lam[i] < 0.0) lam[i] = 0.0;
    M(R, 4, 1, 1) = 1.0; M(R, 4, 2, 1) = u[0]; M(R, 4, 3, 1) = u[1];
    M(R, 4, 4, 1) = phi2 / (g - 1.0);
    M(R, 4, 1, 2) = 0.0; M(R, 4, 2, 2) = nm[1] * rho; M(R, 4, 3, 2) = -nm[0] * rho;
    M(R, 4, 4, 2) = rho * (nm[1] * u[0] - nm[0] * u[1]);
    M(R, 4, 1, 3) = 1.0; M(R, 4, 2, 3) = u[0] + nm[0] * c; M(R, 4, 3, 3) = u[1] + nm[1] * c;
    M(R, 4, 4, 3) = T + phi2 / (g - 1.0) + c * U;
    M(R, 4, 1, 4) = 1.0; M(R, 4, 2, 4) = u[0] - nm[0] * c; M(R, 4, 3, 4) = u[1] - nm[1] * c;
    M(R, 4, 4, 4) = T + phi2 / (g - 1.0) - c * U;
    M(L, 4, 1, 1) = 1.0 - phi2 / (c * c);
    M(L, 4, 2, 1) = -sv * (nm[1] * u[0] - nm[0] * u[1]);
    M(L, 4, 3, 1) = 0.5 * (phi2 / (c * c) - U / c);
    M(L, 4, 4, 1) = 0.5 * (phi2 / (c * c) + U / c);
    M(L, 4, 1, 2) = u[0] / T; M(L, 4, 2, 2) = sv * nm[1];
    M(L, 4, 3, 2) = -0.5 * (u[0] / T - nm[0] / c);
    M(L, 4, 4, 2) = -0.5 * (u[0] / T + nm[0] / c);
    M(L, 4, 1, 3) = u[1] / T; M(L, 4, 2, 3) = -sv * nm[0];
    M(L, 4, 3, 3) = -0.5 * (u[1] / T - nm[1] / c);
    M(L, 4, 4, 3) = -0.5 * (u[1] / T + nm[1] / c);
    M(L, 4, 1, 4) = -1.0 / T; M(L, 4, 2, 4) = 0.0;
    M(L, 4, 3, 4) = 0.5 / T; M(L, 4, 4, 4) = 0.5 / T;
  } else if (nd == 3) {
    lam[0] = U; lam[1] = U; lam[2] = U; lam[3] = U + c; lam[4] = U - c;
    for (int i = 0; i < 5; ++i) lam[i] = arc * lam[i];
    for (int i = 0; i < 5; ++i) if ((double)dirn * lam[i] < 0.0) lam[i] = 0.0;
    const double h = phi2 / (g - 1.0);
    M(R, 5, 1, 1) = nm[0]; M(R, 5, 2, 1) = nm[0] * u[0];
    M(R, 5, 3, 1) = nm[0] * u[1] + rho * nm[2];
    M(R, 5, 4, 1) = nm[0] * u[2] - rho * nm[1];
    M(R, 5, 5, 1) = rho * (nm[2] * u[1] - nm[1] * u[2]) + h * nm[0];
    M(R, 5, 1, 2) = nm[1]; M(R, 5, 2, 2) = nm[1] * u[0] - rho * nm[2];
    M(R, 5, 3, 2) = nm[1] * u[1];
    M(R, 5, 4, 2) = nm[1] * u[2] + rho * nm[0];
    M(R, 5, 5, 2) = rho * (nm[0] * u[2] - nm[2] * u[0]) + h * nm[1];
    M(R, 5, 1, 3) = nm[2]; M(R, 5, 2, 3) = nm[2] * u[0] + rho * nm[1];
    M(R, 5, 3, 3) = nm[2] * u[1] - rho * nm[0];
    M(R, 5, 4, 3) = nm[2] * u[2];
    M(R, 5, 5, 3) = rho * (nm[1] * u[0] - nm[0] * u[1]) + h * nm[2];
    M(R, 5, 1, 4) = 1.0; M(R, 5, 2, 4) = u[0] + nm[0] * c; M(R, 5, 3, 4) = u[1] + nm[1] * c;
    M(R, 5, 4, 4) = u[2] + nm[2] * c; M(R, 5, 5, 4) = T + h + c * U;
    M(R, 5, 1, 5) = 1.0; M(R, 5, 2, 5) = u[0] - nm[0] * c; M(R, 5, 3, 5) = u[1] - nm[1] * c;
    M(R, 5, 4, 5) = u[2] - nm[2] * c; M(R, 5, 5, 5) = T + h - c * U;
    const double a = 1.0 - phi2 / (c * c);
    M(L, 5, 1, 1) = nm[0] * a - sv * (nm[2] * u[1] - nm[1] * u[2]);
    M(L, 5, 2, 1) = nm[1] * a - sv * (nm[0] * u[2] - nm[2] * u[0]);
    M(L, 5, 3, 1) = nm[2] * a - sv * (nm[1] * u[0] - nm[0] * u[1]);
    M(L, 5, 4, 1) = 0.5 * (phi2 / (c * c) - U / c);
    M(L, 5, 5, 1) = 0.5 * (phi2 / (c * c) + U / c);
    M(L, 5, 1, 2) = nm[0] * u[0] / T;
    M(L, 5, 2, 2) = nm[1] * u[0] / T - sv * nm[2];
    M(L, 5, 3, 2) = nm[2] * u[0] / T + sv * nm[1];
    M(L, 5, 4, 2) = -0.5 * (u[0] / T - nm[0] / c);
    M(L, 5, 5, 2) = -0.5 * (u[0] / T + nm[0] / c);
    M(L, 5, 1, 3) = nm[0] * u[1] / T + sv * nm[2];
    M(L, 5, 2, 3) = nm[1] * u[1] / T;
    M(L, 5, 3, 3) = nm[2] * u[1] / T - sv * nm[0];
    M(L, 5, 4, 3) = -0.5 * (u[1] / T - nm[1] / c);
    M(L, 5, 5, 3) = -0.5 * (u[1] / T + nm[1] / c);
    M(L, 5, 1, 4) = nm[0] * u[2] / T - sv * nm[1];
    M(L, 5, 2, 4) = nm[1] * u[2] / T + sv * nm[0];
    M(L, 5, 3, 4) = nm[2] * u[2] / T;
    M(L, 5, 4, 4) = -0.5 * (u[2] / T - nm[2] / c);
    M(L, 5, 5, 4) = -0.5 * (u[2] / T + nm[2] / c);
    M(L, 5, 1, 5) = -nm[0] / T; M(L, 5, 2, 5) = -nm[1] / T; M(L, 5, 3, 5) = -nm[2] / T;
    M(L, 5, 4, 5) = 0.5 / T; M(L, 5, 5, 5) = 0.5 / T;
  } else {
    memset(A, 0, sizeof(double) * 9);
    return;
  }
  for (int j = 1; j <= n; ++j)
    for (int i = 1; i <= n; ++i) {
      double s = M(R, n, i, 1) * lam[0] * M(L, n, 1, j);
      for (int k = 2; k <= n; ++k) s = s + M(R, n, i, k) * lam[k - 1] * M(L, n, k, j);
      M(A, n, i, j) = s;
    }
}

void orc_first_partial_viscous(int nd, const double* q, const double* m, const double* tau,
                               const double* hf, double pl, double g, double sv,
                               const double* u, double T, double* A) {
  const int n = nd + 2;
  double ke = 0.0;
  for (int i = 0; i < nd; ++i) ke += u[i] * u[i];
  const double phi2 = 0.5 * (g - 1.0) * ke;
  double cs[3], ch = 0.0;
  for (int i = 0; i < nd; ++i) {
    double s = 0.0;
    for (int j = 0; j < nd; ++j) s += m[j] * tau[i * nd + j];
    cs[i] = s;
  }
  for (int j = 0; j < nd; ++j) ch += m[j] * hf[j];
  double t1 = 0.0;
  for (int i = 0; i < nd; ++i) t1 += u[i] * cs[i];
  t1 = t1 - ch;
  double us = 0.0;
  for (int i = 0; i < nd; ++i) us += u[i] * cs[i];
  double t2 = pl * g * sv / T * (phi2 / (g - 1.0) - T / g);
  M(A, n, 1, 1) = 0.0;
  for (int i = 0; i < nd; ++i) M(A, n, i + 2, 1) = t2 * cs[i];
  M(A, n, n, 1) = t2 * t1 - sv * us;
  for (int k = 0; k < nd; ++k) {
    t2 = -pl * g * sv / T * u[k];
    M(A, n, 1, k + 2) = 0.0;
    for (int i = 0; i < nd; ++i) M(A, n, i + 2, k + 2) = t2 * cs[i];
    M(A, n, n, k + 2) = t2 * t1 + sv * cs[k];
  }
  t2 = pl * g * sv / T;
  M(A, n, 1, n) = 0.0;
  for (int i = 0; i < nd; ++i) M(A, n, i + 2, n) = t2 * cs[i];
  M(A, n, n, n) = t2 * t1;
  (void)q;
}

void orc_second_partial_viscous(int nd, const double* u, double mu, double lam, double kap,
                                double jac, const double* m1, const double* m2, double* B) {
  const int n = nd + 1;
  double t1 = 0.0, t2 = 0.0, t3 = 0.0;
  for (int i = 0; i < nd; ++i) t1 += m1[i] * m2[i];
  for (int i = 0; i < nd; ++i) t2 += m2[i] * u[i];
  t2 = mu * t2;
  for (int i = 0; i < nd; ++i) t3 += m1[i] * u[i];
  t3 = lam * t3;
  for (int j = 1; j <= nd; ++j) {
    for (int i = 1; i <= nd; ++i) {
      if (i == j)
        M(B, n, i, j) = mu * t1 + (mu + lam) * m1[i - 1] * m2[i - 1];
      else
        M(B, n, i, j) = mu * m1[j - 1] * m2[i - 1] + lam * m1[i - 1] * m2[j - 1];
    }
    M(B, n, n, j) = mu * t1 * u[j - 1] + m1[j - 1] * t2 + m2[j - 1] * t3;
  }
  for (int i = 1; i <= nd; ++i) M(B, n, i, n) = 0.0;
  M(B, n, n, n) = kap * t1;
  for (int k = 0; k < n * n; ++k) B[k] = jac * B[k];
}
