// device.h -- device-side helpers shared by the kernels: point geometry formed from
// the rectilinear 1-D grid arrays, SBP stencil application along a pencil, and the
// pointwise compressible-flow relations (operation order follows
// src/CNSHelperImpl.f90 so that results agree to rounding).
#pragma once
#include "mgd_internal.h"

namespace mgd {

// Grid indices of point p; 32-bit unsigned division (regions hold fewer than 2^31
// points per instance, checked at creation).
template <int ND>
__device__ __forceinline__ void unflat(const Geo& g, long p, int* ijk) {
  const unsigned pu = (unsigned)p;
  const unsigned n0 = (unsigned)g.n[0];
  ijk[0] = (int)(pu % n0);
  if (ND > 1) {
    const unsigned t = pu / n0;
    const unsigned n1 = (unsigned)g.n[1];
    ijk[1] = (int)(t % n1);
    if (ND > 2) ijk[2] = (int)(t / n1);
  }
}

// x / 3 and x / 6 rounded as IEEE division rounds them, in three FMA-unit operations:
// q = x * RN(1/y), remainder r = x - q y (exact with an FMA), q + r RN(1/y) (Markstein's
// correction step).  Checked against IEEE division on 3.8e8 random doubles of all
// exponents with no difference; it may differ only in the sign of an exact zero.
__device__ __forceinline__ double div3(double x) {
  constexpr double z = 1.0 / 3.0;
  const double q = x * z;
  return __fma_rn(__fma_rn(-q, 3.0, x), z, q);
}
__device__ __forceinline__ double div6(double x) {
  constexpr double z = 1.0 / 6.0;
  const double q = x * z;
  return __fma_rn(__fma_rn(-q, 6.0, x), z, q);
}

// metrics(:, (d-1)*nd + d) of a rectilinear grid (src/GridImpl.f90:795-1013)
template <int ND>
__device__ __forceinline__ double metric(const Geo& g, int d, const int* ijk) {
  if (ND == 1) return 1.0;
  if (ND == 2) return d == 0 ? g.h[1][ijk[1]] : g.h[0][ijk[0]];
  if (d == 0) return g.h[1][ijk[1]] * g.h[2][ijk[2]];
  if (d == 1) return g.h[0][ijk[0]] * g.h[2][ijk[2]];
  return g.h[0][ijk[0]] * g.h[1][ijk[1]];
}

// determinant of dx/dxi (before Grid%update inverts it)
template <int ND>
__device__ __forceinline__ double det(const Geo& g, const int* ijk) {
  if (ND == 1) return g.h[0][ijk[0]];
  if (ND == 2) return g.h[0][ijk[0]] * g.h[1][ijk[1]];
  return g.h[0][ijk[0]] * g.h[1][ijk[1]] * g.h[2][ijk[2]];
}

template <int ND>
__device__ __forceinline__ double jacobian(const Geo& g, const int* ijk) {
  return 1.0 / det<ND>(g, ijk);
}

template <int ND>
__device__ __forceinline__ double gnorm(const Geo& g, const int* ijk) {
  double s = 1.0;
  for (int d = 0; d < ND; ++d) s = g.hw[d][ijk[d]] * s;
  return s * det<ND>(g, ijk);
}

__device__ __forceinline__ int wrap(int i, const Dir& dd) {
  if (i < 0) return i + dd.len - dd.off;
  if (i >= dd.len) return i - dd.len + dd.off;
  return i;
}

// Applies operator `op` along a pencil at index i.  f(idx, v) writes the NC values of the
// operand at pencil index idx.  Boundary rows use the closure blocks on non-periodic
// directions (src/StencilOperatorImpl.f90:35-457).
template <int NC, class F>
__device__ __forceinline__ void sbp_apply(const Op& op, const Dir& dd, int i, F&& f, double* out) {
  double v[NC], w[NC];
  for (int c = 0; c < NC; ++c) out[c] = 0.0;
  if (dd.len == 1) return;
  if (!dd.periodic && i < op.bd) {
    for (int col = 0; col < op.bw; ++col) {
      const double b = op.b1[i][col];
      f(col, v);
      for (int c = 0; c < NC; ++c) out[c] += b * v[c];
    }
    return;
  }
  if (!dd.periodic && i >= dd.len - op.bd) {
    const int m = dd.len - 1 - i;
    const int base = dd.len - op.bw;
    for (int col = 0; col < op.bw; ++col) {
      const double b = op.b2[m][col];
      f(base + col, v);
      for (int c = 0; c < NC; ++c) out[c] += b * v[c];
    }
    return;
  }
  if (op.sym == 1) {
    for (int k = 1; k <= op.hw; ++k) {
      f(wrap(i + k, dd), v);
      f(wrap(i - k, dd), w);
      const double ck = op.c[kMaxHW + k];
      for (int c = 0; c < NC; ++c) out[c] += ck * (v[c] - w[c]);
    }
  } else {
    for (int k = 1; k <= op.hw; ++k) {
      f(wrap(i + k, dd), v);
      f(wrap(i - k, dd), w);
      const double ck = op.c[kMaxHW + k];
      for (int c = 0; c < NC; ++c) out[c] += ck * (v[c] + w[c]);
    }
    f(i, v);
    for (int c = 0; c < NC; ++c) out[c] = out[c] + op.c[kMaxHW] * v[c];
  }
}

// Dependent variables at one point (computeDependentVariables, :3-87).
template <int ND>
struct Prim {
  double sv, u[ND], p, T;
};

template <int ND>
__device__ __forceinline__ Prim<ND> prim(const double* q, double g) {
  Prim<ND> s;
  s.sv = 1.0 / q[0];
  double ke = 0.0;
  for (int i = 0; i < ND; ++i) {
    s.u[i] = s.sv * q[i + 1];
    ke += s.u[i] * s.u[i];
  }
  s.p = (g - 1.0) * (q[ND + 1] - 0.5 * q[0] * ke);
  s.T = g * s.p / (g - 1.0) * s.sv;
  return s;
}

// Cartesian inviscid flux along direction d (computeCartesianInviscidFluxes, :563-619).
template <int ND>
__device__ __forceinline__ void inviscid_flux(const double* q, const Prim<ND>& s, int d, double* f) {
  const double ud = s.u[d];
  f[0] = q[d + 1];
  for (int i = 0; i < ND; ++i) {
    if (i == d) f[i + 1] = q[d + 1] * ud + s.p;
    else if (i < d) f[i + 1] = q[i + 1] * ud;
    else f[i + 1] = q[d + 1] * s.u[i];
  }
  f[ND + 1] = ud * (q[ND + 1] + s.p);
}

// Transport variables without a viscosity power law (constant in the temperature).
__device__ __forceinline__ void transport_const(const Geo& g, double* mu, double* la, double* ka) {
  *mu = g.re_inv;
  *la = (g.bulk_ratio - 2.0 / 3.0) * g.re_inv;
  *ka = g.re_inv * g.pr_inv;
}

// Transport variables (computeTransportVariables, :89-177).
__device__ __forceinline__ void transport(const Geo& g, double T, double* mu, double* la,
                                          double* ka) {
  if (g.power_law <= 0.0) {
    transport_const(g, mu, la, ka);
  } else {
    *mu = pow((g.gamma - 1.0) * T, g.power_law) * g.re_inv;
    *la = (g.bulk_ratio - 2.0 / 3.0) * *mu;
    *ka = *mu * g.pr_inv;
  }
}

// Stress tensor from the velocity gradient, in place (computeStressTensor, :406-455);
// t[i*ND + j] holds du_i/dx_j on entry.
template <int ND>
__device__ __forceinline__ void stress(double* t, double mu, double la) {
  if (ND == 1) {
    t[0] = (2.0 * mu + la) * t[0];
  } else if (ND == 2) {
    const double dv = la * (t[0] + t[3]);
    t[0] = 2.0 * mu * t[0] + dv;
    t[1] = mu * (t[1] + t[2]);
    t[2] = t[1];
    t[3] = 2.0 * mu * t[3] + dv;
  } else {
    const double dv = la * (t[0] + t[4] + t[8]);
    t[0] = 2.0 * mu * t[0] + dv;
    t[1] = mu * (t[1] + t[3]);
    t[2] = mu * (t[2] + t[6]);
    t[3] = t[1];
    t[4] = 2.0 * mu * t[4] + dv;
    t[5] = mu * (t[5] + t[7]);
    t[6] = t[2];
    t[7] = t[5];
    t[8] = 2.0 * mu * t[8] + dv;
  }
}

// Cartesian viscous flux along direction d (computeCartesianViscousFluxes, :621-689).
template <int ND>
__device__ __forceinline__ void viscous_flux(const double* u, const double* tau, const double* hf,
                                             int d, double* f) {
  f[0] = 0.0;
  double e = 0.0;
  for (int i = 0; i < ND; ++i) {
    f[i + 1] = tau[i * ND + d];
    e += u[i] * tau[i * ND + d];
  }
  f[ND + 1] = e - hf[d];
}

// Inviscid flux Jacobian (computeJacobianOfInviscidFlux2D/3D, :1098-1444), column-major.
template <int ND>
__device__ __forceinline__ void inviscid_jacobian(const double* m, const Prim<ND>& s, double g,
                                                  double* A) {
  constexpr int n = ND + 2;
#define AJ(i, j) A[((j) - 1) * n + ((i) - 1)]
  double U = 0.0, ke = 0.0;
  for (int i = 0; i < ND; ++i) U += m[i] * s.u[i];
  for (int i = 0; i < ND; ++i) ke += s.u[i] * s.u[i];
  const double phi2 = 0.5 * (g - 1.0) * ke;
  const double T = s.T;
  AJ(1, 1) = 0.0;
  for (int i = 0; i < ND; ++i) AJ(i + 2, 1) = phi2 * m[i] - U * s.u[i];
  AJ(n, 1) = U * ((g - 2.0) / (g - 1.0) * phi2 - T);
  for (int k = 0; k < ND; ++k) {
    AJ(1, k + 2) = m[k];
    for (int i = 0; i < ND; ++i) {
      if (i == k) AJ(i + 2, k + 2) = U - (g - 2.0) * s.u[k] * m[k];
      else AJ(i + 2, k + 2) = s.u[i] * m[k] - (g - 1.0) * s.u[k] * m[i];
    }
    AJ(n, k + 2) = (T + phi2 / (g - 1.0)) * m[k] - (g - 1.0) * U * s.u[k];
  }
  AJ(1, n) = 0.0;
  for (int i = 0; i < ND; ++i) AJ(i + 2, n) = (g - 1.0) * m[i];
  AJ(n, n) = g * U;
#undef AJ
}

// First-partial viscous Jacobian (computeFirstPartialViscousJacobian2D/3D, :2412-2600).
template <int ND>
__device__ __forceinline__ void first_partial_viscous(const double* m, const double* tau,
                                                      const double* hf, double pl, double g,
                                                      const Prim<ND>& s, double* A) {
  constexpr int n = ND + 2;
#define AJ(i, j) A[((j) - 1) * n + ((i) - 1)]
  double ke = 0.0;
  for (int i = 0; i < ND; ++i) ke += s.u[i] * s.u[i];
  const double phi2 = 0.5 * (g - 1.0) * ke;
  double cs[ND], ch = 0.0;
  for (int i = 0; i < ND; ++i) {
    double a = 0.0;
    for (int j = 0; j < ND; ++j) a += m[j] * tau[i * ND + j];
    cs[i] = a;
  }
  for (int j = 0; j < ND; ++j) ch += m[j] * hf[j];
  double us = 0.0;
  for (int i = 0; i < ND; ++i) us += s.u[i] * cs[i];
  const double t1 = us - ch;
  const double sv = s.sv, T = s.T;
  double t2 = pl * g * sv / T * (phi2 / (g - 1.0) - T / g);
  AJ(1, 1) = 0.0;
  for (int i = 0; i < ND; ++i) AJ(i + 2, 1) = t2 * cs[i];
  AJ(n, 1) = t2 * t1 - sv * us;
  for (int k = 0; k < ND; ++k) {
    t2 = -pl * g * sv / T * s.u[k];
    AJ(1, k + 2) = 0.0;
    for (int i = 0; i < ND; ++i) AJ(i + 2, k + 2) = t2 * cs[i];
    AJ(n, k + 2) = t2 * t1 + sv * cs[k];
  }
  t2 = pl * g * sv / T;
  AJ(1, n) = 0.0;
  for (int i = 0; i < ND; ++i) AJ(i + 2, n) = t2 * cs[i];
  AJ(n, n) = t2 * t1;
#undef AJ
}

// Second-partial viscous Jacobian (computeSecondPartialViscousJacobian2D/3D, :2636-2756),
// (ND+1) x (ND+1) column-major, already multiplied by the grid Jacobian.
template <int ND>
__device__ __forceinline__ void second_partial_viscous(const double* u, double mu, double la,
                                                       double ka, double jac, const double* m1,
                                                       const double* m2, double* B) {
  constexpr int n = ND + 1;
  double t1 = 0.0, t2 = 0.0, t3 = 0.0;
  for (int i = 0; i < ND; ++i) t1 += m1[i] * m2[i];
  for (int i = 0; i < ND; ++i) t2 += m2[i] * u[i];
  t2 = mu * t2;
  for (int i = 0; i < ND; ++i) t3 += m1[i] * u[i];
  t3 = la * t3;
  for (int j = 0; j < ND; ++j) {
    for (int i = 0; i < ND; ++i) {
      if (i == j) B[j * n + i] = mu * t1 + (mu + la) * m1[i] * m2[i];
      else B[j * n + i] = mu * m1[j] * m2[i] + la * m1[i] * m2[j];
    }
    B[j * n + ND] = mu * t1 * u[j] + m1[j] * t2 + m2[j] * t3;
  }
  for (int i = 0; i < ND; ++i) B[ND * n + i] = 0.0;
  B[ND * n + ND] = ka * t1;
  for (int k = 0; k < n * n; ++k) B[k] = jac * B[k];
}

}  // namespace mgd
