// kernels.hip -- generic ("plain path") kernels of the magudi RHS on gfx950.
//
// One thread per grid point, instance in blockIdx.y.  Stencil operands are read
// through the cache hierarchy; every nd (1-3), inviscid and viscous, periodic and
// bounded directions are covered.  These are the reference-shaped building blocks;
// the fused 2-D kernels in fused2d.hip replace the hot ones for the benchmark path.
#include <cfloat>
#include <type_traits>

#include "device.h"
#include "kernels.h"

namespace mgd {

__device__ __forceinline__ long gidx() { return (long)blockIdx.x * blockDim.x + threadIdx.x; }

// ------------------------------------------------------------------ state helpers
template <int ND>
__device__ __forceinline__ void load_q(const double* q, long N, long p, double* v) {
  for (int l = 0; l < ND + 2; ++l) v[l] = q[l * N + p];
}

// Primitive variables needed by the viscous terms: w = (u_1..u_nd, T).
template <int ND>
__global__ void __launch_bounds__(kBlock) k_prim(Geo g, const double* __restrict__ q, double* __restrict__ w) {
  const long p = gidx();
  if (p >= g.N) return;
  const long inst = blockIdx.y;
  const double* qi = q + inst * g.nv * g.N;
  double* wi = w + inst * (ND + 1) * g.N;
  double v[ND + 2];
  load_q<ND>(qi, g.N, p, v);
  const Prim<ND> s = prim<ND>(v, g.gamma);
  for (int i = 0; i < ND; ++i) wi[i * g.N + p] = s.u[i];
  wi[ND * g.N + p] = s.T;
}

// Gradient of w along direction d: G[(c*ND + d)] = jacobian * metric_dd * D_d w_c
// (Grid%computeGradient, src/GridImpl.f90:1172-1421, non-curvilinear).
template <int ND>
__global__ void __launch_bounds__(kBlock) k_grad(Geo g, int d, const double* __restrict__ w, double* __restrict__ G) {
  const long p = gidx();
  if (p >= g.N) return;
  const long inst = blockIdx.y;
  const double* wi = w + inst * (ND + 1) * g.N;
  double* Gi = G + inst * (ND + 1) * ND * g.N;
  int ijk[3] = {0, 0, 0};
  unflat<ND>(g, p, ijk);
  const Dir dd = g.dir[d];
  const long base = p - (long)ijk[d] * dd.stride;
  double out[ND + 1];
  sbp_apply<ND + 1>(g.ops[SLOT_FIRST], dd, ijk[d],
                    [&](int idx, double* v) {
                      const long pp = base + (long)idx * dd.stride;
                      for (int c = 0; c <= ND; ++c) v[c] = wi[c * g.N + pp];
                    },
                    out);
  const double jm = jacobian<ND>(g, ijk) * metric<ND>(g, d, ijk);
  for (int c = 0; c <= ND; ++c) Gi[(c * ND + d) * g.N + p] = jm * out[c];
}

// Gradients of w = (u_1..u_3, T) on a fully periodic 3-D grid in one pass: a BX x BY
// column tile marches along z over a chunk of ZC planes.  Each thread keeps w of its
// column for the 9 planes of the z stencil in registers; the plane being
// differentiated is staged with a 4-point x / y halo in LDS, one halo point per thread
// (threads below 8 (BX + BY)).  The halo point of the next plane (and, where the
// registers allow, the column's next lead point) is loaded one iteration before it is
// used, so its latency overlaps a plane's stencil work.
// Same sums as k_prim + k_grad (interior SBP rows, wrap-around indices), so the same
// values.
// STRESS: instead of the 12 gradient components, write what the fluxes need of them at
// the point: the stress tensor (computeStressTensor, src/CNSHelperImpl.f90:406-455) as
// its six distinct components (xx, xy, xz, yy, yz, zz) and the heat flux -kappa grad T
// (src/StateImpl.f90:513-523) in slots 6..8, from the same gradient values in registers
// (visc_point's operations, so the same values); 72 instead of 96 bytes per point.
// POWER: the viscosity follows a power law of the temperature (pow per point); without
// it the transport coefficients are constants and the kernel needs far fewer registers.
template <int BX, int BY, int ZC, bool STRESS = false, bool POWER = true>
__global__ void __launch_bounds__(BX * BY, STRESS ? 2 : 4) k_grad3d_periodic(Geo g, const double* __restrict__ q,
                                                              double* __restrict__ G) {
  constexpr int NT = BX * BY, EXW = BX + 8, EYW = BY + 8, NW = 4, NH = 8 * (BX + BY);
  static_assert(NH <= NT, "one halo point per thread");
  __shared__ double sw[NW][EYW][EXW];
  // XCD-aware block map (as in k_fwd3d_march): block b takes (tile, z chunk) unit
  // (b mod 8) * ceil(units / 8) + b / 8, tiles x-fastest, so each XCD works on one
  // contiguous band and neighbouring tiles' halo lines meet in its L2
  const int ntx = g.n[0] / BX, ntiles = ntx * (g.n[1] / BY);
  const int nunits = ntiles * ((g.n[2] + ZC - 1) / ZC);
  const int per_xcd = (nunits + 7) / 8;
  const int unit = (int)(blockIdx.x % 8) * per_xcd + (int)(blockIdx.x / 8);
  if (unit >= nunits) return;   // padding of the grid to a multiple of 8 (whole block)
  const int tile = unit % ntiles;
  const int x0 = (tile % ntx) * BX, y0 = (tile / ntx) * BY;
  const int z0 = (unit / ntiles) * ZC;
  const long inst = blockIdx.z;
  const double* qi = q + inst * 5 * g.N;
  double* Gi = G + inst * 4 * 3 * g.N;
  const int tx = threadIdx.x % BX, ty = threadIdx.x / BX;
  const int i = x0 + tx, j = y0 + ty;
  const Dir dx = g.dir[0], dy = g.dir[1], dz = g.dir[2];
  const long n01 = (long)g.n[0] * g.n[1];
  auto load_v = [&](int ii, int jj, int kk, double* v) {
    const long p = ii + (long)g.n[0] * jj + n01 * kk;
#pragma unroll
    for (int l = 0; l < 5; ++l) v[l] = qi[l * g.N + p];
  };
  auto to_w = [&](const double* v, double* w) {
    const Prim<3> s = prim<3>(v, g.gamma);
    w[0] = s.u[0]; w[1] = s.u[1]; w[2] = s.u[2]; w[3] = s.T;
  };
  // this thread's halo point: y halo rows over the tile columns, then x halo columns over
  // the tile rows (the corners are never read)
  const int t = threadIdx.x;
  const bool has_halo = t < NH;
  int hx = 0, hy = 0;
  if (t < 8 * BX) {
    const int r = t / BX;
    hy = r < 4 ? r : BY + r;
    hx = 4 + t % BX;
  } else if (has_halo) {
    const int t2 = t - 8 * BX, c = t2 % 8;
    hy = 4 + t2 / 8;
    hx = c < 4 ? c : BX + c;
  }
  const int hi = wrap(x0 + hx - 4, dx), hj = wrap(y0 + hy - 4, dy);
  double Q[9][NW];
  double v[5];
#pragma unroll
  for (int k = 0; k < 8; ++k) {
    load_v(i, j, wrap(z0 - 4 + k, dz), v);
    to_w(v, Q[k]);
  }
  double hv[5];
  if (has_halo) load_v(hi, hj, z0 % dz.len, hv);   // halo, plane z0
  // with constant transport coefficients the registers allow the column's lead point to
  // be loaded a plane ahead too
  constexpr bool kAhead = STRESS && !POWER;
  if (kAhead) load_v(i, j, wrap(z0 + 4, dz), v);   // the column's lead plane z0 + 4
  const Op& op = g.ops[SLOT_FIRST];
  const double c1 = op.c[kMaxHW + 1], c2 = op.c[kMaxHW + 2], c3 = op.c[kMaxHW + 3], c4 = op.c[kMaxHW + 4];
  for (int s = 0; s < ZC; ++s) {
    const int zc = z0 + s;
    if (zc >= g.n[2]) break;   // uniform over the block
    if (!kAhead) load_v(i, j, wrap(zc + 4, dz), v);
    to_w(v, Q[8]);
    double hw[NW];
    if (has_halo) to_w(hv, hw);
    // the next plane's halo point (and lead point): the loads stay in flight through this
    // plane's work
    if (s + 1 < ZC && zc + 1 < g.n[2]) {
      if (kAhead) load_v(i, j, wrap(zc + 5, dz), v);
      if (has_halo) load_v(hi, hj, zc + 1, hv);
    }
    __syncthreads();            // readers of the previous plane are done
#pragma unroll
    for (int c = 0; c < NW; ++c) sw[c][ty + 4][tx + 4] = Q[4][c];
    if (has_halo) {
#pragma unroll
      for (int c = 0; c < NW; ++c) sw[c][hy][hx] = hw[c];
    }
    __syncthreads();
    int ijk[3] = {i, j, zc};
    const long p = i + (long)g.n[0] * j + n01 * zc;
    const double jac = jacobian<3>(g, ijk);
    double gv[NW * 3];
#pragma unroll
    for (int c = 0; c < NW; ++c) {
      const double* row = &sw[c][ty + 4][tx + 4];
      double ox = 0.0, oy = 0.0, oz = 0.0;
      ox += c1 * (row[1] - row[-1]);
      ox += c2 * (row[2] - row[-2]);
      ox += c3 * (row[3] - row[-3]);
      ox += c4 * (row[4] - row[-4]);
      oy += c1 * (row[EXW] - row[-EXW]);
      oy += c2 * (row[2 * EXW] - row[-2 * EXW]);
      oy += c3 * (row[3 * EXW] - row[-3 * EXW]);
      oy += c4 * (row[4 * EXW] - row[-4 * EXW]);
      oz += c1 * (Q[5][c] - Q[3][c]);
      oz += c2 * (Q[6][c] - Q[2][c]);
      oz += c3 * (Q[7][c] - Q[1][c]);
      oz += c4 * (Q[8][c] - Q[0][c]);
      gv[c * 3 + 0] = (jac * metric<3>(g, 0, ijk)) * ox;
      gv[c * 3 + 1] = (jac * metric<3>(g, 1, ijk)) * oy;
      gv[c * 3 + 2] = (jac * metric<3>(g, 2, ijk)) * oz;
      if constexpr (!STRESS) {
#pragma unroll
        for (int d = 0; d < 3; ++d) Gi[(c * 3 + d) * g.N + p] = gv[c * 3 + d];
      }
    }
    if constexpr (STRESS) {
      double mu, la, ka;
      if constexpr (POWER) transport(g, Q[4][3], &mu, &la, &ka);
      else transport_const(g, &mu, &la, &ka);
      stress<3>(gv, mu, la);
      Gi[0 * g.N + p] = gv[0];
      Gi[1 * g.N + p] = gv[1];
      Gi[2 * g.N + p] = gv[2];
      Gi[3 * g.N + p] = gv[4];
      Gi[4 * g.N + p] = gv[5];
      Gi[5 * g.N + p] = gv[8];
#pragma unroll
      for (int d = 0; d < 3; ++d) Gi[(6 + d) * g.N + p] = -ka * gv[9 + d];
    }
#pragma unroll
    for (int k = 0; k < 8; ++k)
#pragma unroll
      for (int c = 0; c < NW; ++c) Q[k][c] = Q[k + 1][c];
  }
}

// Stress tensor and heat flux at one point from the stored gradients.
template <int ND>
__device__ __forceinline__ void visc_point(const Geo& g, const double* Gi, long N, long p,
                                           const Prim<ND>& s, double* tau, double* hf,
                                           double* mu, double* la, double* ka) {
  transport(g, s.T, mu, la, ka);
  for (int i = 0; i < ND; ++i)
    for (int j = 0; j < ND; ++j) tau[i * ND + j] = Gi[(i * ND + j) * N + p];
  stress<ND>(tau, *mu, *la);
  for (int j = 0; j < ND; ++j) hf[j] = -*ka * Gi[(ND * ND + j) * N + p];
}

// Contravariant total flux of direction d at point pp (inviscid minus viscous).
template <int ND>
__device__ __forceinline__ void flux_hat(const Geo& g, const double* qi, const double* Gi, long pp,
                                         int d, double* f) {
  double v[ND + 2];
  load_q<ND>(qi, g.N, pp, v);
  const Prim<ND> s = prim<ND>(v, g.gamma);
  inviscid_flux<ND>(v, s, d, f);
  if (g.viscous) {
    double tau[ND * ND], hf[ND], mu, la, ka, fv[ND + 2];
    visc_point<ND>(g, Gi, g.N, pp, s, tau, hf, &mu, &la, &ka);
    viscous_flux<ND>(s.u, tau, hf, d, fv);
    for (int l = 0; l < ND + 2; ++l) f[l] = f[l] - fv[l];
  }
  int ijk[3] = {0, 0, 0};
  unflat<ND>(g, pp, ijk);
  const double m = metric<ND>(g, d, ijk);
  for (int l = 0; l < ND + 2; ++l) f[l] = m * f[l];
}

// rhs accumulation of -sum_d D_d Fhat_d (computeRhsForward, src/RhsHelperImpl.f90:303-344):
// d == 0 stores D_0 F_0; later directions add; the last one forms 0 - sum.
template <int ND>
__global__ void __launch_bounds__(kBlock) k_flux_div(Geo g, int d, const double* __restrict__ q, const double* __restrict__ G,
                           double* __restrict__ rhs) {
  const long p = gidx();
  if (p >= g.N) return;
  const long inst = blockIdx.y;
  const double* qi = q + inst * g.nv * g.N;
  const double* Gi = G ? G + inst * (ND + 1) * ND * g.N : nullptr;
  double* ri = rhs + inst * g.nv * g.N;
  int ijk[3] = {0, 0, 0};
  unflat<ND>(g, p, ijk);
  const Dir dd = g.dir[d];
  const long base = p - (long)ijk[d] * dd.stride;
  double out[ND + 2];
  sbp_apply<ND + 2>(g.ops[SLOT_FIRST], dd, ijk[d],
                    [&](int idx, double* f) { flux_hat<ND>(g, qi, Gi, base + (long)idx * dd.stride, d, f); },
                    out);
  for (int l = 0; l < ND + 2; ++l) {
    double s = d == 0 ? out[l] : ri[l * g.N + p] + out[l];
    if (d == ND - 1) s = 0.0 - s;
    ri[l * g.N + p] = s;
  }
}

// Contravariant total fluxes of every direction at each point, computed once:
// F[(d*NV + l)] = Fhat_d,l (the same operations as flux_hat, so the same values).
template <int ND>
__global__ void __launch_bounds__(kBlock) k_flux_all(Geo g, const double* __restrict__ q, const double* __restrict__ G,
                           double* __restrict__ F) {
  const long p = gidx();
  if (p >= g.N) return;
  const long inst = blockIdx.y;
  constexpr int NV = ND + 2;
  const double* qi = q + inst * NV * g.N;
  const double* Gi = G ? G + inst * (ND + 1) * ND * g.N : nullptr;
  double* Fi = F + inst * ND * NV * g.N;
  for (int d = 0; d < ND; ++d) {
    double f[NV];
    flux_hat<ND>(g, qi, Gi, p, d, f);
    for (int l = 0; l < NV; ++l) Fi[(d * NV + l) * g.N + p] = f[l];
  }
}

// rhs accumulation of -sum_d D_d Fhat_d from the stored fluxes (same sums as k_flux_div).
template <int ND>
__global__ void __launch_bounds__(kBlock) k_div(Geo g, int d, const double* __restrict__ F, double* __restrict__ rhs) {
  const long p = gidx();
  if (p >= g.N) return;
  const long inst = blockIdx.y;
  constexpr int NV = ND + 2;
  const double* Fd = F + (inst * ND + d) * NV * g.N;
  double* ri = rhs + inst * NV * g.N;
  int ijk[3] = {0, 0, 0};
  unflat<ND>(g, p, ijk);
  const Dir dd = g.dir[d];
  const long base = p - (long)ijk[d] * dd.stride;
  double out[NV];
  sbp_apply<NV>(g.ops[SLOT_FIRST], dd, ijk[d],
                [&](int idx, double* f) {
                  const long pp = base + (long)idx * dd.stride;
                  for (int l = 0; l < NV; ++l) f[l] = Fd[l * g.N + pp];
                },
                out);
  for (int l = 0; l < NV; ++l) {
    double s = d == 0 ? out[l] : ri[l * g.N + p] + out[l];
    if (d == ND - 1) s = 0.0 - s;
    ri[l * g.N + p] = s;
  }
}

// Dissipation, first half: t = D4_d x, scaled by -arcLength_d unless composite
// (addDissipation, src/RhsHelperImpl.f90:58-77).  With composite dissipation the
// term is added to rhs directly.
template <int ND>
__global__ void __launch_bounds__(kBlock) k_diss_a(Geo g, int d, const double* __restrict__ x, double* __restrict__ t,
                         double* __restrict__ rhs, double amount) {
  const long p = gidx();
  if (p >= g.N) return;
  const long inst = blockIdx.y;
  const double* xi = x + inst * g.nv * g.N;
  int ijk[3] = {0, 0, 0};
  unflat<ND>(g, p, ijk);
  const Dir dd = g.dir[d];
  const long base = p - (long)ijk[d] * dd.stride;
  double out[ND + 2];
  sbp_apply<ND + 2>(g.ops[SLOT_DISS], dd, ijk[d],
                    [&](int idx, double* v) { load_q<ND>(xi, g.N, base + (long)idx * dd.stride, v); },
                    out);
  if (g.composite) {
    double* ri = rhs + inst * g.nv * g.N;
    for (int l = 0; l < ND + 2; ++l) ri[l * g.N + p] = ri[l * g.N + p] + amount * out[l];
    return;
  }
  const double arc = fabs(metric<ND>(g, d, ijk));
  double* ti = t + inst * g.nv * g.N;
  for (int l = 0; l < ND + 2; ++l) ti[l * g.N + p] = -arc * out[l];
}

// Dissipation, second half: rhs += amount * H_d^-1 D4_d^T t.
template <int ND>
__global__ void __launch_bounds__(kBlock) k_diss_b(Geo g, int d, const double* __restrict__ t, double* __restrict__ rhs,
                         double amount) {
  const long p = gidx();
  if (p >= g.N) return;
  const long inst = blockIdx.y;
  const double* ti = t + inst * g.nv * g.N;
  double* ri = rhs + inst * g.nv * g.N;
  int ijk[3] = {0, 0, 0};
  unflat<ND>(g, p, ijk);
  const Dir dd = g.dir[d];
  const long base = p - (long)ijk[d] * dd.stride;
  double out[ND + 2];
  sbp_apply<ND + 2>(g.ops[SLOT_DISS_T], dd, ijk[d],
                    [&](int idx, double* v) { load_q<ND>(ti, g.N, base + (long)idx * dd.stride, v); },
                    out);
  const double h = g.hw[d][ijk[d]];
  for (int l = 0; l < ND + 2; ++l) {
    const double tl = out[l] / h;
    ri[l * g.N + p] = ri[l * g.N + p] + amount * tl;
  }
}

// Adjoint: T1_d = D+_d lambda for all nv components (src/RhsHelperImpl.f90:410-415);
// with comp0 = 1 and nc = nv-1 it applies to the viscous diffusion block instead.
template <int ND>
__global__ void __launch_bounds__(kBlock) k_adj_deriv(Geo g, int d, const double* __restrict__ x, long xstride_inst,
                            double* __restrict__ T1, int comp0) {
  const long p = gidx();
  if (p >= g.N) return;
  const long inst = blockIdx.y;
  int ijk[3] = {0, 0, 0};
  unflat<ND>(g, p, ijk);
  const Dir dd = g.dir[d];
  const long base = p - (long)ijk[d] * dd.stride;
  constexpr int NV = ND + 2;
  const double* xi = x + inst * xstride_inst;
  double* Ti = T1 + (inst * ND + d) * NV * g.N;
  if (comp0 == 0) {
    double out[NV];
    sbp_apply<NV>(g.ops[SLOT_ADJ], dd, ijk[d],
                  [&](int idx, double* v) {
                    const long pp = base + (long)idx * dd.stride;
                    for (int l = 0; l < NV; ++l) v[l] = xi[l * g.N + pp];
                  },
                  out);
    for (int l = 0; l < NV; ++l) Ti[l * g.N + p] = out[l];
  } else {
    // operand is T1_d itself (components 1..nv-1), written to the scratch slot x
    double out[NV - 1];
    const double* src = xi + (long)d * NV * g.N;
    sbp_apply<NV - 1>(g.ops[SLOT_ADJ], dd, ijk[d],
                      [&](int idx, double* v) {
                        const long pp = base + (long)idx * dd.stride;
                        for (int l = 0; l < NV - 1; ++l) v[l] = src[(l + 1) * g.N + pp];
                      },
                      out);
    for (int l = 0; l < NV - 1; ++l) Ti[(l + 1) * g.N + p] = out[l];
  }
}

// Adjoint: rhs = sum_d (A_d - Av_d)^T T1_d at each point (src/RhsHelperImpl.f90:431-491).
template <int ND>
__global__ void __launch_bounds__(kBlock) k_adj_jac(Geo g, const double* __restrict__ q, const double* __restrict__ G,
                          const double* __restrict__ T1, double* __restrict__ rhs) {
  const long p = gidx();
  if (p >= g.N) return;
  const long inst = blockIdx.y;
  constexpr int NV = ND + 2;
  const double* qi = q + inst * NV * g.N;
  const double* Ti = T1 + inst * ND * NV * g.N;
  double* ri = rhs + inst * NV * g.N;
  double v[NV];
  load_q<ND>(qi, g.N, p, v);
  const Prim<ND> s = prim<ND>(v, g.gamma);
  int ijk[3] = {0, 0, 0};
  unflat<ND>(g, p, ijk);
  double tau[ND * ND], hf[ND], mu = 0, la = 0, ka = 0;
  if (g.viscous) visc_point<ND>(g, G + inst * (ND + 1) * ND * g.N, g.N, p, s, tau, hf, &mu, &la, &ka);
  double acc[NV];
  for (int l = 0; l < NV; ++l) acc[l] = 0.0;
  for (int d = 0; d < ND; ++d) {
    double m[ND], A[NV * NV];
    for (int i = 0; i < ND; ++i) m[i] = i == d ? metric<ND>(g, d, ijk) : 0.0;
    inviscid_jacobian<ND>(m, s, g.gamma, A);
    if (g.viscous) {
      double B[NV * NV];
      first_partial_viscous<ND>(m, tau, hf, g.power_law, g.gamma, s, B);
      for (int k = 0; k < NV * NV; ++k) A[k] = A[k] - B[k];
    }
    for (int i = 0; i < NV; ++i) {
      double t = 0.0;
      for (int k = 0; k < NV; ++k) t += A[i * NV + k] * Ti[(d * NV + k) * g.N + p];
      acc[i] = acc[i] + t;
    }
  }
  for (int l = 0; l < NV; ++l) ri[l * g.N + p] = acc[l];
}

// The viscous adjoint Jacobian products at one point of a rectilinear grid, one direction
// at a time: AdjPoint holds what they share (the dependent and transport variables, the
// stress tensor and heat flux from the gradient gv of (u_1..u_ND, T) -- gv[c * ND + d] =
// d w_c / d x_d, as k_grad writes it -- and the direction-free scalars of the flux
// Jacobians); adj_visc_dir adds direction d's part for t = T1_d = D-dagger_d lambda:
//   r += (A_d - Av_d)^T t                       (src/RhsHelperImpl.f90:431-491)
//   diff[j] += B_dj^T t (components 1..ND+1)    (src/RhsHelperImpl.f90:509-553)
// Called for d = 0, 1, .. in order, the sums are formed in the order of the reference's
// loops (rhs over d; diff[j] over i for each j).  The metric of a rectilinear grid in
// direction d is h_d e_d: the Jacobian entries (computeJacobianOfInviscidFlux,
// computeFirstPartialViscousJacobian, computeSecondPartialViscousJacobian,
// src/CNSHelperImpl.f90:1098-1444, 2412-2756) are formed one column at a time with the
// products by the metric's zero entries -- exact zeros -- left out, which changes no
// value (at most the sign of an exact zero) and keeps no matrix in registers.  POWER:
// the viscosity power law is not zero (the terms in it are exact zeros otherwise).
template <int ND>
struct AdjPoint {
  Prim<ND> s;
  double tau[ND * ND], hf[ND], mu, la, ka, phi2, jac, h[ND];
  double t2_0, t2_n, t2_k[ND];
};

template <int ND, bool POWER>
__device__ __forceinline__ void adj_point_init(const Geo& g, const double* v, const double* gv,
                                               const int* ijk, AdjPoint<ND>& P) {
  const double gm = g.gamma;
  P.s = prim<ND>(v, gm);
  transport(g, P.s.T, &P.mu, &P.la, &P.ka);
  for (int i = 0; i < ND * ND; ++i) P.tau[i] = gv[i];
  stress<ND>(P.tau, P.mu, P.la);
  for (int j = 0; j < ND; ++j) P.hf[j] = -P.ka * gv[ND * ND + j];
  double ke = 0.0;
  for (int i = 0; i < ND; ++i) ke += P.s.u[i] * P.s.u[i];
  P.phi2 = 0.5 * (gm - 1.0) * ke;
  P.jac = jacobian<ND>(g, ijk);
  for (int d = 0; d < ND; ++d) P.h[d] = metric<ND>(g, d, ijk);
  if (POWER) {
    const double pl = g.power_law, sv = P.s.sv, T = P.s.T;
    P.t2_0 = pl * gm * sv / T * (P.phi2 / (gm - 1.0) - T / gm);
    P.t2_n = pl * gm * sv / T;
    for (int k = 0; k < ND; ++k) P.t2_k[k] = -pl * gm * sv / T * P.s.u[k];
  }
}

// adj_point_init from the stress / heat-flux slots at the point (sv: xx xy xz yy yz zz,
// then the heat flux x y z, as k_grad3d_periodic<.., true> writes them -- stress() and
// -kappa grad T of the same gradient values, so the same tau and hf).
template <bool POWER>
__device__ __forceinline__ void adj_point_init_stress(const Geo& g, const double* v, const double* sv,
                                                      const int* ijk, AdjPoint<3>& P) {
  const double gm = g.gamma;
  P.s = prim<3>(v, gm);
  transport(g, P.s.T, &P.mu, &P.la, &P.ka);
  P.tau[0] = sv[0]; P.tau[1] = sv[1]; P.tau[2] = sv[2];
  P.tau[3] = sv[1]; P.tau[4] = sv[3]; P.tau[5] = sv[4];
  P.tau[6] = sv[2]; P.tau[7] = sv[4]; P.tau[8] = sv[5];
  P.hf[0] = sv[6]; P.hf[1] = sv[7]; P.hf[2] = sv[8];
  double ke = 0.0;
  for (int i = 0; i < 3; ++i) ke += P.s.u[i] * P.s.u[i];
  P.phi2 = 0.5 * (gm - 1.0) * ke;
  P.jac = jacobian<3>(g, ijk);
  for (int d = 0; d < 3; ++d) P.h[d] = metric<3>(g, d, ijk);
  if (POWER) {
    const double pl = g.power_law, svol = P.s.sv, T = P.s.T;
    P.t2_0 = pl * gm * svol / T * (P.phi2 / (gm - 1.0) - T / gm);
    P.t2_n = pl * gm * svol / T;
    for (int k = 0; k < 3; ++k) P.t2_k[k] = -pl * gm * svol / T * P.s.u[k];
  }
}

template <int ND, bool POWER>
__device__ __forceinline__ void adj_visc_dir(const AdjPoint<ND>& P, double gm, int d,
                                             const double* t, double* r, double (*diff)[ND + 1]) {
  constexpr int NV = ND + 2, NW = ND + 1;
  const Prim<ND>& s = P.s;
  const double h = P.h[d], T = s.T, sv = s.sv, phi2 = P.phi2;
  const double U = h * s.u[d];
  double cs[ND];
  for (int i = 0; i < ND; ++i) cs[i] = h * P.tau[i * ND + d];
  const double ch = h * P.hf[d];
  double us = 0.0;
  for (int i = 0; i < ND; ++i) us += s.u[i] * cs[i];
  const double t1v = us - ch;
  // (A_d - Av_d)^T t, column 0 (density)
  {
    double acc = 0.0;
    for (int i = 0; i < ND; ++i) {
      double a = i == d ? phi2 * h - U * s.u[i] : 0.0 - U * s.u[i];
      if (POWER) a = a - P.t2_0 * cs[i];
      acc += a * t[i + 1];
    }
    const double bn = POWER ? P.t2_0 * t1v - sv * us : 0.0 - sv * us;
    acc += (U * ((gm - 2.0) / (gm - 1.0) * phi2 - T) - bn) * t[NV - 1];
    r[0] = r[0] + acc;
  }
  // columns 1..ND (momentum)
  for (int k = 0; k < ND; ++k) {
    double acc = 0.0;
    if (k == d) acc += h * t[0];
    for (int i = 0; i < ND; ++i) {
      double a;
      if (i == k) a = k == d ? U - (gm - 2.0) * s.u[k] * h : U;
      else if (k == d) a = s.u[i] * h;
      else if (i == d) a = 0.0 - (gm - 1.0) * s.u[k] * h;
      else a = 0.0;
      if (POWER) a = a - P.t2_k[k] * cs[i];
      else if (a == 0.0 && i != k && k != d && i != d) continue;
      acc += a * t[i + 1];
    }
    const double ae = k == d ? (T + phi2 / (gm - 1.0)) * h - (gm - 1.0) * U * s.u[k]
                             : 0.0 - (gm - 1.0) * U * s.u[k];
    const double be = POWER ? P.t2_k[k] * t1v + sv * cs[k] : sv * cs[k];
    acc += (ae - be) * t[NV - 1];
    r[k + 1] = r[k + 1] + acc;
  }
  // column ND+1 (energy)
  {
    double acc = 0.0;
    for (int i = 0; i < ND; ++i) {
      double a = i == d ? (gm - 1.0) * h : 0.0;
      if (POWER) a = a - P.t2_n * cs[i];
      else if (i != d) continue;
      acc += a * t[i + 1];
    }
    const double ae = gm * U;
    acc += (POWER ? ae - P.t2_n * t1v : ae) * t[NV - 1];
    r[NV - 1] = r[NV - 1] + acc;
  }
  // second-partial block: B_dj^T t for every j (m1 = h_d e_d, m2 = h_j e_j), times jac
  const double mu = P.mu, la = P.la, ka = P.ka, jac = P.jac;
  for (int j = 0; j < ND; ++j) {
    const double hj = P.h[j];
    if (j == d) {
      const double t1s = h * h;
      const double t2s = mu * (h * s.u[d]);
      const double t3s = la * (h * s.u[d]);
      for (int a = 0; a < ND; ++a) {
        const double dg = a == d ? mu * t1s + (mu + la) * h * h : mu * t1s;
        const double en = a == d ? mu * t1s * s.u[a] + h * t2s + h * t3s : mu * t1s * s.u[a];
        double x = 0.0;
        x += (jac * dg) * t[1 + a];
        x += (jac * en) * t[1 + ND];
        diff[j][a] = diff[j][a] + x;
      }
      diff[j][ND] = diff[j][ND] + (jac * (ka * t1s)) * t[1 + ND];
    } else {
      const double t2s = mu * (hj * s.u[j]);
      const double t3s = la * (h * s.u[d]);
      // column a = d: rows j and ND; column a = j: rows d and ND
      {
        double x = 0.0;
        x += (jac * (mu * h * hj)) * t[1 + j];
        x += (jac * (h * t2s)) * t[1 + ND];
        diff[j][d] = diff[j][d] + x;
      }
      {
        double x = 0.0;
        x += (jac * (la * h * hj)) * t[1 + d];
        x += (jac * (hj * t3s)) * t[1 + ND];
        diff[j][j] = diff[j][j] + x;
      }
    }
  }
}

// All directions at one point: r = sum_d (A_d - Av_d)^T t1_d and diff[j] = sum_i B_ij^T t1_i.
template <int ND>
__device__ __forceinline__ void adj_visc_point(const Geo& g, const double* v, const double* gv,
                                               const int* ijk, const double (*t1)[ND + 2],
                                               double* r, double (*diff)[ND + 1]) {
  for (int l = 0; l < ND + 2; ++l) r[l] = 0.0;
  for (int j = 0; j < ND; ++j)
    for (int a = 0; a <= ND; ++a) diff[j][a] = 0.0;
  AdjPoint<ND> P;
  if (g.power_law != 0.0) {
    adj_point_init<ND, true>(g, v, gv, ijk, P);
    for (int d = 0; d < ND; ++d) adj_visc_dir<ND, true>(P, g.gamma, d, t1[d], r, diff);
  } else {
    adj_point_init<ND, false>(g, v, gv, ijk, P);
    for (int d = 0; d < ND; ++d) adj_visc_dir<ND, false>(P, g.gamma, d, t1[d], r, diff);
  }
}

// Viscous regions: k_adj_jac's rhs and the second-partial block at one point, T1 and q
// read once.  The rhs is written, then components 1..nv-1 of every T1_j are overwritten
// with sum_i B_ij^T T1_i -- the Jacobian part in the same operations as k_adj_jac, so the
// same values.
template <int ND>
__global__ void __launch_bounds__(kBlock) k_adj_jac_visc2(Geo g, const double* __restrict__ q,
                                                         const double* __restrict__ G,
                                                         double* __restrict__ T1,
                                                         double* __restrict__ rhs) {
  const long p = gidx();
  if (p >= g.N) return;
  const long inst = blockIdx.y;
  constexpr int NV = ND + 2, NW = ND + 1;
  const double* qi = q + inst * NV * g.N;
  const double* Gi = G + inst * NW * ND * g.N;
  double* Ti = T1 + inst * ND * NV * g.N;
  double* ri = rhs + inst * NV * g.N;
  double v[NV], gv[NW * ND];
  load_q<ND>(qi, g.N, p, v);
  for (int k = 0; k < NW * ND; ++k) gv[k] = Gi[k * g.N + p];
  int ijk[3] = {0, 0, 0};
  unflat<ND>(g, p, ijk);
  double t1[ND][NV];
  for (int d = 0; d < ND; ++d)
    for (int k = 0; k < NV; ++k) t1[d][k] = Ti[(d * NV + k) * g.N + p];
  double r[NV], diff[ND][NW];
  adj_visc_point<ND>(g, v, gv, ijk, t1, r, diff);
  for (int l = 0; l < NV; ++l) ri[l * g.N + p] = r[l];
  for (int j = 0; j < ND; ++j)
    for (int a = 0; a < NW; ++a) Ti[(j * NV + 1 + a) * g.N + p] = diff[j][a];
}

// Viscous adjoint chain rule at one point (src/RhsHelperImpl.f90:555-570 with sign -1
// and :227-236 with sign +1): t2 = sum_j S_j (components 1..nv-1), applied to r;
// adj_chain_t2 takes t2 formed, v the conserved state at the point.
template <int ND>
__device__ __forceinline__ void adj_chain_t2(const Geo& g, const double* v, double* t2, double sign,
                                             double* r) {
  constexpr int NV = ND + 2;
  const Prim<ND> s = prim<ND>(v, g.gamma);
  t2[ND] = g.gamma * s.sv * t2[ND];
  for (int i = 0; i < ND; ++i) t2[i] = s.sv * t2[i] - s.u[i] * t2[ND];
  for (int l = 1; l < NV; ++l) r[l] = sign > 0 ? r[l] + t2[l - 1] : r[l] - t2[l - 1];
  double su = s.u[0] * t2[0];
  for (int i = 1; i < ND; ++i) su = su + s.u[i] * t2[i];
  const double a = s.sv * v[ND + 1] * t2[ND];
  r[0] = sign > 0 ? r[0] - a - su : r[0] + a + su;
}

template <int ND>
__device__ __forceinline__ void adj_chain_point(const Geo& g, const double* qi, const double* Si,
                                                long p, double sign, double* r) {
  constexpr int NV = ND + 2, NW = ND + 1;
  double v[NV];
  load_q<ND>(qi, g.N, p, v);
  double t2[NW];
  for (int a = 0; a < NW; ++a) {
    double x = Si[(1 + a) * g.N + p];
    for (int j = 1; j < ND; ++j) x = x + Si[(j * NV + 1 + a) * g.N + p];
    t2[a] = x;
  }
  adj_chain_t2<ND>(g, v, t2, sign, r);
}

template <int ND>
__global__ void __launch_bounds__(kBlock) k_adj_chain(Geo g, const double* __restrict__ q, const double* __restrict__ S,
                            double* __restrict__ rhs, double sign) {
  const long p = gidx();
  if (p >= g.N) return;
  const long inst = blockIdx.y;
  constexpr int NV = ND + 2;
  double* ri = rhs + inst * NV * g.N;
  double r[NV];
  for (int l = 0; l < NV; ++l) r[l] = ri[l * g.N + p];
  adj_chain_point<ND>(g, q + inst * NV * g.N, S + inst * ND * NV * g.N, p, sign, r);
  for (int l = 0; l < NV; ++l) ri[l * g.N + p] = r[l];
}

// Viscous far-field adjoint penalty, first half (src/RhsHelperImpl.f90:170-210):
// S_l -= sigma_v * B(dir, l)^T lambda on far-field patch points; S zeroed beforehand.
template <int ND>
__global__ void __launch_bounds__(kBlock) k_ff_visc_adj(Geo g, const Patches* __restrict__ P, const double* __restrict__ q,
                              const double* __restrict__ lam, double* __restrict__ S) {
  const long p = gidx();
  if (p >= g.N) return;
  const long inst = blockIdx.y;
  constexpr int NV = ND + 2, NW = ND + 1;
  int ijk[3] = {0, 0, 0};
  unflat<ND>(g, p, ijk);
  const double* qi = q + inst * NV * g.N;
  const double* li = lam + inst * NV * g.N;
  double* Si = S + inst * ND * NV * g.N;
  for (int ip = 0; ip < P->count; ++ip) {
    const DPatch& pt = P->p[ip];
    if (pt.type != MGD_PATCH_FAR_FIELD) continue;
    bool in = true;
    for (int d = 0; d < ND; ++d) in = in && ijk[d] >= pt.lo[d] && ijk[d] <= pt.hi[d];
    if (!in) continue;
    double v[NV];
    load_q<ND>(qi, g.N, p, v);
    const Prim<ND> s = prim<ND>(v, g.gamma);
    double mu, la, ka;
    transport(g, s.T, &mu, &la, &ka);
    const int dir = abs(pt.normal_dir) - 1;
    const double jac = jacobian<ND>(g, ijk);
    double m1[ND];
    for (int k = 0; k < ND; ++k) m1[k] = k == dir ? metric<ND>(g, dir, ijk) : 0.0;
    for (int l = 0; l < ND; ++l) {
      double m2[ND], B[NW * NW];
      for (int k = 0; k < ND; ++k) m2[k] = k == l ? metric<ND>(g, l, ijk) : 0.0;
      second_partial_viscous<ND>(s.u, mu, la, ka, jac, m1, m2, B);
      for (int a = 0; a < NW; ++a) {
        double t = 0.0;
        for (int k = 0; k < NW; ++k) t += B[a * NW + k] * li[(1 + k) * g.N + p];
        Si[(l * NV + 1 + a) * g.N + p] = Si[(l * NV + 1 + a) * g.N + p] - pt.sigma_visc * t;
      }
    }
  }
}

// Region%computeRhs tail (src/RegionImpl.f90:1969-2023): multiply by the Jacobian, add
// patch terms in bc.dat order, then acoustic sources (forward mode).
// Viscous far-field penalty terms of each far-field patch point, formed before the patch
// pass so that k_finish only adds them in bc.dat order (addFarFieldPenalty,
// src/FarFieldPatchImpl.f90:93-286): forward sigma_v J sum_d (F^v_d - F^v_target,d) m_d,
// adjoint sigma_v J B^T lambda with the first-partial viscous Jacobian B.
template <int ND>
__global__ void __launch_bounds__(kBlock) k_ff_visc_terms(Geo g, int mode, const Patches* __restrict__ P,
                                                          const double* __restrict__ q,
                                                          const double* __restrict__ lam,
                                                          const double* __restrict__ G) {
  const long p = gidx();
  if (p >= g.N) return;
  const long inst = blockIdx.y;
  constexpr int NV = ND + 2;
  const double* qi = q + inst * NV * g.N;
  const double* li = lam ? lam + inst * NV * g.N : nullptr;
  int ijk[3] = {0, 0, 0};
  unflat<ND>(g, p, ijk);
  for (int ip = 0; ip < P->count; ++ip) {
    const DPatch& pt = P->p[ip];
    if (pt.type != MGD_PATCH_FAR_FIELD || !pt.ffv) continue;
    bool in = true;
    for (int d = 0; d < ND; ++d) in = in && ijk[d] >= pt.lo[d] && ijk[d] <= pt.hi[d];
    if (!in) continue;
    int k = ijk[0] - pt.lo[0];
    if (ND > 1) k += pt.ls[0] * (ijk[1] - pt.lo[1]);
    if (ND > 2) k += pt.ls[0] * pt.ls[1] * (ijk[2] - pt.lo[2]);
    const double jac = jacobian<ND>(g, ijk);
    double v[NV];
    load_q<ND>(qi, g.N, p, v);
    const Prim<ND> s = prim<ND>(v, g.gamma);
    double tau[ND * ND], hf[ND], mu, la, ka;
    visc_point<ND>(g, G + inst * (ND + 1) * ND * g.N, g.N, p, s, tau, hf, &mu, &la, &ka);
    const int dir = abs(pt.normal_dir) - 1;
    const double vj = pt.sigma_visc * jac;
    double* out = pt.ffv + inst * NV * pt.np;
    if (mode == MGD_FORWARD) {
      for (int l = 0; l < NV; ++l) {
        double acc = 0.0;
        for (int d = 0; d < ND; ++d) {
          double fv[NV];
          viscous_flux<ND>(s.u, tau, hf, d, fv);
          const double md = d == dir ? metric<ND>(g, dir, ijk) : 0.0;
          acc += (fv[l] - pt.tvflux[((long)d * NV + l) * pt.np + k]) * md;
        }
        out[(long)l * pt.np + k] = vj * acc;
      }
    } else {
      double B[NV * NV], m[ND];
      for (int d = 0; d < ND; ++d) m[d] = d == dir ? metric<ND>(g, dir, ijk) : 0.0;
      first_partial_viscous<ND>(m, tau, hf, g.power_law, g.gamma, s, B);
      for (int i = 0; i < NV; ++i) {
        double t = 0.0;
        for (int c = 0; c < NV; ++c) t += B[i * NV + c] * li[c * g.N + p];
        out[(long)i * pt.np + k] = vj * t;
      }
    }
  }
}

template <int ND>
__global__ void __launch_bounds__(kBlock) k_finish(Geo g, int mode, const Patches* __restrict__ P, const double* __restrict__ q,
                         const double* __restrict__ lam, const double* __restrict__ G,
                         double* __restrict__ rhs, const double* __restrict__ control,
                         long control_inst_stride, const double* __restrict__ aforce,
                         long aforce_inst_stride, double forcing_factor, const Source* __restrict__ src,
                         int nsrc, const double* __restrict__ src_amp, long src_inst_stride,
                         const double* __restrict__ chain_S) {
  const long p = gidx();
  if (p >= g.N) return;
  const long inst = blockIdx.y;
  constexpr int NV = ND + 2;
  const double* qi = q + inst * NV * g.N;
  const double* li = lam ? lam + inst * NV * g.N : nullptr;
  double* ri = rhs + inst * NV * g.N;
  int ijk[3] = {0, 0, 0};
  unflat<ND>(g, p, ijk);
  const double jac = jacobian<ND>(g, ijk);
  double r[NV];
  for (int l = 0; l < NV; ++l) r[l] = ri[l * g.N + p];
  // adjoint, viscous, nothing else between: k_adj_chain's update (sign -1) of this point
  if (chain_S) adj_chain_point<ND>(g, qi, chain_S + inst * ND * NV * g.N, p, -1.0, r);
  for (int l = 0; l < NV; ++l) r[l] = r[l] * jac;
  for (int ip = 0; ip < P->count; ++ip) {
    const DPatch& pt = P->p[ip];
    bool in = true;
    for (int d = 0; d < ND; ++d) in = in && ijk[d] >= pt.lo[d] && ijk[d] <= pt.hi[d];
    if (!in) continue;
    int k = ijk[0] - pt.lo[0];
    if (ND > 1) k += pt.ls[0] * (ijk[1] - pt.lo[1]);
    if (ND > 2) k += pt.ls[0] * pt.ls[1] * (ijk[2] - pt.lo[2]);
    switch (pt.type) {
      case MGD_PATCH_FAR_FIELD: {
        // addFarFieldPenalty (src/FarFieldPatchImpl.f90:93-286); A+ from the target state
        const double* A = pt.aplus + (long)k * NV * NV;
        const double* qt = pt.aplus + (long)pt.np * NV * NV + (long)k * NV;
        const double sj = pt.sigma_inv * jac;
        if (mode == MGD_FORWARD) {
          for (int i = 0; i < NV; ++i) {
            double s = 0.0;
            for (int c = 0; c < NV; ++c) s += A[c * NV + i] * (qi[c * g.N + p] - qt[c]);
            r[i] = r[i] - sj * s;
          }
          if (g.viscous) {
            const double* fv = pt.ffv + inst * NV * pt.np;
            for (int l = 0; l < NV; ++l) r[l] = r[l] + fv[(long)l * pt.np + k];
          }
        } else {
          for (int i = 0; i < NV; ++i) {
            double s = 0.0;
            for (int c = 0; c < NV; ++c) s += A[i * NV + c] * li[c * g.N + p];
            r[i] = r[i] + sj * s;
          }
          if (g.viscous) {
            const double* fv = pt.ffv + inst * NV * pt.np;
            for (int i = 0; i < NV; ++i) r[i] = r[i] - fv[(long)i * pt.np + k];
          }
        }
        break;
      }
      case MGD_PATCH_IMPENETRABLE_WALL:
      case MGD_PATCH_ISOTHERMAL_WALL: {
        // addImpenetrableWallPenalty (src/ImpenetrableWallImpl.f90:60-207)
        double v[NV], m[ND];
        load_q<ND>(qi, g.N, p, v);
        const Prim<ND> s = prim<ND>(v, g.gamma);
        const int dir = abs(pt.normal_dir) - 1;
        for (int d = 0; d < ND; ++d) m[d] = d == dir ? metric<ND>(g, dir, ijk) : 0.0;
        double un = 0.0;
        for (int d = 0; d < ND; ++d) un += v[d + 1] * m[d];
        const double sj = pt.sigma_inv * jac;
        if (mode == MGD_FORWARD) {
          double pen[NV];
          pen[0] = un;
          for (int d = 0; d < ND; ++d) pen[d + 1] = un * s.u[d];
          pen[ND + 1] = un * s.sv * (v[ND + 1] + s.p);
          for (int l = 0; l < NV; ++l) r[l] = r[l] - sj * pen[l];
        } else {
          double A[NV * NV], dp[NV], ke = 0.0;
          inviscid_jacobian<ND>(m, s, g.gamma, A);
          for (int d = 0; d < ND; ++d) ke += s.u[d] * s.u[d];
          dp[0] = 0.5 * ke;
          for (int d = 0; d < ND; ++d) dp[d + 1] = -s.u[d];
          dp[ND + 1] = 1.0;
          for (int l = 0; l < NV; ++l) dp[l] = dp[l] * (g.gamma - 1.0);
          for (int d = 0; d < ND; ++d)
            for (int c = 0; c < NV; ++c) A[c * NV + d + 1] = A[c * NV + d + 1] - m[d] * dp[c];
          for (int i = 0; i < NV; ++i) {
            double t = 0.0;
            for (int c = 0; c < NV; ++c) t += A[i * NV + c] * li[c * g.N + p];
            r[i] = r[i] + sj * t;
          }
        }
        if (pt.type == MGD_PATCH_ISOTHERMAL_WALL && g.viscous) {
          // addIsothermalWallPenalty (src/IsothermalWallImpl.f90:101-330)
          double pen[NV];
          if (mode == MGD_FORWARD) {
            pen[0] = 0.0;
            for (int l = 1; l < NV; ++l) pen[l] = v[l];
            pen[ND + 1] = pen[ND + 1] - v[0] * pt.twall / g.gamma;
            for (int l = 0; l < NV; ++l) pen[l] = jac * pen[l];
            for (int l = 0; l < NV; ++l) r[l] = r[l] - pt.sigma_visc * pen[l];
          } else {
            pen[0] = -li[(ND + 1) * g.N + p] * pt.twall / g.gamma;
            for (int l = 1; l < NV; ++l) pen[l] = li[l * g.N + p];
            for (int l = 0; l < NV; ++l) pen[l] = jac * pen[l];
            for (int l = 0; l < NV; ++l) r[l] = r[l] + pt.sigma_visc * pen[l];
          }
        }
        break;
      }
      case MGD_PATCH_SPONGE: {
        // addDamping (src/SpongePatchImpl.f90:65-166)
        const double sp = pt.strength[k];
        const double* qt = pt.aplus + (long)k * NV;   // target state on the patch
        for (int l = 0; l < NV; ++l) {
          if (mode == MGD_FORWARD) r[l] = r[l] - sp * (qi[l * g.N + p] - qt[l]);
          else r[l] = r[l] + sp * li[l * g.N + p];
        }
        break;
      }
      case MGD_PATCH_ACTUATOR: {
        // updateActuatorPatch (src/ActuatorPatchImpl.f90:108-181): energy equation
        if (mode != MGD_FORWARD || !control) break;
        r[ND + 1] = r[ND + 1] + pt.strength[k] * control[inst * control_inst_stride + k];
        break;
      }
      case MGD_PATCH_COST_TARGET: {
        // addAdjointForcing (src/CostTargetPatchImpl.f90:72-136)
        if (mode == MGD_FORWARD || !aforce) break;
        const double* af = aforce + inst * aforce_inst_stride;
        for (int l = 0; l < NV; ++l) r[l] = r[l] + forcing_factor * af[(long)l * pt.np + k];
        break;
      }
      case MGD_PATCH_KOLMOGOROV: {
        // addKolmogorovForcing (src/KolmogorovForcingPatchImpl.f90:94-178)
        const double f = pt.strength[k];
        if (mode == MGD_FORWARD) r[1] = r[1] + qi[p] * f;
        else r[0] = r[0] - li[g.N + p] * f;
        break;
      }
    }
  }
  if (mode == MGD_FORWARD) {
    // addAcousticSource (src/AcousticSourceImpl.f90:34-64)
    for (int s = 0; s < nsrc; ++s) {
      double rr = 0.0;
      for (int d = 0; d < ND; ++d) {
        const double dx = g.x[d][ijk[d]] - src[s].loc[d];
        rr += dx * dx;
      }
      r[ND + 1] = r[ND + 1] + src_amp[inst * src_inst_stride + s] * exp(-src[s].gfac * rr);
    }
  }
  for (int l = 0; l < NV; ++l) ri[l * g.N + p] = r[l];
}

// ------------------------------------------------------------------ RK4 stages
// substepForwardRK4 updates (src/RK4IntegratorImpl.f90:115-156); q_out2 (optional)
// receives a copy of the new stage state (checkpoint window).
__global__ void __launch_bounds__(kBlock) k_rk_forward(long M, int stage, double dt, double* __restrict__ q,
                             double* __restrict__ b1, double* __restrict__ b2,
                             const double* __restrict__ rhs, double* __restrict__ qcopy) {
  const long i = gidx();
  if (i >= M) return;
  const double r = rhs[i];
  double qn;
  switch (stage) {
    case 1: {
      const double q0 = q[i];
      b1[i] = q0;
      b2[i] = q0 + div6(dt * r);
      qn = q0 + dt * r / 2.0;
      break;
    }
    case 2:
      b2[i] = b2[i] + div3(dt * r);
      qn = b1[i] + dt * r / 2.0;
      break;
    case 3:
      b2[i] = b2[i] + div3(dt * r);
      qn = b1[i] + dt * r;
      break;
    default:
      qn = b2[i] + div6(dt * r);
      break;
  }
  q[i] = qn;
  if (qcopy) qcopy[i] = qn;
}

// substepAdjointRK4 updates (src/RK4IntegratorImpl.f90:215-261)
__global__ void __launch_bounds__(kBlock) k_rk_adjoint(long M, int stage, double dt, double* __restrict__ lam,
                             double* __restrict__ b1, double* __restrict__ b2,
                             const double* __restrict__ rhs) {
  const long i = gidx();
  if (i >= M) return;
  const double r = rhs[i];
  switch (stage) {
    case 4: {
      const double l0 = lam[i];
      b1[i] = l0;
      b2[i] = l0 - div6(dt * r);
      lam[i] = l0 - dt * r / 2.0;
      break;
    }
    case 3:
      b2[i] = b2[i] - div3(dt * r);
      lam[i] = b1[i] - dt * r / 2.0;
      break;
    case 2:
      b2[i] = b2[i] - div3(dt * r);
      lam[i] = b1[i] - dt * r;
      break;
    default:
      lam[i] = b2[i] - div6(dt * r);
      break;
  }
}

__global__ void __launch_bounds__(kBlock) k_copy(long M, const double* __restrict__ a, double* __restrict__ b) {
  const long i = gidx();
  if (i < M) b[i] = a[i];
}

__global__ void __launch_bounds__(kBlock) k_fill(long M, double v, double* __restrict__ b) {
  const long i = gidx();
  if (i < M) b[i] = v;
}

// ------------------------------------------------------------------ patch reductions
// Per-point integrand over one box patch, reduced in two deterministic passes.
template <int ND>
__device__ __forceinline__ long patch_point(const Geo& g, const DPatch& pt, int k, int* ijk) {
  ijk[0] = pt.lo[0] + k % pt.ls[0];
  ijk[1] = ND > 1 ? pt.lo[1] + (k / pt.ls[0]) % pt.ls[1] : 0;
  ijk[2] = ND > 2 ? pt.lo[2] + k / (pt.ls[0] * pt.ls[1]) : 0;
  return ijk[0] + (long)g.n[0] * (ijk[1] + (long)g.n[1] * ijk[2]);
}

constexpr int kRedBlock = 256;

__device__ __forceinline__ double block_sum(double v) {
  __shared__ double sh[kRedBlock];
  sh[threadIdx.x] = v;
  __syncthreads();
  for (int s = kRedBlock / 2; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s) sh[threadIdx.x] += sh[threadIdx.x + s];
    __syncthreads();
  }
  return sh[0];
}

// which 0: AcousticNoise integrand norm * (p - pbar)^2 * m_T on the cost-target patch
//          (computeAcousticNoise + computeQuadratureOnPatches);
// which 1: ThermalActuator sensitivity norm * (lambda_E m_C)^2 on the actuator patch.
template <int ND>
__global__ void __launch_bounds__(kBlock) k_patch_reduce(Geo g, const DPatch* __restrict__ pt_, int which,
                               const double* __restrict__ x, double* __restrict__ partial) {
  const DPatch& pt = *pt_;
  const long inst = blockIdx.y;
  const int k = blockIdx.x * kRedBlock + threadIdx.x;
  double v = 0.0;
  if (k < pt.np) {
    int ijk[3];
    const long p = patch_point<ND>(g, pt, k, ijk);
    const double* xi = x + inst * g.nv * g.N;
    const double nrm = gnorm<ND>(g, ijk);
    if (which == 0) {
      double q[ND + 2];
      load_q<ND>(xi, g.N, p, q);
      const Prim<ND> s = prim<ND>(q, g.gamma);
      const double F = s.p - pt.aplus[k];            // mean pressure on the patch
      v = 1.0 * nrm * (F * F * pt.strength[k]);      // target mollifier
    } else {
      const double F = xi[(ND + 1) * g.N + p] * pt.strength[k] * 1.0;   // control mollifier
      v = 1.0 * nrm * (F * F);
    }
  }
  const double s = block_sum(v);
  if (threadIdx.x == 0) partial[inst * gridDim.x + blockIdx.x] = s;
}

// acc[inst] += weight * sum(partials); out_stage (optional) keeps the raw value.
__global__ void __launch_bounds__(kBlock) k_reduce_final(int nblocks, const double* __restrict__ partial, double weight,
                               double* __restrict__ acc, double* __restrict__ raw) {
  const long inst = blockIdx.x;
  double v = 0.0;
  for (int b = threadIdx.x; b < nblocks; b += kRedBlock) v += partial[inst * nblocks + b];
  const double s = block_sum(v);
  if (threadIdx.x == 0) {
    if (raw) raw[inst] = s;
    if (acc) acc[inst] = acc[inst] + weight * s;
  }
}

// Grid%computeInnerProduct (src/GridImpl.f90:1067-1170): per-block partial sums of
// sum over components of a * norm * b at each point of one instance.
template <int ND>
__global__ void __launch_bounds__(kBlock) k_inner_partial(Geo g, const double* __restrict__ a, const double* __restrict__ b,
                                double* __restrict__ partial) {
  const long inst = blockIdx.y;
  const long p = (long)blockIdx.x * kRedBlock + threadIdx.x;
  double v = 0.0;
  if (p < g.N) {
    int ijk[3];
    unflat<ND>(g, p, ijk);
    const double nrm = gnorm<ND>(g, ijk);
    const double* ai = a + inst * g.nv * g.N;
    const double* bi = b + inst * g.nv * g.N;
    for (int c = 0; c < g.nv; ++c) v += ai[c * g.N + p] * nrm * bi[c * g.N + p];
  }
  const double s = block_sum(v);
  if (threadIdx.x == 0) partial[inst * gridDim.x + blockIdx.x] = s;
}

// computeTimeStepSize / computeCfl (src/CNSHelperImpl.f90:842-982) per instance: the
// advective wave speed J (c |m| + sum_j |u . m_j|), c^2 = (gamma - 1) T, and in viscous
// regions the diffusive one J^2 sum(metrics^2) max(2 mu, kappa); mode 0 reduces
// cfl / speed by min (the time step), mode 1 speed * dt by max (the CFL number).  The
// grid is rectilinear: the reference's row sums and dot products add exact zeros for the
// off-diagonal metrics, so the sums below, without them, are the same numbers.
template <int ND>
__global__ void __launch_bounds__(kBlock) k_wave_speed(Geo g, const double* __restrict__ q, int mode,
                                                       double param, double* __restrict__ partial) {
  __shared__ double sh[kRedBlock];
  const long inst = blockIdx.y;
  const long p = (long)blockIdx.x * kRedBlock + threadIdx.x;
  double v = mode == 0 ? DBL_MAX : 0.0;
  if (p < g.N) {
    int ijk[3];
    unflat<ND>(g, p, ijk);
    double x[ND + 2];
    load_q<ND>(q + inst * (ND + 2) * g.N, g.N, p, x);
    const Prim<ND> s = prim<ND>(x, g.gamma);
    const double jac = jacobian<ND>(g, ijk);
    double m[ND];
    for (int j = 0; j < ND; ++j) m[j] = metric<ND>(g, j, ijk);
    double w = 0.0;
    for (int j = 0; j < ND; ++j) w = w + m[j] * m[j];
    const double c = sqrt((g.gamma - 1.0) * s.T);
    w = c * sqrt(w);
    for (int j = 0; j < ND; ++j) w = w + fabs(s.u[j] * m[j]);
    w = jac * w;
    v = mode == 0 ? param / w : w * param;
    if (g.viscous) {
      double mu, la, ka;
      transport(g, s.T, &mu, &la, &ka);
      double sm = 0.0;
      for (int j = 0; j < ND; ++j) sm += m[j] * m[j];
      const double a = 2.0 * mu;
      const double wd = jac * jac * sm * (a > ka ? a : ka);
      const double vd = mode == 0 ? param / wd : wd * param;
      v = mode == 0 ? (vd < v ? vd : v) : (vd > v ? vd : v);
    }
  }
  sh[threadIdx.x] = v;
  __syncthreads();
  for (int st = kRedBlock / 2; st > 0; st >>= 1) {
    if ((int)threadIdx.x < st) {
      const double o = sh[threadIdx.x + st];
      const double a = sh[threadIdx.x];
      sh[threadIdx.x] = mode == 0 ? (o < a ? o : a) : (o > a ? o : a);
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) partial[inst * gridDim.x + blockIdx.x] = sh[0];
}

// out[inst] = min (mode 0) or max (mode 1) of the instance's partials
__global__ void __launch_bounds__(kBlock) k_minmax_final(int nblocks, const double* __restrict__ partial,
                                                         int mode, double* __restrict__ out) {
  __shared__ double sh[kRedBlock];
  const long inst = blockIdx.x;
  double v = mode == 0 ? DBL_MAX : 0.0;
  for (int b = threadIdx.x; b < nblocks; b += kRedBlock) {
    const double o = partial[inst * nblocks + b];
    v = mode == 0 ? (o < v ? o : v) : (o > v ? o : v);
  }
  sh[threadIdx.x] = v;
  __syncthreads();
  for (int st = kRedBlock / 2; st > 0; st >>= 1) {
    if ((int)threadIdx.x < st) {
      const double o = sh[threadIdx.x + st];
      const double a = sh[threadIdx.x];
      sh[threadIdx.x] = mode == 0 ? (o < a ? o : a) : (o > a ? o : a);
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) out[inst] = sh[0];
}

// checkSolutionLimits (src/SolverImpl.f90:189-304, isVariableWithinRange
// src/GridImpl.f90:1579-1620): an instance whose density or temperature reaches a bound
// (f <= min or f >= max), or is not a number, gets flag[inst] = 1 (a plain store: every
// writer stores the same value).
template <int ND>
__global__ void __launch_bounds__(kBlock) k_check_limits(Geo g, const double* __restrict__ q,
                                                         double rmin, double rmax, double tmin,
                                                         double tmax, double* __restrict__ flag) {
  const long p = gidx();
  if (p >= g.N) return;
  const long inst = blockIdx.y;
  double v[ND + 2];
  load_q<ND>(q + inst * (ND + 2) * g.N, g.N, p, v);
  const Prim<ND> s = prim<ND>(v, g.gamma);
  const bool ok = v[0] > rmin && v[0] < rmax && s.T > tmin && s.T < tmax;
  if (!ok) flag[inst] = 1.0;
}

// Terminal kinetic-energy cost J = sum_p H_p |rho u|^2 / (2 rho) of each instance and its
// pointwise derivative, the adjoint state at the final time (the terminal condition of
// adjoint_nonzero_initial_condition, src/SolverImpl.f90:368-428, 1096-1102):
// lam = (-|rho u|^2 / (2 rho^2), rho u_i / rho, 0).  Per-block partial sums of H_p ke_p.
template <int ND>
__global__ void __launch_bounds__(kBlock) k_kinetic_energy(Geo g, const double* __restrict__ q,
                                                           double* __restrict__ lam,
                                                           double* __restrict__ partial) {
  const long inst = blockIdx.y;
  const long p = (long)blockIdx.x * kRedBlock + threadIdx.x;
  double v = 0.0;
  if (p < g.N) {
    int ijk[3];
    unflat<ND>(g, p, ijk);
    const double* qi = q + inst * (ND + 2) * g.N;
    double* li = lam + inst * (ND + 2) * g.N;
    const double rho = qi[p];
    double m2 = 0.0;
    for (int i = 0; i < ND; ++i) {
      const double m = qi[(1 + i) * g.N + p];
      m2 = m2 + m * m;
      li[(1 + i) * g.N + p] = m / rho;
    }
    li[p] = -0.5 * m2 / (rho * rho);
    li[(ND + 1) * g.N + p] = 0.0;
    v = gnorm<ND>(g, ijk) * (0.5 * m2 / rho);
  }
  const double s = block_sum(v);
  if (threadIdx.x == 0) partial[inst * gridDim.x + blockIdx.x] = s;
}

// d = c + s (a - b) (c may be null: d = s (a - b)) over one instance's state array.
__global__ void __launch_bounds__(kBlock) k_lincomb(long M, const double* __restrict__ a, const double* __restrict__ b,
                          double s, const double* __restrict__ c, double* __restrict__ d) {
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= M) return;
  d[i] = c ? c[i] + s * (a[i] - b[i]) : s * (a[i] - b[i]);
}

// AcousticNoise adjoint forcing on the cost-target patch (src/AcousticNoiseImpl.f90:208-280)
template <int ND>
__global__ void __launch_bounds__(kBlock) k_adj_forcing(Geo g, const DPatch* __restrict__ pt_, const double* __restrict__ q,
                              double* __restrict__ out, int zero) {
  const DPatch& pt = *pt_;
  const long inst = blockIdx.y;
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= pt.np) return;
  double* o = out + inst * (long)(ND + 2) * pt.np;
  if (zero) {
    for (int l = 0; l < ND + 2; ++l) o[(long)l * pt.np + k] = 0.0;
    return;
  }
  int ijk[3];
  const long p = patch_point<ND>(g, pt, k, ijk);
  double v[ND + 2];
  load_q<ND>(q + inst * g.nv * g.N, g.N, p, v);
  const Prim<ND> s = prim<ND>(v, g.gamma);
  const double F = -2.0 * pt.strength[k] * 1.0 * (g.gamma - 1.0) * (s.p - pt.aplus[k]);
  o[(long)(ND + 1) * pt.np + k] = F;
  double ke = 0.0;
  for (int i = 0; i < ND; ++i) {
    o[(long)(i + 1) * pt.np + k] = -s.u[i] * F;
    ke += s.u[i] * s.u[i];
  }
  o[k] = 0.5 * ke * F;
}

// adjoint terminal condition: lambda += factor * forcing on the cost-target patch
// (loadInitialCondition, src/SolverImpl.f90:368-428)
template <int ND>
__global__ void __launch_bounds__(kBlock) k_adj_terminal(Geo g, const DPatch* __restrict__ pt_, const double* __restrict__ af,
                               double factor, double* __restrict__ lam) {
  const DPatch& pt = *pt_;
  const long inst = blockIdx.y;
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= pt.np) return;
  int ijk[3];
  const long p = patch_point<ND>(g, pt, k, ijk);
  const double* a = af + inst * (long)(ND + 2) * pt.np;
  double* li = lam + inst * g.nv * g.N;
  for (int l = 0; l < ND + 2; ++l) {
    const double r = 0.0 + factor * a[(long)l * pt.np + k];
    li[l * g.N + p] = li[l * g.N + p] + r;
  }
}

// ThermalActuator gradient lambda_E * m_C on the actuator patch (:383-443)
template <int ND>
__global__ void __launch_bounds__(kBlock) k_gradient(Geo g, const DPatch* __restrict__ pt_, const double* __restrict__ lam,
                           double* __restrict__ out, long inst_stride) {
  const DPatch& pt = *pt_;
  const long inst = blockIdx.y;
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= pt.np) return;
  int ijk[3];
  const long p = patch_point<ND>(g, pt, k, ijk);
  out[inst * inst_stride + k] = lam[inst * g.nv * g.N + (ND + 1) * g.N + p] * (pt.strength[k] * 1.0);
}

// x and y parts of the flux divergence on a 3-D grid periodic in x and y, one z plane
// per block row: each block stages the contravariant fluxes of its BX x BY points (x
// rows and y columns, with the 4-point halo recomputed from q and G) in LDS and writes
// D_x Fx + D_y Fy to rhs and Fz to its slot of F, for the z divergence (k_deriv_line,
// k_fwd_finish_z or k_div) to finish.  The per-point flux and the derivative sums are
// those of k_flux_all and k_div (same operations in the same order, so the same
// values).
template <int BX, int BY>
__global__ void __launch_bounds__(BX * BY) k_fluxdiv3d_xy(Geo g, const double* __restrict__ q,
                                                           const double* __restrict__ G,
                                                           double* __restrict__ F,
                                                           double* __restrict__ rhs) {
  constexpr int NV = 5, EXW = BX + 8, EYW = BY + 8;
  static_assert(8 * (BX + BY) <= BX * BY, "at most one halo point per thread");
  __shared__ double sfx[NV][BY][EXW];
  __shared__ double sfy[NV][EYW][BX];
  const int ntx = g.n[0] / BX;
  const int x0 = (blockIdx.x % ntx) * BX, y0 = (blockIdx.x / ntx) * BY, z = blockIdx.y;
  const long inst = blockIdx.z;
  const double* qi = q + inst * NV * g.N;
  const double* Gi = g.viscous ? G + inst * 12 * g.N : nullptr;
  const long n01 = (long)g.n[0] * g.n[1];
  // flux of direction d at grid point (ii, jj, z) for each d with want[d]
  auto fluxes = [&](int ii, int jj, const bool* want, double (*out)[NV]) {
    const long pp = ii + (long)g.n[0] * jj + n01 * z;
    double v[NV];
    load_q<3>(qi, g.N, pp, v);
    const Prim<3> s = prim<3>(v, g.gamma);
    double tau[9], hf[3], mu, la, ka;
    if (g.viscous) visc_point<3>(g, Gi, g.N, pp, s, tau, hf, &mu, &la, &ka);
    const int ijk[3] = {ii, jj, z};
#pragma unroll
    for (int d = 0; d < 3; ++d) {
      if (!want[d]) continue;
      double f[NV];
      inviscid_flux<3>(v, s, d, f);
      if (g.viscous) {
        double fv[NV];
        viscous_flux<3>(s.u, tau, hf, d, fv);
#pragma unroll
        for (int l = 0; l < NV; ++l) f[l] = f[l] - fv[l];
      }
      const double m = metric<3>(g, d, ijk);
#pragma unroll
      for (int l = 0; l < NV; ++l) out[d][l] = m * f[l];
    }
  };
  const int tx = threadIdx.x % BX, ty = threadIdx.x / BX;
  const int i = x0 + tx, j = y0 + ty;
  const long p = i + (long)g.n[0] * j + n01 * z;
  {
    const bool all[3] = {true, true, true};
    double f[3][NV];
    fluxes(i, j, all, f);
    double* Fz = F + (inst * 3 + 2) * NV * g.N;
#pragma unroll
    for (int l = 0; l < NV; ++l) {
      sfx[l][ty][tx + 4] = f[0][l];
      sfy[l][ty + 4][tx] = f[1][l];
      Fz[l * g.N + p] = f[2][l];
    }
  }
  {
    // halo: threads [0, 8 BY) take the x halo columns, [8 BY, 8 BY + 8 BX) the y halo rows
    const int e = threadIdx.x;
    double f[3][NV];
    if (e >= 8 * (BX + BY)) {
      // no halo point
    } else if (e < 8 * BY) {
      const int hx = e % 8, hy = e / 8;
      const int ex = hx < 4 ? hx : BX + hx;
      const bool wx[3] = {true, false, false};
      fluxes(wrap(x0 + ex - 4, g.dir[0]), y0 + hy, wx, f);
#pragma unroll
      for (int l = 0; l < NV; ++l) sfx[l][hy][ex] = f[0][l];
    } else {
      const int e2 = e - 8 * BY;
      const int hx = e2 % BX, hy = e2 / BX;
      const int ey = hy < 4 ? hy : BY + hy;
      const bool wy[3] = {false, true, false};
      fluxes(x0 + hx, wrap(y0 + ey - 4, g.dir[1]), wy, f);
#pragma unroll
      for (int l = 0; l < NV; ++l) sfy[l][ey][hx] = f[1][l];
    }
  }
  __syncthreads();
  const Op& op = g.ops[SLOT_FIRST];
  const double c1 = op.c[kMaxHW + 1], c2 = op.c[kMaxHW + 2], c3 = op.c[kMaxHW + 3], c4 = op.c[kMaxHW + 4];
  double* ri = rhs + inst * NV * g.N;
#pragma unroll
  for (int l = 0; l < NV; ++l) {
    const double* rx = &sfx[l][ty][tx + 4];
    const double* cy = &sfy[l][ty + 4][tx];
    double ox = 0.0, oy = 0.0;
    ox += c1 * (rx[1] - rx[-1]);
    ox += c2 * (rx[2] - rx[-2]);
    ox += c3 * (rx[3] - rx[-3]);
    ox += c4 * (rx[4] - rx[-4]);
    oy += c1 * (cy[BX] - cy[-BX]);
    oy += c2 * (cy[2 * BX] - cy[-2 * BX]);
    oy += c3 * (cy[3 * BX] - cy[-3 * BX]);
    oy += c4 * (cy[4 * BX] - cy[-4 * BX]);
    ri[l * g.N + p] = ox + oy;
  }
}

// First-derivative-type operator (SBP 4-8 interior: skew, half width 4) along a
// periodic direction d >= 1 by marching: each thread owns one grid line along d (lanes
// along x, so every load is coalesced) and walks a chunk of kLineChunk points of it,
// keeping the 9-point window of one component in registers -- each operand value is
// read once per chunk instead of once per stencil tap.  Component c of instance i reads
// src[i * src_inst + c * N] and writes dst[i * dst_inst + c * N]; mode 0 stores the
// result, 1 adds it to dst, 2 stores 0 - (dst + result).  The sums are those of
// sbp_apply's interior rows (same order), so the values are the same.
constexpr int kLineChunk = 32;
__global__ void __launch_bounds__(kBlock) k_deriv_line(Geo g, int d, int slot,
                                                      const double* __restrict__ src, long src_inst,
                                                      double* __restrict__ dst, long dst_inst,
                                                      int ncomp, int mode) {
  const Dir dd = g.dir[d];
  const long nlines = g.N / dd.len;
  const long line = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (line >= nlines) return;
  const int inst = blockIdx.z / ncomp, c = blockIdx.z % ncomp;
  // line -> first point: x index fastest, then the remaining direction
  const long n0 = g.n[0];
  const long i = line % n0, o = line / n0;
  const long base = i + o * (d == 1 ? n0 * g.n[1] : n0);
  const double* x = src + inst * src_inst + (long)c * g.N + base;
  double* y = dst + inst * dst_inst + (long)c * g.N + base;
  const long sd = dd.stride;
  const int z0 = blockIdx.y * kLineChunk;
  const int z1 = min(z0 + kLineChunk, dd.len);
  const Op& op = g.ops[slot];
  const double c1 = op.c[kMaxHW + 1], c2 = op.c[kMaxHW + 2], c3 = op.c[kMaxHW + 3],
               c4 = op.c[kMaxHW + 4];
  double Q[9];
#pragma unroll
  for (int k = 0; k < 8; ++k) Q[k] = x[wrap(z0 - 4 + k, dd) * sd];
  // the operand entering the window and the dst value of the next point are loaded one
  // point ahead, so their latency overlaps this point's work
  double nx = x[wrap(z0 + 4, dd) * sd];
  double ny = mode != 0 ? y[z0 * sd] : 0.0;
  for (int z = z0; z < z1; ++z) {
    Q[8] = nx;
    const double yv = ny;
    if (z + 1 < z1) {
      nx = x[wrap(z + 5, dd) * sd];
      if (mode != 0) ny = y[(z + 1) * sd];
    }
    double out = 0.0;
    out += c1 * (Q[5] - Q[3]);
    out += c2 * (Q[6] - Q[2]);
    out += c3 * (Q[7] - Q[1]);
    out += c4 * (Q[8] - Q[0]);
    double* yz = y + z * sd;
    if (mode == 0) {
      *yz = out;
    } else {
      double v = yv + out;
      if (mode == 2) v = 0.0 - v;
      *yz = v;
    }
#pragma unroll
    for (int k = 0; k < 8; ++k) Q[k] = Q[k + 1];
  }
}

// The z divergence of a forward rhs fused with k_finish's forward operations, for 3-D
// regions whose patches are at most one Kolmogorov forcing over the whole domain, with
// no sources and no dissipation.  One thread per z-line as in k_deriv_line, the 9-point
// windows of the five flux components in registers; at each point
// rhs = 0 - (x/y divergence + D_z Fz) (k_deriv_line's add-negate), then the Jacobian
// scaling and the forcing of the x momentum, in k_finish's order.  With rk_stage > 0 the
// point's RK4 stage update follows (k_rk_forward's arithmetic, y = q updated in place,
// copy optional), after all five components are formed since the forcing reads the
// density at the point, and rhs is not written.
__global__ void __launch_bounds__(kBlock) k_fwd_finish_z(Geo g, const double* __restrict__ Fz,
                                                        long f_inst, const double* q,
                                                        const double* __restrict__ kstr,
                                                        double* __restrict__ rhs, int rk_stage,
                                                        double dt, double* y, double* b1,
                                                        double* b2, double* copy) {
  constexpr int ND = 3, NV = 5;
  const Dir dd = g.dir[2];
  const long nlines = g.N / dd.len;
  const long line = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (line >= nlines) return;
  const long inst = blockIdx.z;
  const long n0 = g.n[0];
  const long i = line % n0, o = line / n0;
  const long base = i + o * n0;
  const long sd = dd.stride;
  const double* x = Fz + inst * f_inst + base;
  const double* qi = q + inst * NV * g.N;
  double* ri = rhs + inst * NV * g.N;
  const int z0 = blockIdx.y * kLineChunk;
  const int z1 = min(z0 + kLineChunk, dd.len);
  const Op& op = g.ops[SLOT_FIRST];
  const double c1 = op.c[kMaxHW + 1], c2 = op.c[kMaxHW + 2], c3 = op.c[kMaxHW + 3],
               c4 = op.c[kMaxHW + 4];
  double Q[NV][9];
#pragma unroll
  for (int l = 0; l < NV; ++l)
#pragma unroll
    for (int k = 0; k < 8; ++k) Q[l][k] = x[l * g.N + wrap(z0 - 4 + k, dd) * sd];
  for (int z = z0; z < z1; ++z) {
    const long p = base + z * sd;
    double r[NV];
#pragma unroll
    for (int l = 0; l < NV; ++l) {
      Q[l][8] = x[l * g.N + wrap(z + 4, dd) * sd];
      double out = 0.0;
      out += c1 * (Q[l][5] - Q[l][3]);
      out += c2 * (Q[l][6] - Q[l][2]);
      out += c3 * (Q[l][7] - Q[l][1]);
      out += c4 * (Q[l][8] - Q[l][0]);
      double v = ri[l * g.N + p] + out;
      r[l] = 0.0 - v;
#pragma unroll
      for (int k = 0; k < 8; ++k) Q[l][k] = Q[l][k + 1];
    }
    int ijk[3] = {(int)i, (int)o, z};
    const double jac = jacobian<ND>(g, ijk);
#pragma unroll
    for (int l = 0; l < NV; ++l) r[l] = r[l] * jac;
    if (kstr) r[1] = r[1] + qi[p] * kstr[p];
    if (rk_stage == 0) {
#pragma unroll
      for (int l = 0; l < NV; ++l) ri[l * g.N + p] = r[l];
      continue;
    }
#pragma unroll
    for (int l = 0; l < NV; ++l) {
      const long ii = inst * NV * g.N + l * g.N + p;
      const double rl = r[l];
      double qn;
      switch (rk_stage) {
        case 1: {
          const double q0 = y[ii];
          b1[ii] = q0;
          b2[ii] = q0 + div6(dt * rl);
          qn = q0 + dt * rl / 2.0;
          break;
        }
        case 2:
          b2[ii] = b2[ii] + div3(dt * rl);
          qn = b1[ii] + dt * rl / 2.0;
          break;
        case 3:
          b2[ii] = b2[ii] + div3(dt * rl);
          qn = b1[ii] + dt * rl;
          break;
        default:
          qn = b2[ii] + div6(dt * rl);
          break;
      }
      y[ii] = qn;
      if (copy) copy[ii] = qn;
    }
  }
}

// XCD-aware (tile, z chunk) work unit of block b of a 3-D march: units are dealt to the
// 8 XCDs round-robin by block index, so block b takes unit (b mod 8) * ceil(units / 8) +
// b / 8, tiles x-fastest: each XCD marches one contiguous band, and the halo lines of
// neighbouring tiles are fetched into its own L2 by blocks running beside it.  False for
// the padding blocks of a grid rounded up to a multiple of 8.
template <int BX, int BY, int ZC>
__device__ __forceinline__ bool march_unit(const Geo& g, int& x0, int& y0, int& z0) {
  const int ntx = g.n[0] / BX, ntiles = ntx * (g.n[1] / BY);
  const int nunits = ntiles * ((g.n[2] + ZC - 1) / ZC);
  const int per_xcd = (nunits + 7) / 8;
  const int unit = (int)(blockIdx.x % 8) * per_xcd + (int)(blockIdx.x / 8);
  if (unit >= nunits) return false;
  const int tile = unit % ntiles;
  x0 = (tile % ntx) * BX;
  y0 = (tile / ntx) * BY;
  z0 = (unit / ntiles) * ZC;
  return true;
}

// The x / y halo point of thread t of a BX x BY tile (at most one per thread, 8 (BX + BY)
// of them): [0, 8 BY) the x-halo columns, [8 BY, 8 (BX + BY)) the y-halo rows, none for
// the threads after them.  (hx, hy) is its place in the haloed tile, (hi, hj) its grid
// indices; hmask 1 for an x-halo point, 2 for a y-halo point, 0 for none (then hi, hj are
// the thread's own point, so that every index stays in the grid).
template <int BX, int BY>
__device__ __forceinline__ void march_halo(const Geo& g, int t, int x0, int y0, int& hx, int& hy,
                                           int& hi, int& hj, int& hmask) {
  if (t < 8 * BY) {
    const int c = t % 8;
    hy = t / 8;
    hx = c < 4 ? c : BX + c;
    hi = wrap(x0 + hx - 4, g.dir[0]);
    hj = y0 + hy;
    hmask = 1;
  } else if (t < 8 * (BX + BY)) {
    const int t2 = t - 8 * BY, r = t2 / BX;
    hx = t2 % BX;
    hy = r < 4 ? r : BY + r;
    hi = x0 + hx;
    hj = wrap(y0 + hy - 4, g.dir[1]);
    hmask = 2;
  } else {
    hx = hy = 0;
    hi = x0 + t % BX;
    hj = y0 + t / BX;
    hmask = 0;
  }
}

// The whole forward rhs of a 3-D region periodic in every direction, in one z march
// (computeRhsForward, src/RhsHelperImpl.f90:254-354, with k_finish's operations for at
// most one whole-domain Kolmogorov forcing and no dissipation or sources): a BX x BY
// column tile marches along z over a chunk of ZC planes.  At each lead plane zl every
// thread forms the contravariant fluxes of its tile point (from q and, viscous, the
// stress / heat-flux slots S written by k_grad3d_periodic<.., true>) and of one x- or
// y-halo point; the x and y fluxes go through LDS for D_x Fx + D_y Fy, which waits in a
// 4-plane LDS ring, and Fz enters a 9-plane register window.  Four planes behind, at
// zo = zl - 4, rhs = 0 - ((D_x Fx + D_y Fy) + D_z Fz) -- zo's ring slot is read before
// zl's partial takes it over -- then the Jacobian scaling and the
// forcing; with rk_stage > 0 the RK4 stage update follows and goes to yout, which must
// not be q (other blocks still read q), and rhs is not written; at stage 1 b1 may be
// null (the caller keeps q itself as the state at the start of the step), and at stages
// 1-3 b2 may be null (a recompute that stops before stage 4 needs no accumulator).  The fluxes, sums and
// finishing operations are those of k_fluxdiv3d_xy and k_fwd_finish_z, in the same
// order, so the values are the same; no flux or partial rhs goes through HBM.
// RK: the variant for RK stages (rk_stage > 0), which loads the RK operands of its output
// plane early; the rhs-only variant leaves them out and keeps the registers for the rest.
template <int BX, int BY, int ZC, bool RK>
__global__ void __launch_bounds__(BX * BY, 512 / (BX * BY)) k_fwd3d_march(Geo g, const double* __restrict__ q,
                                                          const double* __restrict__ S,
                                                          const double* __restrict__ kstr,
                                                          double* __restrict__ rhs, int rk_stage,
                                                          double dt, double* __restrict__ yout,
                                                          double* __restrict__ b1,
                                                          double* __restrict__ b2,
                                                          double* __restrict__ copy) {
  constexpr int NV = 5, NT = BX * BY, EXW = BX + 8, EYW = BY + 8;
  static_assert(8 * (BX + BY) <= NT, "at most one halo point per thread");
  __shared__ double sfx[NV][BY][EXW];
  __shared__ double sfy[NV][EYW][BX];
  __shared__ double ring[4][NV][NT];
  int x0, y0, z0;
  if (!march_unit<BX, BY, ZC>(g, x0, y0, z0)) return;
  const int z1 = min(z0 + ZC, g.n[2]);
  const long inst = blockIdx.z;
  const double* qi = q + inst * NV * g.N;
  const double* Si = g.viscous ? S + inst * 12 * g.N : nullptr;
  const long n01 = (long)g.n[0] * g.n[1];
  const int t = threadIdx.x, tx = t % BX, ty = t / BX;
  const int i = x0 + tx, j = y0 + ty;
  int hx, hy, hi, hj, hmask;
  march_halo<BX, BY>(g, t, x0, y0, hx, hy, hi, hj, hmask);
  // operands of one point: q and, viscous, the stress / heat-flux slots the fluxes of the
  // directions in dmask read (tau is symmetric: slot order xx xy xz yy yz zz, then the
  // heat flux x y z); slots not needed are not loaded
  struct Pt {
    double v[NV];
    double sv[9];
  };
  auto load_pt = [&](int ii, int jj, int kk, int dmask, Pt& P) {
    const long pp = ii + (long)g.n[0] * jj + n01 * kk;
    load_q<3>(qi, g.N, pp, P.v);
    if (g.viscous) {
      const int need[9] = {dmask & 1, dmask & 3, dmask & 5, dmask & 2, dmask & 6, dmask & 4,
                           dmask & 1, dmask & 2, dmask & 4};
#pragma unroll
      for (int k = 0; k < 9; ++k) P.sv[k] = need[k] ? Si[k * g.N + pp] : 0.0;
    }
  };
  // contravariant fluxes of the directions in dmask at grid point (ii, jj, kk)
  auto fluxes = [&](int ii, int jj, int kk, int dmask, const Pt& P, double (*out)[NV]) {
    const Prim<3> s = prim<3>(P.v, g.gamma);
    double tau[9], hf[3];
    if (g.viscous) {
      tau[0] = P.sv[0]; tau[1] = P.sv[1]; tau[2] = P.sv[2];
      tau[3] = P.sv[1]; tau[4] = P.sv[3]; tau[5] = P.sv[4];
      tau[6] = P.sv[2]; tau[7] = P.sv[4]; tau[8] = P.sv[5];
      hf[0] = P.sv[6]; hf[1] = P.sv[7]; hf[2] = P.sv[8];
    }
    const int ijk[3] = {ii, jj, kk};
#pragma unroll
    for (int d = 0; d < 3; ++d) {
      if (!(dmask >> d & 1)) continue;
      double f[NV];
      inviscid_flux<3>(P.v, s, d, f);
      if (g.viscous) {
        double fv[NV];
        viscous_flux<3>(s.u, tau, hf, d, fv);
#pragma unroll
        for (int l = 0; l < NV; ++l) f[l] = f[l] - fv[l];
      }
      const double m = metric<3>(g, d, ijk);
#pragma unroll
      for (int l = 0; l < NV; ++l) out[d][l] = m * f[l];
    }
  };
  const Op& op = g.ops[SLOT_FIRST];
  const double c1 = op.c[kMaxHW + 1], c2 = op.c[kMaxHW + 2], c3 = op.c[kMaxHW + 3], c4 = op.c[kMaxHW + 4];
  const Dir dz = g.dir[2];
  double W[9][NV];
#pragma unroll
  for (int k = 0; k < 8; ++k)
#pragma unroll
    for (int l = 0; l < NV; ++l) W[k][l] = 0.0;
  // the operands of the next plane are loaded once this plane's fluxes are formed, so
  // their latency overlaps the divergence and the stage update
  Pt ct, ch, nt, nh;
  load_pt(i, j, wrap(z0 - 4, dz), z0 - 4 >= z0 ? 7 : 4, ct);
  for (int zl = z0 - 4; zl < z1 + 4; ++zl) {      // uniform over the block
    const int kz = wrap(zl, dz);
    const bool xy = zl >= z0 && zl < z1;
    double f[3][NV], fh[3][NV];
    fluxes(i, j, kz, xy ? 7 : 4, ct, f);
#pragma unroll
    for (int l = 0; l < NV; ++l) W[8][l] = f[2][l];
    if (xy && hmask) fluxes(hi, hj, kz, hmask, ch, fh);
    // the operands of the output plane zl - 4 that the finish reads (RK buffers, the
    // Kolmogorov term's density and strength) are loaded with the next plane's, once this
    // plane's fluxes are formed, so their latency overlaps the divergence
    const bool outp = zl >= z0 + 4;
    double ra[NV], rb[NV], rho_k = 0.0, kk_k = 0.0;
    if (outp) {
      const long p = i + (long)g.n[0] * j + n01 * (zl - 4);
      if (kstr) {
        rho_k = qi[p];
        kk_k = kstr[p];
      }
      if constexpr (RK) {
#pragma unroll
        for (int l = 0; l < NV; ++l) {
          const long ii = inst * NV * g.N + l * g.N + p;
          ra[l] = rk_stage == 1 ? q[ii] : (rk_stage == 2 || rk_stage == 3) ? b1[ii] : 0.0;
          rb[l] = rk_stage >= 2 && b2 ? b2[ii] : 0.0;
        }
      }
    }
    // this plane's operands are spent: load the next plane's now
    const int zn = zl + 1;
    const bool nxy = zn >= z0 && zn < z1;
    if (zn < z1 + 4) {
      load_pt(i, j, wrap(zn, dz), nxy ? 7 : 4, nt);
      if (nxy && hmask) load_pt(hi, hj, wrap(zn, dz), hmask, nh);
    }
    if (zl >= z0 + 4) {
      const int zo = zl - 4;
      const long p = i + (long)g.n[0] * j + n01 * zo;
      double r[NV];
#pragma unroll
      for (int l = 0; l < NV; ++l) {
        double out = 0.0;
        out += c1 * (W[5][l] - W[3][l]);
        out += c2 * (W[6][l] - W[2][l]);
        out += c3 * (W[7][l] - W[1][l]);
        out += c4 * (W[8][l] - W[0][l]);
        const double v = ring[zo & 3][l][t] + out;
        r[l] = 0.0 - v;
      }
      int ijk[3] = {i, j, zo};
      const double jac = jacobian<3>(g, ijk);
#pragma unroll
      for (int l = 0; l < NV; ++l) r[l] = r[l] * jac;
      if (kstr) r[1] = r[1] + rho_k * kk_k;
      if (!RK) {
#pragma unroll
        for (int l = 0; l < NV; ++l) rhs[inst * NV * g.N + l * g.N + p] = r[l];
      } else {
#pragma unroll
        for (int l = 0; l < NV; ++l) {
          const long ii = inst * NV * g.N + l * g.N + p;
          const double rl = r[l];
          double qn;
          switch (rk_stage) {
            case 1: {
              const double q0 = ra[l];
              if (b1) b1[ii] = q0;
              if (b2) b2[ii] = q0 + div6(dt * rl);
              qn = q0 + dt * rl / 2.0;
              break;
            }
            case 2:
              if (b2) b2[ii] = rb[l] + div3(dt * rl);
              qn = ra[l] + dt * rl / 2.0;
              break;
            case 3:
              if (b2) b2[ii] = rb[l] + div3(dt * rl);
              qn = ra[l] + dt * rl;
              break;
            default:
              qn = rb[l] + div6(dt * rl);
              break;
          }
          yout[ii] = qn;
          if (copy) copy[ii] = qn;
        }
      }
    }
    if (xy) {
      __syncthreads();          // readers of the previous plane's fluxes are done
#pragma unroll
      for (int l = 0; l < NV; ++l) {
        sfx[l][ty][tx + 4] = f[0][l];
        sfy[l][ty + 4][tx] = f[1][l];
      }
      if (hmask == 1) {
#pragma unroll
        for (int l = 0; l < NV; ++l) sfx[l][hy][hx] = fh[0][l];
      } else if (hmask == 2) {
#pragma unroll
        for (int l = 0; l < NV; ++l) sfy[l][hy][hx] = fh[1][l];
      }
      __syncthreads();
#pragma unroll
      for (int l = 0; l < NV; ++l) {
        const double* rx = &sfx[l][ty][tx + 4];
        const double* cy = &sfy[l][ty + 4][tx];
        double ox = 0.0, oy = 0.0;
        ox += c1 * (rx[1] - rx[-1]);
        ox += c2 * (rx[2] - rx[-2]);
        ox += c3 * (rx[3] - rx[-3]);
        ox += c4 * (rx[4] - rx[-4]);
        oy += c1 * (cy[BX] - cy[-BX]);
        oy += c2 * (cy[2 * BX] - cy[-2 * BX]);
        oy += c3 * (cy[3 * BX] - cy[-3 * BX]);
        oy += c4 * (cy[4 * BX] - cy[-4 * BX]);
        ring[zl & 3][l][t] = ox + oy;   // each thread reads back only its own slots
      }
    }
#pragma unroll
    for (int k = 0; k < 8; ++k)
#pragma unroll
      for (int l = 0; l < NV; ++l) W[k][l] = W[k + 1][l];
    ct = nt;
    if (nxy && hmask) ch = nh;
  }
}

// The z pass of the adjoint's second derivative pass (D-dagger of the viscous block
// T1[z] components 1..4) fused with k_finish's adjoint operations, for 3-D regions whose
// patches are at most one Kolmogorov forcing over the whole domain, with no dissipation
// and no far field (the chain rule then sits in the finishing pass).  One thread per
// z-line as in k_deriv_line, the 9-point windows of the four components in registers.
// At each point t2_a = A_a + (D-dagger_z T1)_a, where A = S_x + S_y was accumulated by
// the x and y passes in S slot 0 (the sum k_finish forms, in its order); then the chain
// rule, the Jacobian scaling and the Kolmogorov term, as k_finish applies them.  With
// rk_stage > 0 the point's RK4 adjoint update follows at once (k_rk_adjoint's
// arithmetic; lam is then y, read at the point before it is written) and rhs is not
// written.
__global__ void __launch_bounds__(kBlock) k_adj_finish_z(Geo g, const double* __restrict__ T1z,
                                                        const double* __restrict__ A, long t_inst,
                                                        const double* __restrict__ q,
                                                        const double* lam,
                                                        const double* __restrict__ kstr,
                                                        double* __restrict__ rhs, int rk_stage,
                                                        double dt, double* y, double* b1,
                                                        double* b2) {
  constexpr int ND = 3, NV = 5, NW = 4;
  const Dir dd = g.dir[2];
  const long nlines = g.N / dd.len;
  const long line = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (line >= nlines) return;
  const long inst = blockIdx.z;
  const long n0 = g.n[0];
  const long i = line % n0, o = line / n0;
  const long base = i + o * n0;
  const long sd = dd.stride;
  const double* x = T1z + inst * t_inst + base;
  const double* ai = A + inst * t_inst + base;
  const double* qi = q + inst * NV * g.N;
  const double* li = lam + inst * NV * g.N;
  double* ri = rhs + inst * NV * g.N;
  const int z0 = blockIdx.y * kLineChunk;
  const int z1 = min(z0 + kLineChunk, dd.len);
  const Op& op = g.ops[SLOT_ADJ];
  const double c1 = op.c[kMaxHW + 1], c2 = op.c[kMaxHW + 2], c3 = op.c[kMaxHW + 3],
               c4 = op.c[kMaxHW + 4];
  double Q[NW][9];
#pragma unroll
  for (int a = 0; a < NW; ++a)
#pragma unroll
    for (int k = 0; k < 8; ++k) Q[a][k] = x[a * g.N + wrap(z0 - 4 + k, dd) * sd];
  for (int z = z0; z < z1; ++z) {
    const long p = base + z * sd;
    double t2[NW];
#pragma unroll
    for (int a = 0; a < NW; ++a) {
      Q[a][8] = x[a * g.N + wrap(z + 4, dd) * sd];
      double out = 0.0;
      out += c1 * (Q[a][5] - Q[a][3]);
      out += c2 * (Q[a][6] - Q[a][2]);
      out += c3 * (Q[a][7] - Q[a][1]);
      out += c4 * (Q[a][8] - Q[a][0]);
      t2[a] = ai[a * g.N + z * sd] + out;
#pragma unroll
      for (int k = 0; k < 8; ++k) Q[a][k] = Q[a][k + 1];
    }
    double v[NV], r[NV];
    load_q<ND>(qi, g.N, p, v);
#pragma unroll
    for (int l = 0; l < NV; ++l) r[l] = ri[l * g.N + p];
    adj_chain_t2<ND>(g, v, t2, -1.0, r);
    int ijk[3] = {(int)i, (int)o, z};
    const double jac = jacobian<ND>(g, ijk);
#pragma unroll
    for (int l = 0; l < NV; ++l) r[l] = r[l] * jac;
    if (kstr) r[0] = r[0] - li[g.N + p] * kstr[p];
    if (rk_stage == 0) {
#pragma unroll
      for (int l = 0; l < NV; ++l) ri[l * g.N + p] = r[l];
    } else {
#pragma unroll
      for (int l = 0; l < NV; ++l) {
        const long ii = inst * NV * g.N + l * g.N + p;
        const double rl = r[l];
        switch (rk_stage) {
          case 4: {
            const double l0 = y[ii];
            b1[ii] = l0;
            b2[ii] = l0 - div6(dt * rl);
            y[ii] = l0 - dt * rl / 2.0;
            break;
          }
          case 3:
            b2[ii] = b2[ii] - div3(dt * rl);
            y[ii] = b1[ii] - dt * rl / 2.0;
            break;
          case 2:
            b2[ii] = b2[ii] - div3(dt * rl);
            y[ii] = b1[ii] - dt * rl;
            break;
          default:
            y[ii] = b2[ii] - div6(dt * rl);
            break;
        }
      }
    }
  }
}

// Element access at a 32-bit byte offset from a wave-uniform base pointer: the address
// form of a global load / store with a scalar base and a 32-bit vector offset, so a
// thread keeps one 32-bit offset per point instead of a 64-bit address per component.
// Callers keep byte offsets below 2^32 (fewer than 2^29 points per component array).
__device__ __forceinline__ double ld8(const double* base, unsigned off) {
  return *reinterpret_cast<const double*>(reinterpret_cast<const char*>(base) + off);
}
__device__ __forceinline__ void st8(double* base, unsigned off, double v) {
  *reinterpret_cast<double*>(reinterpret_cast<char*>(base) + off) = v;
}

// The first half of the adjoint rhs of a viscous 3-D region periodic in every direction,
// in one z march: a BX x BY column tile marches along z over a chunk of ZC planes, each
// thread keeping lambda of its column for the 9 planes of the z stencil in registers; the
// plane being worked on is staged in LDS with its 4-point x / y halo (one halo point per
// thread), so T1_d = D-dagger_d lambda of all three directions is formed at the point
// without going through HBM, one direction at a time, each used as soon as it is formed.
// With q and the stress / heat-flux slots S at the point (k_grad3d_periodic<.., true>),
// adj_visc_dir gives the Jacobian part of the rhs -- written to rhs -- and the
// second-partial block sum_i B_ij^T T1_i, written to components 1..4 of T1_j (where
// k_adj3d_finish_march reads them).  The derivative sums are those of k_deriv_line /
// k_deriv_row and the point operations those of k_adj_jac_visc2, so the values are the
// same.  The halo point of the next plane and the lambda entering the window are loaded
// one plane ahead.  The window of the last NZL components lives in an LDS ring of 9
// planes (each thread reads back only its own column), the others in registers.
template <int BX, int BY, int ZC, bool POWER>
__global__ void __launch_bounds__(BX * BY, 512 / (BX * BY)) k_adj3d_jac_march(Geo g, const double* __restrict__ q,
                                                              const double* __restrict__ lam,
                                                              const double* __restrict__ S,
                                                              double* __restrict__ rhs,
                                                              double* __restrict__ T1) {
  constexpr int NV = 5, NW = 4, EXW = BX + 8, NT = BX * BY, NZL = 2, NZR = NV - NZL;
  static_assert(8 * (BX + BY) <= NT, "at most one halo point per thread");
  __shared__ double slx[NV][BY][EXW];
  __shared__ double sly[NV][BY + 8][BX];
  __shared__ double zring[9][NZL][NT];
  int x0, y0, z0;
  if (!march_unit<BX, BY, ZC>(g, x0, y0, z0)) return;
  const int z1 = min(z0 + ZC, g.n[2]);
  const long inst = blockIdx.z;
  const double* qi = q + inst * NV * g.N;
  const double* li = lam + inst * NV * g.N;
  const double* Si = S + inst * 12 * g.N;
  double* ri = rhs + inst * NV * g.N;
  double* Ti = T1 + inst * 3 * NV * g.N;
  const int t = threadIdx.x, tx = t % BX, ty = t / BX;
  const int i = x0 + tx, j = y0 + ty;
  int hx, hy, hi, hj, hmask;
  march_halo<BX, BY>(g, t, x0, y0, hx, hy, hi, hj, hmask);
  // byte offsets of the thread's column and halo column, and of one plane
  const unsigned n0 = (unsigned)g.n[0], plane8 = 8u * n0 * (unsigned)g.n[1];
  const unsigned col8 = 8u * ((unsigned)i + n0 * (unsigned)j);
  const unsigned hcol8 = 8u * ((unsigned)hi + n0 * (unsigned)hj);
  auto load_l = [&](unsigned off, double* v) {
#pragma unroll
    for (int l = 0; l < NV; ++l) v[l] = ld8(li + l * g.N, off);
  };
  const Op& op = g.ops[SLOT_ADJ];
  const double c1 = op.c[kMaxHW + 1], c2 = op.c[kMaxHW + 2], c3 = op.c[kMaxHW + 3], c4 = op.c[kMaxHW + 4];
  const Dir dz = g.dir[2];
  // window: components [0, NZR) in W, [NZR, NV) in ring slot (plane - (z0 - 4)) mod 9
  double W[9][NZR];
#pragma unroll
  for (int k = 0; k < 8; ++k) {
    double v[NV];
    load_l(col8 + plane8 * (unsigned)wrap(z0 - 4 + k, dz), v);
#pragma unroll
    for (int l = 0; l < NZR; ++l) W[k][l] = v[l];
#pragma unroll
    for (int l = NZR; l < NV; ++l) zring[k][l - NZR][t] = v[l];
  }
  double nl[NV], nh[NV] = {};
  load_l(col8 + plane8 * (unsigned)wrap(z0 + 4, dz), nl);
  if (hmask) load_l(hcol8 + plane8 * (unsigned)z0, nh);
  int rb = 0;                            // ring slot of plane zo - 4
  for (int zo = z0; zo < z1; ++zo) {     // uniform over the block
    double hv[NV], wc[NV];
    const int r8 = rb + 8 < 9 ? rb + 8 : rb - 1;
#pragma unroll
    for (int l = 0; l < NV; ++l) {
      if (l < NZR) W[8][l] = nl[l];
      else zring[r8][l - NZR][t] = nl[l];
      hv[l] = nh[l];
    }
    const int r4 = rb + 4 < 9 ? rb + 4 : rb - 5;
#pragma unroll
    for (int l = 0; l < NV; ++l) wc[l] = l < NZR ? W[4][l] : zring[r4][l - NZR][t];
    if (zo + 1 < z1) {
      load_l(col8 + plane8 * (unsigned)wrap(zo + 5, dz), nl);
      if (hmask) load_l(hcol8 + plane8 * (unsigned)(zo + 1), nh);
    }
    const unsigned p8 = col8 + plane8 * (unsigned)zo;
    double v[NV], sv[9];
#pragma unroll
    for (int l = 0; l < NV; ++l) v[l] = ld8(qi + l * g.N, p8);
#pragma unroll
    for (int k = 0; k < 9; ++k) sv[k] = ld8(Si + k * g.N, p8);
    __syncthreads();            // readers of the previous plane are done
#pragma unroll
    for (int l = 0; l < NV; ++l) {
      slx[l][ty][tx + 4] = wc[l];
      sly[l][ty + 4][tx] = wc[l];
    }
    if (hmask == 1) {
#pragma unroll
      for (int l = 0; l < NV; ++l) slx[l][hy][hx] = hv[l];
    } else if (hmask == 2) {
#pragma unroll
      for (int l = 0; l < NV; ++l) sly[l][hy][hx] = hv[l];
    }
    __syncthreads();
    const int ijk[3] = {i, j, zo};
    AdjPoint<3> P;
    adj_point_init_stress<POWER>(g, v, sv, ijk, P);
    double r[NV], diff[3][NW];
#pragma unroll
    for (int l = 0; l < NV; ++l) r[l] = 0.0;
#pragma unroll
    for (int d = 0; d < 3; ++d)
#pragma unroll
      for (int a = 0; a < NW; ++a) diff[d][a] = 0.0;
#pragma unroll
    for (int d = 0; d < 3; ++d) {
      double t1[NV];
#pragma unroll
      for (int l = 0; l < NV; ++l) {
        double o = 0.0;
        if (d == 0) {
          const double* rx = &slx[l][ty][tx + 4];
          o += c1 * (rx[1] - rx[-1]);
          o += c2 * (rx[2] - rx[-2]);
          o += c3 * (rx[3] - rx[-3]);
          o += c4 * (rx[4] - rx[-4]);
        } else if (d == 1) {
          const double* cy = &sly[l][ty + 4][tx];
          o += c1 * (cy[BX] - cy[-BX]);
          o += c2 * (cy[2 * BX] - cy[-2 * BX]);
          o += c3 * (cy[3 * BX] - cy[-3 * BX]);
          o += c4 * (cy[4 * BX] - cy[-4 * BX]);
        } else if (l < NZR) {
          o += c1 * (W[5][l] - W[3][l]);
          o += c2 * (W[6][l] - W[2][l]);
          o += c3 * (W[7][l] - W[1][l]);
          o += c4 * (W[8][l] - W[0][l]);
        } else {
          auto zw = [&](int k) {
            const int sl = rb + k < 9 ? rb + k : rb + k - 9;
            return zring[sl][l - NZR][t];
          };
          o += c1 * (zw(5) - zw(3));
          o += c2 * (zw(6) - zw(2));
          o += c3 * (zw(7) - zw(1));
          o += c4 * (zw(8) - zw(0));
        }
        t1[l] = o;
      }
      adj_visc_dir<3, POWER>(P, g.gamma, d, t1, r, diff);
    }
#pragma unroll
    for (int l = 0; l < NV; ++l) st8(ri + l * g.N, p8, r[l]);
#pragma unroll
    for (int d = 0; d < 3; ++d)
#pragma unroll
      for (int a = 0; a < NW; ++a) st8(Ti + (d * NV + 1 + a) * g.N, p8, diff[d][a]);
#pragma unroll
    for (int k = 0; k < 8; ++k)
#pragma unroll
      for (int l = 0; l < NZR; ++l) W[k][l] = W[k + 1][l];
    rb = rb + 1 < 9 ? rb + 1 : 0;
  }
}

// The second half, after k_adj3d_jac_march: t2 = (D-dagger_x T1_x + D-dagger_y T1_y) +
// D-dagger_z T1_z over components 1..4 (the x / y operands of the plane staged in LDS with
// their halo, one halo point per thread, the z operand in a 9-plane register window),
// then, on the Jacobian part read back from rhs, k_adj_finish_z's operations: the chain
// rule, the Jacobian scaling and the Kolmogorov term, and with rk_stage > 0 the RK4
// adjoint update in place (y = lambda; only the point's own lambda is read), rhs not
// written.  The sums and their order are those of k_deriv_row, k_deriv_line (x store, y
// add) and k_adj_finish_z, so the values are the same.
template <int BX, int BY, int ZC>
__global__ void __launch_bounds__(BX * BY, 512 / (BX * BY)) k_adj3d_finish_march(
    Geo g, const double* __restrict__ T1, const double* __restrict__ q, const double* lam,
    const double* __restrict__ kstr, double* __restrict__ rhs, int rk_stage, double dt, double* y,
    double* b1, double* b2) {
  constexpr int NV = 5, NW = 4, EXW = BX + 8;
  static_assert(8 * (BX + BY) <= BX * BY, "at most one halo point per thread");
  __shared__ double snx[NW][BY][EXW];
  __shared__ double sny[NW][BY + 8][BX];
  int x0, y0, z0;
  if (!march_unit<BX, BY, ZC>(g, x0, y0, z0)) return;
  const int z1 = min(z0 + ZC, g.n[2]);
  const long inst = blockIdx.z;
  const long io = inst * NV * g.N;
  const double* qi = q + io;
  const double* li = lam + io;
  const double* Tx = T1 + inst * 3 * NV * g.N + g.N;            // T1_x components 1..4
  const double* Ty = Tx + (long)NV * g.N;
  const double* Tz = Ty + (long)NV * g.N;
  double* ri = rhs + io;
  const int t = threadIdx.x, tx = t % BX, ty = t / BX;
  const int i = x0 + tx, j = y0 + ty;
  int hx, hy, hi, hj, hmask;
  march_halo<BX, BY>(g, t, x0, y0, hx, hy, hi, hj, hmask);
  const unsigned n0 = (unsigned)g.n[0], plane8 = 8u * n0 * (unsigned)g.n[1];
  const unsigned col8 = 8u * ((unsigned)i + n0 * (unsigned)j);
  const unsigned hcol8 = 8u * ((unsigned)hi + n0 * (unsigned)hj);
  const double* Th = hmask == 1 ? Tx : Ty;
  auto load4 = [&](const double* src, unsigned off, double* v) {
#pragma unroll
    for (int a = 0; a < NW; ++a) v[a] = ld8(src + a * g.N, off);
  };
  const Op& op = g.ops[SLOT_ADJ];
  const double c1 = op.c[kMaxHW + 1], c2 = op.c[kMaxHW + 2], c3 = op.c[kMaxHW + 3], c4 = op.c[kMaxHW + 4];
  const Dir dz = g.dir[2];
  double Z[9][NW];
#pragma unroll
  for (int k = 0; k < 8; ++k) load4(Tz, col8 + plane8 * (unsigned)wrap(z0 - 4 + k, dz), Z[k]);
  double nz[NW], nx[NW], ny[NW], nh[NW] = {};
  load4(Tz, col8 + plane8 * (unsigned)wrap(z0 + 4, dz), nz);
  load4(Tx, col8 + plane8 * (unsigned)z0, nx);
  load4(Ty, col8 + plane8 * (unsigned)z0, ny);
  if (hmask) load4(Th, hcol8 + plane8 * (unsigned)z0, nh);
  for (int zo = z0; zo < z1; ++zo) {     // uniform over the block
    double xv[NW], yv[NW], hv[NW];
#pragma unroll
    for (int a = 0; a < NW; ++a) {
      Z[8][a] = nz[a];
      xv[a] = nx[a];
      yv[a] = ny[a];
      hv[a] = nh[a];
    }
    if (zo + 1 < z1) {
      const unsigned n8 = col8 + plane8 * (unsigned)(zo + 1);
      load4(Tz, col8 + plane8 * (unsigned)wrap(zo + 5, dz), nz);
      load4(Tx, n8, nx);
      load4(Ty, n8, ny);
      if (hmask) load4(Th, hcol8 + plane8 * (unsigned)(zo + 1), nh);
    }
    const unsigned p8 = col8 + plane8 * (unsigned)zo;
    double v[NV], r[NV];
#pragma unroll
    for (int l = 0; l < NV; ++l) {
      v[l] = ld8(qi + l * g.N, p8);
      r[l] = ld8(ri + l * g.N, p8);
    }
    // the point's RK operands and Kolmogorov-term values, loaded here so that their
    // latency overlaps the derivative work (stage 4 reads lambda before it writes it)
    double ra[NV], rb[NV], lk = 0.0, kk = 0.0;
    if (kstr) {
      lk = ld8(li + g.N, p8);
      kk = ld8(kstr, p8);
    }
#pragma unroll
    for (int l = 0; l < NV; ++l) {
      const long c = io + l * g.N;
      ra[l] = rk_stage == 4 ? ld8(y + c, p8) : (rk_stage == 2 || rk_stage == 3) ? ld8(b1 + c, p8) : 0.0;
      rb[l] = rk_stage >= 1 && rk_stage <= 3 ? ld8(b2 + c, p8) : 0.0;
    }
    __syncthreads();            // readers of the previous plane are done
#pragma unroll
    for (int a = 0; a < NW; ++a) {
      snx[a][ty][tx + 4] = xv[a];
      sny[a][ty + 4][tx] = yv[a];
    }
    if (hmask == 1) {
#pragma unroll
      for (int a = 0; a < NW; ++a) snx[a][hy][hx] = hv[a];
    } else if (hmask == 2) {
#pragma unroll
      for (int a = 0; a < NW; ++a) sny[a][hy][hx] = hv[a];
    }
    __syncthreads();
    double t2[NW];
#pragma unroll
    for (int a = 0; a < NW; ++a) {
      const double* rx = &snx[a][ty][tx + 4];
      const double* cy = &sny[a][ty + 4][tx];
      double ox = 0.0, oy = 0.0, oz = 0.0;
      ox += c1 * (rx[1] - rx[-1]);
      ox += c2 * (rx[2] - rx[-2]);
      ox += c3 * (rx[3] - rx[-3]);
      ox += c4 * (rx[4] - rx[-4]);
      oy += c1 * (cy[BX] - cy[-BX]);
      oy += c2 * (cy[2 * BX] - cy[-2 * BX]);
      oy += c3 * (cy[3 * BX] - cy[-3 * BX]);
      oy += c4 * (cy[4 * BX] - cy[-4 * BX]);
      oz += c1 * (Z[5][a] - Z[3][a]);
      oz += c2 * (Z[6][a] - Z[2][a]);
      oz += c3 * (Z[7][a] - Z[1][a]);
      oz += c4 * (Z[8][a] - Z[0][a]);
      t2[a] = (ox + oy) + oz;
    }
    adj_chain_t2<3>(g, v, t2, -1.0, r);
    const int ijk[3] = {i, j, zo};
    const double jac = jacobian<3>(g, ijk);
#pragma unroll
    for (int l = 0; l < NV; ++l) r[l] = r[l] * jac;
    if (kstr) r[0] = r[0] - lk * kk;
    if (rk_stage == 0) {
#pragma unroll
      for (int l = 0; l < NV; ++l) st8(ri + l * g.N, p8, r[l]);
    } else {
#pragma unroll
      for (int l = 0; l < NV; ++l) {
        const long c = io + l * g.N;
        double* yc = y + c;
        double* b1c = b1 + c;
        double* b2c = b2 + c;
        const double rl = r[l];
        switch (rk_stage) {
          case 4: {
            const double l0 = ra[l];
            st8(b1c, p8, l0);
            st8(b2c, p8, l0 - div6(dt * rl));
            st8(yc, p8, l0 - dt * rl / 2.0);
            break;
          }
          case 3:
            st8(b2c, p8, rb[l] - div3(dt * rl));
            st8(yc, p8, ra[l] - dt * rl / 2.0);
            break;
          case 2:
            st8(b2c, p8, rb[l] - div3(dt * rl));
            st8(yc, p8, ra[l] - dt * rl);
            break;
          default:
            st8(yc, p8, rb[l] - div6(dt * rl));
            break;
        }
      }
    }
#pragma unroll
    for (int k = 0; k < 8; ++k)
#pragma unroll
      for (int a = 0; a < NW; ++a) Z[k][a] = Z[k + 1][a];
  }
}

// The same operator along a periodic x direction: a block stages `segs` (1 or
// kRowSegs) runs of 256 consecutive points of one component -- segments of one row, or
// whole rows when rows are shorter -- with the 4-point wrap-around halo of each row
// segment in LDS, and applies the interior stencil from there; all its loads are issued
// before the one barrier.  Requires rows of a length that divides 256 or is a multiple
// of it, and 256 segs | N.  Modes as for k_deriv_line, plus 3: store 0 - result.
constexpr int kRowSegs = 4;
__global__ void __launch_bounds__(kBlock) k_deriv_row(Geo g, int slot, const double* __restrict__ src,
                                                     long src_inst, double* __restrict__ dst,
                                                     long dst_inst, int ncomp, int mode, int segs) {
  static_assert(kBlock == 256, "one staged point per thread and run");
  constexpr int SW = kBlock + 8 * kBlock / 16;        // (L + 8) per segment, L >= 16
  const Dir dx = g.dir[0];
  const int L = dx.len < kBlock ? dx.len : kBlock;   // segment length
  __shared__ double s[kRowSegs][SW];
  const int inst = blockIdx.z / ncomp, c = blockIdx.z % ncomp;
  const double* x = src + inst * src_inst + (long)c * g.N;
  double* y = dst + inst * dst_inst + (long)c * g.N;
  const int t = threadIdx.x;
  const int seg = t / L, xi = t % L;
#pragma unroll
  for (int u = 0; u < kRowSegs; ++u) {
    if (u >= segs) break;
    const long p = ((long)blockIdx.x * segs + u) * kBlock + t;
    const int xr = (int)(p % dx.len);                 // index along the row
    const long row0 = p - xr;
    double* sg = s[u] + seg * (L + 8);
    sg[4 + xi] = x[p];
    if (xi < 4) sg[xi] = x[row0 + wrap(xr - 4, dx)];
    if (xi >= L - 4) sg[xi + 8] = x[row0 + wrap(xr + 4, dx)];
  }
  __syncthreads();
  const Op& op = g.ops[slot];
  const double c1 = op.c[kMaxHW + 1], c2 = op.c[kMaxHW + 2], c3 = op.c[kMaxHW + 3],
               c4 = op.c[kMaxHW + 4];
#pragma unroll
  for (int u = 0; u < kRowSegs; ++u) {
    if (u >= segs) break;
    const long p = ((long)blockIdx.x * segs + u) * kBlock + t;
    const double* r = s[u] + seg * (L + 8) + 4 + xi;
    double out = 0.0;
    out += c1 * (r[1] - r[-1]);
    out += c2 * (r[2] - r[-2]);
    out += c3 * (r[3] - r[-3]);
    out += c4 * (r[4] - r[-4]);
    if (mode == 0) {
      y[p] = out;
    } else if (mode == 3) {
      y[p] = 0.0 - out;
    } else {
      double v = y[p] + out;
      if (mode == 2) v = 0.0 - v;
      y[p] = v;
    }
  }
}

// Whether k_deriv_row (d = 0) or k_deriv_line (d >= 1) serves direction d: periodic,
// and a grid that fits the kernel.
static bool deriv_line_applies(const Geo& g, int d) {
  if (!g.dir[d].periodic || g.dir[d].len < 16) return false;
  if (d > 0) return true;
  const int n0 = g.dir[0].len;
  return ((n0 % kBlock == 0) || (kBlock % n0 == 0)) && g.N % kBlock == 0;
}

// Launches k_deriv_row (d = 0) or k_deriv_line (d >= 1) where one serves direction d;
// false otherwise (the caller then uses the generic kernel).
static bool launch_deriv_line(const Geo& g, int d, int slot, const double* src, long src_inst,
                              double* dst, long dst_inst, int ncomp, int mode, hipStream_t st) {
  if (!deriv_line_applies(g, d)) return false;
  if (d == 0) {
    const int segs = g.N % ((long)kBlock * kRowSegs) == 0 ? kRowSegs : 1;
    const dim3 grid((unsigned)(g.N / ((long)kBlock * segs)), 1, (unsigned)(g.ninst * ncomp));
    k_deriv_row<<<grid, kBlock, 0, st>>>(g, slot, src, src_inst, dst, dst_inst, ncomp, mode, segs);
    return true;
  }
  const long nlines = g.N / g.dir[d].len;
  const dim3 grid((unsigned)((nlines + kBlock - 1) / kBlock),
                  (unsigned)((g.dir[d].len + kLineChunk - 1) / kLineChunk),
                  (unsigned)(g.ninst * ncomp));
  k_deriv_line<<<grid, kBlock, 0, st>>>(g, d, slot, src, src_inst, dst, dst_inst, ncomp, mode);
  return true;
}

// ------------------------------------------------------------------ launchers
#define MGD_DISPATCH(ND_, ...)                                  \
  switch (ND_) {                                                \
    case 1: { constexpr int ND = 1; __VA_ARGS__; } break;        \
    case 2: { constexpr int ND = 2; __VA_ARGS__; } break;        \
    default: { constexpr int ND = 3; __VA_ARGS__; } break;       \
  }

static dim3 grid_pts(const Geo& g) {
  return dim3((unsigned)((g.N + kBlock - 1) / kBlock), (unsigned)g.ninst);
}

// k_prim + k_grad, or the one-pass periodic 3-D gradient kernel where it applies
// (3-D, every direction periodic, x / y sizes multiples of its 32 x 16 tile).
constexpr int kG3X = 32, kG3Y = 16;
static bool grad3d_applies(const Geo& g) {
  return g.nd == 3 && g.dir[0].periodic && g.dir[1].periodic && g.dir[2].periodic &&
         g.n[0] % kG3X == 0 && g.n[1] % kG3Y == 0 && g.nv == 5;
}

// x / y flux divergence kernel of 3-D grids periodic in x and y with sizes that are
// multiples of its 32 x 16 tile.
constexpr int kFD3X = 32, kFD3Y = 16;
static bool fluxdiv3d_applies(const Geo& g) {
  return g.nd == 3 && g.nv == 5 && g.dir[0].periodic && g.dir[1].periodic &&
         g.n[0] % kFD3X == 0 && g.n[1] % kFD3Y == 0;
}

// grid of the periodic gradient kernel: (tile, z chunk) units padded to a multiple of 8
// The z chunk (planes per work unit) of the 3-D marches, g.zchunk: 32, 64 or 128, and the
// x extent of their tiles, g.mtile: 16 (16 x 16 tiles of 256 threads, two work-groups
// per CU) or 32 (32 x 16 tiles of 512 threads, one per CU: fewer halo points per point) --
// the autotuner's choices; no value depends on them.  f is called with them as
// compile-time values.
// The forward march and the stress pass use (g.zchunk, g.mtile), the adjoint marches
// (g.zchunk_adj, g.mtile_adj).
static int zchunk_of(const Geo& g, bool adj = false) {
  const int z = adj ? g.zchunk_adj : g.zchunk;
  return z == 32 || z == 128 ? z : 64;
}
static int mtile_of(const Geo& g, bool adj = false) {
  return (adj ? g.mtile_adj : g.mtile) == 32 && g.n[0] % 32 == 0 ? 32 : 16;
}
template <class F>
static void with_zchunk(const Geo& g, bool adj, F&& f) {
  switch (zchunk_of(g, adj)) {
    case 32: f(std::integral_constant<int, 32>{}); break;
    case 128: f(std::integral_constant<int, 128>{}); break;
    default: f(std::integral_constant<int, 64>{}); break;
  }
}
template <class F>
static void with_march(const Geo& g, bool adj, F&& f) {
  with_zchunk(g, adj, [&](auto Z) {
    if (mtile_of(g, adj) == 32) f(std::integral_constant<int, 32>{}, Z);
    else f(std::integral_constant<int, 16>{}, Z);
  });
}

static dim3 grad3d_grid(const Geo& g) {
  const int zc = zchunk_of(g);
  const int units = (g.n[0] / kG3X) * (g.n[1] / kG3Y) * ((g.n[2] + zc - 1) / zc);
  return dim3(8 * ((units + 7) / 8), 1, g.ninst);
}

template <int ND>
static void launch_gradients(const Geo& g, const double* q, double* w, double* G, const dim3& grid,
                             hipStream_t st) {
  if (ND == 3 && grad3d_applies(g)) {
    with_zchunk(g, false, [&](auto Z) {
      k_grad3d_periodic<kG3X, kG3Y, decltype(Z)::value><<<grad3d_grid(g), kG3X * kG3Y, 0, st>>>(g, q, G);
    });
    return;
  }
  k_prim<ND><<<grid, kBlock, 0, st>>>(g, q, w);
  for (int d = 0; d < ND; ++d) k_grad<ND><<<grid, kBlock, 0, st>>>(g, d, w, G);
}

// The one-march forward kernel: 3-D, periodic in every direction, x / y sizes multiples
// of its 16 x 16 tile; viscous regions also need the stress kernel (grad3d_applies).
constexpr int kFMX = 16, kFMY = 16;
// grid of the 3-D march kernels: (tile, z chunk) units padded to a multiple of 8
static dim3 march_grid(const Geo& g, bool adj) {
  const int zc = zchunk_of(g, adj);
  const int units = (g.n[0] / mtile_of(g, adj)) * (g.n[1] / kFMY) * ((g.n[2] + zc - 1) / zc);
  return dim3(8 * ((units + 7) / 8), 1, g.ninst);
}
static bool fwd3d_march_applies(const Geo& g) {
  return g.nd == 3 && g.nv == 5 && g.dir[0].periodic && g.dir[1].periodic && g.dir[2].periodic &&
         g.n[0] % kFMX == 0 && g.n[1] % kFMY == 0 && (!g.viscous || grad3d_applies(g));
}

bool fwd3d_march_path(const Geo& g, const FinishInfo& fin, int nsrc) {
  return fin.simple && !g.dissipation && nsrc == 0 && fwd3d_march_applies(g);
}

// The two-march adjoint path: viscous, periodic in every direction with the march tile
// sizes, no dissipation, no far field and at most one whole-domain Kolmogorov forcing (the
// chain rule and the forcing then sit in the finishing march).
static bool adj3d_march_path(const Geo& g, const FinishInfo& fin) {
  return g.viscous && !g.dissipation && !fin.has_farfield && fin.simple &&
         fwd3d_march_applies(g) && g.N < (1L << 29);   // 32-bit byte offsets
}

bool stress_reuse_path(const Geo& g, const FinishInfo& fin, int nsrc) {
  return g.viscous && fwd3d_march_path(g, fin, nsrc) && adj3d_march_path(g, fin);
}

void launch_stress3d(const Geo& g, const double* q, double* S, hipStream_t st) {
  const dim3 gg = grad3d_grid(g);
  with_zchunk(g, false, [&](auto Z) {
    if (g.power_law > 0.0)
      k_grad3d_periodic<kG3X, kG3Y, decltype(Z)::value, true, true><<<gg, kG3X * kG3Y, 0, st>>>(g, q, S);
    else
      k_grad3d_periodic<kG3X, kG3Y, decltype(Z)::value, true, false><<<gg, kG3X * kG3Y, 0, st>>>(g, q, S);
  });
}

void launch_fwd3d_march(const Geo& g, const double* q, const double* S, const double* kstr,
                        double* rhs, const RkStage* rk, hipStream_t st) {
  const dim3 gm = march_grid(g, false);
  if (rk && rk->ev_begin) (void)hipEventRecord(rk->ev_begin, st);
  with_march(g, false, [&](auto X, auto Z) {
    constexpr int BX = decltype(X)::value;
    if (rk)
      k_fwd3d_march<BX, kFMY, decltype(Z)::value, true><<<gm, BX * kFMY, 0, st>>>(
          g, q, S, kstr, rhs, rk->stage, rk->dt, rk->yout, rk->b1, rk->b2, rk->copy);
    else
      k_fwd3d_march<BX, kFMY, decltype(Z)::value, false><<<gm, BX * kFMY, 0, st>>>(
          g, q, S, kstr, rhs, 0, 0.0, nullptr, nullptr, nullptr, nullptr);
  });
  if (rk && rk->ev_end) (void)hipEventRecord(rk->ev_end, st);
}

int launch_rhs_forward(const Geo& g, const Patches* P, const double* q, double* rhs, double* w,
                       double* G, double* t, double* F, const double* control, long control_stride,
                       const Source* src, int nsrc, const double* src_amp, long src_inst_stride,
                       const FinishInfo& fin, hipStream_t st, const RkStage* rk) {
  const dim3 grid = grid_pts(g);
  const int has_farfield = fin.has_farfield;
  if (F && fin.simple && !g.dissipation && nsrc == 0 && fwd3d_march_applies(g)) {
    // the whole rhs in one z march; with an RK stage only out of place (rk->yout)
    if (g.viscous) launch_stress3d(g, q, G, st);
    const bool rk_out = rk && rk->yout;
    launch_fwd3d_march(g, q, G, fin.kolmogorov, rhs, rk_out ? rk : nullptr, st);
    return rk_out ? RK_TO_YOUT : RK_NOT_APPLIED;
  }
  bool finished = false;
  MGD_DISPATCH(g.nd, {
    if (g.viscous) launch_gradients<ND>(g, q, w, G, grid, st);
    if (F && ND == 3 && fluxdiv3d_applies(g)) {
      const dim3 gg((g.n[0] / kFD3X) * (g.n[1] / kFD3Y), g.n[2], g.ninst);
      k_fluxdiv3d_xy<kFD3X, kFD3Y><<<gg, kFD3X * kFD3Y, 0, st>>>(g, q, g.viscous ? G : nullptr, F, rhs);
      // the z divergence also finishes the rhs where nothing else follows it
      const bool fuse = fin.simple && !g.dissipation && nsrc == 0 && deriv_line_applies(g, ND - 1);
      const double* fz = F + (long)(ND - 1) * g.nv * g.N;
      if (fuse) {
        const long nlines = g.N / g.dir[ND - 1].len;
        const dim3 gz((unsigned)((nlines + kBlock - 1) / kBlock),
                      (unsigned)((g.dir[ND - 1].len + kLineChunk - 1) / kLineChunk), (unsigned)g.ninst);
        k_fwd_finish_z<<<gz, kBlock, 0, st>>>(g, fz, (long)ND * g.nv * g.N, q, fin.kolmogorov, rhs,
                                              rk ? rk->stage : 0, rk ? rk->dt : 0.0,
                                              rk ? rk->y : nullptr, rk ? rk->b1 : nullptr,
                                              rk ? rk->b2 : nullptr, rk ? rk->copy : nullptr);
        finished = true;
      } else if (!launch_deriv_line(g, ND - 1, SLOT_FIRST, fz, (long)ND * g.nv * g.N, rhs,
                                    (long)g.nv * g.N, g.nv, 2, st)) {
        k_div<ND><<<grid, kBlock, 0, st>>>(g, ND - 1, F, rhs);
      }
    } else if (F) {
      k_flux_all<ND><<<grid, kBlock, 0, st>>>(g, q, g.viscous ? G : nullptr, F);
      for (int d = 0; d < ND; ++d) {
        const int mode = d == 0 ? (ND == 1 ? 3 : 0) : d == ND - 1 ? 2 : 1;
        if (!launch_deriv_line(g, d, SLOT_FIRST, F + (long)d * g.nv * g.N,
                                         (long)ND * g.nv * g.N, rhs, (long)g.nv * g.N, g.nv,
                                         mode, st))
          k_div<ND><<<grid, kBlock, 0, st>>>(g, d, F, rhs);
      }
    } else {
      for (int d = 0; d < ND; ++d)
        k_flux_div<ND><<<grid, kBlock, 0, st>>>(g, d, q, g.viscous ? G : nullptr, rhs);
    }
    if (g.dissipation) {
      for (int d = 0; d < ND; ++d) {
        k_diss_a<ND><<<grid, kBlock, 0, st>>>(g, d, q, t, rhs, g.diss_amount);
        if (!g.composite) k_diss_b<ND><<<grid, kBlock, 0, st>>>(g, d, t, rhs, g.diss_amount);
      }
    }
    if (g.viscous && has_farfield)
      k_ff_visc_terms<ND><<<grid, kBlock, 0, st>>>(g, MGD_FORWARD, P, q, nullptr, G);
    if (!finished)
      k_finish<ND><<<grid, kBlock, 0, st>>>(g, MGD_FORWARD, P, q, nullptr, G, rhs, control,
                                            control_stride, nullptr, 0, 0.0, src, nsrc, src_amp,
                                            src_inst_stride, nullptr);
  });
  return finished && rk != nullptr ? RK_IN_PLACE : RK_NOT_APPLIED;
}

bool launch_rhs_adjoint(const Geo& g, const Patches* P, const double* q, const double* lam,
                        double* rhs, double* w, double* G, double* t, double* T1, double* S,
                        const double* aforce, long aforce_stride, double forcing_factor,
                        const FinishInfo& fin, hipStream_t st, const RkStage* rk) {
  const dim3 grid = grid_pts(g);
  const int has_farfield = fin.has_farfield;
  if (adj3d_march_path(g, fin)) {
    // stress / heat-flux pass (9 slots of G; skipped where the caller's G already holds
    // them), then the two marches; the second finishes the rhs (and the stage)
    (void)w;
    (void)grid;
    if (!(rk && rk->stress_ready)) launch_stress3d(g, q, G, st);
    const dim3 gm = march_grid(g, true);
    with_march(g, true, [&](auto X, auto Z) {
      constexpr int BX = decltype(X)::value, ZC = decltype(Z)::value;
      if (g.power_law != 0.0)
        k_adj3d_jac_march<BX, kFMY, ZC, true><<<gm, BX * kFMY, 0, st>>>(g, q, lam, G, rhs, T1);
      else
        k_adj3d_jac_march<BX, kFMY, ZC, false><<<gm, BX * kFMY, 0, st>>>(g, q, lam, G, rhs, T1);
      k_adj3d_finish_march<BX, kFMY, ZC><<<gm, BX * kFMY, 0, st>>>(
          g, T1, q, lam, fin.kolmogorov, rhs, rk ? rk->stage : 0, rk ? rk->dt : 0.0,
          rk ? rk->y : nullptr, rk ? rk->b1 : nullptr, rk ? rk->b2 : nullptr);
    });
    return rk != nullptr;
  }
  bool finished = false;
  MGD_DISPATCH(g.nd, {
    constexpr int NV = ND + 2;
    if (g.viscous) launch_gradients<ND>(g, q, w, G, grid, st);
    for (int d = 0; d < ND; ++d)
      if (!launch_deriv_line(g, d, SLOT_ADJ, lam, (long)NV * g.N, T1 + (long)d * NV * g.N,
                             (long)ND * NV * g.N, NV, 0, st))
        k_adj_deriv<ND><<<grid, kBlock, 0, st>>>(g, d, lam, (long)NV * g.N, T1, 0);
    const bool chain_in_finish = g.viscous && !g.dissipation && !has_farfield;
    if (g.viscous)
      k_adj_jac_visc2<ND><<<grid, kBlock, 0, st>>>(g, q, G, T1, rhs);
    else
      k_adj_jac<ND><<<grid, kBlock, 0, st>>>(g, q, G, T1, rhs);
    // the z pass of the second derivative pass also finishes the rhs where nothing but a
    // whole-domain Kolmogorov forcing follows (the x and y passes accumulate in S slot 0)
    const bool fuse = ND == 3 && g.viscous && chain_in_finish && fin.simple &&
                      deriv_line_applies(g, 0) && deriv_line_applies(g, 1) &&
                      deriv_line_applies(g, 2);
    if (fuse) {
      const long ts = (long)ND * NV * g.N;
      launch_deriv_line(g, 0, SLOT_ADJ, T1 + (long)1 * g.N, ts, S + (long)1 * g.N, ts, NV - 1, 0, st);
      launch_deriv_line(g, 1, SLOT_ADJ, T1 + (long)(NV + 1) * g.N, ts, S + (long)1 * g.N, ts, NV - 1,
                        1, st);
      const long nlines = g.N / g.dir[ND - 1].len;
      const dim3 gz((unsigned)((nlines + kBlock - 1) / kBlock),
                    (unsigned)((g.dir[ND - 1].len + kLineChunk - 1) / kLineChunk), (unsigned)g.ninst);
      k_adj_finish_z<<<gz, kBlock, 0, st>>>(g, T1 + (long)((ND - 1) * NV + 1) * g.N, S + (long)1 * g.N,
                                            ts, q, lam, fin.kolmogorov, rhs, rk ? rk->stage : 0,
                                            rk ? rk->dt : 0.0, rk ? rk->y : nullptr,
                                            rk ? rk->b1 : nullptr, rk ? rk->b2 : nullptr);
      finished = true;
    } else if (g.viscous) {
      for (int d = 0; d < ND; ++d)
        if (!launch_deriv_line(g, d, SLOT_ADJ, T1 + (long)(d * NV + 1) * g.N, (long)ND * NV * g.N,
                               S + (long)(d * NV + 1) * g.N, (long)ND * NV * g.N, NV - 1, 0, st))
          k_adj_deriv<ND><<<grid, kBlock, 0, st>>>(g, d, T1, (long)ND * NV * g.N, S, 1);
      // with no rhs term between them, k_finish applies the chain rule itself
      if (!chain_in_finish) k_adj_chain<ND><<<grid, kBlock, 0, st>>>(g, q, S, rhs, -1.0);
    }
    if (g.dissipation) {
      for (int d = 0; d < ND; ++d) {
        k_diss_a<ND><<<grid, kBlock, 0, st>>>(g, d, lam, t, rhs, -g.diss_amount);
        if (!g.composite) k_diss_b<ND><<<grid, kBlock, 0, st>>>(g, d, t, rhs, -g.diss_amount);
      }
    }
    if (g.viscous && has_farfield) {
      const long M = (long)g.ninst * ND * NV * g.N;
      k_fill<<<(unsigned)((M + kBlock - 1) / kBlock), kBlock, 0, st>>>(M, 0.0, T1);
      k_ff_visc_adj<ND><<<grid, kBlock, 0, st>>>(g, P, q, lam, T1);
      for (int d = 0; d < ND; ++d)
        if (!launch_deriv_line(g, d, SLOT_ADJ, T1 + (long)(d * NV + 1) * g.N, (long)ND * NV * g.N,
                               S + (long)(d * NV + 1) * g.N, (long)ND * NV * g.N, NV - 1, 0, st))
          k_adj_deriv<ND><<<grid, kBlock, 0, st>>>(g, d, T1, (long)ND * NV * g.N, S, 1);
      k_adj_chain<ND><<<grid, kBlock, 0, st>>>(g, q, S, rhs, +1.0);
    }
    if (g.viscous && has_farfield)
      k_ff_visc_terms<ND><<<grid, kBlock, 0, st>>>(g, MGD_ADJOINT, P, q, lam, G);
    if (!finished)
      k_finish<ND><<<grid, kBlock, 0, st>>>(g, MGD_ADJOINT, P, q, lam, G, rhs, nullptr, 0, aforce,
                                            aforce_stride, forcing_factor, nullptr, 0, nullptr, 0,
                                            chain_in_finish ? S : nullptr);
  });
  return finished && rk != nullptr;
}

void launch_rk_forward(long M, int stage, double dt, double* q, double* b1, double* b2,
                       const double* rhs, double* qcopy, hipStream_t st) {
  k_rk_forward<<<(unsigned)((M + kBlock - 1) / kBlock), kBlock, 0, st>>>(M, stage, dt, q, b1, b2, rhs,
                                                                        qcopy);
}

void launch_rk_adjoint(long M, int stage, double dt, double* lam, double* b1, double* b2,
                       const double* rhs, hipStream_t st) {
  k_rk_adjoint<<<(unsigned)((M + kBlock - 1) / kBlock), kBlock, 0, st>>>(M, stage, dt, lam, b1, b2,
                                                                         rhs);
}

void launch_copy(long M, const double* a, double* b, hipStream_t st) {
  k_copy<<<(unsigned)((M + kBlock - 1) / kBlock), kBlock, 0, st>>>(M, a, b);
}

void launch_fill(long M, double v, double* b, hipStream_t st) {
  k_fill<<<(unsigned)((M + kBlock - 1) / kBlock), kBlock, 0, st>>>(M, v, b);
}

int reduce_blocks(int np) { return (np + kRedBlock - 1) / kRedBlock; }

void launch_patch_reduce(const Geo& g, const DPatch* pt, int np, int which, const double* x,
                         double* partial, double weight, double* acc, double* raw, hipStream_t st) {
  const int nb = reduce_blocks(np);
  MGD_DISPATCH(g.nd, {
    k_patch_reduce<ND><<<dim3(nb, g.ninst), kRedBlock, 0, st>>>(g, pt, which, x, partial);
  });
  k_reduce_final<<<g.ninst, kRedBlock, 0, st>>>(nb, partial, weight, acc, raw);
}

void launch_inner_product(const Geo& g, const double* a, const double* b, double* partial,
                          double* out, hipStream_t st) {
  const int nb = (int)((g.N + kRedBlock - 1) / kRedBlock);
  MGD_DISPATCH(g.nd, {
    k_inner_partial<ND><<<dim3(nb, g.ninst), kRedBlock, 0, st>>>(g, a, b, partial);
  });
  k_reduce_final<<<g.ninst, kRedBlock, 0, st>>>(nb, partial, 1.0, nullptr, out);
}

void launch_wave_speed(const Geo& g, const double* q, int mode, double param, double* partial,
                       double* out, hipStream_t st) {
  const int nb = (int)((g.N + kRedBlock - 1) / kRedBlock);
  MGD_DISPATCH(g.nd, {
    k_wave_speed<ND><<<dim3(nb, g.ninst), kRedBlock, 0, st>>>(g, q, mode, param, partial);
  });
  k_minmax_final<<<g.ninst, kRedBlock, 0, st>>>(nb, partial, mode, out);
}

void launch_check_limits(const Geo& g, const double* q, const double* range, double* flag,
                         hipStream_t st) {
  MGD_DISPATCH(g.nd, {
    k_check_limits<ND><<<grid_pts(g), kBlock, 0, st>>>(g, q, range[0], range[1], range[2], range[3],
                                                       flag);
  });
}

void launch_kinetic_energy(const Geo& g, const double* q, double* lam, double* partial, double* J,
                           hipStream_t st) {
  const int nb = (int)((g.N + kRedBlock - 1) / kRedBlock);
  MGD_DISPATCH(g.nd, {
    k_kinetic_energy<ND><<<dim3(nb, g.ninst), kRedBlock, 0, st>>>(g, q, lam, partial);
  });
  k_reduce_final<<<g.ninst, kRedBlock, 0, st>>>(nb, partial, 1.0, nullptr, J);
}

void launch_lincomb(long M, const double* a, const double* b, double s, const double* c, double* d,
                    hipStream_t st) {
  k_lincomb<<<(unsigned)((M + kBlock - 1) / kBlock), kBlock, 0, st>>>(M, a, b, s, c, d);
}

void launch_adj_forcing(const Geo& g, const DPatch* pt, int np, const double* q, double* out,
                        int zero, hipStream_t st) {
  MGD_DISPATCH(g.nd, {
    k_adj_forcing<ND><<<dim3((np + kBlock - 1) / kBlock, g.ninst), kBlock, 0, st>>>(g, pt, q, out, zero);
  });
}

void launch_adj_terminal(const Geo& g, const DPatch* pt, int np, const double* af, double factor,
                         double* lam, hipStream_t st) {
  MGD_DISPATCH(g.nd, {
    k_adj_terminal<ND><<<dim3((np + kBlock - 1) / kBlock, g.ninst), kBlock, 0, st>>>(g, pt, af, factor, lam);
  });
}

void launch_gradient(const Geo& g, const DPatch* pt, int np, const double* lam, double* out,
                     long inst_stride, hipStream_t st) {
  MGD_DISPATCH(g.nd, {
    k_gradient<ND><<<dim3((np + kBlock - 1) / kBlock, g.ninst), kBlock, 0, st>>>(g, pt, lam, out,
                                                                               inst_stride);
  });
}

// Stencil operator on a device field (StencilOperator%apply), ncomp components.
__global__ void __launch_bounds__(kBlock) k_apply(Geo g, int slot, int d, int ncomp, const double* __restrict__ x,
                        double* __restrict__ y) {
  const long p = gidx();
  if (p >= g.N) return;
  const int c = blockIdx.y;
  int ijk[3] = {0, 0, 0};
  ijk[0] = (int)(p % g.n[0]);
  ijk[1] = (int)((p / g.n[0]) % g.n[1]);
  ijk[2] = (int)(p / ((long)g.n[0] * g.n[1]));
  const Dir dd = g.dir[d];
  const long base = p - (long)ijk[d] * dd.stride;
  const double* xc = x + (long)c * g.N;
  double out[1];
  sbp_apply<1>(g.ops[slot], dd, ijk[d],
               [&](int idx, double* v) { v[0] = xc[base + (long)idx * dd.stride]; }, out);
  y[(long)c * g.N + p] = out[0];
}

void launch_apply(const Geo& g, int slot, int d, int ncomp, const double* x, double* y,
                  hipStream_t st) {
  k_apply<<<dim3((unsigned)((g.N + kBlock - 1) / kBlock), ncomp), kBlock, 0, st>>>(g, slot, d, ncomp,
                                                                                 x, y);
}

}  // namespace mgd
