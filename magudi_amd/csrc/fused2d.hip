// fused2d.hip -- one-launch-per-RK-stage kernels for 2-D inviscid flows on gfx950.
//
// A work-group owns a TX x TY tile of grid points of one instance.  It stages the
// operand (q for the forward stage, lambda for the adjoint stage) plus a 4-point halo
// in LDS, forms the contravariant fluxes once per staged point, and computes for its
// points: the SBP 4-8 flux divergence in x and y, the artificial dissipation
// D4^T diag(-arc) D4 in x and y, the Jacobian scaling, every patch term
// (far-field SAT, sponge, actuator, cost target, Kolmogorov), the acoustic source
// and the RK4 stage update -- so each RK stage reads its operand once and writes
// its outputs once.  Points within the SBP boundary closures of a non-periodic
// direction (8 rows, 12 for the adjoint operator) take the closure rows from global
// memory through the cache; the LDS path covers everything else.
//
// Per-stage functional / sensitivity partial sums are written per tile into
// [stage][tile] slots and reduced in a fixed order at the end of a run, so results
// are deterministic and independent of launch order.
#include "device.h"
#include "fused2d.h"

namespace mgd {

constexpr int TX = 32, TY = 16, HX = 4, HY = 4;
constexpr int EX = TX + 2 * HX, EY = TY + 2 * HY;
constexpr int NT = 256;
constexpr int PPT = TX * TY / NT;   // points per thread (2)
constexpr int NV2 = 4;
// LDS layout (doubles): operand [4][EY][EX], then scratch reused for the fluxes and
// for the dissipation intermediates.
constexpr int SQ = NV2 * EY * EX;                       // 3840
constexpr int SFX = NV2 * TY * EX, SFY = NV2 * EY * TX;  // 2560, 3072
constexpr int STX = NV2 * TY * (TX + 4), STY = NV2 * (TY + 4) * TX;   // 2304, 2560
constexpr int SSCR = (SFX + SFY) > (STX + STY) ? (SFX + SFY) : (STX + STY);

__device__ __forceinline__ int zone_lo(const Dir& dd, int i, int w) {
  // 1 when i lies in a closure zone of width w of a non-periodic direction
  return !dd.periodic && (i < w || i >= dd.len - w);
}

// maps a tile-relative extended index to a global index; returns -1 outside the domain
__device__ __forceinline__ int ext_index(const Dir& dd, int gi) {
  if (gi >= 0 && gi < dd.len) return gi;
  if (!dd.periodic) return -1;
  return gi < 0 ? gi + dd.len - dd.off : gi - dd.len + dd.off;
}

// Dissipation D4^T(-arc D4 x) / H (or the composite operator) along d at one point,
// operands from global memory (closure zones).
__device__ __forceinline__ void diss_global(const Geo& g, const double* xi, int d, const int* ijk,
                                            double* out) {
  const Dir dd = g.dir[d];
  const long p = ijk[0] + (long)g.n[0] * ijk[1];
  const long base = p - (long)ijk[d] * dd.stride;
  if (g.composite) {
    sbp_apply<NV2>(g.ops[SLOT_DISS], dd, ijk[d],
                   [&](int idx, double* v) {
                     const long pp = base + (long)idx * dd.stride;
                     for (int l = 0; l < NV2; ++l) v[l] = xi[l * g.N + pp];
                   },
                   out);
    return;
  }
  sbp_apply<NV2>(g.ops[SLOT_DISS_T], dd, ijk[d],
                 [&](int idx, double* t) {
                   sbp_apply<NV2>(g.ops[SLOT_DISS], dd, idx,
                                  [&](int idx2, double* v) {
                                    const long pp = base + (long)idx2 * dd.stride;
                                    for (int l = 0; l < NV2; ++l) v[l] = xi[l * g.N + pp];
                                  },
                                  t);
                   int jk[3] = {ijk[0], ijk[1], 0};
                   jk[d] = idx;
                   const double arc = fabs(metric<2>(g, d, jk));
                   for (int l = 0; l < NV2; ++l) t[l] = -arc * t[l];
                 },
                 out);
  const double h = g.hw[d][ijk[d]];
  for (int l = 0; l < NV2; ++l) out[l] = out[l] / h;
}

// contravariant inviscid flux along d at a global point
__device__ __forceinline__ void fhat_global(const Geo& g, const double* qi, long pp, int d,
                                            const int* jk, double* f) {
  double v[NV2];
  for (int l = 0; l < NV2; ++l) v[l] = qi[l * g.N + pp];
  const Prim<2> s = prim<2>(v, g.gamma);
  inviscid_flux<2>(v, s, d, f);
  const double m = metric<2>(g, d, jk);
  for (int l = 0; l < NV2; ++l) f[l] = m * f[l];
}

__device__ __forceinline__ double block_sum256(double v, double* red) {
  // red: NT doubles of LDS; the first barrier retires earlier readers of that space
  __syncthreads();
  red[threadIdx.x] = v;
  __syncthreads();
  for (int s = NT / 2; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
    __syncthreads();
  }
  const double r = red[0];
  __syncthreads();
  return r;
}

// -------------------------------------------------------------------------- forward
__global__ void __launch_bounds__(NT) k_fwd_stage2d(Geo g, const Patches* __restrict__ P,
                                                    FwdStage a) {
  __shared__ double sq[SQ];
  __shared__ double sc[SSCR];
  const int x0 = blockIdx.x * TX, y0 = blockIdx.y * TY;
  const long inst = blockIdx.z;
  const long NVN = (long)NV2 * g.N;
  const double* qi = a.qin + inst * NVN;
  const Dir dx = g.dir[0], dy = g.dir[1];
  const double gam = g.gamma;
  double* sfx = sc;
  double* sfy = sc + SFX;
  // ---- stage the operand and the fluxes
  for (int e = threadIdx.x; e < EX * EY; e += NT) {
    const int ex = e % EX, ey = e / EX;
    const int gi = ext_index(dx, x0 - HX + ex), gj = ext_index(dy, y0 - HY + ey);
    if (gi < 0 || gj < 0) continue;
    const long pp = gi + (long)g.n[0] * gj;
    double v[NV2];
    for (int l = 0; l < NV2; ++l) {
      v[l] = qi[l * g.N + pp];
      sq[(l * EY + ey) * EX + ex] = v[l];
    }
    const Prim<2> s = prim<2>(v, gam);
    if (ey >= HY && ey < HY + TY) {
      double f[NV2];
      inviscid_flux<2>(v, s, 0, f);
      const double m = g.h[1][gj];
      for (int l = 0; l < NV2; ++l) sfx[(l * TY + (ey - HY)) * EX + ex] = m * f[l];
    }
    if (ex >= HX && ex < HX + TX) {
      double f[NV2];
      inviscid_flux<2>(v, s, 1, f);
      const double m = g.h[0][gi];
      for (int l = 0; l < NV2; ++l) sfy[(l * EY + ey) * TX + (ex - HX)] = m * f[l];
    }
  }
  __syncthreads();
  const Op& o1 = g.ops[SLOT_FIRST];
  double r[PPT][NV2];
  bool own[PPT];
  int pi[PPT], pj[PPT];
  for (int k = 0; k < PPT; ++k) {
    const int tx = threadIdx.x % TX, ty = threadIdx.x / TX + k * (NT / TX);
    pi[k] = x0 + tx;
    pj[k] = y0 + ty;
    own[k] = pi[k] < dx.len && pj[k] < dy.len;
    if (!own[k]) continue;
    double dfx[NV2], dfy[NV2];
    const int jk[3] = {pi[k], pj[k], 0};
    if (zone_lo(dx, pi[k], o1.bd)) {
      const long base = (long)g.n[0] * pj[k];
      sbp_apply<NV2>(o1, dx, pi[k],
                     [&](int idx, double* f) {
                       const int jj[3] = {idx, pj[k], 0};
                       fhat_global(g, qi, base + idx, 0, jj, f);
                     },
                     dfx);
    } else {
      for (int l = 0; l < NV2; ++l) dfx[l] = 0.0;
      const double* row = sfx + ty * EX + (tx + HX);
      for (int kk = 1; kk <= 4; ++kk) {
        const double c = o1.c[kMaxHW + kk];
        for (int l = 0; l < NV2; ++l)
          dfx[l] += c * (row[l * TY * EX + kk] - row[l * TY * EX - kk]);
      }
    }
    if (zone_lo(dy, pj[k], o1.bd)) {
      sbp_apply<NV2>(o1, dy, pj[k],
                     [&](int idx, double* f) {
                       const int jj[3] = {pi[k], idx, 0};
                       fhat_global(g, qi, pi[k] + (long)g.n[0] * idx, 1, jj, f);
                     },
                     dfy);
    } else {
      for (int l = 0; l < NV2; ++l) dfy[l] = 0.0;
      const double* col = sfy + (ty + HY) * TX + tx;
      for (int kk = 1; kk <= 4; ++kk) {
        const double c = o1.c[kMaxHW + kk];
        for (int l = 0; l < NV2; ++l)
          dfy[l] += c * (col[l * EY * TX + kk * TX] - col[l * EY * TX - kk * TX]);
      }
    }
    for (int l = 0; l < NV2; ++l) r[k][l] = 0.0 - (dfx[l] + dfy[l]);
    (void)jk;
  }
  if (g.dissipation) {
    __syncthreads();   // fluxes consumed; reuse the scratch for the dissipation
    const Op& od = g.ops[SLOT_DISS];
    double* stx = sc;          // [4][TY][TX+4]
    double* sty = sc + STX;    // [4][TY+4][TX]
    if (!g.composite) {
      for (int e = threadIdx.x; e < TY * (TX + 4); e += NT) {
        const int ex = e % (TX + 4) + 2, ey = e / (TX + 4) + HY;   // extended coords
        const int gj = ext_index(dy, y0 - HY + ey);
        const double arc = gj >= 0 ? fabs(g.h[1][gj]) : 0.0;
        for (int l = 0; l < NV2; ++l) {
          const double* row = sq + (l * EY + ey) * EX + ex;
          double s = 0.0;
          for (int kk = 1; kk <= 2; ++kk) s += od.c[kMaxHW + kk] * (row[kk] + row[-kk]);
          s = s + od.c[kMaxHW] * row[0];
          stx[(l * TY + (ey - HY)) * (TX + 4) + (ex - 2)] = -arc * s;
        }
      }
      for (int e = threadIdx.x; e < (TY + 4) * TX; e += NT) {
        const int ex = e % TX + HX, ey = e / TX + 2;
        const int gi = ext_index(dx, x0 - HX + ex);
        const double arc = gi >= 0 ? fabs(g.h[0][gi]) : 0.0;
        for (int l = 0; l < NV2; ++l) {
          const double* col = sq + (l * EY + ey) * EX + ex;
          double s = 0.0;
          for (int kk = 1; kk <= 2; ++kk) s += od.c[kMaxHW + kk] * (col[kk * EX] + col[-kk * EX]);
          s = s + od.c[kMaxHW] * col[0];
          sty[(l * (TY + 4) + (ey - 2)) * TX + (ex - HX)] = -arc * s;
        }
      }
      __syncthreads();
    }
    const Op& ot = g.ops[SLOT_DISS_T];
    for (int k = 0; k < PPT; ++k) {
      if (!own[k]) continue;
      const int tx = threadIdx.x % TX, ty = threadIdx.x / TX + k * (NT / TX);
      const int jk[3] = {pi[k], pj[k], 0};
      for (int d = 0; d < 2; ++d) {
        double dd_[NV2];
        const Dir& dr = g.dir[d];
        const int ii = d == 0 ? pi[k] : pj[k];
        if (zone_lo(dr, ii, 8)) {
          diss_global(g, qi, d, jk, dd_);
        } else if (g.composite) {
          for (int l = 0; l < NV2; ++l) {
            const double* b = d == 0 ? sq + (l * EY + ty + HY) * EX + tx + HX
                                     : sq + (l * EY + ty + HY) * EX + tx + HX;
            const int st = d == 0 ? 1 : EX;
            double s = 0.0;
            for (int kk = 1; kk <= 4; ++kk) s += od.c[kMaxHW + kk] * (b[kk * st] + b[-kk * st]);
            dd_[l] = s + od.c[kMaxHW] * b[0];
          }
        } else {
          for (int l = 0; l < NV2; ++l) {
            const double* b = d == 0 ? stx + (l * TY + ty) * (TX + 4) + tx + 2
                                     : sty + (l * (TY + 4) + ty + 2) * TX + tx;
            const int st = d == 0 ? 1 : TX;
            double s = 0.0;
            for (int kk = 1; kk <= 2; ++kk) s += ot.c[kMaxHW + kk] * (b[kk * st] + b[-kk * st]);
            s = s + ot.c[kMaxHW] * b[0];
            dd_[l] = s / 1.0;
          }
        }
        for (int l = 0; l < NV2; ++l) r[k][l] = r[k][l] + g.diss_amount * dd_[l];
      }
    }
  }
  // ---- Jacobian, patches, sources, RK4 update, functional partials
  double jpart = 0.0;
  for (int k = 0; k < PPT; ++k) {
    if (!own[k]) continue;
    const int jk[3] = {pi[k], pj[k], 0};
    const long p = pi[k] + (long)g.n[0] * pj[k];
    const double jac = jacobian<2>(g, jk);
    double* rr = r[k];
    for (int l = 0; l < NV2; ++l) rr[l] = rr[l] * jac;
    double qp[NV2];
    for (int l = 0; l < NV2; ++l) qp[l] = qi[l * g.N + p];
    for (int ip = 0; ip < P->count; ++ip) {
      const DPatch& pt = P->p[ip];
      if (pi[k] < pt.lo[0] || pi[k] > pt.hi[0] || pj[k] < pt.lo[1] || pj[k] > pt.hi[1]) continue;
      const int kp = (pi[k] - pt.lo[0]) + pt.ls[0] * (pj[k] - pt.lo[1]);
      switch (pt.type) {
        case MGD_PATCH_FAR_FIELD: {
          const double* A = pt.aplus + (long)kp * 16;
          const double* qt = pt.aplus + (long)pt.np * 16 + (long)kp * 4;
          const double sj = pt.sigma_inv * jac;
          for (int i = 0; i < NV2; ++i) {
            double s = 0.0;
            for (int c = 0; c < NV2; ++c) s += A[c * 4 + i] * (qp[c] - qt[c]);
            rr[i] = rr[i] - sj * s;
          }
          break;
        }
        case MGD_PATCH_SPONGE: {
          const double sp = pt.strength[kp];
          const double* qt = pt.aplus + (long)kp * 4;
          for (int l = 0; l < NV2; ++l) rr[l] = rr[l] - sp * (qp[l] - qt[l]);
          break;
        }
        case MGD_PATCH_ACTUATOR:
          if (a.control) rr[3] = rr[3] + pt.strength[kp] * a.control[inst * a.control_stride + kp];
          break;
        case MGD_PATCH_KOLMOGOROV:
          rr[1] = rr[1] + qp[0] * pt.strength[kp];
          break;
        default:
          break;
      }
    }
    for (int s = 0; s < a.nsrc; ++s) {
      const double dx0 = g.x[0][pi[k]] - a.src[s].loc[0];
      const double dy0 = g.x[1][pj[k]] - a.src[s].loc[1];
      double rr2 = 0.0;
      rr2 += dx0 * dx0;
      rr2 += dy0 * dy0;
      rr[3] = rr[3] + a.src_amp[s] * exp(-a.src[s].gfac * rr2);
    }
    // RK4 stage update (substepForwardRK4, src/RK4IntegratorImpl.f90:115-156)
    double qn[NV2];
    const double dt = a.dt;
    double* b2 = a.b2 + inst * NVN;
    const double* q0 = a.q0 + inst * NVN;
    for (int l = 0; l < NV2; ++l) {
      const long o = l * g.N + p;
      const double rl = rr[l];
      switch (a.stage) {
        case 1:
          b2[o] = qp[l] + dt * rl / 6.0;
          qn[l] = qp[l] + dt * rl / 2.0;
          break;
        case 2:
          b2[o] = b2[o] + dt * rl / 3.0;
          qn[l] = q0[o] + dt * rl / 2.0;
          break;
        case 3:
          b2[o] = b2[o] + dt * rl / 3.0;
          qn[l] = q0[o] + dt * rl;
          break;
        default:
          qn[l] = b2[o] + dt * rl / 6.0;
          break;
      }
      a.qout[inst * NVN + o] = qn[l];
      if (a.rhs_out) a.rhs_out[inst * NVN + o] = rl;
    }
    if (a.jpartial && a.ip_cost >= 0) {
      const DPatch& pt = P->p[a.ip_cost];
      if (pi[k] >= pt.lo[0] && pi[k] <= pt.hi[0] && pj[k] >= pt.lo[1] && pj[k] <= pt.hi[1]) {
        const int kp = (pi[k] - pt.lo[0]) + pt.ls[0] * (pj[k] - pt.lo[1]);
        const Prim<2> s = prim<2>(qn, gam);
        const double F = s.p - pt.aplus[kp];
        jpart += 1.0 * gnorm<2>(g, jk) * (F * F * pt.strength[kp]);
      }
    }
  }
  if (a.jpartial && a.ip_cost >= 0) {
    const int rank = a.tile_rank[blockIdx.y * gridDim.x + blockIdx.x];
    if (rank >= 0) {
      const double s = block_sum256(jpart, sc);
      if (threadIdx.x == 0) a.jpartial[(inst * a.nstage_slots + a.stage_slot) * a.ntiles + rank] = s;
    }
  }
}

// -------------------------------------------------------------------------- adjoint
__global__ void __launch_bounds__(NT) k_adj_stage2d(Geo g, const Patches* __restrict__ P,
                                                    AdjStage a) {
  __shared__ double sq[SQ];
  __shared__ double sc[SSCR];
  const int x0 = blockIdx.x * TX, y0 = blockIdx.y * TY;
  const long inst = blockIdx.z;
  const long NVN = (long)NV2 * g.N;
  const double* li = a.lam + inst * NVN;
  const double* qi = a.q + inst * NVN;
  const Dir dx = g.dir[0], dy = g.dir[1];
  const double gam = g.gamma;
  for (int e = threadIdx.x; e < EX * EY; e += NT) {
    const int ex = e % EX, ey = e / EX;
    const int gi = ext_index(dx, x0 - HX + ex), gj = ext_index(dy, y0 - HY + ey);
    if (gi < 0 || gj < 0) continue;
    const long pp = gi + (long)g.n[0] * gj;
    for (int l = 0; l < NV2; ++l) sq[(l * EY + ey) * EX + ex] = li[l * g.N + pp];
  }
  __syncthreads();
  const Op& oa = g.ops[SLOT_ADJ];
  double r[PPT][NV2];
  bool own[PPT];
  int pi[PPT], pj[PPT];
  for (int k = 0; k < PPT; ++k) {
    const int tx = threadIdx.x % TX, ty = threadIdx.x / TX + k * (NT / TX);
    pi[k] = x0 + tx;
    pj[k] = y0 + ty;
    own[k] = pi[k] < dx.len && pj[k] < dy.len;
    if (!own[k]) continue;
    const long p = pi[k] + (long)g.n[0] * pj[k];
    double T[2][NV2];
    if (zone_lo(dx, pi[k], oa.bd)) {
      const long base = (long)g.n[0] * pj[k];
      sbp_apply<NV2>(oa, dx, pi[k],
                     [&](int idx, double* v) {
                       for (int l = 0; l < NV2; ++l) v[l] = li[l * g.N + base + idx];
                     },
                     T[0]);
    } else {
      for (int l = 0; l < NV2; ++l) {
        const double* row = sq + (l * EY + ty + HY) * EX + tx + HX;
        double s = 0.0;
        for (int kk = 1; kk <= 4; ++kk) s += oa.c[kMaxHW + kk] * (row[kk] - row[-kk]);
        T[0][l] = s;
      }
    }
    if (zone_lo(dy, pj[k], oa.bd)) {
      sbp_apply<NV2>(oa, dy, pj[k],
                     [&](int idx, double* v) {
                       for (int l = 0; l < NV2; ++l) v[l] = li[l * g.N + pi[k] + (long)g.n[0] * idx];
                     },
                     T[1]);
    } else {
      for (int l = 0; l < NV2; ++l) {
        const double* col = sq + (l * EY + ty + HY) * EX + tx + HX;
        double s = 0.0;
        for (int kk = 1; kk <= 4; ++kk) s += oa.c[kMaxHW + kk] * (col[kk * EX] - col[-kk * EX]);
        T[1][l] = s;
      }
    }
    double v[NV2];
    for (int l = 0; l < NV2; ++l) v[l] = qi[l * g.N + p];
    const Prim<2> s = prim<2>(v, gam);
    const int jk[3] = {pi[k], pj[k], 0};
    double acc[NV2] = {0.0, 0.0, 0.0, 0.0};
    for (int d = 0; d < 2; ++d) {
      double m[2], A[16];
      m[0] = d == 0 ? metric<2>(g, 0, jk) : 0.0;
      m[1] = d == 1 ? metric<2>(g, 1, jk) : 0.0;
      inviscid_jacobian<2>(m, s, gam, A);
      for (int i = 0; i < NV2; ++i) {
        double t = 0.0;
        for (int c = 0; c < NV2; ++c) t += A[i * 4 + c] * T[d][c];
        acc[i] = acc[i] + t;
      }
    }
    for (int l = 0; l < NV2; ++l) r[k][l] = acc[l];
  }
  if (g.dissipation) {
    const Op& od = g.ops[SLOT_DISS];
    const Op& ot = g.ops[SLOT_DISS_T];
    double* stx = sc;
    double* sty = sc + STX;
    if (!g.composite) {
      for (int e = threadIdx.x; e < TY * (TX + 4); e += NT) {
        const int ex = e % (TX + 4) + 2, ey = e / (TX + 4) + HY;
        const int gj = ext_index(dy, y0 - HY + ey);
        const double arc = gj >= 0 ? fabs(g.h[1][gj]) : 0.0;
        for (int l = 0; l < NV2; ++l) {
          const double* row = sq + (l * EY + ey) * EX + ex;
          double s = 0.0;
          for (int kk = 1; kk <= 2; ++kk) s += od.c[kMaxHW + kk] * (row[kk] + row[-kk]);
          s = s + od.c[kMaxHW] * row[0];
          stx[(l * TY + (ey - HY)) * (TX + 4) + (ex - 2)] = -arc * s;
        }
      }
      for (int e = threadIdx.x; e < (TY + 4) * TX; e += NT) {
        const int ex = e % TX + HX, ey = e / TX + 2;
        const int gi = ext_index(dx, x0 - HX + ex);
        const double arc = gi >= 0 ? fabs(g.h[0][gi]) : 0.0;
        for (int l = 0; l < NV2; ++l) {
          const double* col = sq + (l * EY + ey) * EX + ex;
          double s = 0.0;
          for (int kk = 1; kk <= 2; ++kk) s += od.c[kMaxHW + kk] * (col[kk * EX] + col[-kk * EX]);
          s = s + od.c[kMaxHW] * col[0];
          sty[(l * (TY + 4) + (ey - 2)) * TX + (ex - HX)] = -arc * s;
        }
      }
      __syncthreads();
    }
    for (int k = 0; k < PPT; ++k) {
      if (!own[k]) continue;
      const int tx = threadIdx.x % TX, ty = threadIdx.x / TX + k * (NT / TX);
      const int jk[3] = {pi[k], pj[k], 0};
      for (int d = 0; d < 2; ++d) {
        double dd_[NV2];
        const int ii = d == 0 ? pi[k] : pj[k];
        if (zone_lo(g.dir[d], ii, 8)) {
          diss_global(g, li, d, jk, dd_);
        } else if (g.composite) {
          for (int l = 0; l < NV2; ++l) {
            const double* b = sq + (l * EY + ty + HY) * EX + tx + HX;
            const int st = d == 0 ? 1 : EX;
            double s = 0.0;
            for (int kk = 1; kk <= 4; ++kk) s += od.c[kMaxHW + kk] * (b[kk * st] + b[-kk * st]);
            dd_[l] = s + od.c[kMaxHW] * b[0];
          }
        } else {
          for (int l = 0; l < NV2; ++l) {
            const double* b = d == 0 ? stx + (l * TY + ty) * (TX + 4) + tx + 2
                                     : sty + (l * (TY + 4) + ty + 2) * TX + tx;
            const int st = d == 0 ? 1 : TX;
            double s = 0.0;
            for (int kk = 1; kk <= 2; ++kk) s += ot.c[kMaxHW + kk] * (b[kk * st] + b[-kk * st]);
            s = s + ot.c[kMaxHW] * b[0];
            dd_[l] = s / 1.0;
          }
        }
        for (int l = 0; l < NV2; ++l) r[k][l] = r[k][l] + (-g.diss_amount) * dd_[l];
      }
    }
  }
  double spart = 0.0;
  for (int k = 0; k < PPT; ++k) {
    if (!own[k]) continue;
    const int jk[3] = {pi[k], pj[k], 0};
    const long p = pi[k] + (long)g.n[0] * pj[k];
    const double jac = jacobian<2>(g, jk);
    double* rr = r[k];
    for (int l = 0; l < NV2; ++l) rr[l] = rr[l] * jac;
    double lp[NV2];
    for (int l = 0; l < NV2; ++l) lp[l] = li[l * g.N + p];
    for (int ip = 0; ip < P->count; ++ip) {
      const DPatch& pt = P->p[ip];
      if (pi[k] < pt.lo[0] || pi[k] > pt.hi[0] || pj[k] < pt.lo[1] || pj[k] > pt.hi[1]) continue;
      const int kp = (pi[k] - pt.lo[0]) + pt.ls[0] * (pj[k] - pt.lo[1]);
      switch (pt.type) {
        case MGD_PATCH_FAR_FIELD: {
          const double* A = pt.aplus + (long)kp * 16;
          const double sj = pt.sigma_inv * jac;
          for (int i = 0; i < NV2; ++i) {
            double s = 0.0;
            for (int c = 0; c < NV2; ++c) s += A[i * 4 + c] * lp[c];
            rr[i] = rr[i] + sj * s;
          }
          break;
        }
        case MGD_PATCH_SPONGE: {
          const double sp = pt.strength[kp];
          for (int l = 0; l < NV2; ++l) rr[l] = rr[l] + sp * lp[l];
          break;
        }
        case MGD_PATCH_COST_TARGET: {
          // adjoint forcing of the acoustic-noise functional from the forward state at
          // this point (src/AcousticNoiseImpl.f90:208-280), scaled by the stage factor
          double v[NV2];
          for (int l = 0; l < NV2; ++l) v[l] = qi[l * g.N + p];
          const Prim<2> s = prim<2>(v, gam);
          double af[NV2];
          if (a.zero_forcing) {
            for (int l = 0; l < NV2; ++l) af[l] = 0.0;
          } else {
            const double F = -2.0 * pt.strength[kp] * 1.0 * (gam - 1.0) * (s.p - pt.aplus[kp]);
            af[3] = F;
            af[1] = -s.u[0] * F;
            af[2] = -s.u[1] * F;
            double ke = 0.0;
            ke += s.u[0] * s.u[0];
            ke += s.u[1] * s.u[1];
            af[0] = 0.5 * ke * F;
          }
          for (int l = 0; l < NV2; ++l) rr[l] = rr[l] + a.forcing_factor * af[l];
          break;
        }
        case MGD_PATCH_ACTUATOR: {
          // ThermalActuator gradient and sensitivity on the pre-update adjoint state
          const double g1 = lp[3] * (pt.strength[kp] * 1.0);
          if (a.grad) a.grad[inst * a.grad_stride + kp] = g1;
          const double F = lp[3] * pt.strength[kp] * 1.0;
          spart += 1.0 * gnorm<2>(g, jk) * (F * F);
          break;
        }
        case MGD_PATCH_KOLMOGOROV:
          rr[0] = rr[0] - lp[1] * pt.strength[kp];
          break;
      }
    }
    // RK4 adjoint stage update (substepAdjointRK4, src/RK4IntegratorImpl.f90:215-261)
    const double dt = a.dt;
    const double* b1 = a.l0 + inst * NVN;
    double* b2 = a.b2 + inst * NVN;
    for (int l = 0; l < NV2; ++l) {
      const long o = l * g.N + p;
      const double rl = rr[l];
      double ln;
      switch (a.stage) {
        case 4:
          b2[o] = lp[l] - dt * rl / 6.0;
          ln = lp[l] - dt * rl / 2.0;
          break;
        case 3:
          b2[o] = b2[o] - dt * rl / 3.0;
          ln = b1[o] - dt * rl / 2.0;
          break;
        case 2:
          b2[o] = b2[o] - dt * rl / 3.0;
          ln = b1[o] - dt * rl;
          break;
        default:
          ln = b2[o] - dt * rl / 6.0;
          break;
      }
      a.lout[inst * NVN + o] = ln;
      if (a.rhs_out) a.rhs_out[inst * NVN + o] = rl;
    }
  }
  if (a.spartial && a.ip_act >= 0) {
    const int rank = a.tile_rank[blockIdx.y * gridDim.x + blockIdx.x];
    if (rank >= 0) {
      const double s = block_sum256(spart, sc);
      if (threadIdx.x == 0) a.spartial[(inst * a.nstage_slots + a.stage_slot) * a.ntiles + rank] = s;
    }
  }
}

// Per-stage sums of the tile partials: one block per (stage, instance).
__global__ void k_stage_sums(const double* __restrict__ partial, int nstages, int ntiles,
                             double* __restrict__ sums) {
  __shared__ double red[NT];
  const int s = blockIdx.x;
  const long inst = blockIdx.y;
  double v = 0.0;
  for (int t = threadIdx.x; t < ntiles; t += NT) v += partial[(inst * nstages + s) * ntiles + t];
  const double sum = block_sum256(v, red);
  if (threadIdx.x == 0) sums[inst * nstages + s] = sum;
}

// acc[inst] += w[s] * sums[s] over the stages in time-marching order (the running time
// quadrature of Solver%runForward / runAdjoint); reverse = adjoint order.
__global__ void k_stage_accumulate(const double* __restrict__ sums, int nstages,
                                   const double* __restrict__ w, double* __restrict__ acc,
                                   int ninst, int reverse) {
  const int inst = blockIdx.x * blockDim.x + threadIdx.x;
  if (inst >= ninst) return;
  double total = acc[inst];
  for (int s0 = 0; s0 < nstages; ++s0) {
    const int s = reverse ? nstages - 1 - s0 : s0;
    total = total + w[s] * sums[(long)inst * nstages + s];
  }
  acc[inst] = total;
}

dim3 fused2d_grid(const Geo& g) {
  return dim3((unsigned)((g.n[0] + TX - 1) / TX), (unsigned)((g.n[1] + TY - 1) / TY),
              (unsigned)g.ninst);
}

int fused2d_tiles(const Geo& g) { return ((g.n[0] + TX - 1) / TX) * ((g.n[1] + TY - 1) / TY); }

// tile -> rank among the tiles intersecting the box [lo, hi] (or -1)
int fused2d_tile_ranks(const Geo& g, const int* lo, const int* hi, std::vector<int>& rank) {
  const int nbx = (g.n[0] + TX - 1) / TX, nby = (g.n[1] + TY - 1) / TY;
  rank.assign((size_t)nbx * nby, -1);
  int c = 0;
  for (int by = 0; by < nby; ++by)
    for (int bx = 0; bx < nbx; ++bx) {
      const int xa = bx * TX, xb = std::min(g.n[0], xa + TX) - 1;
      const int ya = by * TY, yb = std::min(g.n[1], ya + TY) - 1;
      if (xb < lo[0] || xa > hi[0] || yb < lo[1] || ya > hi[1]) continue;
      rank[(size_t)by * nbx + bx] = c++;
    }
  return c;
}

void launch_fwd_stage2d(const Geo& g, const Patches* P, const FwdStage& a, hipStream_t st) {
  k_fwd_stage2d<<<fused2d_grid(g), NT, 0, st>>>(g, P, a);
}

void launch_adj_stage2d(const Geo& g, const Patches* P, const AdjStage& a, hipStream_t st) {
  k_adj_stage2d<<<fused2d_grid(g), NT, 0, st>>>(g, P, a);
}

void launch_stage_quadrature(const double* partial, int nstages, int ntiles, const double* w,
                             double* stage_sums, double* acc, int ninst, int reverse,
                             hipStream_t st) {
  if (nstages <= 0) return;
  k_stage_sums<<<dim3(nstages, ninst), NT, 0, st>>>(partial, nstages, ntiles, stage_sums);
  k_stage_accumulate<<<(ninst + 63) / 64, 64, 0, st>>>(stage_sums, nstages, w, acc, ninst, reverse);
}

}  // namespace mgd
