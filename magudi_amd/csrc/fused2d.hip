// fused2d.hip -- one-launch-per-RK-stage kernels for 2-D inviscid flows on gfx950.
//
// A work-group owns one tile of at most TX x TY grid points of one instance.  It
// stages the operand (q for a forward stage, lambda for an adjoint stage) with a
// 4-point halo in LDS, forms the contravariant fluxes once per staged point, and
// computes for its points: the SBP 4-8 flux divergence in x and y, the artificial
// dissipation D4^T diag(-arc) D4 (or the composite operator) in x and y, the Jacobian
// scaling, the patch terms (far-field SAT, sponge, actuator, cost target, Kolmogorov),
// the acoustic source and the RK4 stage update -- each RK stage reads its operands
// once and writes its outputs once.
//
// Tiles come from a balanced partition of each direction (host side), so on a
// non-periodic direction the first and last tiles hold at least 12 points: every
// boundary-closure row then lies in one tile and its closure stencil (up to 16 wide
// for the adjoint operator) reads staged LDS data like the interior stencil does.
// Each tile carries a bit mask of the patches that intersect it.
//
// The functional / sensitivity quadrature terms are written per patch point and stage
// and reduced by a fixed-order tree, so J and the sensitivity are deterministic and
// independent of the tile shape and of scheduling.
#include "device.h"
#include "fused2d.h"

namespace mgd {

constexpr int HX = 4, HY = 4;     // halo of the staged operand
constexpr int NV2 = 4;
constexpr int kMinEdgeTile = 12;    // closure depth of the adjoint 4-8 operator
constexpr int kRedThreads = 256;    // block size of the reduction kernels
constexpr int PPT = 1;              // points per thread

// Tile shape: TX x TY points, one point per thread.  LDS (doubles): operand
// [4][EY][EX]; scratch for the fluxes, reused for the dissipation intermediates and the
// final reduction.
template <int TX_, int TY_>
struct Shape {
  static constexpr int TX = TX_, TY = TY_;
  static constexpr int EX = TX + 2 * HX, EY = TY + 2 * HY;
  static constexpr int NT = TX * TY;
  static constexpr int kStageIters = (EX * EY + NT - 1) / NT;   // staged points per thread
  static constexpr int SQ = NV2 * EY * EX;
  static constexpr int SFX = NV2 * TY * EX, SFY = NV2 * EY * TX;
  static constexpr int STX = NV2 * TY * (TX + 4), STY = NV2 * (TY + 4) * TX;
  static constexpr int SSCR = (SFX + SFY) > (STX + STY) ? (SFX + SFY) : (STX + STY);
  // waves per SIMD that the LDS use allows (2 blocks of 512 threads or 3 of 256 per CU):
  // the register budget is sized to it
  static constexpr int kWavesPerEU = NT >= 512 ? 4 : 3;
};

struct TileGeo {
  int x0, y0, sx, sy;   // origin and size of the tile
  int ex0, ey0;         // global index of LDS column / row 0 (origin - halo)
};

// Tile of this launch block (XCD band map), or -1; its geometry from the balanced
// partition, with no memory access.
__device__ __forceinline__ int block_tile(const Tiling& T) {
  const int b = blockIdx.x;
  const int tile = (b % kNumXcd) * T.band + b / kNumXcd;
  return tile < T.nbx * T.nby ? tile : -1;
}

__device__ __forceinline__ TileGeo tile_geo(const Tiling& T, int tile) {
  const int bx = tile % T.nbx, by = tile / T.nbx;
  TileGeo t;
  t.x0 = bx * T.qx + min(bx, T.rx);
  t.sx = T.qx + (bx < T.rx ? 1 : 0);
  t.y0 = by * T.qy + min(by, T.ry);
  t.sy = T.qy + (by < T.ry ? 1 : 0);
  t.ex0 = t.x0 - HX;
  t.ey0 = t.y0 - HY;
  return t;
}

// Interior stencil coefficients of the operators a stage kernel applies, read once at
// kernel entry: d1[k] = c_k of the first-derivative-type operator (k = 1..4), dd / dt:
// c_0..c_2 of the dissipation operator and of its transpose.
struct Coefs {
  double d1[5], dd[3], dt[3];
};

__device__ __forceinline__ Coefs load_coefs(const Geo& g, int first_slot) {
  Coefs c;
  const Op& o1 = g.ops[first_slot];
  const Op& od = g.ops[SLOT_DISS];
  const Op& ot = g.ops[SLOT_DISS_T];
  c.d1[0] = 0.0;
#pragma unroll
  for (int k = 1; k <= 4; ++k) c.d1[k] = o1.c[kMaxHW + k];
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    c.dd[k] = od.c[kMaxHW + k];
    c.dt[k] = ot.c[kMaxHW + k];
  }
  return c;
}

// The 1-D grid values of a tile's columns (x) and rows (y) -- h, the norm weight and the
// coordinate -- held in LDS for the whole block: one thread loads each value at kernel
// entry, so the later phases read them from LDS instead of each lane fetching its own
// copy from memory.  Layout: [h | norm weight | coordinate], each TX column values
// then TY row values.
template <class S>
struct TileGrid {
  static constexpr int kRow = S::TX + S::TY;
  static constexpr int kSize = 3 * kRow;
  static_assert(kSize <= S::NT, "one grid value per thread");
  const double* s;
  __device__ double h0(int tx) const { return s[tx]; }
  __device__ double h1(int ty) const { return s[S::TX + ty]; }
  __device__ double hw0(int tx) const { return s[kRow + tx]; }
  __device__ double hw1(int ty) const { return s[kRow + S::TX + ty]; }
  __device__ double x0(int tx) const { return s[2 * kRow + tx]; }
  __device__ double x1(int ty) const { return s[2 * kRow + S::TX + ty]; }
};

// This thread's grid value (threads below TileGrid::kSize), to be stored at its index.
template <class S>
__device__ __forceinline__ double load_tile_grid(const Geo& g, const TileGeo& t) {
  constexpr int kRow = TileGrid<S>::kRow;
  const int e = threadIdx.x;
  if (e >= TileGrid<S>::kSize) return 0.0;
  const int what = e / kRow, i = e % kRow;
  const int d = i < S::TX ? 0 : 1;
  const int k = d == 0 ? i : i - S::TX;
  if (k >= (d == 0 ? t.sx : t.sy)) return 0.0;
  const int idx = (d == 0 ? t.x0 : t.y0) + k;
  // constant indices into the kernel-argument arrays (a run-time index would copy the
  // argument block to scratch memory)
  const double* src = what == 0 ? (d == 0 ? g.h[0] : g.h[1])
                    : what == 1 ? (d == 0 ? g.hw[0] : g.hw[1])
                                : (d == 0 ? g.x[0] : g.x[1]);
  return src[idx];
}

__device__ __forceinline__ bool closure(const Dir& dd, int i, int w) {
  return !dd.periodic && (i < w || i >= dd.len - w);
}

// global index of an extended (halo) index, or -1 outside a non-periodic domain
__device__ __forceinline__ int ext_index(const Dir& dd, int gi) {
  if (gi >= 0 && gi < dd.len) return gi;
  if (!dd.periodic) return -1;
  return gi < 0 ? gi + dd.len - dd.off : gi - dd.len + dd.off;
}

// Staged operand of this tile (all four components) in registers: stage_load issues
// every load of this thread before any is consumed; stage_store writes them to sq and,
// for the forward stage, the contravariant fluxes of the tile rows (x) and tile columns
// (y) to sfx / sfy.
template <class S>
struct StageRegs {
  double v[S::kStageIters][NV2];
};

// Staging slot `it` of this thread: extended-layout indices (ex, ey), the grid point
// (gi, gj) and whether the slot holds a staged point (the corners and the columns /
// rows beyond a partial tile are skipped).
template <class S>
__device__ __forceinline__ bool stage_slot(const Geo& g, const TileGeo& t, int it, int* ex,
                                           int* ey, int* gi, int* gj) {
  constexpr int EX = S::EX, EY = S::EY;
  const int e = threadIdx.x + it * S::NT;
  *ex = e % EX;
  *ey = e / EX;
  *gi = ext_index(g.dir[0], t.ex0 + *ex);
  *gj = ext_index(g.dir[1], t.ey0 + *ey);
  const bool rowx = *ey >= HY && *ey < HY + t.sy;
  const bool coly = *ex >= HX && *ex < HX + t.sx;
  return e < EX * EY && *ex < t.sx + 2 * HX && *ey < t.sy + 2 * HY && *gi >= 0 && *gj >= 0 &&
         (rowx || coly);
}

template <class S>
__device__ __forceinline__ void stage_load(const Geo& g, const TileGeo& t, const double* xi,
                                           StageRegs<S>& r) {
#pragma unroll
  for (int it = 0; it < S::kStageIters; ++it) {
    int ex, ey, gi, gj;
    if (stage_slot<S>(g, t, it, &ex, &ey, &gi, &gj)) {
      const long pp = gi + (long)g.n[0] * gj;
#pragma unroll
      for (int l = 0; l < NV2; ++l) r.v[it][l] = xi[l * g.N + pp];
    }
  }
}

template <class S, bool FLUX>
__device__ __forceinline__ void stage_store(const Geo& g, const TileGeo& t, const StageRegs<S>& r,
                                            const TileGrid<S>& tg, double* sq, double* sfx,
                                            double* sfy) {
  constexpr int TX = S::TX, TY = S::TY, EX = S::EX, EY = S::EY;
#pragma unroll
  for (int it = 0; it < S::kStageIters; ++it) {
    int ex, ey, gi, gj;
    if (!stage_slot<S>(g, t, it, &ex, &ey, &gi, &gj)) continue;
#pragma unroll
    for (int l = 0; l < NV2; ++l) sq[(l * EY + ey) * EX + ex] = r.v[it][l];
    if (FLUX) {
      const Prim<2> s = prim<2>(r.v[it], g.gamma);
      if (ey >= HY && ey < HY + t.sy) {
        double f[NV2];
        inviscid_flux<2>(r.v[it], s, 0, f);
        const double m = tg.h1(ey - HY);   // metric<2>(g, 0, .) of the row
#pragma unroll
        for (int l = 0; l < NV2; ++l) sfx[(l * TY + (ey - HY)) * EX + ex] = m * f[l];
      }
      if (ex >= HX && ex < HX + t.sx) {
        double f[NV2];
        inviscid_flux<2>(r.v[it], s, 1, f);
        const double m = tg.h0(ex - HX);   // metric<2>(g, 1, .) of the column
#pragma unroll
        for (int l = 0; l < NV2; ++l) sfy[(l * EY + ey) * TX + (ex - HX)] = m * f[l];
      }
    }
  }
}

// Boundary-closure row of an operator with compile-time block size BW x BD (SBP 4-8:
// first derivative 12 x 8, adjoint 16 x 12, dissipation transpose 7 x 5, composite
// 12 x 8).  Unrolled so the coefficient loads issue together.  f(c) returns the NV2
// operand values at closure column c (0-based from the boundary, mirrored on the right).
template <int BW, int BD, class F>
__device__ __forceinline__ void closure_row(const Op& op, int i, int len, F&& f, double* out) {
  const bool left = i < BD;
  const int m = left ? i : len - 1 - i;
  const double* b = left ? op.b1[m] : op.b2[m];
  const int base = left ? 0 : len - BW;
#pragma unroll
  for (int l = 0; l < NV2; ++l) out[l] = 0.0;
#pragma unroll
  for (int c = 0; c < BW; ++c) {
    double v[NV2];
    f(base + c, v);
    const double bc = b[c];
#pragma unroll
    for (int l = 0; l < NV2; ++l) out[l] += bc * v[l];
  }
}

// First-derivative-type operator (BW x BD closure) along x or y at a point of the tile,
// operand in LDS; row(l, e) returns component l at extended index e.
template <int BW, int BD, class Row>
__device__ __forceinline__ void deriv_lds(const Op& op, const double* ck, const Dir& dd, int i,
                                          int e0, Row&& row, double* out) {
  if (!dd.periodic && (i < BD || i >= dd.len - BD)) {
    closure_row<BW, BD>(op, i, dd.len,
                        [&](int idx, double* v) {
#pragma unroll
                          for (int l = 0; l < NV2; ++l) v[l] = row(l, idx - e0);
                        },
                        out);
    return;
  }
  const int e = i - e0;
#pragma unroll
  for (int l = 0; l < NV2; ++l) {
    double s = 0.0;
#pragma unroll
    for (int k = 1; k <= 4; ++k) s += ck[k] * (row(l, e + k) - row(l, e - k));
    out[l] = s;
  }
}

// D4 x at index i along a direction, with magudi's boundary rows: on a non-periodic
// direction rows 0,1 (and the last two) repeat the stencil centred two points inside
// (src/StencilOperatorImpl.f90:1928-1940), summed in the closure-row order.
template <class Row>
__device__ __forceinline__ double d4_at(const Dir& dd, int i, double c0, double c1, double c2,
                                        Row&& x) {
  if (!dd.periodic && (i < 2 || i >= dd.len - 2)) {
    const int ctr = i < 2 ? 2 : dd.len - 3;
    double s = 0.0;
    s += c2 * x(ctr - 2);
    s += c1 * x(ctr - 1);
    s += c0 * x(ctr);
    s += c1 * x(ctr + 1);
    s += c2 * x(ctr + 2);
    return s;
  }
  double s = 0.0;
  s += c1 * (x(i + 1) + x(i - 1));
  s += c2 * (x(i + 2) + x(i - 2));
  return s + c0 * x(i);
}

// t = -arc D4 x on the tile rows (x direction, columns x0-2 .. x0+sx+1) and tile columns
// (y direction, rows y0-2 .. y0+sy+1), for staged points inside the domain.
template <class S>
__device__ __forceinline__ void diss_intermediates(const Geo& g, const TileGeo& t, const Coefs& cf,
                                                   const TileGrid<S>& tg, const double* sq,
                                                   double* stx, double* sty) {
  constexpr int TX = S::TX, TY = S::TY, EX = S::EX, EY = S::EY, NT = S::NT;
  const double c0 = cf.dd[0], c1 = cf.dd[1], c2 = cf.dd[2];
  const Dir dx = g.dir[0], dy = g.dir[1];
  const int wx = t.sx + 4;
#pragma unroll
  for (int it = 0; it < (TY * (TX + 4) + NT - 1) / NT; ++it) {
    const int e = threadIdx.x + it * NT;
    if (e >= TY * (TX + 4)) continue;
    const int ex = e % (TX + 4) + 2, ey = e / (TX + 4) + HY;
    if (ex - 2 >= wx || ey - HY >= t.sy) continue;
    const int gi = t.ex0 + ex;
    if (!dx.periodic && (gi < 0 || gi >= dx.len)) continue;
    const double arc = fabs(tg.h1(ey - HY));
#pragma unroll
    for (int l = 0; l < NV2; ++l) {
      const double* r = sq + (l * EY + ey) * EX;
      const double s = d4_at(dx, gi, c0, c1, c2, [&](int k) { return r[k - t.ex0]; });
      stx[(l * TY + (ey - HY)) * (TX + 4) + (ex - 2)] = -arc * s;
    }
  }
  const int hy = t.sy + 4;
#pragma unroll
  for (int it = 0; it < ((TY + 4) * TX + NT - 1) / NT; ++it) {
    const int e = threadIdx.x + it * NT;
    if (e >= (TY + 4) * TX) continue;
    const int ex = e % TX + HX, ey = e / TX + 2;
    if (ex - HX >= t.sx || ey - 2 >= hy) continue;
    const int gj = t.ey0 + ey;
    if (!dy.periodic && (gj < 0 || gj >= dy.len)) continue;
    const double arc = fabs(tg.h0(ex - HX));
#pragma unroll
    for (int l = 0; l < NV2; ++l) {
      const double* c = sq + l * EY * EX + ex;
      const double s = d4_at(dy, gj, c0, c1, c2, [&](int k) { return c[(k - t.ey0) * EX]; });
      sty[(l * (TY + 4) + (ey - 2)) * TX + (ex - HX)] = -arc * s;
    }
  }
}

// Adds amount * (dissipation in x then in y) at one point: H^-1 D4^T t (closure rows of
// D4^T near non-periodic boundaries) or the composite operator on the staged operand.
template <class S>
__device__ __forceinline__ void add_dissipation(const Geo& g, const TileGeo& t, const Coefs& cf,
                                                const TileGrid<S>& tg, const double* sq,
                                                const double* stx, const double* sty, int tx,
                                                int ty, int gi, int gj, double amount, double* r) {
  constexpr int TX = S::TX, TY = S::TY, EX = S::EX, EY = S::EY;
  const Op& od = g.ops[SLOT_DISS];
  const Op& ot = g.ops[SLOT_DISS_T];
#pragma unroll
  for (int d = 0; d < 2; ++d) {
    double dv[NV2];
    const Dir& dr = g.dir[d];
    const int i = d == 0 ? gi : gj;
    if (g.composite) {
      auto row = [&](int l, int e) {
        return d == 0 ? sq[(l * EY + ty + HY) * EX + e] : sq[(l * EY + e) * EX + tx + HX];
      };
      const int e0 = d == 0 ? t.ex0 : t.ey0;
      if (!dr.periodic && (i < 8 || i >= dr.len - 8)) {
        closure_row<12, 8>(od, i, dr.len,
                           [&](int idx, double* v) {
#pragma unroll
                             for (int l = 0; l < NV2; ++l) v[l] = row(l, idx - e0);
                           },
                           dv);
      } else {
        const int e = i - e0;
#pragma unroll
        for (int l = 0; l < NV2; ++l) {
          double s = 0.0;
#pragma unroll
          for (int k = 1; k <= 4; ++k) s += od.c[kMaxHW + k] * (row(l, e + k) + row(l, e - k));
          dv[l] = s + od.c[kMaxHW] * row(l, e);
        }
      }
    } else {
      // t index: x uses columns from x0-2, y uses rows from y0-2
      auto tv = [&](int l, int idx) {
        return d == 0 ? stx[(l * TY + ty) * (TX + 4) + (idx - (t.x0 - 2))]
                      : sty[(l * (TY + 4) + (idx - (t.y0 - 2))) * TX + tx];
      };
      if (!dr.periodic && (i < 5 || i >= dr.len - 5)) {
        closure_row<7, 5>(ot, i, dr.len,
                          [&](int idx, double* v) {
#pragma unroll
                            for (int l = 0; l < NV2; ++l) v[l] = tv(l, idx);
                          },
                          dv);
      } else {
        const double c0 = cf.dt[0], c1 = cf.dt[1], c2 = cf.dt[2];
#pragma unroll
        for (int l = 0; l < NV2; ++l) {
          double s = 0.0;
          s += c1 * (tv(l, i + 1) + tv(l, i - 1));
          s += c2 * (tv(l, i + 2) + tv(l, i - 2));
          dv[l] = s + c0 * tv(l, i);
        }
      }
      // H^-1: the norm weight is exactly 1 away from the first / last boundary rows, where
      // the division is the identity and is skipped
      const double h = d == 0 ? tg.hw0(tx) : tg.hw1(ty);
      if (h != 1.0) {
#pragma unroll
        for (int l = 0; l < NV2; ++l) dv[l] = dv[l] / h;
      }
    }
#pragma unroll
    for (int l = 0; l < NV2; ++l) r[l] = r[l] + amount * dv[l];
  }
}

// Block sum: wave-level shuffles, then one LDS exchange of the per-wave sums.
template <int NT>
__device__ __forceinline__ double block_sum(double v, double* red) {
  for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
  __syncthreads();   // retire earlier readers of the scratch space
  const int w = threadIdx.x / 64;
  if ((threadIdx.x & 63) == 0) red[w] = v;
  __syncthreads();
  double s = 0.0;
  if (threadIdx.x == 0)
    for (int k = 0; k < NT / 64; ++k) s += red[k];
  return s;
}

__device__ __forceinline__ bool in_patch(const DPatch& pt, int i, int j) {
  return i >= pt.lo[0] && i <= pt.hi[0] && j >= pt.lo[1] && j <= pt.hi[1];
}

// -------------------------------------------------------------------------- forward
template <class S>
__global__ void __launch_bounds__(S::NT, S::kWavesPerEU) k_fwd_stage2d(Geo g, const Patches* __restrict__ P,
                                                       Tiling T, FwdStage a) {
  constexpr int TX = S::TX, TY = S::TY, EX = S::EX, EY = S::EY, NT = S::NT;
  constexpr int SFX = S::SFX, STX = S::STX;
  __shared__ double sq[S::SQ];
  __shared__ double sc[S::SSCR];
  __shared__ double sgrid[TileGrid<S>::kSize];
  const int tile = block_tile(T);
  if (tile < 0) return;   // block beyond the tiles of its XCD (uniform per block)
  const TileGeo t = tile_geo(T, tile);
  const long inst = blockIdx.y;
  const long NVN = (long)NV2 * g.N;
  const double* qi = a.qin + inst * NVN;
  double* sfx = sc;
  double* sfy = sc + SFX;
  const Coefs cf = load_coefs(g, SLOT_FIRST);
  const double gv = load_tile_grid<S>(g, t);
  const TileGrid<S> tg{sgrid};
  const unsigned pmask = T.pmask[tile];
  // RK operands of this thread's point, fetched before the stencil work
  double pq0[NV2], pb2[NV2];
  {
    const int tx0 = threadIdx.x % TX, ty0 = threadIdx.x / TX;
    if (tx0 < t.sx && ty0 < t.sy) {
      const long o = (t.x0 + tx0) + (long)g.n[0] * (t.y0 + ty0);
#pragma unroll
      for (int l = 0; l < NV2; ++l) {
        pq0[l] = (a.stage == 2 || a.stage == 3) ? a.q0[inst * NVN + l * g.N + o] : 0.0;
        pb2[l] = a.stage >= 2 ? a.b2[inst * NVN + l * g.N + o] : 0.0;
      }
    }
  }
  StageRegs<S> sr;
  stage_load<S>(g, t, qi, sr);
  // the grid values are needed by the flux metrics; this barrier waits for them only
  // (the operand loads stay in flight)
  if (threadIdx.x < TileGrid<S>::kSize) sgrid[threadIdx.x] = gv;
  __syncthreads();
  stage_store<S, true>(g, t, sr, tg, sq, sfx, sfy);
  __syncthreads();
  const Op& o1 = g.ops[SLOT_FIRST];
  double r[PPT][NV2];
  const int tx = threadIdx.x % TX;
#pragma unroll
  for (int k = 0; k < PPT; ++k) {
    const int ty = threadIdx.x / TX + k * (NT / TX);
    if (tx >= t.sx || ty >= t.sy) continue;
    const int gi = t.x0 + tx, gj = t.y0 + ty;
    double dfx[NV2], dfy[NV2];
    deriv_lds<12, 8>(o1, cf.d1, g.dir[0], gi, t.ex0,
              [&](int l, int e) { return sfx[(l * TY + ty) * EX + e]; }, dfx);
    deriv_lds<12, 8>(o1, cf.d1, g.dir[1], gj, t.ey0,
              [&](int l, int e) { return sfy[(l * EY + e) * TX + tx]; }, dfy);
#pragma unroll
    for (int l = 0; l < NV2; ++l) r[k][l] = 0.0 - (dfx[l] + dfy[l]);
  }
  double* stx = sc;
  double* sty = sc + STX;
  if (g.dissipation) {
    __syncthreads();   // fluxes consumed
    if (!g.composite) {
      diss_intermediates<S>(g, t, cf, tg, sq, stx, sty);
      __syncthreads();
    }
  }
#pragma unroll
  for (int k = 0; k < PPT; ++k) {
    const int ty = threadIdx.x / TX + k * (NT / TX);
    if (tx >= t.sx || ty >= t.sy) continue;
    const int gi = t.x0 + tx, gj = t.y0 + ty;
    if (g.dissipation)
      add_dissipation<S>(g, t, cf, tg, sq, stx, sty, tx, ty, gi, gj, g.diss_amount, r[k]);
    const long p = gi + (long)g.n[0] * gj;
    const double jac = 1.0 / (tg.h0(tx) * tg.h1(ty));   // jacobian<2>
    double qp[NV2];
#pragma unroll
    for (int l = 0; l < NV2; ++l) {
      r[k][l] = r[k][l] * jac;
      qp[l] = sq[(l * EY + ty + HY) * EX + tx + HX];
    }
    unsigned m = pmask;
    while (m) {
      const int ip = __ffs(m) - 1;
      m &= m - 1;
      const DPatch& pt = P->p[ip];
      if (!in_patch(pt, gi, gj)) continue;
      const int kp = (gi - pt.lo[0]) + pt.ls[0] * (gj - pt.lo[1]);
      switch (pt.type) {
        case MGD_PATCH_FAR_FIELD: {
          const double* A = pt.aplus + (long)kp * 16;
          const double* qt = pt.aplus + (long)pt.np * 16 + (long)kp * 4;
          const double sj = pt.sigma_inv * jac;
#pragma unroll
          for (int i = 0; i < NV2; ++i) {
            double s = 0.0;
#pragma unroll
            for (int c = 0; c < NV2; ++c) s += A[c * 4 + i] * (qp[c] - qt[c]);
            r[k][i] = r[k][i] - sj * s;
          }
          break;
        }
        case MGD_PATCH_SPONGE: {
          const double sp = pt.strength[kp];
          const double* qt = pt.aplus + (long)kp * 4;
#pragma unroll
          for (int l = 0; l < NV2; ++l) r[k][l] = r[k][l] - sp * (qp[l] - qt[l]);
          break;
        }
        case MGD_PATCH_ACTUATOR:
          if (a.control) r[k][3] = r[k][3] + pt.strength[kp] * a.control[inst * a.control_stride + kp];
          break;
        case MGD_PATCH_KOLMOGOROV:
          r[k][1] = r[k][1] + qp[0] * pt.strength[kp];
          break;
        default:
          break;
      }
    }
    for (int s = 0; s < a.nsrc; ++s) {
      const double dx0 = tg.x0(tx) - a.src[s].loc[0];
      const double dy0 = tg.x1(ty) - a.src[s].loc[1];
      double rr2 = 0.0;
      rr2 += dx0 * dx0;
      rr2 += dy0 * dy0;
      r[k][3] = r[k][3] + a.src_amp[inst * a.src_inst_stride + s] * exp(-a.src[s].gfac * rr2);
    }
    // RK4 stage update (substepForwardRK4, src/RK4IntegratorImpl.f90:115-156)
    const double dt = a.dt;
    double* b2 = a.b2 + inst * NVN;
    double qn[NV2];
#pragma unroll
    for (int l = 0; l < NV2; ++l) {
      const long o = l * g.N + p;
      const double rl = r[k][l];
      if (a.stage == 1) {
        b2[o] = qp[l] + div6(dt * rl);
        qn[l] = qp[l] + dt * rl / 2.0;
      } else if (a.stage == 2) {
        b2[o] = pb2[l] + div3(dt * rl);
        qn[l] = pq0[l] + dt * rl / 2.0;
      } else if (a.stage == 3) {
        b2[o] = pb2[l] + div3(dt * rl);
        qn[l] = pq0[l] + dt * rl;
      } else {
        qn[l] = pb2[l] + div6(dt * rl);
      }
      a.qout[inst * NVN + o] = qn[l];
      if (a.rhs_out) a.rhs_out[inst * NVN + o] = rl;
    }
    if (a.jpartial && a.ip_cost >= 0 && ((pmask >> a.ip_cost) & 1u)) {
      const DPatch& pt = P->p[a.ip_cost];
      if (in_patch(pt, gi, gj)) {
        const int kp = (gi - pt.lo[0]) + pt.ls[0] * (gj - pt.lo[1]);
        const Prim<2> s = prim<2>(qn, g.gamma);
        const double F = s.p - pt.aplus[kp];
        const double gn = (tg.hw1(ty) * (tg.hw0(tx) * 1.0)) * (tg.h0(tx) * tg.h1(ty));   // gnorm<2>
        a.jpartial[(inst * a.nslots + a.slot) * a.npts + kp] = 1.0 * gn * (F * F * pt.strength[kp]);
      }
    }
  }
}

// -------------------------------------------------------------------------- adjoint
template <class S>
__global__ void __launch_bounds__(S::NT, S::kWavesPerEU) k_adj_stage2d(Geo g, const Patches* __restrict__ P,
                                                       Tiling T, AdjStage a) {
  constexpr int TX = S::TX, EX = S::EX, EY = S::EY, NT = S::NT;
  constexpr int STX = S::STX;
  __shared__ double sq[S::SQ];
  __shared__ double sc[S::SSCR];
  __shared__ double sgrid[TileGrid<S>::kSize];
  const int tile = block_tile(T);
  if (tile < 0) return;   // block beyond the tiles of its XCD (uniform per block)
  const TileGeo t = tile_geo(T, tile);
  const long inst = blockIdx.y;
  const long NVN = (long)NV2 * g.N;
  const double* li = a.lam + inst * NVN;
  const double* qi = a.q + inst * NVN;
  const Coefs cf = load_coefs(g, SLOT_ADJ);
  const double gv = load_tile_grid<S>(g, t);
  const TileGrid<S> tg{sgrid};
  const unsigned pmask = T.pmask[tile];
  // forward stage state at this thread's point, fetched before the stencil work
  double pq[NV2];
  {
    const int tx0 = threadIdx.x % TX, ty0 = threadIdx.x / TX;
    if (tx0 < t.sx && ty0 < t.sy) {
      const long o = (t.x0 + tx0) + (long)g.n[0] * (t.y0 + ty0);
#pragma unroll
      for (int l = 0; l < NV2; ++l) pq[l] = qi[l * g.N + o];
    }
  }
  {
    StageRegs<S> sr;
    stage_load<S>(g, t, li, sr);
    if (threadIdx.x < TileGrid<S>::kSize) sgrid[threadIdx.x] = gv;
    stage_store<S, false>(g, t, sr, tg, sq, nullptr, nullptr);
  }
  double* stx = sc;
  double* sty = sc + STX;
  if (g.dissipation && !g.composite) {
    __syncthreads();
    diss_intermediates<S>(g, t, cf, tg, sq, stx, sty);
  }
  __syncthreads();
  const Op& oa = g.ops[SLOT_ADJ];
  const int tx = threadIdx.x % TX;
#pragma unroll
  for (int k = 0; k < PPT; ++k) {
    const int ty = threadIdx.x / TX + k * (NT / TX);
    if (tx >= t.sx || ty >= t.sy) continue;
    const int gi = t.x0 + tx, gj = t.y0 + ty;
    double T1[2][NV2];
    deriv_lds<16, 12>(oa, cf.d1, g.dir[0], gi, t.ex0,
              [&](int l, int e) { return sq[(l * EY + ty + HY) * EX + e]; }, T1[0]);
    deriv_lds<16, 12>(oa, cf.d1, g.dir[1], gj, t.ey0,
              [&](int l, int e) { return sq[(l * EY + e) * EX + tx + HX]; }, T1[1]);
    const long p = gi + (long)g.n[0] * gj;
    const Prim<2> s = prim<2>(pq, g.gamma);
    double r[NV2] = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
    for (int d = 0; d < 2; ++d) {
      double m[2], A[16];
      m[0] = d == 0 ? tg.h1(ty) : 0.0;   // metric<2>(g, 0, .)
      m[1] = d == 1 ? tg.h0(tx) : 0.0;   // metric<2>(g, 1, .)
      inviscid_jacobian<2>(m, s, g.gamma, A);
#pragma unroll
      for (int i = 0; i < NV2; ++i) {
        double acc = 0.0;
#pragma unroll
        for (int c = 0; c < NV2; ++c) acc += A[i * 4 + c] * T1[d][c];
        r[i] = r[i] + acc;
      }
    }
    // RK operands, fetched here so that they are not live through the derivative work
    double pl0[NV2], pb2[NV2];
#pragma unroll
    for (int l = 0; l < NV2; ++l) {
      pl0[l] = (a.stage == 3 || a.stage == 2) ? a.l0[inst * NVN + l * g.N + p] : 0.0;
      pb2[l] = a.stage <= 3 ? a.b2[inst * NVN + l * g.N + p] : 0.0;
    }
    if (g.dissipation)
      add_dissipation<S>(g, t, cf, tg, sq, stx, sty, tx, ty, gi, gj, -g.diss_amount, r);
    const double jac = 1.0 / (tg.h0(tx) * tg.h1(ty));   // jacobian<2>
    double lp[NV2];
#pragma unroll
    for (int l = 0; l < NV2; ++l) {
      r[l] = r[l] * jac;
      lp[l] = sq[(l * EY + ty + HY) * EX + tx + HX];
    }
    unsigned m = pmask;
    while (m) {
      const int ip = __ffs(m) - 1;
      m &= m - 1;
      const DPatch& pt = P->p[ip];
      if (!in_patch(pt, gi, gj)) continue;
      const int kp = (gi - pt.lo[0]) + pt.ls[0] * (gj - pt.lo[1]);
      switch (pt.type) {
        case MGD_PATCH_FAR_FIELD: {
          const double* A = pt.aplus + (long)kp * 16;
          const double sj = pt.sigma_inv * jac;
#pragma unroll
          for (int i = 0; i < NV2; ++i) {
            double acc = 0.0;
#pragma unroll
            for (int c = 0; c < NV2; ++c) acc += A[i * 4 + c] * lp[c];
            r[i] = r[i] + sj * acc;
          }
          break;
        }
        case MGD_PATCH_SPONGE: {
          const double sp = pt.strength[kp];
#pragma unroll
          for (int l = 0; l < NV2; ++l) r[l] = r[l] + sp * lp[l];
          break;
        }
        case MGD_PATCH_COST_TARGET: {
          // acoustic-noise adjoint forcing from the forward state at this point
          // (src/AcousticNoiseImpl.f90:208-280), scaled by the stage factor
          double af[NV2] = {0.0, 0.0, 0.0, 0.0};
          if (!a.zero_forcing) {
            const double F = -2.0 * pt.strength[kp] * 1.0 * (g.gamma - 1.0) * (s.p - pt.aplus[kp]);
            af[3] = F;
            af[1] = -s.u[0] * F;
            af[2] = -s.u[1] * F;
            double ke = 0.0;
            ke += s.u[0] * s.u[0];
            ke += s.u[1] * s.u[1];
            af[0] = 0.5 * ke * F;
          }
#pragma unroll
          for (int l = 0; l < NV2; ++l) r[l] = r[l] + a.forcing_factor * af[l];
          break;
        }
        case MGD_PATCH_ACTUATOR: {
          // ThermalActuator gradient and sensitivity on the pre-update adjoint state
          if (a.grad) a.grad[inst * a.grad_stride + kp] = lp[3] * (pt.strength[kp] * 1.0);
          if (a.spartial) {
            const double F = lp[3] * pt.strength[kp] * 1.0;
            const double gn = (tg.hw1(ty) * (tg.hw0(tx) * 1.0)) * (tg.h0(tx) * tg.h1(ty));   // gnorm<2>
            a.spartial[(inst * a.nslots + a.slot) * a.npts + kp] = 1.0 * gn * (F * F);
          }
          break;
        }
        case MGD_PATCH_KOLMOGOROV:
          r[0] = r[0] - lp[1] * pt.strength[kp];
          break;
      }
    }
    // RK4 adjoint stage update (substepAdjointRK4, src/RK4IntegratorImpl.f90:215-261)
    const double dt = a.dt;
    double* b2 = a.b2 + inst * NVN;
#pragma unroll
    for (int l = 0; l < NV2; ++l) {
      const long o = l * g.N + p;
      const double rl = r[l];
      double ln;
      if (a.stage == 4) {
        b2[o] = lp[l] - div6(dt * rl);
        ln = lp[l] - dt * rl / 2.0;
      } else if (a.stage == 3) {
        b2[o] = pb2[l] - div3(dt * rl);
        ln = pl0[l] - dt * rl / 2.0;
      } else if (a.stage == 2) {
        b2[o] = pb2[l] - div3(dt * rl);
        ln = pl0[l] - dt * rl;
      } else {
        ln = pb2[l] - div6(dt * rl);
      }
      a.lout[inst * NVN + o] = ln;
      if (a.rhs_out) a.rhs_out[inst * NVN + o] = rl;
    }
  }
}

// Stage sums of the per-point quadrature terms of one chunk of stages: one block of
// kSumThreads per (stage, instance), a fixed-order tree over the points.
constexpr int kSumThreads = 1024;
__global__ void __launch_bounds__(kSumThreads) k_stage_sums(const double* __restrict__ partial,
                                                            int nslots, int npts,
                                                            double* __restrict__ sums,
                                                            int nstages, int base) {
  __shared__ double red[kSumThreads / 64];
  const int s = blockIdx.x;
  const long inst = blockIdx.y;
  const double* src = partial + (inst * nslots + s) * (long)npts;
  double v = 0.0;
  for (int k = threadIdx.x; k < npts; k += kSumThreads) v += src[k];
  const double sum = block_sum<kSumThreads>(v, red);
  if (threadIdx.x == 0) sums[inst * nstages + base + s] = sum;
}

// acc[inst] += w[s] * sums[s] over the stages in time-marching order (the running time
// quadrature of Solver%runForward / runAdjoint); reverse = adjoint order.  One block
// per instance: the products are formed in parallel into LDS, one chunk at a time, and
// added by one thread in stage order (the order of the running sum).
__global__ void __launch_bounds__(kRedThreads) k_stage_accumulate(
    const double* __restrict__ sums, int nstages, const double* __restrict__ w,
    double* __restrict__ acc, int reverse) {
  constexpr int kChunk = 2048;
  __shared__ double prod[kChunk];
  const long inst = blockIdx.x;
  double total = acc[inst];
  for (int c0 = 0; c0 < nstages; c0 += kChunk) {
    const int n = min(kChunk, nstages - c0);
    __syncthreads();
    for (int k = threadIdx.x; k < n; k += kRedThreads) {
      const int s = reverse ? nstages - 1 - (c0 + k) : c0 + k;
      prod[k] = w[s] * sums[inst * nstages + s];
    }
    __syncthreads();
    if (threadIdx.x == 0) {
#pragma unroll 8
      for (int k = 0; k < n; ++k) total = total + prod[k];
    }
  }
  if (threadIdx.x == 0) acc[inst] = total;
}

// ------------------------------------------------------------------------ host side
using S32x16 = Shape<32, 16>;
using S16x32 = Shape<16, 32>;
using S16x16 = Shape<16, 16>;
using S24x16 = Shape<24, 16>;

static const int kShapeDims[kNumTileShapes][2] = {{32, 16}, {16, 32}, {16, 16}, {24, 16}};

int fused2d_shape_dims(int shape, int* tx, int* ty) {
  if (shape < 0 || shape >= kNumTileShapes) return MGD_ERR_ARG;
  *tx = kShapeDims[shape][0];
  *ty = kShapeDims[shape][1];
  return MGD_OK;
}

// Balanced partition of len points into ceil(len / T) tiles: tile b starts at
// b q + min(b, r) and holds q + (b < r) points.
static void balanced(int len, int T, std::vector<int>& org, std::vector<int>& size, int* q,
                     int* r) {
  const int n = (len + T - 1) / T;
  *q = len / n;
  *r = len % n;
  for (int b = 0; b < n; ++b) {
    org.push_back(b * *q + std::min(b, *r));
    size.push_back(*q + (b < *r ? 1 : 0));
  }
}

int fused2d_make_tiling(const Geo& g, const Patches& hp, int shape, HostTiling* out) {
  int tx, ty;
  if (fused2d_shape_dims(shape, &tx, &ty) != MGD_OK) return MGD_ERR_ARG;
  out->shape = shape;
  out->org.clear();
  out->size.clear();
  balanced(g.n[0], tx, out->org, out->size, &out->qx, &out->rx);
  out->nbx = (int)out->org.size();
  balanced(g.n[1], ty, out->org, out->size, &out->qy, &out->ry);
  out->nby = (int)out->org.size() - out->nbx;
  for (int d = 0; d < 2; ++d) {
    if (g.dir[d].periodic) continue;
    const int first = d == 0 ? 0 : out->nbx;
    const int last = d == 0 ? out->nbx - 1 : out->nbx + out->nby - 1;
    if (out->size[first] < kMinEdgeTile || out->size[last] < kMinEdgeTile) return MGD_ERR_UNSUPPORTED;
  }
  out->pmask.assign((size_t)out->nbx * out->nby, 0u);
  for (int by = 0; by < out->nby; ++by)
    for (int bx = 0; bx < out->nbx; ++bx) {
      const int xa = out->org[bx], xb = xa + out->size[bx] - 1;
      const int ya = out->org[out->nbx + by], yb = ya + out->size[out->nbx + by] - 1;
      unsigned m = 0;
      for (int ip = 0; ip < hp.count; ++ip) {
        const DPatch& pt = hp.p[ip];
        if (xb < pt.lo[0] || xa > pt.hi[0] || yb < pt.lo[1] || ya > pt.hi[1]) continue;
        m |= 1u << ip;
      }
      out->pmask[(size_t)by * out->nbx + bx] = m;
    }
  // XCD-aware block -> tile map.  Work-groups are dealt to the XCDs round-robin by
  // linear block index, and each XCD has its own L2: XCD x receives blocks x, x + 8, ...
  // which are mapped to one contiguous band of tiles (row-major), so neighbouring tiles
  // -- whose halos overlap -- share an L2.  Bands hold ceil(ntiles / 8) tiles; the spare
  // blocks of the last band map to -1 and exit at once.
  const int ntiles = out->nbx * out->nby;
  const int band = (ntiles + kNumXcd - 1) / kNumXcd;
  out->band = band;
  out->nblk = band * kNumXcd;
  out->tile_of_block.assign((size_t)out->nblk, -1);
  for (int b = 0; b < out->nblk; ++b) {
    const int pos = (b % kNumXcd) * band + b / kNumXcd;
    out->tile_of_block[(size_t)b] = pos < ntiles ? pos : -1;
  }
  return MGD_OK;
}

template <class S>
static void fwd_launch(const Geo& g, const Patches* P, const Tiling& T, const FwdStage& a,
                       hipStream_t st) {
  k_fwd_stage2d<S><<<dim3(T.nblk, g.ninst), S::NT, 0, st>>>(g, P, T, a);
}

template <class S>
static void adj_launch(const Geo& g, const Patches* P, const Tiling& T, const AdjStage& a,
                       hipStream_t st) {
  k_adj_stage2d<S><<<dim3(T.nblk, g.ninst), S::NT, 0, st>>>(g, P, T, a);
}

void launch_fwd_stage2d(const Geo& g, const Patches* P, const Tiling& T, const FwdStage& a,
                        hipStream_t st) {
  switch (T.shape) {
    case 0: fwd_launch<S32x16>(g, P, T, a, st); break;
    case 1: fwd_launch<S16x32>(g, P, T, a, st); break;
    case 2: fwd_launch<S16x16>(g, P, T, a, st); break;
    default: fwd_launch<S24x16>(g, P, T, a, st); break;
  }
}

void launch_adj_stage2d(const Geo& g, const Patches* P, const Tiling& T, const AdjStage& a,
                        hipStream_t st) {
  switch (T.shape) {
    case 0: adj_launch<S32x16>(g, P, T, a, st); break;
    case 1: adj_launch<S16x32>(g, P, T, a, st); break;
    case 2: adj_launch<S16x16>(g, P, T, a, st); break;
    default: adj_launch<S24x16>(g, P, T, a, st); break;
  }
}

void launch_stage_reduce(const double* partial, int nslots, int nfill, int npts, double* sums,
                         int nstages, int base, int ninst, hipStream_t st) {
  if (nfill <= 0 || npts <= 0) return;
  k_stage_sums<<<dim3(nfill, ninst), kSumThreads, 0, st>>>(partial, nslots, npts, sums, nstages,
                                                           base);
}

void launch_stage_accumulate(const double* sums, int nstages, const double* w, double* acc,
                             int ninst, int reverse, hipStream_t st) {
  if (nstages <= 0) return;
  k_stage_accumulate<<<ninst, kRedThreads, 0, st>>>(sums, nstages, w, acc, reverse);
}

}  // namespace mgd
