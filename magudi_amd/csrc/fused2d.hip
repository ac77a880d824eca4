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

// Tile shape: TX x TY points and CS threads per point.  Every thread stages operand
// points; in the stencil and point-wise phases each of the CS threads of a point forms
// NC = 4 / CS of its components (thread group h = threadIdx / NP owns components
// h NC .. h NC + NC - 1; a group is a whole number of waves), so CS = 2 gives a tile
// twice the waves, each with half the stencil work.  LDS (doubles): operand [4][EY][EX];
// scratch for the fluxes, reused for the dissipation intermediates.
template <int TX_, int TY_, int CS_ = 1>
struct Shape {
  static constexpr int TX = TX_, TY = TY_, CS = CS_;
  static constexpr int NP = TX * TY, NT = NP * CS, NC = NV2 / CS;
  static_assert(NP % 64 == 0 && NV2 % CS == 0, "component groups are whole waves");
  static constexpr int EX = TX + 2 * HX, EY = TY + 2 * HY;
  // staged points: the tile rows over the extended width, plus the halo rows over the tile
  // columns (the halo corners are never read)
  static constexpr int kStaged = TY * EX + 2 * HY * TX;
  static constexpr int kStageIters = (kStaged + NT - 1) / NT;   // staged points per thread
  static constexpr int SQ = NV2 * EY * EX;
  static constexpr int SFX = NV2 * TY * EX, SFY = NV2 * EY * TX;
  static constexpr int STX = NV2 * TY * (TX + 4), STY = NV2 * (TY + 4) * TX;
  static constexpr int SSCR = (SFX + SFY) > (STX + STY) ? (SFX + SFY) : (STX + STY);
  static_assert(CS == 1 || SSCR >= 2 * NV2 * NP, "the scratch space holds the adjoint exchange");
  // waves per SIMD that the LDS use allows (160 KiB per CU); the register budget is
  // sized to it
  static constexpr int kLdsBytes = 8 * (SQ + SSCR + 3 * (TX + TY));
  static constexpr int kBlocksPerCu = (160 * 1024) / kLdsBytes;
  static constexpr int kWaves = kBlocksPerCu * NT / 64 / 4;
  static constexpr int kWavesPerEU = kWaves < 1 ? 1 : kWaves > 8 ? 8 : kWaves;
  // forward stage: form the dissipation intermediates in registers alongside the flux
  // derivatives (where the register budget has room for them)
  static constexpr bool kDissRegs = kWavesPerEU <= 3 || CS > 1;
  // the tile splits into 8 x 8 blocks of points, one wave each (see point_of_lane)
  static constexpr bool kQuads = TX % 8 == 0 && TY % 8 == 0;
};

struct TileGeo {
  int x0, y0, sx, sy;   // origin and size of the tile
  int ex0, ey0;         // global index of LDS column / row 0 (origin - halo)
};

// Tile of this launch block (XCD band map), or -1; its geometry from the balanced
// partition, with no memory access.
__device__ __forceinline__ int block_tile(const Tiling& T, int b) {
  const int tile = (b % kNumXcd) * T.band + b / kNumXcd;
  return tile < T.nbx * T.nby ? tile : -1;
}

// Point (tx, ty) of a thread of the forward stage (lane = index within its component
// group).  Row-major in general: a wave covers 64 / TX whole rows.  In a tile on a
// non-periodic x boundary the boundary-closure columns (the first / last 8 points) would
// split every such wave into closure and interior lanes, and a wave runs both paths;
// there each wave covers an 8 x 8 block of points instead, so the closure columns of the
// first-derivative operator fill whole waves.  Which thread forms a point changes no
// value.
template <class S>
__device__ __forceinline__ void point_of_lane(const Geo& g, const Tiling& T, int tile, int lane,
                                              int* tx, int* ty) {
  const int bx = tile % T.nbx;
  if (S::kQuads && !g.dir[0].periodic && (bx == 0 || bx == T.nbx - 1)) {
    const int l = lane & 63, w = lane >> 6;
    constexpr int QX = S::TX / 8;
    *tx = (l & 7) + 8 * (w % QX);
    *ty = (l >> 3) + 8 * (w / QX);
  } else {
    *tx = lane % S::TX;
    *ty = lane / S::TX;
  }
}

__device__ __forceinline__ TileGeo tile_geo(const Tiling& T, int tile) {
  const int bx = tile % T.nbx, by = tile / T.nbx;
  TileGeo t;
  t.x0 = bx * T.qx + min(bx, T.lx) + max(0, bx - (T.nbx - T.rx));
  t.sx = T.qx + (bx < T.lx || bx >= T.nbx - T.rx ? 1 : 0);
  t.y0 = by * T.qy + min(by, T.ly) + max(0, by - (T.nby - T.ry));
  t.sy = T.qy + (by < T.ly || by >= T.nby - T.ry ? 1 : 0);
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
// (gi, gj) and whether the slot holds a staged point.  Slots enumerate the tile rows
// over the extended width first, then the halo rows over the tile columns (the halo
// rows below a partial tile follow its last row); slots beyond a partial tile, and
// halo points outside a non-periodic domain, hold none.
template <class S>
__device__ __forceinline__ bool stage_slot(const Geo& g, const TileGeo& t, int it, int* ex,
                                           int* ey, int* gi, int* gj) {
  constexpr int TX = S::TX, TY = S::TY, EX = S::EX;
  const int e = threadIdx.x + it * S::NT;
  bool in;
  if (e < TY * EX) {
    const int r = e / EX;
    *ex = e % EX;
    *ey = HY + r;
    in = r < t.sy && *ex < t.sx + 2 * HX;
  } else {
    const int e2 = e - TY * EX;
    const int hr = e2 / TX, c = e2 % TX;
    *ex = HX + c;
    *ey = hr < HY ? hr : HY + t.sy + (hr - HY);
    in = e2 < 2 * HY * TX && c < t.sx;
  }
  *gi = ext_index(g.dir[0], t.ex0 + *ex);
  *gj = ext_index(g.dir[1], t.ey0 + *ey);
  return in && *gi >= 0 && *gj >= 0;
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
// 12 x 8) for NC components.  Unrolled so the coefficient loads issue together.  f(c, v)
// writes the NC operand values at closure column c (0-based from the boundary,
// mirrored on the right).
template <int BW, int BD, int NC, class F>
__device__ __forceinline__ void closure_row(const Op& op, int i, int len, F&& f, double* out) {
  const bool left = i < BD;
  const int m = left ? i : len - 1 - i;
  const double* b = left ? op.b1[m] : op.b2[m];
  const int base = left ? 0 : len - BW;
#pragma unroll
  for (int j = 0; j < NC; ++j) out[j] = 0.0;
#pragma unroll
  for (int c = 0; c < BW; ++c) {
    double v[NC];
    f(base + c, v);
    const double bc = b[c];
#pragma unroll
    for (int j = 0; j < NC; ++j) out[j] += bc * v[j];
  }
}

// First-derivative-type operator (BW x BD closure) along x or y at a point of the tile
// for components c0 .. c0 + NC - 1, operand in LDS; row(l, e) returns component l at
// extended index e.
template <int BW, int BD, int NC, class Row>
__device__ __forceinline__ void deriv_lds(const Op& op, const double* ck, const Dir& dd, int i,
                                          int e0, int c0, Row&& row, double* out) {
  if (!dd.periodic && (i < BD || i >= dd.len - BD)) {
    closure_row<BW, BD, NC>(op, i, dd.len,
                            [&](int idx, double* v) {
#pragma unroll
                              for (int j = 0; j < NC; ++j) v[j] = row(c0 + j, idx - e0);
                            },
                            out);
    return;
  }
  const int e = i - e0;
#pragma unroll
  for (int j = 0; j < NC; ++j) {
    double s = 0.0;
#pragma unroll
    for (int k = 1; k <= 4; ++k) s += ck[k] * (row(c0 + j, e + k) - row(c0 + j, e - k));
    out[j] = s;
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

// Slots of diss_intermediates_to per thread, and the intermediates of one thread held in
// registers (so that they can be formed before the flux space they are stored over is
// free).  With one thread per point the x positions (plane PX) and then the y positions
// (plane PY) are dealt to the threads together, slot it of lane holding position
// lane + it NP: every thread forms at most NS of them (3 for 16 x 16 tiles, where
// dealing the planes one after the other gives the first wave 4 and the others 2).  With
// component groups (CS > 1) the planes are dealt one after the other (slots 0 .. IX - 1
// for x, IX .. for y), which keeps their register budget.
template <class S>
struct DissSlots {
  static constexpr int PX = S::TY * (S::TX + 4), PY = (S::TY + 4) * S::TX;   // planes
  static constexpr bool kJoint = S::CS == 1;
  static constexpr int IX = (PX + S::NP - 1) / S::NP, IY = (PY + S::NP - 1) / S::NP;
  static constexpr int NS = kJoint ? (PX + PY + S::NP - 1) / S::NP : IX + IY;
  double v[NS][S::NC];
  int base[NS];   // LDS index in its plane, or -1
  bool y[NS];     // position of the y plane
};

// t = -arc D4 x on the tile rows (x direction, columns x0-2 .. x0+sx+1) and tile columns
// (y direction, rows y0-2 .. y0+sy+1), for staged points inside the domain, components
// c0 .. c0 + NC - 1; the positions are dealt to the NP threads of the component group
// (lane = index within the group) as DissSlots describes.  Each value goes to
// sink_x / sink_y (it, j, base, v): slot `it` of this thread, component c0 + j, LDS
// index base + (c0 + j) * plane of the x or y plane.
template <class S, class SinkX, class SinkY>
__device__ __forceinline__ void diss_intermediates_to(const Geo& g, const TileGeo& t,
                                                      const Coefs& cf, const TileGrid<S>& tg,
                                                      const double* sq, int lane, int c0,
                                                      SinkX&& sink_x, SinkY&& sink_y) {
  constexpr int TX = S::TX, EX = S::EX, EY = S::EY, NP = S::NP, NC = S::NC;
  constexpr int PX = DissSlots<S>::PX, PY = DissSlots<S>::PY;
  const double k0 = cf.dd[0], k1 = cf.dd[1], k2 = cf.dd[2];
  const Dir dx = g.dir[0], dy = g.dir[1];
  const int wx = t.sx + 4, hy = t.sy + 4;
  constexpr bool kJoint = DissSlots<S>::kJoint;
  constexpr int IX = DissSlots<S>::IX;
#pragma unroll
  for (int it = 0; it < DissSlots<S>::NS; ++it) {
    // joint: position lane + it NP of the concatenated planes; else x slots, then y
    const int e = kJoint ? lane + it * NP : it < IX ? lane + it * NP : PX + lane + (it - IX) * NP;
    if (e >= PX + PY || (!kJoint && it < IX && e >= PX)) continue;
    // slot it holds positions it NP .. it NP + NP - 1: only a slot that straddles the
    // end of the x plane needs both branches (it is a constant once unrolled)
    if (kJoint ? ((it + 1) * NP <= PX || (it * NP < PX && e < PX)) : it < IX) {
      const int ex = e % (TX + 4) + 2, ey = e / (TX + 4) + HY;
      if (ex - 2 >= wx || ey - HY >= t.sy) continue;
      const int gi = t.ex0 + ex;
      if (!dx.periodic && (gi < 0 || gi >= dx.len)) continue;
      const double arc = fabs(tg.h1(ey - HY));
#pragma unroll
      for (int j = 0; j < NC; ++j) {
        const int l = c0 + j;
        const double* r = sq + (l * EY + ey) * EX;
        const double v = d4_at(dx, gi, k0, k1, k2, [&](int k) { return r[k - t.ex0]; });
        sink_x(it, j, (ey - HY) * (TX + 4) + (ex - 2), -arc * v);
      }
    } else {
      const int e2 = e - PX;
      const int ex = e2 % TX + HX, ey = e2 / TX + 2;
      if (ex - HX >= t.sx || ey - 2 >= hy) continue;
      const int gj = t.ey0 + ey;
      if (!dy.periodic && (gj < 0 || gj >= dy.len)) continue;
      const double arc = fabs(tg.h0(ex - HX));
#pragma unroll
      for (int j = 0; j < NC; ++j) {
        const int l = c0 + j;
        const double* c = sq + l * EY * EX + ex;
        const double v = d4_at(dy, gj, k0, k1, k2, [&](int k) { return c[(k - t.ey0) * EX]; });
        sink_y(it, j, (ey - 2) * TX + (ex - HX), -arc * v);
      }
    }
  }
}

// The intermediates straight into LDS.
template <class S>
__device__ __forceinline__ void diss_intermediates(const Geo& g, const TileGeo& t, const Coefs& cf,
                                                   const TileGrid<S>& tg, const double* sq,
                                                   double* stx, double* sty, int lane, int c0) {
  diss_intermediates_to<S>(
      g, t, cf, tg, sq, lane, c0,
      [&](int, int j, int base, double v) { stx[(c0 + j) * DissSlots<S>::PX + base] = v; },
      [&](int, int j, int base, double v) { sty[(c0 + j) * DissSlots<S>::PY + base] = v; });
}

// The intermediates into registers, then (diss_store) into LDS.
template <class S>
__device__ __forceinline__ void diss_compute(const Geo& g, const TileGeo& t, const Coefs& cf,
                                             const TileGrid<S>& tg, const double* sq, int lane,
                                             int c0, DissSlots<S>& d) {
#pragma unroll
  for (int it = 0; it < DissSlots<S>::NS; ++it) {
    d.base[it] = -1;
    d.y[it] = false;
  }
  diss_intermediates_to<S>(
      g, t, cf, tg, sq, lane, c0,
      [&](int it, int j, int base, double v) { d.base[it] = base; d.y[it] = false; d.v[it][j] = v; },
      [&](int it, int j, int base, double v) { d.base[it] = base; d.y[it] = true; d.v[it][j] = v; });
}

template <class S>
__device__ __forceinline__ void diss_store(const DissSlots<S>& d, int c0, double* stx, double* sty) {
#pragma unroll
  for (int it = 0; it < DissSlots<S>::NS; ++it) {
    if (d.base[it] < 0) continue;
    if (d.y[it]) {
#pragma unroll
      for (int j = 0; j < S::NC; ++j) sty[(c0 + j) * DissSlots<S>::PY + d.base[it]] = d.v[it][j];
    } else {
#pragma unroll
      for (int j = 0; j < S::NC; ++j) stx[(c0 + j) * DissSlots<S>::PX + d.base[it]] = d.v[it][j];
    }
  }
}

// Dissipation terms at one point for components c0 .. c0 + NC - 1: w[d][j] = amount *
// (H^-1 D4^T t) along direction d (closure rows of D4^T near non-periodic boundaries),
// or amount * (composite operator on the staged operand).  The caller adds w[0] then
// w[1].
template <class S>
__device__ __forceinline__ void diss_terms(const Geo& g, const TileGeo& t, const Coefs& cf,
                                           const TileGrid<S>& tg, const double* sq,
                                           const double* stx, const double* sty, int tx, int ty,
                                           int gi, int gj, int c0, double amount,
                                           double (*w)[S::NC]) {
  constexpr int TX = S::TX, TY = S::TY, EX = S::EX, EY = S::EY, NC = S::NC;
  const Op& od = g.ops[SLOT_DISS];
  const Op& ot = g.ops[SLOT_DISS_T];
#pragma unroll
  for (int d = 0; d < 2; ++d) {
    double dv[NC];
    const Dir& dr = g.dir[d];
    const int i = d == 0 ? gi : gj;
    if (g.composite) {
      auto row = [&](int l, int e) {
        return d == 0 ? sq[(l * EY + ty + HY) * EX + e] : sq[(l * EY + e) * EX + tx + HX];
      };
      const int e0 = d == 0 ? t.ex0 : t.ey0;
      if (!dr.periodic && (i < 8 || i >= dr.len - 8)) {
        closure_row<12, 8, NC>(od, i, dr.len,
                               [&](int idx, double* v) {
#pragma unroll
                                 for (int j = 0; j < NC; ++j) v[j] = row(c0 + j, idx - e0);
                               },
                               dv);
      } else {
        const int e = i - e0;
#pragma unroll
        for (int j = 0; j < NC; ++j) {
          double v = 0.0;
#pragma unroll
          for (int k = 1; k <= 4; ++k)
            v += od.c[kMaxHW + k] * (row(c0 + j, e + k) + row(c0 + j, e - k));
          dv[j] = v + od.c[kMaxHW] * row(c0 + j, e);
        }
      }
    } else {
      // t index: x uses columns from x0-2, y uses rows from y0-2
      auto tv = [&](int l, int idx) {
        return d == 0 ? stx[(l * TY + ty) * (TX + 4) + (idx - (t.x0 - 2))]
                      : sty[(l * (TY + 4) + (idx - (t.y0 - 2))) * TX + tx];
      };
      if (!dr.periodic && (i < 5 || i >= dr.len - 5)) {
        closure_row<7, 5, NC>(ot, i, dr.len,
                              [&](int idx, double* v) {
#pragma unroll
                                for (int j = 0; j < NC; ++j) v[j] = tv(c0 + j, idx);
                              },
                              dv);
      } else {
        const double k0 = cf.dt[0], k1 = cf.dt[1], k2 = cf.dt[2];
#pragma unroll
        for (int j = 0; j < NC; ++j) {
          double v = 0.0;
          v += k1 * (tv(c0 + j, i + 1) + tv(c0 + j, i - 1));
          v += k2 * (tv(c0 + j, i + 2) + tv(c0 + j, i - 2));
          dv[j] = v + k0 * tv(c0 + j, i);
        }
      }
      // H^-1: the norm weight is exactly 1 away from the first / last boundary rows, where
      // the division is the identity and is skipped
      const double h = d == 0 ? tg.hw0(tx) : tg.hw1(ty);
      if (h != 1.0) {
#pragma unroll
        for (int j = 0; j < NC; ++j) dv[j] = dv[j] / h;
      }
    }
#pragma unroll
    for (int j = 0; j < NC; ++j) w[d][j] = amount * dv[j];
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

// x[c0 + j] of a register array with the wave-uniform group offset c0 (a multiple of
// NC), formed by selects on constant indices (a run-time register index would move the
// array to scratch memory).
template <int NC>
__device__ __forceinline__ double comp(const double* x, int c0, int j) {
  double v = x[j];
#pragma unroll
  for (int k = 1; k * NC + j < NV2; ++k)
    if (c0 == k * NC) v = x[k * NC + j];
  return v;
}

// Rows C0 .. C0 + NC - 1 of sum_d A_d^T-side products of the adjoint rhs:
// r[j] = sum_d sum_c A_d[C0 + j][c] T1[d][c] (computeRhsAdjoint's Jacobian part for 2-D
// Euler), with the row offset a compile-time constant so only those rows are formed.
template <int C0, int NC>
__device__ __forceinline__ void adj_jacobian_rows(const double* mx, const Prim<2>& s, double gamma,
                                                  const double (*T1)[NV2], double* r) {
#pragma unroll
  for (int j = 0; j < NC; ++j) r[j] = 0.0;
#pragma unroll
  for (int d = 0; d < 2; ++d) {
    double m[2], A[16];
    m[0] = d == 0 ? mx[0] : 0.0;
    m[1] = d == 1 ? mx[1] : 0.0;
    inviscid_jacobian<2>(m, s, gamma, A);
#pragma unroll
    for (int j = 0; j < NC; ++j) {
      double acc = 0.0;
#pragma unroll
      for (int c = 0; c < NV2; ++c) acc += A[(C0 + j) * 4 + c] * T1[d][c];
      r[j] = r[j] + acc;
    }
  }
}

__device__ __forceinline__ bool in_patch(const DPatch& pt, int i, int j) {
  return i >= pt.lo[0] && i <= pt.hi[0] && j >= pt.lo[1] && j <= pt.hi[1];
}

// The LDS of a stage body: the staged operand, the flux / scratch space and the tile's
// grid values.  One allocation per kernel, shared by the forward and adjoint bodies where
// one kernel runs both (k_dual_stage2d).
template <class S>
struct StageLds {
  double q[S::SQ];
  double c[S::SSCR];
  double grid[TileGrid<S>::kSize];
};

template <class S>
__device__ __forceinline__ StageLds<S>& stage_lds() {
  __shared__ StageLds<S> lds;
  return lds;
}

// -------------------------------------------------------------------------- forward
// One forward stage on tile block `blk` of instance `inst` (uniform per work-group).
template <class S>
__device__ __forceinline__ void fwd_stage_body(const Geo& g, const Patches* __restrict__ P,
                                               const Tiling& T, const FwdStage& a, int blk,
                                               long inst) {
  constexpr int TX = S::TX, TY = S::TY, EX = S::EX, EY = S::EY, NP = S::NP, NC = S::NC;
  constexpr int SFX = S::SFX, STX = S::STX;
  StageLds<S>& lds = stage_lds<S>();
  double(&sq)[S::SQ] = lds.q;
  double(&sc)[S::SSCR] = lds.c;
  double(&sgrid)[TileGrid<S>::kSize] = lds.grid;
  const int tile = block_tile(T, blk);
  if (tile < 0) return;   // block beyond the tiles of its XCD (uniform per block)
  const TileGeo t = tile_geo(T, tile);
  const long NVN = (long)NV2 * g.N;
  const double* qi = a.qin + inst * NVN;
  double* sfx = sc;
  double* sfy = sc + SFX;
  const Coefs cf = load_coefs(g, SLOT_FIRST);
  const double gv = load_tile_grid<S>(g, t);
  const TileGrid<S> tg{sgrid};
  const unsigned pmask = T.pmask[tile];
  // this thread's point and component group
  const int lane = threadIdx.x % NP, c0 = (threadIdx.x / NP) * NC;
  int tx, ty;
  point_of_lane<S>(g, T, tile, lane, &tx, &ty);
  const bool mine = tx < t.sx && ty < t.sy;
  const int gi = t.x0 + tx, gj = t.y0 + ty;
  const long p = gi + (long)g.n[0] * gj;
  StageRegs<S> sr;
  stage_load<S>(g, t, qi, sr);
  // the grid values are needed by the flux metrics; this barrier waits for them only
  // (the operand loads stay in flight)
  if (threadIdx.x < TileGrid<S>::kSize) sgrid[threadIdx.x] = gv;
  __syncthreads();
  stage_store<S, true>(g, t, sr, tg, sq, sfx, sfy);
  __syncthreads();
  const Op& o1 = g.ops[SLOT_FIRST];
  // the dissipation intermediates are formed from the staged operand together with the
  // flux derivatives and stored once the flux space they overwrite is free
  DissSlots<S> dsl;
  const bool two_level = g.dissipation && !g.composite;
  if (S::kDissRegs && two_level) diss_compute<S>(g, t, cf, tg, sq, lane, c0, dsl);
  double r[NC];
  if (mine) {
    double dfx[NC], dfy[NC];
    deriv_lds<12, 8, NC>(o1, cf.d1, g.dir[0], gi, t.ex0, c0,
                         [&](int l, int e) { return sfx[(l * TY + ty) * EX + e]; }, dfx);
    deriv_lds<12, 8, NC>(o1, cf.d1, g.dir[1], gj, t.ey0, c0,
                         [&](int l, int e) { return sfy[(l * EY + e) * TX + tx]; }, dfy);
#pragma unroll
    for (int j = 0; j < NC; ++j) r[j] = 0.0 - (dfx[j] + dfy[j]);
  }
  // RK operands of this thread's components, fetched after the flux derivatives: every
  // block issues its operand loads at once at entry, and these (two thirds of the entry
  // bytes in stages 2 and 3) would lengthen that burst; here they arrive during the
  // dissipation terms
  double pq0[NC], pb2[NC];
  if (mine) {
#pragma unroll
    for (int j = 0; j < NC; ++j) {
      const long o = (c0 + j) * g.N + p;
      pq0[j] = (a.stage == 2 || a.stage == 3) ? a.q0[inst * NVN + o] : 0.0;
      pb2[j] = a.stage >= 2 ? a.b2[inst * NVN + o] : 0.0;
    }
  }
  double* stx = sc;
  double* sty = sc + STX;
  if (two_level) {
    __syncthreads();   // fluxes consumed
    if (S::kDissRegs)
      diss_store<S>(dsl, c0, stx, sty);
    else
      diss_intermediates<S>(g, t, cf, tg, sq, stx, sty, lane, c0);
    __syncthreads();
  }
  double qn[NC];
  if (mine) {
    if (g.dissipation) {
      double w[2][NC];
      diss_terms<S>(g, t, cf, tg, sq, stx, sty, tx, ty, gi, gj, c0, g.diss_amount, w);
#pragma unroll
      for (int d = 0; d < 2; ++d)
#pragma unroll
        for (int j = 0; j < NC; ++j) r[j] = r[j] + w[d][j];
    }
    const double jac = 1.0 / (tg.h0(tx) * tg.h1(ty));   // jacobian<2>
    double qp[NV2];
#pragma unroll
    for (int l = 0; l < NV2; ++l) qp[l] = sq[(l * EY + ty + HY) * EX + tx + HX];
#pragma unroll
    for (int j = 0; j < NC; ++j) r[j] = r[j] * jac;
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
          for (int j = 0; j < NC; ++j) {
            const int i = c0 + j;
            double v = 0.0;
#pragma unroll
            for (int c = 0; c < NV2; ++c) v += A[c * 4 + i] * (qp[c] - qt[c]);
            r[j] = r[j] - sj * v;
          }
          break;
        }
        case MGD_PATCH_SPONGE: {
          const double sp = pt.strength[kp];
          const double* qt = pt.aplus + (long)kp * 4;
#pragma unroll
          for (int j = 0; j < NC; ++j) r[j] = r[j] - sp * (comp<NC>(qp, c0, j) - qt[c0 + j]);
          break;
        }
        case MGD_PATCH_ACTUATOR:
#pragma unroll
          for (int j = 0; j < NC; ++j)
            if (c0 + j == 3 && a.control)
              r[j] = r[j] + pt.strength[kp] * a.control[inst * a.control_stride + kp];
          break;
        case MGD_PATCH_KOLMOGOROV:
#pragma unroll
          for (int j = 0; j < NC; ++j)
            if (c0 + j == 1) r[j] = r[j] + qp[0] * pt.strength[kp];
          break;
        default:
          break;
      }
    }
#pragma unroll
    for (int j = 0; j < NC; ++j) {
      if (c0 + j != 3) continue;
      for (int s = 0; s < a.nsrc; ++s) {
        const double dx0 = tg.x0(tx) - a.src[s].loc[0];
        const double dy0 = tg.x1(ty) - a.src[s].loc[1];
        double rr2 = 0.0;
        rr2 += dx0 * dx0;
        rr2 += dy0 * dy0;
        r[j] = r[j] + a.src_amp[inst * a.src_inst_stride + s] * exp(-a.src[s].gfac * rr2);
      }
    }
    // RK4 stage update (substepForwardRK4, src/RK4IntegratorImpl.f90:115-156).  The
    // stage outputs are written with non-temporal stores: they stream to HBM during the
    // launch instead of waiting in L2 for the write-back at its end
    const double dt = a.dt;
    double* b2 = a.b2 + inst * NVN;
#pragma unroll
    for (int j = 0; j < NC; ++j) {
      const int l = c0 + j;
      const long o = l * g.N + p;
      const double rl = r[j];
      if (a.stage == 1) {
        const double ql = comp<NC>(qp, c0, j);
        __builtin_nontemporal_store(ql + div6(dt * rl), &b2[o]);
        qn[j] = ql + dt * rl / 2.0;
      } else if (a.stage == 2) {
        __builtin_nontemporal_store(pb2[j] + div3(dt * rl), &b2[o]);
        qn[j] = pq0[j] + dt * rl / 2.0;
      } else if (a.stage == 3) {
        __builtin_nontemporal_store(pb2[j] + div3(dt * rl), &b2[o]);
        qn[j] = pq0[j] + dt * rl;
      } else {
        qn[j] = pb2[j] + div6(dt * rl);
      }
      __builtin_nontemporal_store(qn[j], &a.qout[inst * NVN + o]);
      if (a.rhs_out) a.rhs_out[inst * NVN + o] = rl;
    }
  }
  // functional term of the new stage state at cost-target points (needs all components:
  // with CS > 1 they are exchanged through the operand space)
  if (a.jpartial && a.ip_cost >= 0 && ((pmask >> a.ip_cost) & 1u)) {   // uniform per block
    double qa[NV2];
    if (S::CS == 1) {
#pragma unroll
      for (int j = 0; j < NC; ++j) qa[j] = qn[j];
    } else {
      __syncthreads();   // every read of the staged operand is done
      if (mine) {
#pragma unroll
        for (int j = 0; j < NC; ++j) sq[(c0 + j) * NP + lane] = qn[j];
      }
      __syncthreads();
#pragma unroll
      for (int l = 0; l < NV2; ++l) qa[l] = sq[l * NP + lane];
    }
    const DPatch& pt = P->p[a.ip_cost];
    if (mine && c0 == 0 && in_patch(pt, gi, gj)) {
      const int kp = (gi - pt.lo[0]) + pt.ls[0] * (gj - pt.lo[1]);
      const Prim<2> s = prim<2>(qa, g.gamma);
      const double F = s.p - pt.aplus[kp];
      const double gn = (tg.hw1(ty) * (tg.hw0(tx) * 1.0)) * (tg.h0(tx) * tg.h1(ty));   // gnorm<2>
      a.jpartial[(inst * a.nslots + a.slot) * a.npts + kp] = 1.0 * gn * (F * F * pt.strength[kp]);
    }
  }
}

// -------------------------------------------------------------------------- adjoint
// One adjoint stage on tile block `blk` of instance `inst` (uniform per work-group).
template <class S>
__device__ __forceinline__ void adj_stage_body(const Geo& g, const Patches* __restrict__ P,
                                               const Tiling& T, const AdjStage& a, int blk,
                                               long inst) {
  constexpr int TX = S::TX, EX = S::EX, EY = S::EY, NP = S::NP, NC = S::NC;
  constexpr int STX = S::STX;
  StageLds<S>& lds = stage_lds<S>();
  double(&sq)[S::SQ] = lds.q;
  double(&sc)[S::SSCR] = lds.c;
  double(&sgrid)[TileGrid<S>::kSize] = lds.grid;
  const int tile = block_tile(T, blk);
  if (tile < 0) return;   // block beyond the tiles of its XCD (uniform per block)
  const TileGeo t = tile_geo(T, tile);
  const long NVN = (long)NV2 * g.N;
  const double* li = a.lam + inst * NVN;
  const double* qi = a.q + inst * NVN;
  const Coefs cf = load_coefs(g, SLOT_ADJ);
  const double gv = load_tile_grid<S>(g, t);
  const TileGrid<S> tg{sgrid};
  const unsigned pmask = T.pmask[tile];
  const int lane = threadIdx.x % NP, c0 = (threadIdx.x / NP) * NC;
  // row-major points (the 8 x 8 wave blocks of point_of_lane do not pay here: the
  // adjoint closure is 12 deep and splits the second block column)
  const int tx = lane % TX, ty = lane / TX;
  const bool mine = tx < t.sx && ty < t.sy;
  const int gi = t.x0 + tx, gj = t.y0 + ty;
  const long p = gi + (long)g.n[0] * gj;
  // forward stage state at this thread's point, fetched before the stencil work
  double pq[NV2];
  if (mine) {
#pragma unroll
    for (int l = 0; l < NV2; ++l) pq[l] = qi[l * g.N + p];
  }
  {
    StageRegs<S> sr;
    stage_load<S>(g, t, li, sr);
    if (threadIdx.x < TileGrid<S>::kSize) sgrid[threadIdx.x] = gv;
    stage_store<S, false>(g, t, sr, tg, sq, nullptr, nullptr);
  }
  double* stx = sc;
  double* sty = sc + STX;
  __syncthreads();
  // the dissipation intermediates (into the scratch space) and D+ lambda in one phase
  const bool two_level = g.dissipation && !g.composite;
  if (two_level) diss_intermediates<S>(g, t, cf, tg, sq, stx, sty, lane, c0);
  const Op& oa = g.ops[SLOT_ADJ];
  double T1[2][NV2];   // D+ lambda along x and y, all components (own ones first)
  double w[2][NC];     // dissipation terms of this thread's components
  if (mine) {
    double u[2][NC];
    deriv_lds<16, 12, NC>(oa, cf.d1, g.dir[0], gi, t.ex0, c0,
                          [&](int l, int e) { return sq[(l * EY + ty + HY) * EX + e]; }, u[0]);
    deriv_lds<16, 12, NC>(oa, cf.d1, g.dir[1], gj, t.ey0, c0,
                          [&](int l, int e) { return sq[(l * EY + e) * EX + tx + HX]; }, u[1]);
#pragma unroll
    for (int d = 0; d < 2; ++d)
#pragma unroll
      for (int j = 0; j < NC; ++j) T1[d][j] = u[d][j];
  }
  if (two_level) __syncthreads();   // intermediates complete
  // with component groups the dissipation terms are formed before the exchange
  if (S::CS > 1 && mine && g.dissipation)
    diss_terms<S>(g, t, cf, tg, sq, stx, sty, tx, ty, gi, gj, c0, -g.diss_amount, w);
  if (S::CS > 1) {
    // every component of D+ lambda is needed by the Jacobian products: exchange them
    // through the scratch space (the dissipation intermediates are no longer read)
    __syncthreads();
    if (mine) {
#pragma unroll
      for (int d = 0; d < 2; ++d)
#pragma unroll
        for (int j = 0; j < NC; ++j) sc[(d * NV2 + c0 + j) * NP + lane] = T1[d][j];
    }
    __syncthreads();
    if (mine) {
#pragma unroll
      for (int d = 0; d < 2; ++d)
#pragma unroll
        for (int l = 0; l < NV2; ++l) T1[d][l] = sc[(d * NV2 + l) * NP + lane];
    }
  }
  if (!mine) return;
  const Prim<2> s = prim<2>(pq, g.gamma);
  double r[NC];
  const double mx[2] = {tg.h1(ty), tg.h0(tx)};   // metric<2>(g, 0, .), metric<2>(g, 1, .)
  if (c0 == 0)
    adj_jacobian_rows<0, NC>(mx, s, g.gamma, T1, r);
  else
    adj_jacobian_rows<(S::CS > 1 ? NC : 0), NC>(mx, s, g.gamma, T1, r);
  // RK operands, fetched here so that they are not live through the derivative work
  double pl0[NC], pb2[NC];
#pragma unroll
  for (int j = 0; j < NC; ++j) {
    const long o = (c0 + j) * g.N + p;
    pl0[j] = (a.stage == 3 || a.stage == 2) ? a.l0[inst * NVN + o] : 0.0;
    pb2[j] = a.stage <= 3 ? a.b2[inst * NVN + o] : 0.0;
  }
  if (g.dissipation) {
    if (S::CS == 1)
      diss_terms<S>(g, t, cf, tg, sq, stx, sty, tx, ty, gi, gj, c0, -g.diss_amount, w);
#pragma unroll
    for (int d = 0; d < 2; ++d)
#pragma unroll
      for (int j = 0; j < NC; ++j) r[j] = r[j] + w[d][j];
  }
  const double jac = 1.0 / (tg.h0(tx) * tg.h1(ty));   // jacobian<2>
#pragma unroll
  for (int j = 0; j < NC; ++j) r[j] = r[j] * jac;
  double lp[NV2];
#pragma unroll
  for (int l = 0; l < NV2; ++l) lp[l] = sq[(l * EY + ty + HY) * EX + tx + HX];
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
        for (int j = 0; j < NC; ++j) {
          const int i = c0 + j;
          double acc = 0.0;
#pragma unroll
          for (int c = 0; c < NV2; ++c) acc += A[i * 4 + c] * lp[c];
          r[j] = r[j] + sj * acc;
        }
        break;
      }
      case MGD_PATCH_SPONGE: {
        const double sp = pt.strength[kp];
#pragma unroll
        for (int j = 0; j < NC; ++j) r[j] = r[j] + sp * comp<NC>(lp, c0, j);
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
        for (int j = 0; j < NC; ++j) r[j] = r[j] + a.forcing_factor * comp<NC>(af, c0, j);
        break;
      }
      case MGD_PATCH_ACTUATOR: {
        // ThermalActuator gradient and sensitivity on the pre-update adjoint state,
        // written by the group that owns the energy component
        if (c0 + NC - 1 == 3) {
          if (a.grad) a.grad[inst * a.grad_stride + kp] = lp[3] * (pt.strength[kp] * 1.0);
          if (a.spartial) {
            const double F = lp[3] * pt.strength[kp] * 1.0;
            const double gn = (tg.hw1(ty) * (tg.hw0(tx) * 1.0)) * (tg.h0(tx) * tg.h1(ty));   // gnorm<2>
            a.spartial[(inst * a.nslots + a.slot) * a.npts + kp] = 1.0 * gn * (F * F);
          }
        }
        break;
      }
      case MGD_PATCH_KOLMOGOROV:
#pragma unroll
        for (int j = 0; j < NC; ++j)
          if (c0 + j == 0) r[j] = r[j] - lp[1] * pt.strength[kp];
        break;
    }
  }
  // RK4 adjoint stage update (substepAdjointRK4, src/RK4IntegratorImpl.f90:215-261),
  // non-temporal stores as in the forward stage
  const double dt = a.dt;
  double* b2 = a.b2 + inst * NVN;
#pragma unroll
  for (int j = 0; j < NC; ++j) {
    const long o = (c0 + j) * g.N + p;
    const double rl = r[j];
    double ln;
    if (a.stage == 4) {
      const double ll = comp<NC>(lp, c0, j);
      __builtin_nontemporal_store(ll - div6(dt * rl), &b2[o]);
      ln = ll - dt * rl / 2.0;
    } else if (a.stage == 3) {
      __builtin_nontemporal_store(pb2[j] - div3(dt * rl), &b2[o]);
      ln = pl0[j] - dt * rl / 2.0;
    } else if (a.stage == 2) {
      __builtin_nontemporal_store(pb2[j] - div3(dt * rl), &b2[o]);
      ln = pl0[j] - dt * rl;
    } else {
      ln = pb2[j] - div6(dt * rl);
    }
    __builtin_nontemporal_store(ln, &a.lout[inst * NVN + o]);
    if (a.rhs_out) a.rhs_out[inst * NVN + o] = rl;
  }
}

template <class S>
__global__ void __launch_bounds__(S::NT, S::kWavesPerEU) k_fwd_stage2d(Geo g, const Patches* __restrict__ P,
                                                       Tiling T, FwdStage a) {
  fwd_stage_body<S>(g, P, T, a, blockIdx.x, blockIdx.y);
}

template <class S>
__global__ void __launch_bounds__(S::NT, S::kWavesPerEU) k_adj_stage2d(Geo g, const Patches* __restrict__ P,
                                                       Tiling T, AdjStage a) {
  adj_stage_body<S>(g, P, T, a, blockIdx.x, blockIdx.y);
}

// One adjoint stage and one forward stage of the same region in one launch: blocks
// [0, nblk) run the adjoint stage's tiles, blocks [nblk, 2 nblk) the forward stage's (the
// window recompute beside the adjoint of the next window: independent states, so the
// same values as two launches).  nblk is a multiple of 8, so a block's XCD -- its index
// mod 8 -- is the one its tile is dealt to either way.
template <class S>
__global__ void __launch_bounds__(S::NT, S::kWavesPerEU) k_dual_stage2d(Geo g, const Patches* __restrict__ P,
                                                        Tiling T, AdjStage a, FwdStage f) {
  if ((int)blockIdx.x < T.nblk) adj_stage_body<S>(g, P, T, a, blockIdx.x, blockIdx.y);
  else fwd_stage_body<S>(g, P, T, f, (int)blockIdx.x - T.nblk, blockIdx.y);
}

// Several regions in one launch: launch block k belongs to the member m with
// mem[m].block0 <= k < mem[m + 1].block0 (members in block order), which sees it as tile
// block (k - block0) mod nblk of instance (k - block0) / nblk.  args[m]: the stage of m.
__device__ __forceinline__ int group_member(const GroupMember* mem, int nmem, int* blk,
                                            long* inst) {
  int m = 0;
  while (m + 1 < nmem && (int)blockIdx.x >= mem[m + 1].block0) ++m;
  const int lb = (int)blockIdx.x - mem[m].block0;
  *blk = lb % mem[m].T.nblk;
  *inst = lb / mem[m].T.nblk;
  return m;
}

// Copies one trivially copyable descriptor from device memory into LDS, word by word
// across the work-group (the caller synchronises).
template <class T>
__device__ __forceinline__ void copy_to_lds(const T* src, T* dst) {
  static_assert(sizeof(T) % 4 == 0, "descriptor of whole words");
  const int* s = reinterpret_cast<const int*>(src);
  int* d = reinterpret_cast<int*>(dst);
  for (int i = threadIdx.x; i < (int)(sizeof(T) / 4); i += blockDim.x) d[i] = s[i];
}

// The member's descriptors are staged in LDS: the body reads them many times, and LDS
// cannot alias the global stores of the stage, so no read is repeated after a store.
template <class S>
__global__ void __launch_bounds__(S::NT, S::kWavesPerEU)
    k_fwd_stage2d_group(const GroupMember* __restrict__ mem, int nmem, const FwdStage* __restrict__ args) {
  __shared__ GroupMember sm;
  __shared__ FwdStage sa;
  int blk;
  long inst;
  const int m = group_member(mem, nmem, &blk, &inst);
  copy_to_lds(&mem[m], &sm);
  copy_to_lds(&args[m], &sa);
  __syncthreads();
  fwd_stage_body<S>(sm.g, sm.P, sm.T, sa, blk, inst);
}

template <class S>
__global__ void __launch_bounds__(S::NT, S::kWavesPerEU)
    k_adj_stage2d_group(const GroupMember* __restrict__ mem, int nmem, const AdjStage* __restrict__ args) {
  __shared__ GroupMember sm;
  __shared__ AdjStage sa;
  int blk;
  long inst;
  const int m = group_member(mem, nmem, &blk, &inst);
  copy_to_lds(&mem[m], &sm);
  copy_to_lds(&args[m], &sa);
  __syncthreads();
  adj_stage_body<S>(sm.g, sm.P, sm.T, sa, blk, inst);
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
using S16x16c2 = Shape<16, 16, 2>;
using S16x12c2 = Shape<16, 12, 2>;

static const int kShapeDims[kNumTileShapes][2] = {{32, 16}, {16, 32}, {16, 16}, {24, 16},
                                                  {16, 16}, {16, 12}};

int fused2d_shape_dims(int shape, int* tx, int* ty) {
  if (shape < 0 || shape >= kNumTileShapes) return MGD_ERR_ARG;
  *tx = kShapeDims[shape][0];
  *ty = kShapeDims[shape][1];
  return MGD_OK;
}

// Balanced partition of len points into n = ceil(len / T) tiles of q or q + 1 points,
// the larger ones at both ends (l leading, r trailing: the edge tiles are the largest,
// as the boundary closures need): tile b holds q + (b < l || b >= n - r) points.
static void balanced(int len, int T, std::vector<int>& org, std::vector<int>& size, int* q,
                     int* l, int* r) {
  const int n = (len + T - 1) / T;
  const int extra = len % n;
  *q = len / n;
  *l = (extra + 1) / 2;
  *r = extra / 2;
  int o = 0;
  for (int b = 0; b < n; ++b) {
    const int sz = *q + (b < *l || b >= n - *r ? 1 : 0);
    org.push_back(o);
    size.push_back(sz);
    o += sz;
  }
}

int fused2d_make_tiling(const Geo& g, const Patches& hp, int shape, HostTiling* out) {
  int tx, ty;
  if (fused2d_shape_dims(shape, &tx, &ty) != MGD_OK) return MGD_ERR_ARG;
  out->shape = shape;
  out->org.clear();
  out->size.clear();
  balanced(g.n[0], tx, out->org, out->size, &out->qx, &out->lx, &out->rx);
  out->nbx = (int)out->org.size();
  balanced(g.n[1], ty, out->org, out->size, &out->qy, &out->ly, &out->ry);
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
static void fwd_group_launch(const GroupMember* mem, int nmem, int nblocks, const FwdStage* args,
                             hipStream_t st) {
  k_fwd_stage2d_group<S><<<nblocks, S::NT, 0, st>>>(mem, nmem, args);
}

template <class S>
static void adj_group_launch(const GroupMember* mem, int nmem, int nblocks, const AdjStage* args,
                             hipStream_t st) {
  k_adj_stage2d_group<S><<<nblocks, S::NT, 0, st>>>(mem, nmem, args);
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

template <class S>
static void dual_launch(const Geo& g, const Patches* P, const Tiling& T, const AdjStage& a,
                        const FwdStage& f, hipStream_t st) {
  k_dual_stage2d<S><<<dim3(2 * T.nblk, g.ninst), S::NT, 0, st>>>(g, P, T, a, f);
}

void launch_dual_stage2d(const Geo& g, const Patches* P, const Tiling& T, const AdjStage& a,
                         const FwdStage& f, hipStream_t st) {
  switch (T.shape) {
    case 0: dual_launch<S32x16>(g, P, T, a, f, st); break;
    case 1: dual_launch<S16x32>(g, P, T, a, f, st); break;
    case 2: dual_launch<S16x16>(g, P, T, a, f, st); break;
    case 3: dual_launch<S24x16>(g, P, T, a, f, st); break;
    case 4: dual_launch<S16x16c2>(g, P, T, a, f, st); break;
    default: dual_launch<S16x12c2>(g, P, T, a, f, st); break;
  }
}

void launch_fwd_stage2d(const Geo& g, const Patches* P, const Tiling& T, const FwdStage& a,
                        hipStream_t st) {
  switch (T.shape) {
    case 0: fwd_launch<S32x16>(g, P, T, a, st); break;
    case 1: fwd_launch<S16x32>(g, P, T, a, st); break;
    case 2: fwd_launch<S16x16>(g, P, T, a, st); break;
    case 3: fwd_launch<S24x16>(g, P, T, a, st); break;
    case 4: fwd_launch<S16x16c2>(g, P, T, a, st); break;
    default: fwd_launch<S16x12c2>(g, P, T, a, st); break;
  }
}

void launch_adj_stage2d(const Geo& g, const Patches* P, const Tiling& T, const AdjStage& a,
                        hipStream_t st) {
  switch (T.shape) {
    case 0: adj_launch<S32x16>(g, P, T, a, st); break;
    case 1: adj_launch<S16x32>(g, P, T, a, st); break;
    case 2: adj_launch<S16x16>(g, P, T, a, st); break;
    case 3: adj_launch<S24x16>(g, P, T, a, st); break;
    case 4: adj_launch<S16x16c2>(g, P, T, a, st); break;
    default: adj_launch<S16x12c2>(g, P, T, a, st); break;
  }
}

int launch_fwd_stage2d_group(int shape, const GroupMember* mem, int nmem, int nblocks,
                             const FwdStage* args, hipStream_t st) {
  if (shape != kGroupShape) return MGD_ERR_UNSUPPORTED;
  fwd_group_launch<S16x16>(mem, nmem, nblocks, args, st);
  return MGD_OK;
}

int launch_adj_stage2d_group(int shape, const GroupMember* mem, int nmem, int nblocks,
                             const AdjStage* args, hipStream_t st) {
  if (shape != kGroupShape) return MGD_ERR_UNSUPPORTED;
  adj_group_launch<S16x16>(mem, nmem, nblocks, args, st);
  return MGD_OK;
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
