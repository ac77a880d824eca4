// mgd_internal.h -- internal types of the MI355X magudi RHS library.
//
// Layout in HBM (all FP64):
//   state-like arrays  x[(inst * nv + l) * N + p],   p = i + nx * (j + ny * k)
//   grid               rectilinear: per direction d, 1-D arrays over i_d of
//                      h_d = dx_d/dxi_d, the SBP norm weight Hw_d and the coordinate;
//                      metrics, Jacobian, norm and arc lengths are products of these
//                      and are formed in registers, never stored per point.
//   patches            small per-patch-point arrays (sponge strength, A+ blocks, ...).
#pragma once
#include <hip/hip_runtime.h>
#include <cstdint>
#include <string>
#include <vector>

#include "../../include/magudi_hip.h"

namespace mgd {

constexpr int kMaxBW = 16;   // widest boundary block (adjoint 4-8 first derivative: 16 x 12)
constexpr int kMaxBD = 12;
constexpr int kMaxHW = 4;    // interior half width (SBP 4-8)
constexpr int kMaxPatches = 32;

// One SBP stencil operator (t_StencilOperator).  b1[m][c] = rhsBoundary1(c+1, m+1).
struct Op {
  int sym;   // 0 symmetric, 1 skew-symmetric
  int hw;    // interiorWidth / 2
  int bw;    // boundaryWidth
  int bd;    // boundaryDepth
  double c[2 * kMaxHW + 1];   // rhsInterior(k) at c[k + kMaxHW]
  double b1[kMaxBD][kMaxBW];
  double b2[kMaxBD][kMaxBW];
  double norm[kMaxBD];
  int nnorm;
};

enum OpKind { OP_FIRST = 0, OP_SECOND = 1, OP_DISS = 2, OP_DISS_T = 3, OP_COMPOSITE = 4 };
enum OpSlot { SLOT_FIRST = 0, SLOT_ADJ = 1, SLOT_DISS = 2, SLOT_DISS_T = 3, SLOT_SECOND = 4,
              kNumSlots = 5 };

int op_setup(Op* op, int scheme, int kind);     // sbp.cpp
void op_adjoint(const Op& op, Op* adj);         // sbp.cpp
// Host 1-D application along a contiguous line (grid set-up only).
void op_apply_line(const Op& op, int len, int periodicity, const double* in, double* out);

// Geometry of one direction as seen by the kernels.
struct Dir {
  int len;        // points along the direction
  int periodic;   // 0 or 1
  int off;        // 1 for OVERLAP periodicity, else 0
  long stride;    // flat-index stride
};

// Everything a kernel needs, passed by value (well under the 4 KB argument limit).
struct Geo {
  int nd, nv, ninst;
  int n[3];
  long N;
  Dir dir[3];
  const double* h[3];       // dx_d/dxi_d along direction d
  const double* hw[3];      // SBP norm weight along direction d (1 in the interior)
  const double* x[3];       // coordinate along direction d
  const Op* ops;            // kNumSlots operators (direction independent)
  double gamma, re_inv, pr_inv, power_law, bulk_ratio, diss_amount;
  int viscous, dissipation, composite;
  int zchunk;               // z planes per work unit of the forward 3-D march and the
                            // stress pass (32, 64 or 128)
  int mtile;                // x extent of the forward march tiles (16 or 32; 16 rows)
  int zchunk_adj, mtile_adj;  // the same for the adjoint marches
};

// Device patch descriptor.
struct DPatch {
  int type;
  int normal_dir;
  int lo[3], hi[3], ls[3];
  int np;
  double sigma_inv, sigma_visc;
  double twall;             // isothermal wall temperature
  const double* strength;   // sponge strength, Kolmogorov force, or the actuator / cost-target mollifier
  const double* aplus;      // far field: A+ blocks (np x nv x nv, column major) then the target state (np x nv);
                            // sponge: target state (np x nv); cost target: mean pressure (np)
  const double* tvflux;     // far field: target viscous fluxes (nd x nv x np)
  double* ffv;              // far field, viscous: the viscous penalty term of each point
                            // (instance x nv x np), formed before the patch pass
};

struct Patches {
  int count;
  DPatch p[kMaxPatches];
};

struct Source {
  double loc[3], amplitude, omega, gfac, phase;
};

}  // namespace mgd

// ---------------------------------------------------------------- error state
namespace mgd {
void set_error(const std::string& msg);
}
#define MGD_HIP(call)                                                                   \
  do {                                                                                  \
    hipError_t e_ = (call);                                                             \
    if (e_ != hipSuccess) {                                                             \
      mgd::set_error(std::string(#call) + ": " + hipGetErrorString(e_));                 \
      return MGD_ERR_HIP;                                                               \
    }                                                                                   \
  } while (0)
