// fused2d.h -- arguments and launchers of the fused 2-D RK-stage kernels (fused2d.hip).
#pragma once
#include <vector>

#include "mgd_internal.h"

namespace mgd {

// Tile partition of a 2-D grid (device view): tiles of at most 32 x 16 points with
// origins org[0..nbx) along x and org[nbx..nbx+nby) along y; pmask: patches per tile.
struct Tiling {
  int nbx, nby;
  const int* org;
  const int* size;
  const unsigned* pmask;
};

struct HostTiling {
  int nbx = 0, nby = 0;
  std::vector<int> org, size;
  std::vector<unsigned> pmask;
};

// One forward RK4 stage: rhs at `qin`, buffer update of `b2`, new stage state to `qout`.
// q0 is the state at the start of the time step (RK buffer1).
struct FwdStage {
  const double* qin;
  const double* q0;
  double* b2;
  double* qout;
  double* rhs_out;          // optional copy of the rhs (tests)
  int stage;
  double dt;
  const double* control;    // actuator forcing of this stage (per instance stride below)
  long control_stride;
  const Source* src;
  const double* src_amp;    // amplitude of each source at this stage's time
  int nsrc;
  double* jpartial;         // [inst][stage_slot][tile rank] functional partial sums, or null
  const int* tile_rank;
  int ntiles, nstage_slots, stage_slot, ip_cost;
};

// One adjoint RK4 stage (substepAdjointRK4): rhs at (q, lam), new adjoint state to lout.
// l0 is the adjoint state at the start of the reversed step (RK buffer1).
struct AdjStage {
  const double* lam;
  const double* l0;
  const double* q;
  double* b2;
  double* lout;
  double* rhs_out;
  int stage;
  double dt;
  double forcing_factor;
  int zero_forcing;
  double* grad;             // gradient of this stage (per instance stride below), or null
  long grad_stride;
  double* spartial;         // [inst][stage_slot][tile rank] sensitivity partials, or null
  const int* tile_rank;
  int ntiles, nstage_slots, stage_slot, ip_act;
};

// Balanced partition; MGD_ERR_UNSUPPORTED when an edge tile of a non-periodic direction
// would hold fewer points than the deepest boundary closure.
int fused2d_make_tiling(const Geo& g, const Patches& hp, HostTiling* out);
int fused2d_tile_ranks(const HostTiling& T, const int* lo, const int* hi, std::vector<int>& rank);
void launch_fwd_stage2d(const Geo& g, const Patches* P, const Tiling& T, const FwdStage& a,
                        hipStream_t st);
void launch_adj_stage2d(const Geo& g, const Patches* P, const Tiling& T, const AdjStage& a,
                        hipStream_t st);
void launch_stage_quadrature(const double* partial, int nstages, int ntiles, const double* w,
                             double* stage_sums, double* acc, int ninst, int reverse,
                             hipStream_t st);

}  // namespace mgd
