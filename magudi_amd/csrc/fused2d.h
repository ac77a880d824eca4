// fused2d.h -- arguments and launchers of the fused 2-D RK-stage kernels (fused2d.hip).
#pragma once
#include <vector>

#include "mgd_internal.h"

namespace mgd {

// Tile shapes of the fused kernels (points along x, y): 0: 32 x 16, 1: 16 x 32,
// 2: 16 x 16, 3: 24 x 16, one thread per point; 4: 16 x 16 and 5: 16 x 12 with two
// threads per point (each forming two of the four components).
constexpr int kNumTileShapes = 6;
constexpr int kNumXcd = 8;   // XCDs of an MI355X; work-groups are dealt to them round-robin

// Tile partition of a 2-D grid (device view): nbx x nby tiles of at most TX x TY points
// from a balanced partition of each direction whose larger tiles sit at both ends --
// along x, tile b holds qx points, plus one when b < lx or b >= nbx - rx, and starts at
// b qx + min(b, lx) + max(0, b - (nbx - rx)) (likewise along y) -- and pmask: patches
// per tile (tile index bx + nbx * by).  Launch block k works on tile
// (k % kNumXcd) band + k / kNumXcd, or on none when that is beyond the last tile.
struct Tiling {
  int shape;
  int nbx, nby, nblk, band;
  int qx, lx, rx, qy, ly, ry;
  const unsigned* pmask;
};

struct HostTiling {
  int shape = 0;
  int nbx = 0, nby = 0, nblk = 0, band = 0;
  int qx = 0, lx = 0, rx = 0, qy = 0, ly = 0, ry = 0;
  std::vector<int> org, size;          // tile origins / sizes: x tiles, then y tiles
  std::vector<unsigned> pmask;
  std::vector<int> tile_of_block;      // launch block -> tile, or -1
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
  long src_inst_stride;     // entries between instances of src_amp (0: shared)
  int nsrc;
  double* jpartial;         // [inst][slot][cost-patch point] functional terms, or null
  int npts, nslots, slot, ip_cost;
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
  double* spartial;         // [inst][slot][actuator-patch point] sensitivity terms, or null
  int npts, nslots, slot, ip_act;
};

// A member region of a multi-region launch (device table): its geometry, patches and
// tiling, and its first launch block (members take consecutive ranges of nblk x
// instances blocks).
struct GroupMember {
  Geo g;
  const Patches* P;
  Tiling T;
  int block0;
};
constexpr int kGroupShape = 2;     // tile shape of multi-region launches (16 x 16)
constexpr int kMaxGroup = 64;      // members of one multi-region launch

// Balanced partition; MGD_ERR_UNSUPPORTED when an edge tile of a non-periodic direction
// would hold fewer points than the deepest boundary closure.
int fused2d_make_tiling(const Geo& g, const Patches& hp, int shape, HostTiling* out);
int fused2d_shape_dims(int shape, int* tx, int* ty);
void launch_fwd_stage2d(const Geo& g, const Patches* P, const Tiling& T, const FwdStage& a,
                        hipStream_t st);
void launch_adj_stage2d(const Geo& g, const Patches* P, const Tiling& T, const AdjStage& a,
                        hipStream_t st);
// One adjoint stage and one forward stage of independent states in one launch (the same
// values as the two launches).
void launch_dual_stage2d(const Geo& g, const Patches* P, const Tiling& T, const AdjStage& a,
                         const FwdStage& f, hipStream_t st);
// One stage of every member of a group in one launch (args: one stage per member, in
// device memory); MGD_ERR_UNSUPPORTED unless shape is kGroupShape.
int launch_fwd_stage2d_group(int shape, const GroupMember* mem, int nmem, int nblocks,
                             const FwdStage* args, hipStream_t st);
int launch_adj_stage2d_group(int shape, const GroupMember* mem, int nmem, int nblocks,
                             const AdjStage* args, hipStream_t st);
// Stage sums of per-point quadrature terms: sums[inst][base + s] = sum over the points of
// partial[inst][s][point] for s < nfill (a fixed-order tree, independent of the tiling).
void launch_stage_reduce(const double* partial, int nslots, int nfill, int npts, double* sums,
                         int nstages, int base, int ninst, hipStream_t st);
// acc[inst] += sum_s w[s] * sums[inst][s], in stage order (reverse: descending).
void launch_stage_accumulate(const double* sums, int nstages, const double* w, double* acc,
                             int ninst, int reverse, hipStream_t st);

}  // namespace mgd
