// region.cpp -- host side of the MI355X magudi RHS library: the C ABI of
// include/magudi_hip.h, grid set-up, patch set-up, HBM-resident state and the
// forward / adjoint drivers (Solver%runForward / runAdjoint with uniform
// checkpointing held in HBM).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <limits>
#include <memory>
#include <mutex>
#include <string>
#include <vector>

#include "fused2d.h"
#include "kernels.h"
#include "mgd_internal.h"

namespace mgd {

static thread_local std::string g_err;
void set_error(const std::string& msg) { g_err = msg; }

namespace {

const double kRkNorm[4] = {1.0 / 6.0, 1.0 / 3.0, 1.0 / 3.0, 1.0 / 6.0};

struct DevBuf {
  double* p = nullptr;
  size_t n = 0;
  ~DevBuf() { if (p) (void)hipFree(p); }
  int alloc(size_t count) {
    if (n >= count && p) return MGD_OK;
    if (p) (void)hipFree(p);
    p = nullptr;
    n = 0;
    if (count == 0) return MGD_OK;
    hipError_t e = hipMalloc(&p, count * sizeof(double));
    if (e != hipSuccess) {
      set_error(std::string("hipMalloc of ") + std::to_string(count * sizeof(double)) +
                " bytes: " + hipGetErrorString(e));
      p = nullptr;
      return MGD_ERR_HIP;
    }
    n = count;
    return MGD_OK;
  }
  void release() {
    if (p) (void)hipFree(p);
    p = nullptr;
    n = 0;
  }
  size_t bytes() const { return n * sizeof(double); }
};

struct HostPatch {
  mgd_patch_desc pd;
  int lo[3], hi[3], ls[3], np;
  std::vector<double> strength, data, tvflux;
  double sig_inv = 0.0, sig_visc = 0.0, twall = 0.0;   // penalty amounts set at finalize
  DevBuf d_strength, d_data, d_tvflux, d_ffv;
};

}  // namespace
}  // namespace mgd

struct GraphSlot {
  hipGraphExec_t exec = nullptr;
  std::vector<uintptr_t> key;
};

struct mgd_region {
  mgd_region_desc desc;
  int nd, nv, ninst;
  long N;
  std::vector<double> coords;            // nd x N (host copy)
  std::vector<double> line[3];           // coordinate along each direction
  std::vector<double> h[3], hw[3];
  mgd::Op ops[mgd::kNumSlots];
  mgd::Op first_host;
  mgd::DevBuf d_h[3], d_hw[3], d_x[3], d_ops;
  mgd::Geo geo;
  std::vector<std::unique_ptr<mgd::HostPatch>> patches;
  std::vector<mgd::Source> sources;
  std::vector<double> target, ctrl_moll, tgt_moll;
  std::vector<double> mean_p;            // mean pressure field (mean_pressure_file), or empty
  bool finalized = false;
  mgd::Patches hpatches;
  mgd::DevBuf d_patches, d_sources;
  int ip_cost = -1, ip_act = -1, has_farfield = 0;
  mgd::FinishInfo fin;                   // the forward finishing pass (set at finalize)
  // HBM-resident state
  mgd::DevBuf q, lam, rhs, b1, b2, b1a, b2a, t, w, G, T1, S;
  mgd::DevBuf qs, aforce, control, grad, acc_J, acc_S, partial, src_amp, init;
  mgd::DevBuf snaps, window;
  // fused 2-D path: ping-pong stage buffers, per-point quadrature terms of one chunk of
  // stages, stage sums, quadrature weights
  mgd::DevBuf qy, ly, l0, jpart, spart, ssum, wts;
  int chunk = 0;                         // stages per chunk of quadrature terms
  int wts_nst = -1;                      // stages and dt of the quadrature weights held
  double wts_dt = 0.0;
  mgd::HostTiling htile;
  mgd::Tiling tiling{};
  mgd::DevBuf d_tile_mask;
  // the 16 x 16 tiling of group launches, held beside the region's own (built at the first
  // group run; the two are swapped for each group run)
  mgd::HostTiling g_htile;
  mgd::Tiling g_tiling{};
  mgd::DevBuf g_tile_mask;
  bool g_built = false, g_swapped = false;
  bool tiling_ok = false;
  int tile_shape = -1;                   // MGD_OPT_TILE_SHAPE: -1 autotune, else fixed
  bool tuned = false;
  // the fused adjoint run shares each recompute stage's launch with an adjoint stage
  // (k_dual_stage2d) or launches them one after the other (the autotuner's choice; with a
  // fixed tile shape: shared for one instance, where single launches leave the GPU half idle)
  bool dual = true;
  int opt_zchunk = 0;                    // MGD_OPT_Z_CHUNK: 0 autotune, else 32, 64 or 128
  int opt_mtile = 0;                     // MGD_OPT_MARCH_TILE: 0 autotune, else 16 or 32
  bool zc_tuned = false;
  int force_plain = 0;
  int aforce_zero = 0;
  long ctrl_stride = 0;                  // control entries per instance
  hipStream_t stream = nullptr;
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  // window recompute of the fused adjoint run: the second window buffer, its RK buffer
  // and the scratch state
  mgd::DevBuf window2, b2r, qsr;
  // MGD_OPT_RESIDENT_STATES: the forward pass keeps every RK stage state (4 nsteps + 1
  // states) in HBM and the adjoint pass reads them instead of recomputing windows
  bool resident = false;                 // this run's choice
  int opt_resident = -1;                 // -1 where every stage state fits, 0 off, 1 on
  bool group_run = false;                // preparing as a member of a group launch
  bool group_resident = false;           // ... whose members keep every stage state
  mgd::DevBuf wall;
  // the last checkpointed forward pass left the stage states of its last window in
  // `window` (the adjoint pass then starts without recomputing that window)
  bool last_window = false;
  // checkpointed forward run the adjoint pass marches back from
  struct {
    bool valid = false;
    double t0 = 0, dt = 0;
    int nsteps = 0, K = 1;
    std::vector<double> control;         // actuator forcing of every stage (host copy)
  } ck;
  mgd::DevBuf qT, scr;                   // final forward state; scratch state (MGD_SCRATCH)
  mgd::DevBuf q_alt, qs_alt;             // second state buffers of out-of-place RK stages
  // third state buffers of the 3-D z-march stages: stages 2 and 3 write there, so the
  // state at the start of the step stays where stage 1 read it and no copy of it is made
  mgd::DevBuf q_alt2, qs_alt2;
  mgd::DevBuf lam_alt;                   // second adjoint buffer of out-of-place adjoint stages
  int mig_loaded = -1;                   // snapshot step whose window mgd_checkpoint_migrate holds
  // window cache of the generic-kernel path: the checkpointed forward pass also keeps the
  // stage states of its last `ncache` windows (the first starting at step cache_first) in
  // HBM that nothing else holds, and the adjoint passes read them instead of recomputing
  // those windows (the recompute gives the same values)
  static constexpr int kMaxCache = 8;
  mgd::DevBuf wcache[kMaxCache];
  int ncache = 0, cache_first = 0;
  int opt_cache = -1;                    // MGD_OPT_WINDOW_CACHE: -1 as memory allows, else at most n
  // step store of the generic-kernel path: the checkpointed forward pass keeps the state at
  // the start of every step (sstore, steps 0 .. nsteps-1) and the three inner stage states
  // of its last `nscache` steps (scache, from step scache_first on); the adjoint pass
  // recomputes the inner stage states of the other steps from the stored state, three
  // stages per step into `window`, instead of whole windows from snapshots
  int opt_step = -1;                     // MGD_OPT_STEP_STORE: -1 where it fits, 0 off, 1 on
  bool step_mode = false;
  mgd::DevBuf sstore, scache;
  int nscache = 0, scache_first = 0;
  int rec_step = -1;                     // step whose inner stage states `window` holds
  // step mode on the 3-D march paths: the stress / heat-flux slots of the recomputed
  // stages 1-3 (three gradient buffers), which the adjoint stages about the same states
  // read instead of running the stress pass again
  mgd::DevBuf sbuf;
  bool sreuse = false;
  // per-launch timer of the z-march forward kernel (mgd_kernel_timer): event pairs
  bool timing = false;
  std::vector<hipEvent_t> tev;
  long tcount = 0, tcopy = 0;
  // solution limits (enable_solution_limits): rho_min, rho_max, T_min, T_max; checked on the
  // input of every forward stage of a run while `checking` (not in recomputes)
  int lim_on = 0;
  double lim[4] = {0, 0, 0, 0};
  bool checking = false;
  mgd::DevBuf d_flag;                    // per-instance violation flags
  mgd::DevBuf ipart, iout;               // inner-product partial sums and results
  GraphSlot graphs[2];                   // captured runs: forward, forward + adjoint
  std::vector<double> toff;              // per-instance time offsets (empty: none)
  long src_inst_stride = 0;              // source table entries per instance (0: shared)
  double last_ms[3] = {0, 0, 0};
  int last_launch[3] = {0, 0, 0};
  int device = 0;                        // HIP device the region was created on
};

using namespace mgd;

// Regions of different sizes run in lock step, one launch per RK stage for all of them
// (mgd_group_*): the stage arguments of every member and stage are built on the host
// once per run and read by the kernels from device tables.
struct mgd_group {
  std::vector<mgd_region*> m;
  hipStream_t st = nullptr, st_rec = nullptr;
  hipEvent_t ev_fwd = nullptr, ev_ready[2] = {nullptr, nullptr}, ev_free[2] = {nullptr, nullptr};
  mgd::DevBuf dmem, dfwd, drec, dadj;   // member table; forward, recompute, adjoint stage tables
  int nblocks = 0;
};

namespace {

long M_of(const mgd_region* r) { return (long)r->nv * r->N; }

// Grid%computeCoordinateDerivatives along one rectilinear line (src/GridImpl.f90:621-744)
void coordinate_derivative(const Op& op, int len, int periodicity, double plen,
                           const std::vector<double>& x, std::vector<double>& out) {
  out.assign(len, 0.0);
  if (periodicity != MGD_PERIODIC_PLANE) {
    op_apply_line(op, len, periodicity, x.data(), out.data());
    return;
  }
  // PLANE: ghost coordinates shifted by the periodic length, interior stencil everywhere
  const int hw = op.hw;
  std::vector<double> gx(len + 2 * hw);
  for (int i = -hw; i < len + hw; ++i) {
    const int w = i < 0 ? i + len : (i >= len ? i - len : i);
    double v = x[w];
    if (i < 0) v = v - plen;
    else if (i >= len) v = v + plen;
    gx[i + hw] = v;
  }
  for (int i = 0; i < len; ++i) {
    double s = 0.0;
    for (int k = 1; k <= hw; ++k) s += op.c[kMaxHW + k] * (gx[i + k + hw] - gx[i - k + hw]);
    out[i] = s;
  }
}

long stride_of(const mgd_region* r, int d) {
  return d == 0 ? 1 : (d == 1 ? r->desc.n[0] : (long)r->desc.n[0] * r->desc.n[1]);
}

// Incoming part of the inviscid flux Jacobian from the target state
// (computeIncomingJacobianOfInviscidFlux2D/3D, src/CNSHelperImpl.f90:1669-2342).
void incoming_jacobian(int nd, const double* q, const double* metrics, double g, int dirn, double* A) {
  const int n = nd + 2;
  double R[25] = {0}, L[25] = {0}, lam[5] = {0}, nm[3] = {0}, u[3] = {0};
  auto Rm = [&](int i, int j) -> double& { return R[(j - 1) * n + (i - 1)]; };
  auto Lm = [&](int i, int j) -> double& { return L[(j - 1) * n + (i - 1)]; };
  const double sv = 1.0 / q[0];
  double ke = 0.0;
  for (int i = 0; i < nd; ++i) u[i] = sv * q[i + 1];
  for (int i = 0; i < nd; ++i) ke += u[i] * u[i];
  const double T = g * (sv * q[nd + 1] - 0.5 * ke);
  double arc = 0.0;
  for (int i = 0; i < nd; ++i) arc += metrics[i] * metrics[i];
  arc = std::sqrt(arc);
  for (int i = 0; i < nd; ++i) nm[i] = metrics[i] / arc;
  double U = 0.0;
  for (int i = 0; i < nd; ++i) U += nm[i] * u[i];
  const double c = std::sqrt((g - 1.0) * T);
  const double phi2 = 0.5 * (g - 1.0) * ke;
  const double rho = q[0];
  const double h = phi2 / (g - 1.0);
  if (nd == 2) {
    lam[0] = U; lam[1] = U; lam[2] = U + c; lam[3] = U - c;
    Rm(1, 1) = 1.0; Rm(2, 1) = u[0]; Rm(3, 1) = u[1]; Rm(4, 1) = h;
    Rm(1, 2) = 0.0; Rm(2, 2) = nm[1] * rho; Rm(3, 2) = -nm[0] * rho;
    Rm(4, 2) = rho * (nm[1] * u[0] - nm[0] * u[1]);
    Rm(1, 3) = 1.0; Rm(2, 3) = u[0] + nm[0] * c; Rm(3, 3) = u[1] + nm[1] * c; Rm(4, 3) = T + h + c * U;
    Rm(1, 4) = 1.0; Rm(2, 4) = u[0] - nm[0] * c; Rm(3, 4) = u[1] - nm[1] * c; Rm(4, 4) = T + h - c * U;
    Lm(1, 1) = 1.0 - phi2 / (c * c);
    Lm(2, 1) = -sv * (nm[1] * u[0] - nm[0] * u[1]);
    Lm(3, 1) = 0.5 * (phi2 / (c * c) - U / c);
    Lm(4, 1) = 0.5 * (phi2 / (c * c) + U / c);
    Lm(1, 2) = u[0] / T; Lm(2, 2) = sv * nm[1];
    Lm(3, 2) = -0.5 * (u[0] / T - nm[0] / c); Lm(4, 2) = -0.5 * (u[0] / T + nm[0] / c);
    Lm(1, 3) = u[1] / T; Lm(2, 3) = -sv * nm[0];
    Lm(3, 3) = -0.5 * (u[1] / T - nm[1] / c); Lm(4, 3) = -0.5 * (u[1] / T + nm[1] / c);
    Lm(1, 4) = -1.0 / T; Lm(2, 4) = 0.0; Lm(3, 4) = 0.5 / T; Lm(4, 4) = 0.5 / T;
  } else if (nd == 3) {
    lam[0] = U; lam[1] = U; lam[2] = U; lam[3] = U + c; lam[4] = U - c;
    Rm(1, 1) = nm[0]; Rm(2, 1) = nm[0] * u[0]; Rm(3, 1) = nm[0] * u[1] + rho * nm[2];
    Rm(4, 1) = nm[0] * u[2] - rho * nm[1];
    Rm(5, 1) = rho * (nm[2] * u[1] - nm[1] * u[2]) + h * nm[0];
    Rm(1, 2) = nm[1]; Rm(2, 2) = nm[1] * u[0] - rho * nm[2]; Rm(3, 2) = nm[1] * u[1];
    Rm(4, 2) = nm[1] * u[2] + rho * nm[0];
    Rm(5, 2) = rho * (nm[0] * u[2] - nm[2] * u[0]) + h * nm[1];
    Rm(1, 3) = nm[2]; Rm(2, 3) = nm[2] * u[0] + rho * nm[1]; Rm(3, 3) = nm[2] * u[1] - rho * nm[0];
    Rm(4, 3) = nm[2] * u[2];
    Rm(5, 3) = rho * (nm[1] * u[0] - nm[0] * u[1]) + h * nm[2];
    Rm(1, 4) = 1.0; Rm(2, 4) = u[0] + nm[0] * c; Rm(3, 4) = u[1] + nm[1] * c;
    Rm(4, 4) = u[2] + nm[2] * c; Rm(5, 4) = T + h + c * U;
    Rm(1, 5) = 1.0; Rm(2, 5) = u[0] - nm[0] * c; Rm(3, 5) = u[1] - nm[1] * c;
    Rm(4, 5) = u[2] - nm[2] * c; Rm(5, 5) = T + h - c * U;
    const double a = 1.0 - phi2 / (c * c);
    Lm(1, 1) = nm[0] * a - sv * (nm[2] * u[1] - nm[1] * u[2]);
    Lm(2, 1) = nm[1] * a - sv * (nm[0] * u[2] - nm[2] * u[0]);
    Lm(3, 1) = nm[2] * a - sv * (nm[1] * u[0] - nm[0] * u[1]);
    Lm(4, 1) = 0.5 * (phi2 / (c * c) - U / c);
    Lm(5, 1) = 0.5 * (phi2 / (c * c) + U / c);
    Lm(1, 2) = nm[0] * u[0] / T; Lm(2, 2) = nm[1] * u[0] / T - sv * nm[2];
    Lm(3, 2) = nm[2] * u[0] / T + sv * nm[1];
    Lm(4, 2) = -0.5 * (u[0] / T - nm[0] / c); Lm(5, 2) = -0.5 * (u[0] / T + nm[0] / c);
    Lm(1, 3) = nm[0] * u[1] / T + sv * nm[2]; Lm(2, 3) = nm[1] * u[1] / T;
    Lm(3, 3) = nm[2] * u[1] / T - sv * nm[0];
    Lm(4, 3) = -0.5 * (u[1] / T - nm[1] / c); Lm(5, 3) = -0.5 * (u[1] / T + nm[1] / c);
    Lm(1, 4) = nm[0] * u[2] / T - sv * nm[1]; Lm(2, 4) = nm[1] * u[2] / T + sv * nm[0];
    Lm(3, 4) = nm[2] * u[2] / T;
    Lm(4, 4) = -0.5 * (u[2] / T - nm[2] / c); Lm(5, 4) = -0.5 * (u[2] / T + nm[2] / c);
    Lm(1, 5) = -nm[0] / T; Lm(2, 5) = -nm[1] / T; Lm(3, 5) = -nm[2] / T;
    Lm(4, 5) = 0.5 / T; Lm(5, 5) = 0.5 / T;
  } else {
    for (int k = 0; k < n * n; ++k) A[k] = 0.0;
    return;
  }
  for (int i = 0; i < n; ++i) lam[i] = arc * lam[i];
  for (int i = 0; i < n; ++i) if ((double)dirn * lam[i] < 0.0) lam[i] = 0.0;
  for (int j = 1; j <= n; ++j)
    for (int i = 1; i <= n; ++i) {
      double s = Rm(i, 1) * lam[0] * Lm(1, j);
      for (int k = 2; k <= n; ++k) s = s + Rm(i, k) * lam[k - 1] * Lm(k, j);
      A[(j - 1) * n + (i - 1)] = s;
    }
}

int upload(DevBuf& b, const std::vector<double>& v) {
  if (b.alloc(std::max<size_t>(v.size(), 1))) return MGD_ERR_HIP;
  if (!v.empty()) MGD_HIP(hipMemcpy(b.p, v.data(), v.size() * sizeof(double), hipMemcpyHostToDevice));
  return MGD_OK;
}

double metric_host(const mgd_region* r, int d, const int* ijk) {
  const int nd = r->nd;
  if (nd == 1) return 1.0;
  if (nd == 2) return d == 0 ? r->h[1][ijk[1]] : r->h[0][ijk[0]];
  if (d == 0) return r->h[1][ijk[1]] * r->h[2][ijk[2]];
  if (d == 1) return r->h[0][ijk[0]] * r->h[2][ijk[2]];
  return r->h[0][ijk[0]] * r->h[1][ijk[1]];
}

void unflat_host(const mgd_region* r, long p, int* ijk) {
  ijk[0] = (int)(p % r->desc.n[0]);
  ijk[1] = (int)((p / r->desc.n[0]) % r->desc.n[1]);
  ijk[2] = (int)(p / ((long)r->desc.n[0] * r->desc.n[1]));
}

long patch_point(const mgd_region* r, const HostPatch& P, int k) {
  const int a = k % P.ls[0], b = (k / P.ls[0]) % P.ls[1], c = k / (P.ls[0] * P.ls[1]);
  return (P.lo[0] + a) + (long)r->desc.n[0] * ((P.lo[1] + b) + (long)r->desc.n[1] * (P.lo[2] + c));
}

// viscous fluxes of the target state at the far-field patch points (host; set-up only)
void target_viscous_fluxes(mgd_region* r, HostPatch& P) {
  const int nd = r->nd, nv = r->nv;
  const long N = r->N;
  const double g = r->desc.gamma;
  std::vector<double> w((size_t)(nd + 1) * N);
  for (long p = 0; p < N; ++p) {
    const double sv = 1.0 / r->target[p];
    double ke = 0.0;
    for (int i = 0; i < nd; ++i) {
      const double u = sv * r->target[(i + 1) * N + p];
      w[i * N + p] = u;
      ke += u * u;
    }
    const double pr = (g - 1.0) * (r->target[(nd + 1) * N + p] - 0.5 * r->target[p] * ke);
    w[nd * N + p] = g * pr / (g - 1.0) * sv;
  }
  const double re_inv = r->geo.re_inv, pr_inv = r->geo.pr_inv;
  P.tvflux.assign((size_t)nd * nv * P.np, 0.0);
  for (int k = 0; k < P.np; ++k) {
    const long p = patch_point(r, P, k);
    int ijk[3];
    unflat_host(r, p, ijk);
    double grad[4][3];
    for (int d = 0; d < nd; ++d) {
      const int len = r->desc.n[d];
      const long st = stride_of(r, d);
      const long base = p - (long)ijk[d] * st;
      std::vector<double> in(len), out(len);
      const double jm = (1.0 / (nd == 1 ? r->h[0][ijk[0]]
                                : nd == 2 ? r->h[0][ijk[0]] * r->h[1][ijk[1]]
                                          : r->h[0][ijk[0]] * r->h[1][ijk[1]] * r->h[2][ijk[2]])) *
                        metric_host(r, d, ijk);
      for (int c = 0; c <= nd; ++c) {
        for (int i = 0; i < len; ++i) in[i] = w[c * N + base + i * st];
        op_apply_line(r->first_host, len, r->desc.periodicity[d], in.data(), out.data());
        grad[c][d] = jm * out[ijk[d]];
      }
    }
    const double T = w[nd * N + p];
    double mu, la, ka;
    if (r->desc.power_law <= 0.0) {
      mu = re_inv; la = (r->desc.bulk_ratio - 2.0 / 3.0) * re_inv; ka = re_inv * pr_inv;
    } else {
      mu = std::pow((g - 1.0) * T, r->desc.power_law) * re_inv;
      la = (r->desc.bulk_ratio - 2.0 / 3.0) * mu;
      ka = mu * pr_inv;
    }
    double tau[9];
    for (int i = 0; i < nd; ++i)
      for (int j = 0; j < nd; ++j) tau[i * nd + j] = grad[i][j];
    if (nd == 2) {
      const double dv = la * (tau[0] + tau[3]);
      tau[0] = 2.0 * mu * tau[0] + dv; tau[1] = mu * (tau[1] + tau[2]); tau[2] = tau[1];
      tau[3] = 2.0 * mu * tau[3] + dv;
    } else if (nd == 3) {
      const double dv = la * (tau[0] + tau[4] + tau[8]);
      tau[0] = 2.0 * mu * tau[0] + dv; tau[1] = mu * (tau[1] + tau[3]);
      tau[2] = mu * (tau[2] + tau[6]); tau[3] = tau[1]; tau[4] = 2.0 * mu * tau[4] + dv;
      tau[5] = mu * (tau[5] + tau[7]); tau[6] = tau[2]; tau[7] = tau[5];
      tau[8] = 2.0 * mu * tau[8] + dv;
    } else {
      tau[0] = (2.0 * mu + la) * tau[0];
    }
    double hf[3];
    for (int j = 0; j < nd; ++j) hf[j] = -ka * grad[nd][j];
    for (int d = 0; d < nd; ++d) {
      P.tvflux[((size_t)d * nv + 0) * P.np + k] = 0.0;
      double e = 0.0;
      for (int i = 0; i < nd; ++i) {
        P.tvflux[((size_t)d * nv + 1 + i) * P.np + k] = tau[i * nd + d];
        e += w[i * N + p] * tau[i * nd + d];
      }
      P.tvflux[((size_t)d * nv + nd + 1) * P.np + k] = e - hf[d];
    }
  }
}

// Device array of a state id (MGD_CONSERVED ... MGD_SCRATCH); `create` allocates the
// arrays that exist only on demand (initial, final, scratch).
DevBuf* state_buf(mgd_region* r, int which, bool create) {
  DevBuf* b = which == MGD_CONSERVED ? &r->q : which == MGD_ADJOINT_VARS ? &r->lam
            : which == MGD_RHS ? &r->rhs : which == MGD_INITIAL_STATE ? &r->init
            : which == MGD_FINAL_STATE ? &r->qT : which == MGD_SCRATCH ? &r->scr : nullptr;
  if (!b) return nullptr;
  if (!b->p && create && b->alloc((size_t)r->ninst * r->nv * r->N)) return nullptr;
  return b->p ? b : nullptr;
}

int ensure_state(mgd_region* r) {
  const size_t M = (size_t)r->ninst * r->nv * r->N;
  if (r->q.alloc(M) || r->lam.alloc(M) || r->rhs.alloc(M) || r->b1.alloc(M) || r->b2.alloc(M) ||
      r->T1.alloc(M * r->nd))
    return MGD_ERR_HIP;
  if (r->desc.dissipation && !r->desc.composite_dissipation && r->t.alloc(M)) return MGD_ERR_HIP;
  // 3-D regions: the z-march rhs kernel writes RK stage states out of place
  if (r->nd == 3 && (r->q_alt.alloc(M) || r->q_alt2.alloc(M))) return MGD_ERR_HIP;
  if (r->desc.viscous) {
    if (r->w.alloc((size_t)r->ninst * (r->nd + 1) * r->N) ||
        r->G.alloc((size_t)r->ninst * (r->nd + 1) * r->nd * r->N))
      return MGD_ERR_HIP;
  }
  return MGD_OK;
}

int ensure_adjoint(mgd_region* r) {
  const size_t M = (size_t)r->ninst * r->nv * r->N;
  if (r->b1a.alloc(M) || r->b2a.alloc(M) || r->T1.alloc(M * r->nd)) return MGD_ERR_HIP;
  if (r->desc.viscous && r->S.alloc(M * r->nd)) return MGD_ERR_HIP;
  return MGD_OK;
}

int n_act(const mgd_region* r) { return r->ip_act < 0 ? 0 : r->patches[r->ip_act]->np; }
int n_cost(const mgd_region* r) { return r->ip_cost < 0 ? 0 : r->patches[r->ip_cost]->np; }
const DPatch* dpatch(const mgd_region* r, int ip) {
  return reinterpret_cast<const Patches*>(r->d_patches.p)->p + ip;
}
const Patches* dpatches(const mgd_region* r) { return reinterpret_cast<const Patches*>(r->d_patches.p); }

// stage times of substepForwardRK4: stage 1 at t, 2 and 3 at t + dt/2, 4 at t + dt
std::vector<double> stage_times(double t0, double dt, int nsteps);

// Source amplitudes a = amplitude cos(omega t + phase), laid out [instance][time][source],
// instance i at the times `times` shifted by its time offset.  With no offsets the
// instances share one table (inst_stride 0).
void source_table(mgd_region* r, const std::vector<double>& times, std::vector<double>& out) {
  const int ns = (int)r->sources.size();
  const size_t nt = times.size();
  const bool shifted = !r->toff.empty();
  const int ni = shifted ? r->ninst : 1;
  out.assign((size_t)ni * nt * std::max(ns, 1), 0.0);
  for (int i = 0; i < ni; ++i)
    for (size_t k = 0; k < nt; ++k)
      for (int s = 0; s < ns; ++s) {
        const double t = shifted ? times[k] + r->toff[(size_t)i] : times[k];
        out[((size_t)i * nt + k) * ns + s] =
            r->sources[s].amplitude * std::cos(r->sources[s].omega * t + r->sources[s].phase);
      }
  r->src_inst_stride = shifted ? (long)(nt * ns) : 0;
}

// Stage times of a run: per instance when the instances carry time offsets (the offsets
// are applied to the start time, then the stages are accumulated as in substepForwardRK4).
void run_source_table(mgd_region* r, double t0, double dt, int nsteps, std::vector<double>& out) {
  if (r->toff.empty()) { source_table(r, stage_times(t0, dt, nsteps), out); return; }
  const int ns = (int)r->sources.size();
  const size_t nt = 4 * (size_t)nsteps;
  out.assign((size_t)r->ninst * nt * std::max(ns, 1), 0.0);
  for (int i = 0; i < r->ninst; ++i) {
    const std::vector<double> t = stage_times(t0 + r->toff[(size_t)i], dt, nsteps);
    for (size_t k = 0; k < nt; ++k)
      for (int s = 0; s < ns; ++s)
        out[((size_t)i * nt + k) * ns + s] =
            r->sources[s].amplitude * std::cos(r->sources[s].omega * t[k] + r->sources[s].phase);
  }
  r->src_inst_stride = (long)(nt * ns);
}

std::vector<double> stage_times(double t0, double dt, int nsteps) {
  std::vector<double> t(4 * (size_t)nsteps);
  double time = t0;
  for (int s = 0; s < nsteps; ++s) {
    t[4 * s + 0] = time;
    time = time + dt / 2.0;
    t[4 * s + 1] = time;
    t[4 * s + 2] = time;
    time = time + dt / 2.0;
    t[4 * s + 3] = time;
  }
  return t;
}

// checkSolutionLimits before a forward stage of a run (src/SolverImpl.f90:814-819)
void check_stage_input(mgd_region* r, const double* q) {
  if (r->lim_on && r->checking) launch_check_limits(r->geo, q, r->lim, r->d_flag.p, r->stream);
}

// One forward RK4 substep on the state in `qb` (all instances).  Where the rhs kernels
// write the new stage state out of place (the 3-D z march) it goes to `alt` and the two
// buffers trade places, so qb holds the new state either way.  With the third buffer of
// the pair (q_alt2, qs_alt2) stages 2 and 3 write there instead and read the state at the
// start of the step from `alt`, where stage 1 left it, so stage 1 writes no copy of it
// (buffers A B C: stage 1 A -> B, 2 B -> C, 3 C -> B, 4 B -> A, reading A in 2 and 3);
// every RK4 step has four stages, so after whole steps each buffer is back where it
// started.
void forward_stage(mgd_region* r, DevBuf& qb, DevBuf& alt, double* b1, double* b2, int sidx,
                   int stage, double dt, double* qcopy, double* Gbuf = nullptr) {
  double* q = qb.p;
  check_stage_input(r, q);
  const int na = n_act(r);
  const double* ctrl = nullptr;
  long cstride = 0;
  if (r->control.p && na > 0) {
    ctrl = r->control.p + (long)sidx * na;
    cstride = r->ctrl_stride;
  }
  const int ns = (int)r->sources.size();
  // the rhs launcher applies the RK update itself where its last kernel finishes the rhs
  const long M = (long)r->ninst * M_of(r);
  RkStage rk{stage, dt, q, b1, b2, qcopy};
  if (alt.p && (long)alt.n >= M) rk.yout = alt.p;   // allocated by the run preparation
  DevBuf* alt2 = &qb == &r->q ? &r->q_alt2 : (&qb == &r->qs ? &r->qs_alt2 : nullptr);
  const bool rot = rk.yout && alt2 && alt2->p && (long)alt2->n >= M && r->T1.p &&
                   fwd3d_march_path(r->geo, r->fin, (int)r->sources.size());
  if (rot && stage == 1) {
    rk.b1 = nullptr;
  } else if (rot && (stage == 2 || stage == 3)) {
    rk.b1 = alt.p;
    rk.yout = alt2->p;
  }
  if (r->timing) {
    const size_t k = 2 * (size_t)r->tcount;
    while (r->tev.size() < k + 2) {
      hipEvent_t e = nullptr;
      if (hipEventCreate(&e) != hipSuccess) { r->timing = false; break; }
      r->tev.push_back(e);
    }
    if (r->timing) {
      rk.ev_begin = r->tev[k];
      rk.ev_end = r->tev[k + 1];
    }
  }
  const int done = launch_rhs_forward(r->geo, dpatches(r), q, r->rhs.p, r->w.p,
                                      Gbuf ? Gbuf : r->G.p, r->t.p,
                                      r->T1.p, ctrl, cstride,
                                      reinterpret_cast<const Source*>(r->d_sources.p), ns,
                                      r->src_amp.p + (long)sidx * std::max(ns, 1),
                                      r->src_inst_stride, r->fin, r->stream, &rk);
  if (done == RK_TO_YOUT) {
    DevBuf& out = rot && (stage == 2 || stage == 3) ? *alt2 : alt;
    std::swap(qb.p, out.p);
    std::swap(qb.n, out.n);
    if (rk.ev_begin) {          // the z march ran between the two events
      ++r->tcount;
      if (qcopy) ++r->tcopy;
    }
  } else if (done == RK_NOT_APPLIED) {
    launch_rk_forward(M, stage, dt, q, b1, b2, r->rhs.p, qcopy, r->stream);
  }
}

// Whether the generic-kernel path's forward stages run on the 3-D z march (out of place:
// the stage writes its output wherever the caller says, and reads the state at the start
// of the step from where stage 1 read it).
bool march_stages(const mgd_region* r) {
  return r->nd == 3 && r->T1.p && fwd3d_march_path(r->geo, r->fin, (int)r->sources.size());
}

// One forward RK4 stage on the 3-D z march from the state at `in` to `out` (march_stages):
// `start` is the state at the start of the step (the input of stage 1, read by stages 2
// and 3), b2 the RK accumulator.  out must differ from in and start (other blocks still
// read in).  The step store passes its own slots, so no stage output is written twice;
// accumulate = false leaves out the RK accumulator (stages 1-3 of a recompute that stops
// before stage 4).
void march_stage(mgd_region* r, const double* in, double* out, const double* start, int sidx,
                 int stage, double dt, double* Gbuf = nullptr, bool accumulate = true) {
  check_stage_input(r, in);
  const int na = n_act(r);
  const double* ctrl = nullptr;
  long cstride = 0;
  if (r->control.p && na > 0) {
    ctrl = r->control.p + (long)sidx * na;
    cstride = r->ctrl_stride;
  }
  const int ns = (int)r->sources.size();
  RkStage rk{stage, dt, const_cast<double*>(in),
             stage == 2 || stage == 3 ? const_cast<double*>(start) : nullptr,
             accumulate || stage == 4 ? r->b2.p : nullptr, nullptr, out};
  if (r->timing) {
    const size_t k = 2 * (size_t)r->tcount;
    while (r->tev.size() < k + 2) {
      hipEvent_t e = nullptr;
      if (hipEventCreate(&e) != hipSuccess) { r->timing = false; break; }
      r->tev.push_back(e);
    }
    if (r->timing) {
      rk.ev_begin = r->tev[k];
      rk.ev_end = r->tev[k + 1];
    }
  }
  const int done = launch_rhs_forward(r->geo, dpatches(r), in, r->rhs.p, r->w.p,
                                      Gbuf ? Gbuf : r->G.p, r->t.p, r->T1.p, ctrl, cstride,
                                      reinterpret_cast<const Source*>(r->d_sources.p), ns,
                                      r->src_amp.p + (long)sidx * std::max(ns, 1),
                                      r->src_inst_stride, r->fin, r->stream, &rk);
  if (done == RK_TO_YOUT && rk.ev_begin) ++r->tcount;
}

bool has_walls(const mgd_region* r) {
  for (const auto& P : r->patches)
    if (P->pd.type == MGD_PATCH_IMPENETRABLE_WALL || P->pd.type == MGD_PATCH_ISOTHERMAL_WALL) return true;
  return false;
}

// The fused 2-D stage kernels cover inviscid 2-D regions with far-field, sponge,
// actuator, cost-target and Kolmogorov patches; walls run on the generic kernels, and so
// do OVERLAP-periodic directions: the tile's periodic halo is a shifted copy of the tile
// frame, in which the neighbour of the last distinct point is the first point rather
// than the repeated last grid line the reference's ghost exchange serves (the two differ
// when the repeated line holds values of its own).
bool fused_ok(const mgd_region* r) {
  bool overlap = false;
  for (int d = 0; d < r->nd; ++d) overlap = overlap || r->desc.periodicity[d] == MGD_PERIODIC_OVERLAP;
  return !r->force_plain && r->nd == 2 && !r->desc.viscous && r->tiling_ok && !has_walls(r) &&
         !overlap;
}

int upload_ints(DevBuf& b, const void* data, size_t bytes) {
  if (b.alloc(bytes / sizeof(double) + 1)) return MGD_ERR_HIP;
  MGD_HIP(hipMemcpy(b.p, data, bytes, hipMemcpyHostToDevice));
  return MGD_OK;
}

// Installs the tiling of one tile shape (host and device views).
// A tiling of `shape` for the region: the host partition in H, its patch masks uploaded to
// `mask`, the launch descriptor in T.
int make_tiling(mgd_region* r, int shape, HostTiling& Hout, Tiling& T, DevBuf& mask) {
  HostTiling H;
  if (fused2d_make_tiling(r->geo, r->hpatches, shape, &H) != MGD_OK) return MGD_ERR_UNSUPPORTED;
  if (upload_ints(mask, H.pmask.data(), H.pmask.size() * sizeof(unsigned))) return MGD_ERR_HIP;
  Hout = H;
  T.shape = H.shape;
  T.nbx = H.nbx;
  T.nby = H.nby;
  T.nblk = H.nblk;
  T.band = H.band;
  T.qx = H.qx;
  T.lx = H.lx;
  T.rx = H.rx;
  T.qy = H.qy;
  T.ly = H.ly;
  T.ry = H.ry;
  T.pmask = reinterpret_cast<const unsigned*>(mask.p);
  return MGD_OK;
}

int install_tiling(mgd_region* r, int shape) {
  return make_tiling(r, shape, r->htile, r->tiling, r->d_tile_mask);
}

// Exchanges the region's tiling with the group tiling held beside it.
void swap_tilings(mgd_region* r) {
  std::swap(r->htile, r->g_htile);
  std::swap(r->tiling, r->g_tiling);
  std::swap(r->d_tile_mask.p, r->g_tile_mask.p);
  std::swap(r->d_tile_mask.n, r->g_tile_mask.n);
  r->g_swapped = !r->g_swapped;
}

int setup_fused(mgd_region* r) {
  const size_t M = (size_t)r->ninst * r->nv * r->N;
  r->tiling_ok = false;
  r->tuned = false;
  // the first shape that fits (32 x 16 unless the grid is too small for it); the autotuner
  // may replace it before the first run
  int rc = MGD_ERR_UNSUPPORTED;
  if (r->tile_shape >= 0) {
    rc = install_tiling(r, r->tile_shape);
  } else {
    for (int sh = 0; sh < kNumTileShapes && rc == MGD_ERR_UNSUPPORTED; ++sh) rc = install_tiling(r, sh);
  }
  if (rc == MGD_ERR_UNSUPPORTED) return MGD_OK;   // plain path
  if (rc != MGD_OK) return rc;
  if (r->qy.alloc(M)) return MGD_ERR_HIP;
  r->tiling_ok = true;
  return MGD_OK;
}

constexpr int kChunkStages = 128;   // stages of quadrature terms held before their reduction

int prepare_quadrature(mgd_region* r, int nst, double dt) {
  // the weights depend on nst and dt only: uploaded again when either changes
  if (r->wts_nst != nst || r->wts_dt != dt || !r->wts.p) {
    std::vector<double> w((size_t)nst);
    for (int s = 0; s < nst; ++s) w[s] = kRkNorm[s % 4] * dt;
    r->wts_nst = -1;
    if (upload(r->wts, w)) return MGD_ERR_HIP;
    r->wts_nst = nst;
    r->wts_dt = dt;
  }
  if (r->ssum.alloc((size_t)r->ninst * std::max(nst, 1))) return MGD_ERR_HIP;
  r->chunk = std::max(1, std::min(nst, kChunkStages));
  const int nc = n_cost(r), na = n_act(r);
  if (nc > 0 && r->jpart.alloc((size_t)r->ninst * r->chunk * nc)) return MGD_ERR_HIP;
  if (na > 0 && r->spart.alloc((size_t)r->ninst * r->chunk * na)) return MGD_ERR_HIP;
  return MGD_OK;
}

// Stage sums of the chunk holding stage sidx once that chunk is complete: forward order
// completes a chunk at its last stage, adjoint order at its first.
void reduce_chunk(mgd_region* r, const DevBuf& part, int npts, int nst, int sidx, bool reverse) {
  if (npts <= 0) return;
  const int base = (sidx / r->chunk) * r->chunk;
  const int nfill = std::min(r->chunk, nst - base);
  const bool done = reverse ? sidx == base : sidx == base + nfill - 1;
  if (done) launch_stage_reduce(part.p, r->chunk, nfill, npts, r->ssum.p, nst, base, r->ninst, r->stream);
}

FwdStage fwd_args(mgd_region* r, const double* qin, const double* q0, double* b2, double* qout,
                  int stage, double dt, int sidx, int nst, bool with_j) {
  FwdStage a;
  std::memset(&a, 0, sizeof a);
  a.qin = qin; a.q0 = q0; a.b2 = b2; a.qout = qout; a.stage = stage; a.dt = dt;
  const int na = n_act(r);
  if (r->control.p && na > 0) {
    a.control = r->control.p + (long)sidx * na;
    a.control_stride = r->ctrl_stride;
  }
  const int ns = (int)r->sources.size();
  a.src = reinterpret_cast<const Source*>(r->d_sources.p);
  a.nsrc = ns;
  a.src_amp = r->src_amp.p + (long)sidx * std::max(ns, 1);
  a.src_inst_stride = r->src_inst_stride;
  a.ip_cost = r->ip_cost;
  (void)nst;
  if (with_j && n_cost(r) > 0) {
    a.jpartial = r->jpart.p;
    a.npts = n_cost(r);
    a.nslots = r->chunk;
    a.slot = sidx % r->chunk;
  }
  return a;
}

// one RK4 time step of the fused path: levels q0 -> q0, stage states in x / y; with
// `outs`, the outputs of stages 1-4 go to outs[0..3] instead (window recompute)
void fused_forward_step(mgd_region* r, double* q0, double* x, double* y, double* b2, int s, double dt,
                        int nst, bool with_j, double* const* outs, hipStream_t st = nullptr) {
  if (!st) st = r->stream;
  const double* qin[4] = {q0, x, y, x};
  double* qo[4] = {x, y, x, q0};
  for (int i = 1; i <= 4; ++i) {
    double* o = outs ? outs[i - 1] : qo[i - 1];
    const double* in = outs ? (i == 1 ? q0 : outs[i - 2]) : qin[i - 1];
    if (st == r->stream) check_stage_input(r, in);
    FwdStage a = fwd_args(r, in, q0, b2, o, i, dt, 4 * s + i - 1, nst, with_j);
    launch_fwd_stage2d(r->geo, dpatches(r), r->tiling, a, st);
    if (with_j) reduce_chunk(r, r->jpart, n_cost(r), nst, 4 * s + i - 1, false);
  }
}

// Autotuner: times one forward and one adjoint RK4 step (8 stage launches, the adjoint
// ones shared with recompute stages as in a run) with each
// tile shape that fits the grid and keeps the fastest.  J, the sensitivity and the
// gradient do not depend on the choice (per-point quadrature terms, fixed-order sums),
// nor does any rhs value.  Uses scratch buffers only (qs, qy, ly, l0, b2, qsr, b2r); the
// conserved and adjoint states are not touched.
int tune_tiling(mgd_region* r) {
  r->tuned = true;
  const long M = (long)r->ninst * M_of(r);
  if (r->qs.alloc(M) || r->qy.alloc(M) || r->ly.alloc(M) || r->l0.alloc(M) || r->qsr.alloc(M) ||
      r->b2r.alloc(M))
    return MGD_ERR_HIP;
  std::vector<double> amp;
  source_table(r, std::vector<double>{0.0, 0.5e-3, 0.5e-3, 1e-3}, amp);
  if (upload(r->src_amp, amp)) return MGD_ERR_HIP;
  if (r->control.p) r->control.alloc(0);
  const double dt = 1e-3;
  // one forward step, then one adjoint step with one recompute step as an adjoint run
  // issues them: each adjoint stage sharing its launch with a recompute stage (dual) or
  // the two launched one after the other
  auto step_pair = [&](bool dual) {
    fused_forward_step(r, r->qs.p, r->qy.p, r->ly.p, r->b2.p, 0, dt, 1, false, nullptr);
    for (int i = 4; i >= 1; --i) {
      AdjStage a;
      std::memset(&a, 0, sizeof a);
      a.lam = i == 4 ? r->l0.p : (i == 3 ? r->ly.p : (i == 2 ? r->qy.p : r->ly.p));
      a.lout = i == 4 ? r->ly.p : (i == 3 ? r->qy.p : (i == 2 ? r->ly.p : r->l0.p));
      a.l0 = r->l0.p; a.q = r->qs.p; a.b2 = r->b2.p;
      a.stage = i; a.dt = dt; a.forcing_factor = 1.0; a.ip_act = r->ip_act;
      const FwdStage f = fwd_args(r, r->qs.p, r->qs.p, r->b2r.p, r->qsr.p, i, dt, 0, 1, false);
      if (dual) {
        launch_dual_stage2d(r->geo, dpatches(r), r->tiling, a, f, r->stream);
      } else {
        launch_adj_stage2d(r->geo, dpatches(r), r->tiling, a, r->stream);
        launch_fwd_stage2d(r->geo, dpatches(r), r->tiling, f, r->stream);
      }
    }
  };
  // three interleaved rounds over (shape, dual or not); each keeps its fastest round
  float best_ms[kNumTileShapes][2];
  bool fits[kNumTileShapes];
  for (int sh = 0; sh < kNumTileShapes; ++sh) { best_ms[sh][0] = best_ms[sh][1] = 0.f; fits[sh] = false; }
  for (int round = 0; round < 3; ++round) {
    for (int sh = 0; sh < kNumTileShapes; ++sh) {
      if (round > 0 && !fits[sh]) continue;
      const int rc = install_tiling(r, sh);
      if (rc == MGD_ERR_UNSUPPORTED) continue;
      if (rc != MGD_OK) return rc;
      for (int dual = 0; dual < 2; ++dual) {
        launch_copy(M, r->q.p, r->qs.p, r->stream);
        launch_copy(M, r->q.p, r->l0.p, r->stream);
        step_pair(dual);   // warm-up
        MGD_HIP(hipEventRecord(r->ev0, r->stream));
        for (int k = 0; k < 2; ++k) step_pair(dual);
        MGD_HIP(hipEventRecord(r->ev1, r->stream));
        MGD_HIP(hipGetLastError());
        MGD_HIP(hipEventSynchronize(r->ev1));
        float ms = 0.f;
        MGD_HIP(hipEventElapsedTime(&ms, r->ev0, r->ev1));
        if (!fits[sh] || ms < best_ms[sh][dual]) best_ms[sh][dual] = ms;
      }
      fits[sh] = true;
    }
  }
  int best = -1, best_dual = 1;
  for (int sh = 0; sh < kNumTileShapes; ++sh)
    for (int dual = 0; dual < 2; ++dual)
      if (fits[sh] && (best < 0 || best_ms[sh][dual] < best_ms[best][best_dual])) {
        best = sh;
        best_dual = dual;
      }
  if (best < 0) { set_error("no tile shape fits the grid"); return MGD_ERR_UNSUPPORTED; }
  r->dual = best_dual != 0;
  return install_tiling(r, best);
}

// The shape of the 3-D marches' work units -- the z chunk (32, 64 or 128 planes) and the
// tile width (16 or 32 points, 16 rows) -- unless fixed by option: one forward RK4 step
// (stress pass and march per stage, out of place) and one adjoint rhs (stress pass and
// the two marches, where they serve the region) timed with HIP events for each shape,
// three interleaved rounds, the fastest round of each compared; the forward march (with
// the stress pass) and the adjoint marches each keep their own.  No value depends on the
// shape.
int tune_march(mgd_region* r) {
  r->zc_tuned = true;
  const long M = (long)r->ninst * M_of(r);
  if (int rc = ensure_adjoint(r)) return rc;
  if (r->qs.alloc(M) || r->qy.alloc(M) || r->qs_alt2.alloc(M)) return MGD_ERR_HIP;
  const double dt = 1e-3;
  // the state copies, one forward RK4 step, then (after event `mid`, if given) one adjoint rhs
  auto round_trip = [&](hipEvent_t mid = nullptr) {
    launch_copy(M, r->q.p, r->qs.p, r->stream);
    launch_copy(M, r->q.p, r->b2.p, r->stream);
    launch_copy(M, r->q.p, r->lam.p, r->stream);
    // the stages as the runs issue them: qs -> qy -> qs_alt2 -> qy -> qs, stages 2 and 3
    // reading the state at the start of the step from qs
    double* const in[4] = {r->qs.p, r->qy.p, r->qs_alt2.p, r->qy.p};
    double* const out[4] = {r->qy.p, r->qs_alt2.p, r->qy.p, r->qs.p};
    for (int i = 1; i <= 4; ++i) {
      RkStage rk{i, dt, in[i - 1], i == 1 ? nullptr : r->qs.p, r->b2.p, nullptr, out[i - 1]};
      if (r->geo.viscous) launch_stress3d(r->geo, rk.y, r->G.p, r->stream);
      launch_fwd3d_march(r->geo, rk.y, r->G.p, r->fin.kolmogorov, r->rhs.p, &rk, r->stream);
    }
    if (mid) (void)hipEventRecord(mid, r->stream);
    launch_rhs_adjoint(r->geo, dpatches(r), r->qs.p, r->lam.p, r->rhs.p, r->w.p, r->G.p, r->t.p,
                       r->T1.p, r->S.p, nullptr, 0, 1.0, r->fin, r->stream);
  };
  std::vector<std::pair<int, int>> shapes;   // (tile width, z chunk)
  for (int mt : {16, 32}) {
    if (r->opt_mtile != 0 && mt != r->opt_mtile) continue;
    if (mt == 32 && r->geo.n[0] % 32 != 0) continue;
    for (int zc : {32, 64, 128})
      if (r->opt_zchunk == 0 || zc == r->opt_zchunk) shapes.emplace_back(mt, zc);
  }
  if (shapes.empty()) shapes.emplace_back(16, 64);
  // the forward step and the adjoint rhs are timed apart and each keeps its own shape
  hipEvent_t evm = nullptr;
  MGD_HIP(hipEventCreate(&evm));
  std::vector<float> best_f(shapes.size(), 0.f), best_a(shapes.size(), 0.f);
  int rc = MGD_OK;
  for (int round = 0; round < 3 && rc == MGD_OK; ++round)
    for (size_t k = 0; k < shapes.size(); ++k) {
      r->geo.mtile = r->geo.mtile_adj = shapes[k].first;
      r->geo.zchunk = r->geo.zchunk_adj = shapes[k].second;
      round_trip();   // warm-up
      float ms_f = 0.f, ms_a = 0.f;
      if (hipEventRecord(r->ev0, r->stream) != hipSuccess) { rc = MGD_ERR_HIP; break; }
      round_trip(evm);
      if (hipEventRecord(r->ev1, r->stream) != hipSuccess || hipGetLastError() != hipSuccess ||
          hipEventSynchronize(r->ev1) != hipSuccess ||
          hipEventElapsedTime(&ms_f, r->ev0, evm) != hipSuccess ||
          hipEventElapsedTime(&ms_a, evm, r->ev1) != hipSuccess) {
        rc = MGD_ERR_HIP;
        break;
      }
      if (round == 0 || ms_f < best_f[k]) best_f[k] = ms_f;
      if (round == 0 || ms_a < best_a[k]) best_a[k] = ms_a;
    }
  (void)hipEventDestroy(evm);
  if (rc != MGD_OK) { set_error("HIP error in the march autotuner"); return rc; }
  size_t kf = 0, ka = 0;
  for (size_t k = 1; k < shapes.size(); ++k) {
    if (best_f[k] < best_f[kf]) kf = k;
    if (best_a[k] < best_a[ka]) ka = k;
  }
  r->geo.mtile = shapes[kf].first;
  r->geo.zchunk = shapes[kf].second;
  r->geo.mtile_adj = shapes[ka].first;
  r->geo.zchunk_adj = shapes[ka].second;
  return MGD_OK;
}

// Runs the autotuners once per region unless fixed by option: the tile shape of the fused
// 2-D kernels, the z chunk of the 3-D marches.
int maybe_tune(mgd_region* r) {
  if (r->nd == 3 && !r->zc_tuned && (r->opt_zchunk == 0 || r->opt_mtile == 0) &&
      fwd3d_march_path(r->geo, r->fin, (int)r->sources.size())) {
    if (int rc = ensure_state(r)) return rc;
    if (int rc = tune_march(r)) return rc;
  }
  if (!fused_ok(r) || r->tuned || r->tile_shape >= 0) return MGD_OK;
  return tune_tiling(r);
}

// Solver%runAdjoint on the fused path, split in its two passes.  Forward pass: snapshot
// every K steps.  Adjoint pass: marches back window by window (window w: steps
// [wK, min((w+1)K, nsteps))).  A window's stage states are recomputed from its snapshot
// into one of two window buffers, each recompute stage of window w-1 in the same launch
// as an adjoint stage of window w (k_dual_stage2d), so the two fill the GPU together
// (UniformCheckpointer%migrateTo semantics: the same states, each recomputed once).  Window slot 4m holds the state at step wK+m
// and slots 4m+1..4m+3 the inputs of its stages 2-4.  The forward pass writes the stage
// states of the last window straight into the first window buffer (the same launches,
// other output addresses), so the adjoint pass starts at once; windows then alternate
// between the two buffers from the last one down.
int fused_forward_pass(mgd_region* r, double dt, int nsteps, int K) {
  const int nc = n_cost(r), nst = 4 * nsteps;
  const long M = (long)r->ninst * M_of(r);
  if (r->resident) {
    // every stage state into the resident buffer: state 4 s + k is the input of stage
    // k + 1 of step s, state 4 nsteps the final state
    double* W = r->wall.p;
    launch_copy(M, r->q.p, W, r->stream);
    for (int s = 0; s < nsteps; ++s) {
      double* outs[4];
      for (int k = 0; k < 4; ++k) outs[k] = W + (4L * s + 1 + k) * M;
      fused_forward_step(r, W + 4L * s * M, nullptr, nullptr, r->b2.p, s, dt, nst, true, outs);
    }
    launch_copy(M, W + 4L * nsteps * M, r->q.p, r->stream);
    if (nc > 0) launch_stage_accumulate(r->ssum.p, nst, r->wts.p, r->acc_J.p, r->ninst, 0, r->stream);
    return MGD_OK;
  }
  if (r->step_mode) {
    // step store: each stage writes its output where it is kept -- the state at the start
    // of step s + 1 into the store (the final state into q), the inner stage states of a
    // held step into the step cache, the others into two slots of `window`
    r->rec_step = -1;
    launch_copy(M, r->q.p, r->sstore.p, r->stream);
    for (int s = 0; s < nsteps; ++s) {
      const bool held = s >= r->scache_first;
      double* outs[4];
      for (int k = 0; k < 3; ++k)
        outs[k] = held ? r->scache.p + ((long)(s - r->scache_first) * 3 + k) * M
                       : r->window.p + (long)(k % 2) * M;
      outs[3] = s + 1 < nsteps ? r->sstore.p + (long)(s + 1) * M : r->q.p;
      fused_forward_step(r, r->sstore.p + (long)s * M, nullptr, nullptr, r->b2.p, s, dt, nst, true, outs);
    }
    if (nc > 0) launch_stage_accumulate(r->ssum.p, nst, r->wts.p, r->acc_J.p, r->ninst, 0, r->stream);
    return MGD_OK;
  }
  const int t0last = ((nsteps + K - 1) / K - 1) * K;
  double* W = r->window.p;
  launch_copy(M, r->q.p, r->snaps.p, r->stream);
  for (int s = 0; s < nsteps; ++s) {
    if (s < t0last) {
      fused_forward_step(r, r->q.p, r->b1.p, r->qy.p, r->b2.p, s, dt, nst, true, nullptr);
    } else {
      if (s == t0last) launch_copy(M, r->q.p, W, r->stream);
      const long base = 4L * (s - t0last);
      double* outs[4];
      for (int k = 0; k < 4; ++k) outs[k] = W + (base + 1 + k) * M;
      if (s == nsteps - 1) outs[3] = r->q.p;   // final state
      fused_forward_step(r, W + base * M, nullptr, nullptr, r->b2.p, s, dt, nst, true, outs);
    }
    if ((s + 1) % K == 0) launch_copy(M, r->q.p, r->snaps.p + (long)((s + 1) / K) * M, r->stream);
  }
  if (nc > 0) launch_stage_accumulate(r->ssum.p, nst, r->wts.p, r->acc_J.p, r->ninst, 0, r->stream);
  r->last_window = true;
  return MGD_OK;
}

int fused_adjoint_pass(mgd_region* r, double dt, int nsteps, int K, bool terminal, bool want_grad,
                       bool last_window) {
  const int na = n_act(r), nc = n_cost(r), nst = 4 * nsteps;
  const long M = (long)r->ninst * M_of(r);
  double* wbuf[2] = {r->window.p, r->window2.p};
  const int nw = (nsteps + K - 1) / K;
  auto buf = [&](int w) { return (nw - 1 - w) % 2; };   // window buffer of window w
  // stage j of the recompute of window w into its buffer (slot 4m: the state at step
  // wK + m, slots 4m + 1 .. 4m + 3 the inputs of its stages 2-4; the state at the next
  // snapshot goes to qsr, which nothing reads)
  auto recompute_args = [&](int w, int j) {
    double* W = wbuf[buf(w)];
    const int t0w = w * K, t1w = std::min(nsteps, t0w + K);
    const int t = t0w + j / 4, i = j % 4 + 1;
    const long base = 4L * (t - t0w);
    double* q0 = W + base * M;
    double* out = t == t1w - 1 && i == 4 ? r->qsr.p : W + (base + i) * M;
    const double* in = i == 1 ? q0 : W + (base + i - 1) * M;
    return fwd_args(r, in, q0, r->b2r.p, out, i, dt, 4 * t + i - 1, nst, false);
  };
  auto recompute_stages = [&](int w) { return 4 * (std::min(nsteps, w * K + K) - w * K); };
  // terminal condition (loadInitialCondition ADJOINT, src/SolverImpl.f90:368-428): zero or
  // the given adjoint state, plus the cost forcing of the final state with factor -dt/6
  double* L0 = r->lam.p;
  double* LX = r->ly.p;
  double* LY = r->l0.p;
  if (!terminal) launch_fill(M, 0.0, L0, r->stream);
  if (nc > 0) {
    launch_adj_forcing(r->geo, dpatch(r, r->ip_cost), nc, r->qT.p, r->aforce.p, 0, r->stream);
    launch_adj_terminal(r->geo, dpatch(r, r->ip_cost), nc, r->aforce.p, -dt / 6.0, L0, r->stream);
  }
  // adjoint stage i of step t about the forward stage state qcur
  auto adj_args_q = [&](const double* qcur, int t, int i) {
    const int sidx = 4 * t + i - 1;
    AdjStage a;
    std::memset(&a, 0, sizeof a);
    a.lam = i == 4 ? L0 : (i == 3 ? LX : (i == 2 ? LY : LX));
    a.lout = i == 4 ? LX : (i == 3 ? LY : (i == 2 ? LX : L0));
    a.l0 = L0;
    a.q = qcur;
    a.b2 = r->b2.p;
    a.stage = i;
    a.dt = dt;
    a.forcing_factor = i == 4 ? 2.0 : (i == 2 ? 0.5 : 1.0);
    a.zero_forcing = t == 0 && i == 1;
    a.ip_act = r->ip_act;
    if (na > 0) {
      if (want_grad) {
        a.grad = r->grad.p + (long)sidx * na;
        a.grad_stride = (long)nst * na;
      }
      a.spartial = r->spart.p;
      a.npts = na;
      a.nslots = r->chunk;
      a.slot = sidx % r->chunk;
    }
    return a;
  };
  auto adj_args = [&](const double* W, int t0w, int t, int i) {
    return adj_args_q(W + (4L * (t - t0w) + (i - 1)) * M, t, i);
  };
  if (r->step_mode) {
    // step store: the stage states of step t are its stored state and its inner states
    // from the step cache or, for a step not held, recomputed into rbuf[t % 2]; the three
    // recompute stages of step t - 1 share their launches with the adjoint stages 4, 3, 2
    // of step t (or follow them), into the other buffer
    double* const rbuf[2] = {r->window.p, r->window2.p};
    auto held = [&](int t) { return t >= r->scache_first; };
    auto rec_args = [&](int t, int i) {
      double* R = rbuf[t % 2];
      const double* start = r->sstore.p + (long)t * M;
      const double* in = i == 1 ? start : R + (long)(i - 2) * M;
      return fwd_args(r, in, start, r->b2r.p, R + (long)(i - 1) * M, i, dt, 4 * t + i - 1, nst, false);
    };
    if (!held(nsteps - 1))
      for (int i = 1; i <= 3; ++i)
        launch_fwd_stage2d(r->geo, dpatches(r), r->tiling, rec_args(nsteps - 1, i), r->stream);
    for (int t = nsteps - 1; t >= 0; --t) {
      const double* qst[4];
      qst[0] = r->sstore.p + (long)t * M;
      for (int k = 1; k < 4; ++k)
        qst[k] = held(t) ? r->scache.p + ((long)(t - r->scache_first) * 3 + k - 1) * M
                         : rbuf[t % 2] + (long)(k - 1) * M;
      const int nrec = t >= 1 && !held(t - 1) ? 3 : 0;
      int j = 0;
      for (int i = 4; i >= 1; --i) {
        const AdjStage a = adj_args_q(qst[i - 1], t, i);
        if (j < nrec && r->dual) {
          ++j;
          launch_dual_stage2d(r->geo, dpatches(r), r->tiling, a, rec_args(t - 1, j), r->stream);
        } else {
          launch_adj_stage2d(r->geo, dpatches(r), r->tiling, a, r->stream);
          if (j < nrec) {
            ++j;
            launch_fwd_stage2d(r->geo, dpatches(r), r->tiling, rec_args(t - 1, j), r->stream);
          }
        }
        reduce_chunk(r, r->spart, na, nst, 4 * t + i - 1, true);
      }
    }
    r->rec_step = -1;   // `window` holds the inner states of another step than before
  } else if (r->resident) {
    for (int t = nsteps - 1; t >= 0; --t)
      for (int i = 4; i >= 1; --i) {
        launch_adj_stage2d(r->geo, dpatches(r), r->tiling, adj_args(r->wall.p, 0, t, i), r->stream);
        reduce_chunk(r, r->spart, na, nst, 4 * t + i - 1, true);
      }
  } else {
    // window by window from the last one (its stage states written by the forward pass,
    // or recomputed first); each adjoint stage of window w shares its launch with the
    // matching stage of the recompute of window w-1 (the other window buffer) or is
    // followed by it, the recompute's stages left over running alone
    if (!last_window) {
      launch_copy(M, r->snaps.p + (long)(nw - 1) * M, wbuf[buf(nw - 1)], r->stream);
      for (int j = 0; j < recompute_stages(nw - 1); ++j)
        launch_fwd_stage2d(r->geo, dpatches(r), r->tiling, recompute_args(nw - 1, j), r->stream);
    }
    for (int w = nw - 1; w >= 0; --w) {
      const int t0w = w * K, t1w = std::min(nsteps, t0w + K);
      const int nrec = w >= 1 ? recompute_stages(w - 1) : 0;
      if (nrec > 0) launch_copy(M, r->snaps.p + (long)(w - 1) * M, wbuf[buf(w - 1)], r->stream);
      int j = 0;
      for (int t = t1w - 1; t >= t0w; --t)
        for (int i = 4; i >= 1; --i) {
          const AdjStage a = adj_args(wbuf[buf(w)], t0w, t, i);
          if (j < nrec && r->dual) {
            launch_dual_stage2d(r->geo, dpatches(r), r->tiling, a, recompute_args(w - 1, j++), r->stream);
          } else {
            launch_adj_stage2d(r->geo, dpatches(r), r->tiling, a, r->stream);
            if (j < nrec) launch_fwd_stage2d(r->geo, dpatches(r), r->tiling, recompute_args(w - 1, j++), r->stream);
          }
          reduce_chunk(r, r->spart, na, nst, 4 * t + i - 1, true);
        }
      for (; j < nrec; ++j)
        launch_fwd_stage2d(r->geo, dpatches(r), r->tiling, recompute_args(w - 1, j), r->stream);
    }
  }
  if (na > 0) launch_stage_accumulate(r->ssum.p, nst, r->wts.p, r->acc_S.p, r->ninst, 1, r->stream);
  return MGD_OK;
}

// Plain path of the two passes (generic per-operator kernels).
// Stage states of the window starting at step lt, when the last checkpointed forward pass
// kept them; null otherwise.
double* cached_window(mgd_region* r, int lt) {
  if (r->ncache <= 0 || lt < r->cache_first) return nullptr;
  const int k = (lt - r->cache_first) / std::max(r->ck.K, 1);
  return k < r->ncache ? r->wcache[k].p : nullptr;
}

// Device memory a new storage plan of a checkpointed run may take: what is free plus the
// storage held for earlier runs, less a reserve of max(8 GiB, 1/32 of the device) for
// what the runtime and other allocations of the process take later.
size_t storage_budget(mgd_region* r) {
  size_t fr = 0, tot = 0;
  if (hipMemGetInfo(&fr, &tot) != hipSuccess) return 0;
  size_t held = r->window.bytes() + r->window2.bytes() + r->snaps.bytes() + r->sstore.bytes() +
                r->scache.bytes() + r->sbuf.bytes() + r->wall.bytes();
  for (int k = 0; k < mgd_region::kMaxCache; ++k) held += r->wcache[k].bytes();
  const size_t reserve = std::max((size_t)8 << 30, tot / 32);
  return fr + held > reserve ? fr + held - reserve : 0;
}

// How many of the last windows of a checkpointed run the forward pass keeps: as many as
// MGD_OPT_WINDOW_CACHE allows and, by default, as fit in storage_budget (the generic-kernel
// path only; the fused 2-D path keeps its last window itself).
int plan_window_cache(mgd_region* r, int nsteps, int K, long M) {
  r->ncache = 0;
  const int nwin = (nsteps + K - 1) / K;
  int want = r->opt_cache < 0 ? nwin : std::min(r->opt_cache, nwin);
  want = std::min(want, mgd_region::kMaxCache);
  if (fused_ok(r)) want = 0;
  const size_t bytes = (size_t)4 * K * M * sizeof(double);
  for (int k = want; k < mgd_region::kMaxCache; ++k) r->wcache[k].release();
  if (want > 0 && r->opt_cache < 0) {
    // counts the window buffer and snapshots of this run as taken
    const size_t budget = storage_budget(r);
    const size_t taken = r->window.bytes() + r->snaps.bytes();
    const size_t avail = budget > taken ? budget - taken : 0;
    want = std::min(want, (int)(avail / bytes));
    for (int k = want; k < mgd_region::kMaxCache; ++k) r->wcache[k].release();
  }
  int n = 0;
  while (n < want && r->wcache[n].alloc((size_t)4 * K * M) == MGD_OK) ++n;
  if (n < want) (void)hipGetLastError();   // fewer windows than planned: the rest recompute
  r->ncache = n;
  r->cache_first = (nwin - n) * K;
  return MGD_OK;
}

// Storage of a checkpointed run on the generic-kernel path, and on the fused 2-D path where
// the stage states are not resident.  Step mode (MGD_OPT_STEP_STORE 1, or -1 with
// MGD_OPT_WINDOW_CACHE at its default, where the states of every step and the recompute
// buffers fit in storage_budget): the state at the start of every step, and the inner
// stage states of as many of the last steps as the rest of the budget holds
// (MGD_OPT_WINDOW_CACHE n >= 0: at most n save_interval steps); the adjoint pass recomputes
// three stages per step not held, against 4 save_interval - 1 per window in window mode.
// The recompute buffer holds three states; the fused path has two (the recompute of one
// step shares its launches with the adjoint stages of the next).  Window mode (always for
// the fused members of a group launch): snapshots every K steps, a window buffer and, on
// the generic kernels, whole cached windows.
int plan_plain_storage(mgd_region* r, int nsteps, int K, long M) {
  const size_t S = (size_t)M * sizeof(double);
  const bool fused = fused_ok(r);
  const size_t nrec = fused ? 6 : 3;   // states of the recompute buffers
  r->rec_step = -1;
  r->step_mode = false;
  const bool allowed = !(fused && r->group_run);
  bool step = r->opt_step == 1 && allowed;
  size_t budget = 0;
  if (r->opt_step != 0 && allowed) {
    budget = storage_budget(r);
    if (r->opt_step < 0) step = r->opt_cache < 0 && (size_t)(nsteps + nrec) * S <= budget;
  }
  size_t fixed = (size_t)(nsteps + nrec) * S;
  if (step) {
    for (int k = 0; k < mgd_region::kMaxCache; ++k) r->wcache[k].release();
    r->ncache = 0;
    r->snaps.release();
    if (r->window.n != (size_t)3 * M) r->window.release();
    if (fused && r->window2.n != (size_t)3 * M) r->window2.release();
    if (r->sstore.n != (size_t)nsteps * M) r->sstore.release();
    if (r->window.alloc((size_t)3 * M) || (fused && r->window2.alloc((size_t)3 * M)) ||
        r->sstore.alloc((size_t)nsteps * M)) {
      // taken meanwhile (regions of a group plan one after another): window mode
      (void)hipGetLastError();
      step = false;
    }
  }
  r->sreuse = false;
  if (step && !fused && stress_reuse_path(r->geo, r->fin, (int)r->sources.size()) && r->G.n > 0 &&
      fixed + 3 * r->G.bytes() <= budget) {
    if (r->sbuf.alloc(3 * r->G.n) == MGD_OK) {
      r->sreuse = true;
      fixed += 3 * r->G.bytes();
    } else {
      (void)hipGetLastError();
    }
  }
  if (!r->sreuse) r->sbuf.release();
  if (step) {
    const size_t rest = budget > fixed ? budget - fixed : 0;
    int nsc = (int)std::min<size_t>((size_t)nsteps, rest / (3 * S));
    if (r->opt_cache >= 0) nsc = std::min(nsc, r->opt_cache * K);
    // a cache held from an earlier run stays unless the plan is well above it (the free
    // memory moves a little between runs; reallocating ~100 GB each run would cost more
    // than the few steps it adds)
    const size_t step3 = (size_t)3 * M;
    const int have = r->scache.n % step3 == 0 ? (int)(r->scache.n / step3) : -1;
    if (have > 0 && have <= nsteps && (r->opt_cache < 0 || have <= r->opt_cache * K) &&
        nsc <= have + std::max(2, have / 16))
      nsc = have;
    if (r->scache.n != (size_t)3 * nsc * M) r->scache.release();
    while (nsc > 0 && r->scache.alloc((size_t)3 * nsc * M) != MGD_OK) {
      (void)hipGetLastError();             // fewer steps than planned: the rest recompute
      nsc /= 2;
    }
    r->nscache = nsc;
    r->scache_first = nsteps - nsc;
    r->step_mode = true;
    return MGD_OK;
  }
  r->sstore.release();
  r->scache.release();
  r->sbuf.release();
  r->nscache = 0;
  const int nsnap = nsteps / K + 1;
  if (r->snaps.alloc((size_t)nsnap * M) || r->window.alloc((size_t)4 * K * M)) return MGD_ERR_HIP;
  return plan_window_cache(r, nsteps, K, M);
}

// Step mode: the inner stage states Y2, Y3, Y4 of step t (the outputs of its stages 1-3)
// from the stored state of step t into `window`; the scratch state buffers keep their
// places.
void recompute_step(mgd_region* r, int t, double dt) {
  const long M = (long)r->ninst * M_of(r);
  if (fused_ok(r)) {
    // fused forward stages 1-3 from the stored state straight into the three slots of
    // `window` (the accumulator b2r is scratch: no stage 4 follows)
    const double* start = r->sstore.p + (long)t * M;
    for (int st = 1; st <= 3; ++st) {
      const FwdStage a = fwd_args(r, st == 1 ? start : r->window.p + (long)(st - 2) * M, start,
                                  r->b2r.p, r->window.p + (long)(st - 1) * M, st, dt,
                                  4 * t + st - 1, 4 * r->ck.nsteps, false);
      launch_fwd_stage2d(r->geo, dpatches(r), r->tiling, a, r->stream);
    }
    r->rec_step = t;
    return;
  }
  if (march_stages(r)) {
    // from the stored state straight into the three slots of `window`
    const double* start = r->sstore.p + (long)t * M;
    for (int st = 1; st <= 3; ++st)
      march_stage(r, st == 1 ? start : r->window.p + (long)(st - 2) * M,
                  r->window.p + (long)(st - 1) * M, start, 4 * t + st - 1, st, dt,
                  r->sreuse ? r->sbuf.p + (long)(st - 1) * (long)r->G.n : nullptr, false);
    r->rec_step = t;
    return;
  }
  launch_copy(M, r->sstore.p + (long)t * M, r->qs.p, r->stream);
  double* const p0 = r->qs.p;
  for (int st = 1; st <= 3; ++st)
    forward_stage(r, r->qs, r->qs_alt, r->b1.p, r->b2.p, 4 * t + st - 1, st, dt,
                  r->window.p + (long)(st - 1) * M,
                  r->sreuse ? r->sbuf.p + (long)(st - 1) * (long)r->G.n : nullptr);
  if (r->qs.p != p0) {
    std::swap(r->qs.p, r->qs_alt.p);
    std::swap(r->qs.n, r->qs_alt.n);
  }
  r->rec_step = t;
}

// Step mode: forward stage state `stage` (1: the state at the start of step t, 2-4: the
// outputs of its stages 1-3), recomputing the step's inner states where they are not held.
const double* step_stage_state(mgd_region* r, int t, int stage, double dt, bool recompute) {
  const long M = (long)r->ninst * M_of(r);
  if (stage == 1) return r->sstore.p + (long)t * M;
  if (t >= r->scache_first) return r->scache.p + ((long)(t - r->scache_first) * 3 + stage - 2) * M;
  if (recompute) recompute_step(r, t, dt);
  return r->window.p + (long)(stage - 2) * M;
}

int plain_forward_pass(mgd_region* r, double dt, int nsteps, int K) {
  const int nc = n_cost(r);
  const long M = (long)r->ninst * M_of(r);
  if (r->step_mode && march_stages(r)) {
    // the 3-D marches write each stage state straight where it is kept: the state at the
    // start of step s + 1 into the store (the final state into q), the inner stage states
    // of a held step into the step cache, the others into the two scratch state buffers
    r->rec_step = -1;
    launch_copy(M, r->q.p, r->sstore.p, r->stream);
    for (int s = 0; s < nsteps; ++s) {
      const double* start = r->sstore.p + (long)s * M;
      const bool held = s >= r->scache_first;
      double* y[5];
      y[1] = const_cast<double*>(start);
      for (int k = 2; k <= 4; ++k)
        y[k] = held ? r->scache.p + ((long)(s - r->scache_first) * 3 + k - 2) * M
                    : (k == 3 ? r->q_alt2.p : r->q_alt.p);
      double* const y5 = s + 1 < nsteps ? r->sstore.p + (long)(s + 1) * M : r->q.p;
      for (int i = 1; i <= 4; ++i) {
        double* out = i < 4 ? y[i + 1] : y5;
        march_stage(r, y[i], out, start, 4 * s + i - 1, i, dt);
        if (nc > 0)
          launch_patch_reduce(r->geo, dpatch(r, r->ip_cost), nc, 0, out, r->partial.p,
                              kRkNorm[i - 1] * dt, r->acc_J.p, nullptr, r->stream);
      }
    }
    return MGD_OK;
  }
  if (r->step_mode) {
    // the state at the start of every step (stage 4 of the step before writes it) and the
    // inner stage states of the last nscache steps
    r->rec_step = -1;
    launch_copy(M, r->q.p, r->sstore.p, r->stream);
    for (int s = 0; s < nsteps; ++s)
      for (int i = 1; i <= 4; ++i) {
        double* copy = nullptr;
        if (i < 4 && s >= r->scache_first)
          copy = r->scache.p + ((long)(s - r->scache_first) * 3 + i - 1) * M;
        else if (i == 4 && s + 1 < nsteps)
          copy = r->sstore.p + (long)(s + 1) * M;
        forward_stage(r, r->q, r->q_alt, r->b1.p, r->b2.p, 4 * s + i - 1, i, dt, copy);
        if (nc > 0)
          launch_patch_reduce(r->geo, dpatch(r, r->ip_cost), nc, 0, r->q.p, r->partial.p,
                              kRkNorm[i - 1] * dt, r->acc_J.p, nullptr, r->stream);
      }
    return MGD_OK;
  }
  // forward run with snapshots at t0 and every K steps (showProgress .q files); in the
  // cached windows every stage state also goes to the window's cache
  launch_copy(M, r->q.p, r->snaps.p, r->stream);
  for (int s = 0; s < nsteps; ++s) {
    const int lt = s - s % K;
    double* cw = cached_window(r, lt);
    if (cw && s == lt) launch_copy(M, r->q.p, cw, r->stream);
    for (int i = 1; i <= 4; ++i) {
      const int idx = 4 * (s - lt) + i;
      forward_stage(r, r->q, r->q_alt, r->b1.p, r->b2.p, 4 * s + i - 1, i, dt,
                    cw && idx < 4 * K ? cw + (long)idx * M : nullptr);
      if (nc > 0)
        launch_patch_reduce(r->geo, dpatch(r, r->ip_cost), nc, 0, r->q.p, r->partial.p,
                            kRkNorm[i - 1] * dt, r->acc_J.p, nullptr, r->stream);
    }
    if ((s + 1) % K == 0) launch_copy(M, r->q.p, r->snaps.p + (long)((s + 1) / K) * M, r->stream);
  }
  return MGD_OK;
}

// One adjoint RK4 stage (stage i of step t) of the generic-kernel path, linearised about
// the forward stage state qcur.
void plain_adjoint_stage(mgd_region* r, double dt, int nsteps, int t, int i, const double* qcur,
                         bool want_grad, double* stress = nullptr) {
  const int na = n_act(r), nc = n_cost(r), nst = 4 * nsteps;
  const long M = (long)r->ninst * M_of(r);
  const int sidx = 4 * t + i - 1;
  if (na > 0) {
    if (want_grad)
      launch_gradient(r->geo, dpatch(r, r->ip_act), na, r->lam.p, r->grad.p + (long)sidx * na,
                      (long)nst * na, r->stream);
    launch_patch_reduce(r->geo, dpatch(r, r->ip_act), na, 1, r->lam.p, r->partial.p,
                        kRkNorm[i - 1] * dt, r->acc_S.p, nullptr, r->stream);
  }
  if (nc > 0)
    launch_adj_forcing(r->geo, dpatch(r, r->ip_cost), nc, qcur, r->aforce.p, t == 0 && i == 1,
                       r->stream);
  const double factor = i == 4 ? 2.0 : (i == 2 ? 0.5 : 1.0);
  // the rhs launcher applies the RK update itself where its last kernel finishes the rhs
  RkStage rk{i, dt, r->lam.p, r->b1a.p, r->b2a.p};
  rk.stress_ready = stress != nullptr;
  if (!launch_rhs_adjoint(r->geo, dpatches(r), qcur, r->lam.p, r->rhs.p, r->w.p,
                          stress ? stress : r->G.p, r->t.p, r->T1.p, r->S.p,
                          nc > 0 ? r->aforce.p : nullptr, (long)r->nv * nc, factor, r->fin,
                          r->stream, &rk))
    launch_rk_adjoint(M, i, dt, r->lam.p, r->b1a.p, r->b2a.p, r->rhs.p, r->stream);
}

int plain_adjoint_pass(mgd_region* r, double dt, int nsteps, int K, bool terminal, bool want_grad) {
  const int nc = n_cost(r);
  const long M = (long)r->ninst * M_of(r);
  // adjoint terminal condition (loadInitialCondition ADJOINT, src/SolverImpl.f90:368-428)
  if (!terminal) launch_fill(M, 0.0, r->lam.p, r->stream);
  if (nc > 0) {
    launch_adj_forcing(r->geo, dpatch(r, r->ip_cost), nc, r->qT.p, r->aforce.p, 0, r->stream);
    launch_adj_terminal(r->geo, dpatch(r, r->ip_cost), nc, r->aforce.p, -dt / 6.0, r->lam.p, r->stream);
  }
  if (r->step_mode) {
    // every step's inner stage states are recomputed in this pass (a captured pass replays
    // the recompute too), the held ones read from the step cache
    for (int t = nsteps - 1; t >= 0; --t) {
      // a recomputed step left the stress slots of its stage states 1-3 in sbuf
      const bool reuse = r->sreuse && t < r->scache_first;
      for (int i = 4; i >= 1; --i)
        plain_adjoint_stage(r, dt, nsteps, t, i, step_stage_state(r, t, i, dt, i == 4), want_grad,
                            reuse && i < 4 ? r->sbuf.p + (long)(i - 1) * (long)r->G.n : nullptr);
    }
    r->rec_step = -1;
    return MGD_OK;
  }
  int loaded = -1;
  const double* base = r->window.p;      // stage states of the loaded window
  for (int t = nsteps - 1; t >= 0; --t) {
    for (int i = 4; i >= 1; --i) {
      const int ts = i == 1 ? t : t + 1;
      const int st = i == 1 ? 4 : i - 1;
      const bool need = loaded == -1 || ts < loaded || ts > loaded + K ||
                        (ts == loaded && st < 4) || (ts == loaded + K && st == 4);
      if (need) {
        int lt;
        if (ts % K == 0 && st == 4) lt = ts;
        else if (ts % K == 0) lt = ts - K;
        else lt = ts - ts % K;
        loaded = lt;
        const double* cw = lt < nsteps ? cached_window(r, lt) : nullptr;
        base = cw ? cw : r->window.p;
        if (!cw) {                       // else kept by the forward pass
          launch_copy(M, r->snaps.p + (long)(lt / K) * M, r->qs.p, r->stream);
          launch_copy(M, r->qs.p, r->window.p, r->stream);
          if (loaded != nsteps) {
            int idx = 1;
            for (int s2 = loaded + 1; s2 <= loaded + K && s2 <= nsteps; ++s2)
              for (int st2 = 1; st2 <= 4; ++st2) {
                // the window's last stage gives the next snapshot, which no stage reads
                if (s2 == loaded + K && st2 == 4) break;
                forward_stage(r, r->qs, r->qs_alt, r->b1.p, r->b2.p, 4 * (s2 - 1) + st2 - 1, st2,
                              dt, r->window.p + (long)idx * M);
                ++idx;
              }
          }
        }
      }
      const int widx = (ts - 1 - loaded) * 4 + st;
      plain_adjoint_stage(r, dt, nsteps, t, i, base + (long)widx * M, want_grad);
    }
  }
  return MGD_OK;
}

// Uploads the per-run inputs (actuator control of every stage, source amplitudes at the
// stage times) of the checkpointed run held in r->ck.
int upload_run_inputs(mgd_region* r) {
  const int na = n_act(r), nst = 4 * r->ck.nsteps;
  if (!r->ck.control.empty() && na > 0) {
    if (upload(r->control, r->ck.control)) return MGD_ERR_HIP;
    r->ctrl_stride = (long)nst * na;
  } else if (r->control.p) {
    r->control.alloc(0);
  }
  std::vector<double> amp;
  run_source_table(r, r->ck.t0, r->ck.dt, r->ck.nsteps, amp);
  if (upload(r->src_amp, amp)) return MGD_ERR_HIP;
  return MGD_OK;
}

// Forward pass with HBM snapshots (first half of Solver%runAdjoint; the snapshots play
// the part of the .q files the reference writes every save_interval steps).  Split in a
// preparation (allocations and host uploads) and the device work, which can be captured
// into a HIP graph.
int prepare_forward_ck(mgd_region* r, double t0, double dt, int nsteps, int K, const double* control) {
  if (int rc = ensure_adjoint(r)) return rc;
  if (int rc = maybe_tune(r)) return rc;
  const int na = n_act(r), nst = 4 * nsteps;
  const long M = (long)r->ninst * M_of(r);
  r->ck.valid = false;
  r->last_window = false;
  r->mig_loaded = -1;
  r->ck.t0 = t0;
  r->ck.dt = dt;
  r->ck.nsteps = nsteps;
  r->ck.K = K;
  if (control && na > 0) r->ck.control.assign(control, control + (size_t)r->ninst * nst * na);
  else r->ck.control.clear();
  if (int rc = upload_run_inputs(r)) return rc;
  if (r->qs.alloc(M) || r->qT.alloc(M) || (r->nd == 3 && (r->qs_alt.alloc(M) || r->qs_alt2.alloc(M))))
    return MGD_ERR_HIP;
  if (!fused_ok(r)) {
    if (int rc = plan_plain_storage(r, nsteps, K, M)) return rc;
  } else {
    const int nsnap = nsteps / K + 1;
    r->step_mode = false;
    // every stage state resident (MGD_OPT_RESIDENT_STATES): asked for, or by default where
    // they fit in storage_budget (the members of a group launch: where every member asks
    // for it); otherwise the step store or snapshots and windows, as on the generic kernels
    const size_t need = (size_t)(4L * nsteps + 1) * M;
    bool res = r->opt_resident == 1;
    if (r->group_run) res = r->group_resident;
    else if (r->opt_resident < 0) res = need * sizeof(double) <= storage_budget(r);
    if (res) {
      r->sstore.release();
      r->scache.release();
      r->sbuf.release();
    } else {
      r->wall.release();
    }
    if (res && r->wall.alloc(need)) {
      if (r->opt_resident == 1) {
        set_error("not enough device memory to keep every stage state resident");
        return MGD_ERR_HIP;
      }
      (void)hipGetLastError();
      res = false;
    }
    r->resident = res;
    if (res) {
      if (r->snaps.alloc((size_t)nsnap * M) || r->window.alloc((size_t)4 * K * M)) return MGD_ERR_HIP;
      if (int rc = plan_window_cache(r, nsteps, K, M)) return rc;
    } else {
      if (int rc = plan_plain_storage(r, nsteps, K, M)) return rc;
    }
  }
  if (fused_ok(r)) {
    if (int rc = prepare_quadrature(r, nst, dt)) return rc;
  }
  return MGD_OK;
}

int enqueue_forward_ck(mgd_region* r) {
  const long M = (long)r->ninst * M_of(r);
  launch_fill(r->ninst, 0.0, r->acc_J.p, r->stream);
  if (r->lim_on) launch_fill(r->ninst, 0.0, r->d_flag.p, r->stream);
  r->checking = true;
  const int rc = fused_ok(r) ? fused_forward_pass(r, r->ck.dt, r->ck.nsteps, r->ck.K)
                             : plain_forward_pass(r, r->ck.dt, r->ck.nsteps, r->ck.K);
  r->checking = false;
  if (rc) return rc;
  launch_copy(M, r->q.p, r->qT.p, r->stream);   // final state: terminal forcing, MGD_FINAL_STATE
  return MGD_OK;
}

// Adjoint pass from the snapshots of the last checkpointed forward pass.  inputs_fresh:
// the run's inputs were uploaded by prepare_forward_ck just before (a group run), so they
// are not uploaded again.
int prepare_adjoint_ck(mgd_region* r, bool inputs_fresh = false) {
  if (!r->ck.valid) { set_error("no checkpointed forward run to march back from"); return MGD_ERR_STATE; }
  const int na = n_act(r), nst = 4 * r->ck.nsteps;
  const long M = (long)r->ninst * M_of(r);
  if (!inputs_fresh) {
    if (int rc = upload_run_inputs(r)) return rc;
  }
  if (na > 0 && r->grad.alloc((size_t)r->ninst * nst * na)) return MGD_ERR_HIP;
  if (fused_ok(r)) {
    if (int rc = prepare_quadrature(r, nst, r->ck.dt)) return rc;
    if (r->ly.alloc(M) || r->l0.alloc(M) ||
        r->window2.alloc((size_t)(r->step_mode ? 3 : 4 * r->ck.K) * M) || r->b2r.alloc(M) ||
        r->qsr.alloc(M))
      return MGD_ERR_HIP;
  }
  return MGD_OK;
}

int enqueue_adjoint_ck(mgd_region* r, bool terminal, bool want_grad) {
  // the window buffers are rewritten by every adjoint pass: the last window written by
  // the forward pass serves the first adjoint pass after it only
  const bool last_window = r->last_window;
  r->last_window = false;
  launch_fill(r->ninst, 0.0, r->acc_S.p, r->stream);
  return fused_ok(r) ? fused_adjoint_pass(r, r->ck.dt, r->ck.nsteps, r->ck.K, terminal, want_grad, last_window)
                     : plain_adjoint_pass(r, r->ck.dt, r->ck.nsteps, r->ck.K, terminal, want_grad);
}

// Forward run without snapshots (Solver%runForward): device work of one run.
int enqueue_forward_run(mgd_region* r, double dt, int nsteps) {
  const int nst = 4 * nsteps;
  launch_fill(r->ninst, 0.0, r->acc_J.p, r->stream);
  if (r->lim_on) launch_fill(r->ninst, 0.0, r->d_flag.p, r->stream);
  r->checking = true;
  if (fused_ok(r)) {
    for (int s = 0; s < nsteps; ++s)
      fused_forward_step(r, r->q.p, r->b1.p, r->qy.p, r->b2.p, s, dt, nst, true, nullptr);
    if (n_cost(r) > 0)
      launch_stage_accumulate(r->ssum.p, nst, r->wts.p, r->acc_J.p, r->ninst, 0, r->stream);
    r->checking = false;
    return MGD_OK;
  }
  for (int s = 0; s < nsteps; ++s)
    for (int i = 1; i <= 4; ++i) {
      const int sidx = 4 * s + i - 1;
      forward_stage(r, r->q, r->q_alt, r->b1.p, r->b2.p, sidx, i, dt, nullptr);
      if (r->ip_cost >= 0)
        launch_patch_reduce(r->geo, dpatch(r, r->ip_cost), n_cost(r), 0, r->q.p, r->partial.p,
                            kRkNorm[i - 1] * dt, r->acc_J.p, nullptr, r->stream);
    }
  r->checking = false;
  return MGD_OK;
}

// After a run's forward pass: instances that reached a solution limit get the cost
// HUGE (costFunctional = HUGE(0.0_wp), src/SolverImpl.f90:814-819) and, in adjoint runs,
// the sensitivity HUGE; the return value is MGD_SOLUTION_LIMIT if any did.
int limits_outcome(mgd_region* r, double* J, double* sensitivity) {
  if (!r->lim_on) return MGD_OK;
  std::vector<double> f(r->ninst, 0.0);
  MGD_HIP(hipMemcpy(f.data(), r->d_flag.p, sizeof(double) * r->ninst, hipMemcpyDeviceToHost));
  int rc = MGD_OK;
  for (int i = 0; i < r->ninst; ++i)
    if (f[i] != 0.0) {
      rc = MGD_SOLUTION_LIMIT;
      if (J) J[i] = std::numeric_limits<double>::max();
      if (sensitivity) sensitivity[i] = std::numeric_limits<double>::max();
    }
  if (rc) set_error("a solution limit was reached (density or temperature out of range)");
  return rc;
}

// HIP graph of a run's device work, replayed while the run parameters and every device
// buffer the work touches stay the same (the key); inputs that change between runs
// (control, source amplitudes, states) live in those buffers and are re-uploaded before
// each replay.  Captured work holds kernels, copies, fills and event waits only.
std::vector<uintptr_t> graph_key(mgd_region* r, int kind, std::initializer_list<uintptr_t> extra) {
  std::vector<uintptr_t> k = {(uintptr_t)kind, (uintptr_t)r->tiling.shape, (uintptr_t)fused_ok(r),
                              (uintptr_t)r->ctrl_stride, (uintptr_t)r->src_inst_stride,
                              (uintptr_t)r->chunk, (uintptr_t)r->ck.nsteps, (uintptr_t)r->ck.K,
                              (uintptr_t)r->resident, (uintptr_t)r->dual};
  const DevBuf* bufs[] = {&r->q, &r->lam, &r->rhs, &r->b1, &r->b2, &r->b1a, &r->b2a, &r->t, &r->w,
                          &r->G, &r->T1, &r->S, &r->qs, &r->aforce, &r->control, &r->grad,
                          &r->acc_J, &r->acc_S, &r->partial, &r->src_amp, &r->snaps, &r->window,
                          &r->qy, &r->ly, &r->l0, &r->jpart, &r->spart, &r->ssum, &r->wts,
                          &r->window2, &r->b2r, &r->qsr, &r->qT, &r->d_tile_mask, &r->wall};
  for (const DevBuf* b : bufs) k.push_back((uintptr_t)b->p);
  k.push_back((uintptr_t)r->q_alt.p);
  k.push_back((uintptr_t)r->qs_alt.p);
  k.push_back((uintptr_t)r->q_alt2.p);
  k.push_back((uintptr_t)r->qs_alt2.p);
  k.push_back((uintptr_t)r->ncache);
  k.push_back((uintptr_t)r->cache_first);
  for (int c = 0; c < r->ncache; ++c) k.push_back((uintptr_t)r->wcache[c].p);
  k.push_back((uintptr_t)r->step_mode);
  k.push_back((uintptr_t)r->nscache);
  k.push_back((uintptr_t)r->sstore.p);
  k.push_back((uintptr_t)r->scache.p);
  k.push_back((uintptr_t)r->sreuse);
  k.push_back((uintptr_t)r->sbuf.p);
  for (uintptr_t e : extra) k.push_back(e);
  return k;
}

template <class F>
int run_device_work(mgd_region* r, bool use_graph, int slot, const std::vector<uintptr_t>& key,
                    F&& enqueue) {
  if (!use_graph) return enqueue();
  r->timing = false;                     // the per-launch timer is for eager runs only
  GraphSlot& gs = r->graphs[slot];
  if (!gs.exec || gs.key != key) {
    if (gs.exec) {
      (void)hipGraphExecDestroy(gs.exec);
      gs.exec = nullptr;
    }
    MGD_HIP(hipStreamBeginCapture(r->stream, hipStreamCaptureModeThreadLocal));
    const int rc = enqueue();
    hipGraph_t g = nullptr;
    const hipError_t e = hipStreamEndCapture(r->stream, &g);
    if (rc) {
      if (g) (void)hipGraphDestroy(g);
      return rc;
    }
    MGD_HIP(e);
    const hipError_t ei = hipGraphInstantiate(&gs.exec, g, nullptr, nullptr, 0);
    (void)hipGraphDestroy(g);
    MGD_HIP(ei);
    gs.key = key;
  }
  MGD_HIP(hipGraphLaunch(gs.exec, r->stream));
  return MGD_OK;
}

int finish_timed(mgd_region* r) {
  MGD_HIP(hipEventRecord(r->ev1, r->stream));
  MGD_HIP(hipGetLastError());
  MGD_HIP(hipStreamSynchronize(r->stream));
  float ms = 0.f;
  MGD_HIP(hipEventElapsedTime(&ms, r->ev0, r->ev1));
  r->last_ms[0] = ms;
  return MGD_OK;
}

int copy_adjoint_outputs(mgd_region* r, double* gradient, double* sensitivity) {
  const int na = n_act(r), nst = 4 * r->ck.nsteps;
  if (sensitivity) MGD_HIP(hipMemcpy(sensitivity, r->acc_S.p, sizeof(double) * r->ninst, hipMemcpyDeviceToHost));
  if (gradient && na > 0)
    MGD_HIP(hipMemcpy(gradient, r->grad.p, sizeof(double) * r->ninst * nst * na, hipMemcpyDeviceToHost));
  return MGD_OK;
}

// ---------------------------------------------------------------- region groups
// Stage tables of a group run: for stage index sidx and member m, entry sidx * n + m.  The
// stage arguments of each member are those of fused_forward_pass / fused_adjoint_pass
// (stage states ping-pong in the forward pass and every window is recomputed), so every
// launch computes what the member's own launch would.
int group_tables(mgd_group* G, double dt0, const double* dts, int nsteps, int K, bool want_grad) {
  (void)dt0;
  const int n = (int)G->m.size(), nst = 4 * nsteps;
  std::vector<FwdStage> fwd((size_t)nst * n), rec((size_t)nst * n);
  std::vector<AdjStage> adj((size_t)nst * n);
  std::vector<GroupMember> mem(n);
  int block0 = 0;
  for (int k = 0; k < n; ++k) {
    mgd_region* r = G->m[k];
    const double dt = dts[k];
    const long M = (long)r->ninst * M_of(r);
    const int na = n_act(r);
    mem[k].g = r->geo;
    mem[k].P = dpatches(r);
    mem[k].T = r->tiling;
    mem[k].block0 = block0;
    block0 += r->tiling.nblk * r->ninst;
    // forward pass: levels in q, stage states in b1 / qy (fused_forward_step)
    const double* qin[4] = {r->q.p, r->b1.p, r->qy.p, r->b1.p};
    double* qo[4] = {r->b1.p, r->qy.p, r->b1.p, r->q.p};
    // resident members: every stage state in `wall` (state 4 t + i - 1 the input of stage
    // i of step t), written by the forward stages and read by the adjoint stages; no
    // recompute
    double* wbuf[2] = {r->window.p, r->window2.p};
    for (int t = 0; t < nsteps; ++t) {
      const int w = t / K, t0w = w * K, t1w = std::min(nsteps, t0w + K);
      double* W = r->resident ? r->wall.p : wbuf[w % 2];
      const long base = r->resident ? 4L * t : 4L * (t - t0w);
      double* outs[4];
      for (int i = 0; i < 4; ++i) outs[i] = W + (base + 1 + i) * M;
      if (t == t1w - 1 && !r->resident) outs[3] = r->qsr.p;
      for (int i = 1; i <= 4; ++i) {
        const int sidx = 4 * t + i - 1;
        const double* in_w = i == 1 ? W + base * M : outs[i - 2];
        fwd[(size_t)sidx * n + k] =
            r->resident ? fwd_args(r, in_w, W + base * M, r->b2.p, outs[i - 1], i, dt, sidx, nst, true)
                        : fwd_args(r, qin[i - 1], r->q.p, r->b2.p, qo[i - 1], i, dt, sidx, nst, true);
        rec[(size_t)sidx * n + k] =
            fwd_args(r, in_w, W + base * M, r->b2r.p, r->resident ? r->qsr.p : outs[i - 1], i, dt,
                     sidx, nst, false);
        AdjStage a;
        std::memset(&a, 0, sizeof a);
        double* L0 = r->lam.p;
        double* LX = r->ly.p;
        double* LY = r->l0.p;
        a.lam = i == 4 ? L0 : (i == 3 ? LX : (i == 2 ? LY : LX));
        a.lout = i == 4 ? LX : (i == 3 ? LY : (i == 2 ? LX : L0));
        a.l0 = L0;
        a.q = W + (base + (i - 1)) * M;
        a.b2 = r->b2.p;
        a.stage = i;
        a.dt = dt;
        a.forcing_factor = i == 4 ? 2.0 : (i == 2 ? 0.5 : 1.0);
        a.zero_forcing = t == 0 && i == 1;
        a.ip_act = r->ip_act;
        if (na > 0) {
          if (want_grad) {
            a.grad = r->grad.p + (long)sidx * na;
            a.grad_stride = (long)nst * na;
          }
          a.spartial = r->spart.p;
          a.npts = na;
          a.nslots = r->chunk;
          a.slot = sidx % r->chunk;
        }
        adj[(size_t)sidx * n + k] = a;
      }
    }
  }
  G->nblocks = block0;
  if (upload_ints(G->dmem, mem.data(), mem.size() * sizeof(GroupMember)) ||
      upload_ints(G->dfwd, fwd.data(), fwd.size() * sizeof(FwdStage)) ||
      upload_ints(G->drec, rec.data(), rec.size() * sizeof(FwdStage)) ||
      upload_ints(G->dadj, adj.data(), adj.size() * sizeof(AdjStage)))
    return MGD_ERR_HIP;
  return MGD_OK;
}

const FwdStage* stage_table(const DevBuf& b, int sidx, int n) {
  return reinterpret_cast<const FwdStage*>(b.p) + (size_t)sidx * n;
}

// fused_forward_pass + the enqueue_forward_ck bookkeeping of every member, one launch
// per stage for all members, on the group stream (every member's stream is the group's
// for the duration of the run).
int group_forward_pass(mgd_group* G, int nsteps, int K) {
  const int n = (int)G->m.size(), nst = 4 * nsteps;
  const GroupMember* mem = reinterpret_cast<const GroupMember*>(G->dmem.p);
  for (mgd_region* r : G->m) {
    launch_fill(r->ninst, 0.0, r->acc_J.p, G->st);
    launch_copy((long)r->ninst * M_of(r), r->q.p, r->resident ? r->wall.p : r->snaps.p, G->st);
  }
  for (int s = 0; s < nsteps; ++s) {
    for (int i = 1; i <= 4; ++i) {
      const int sidx = 4 * s + i - 1;
      if (int rc = launch_fwd_stage2d_group(kGroupShape, mem, n, G->nblocks,
                                            stage_table(G->dfwd, sidx, n), G->st))
        return rc;
      for (mgd_region* r : G->m) reduce_chunk(r, r->jpart, n_cost(r), nst, sidx, false);
    }
    if ((s + 1) % K == 0)
      for (mgd_region* r : G->m) {
        if (r->resident) continue;
        const long M = (long)r->ninst * M_of(r);
        launch_copy(M, r->q.p, r->snaps.p + (long)((s + 1) / K) * M, G->st);
      }
  }
  for (mgd_region* r : G->m) {
    const long M = (long)r->ninst * M_of(r);
    if (r->resident) launch_copy(M, r->wall.p + 4L * nsteps * M, r->q.p, G->st);   // final state
    if (n_cost(r) > 0)
      launch_stage_accumulate(r->ssum.p, nst, r->wts.p, r->acc_J.p, r->ninst, 0, G->st);
    launch_copy((long)r->ninst * M_of(r), r->q.p, r->qT.p, G->st);
  }
  return MGD_OK;
}

// fused_adjoint_pass of every member (terminal condition zero), one launch per stage for
// all members: the members' resident stage states read in place, or their windows
// recomputed on the group's second stream.
int group_adjoint_pass(mgd_group* G, const double* dts, int nsteps, int K) {
  const int n = (int)G->m.size(), nst = 4 * nsteps;
  const GroupMember* mem = reinterpret_cast<const GroupMember*>(G->dmem.p);
  const bool resident = G->m[0]->resident;   // the same for every member
  const int nw = resident ? 1 : (nsteps + K - 1) / K;
  if (resident) K = nsteps;                  // one window: the resident states
  for (mgd_region* r : G->m) launch_fill(r->ninst, 0.0, r->acc_S.p, G->st);
  MGD_HIP(hipEventRecord(G->ev_fwd, G->st));
  MGD_HIP(hipStreamWaitEvent(G->st_rec, G->ev_fwd, 0));
  auto recompute = [&](int w) -> int {
    const int t0w = w * K, t1w = std::min(nsteps, t0w + K);
    for (mgd_region* r : G->m) {
      const long M = (long)r->ninst * M_of(r);
      launch_copy(M, r->snaps.p + (long)w * M, (w % 2 ? r->window2.p : r->window.p), G->st_rec);
    }
    for (int t = t0w; t < t1w; ++t)
      for (int i = 1; i <= 4; ++i)
        if (int rc = launch_fwd_stage2d_group(kGroupShape, mem, n, G->nblocks,
                                              stage_table(G->drec, 4 * t + i - 1, n), G->st_rec))
          return rc;
    return MGD_OK;
  };
  if (!resident) {
    if (int rc = recompute(nw - 1)) return rc;
    MGD_HIP(hipEventRecord(G->ev_ready[(nw - 1) % 2], G->st_rec));
  }
  for (int k = 0; k < n; ++k) {
    mgd_region* r = G->m[k];
    const long M = (long)r->ninst * M_of(r);
    const int nc = n_cost(r);
    launch_fill(M, 0.0, r->lam.p, G->st);
    if (nc > 0) {
      launch_adj_forcing(r->geo, dpatch(r, r->ip_cost), nc, r->qT.p, r->aforce.p, 0, G->st);
      launch_adj_terminal(r->geo, dpatch(r, r->ip_cost), nc, r->aforce.p, -dts[k] / 6.0,
                          r->lam.p, G->st);
    }
  }
  const AdjStage* adj = reinterpret_cast<const AdjStage*>(G->dadj.p);
  for (int w = nw - 1; w >= 0; --w) {
    if (w >= 1) {
      if (w + 1 < nw) MGD_HIP(hipStreamWaitEvent(G->st_rec, G->ev_free[(w + 1) % 2], 0));
      if (int rc = recompute(w - 1)) return rc;
      MGD_HIP(hipEventRecord(G->ev_ready[(w - 1) % 2], G->st_rec));
    }
    if (!resident) MGD_HIP(hipStreamWaitEvent(G->st, G->ev_ready[w % 2], 0));
    const int t0w = w * K, t1w = std::min(nsteps, t0w + K);
    for (int t = t1w - 1; t >= t0w; --t)
      for (int i = 4; i >= 1; --i) {
        const int sidx = 4 * t + i - 1;
        if (int rc = launch_adj_stage2d_group(kGroupShape, mem, n, G->nblocks,
                                              adj + (size_t)sidx * n, G->st))
          return rc;
        for (mgd_region* r : G->m) reduce_chunk(r, r->spart, n_act(r), nst, sidx, true);
      }
    if (!resident) MGD_HIP(hipEventRecord(G->ev_free[w % 2], G->st));
  }
  for (mgd_region* r : G->m)
    if (n_act(r) > 0)
      launch_stage_accumulate(r->ssum.p, nst, r->wts.p, r->acc_S.p, r->ninst, 1, G->st);
  return MGD_OK;
}

// normalizeKolmogorovForcing (src/RegionImpl.f90:910-999, called at :1477): the forces of
// all Kolmogorov patches summed per grid point, and the number of patches covering each
// point, are integrated with the norm over the union of those patches
// (computeQuadratureOnPatches, src/PatchFactoryImpl.f90:376-444); every patch's force then
// loses globalForcing / globalVolume, so the forcing has zero quadrature mean.
void normalize_kolmogorov(mgd_region* r) {
  const long N = r->N;
  const int nd = r->nd;
  bool any = false;
  for (const auto& P : r->patches) any = any || P->pd.type == MGD_PATCH_KOLMOGOROV;
  if (!any) return;
  std::vector<double> force(N, 0.0), count(N, 0.0), mask(N, 0.0);
  for (const auto& P : r->patches) {
    if (P->pd.type != MGD_PATCH_KOLMOGOROV) continue;
    for (int k = 0; k < P->np; ++k) {
      const long p = patch_point(r, *P, k);
      force[p] = force[p] + P->strength[k];
      count[p] = count[p] + 1.0;
      mask[p] = 1.0;
    }
  }
  double volume = 0.0, forcing = 0.0;
  for (long p = 0; p < N; ++p) {
    int ijk[3];
    unflat_host(r, p, ijk);
    double det = r->h[0][ijk[0]];
    if (nd > 1) det = det * r->h[1][ijk[1]];
    if (nd > 2) det = det * r->h[2][ijk[2]];
    double w = 1.0;
    for (int d = 0; d < nd; ++d) w = r->hw[d][ijk[d]] * w;
    const double norm = w * det;
    volume += mask[p] * norm * count[p];
    forcing += mask[p] * norm * force[p];
  }
  const double mean = forcing / volume;
  for (auto& P : r->patches) {
    if (P->pd.type != MGD_PATCH_KOLMOGOROV) continue;
    for (int k = 0; k < P->np; ++k) P->strength[k] = P->strength[k] - mean;
  }
}

// Scratch of the allocation-free entry point (mgd_inner_product_noalloc): the per-block
// partial sums of the inner product live in this fixed device array instead of a buffer
// the region allocates.  Each device holds its own copy at its own address, looked up
// once per device with hipGetSymbolAddress when a region is created there; host code
// uses only that address, never the array's name.  Calls that overlap on one device
// share the array and race.
constexpr long kScratchDoubles = 1L << 18;   // 2 MiB: 2^18 blocks of 256 points
__device__ __align__(256) double scratch[kScratchDoubles];
constexpr int kMaxDevices = 64;
double* g_scratch_ptr[kMaxDevices] = {};
std::mutex g_scratch_mu;

// Looks up the scratch array's device address on the current device (once per device).
int scratch_setup(int* device) {
  int dev = 0;
  MGD_HIP(hipGetDevice(&dev));
  if (dev < 0 || dev >= kMaxDevices) { set_error("device index beyond the scratch table"); return MGD_ERR_ARG; }
  *device = dev;
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (g_scratch_ptr[dev]) return MGD_OK;
  void* ptr = nullptr;
  hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
  if (e != hipSuccess || !ptr) {
    set_error(std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
    return MGD_ERR_HIP;
  }
  g_scratch_ptr[dev] = static_cast<double*>(ptr);
  return MGD_OK;
}

}  // namespace

extern "C" {

const char* mgd_last_error(void) { return g_err.c_str(); }

int mgd_device_count(int* count) {
  int c = 0;
  hipError_t e = hipGetDeviceCount(&c);
  *count = e == hipSuccess ? c : 0;
  return MGD_OK;
}

int mgd_region_create(const mgd_region_desc* d, const double* coords, mgd_region** out) {
  if (!d || !coords || !out) { set_error("null argument"); return MGD_ERR_ARG; }
  if (d->nd < 1 || d->nd > 3) { set_error("nd must be 1, 2 or 3"); return MGD_ERR_ARG; }
  if (d->scheme != 0) { set_error("only the SBP 4-8 scheme is built"); return MGD_ERR_UNSUPPORTED; }
  auto r = std::make_unique<mgd_region>();
  r->desc = *d;
  r->nd = d->nd;
  r->nv = d->nd + 2;
  r->ninst = std::max(1, d->n_instances);
  r->dual = r->ninst == 1;
  for (int i = d->nd; i < 3; ++i) { r->desc.n[i] = 1; r->desc.periodicity[i] = MGD_PERIODIC_NONE; }
  for (int i = 0; i < d->nd; ++i)
    if (r->desc.periodicity[i] != MGD_PERIODIC_NONE && r->desc.n[i] < 9) {
      set_error("periodic directions need at least 9 points for the SBP 4-8 interior stencil");
      return MGD_ERR_ARG;
    }
  for (int i = 0; i < d->nd; ++i)
    if (r->desc.n[i] < 2 * 12 + 1 && r->desc.periodicity[i] == MGD_PERIODIC_NONE) {
      set_error("non-periodic directions need at least 25 points for the SBP 4-8 adjoint closure");
      return MGD_ERR_ARG;
    }
  r->N = (long)r->desc.n[0] * r->desc.n[1] * r->desc.n[2];
  if (r->N >= (1L << 31)) { set_error("a region holds fewer than 2^31 grid points"); return MGD_ERR_ARG; }
  const long N = r->N;
  r->coords.assign(coords, coords + (size_t)N * r->nd);
  // rectilinear check: coordinate d depends on index d only
  for (int d0 = 0; d0 < r->nd; ++d0) {
    const int len = r->desc.n[d0];
    const long st = stride_of(r.get(), d0);
    r->line[d0].resize(len);
    for (int i = 0; i < len; ++i) r->line[d0][i] = r->coords[d0 * N + i * st];
    for (long p = 0; p < N; ++p) {
      int ijk[3];
      unflat_host(r.get(), p, ijk);
      if (r->coords[d0 * N + p] != r->line[d0][ijk[d0]]) {
        set_error("grid is not rectilinear (curvilinear_domain grids are not built)");
        return MGD_ERR_UNSUPPORTED;
      }
    }
  }
  // operators (Grid%setupSpatialDiscretization, src/GridImpl.f90:487-619)
  Op first, diss, dissT, adj, second;
  op_setup(&first, d->scheme, OP_FIRST);
  op_setup(&second, d->scheme, OP_SECOND);
  op_setup(&diss, d->scheme, d->composite_dissipation ? OP_COMPOSITE : OP_DISS);
  op_setup(&dissT, d->scheme, OP_DISS_T);
  op_adjoint(first, &adj);
  r->first_host = first;
  r->ops[SLOT_FIRST] = first;
  r->ops[SLOT_ADJ] = adj;
  r->ops[SLOT_DISS] = diss;
  r->ops[SLOT_DISS_T] = dissT;
  r->ops[SLOT_SECOND] = second;   // grid%secondDerivative (src/GridImpl.f90:30-31, 557-567)
  // Grid%update for a rectilinear grid: per-direction coordinate derivative and norm weight
  for (int d0 = 0; d0 < 3; ++d0) {
    const int len = r->desc.n[d0];
    if (d0 >= r->nd) {
      r->h[d0].assign(1, 1.0);
      r->hw[d0].assign(1, 1.0);
      r->line[d0].assign(1, 0.0);
      continue;
    }
    coordinate_derivative(first, len, r->desc.periodicity[d0], r->desc.periodic_length[d0],
                          r->line[d0], r->h[d0]);
    r->hw[d0].assign(len, 1.0);
    if (r->desc.periodicity[d0] == MGD_PERIODIC_NONE)
      for (int i = 0; i < first.bd; ++i) {
        r->hw[d0][i] = first.norm[i];
        r->hw[d0][len - 1 - i] = first.norm[i];
      }
  }
  for (int d0 = 0; d0 < 3; ++d0) {
    if (upload(r->d_h[d0], r->h[d0]) || upload(r->d_hw[d0], r->hw[d0]) ||
        upload(r->d_x[d0], r->line[d0]))
      return MGD_ERR_HIP;
  }
  const size_t ops_bytes = sizeof(r->ops);   // DevBuf counts doubles
  if (r->d_ops.alloc(ops_bytes / sizeof(double) + 1)) return MGD_ERR_HIP;
  MGD_HIP(hipMemcpy(r->d_ops.p, r->ops, sizeof(r->ops), hipMemcpyHostToDevice));
  Geo& g = r->geo;
  std::memset(&g, 0, sizeof g);
  g.nd = r->nd; g.nv = r->nv; g.ninst = r->ninst; g.N = N;
  for (int d0 = 0; d0 < 3; ++d0) {
    g.n[d0] = r->desc.n[d0];
    g.dir[d0].len = r->desc.n[d0];
    g.dir[d0].periodic = r->desc.periodicity[d0] != MGD_PERIODIC_NONE;
    g.dir[d0].off = r->desc.periodicity[d0] == MGD_PERIODIC_OVERLAP ? 1 : 0;
    g.dir[d0].stride = stride_of(r.get(), d0);
    g.h[d0] = r->d_h[d0].p;
    g.hw[d0] = r->d_hw[d0].p;
    g.x[d0] = r->d_x[d0].p;
  }
  g.ops = reinterpret_cast<const Op*>(r->d_ops.p);
  g.gamma = d->gamma;
  if (d->viscous) {
    g.re_inv = d->reynolds > 0.0 ? 1.0 / d->reynolds : 0.0;
    g.pr_inv = d->prandtl > 0.0 ? 1.0 / d->prandtl : 0.0;
  }
  g.power_law = d->power_law;
  g.bulk_ratio = d->bulk_ratio;
  g.diss_amount = d->dissipation_amount;
  g.viscous = d->viscous;
  g.dissipation = d->dissipation;
  g.composite = d->composite_dissipation;
  g.zchunk = g.zchunk_adj = 64;
  g.mtile = g.mtile_adj = 16;
  if (int rc = scratch_setup(&r->device)) return rc;
  MGD_HIP(hipStreamCreate(&r->stream));
  MGD_HIP(hipEventCreate(&r->ev0));
  MGD_HIP(hipEventCreate(&r->ev1));
  r->target.assign((size_t)r->nv * N, 0.0);
  r->ctrl_moll.assign(N, 0.0);
  r->tgt_moll.assign(N, 0.0);
  if (int rc = ensure_state(r.get())) return rc;
  *out = r.release();
  return MGD_OK;
}

int mgd_region_destroy(mgd_region* r) {
  if (!r) return MGD_OK;
  if (r->stream) (void)hipStreamSynchronize(r->stream);
  for (GraphSlot& gs : r->graphs)
    if (gs.exec) (void)hipGraphExecDestroy(gs.exec);
  if (r->ev0) (void)hipEventDestroy(r->ev0);
  if (r->ev1) (void)hipEventDestroy(r->ev1);
  for (hipEvent_t e : r->tev) (void)hipEventDestroy(e);
  if (r->stream) (void)hipStreamDestroy(r->stream);
  delete r;
  return MGD_OK;
}

int mgd_region_add_patch(mgd_region* r, const mgd_patch_desc* pd) {
  if (!r || !pd) { set_error("null argument"); return MGD_ERR_ARG; }
  if ((int)r->patches.size() >= kMaxPatches) { set_error("too many patches"); return MGD_ERR_ARG; }
  if (pd->type < MGD_PATCH_FAR_FIELD || pd->type > MGD_PATCH_ISOTHERMAL_WALL) {
    set_error("unsupported patch type");
    return MGD_ERR_UNSUPPORTED;
  }
  auto P = std::make_unique<HostPatch>();
  P->pd = *pd;
  P->np = 1;
  for (int i = 0; i < 3; ++i) {
    int a = pd->extent[2 * i], b = pd->extent[2 * i + 1];
    if (i >= r->nd) { a = 1; b = 1; }
    if (a < 0) a += r->desc.n[i] + 1;
    if (b < 0) b += r->desc.n[i] + 1;
    if (a < 1 || b > r->desc.n[i] || b < a) { set_error("invalid patch extent"); return MGD_ERR_ARG; }
    P->lo[i] = a - 1;
    P->hi[i] = b - 1;
    P->ls[i] = b - a + 1;
    P->np *= P->ls[i];
  }
  if (pd->type == MGD_PATCH_FAR_FIELD && (pd->normal_dir == 0 || std::abs(pd->normal_dir) > r->nd)) {
    set_error("far-field patch needs a normal direction");
    return MGD_ERR_ARG;
  }
  if (pd->type == MGD_PATCH_COST_TARGET) {
    if (r->ip_cost >= 0) { set_error("one COST_TARGET patch is supported"); return MGD_ERR_UNSUPPORTED; }
    r->ip_cost = (int)r->patches.size();
  }
  if (pd->type == MGD_PATCH_ACTUATOR) {
    if (r->ip_act >= 0) { set_error("one ACTUATOR patch is supported"); return MGD_ERR_UNSUPPORTED; }
    r->ip_act = (int)r->patches.size();
  }
  if (pd->type == MGD_PATCH_FAR_FIELD) r->has_farfield = 1;
  r->patches.push_back(std::move(P));
  r->finalized = false;
  return MGD_OK;
}

int mgd_region_set_field(mgd_region* r, int which, const double* v) {
  if (!r || !v) { set_error("null argument"); return MGD_ERR_ARG; }
  const long N = r->N;
  switch (which) {
    case MGD_FIELD_TARGET_STATE: r->target.assign(v, v + (size_t)r->nv * N); break;
    case MGD_FIELD_CONTROL_MOLLIFIER: r->ctrl_moll.assign(v, v + N); break;
    case MGD_FIELD_TARGET_MOLLIFIER: r->tgt_moll.assign(v, v + N); break;
    case MGD_FIELD_MEAN_PRESSURE: r->mean_p.assign(v, v + N); break;
    default: set_error("unknown field"); return MGD_ERR_ARG;
  }
  r->finalized = false;
  return MGD_OK;
}

int mgd_region_add_source(mgd_region* r, const double* loc, double amplitude, double frequency,
                          double radius, double phase) {
  if (!r || !loc || radius <= 0.0) { set_error("bad acoustic source"); return MGD_ERR_ARG; }
  Source s;
  const double pi = 4.0 * std::atan(1.0);
  for (int i = 0; i < 3; ++i) s.loc[i] = i < r->nd ? loc[i] : 0.0;
  s.amplitude = amplitude;
  s.omega = 2.0 * pi * frequency;
  s.gfac = 9.0 / (2.0 * (radius * radius));
  s.phase = phase;
  r->sources.push_back(s);
  r->finalized = false;
  return MGD_OK;
}

int mgd_region_finalize(mgd_region* r) {
  if (!r) { set_error("null region"); return MGD_ERR_ARG; }
  const int nd = r->nd, nv = r->nv;
  const long N = r->N;
  const double g = r->desc.gamma;
  std::memset(&r->hpatches, 0, sizeof r->hpatches);
  r->hpatches.count = (int)r->patches.size();
  for (size_t ip = 0; ip < r->patches.size(); ++ip) {
    HostPatch& P = *r->patches[ip];
    P.strength.clear();
    P.data.clear();
    double sig_inv = 0.0, sig_visc = 0.0, twall = 0.0;
    switch (P.pd.type) {
      case MGD_PATCH_IMPENETRABLE_WALL:
      case MGD_PATCH_ISOTHERMAL_WALL: {
        // setupImpenetrableWall (src/ImpenetrableWallImpl.f90:3-40) and setupIsothermalWall
        // (src/IsothermalWallImpl.f90:3-80): sign(amount, normal) / H00; viscous amount
        // / H00 / Re; the wall temperature is the patch option
        const double h0 = r->first_host.norm[0];
        sig_inv = std::copysign(P.pd.p[0], (double)P.pd.normal_dir) / h0;
        if (P.pd.type == MGD_PATCH_ISOTHERMAL_WALL && r->desc.viscous)
          sig_visc = P.pd.p[1] / h0 * r->geo.re_inv;
        twall = P.pd.p[2];
        break;
      }
      case MGD_PATCH_FAR_FIELD: {
        const int dir = std::abs(P.pd.normal_dir) - 1;
        const double h0 = r->first_host.norm[0];
        sig_inv = std::copysign(P.pd.p[0], (double)P.pd.normal_dir) / h0;
        sig_visc = r->desc.viscous ? std::copysign(P.pd.p[1], (double)P.pd.normal_dir) / h0 : 0.0;
        P.data.assign((size_t)P.np * nv * nv + (size_t)P.np * nv, 0.0);
        for (int k = 0; k < P.np; ++k) {
          const long p = patch_point(r, P, k);
          int ijk[3];
          unflat_host(r, p, ijk);
          double qt[5], m[3] = {0, 0, 0};
          for (int l = 0; l < nv; ++l) qt[l] = r->target[l * N + p];
          m[dir] = metric_host(r, dir, ijk);
          incoming_jacobian(nd, qt, m, g, P.pd.normal_dir, &P.data[(size_t)k * nv * nv]);
          for (int l = 0; l < nv; ++l) P.data[(size_t)P.np * nv * nv + (size_t)k * nv + l] = qt[l];
        }
        if (r->desc.viscous) target_viscous_fluxes(r, P);
        break;
      }
      case MGD_PATCH_SPONGE: {
        // computeSpongeStrengths (src/PatchFactoryImpl.f90:161-374) on a rectilinear line
        const int dir = std::abs(P.pd.normal_dir) - 1;
        if (dir < 0 || dir >= nd) { set_error("sponge needs a normal direction"); return MGD_ERR_ARG; }
        const std::vector<double>& hd = r->h[dir];
        std::vector<double> cli(hd.size());
        for (size_t i = 0; i < hd.size(); ++i) cli[i] = std::sqrt(hd[i] * hd[i]);
        const int mn = P.lo[dir], mx = P.hi[dir];
        P.strength.resize(P.np);
        P.data.resize((size_t)P.np * nv);
        for (int k = 0; k < P.np; ++k) {
          const long p = patch_point(r, P, k);
          int ijk[3];
          unflat_host(r, p, ijk);
          const int i = ijk[dir];
          double a = 0.0, b = 0.0;
          if (P.pd.normal_dir > 0) {
            for (int c = mn; c <= i - 1; ++c) a += cli[c];
            for (int c = mn; c <= mx - 1; ++c) b += cli[c];
          } else {
            for (int c = i + 1; c <= mx; ++c) a += cli[c];
            for (int c = mn + 1; c <= mx; ++c) b += cli[c];
          }
          const double s = a / b;
          P.strength[k] = P.pd.p[0] * std::pow(1.0 - s, P.pd.p[1]);
          for (int l = 0; l < nv; ++l) P.data[(size_t)k * nv + l] = r->target[l * N + p];
        }
        break;
      }
      case MGD_PATCH_ACTUATOR: {
        P.strength.resize(P.np);
        for (int k = 0; k < P.np; ++k) P.strength[k] = r->ctrl_moll[patch_point(r, P, k)];
        break;
      }
      case MGD_PATCH_COST_TARGET: {
        // mean pressure: the mean_pressure_file field when given, else from the target
        // state (src/AcousticNoiseImpl.f90:51-71)
        P.strength.resize(P.np);
        P.data.resize(P.np);
        for (int k = 0; k < P.np; ++k) {
          const long p = patch_point(r, P, k);
          P.strength[k] = r->tgt_moll[p];
          if (!r->mean_p.empty()) {
            P.data[k] = r->mean_p[p];
            continue;
          }
          double s = 0.0;
          for (int i = 0; i < nd; ++i) s += r->target[(i + 1) * N + p] * r->target[(i + 1) * N + p];
          P.data[k] = (g - 1.0) * (r->target[(nd + 1) * N + p] - 0.5 * s / r->target[p]);
        }
        break;
      }
      case MGD_PATCH_KOLMOGOROV: {
        const double pi = 4.0 * std::atan(1.0);
        const int wn = P.pd.p[0] > 0 ? (int)P.pd.p[0] : 0;
        P.strength.resize(P.np);
        for (int k = 0; k < P.np; ++k) {
          const long p = patch_point(r, P, k);
          const double y = nd > 1 ? r->coords[1 * N + p] : 0.0;
          P.strength[k] = P.pd.p[1] * std::sin(2.0 * pi * wn * y);
        }
        break;
      }
    }
    P.sig_inv = sig_inv;
    P.sig_visc = sig_visc;
    P.twall = twall;
  }
  normalize_kolmogorov(r);
  for (size_t ip = 0; ip < r->patches.size(); ++ip) {
    HostPatch& P = *r->patches[ip];
    if (upload(P.d_strength, P.strength) || upload(P.d_data, P.data) || upload(P.d_tvflux, P.tvflux))
      return MGD_ERR_HIP;
    DPatch& D = r->hpatches.p[ip];
    D.type = P.pd.type;
    D.normal_dir = P.pd.normal_dir;
    for (int i = 0; i < 3; ++i) { D.lo[i] = P.lo[i]; D.hi[i] = P.hi[i]; D.ls[i] = P.ls[i]; }
    D.np = P.np;
    D.sigma_inv = P.sig_inv;
    D.sigma_visc = P.sig_visc;
    D.twall = P.twall;
    D.strength = P.d_strength.p;
    D.aplus = P.d_data.p;
    D.tvflux = P.d_tvflux.p;
    D.ffv = nullptr;
    if (P.pd.type == MGD_PATCH_FAR_FIELD && r->desc.viscous) {
      if (P.d_ffv.alloc((size_t)r->ninst * nv * P.np)) return MGD_ERR_HIP;
      D.ffv = P.d_ffv.p;
    }
  }
  // the forward finishing pass: far field present; at most one Kolmogorov forcing over
  // the whole domain and no other patch
  r->fin.has_farfield = r->has_farfield;
  r->fin.simple = 1;
  r->fin.kolmogorov = nullptr;
  for (int ip = 0; ip < r->hpatches.count; ++ip) {
    const DPatch& D = r->hpatches.p[ip];
    bool whole = D.type == MGD_PATCH_KOLMOGOROV && !r->fin.kolmogorov;
    for (int d = 0; d < 3 && whole; ++d)
      whole = D.lo[d] == 0 && D.hi[d] == (d < nd ? r->desc.n[d] - 1 : 0);
    if (whole) r->fin.kolmogorov = D.strength;
    else r->fin.simple = 0;
  }
  if (r->d_patches.alloc(sizeof(Patches) / sizeof(double) + 1)) return MGD_ERR_HIP;
  MGD_HIP(hipMemcpy(r->d_patches.p, &r->hpatches, sizeof(Patches), hipMemcpyHostToDevice));
  if (r->d_sources.alloc((sizeof(Source) * std::max<size_t>(1, r->sources.size())) / sizeof(double) + 1))
    return MGD_ERR_HIP;
  if (!r->sources.empty())
    MGD_HIP(hipMemcpy(r->d_sources.p, r->sources.data(), sizeof(Source) * r->sources.size(),
                      hipMemcpyHostToDevice));
  if (r->acc_J.alloc(r->ninst) || r->acc_S.alloc(r->ninst)) return MGD_ERR_HIP;
  const int npr = std::max(reduce_blocks(std::max(n_cost(r), 1)), reduce_blocks(std::max(n_act(r), 1)));
  if (r->partial.alloc((size_t)npr * r->ninst)) return MGD_ERR_HIP;
  if (n_cost(r) > 0 && r->aforce.alloc((size_t)r->ninst * nv * n_cost(r))) return MGD_ERR_HIP;
  if (r->nd == 2 && !r->desc.viscous)
    if (int rc = setup_fused(r)) return rc;
  r->finalized = true;
  return MGD_OK;
}

int mgd_region_set_solution_limits(mgd_region* r, int enable, const double* range) {
  if (!r || (enable && !range)) { set_error("bad argument"); return MGD_ERR_ARG; }
  if (enable && !(range[0] < range[1] && range[2] < range[3])) {
    set_error("solution limits need minimum < maximum for density and temperature");
    return MGD_ERR_ARG;
  }
  r->lim_on = enable != 0;
  if (r->lim_on) {
    for (int k = 0; k < 4; ++k) r->lim[k] = range[k];
    if (r->d_flag.alloc(r->ninst)) return MGD_ERR_HIP;
    launch_fill(r->ninst, 0.0, r->d_flag.p, r->stream);
    MGD_HIP(hipStreamSynchronize(r->stream));
  }
  return MGD_OK;
}

// computeTimeStepSize (mode 0) / computeCfl (mode 1) of the current conserved state
int wave_speed_reduce(mgd_region* r, int mode, double param, double* out) {
  if (!r || !r->finalized || !out) { set_error("bad argument"); return MGD_ERR_ARG; }
  if (!(param > 0.0)) { set_error(mode == 0 ? "cfl must be positive" : "dt must be positive"); return MGD_ERR_ARG; }
  const long nb = (r->N + 255) / 256;
  if (r->ipart.alloc((size_t)nb * r->ninst) || r->iout.alloc(r->ninst)) return MGD_ERR_HIP;
  launch_wave_speed(r->geo, r->q.p, mode, param, r->ipart.p, r->iout.p, r->stream);
  MGD_HIP(hipGetLastError());
  MGD_HIP(hipMemcpyAsync(out, r->iout.p, sizeof(double) * r->ninst, hipMemcpyDeviceToHost, r->stream));
  MGD_HIP(hipStreamSynchronize(r->stream));
  return MGD_OK;
}

int mgd_time_step_size(mgd_region* r, double cfl, double* dt) { return wave_speed_reduce(r, 0, cfl, dt); }

int mgd_cfl(mgd_region* r, double dt, double* cfl) { return wave_speed_reduce(r, 1, dt, cfl); }

int mgd_check_solution_limits(mgd_region* r, int* crashed) {
  if (!r || !r->finalized || !crashed) { set_error("bad argument"); return MGD_ERR_ARG; }
  if (!r->lim_on) {
    for (int i = 0; i < r->ninst; ++i) crashed[i] = 0;
    return MGD_OK;
  }
  launch_fill(r->ninst, 0.0, r->d_flag.p, r->stream);
  launch_check_limits(r->geo, r->q.p, r->lim, r->d_flag.p, r->stream);
  MGD_HIP(hipGetLastError());
  std::vector<double> f(r->ninst, 0.0);
  MGD_HIP(hipMemcpyAsync(f.data(), r->d_flag.p, sizeof(double) * r->ninst, hipMemcpyDeviceToHost,
                         r->stream));
  MGD_HIP(hipStreamSynchronize(r->stream));
  int any = 0;
  for (int i = 0; i < r->ninst; ++i) {
    crashed[i] = f[i] != 0.0;
    any |= crashed[i];
  }
  return any ? MGD_SOLUTION_LIMIT : MGD_OK;
}

int mgd_region_set_time_offsets(mgd_region* r, const double* offsets) {
  if (!r) { set_error("null region"); return MGD_ERR_ARG; }
  if (offsets) r->toff.assign(offsets, offsets + r->ninst);
  else r->toff.clear();
  return MGD_OK;
}

int mgd_region_get_march_shape(mgd_region* r, int* shape) {
  if (!r || !shape) { set_error("null argument"); return MGD_ERR_ARG; }
  shape[0] = r->geo.mtile;
  shape[1] = r->geo.zchunk;
  shape[2] = r->geo.mtile_adj;
  shape[3] = r->geo.zchunk_adj;
  return MGD_OK;
}

int mgd_region_get_checkpoint_plan(mgd_region* r, int* plan) {
  if (!r || !plan) { set_error("null argument"); return MGD_ERR_ARG; }
  if (!r->ck.valid && r->ck.nsteps == 0) { set_error("no checkpointed forward run prepared"); return MGD_ERR_STATE; }
  if (r->step_mode) {
    plan[0] = 1;
    plan[1] = r->nscache;
    plan[2] = r->scache_first;
  } else if (fused_ok(r)) {
    plan[0] = 2;
    plan[1] = r->resident ? 4 * r->ck.nsteps + 1 : 0;
    plan[2] = 0;
  } else {
    plan[0] = 0;
    plan[1] = r->ncache;
    plan[2] = r->cache_first;
  }
  return MGD_OK;
}

int mgd_region_get_tile_shape(mgd_region* r, int* shape) {
  if (!r || !shape) { set_error("null argument"); return MGD_ERR_ARG; }
  *shape = fused_ok(r) ? r->tiling.shape : -1;
  return MGD_OK;
}

int mgd_region_set_option(mgd_region* r, int option, int value) {
  if (!r) { set_error("null region"); return MGD_ERR_ARG; }
  switch (option) {
    case MGD_OPT_FORCE_PLAIN: r->force_plain = value != 0; return MGD_OK;
    case MGD_OPT_RESIDENT_STATES:
      if (value < -1 || value > 1) {
        set_error("resident states: -1 (where they fit), 0 (off) or 1 (on)");
        return MGD_ERR_ARG;
      }
      r->opt_resident = value;
      return MGD_OK;
    case MGD_OPT_WINDOW_CACHE: r->opt_cache = value < 0 ? -1 : value; return MGD_OK;
    case MGD_OPT_STEP_STORE:
      if (value < -1 || value > 1) {
        set_error("step store: -1 (where it fits), 0 (off) or 1 (on)");
        return MGD_ERR_ARG;
      }
      r->opt_step = value;
      return MGD_OK;
    case MGD_OPT_Z_CHUNK:
      if (value != 0 && value != 32 && value != 64 && value != 128) {
        set_error("z chunk: 0 (autotune), 32, 64 or 128");
        return MGD_ERR_ARG;
      }
      r->opt_zchunk = value;
      r->geo.zchunk = r->geo.zchunk_adj = value == 0 ? 64 : value;
      r->zc_tuned = value != 0 && r->opt_mtile != 0;
      return MGD_OK;
    case MGD_OPT_MARCH_TILE:
      if (value != 0 && value != 16 && value != 32) {
        set_error("march tile: 0 (autotune), 16 or 32");
        return MGD_ERR_ARG;
      }
      if (value == 32 && r->nd == 3 && r->geo.n[0] % 32 != 0) {
        set_error("march tile 32 needs an x size that is a multiple of 32");
        return MGD_ERR_ARG;
      }
      r->opt_mtile = value;
      r->geo.mtile = r->geo.mtile_adj = value == 0 ? 16 : value;
      r->zc_tuned = value != 0 && r->opt_zchunk != 0;
      return MGD_OK;
    case MGD_OPT_TILE_SHAPE: {
      if (value < -1 || value >= kNumTileShapes) { set_error("tile shape must be -1 .. 5"); return MGD_ERR_ARG; }
      r->tile_shape = value;
      r->tuned = false;
      if (r->finalized && r->nd == 2) {
        if (value >= 0 && r->tiling_ok) {
          const int rc = install_tiling(r, value);
          if (rc == MGD_ERR_UNSUPPORTED) { set_error("tile shape does not fit the grid"); return rc; }
          return rc;
        }
      }
      return MGD_OK;
    }
  }
  set_error("unknown option");
  return MGD_ERR_ARG;
}

int mgd_state_put(mgd_region* r, int which, int inst, const double* host) {
  if (!r || !host || inst < 0 || inst >= r->ninst) { set_error("bad argument"); return MGD_ERR_ARG; }
  DevBuf* b = state_buf(r, which, true);
  if (!b) { set_error("unknown state array"); return MGD_ERR_ARG; }
  const size_t M = (size_t)M_of(r);
  MGD_HIP(hipMemcpyAsync(b->p + inst * M, host, M * sizeof(double), hipMemcpyHostToDevice, r->stream));
  MGD_HIP(hipStreamSynchronize(r->stream));
  return MGD_OK;
}

int mgd_state_get(mgd_region* r, int which, int inst, double* host) {
  if (!r || !host || inst < 0 || inst >= r->ninst) { set_error("bad argument"); return MGD_ERR_ARG; }
  DevBuf* b = state_buf(r, which, false);
  if (!b) { set_error("unknown or empty state array"); return MGD_ERR_ARG; }
  const size_t M = (size_t)M_of(r);
  MGD_HIP(hipMemcpyAsync(host, b->p + inst * M, M * sizeof(double), hipMemcpyDeviceToHost, r->stream));
  MGD_HIP(hipStreamSynchronize(r->stream));
  return MGD_OK;
}

int mgd_state_copy(mgd_region* r, int src, int dst) {
  if (!r) { set_error("null region"); return MGD_ERR_ARG; }
  DevBuf* a = state_buf(r, src, false);
  DevBuf* b = state_buf(r, dst, true);
  if (!a || !b) { set_error("unknown or empty state array"); return MGD_ERR_ARG; }
  launch_copy((long)r->ninst * M_of(r), a->p, b->p, r->stream);
  MGD_HIP(hipGetLastError());
  return MGD_OK;
}

int mgd_state_lincomb(mgd_region* r, int dst, int dst_inst, int a, int a_inst, int b, int b_inst,
                      double s, int c, int c_inst) {
  if (!r || !r->finalized) { set_error("region not finalized"); return MGD_ERR_STATE; }
  auto ok = [&](int i) { return i >= 0 && i < r->ninst; };
  if (!ok(dst_inst) || !ok(a_inst) || !ok(b_inst) || (c >= 0 && !ok(c_inst))) {
    set_error("instance out of range");
    return MGD_ERR_ARG;
  }
  DevBuf* A = state_buf(r, a, false);
  DevBuf* B = state_buf(r, b, false);
  DevBuf* Cb = c >= 0 ? state_buf(r, c, false) : nullptr;
  DevBuf* D = state_buf(r, dst, true);
  if (!A || !B || !D || (c >= 0 && !Cb)) { set_error("unknown or empty state array"); return MGD_ERR_ARG; }
  const long M = M_of(r);
  launch_lincomb(M, A->p + a_inst * M, B->p + b_inst * M, s, Cb ? Cb->p + c_inst * M : nullptr,
                 D->p + dst_inst * M, r->stream);
  MGD_HIP(hipGetLastError());
  MGD_HIP(hipStreamSynchronize(r->stream));
  return MGD_OK;
}

int mgd_inner_product(mgd_region* r, int a, int b, double* out) {
  if (!r || !r->finalized || !out) { set_error("bad argument"); return MGD_ERR_ARG; }
  DevBuf* A = state_buf(r, a, false);
  DevBuf* B = state_buf(r, b, false);
  if (!A || !B) { set_error("unknown or empty state array"); return MGD_ERR_ARG; }
  const long nb = (r->N + 255) / 256;
  if (r->ipart.alloc((size_t)nb * r->ninst) || r->iout.alloc(r->ninst)) return MGD_ERR_HIP;
  launch_inner_product(r->geo, A->p, B->p, r->ipart.p, r->iout.p, r->stream);
  MGD_HIP(hipGetLastError());
  MGD_HIP(hipMemcpyAsync(out, r->iout.p, sizeof(double) * r->ninst, hipMemcpyDeviceToHost, r->stream));
  MGD_HIP(hipStreamSynchronize(r->stream));
  return MGD_OK;
}

int mgd_inner_product_noalloc(mgd_region* r, int a, int b, double* out_device) {
  if (!r || !r->finalized || !out_device) { set_error("bad argument"); return MGD_ERR_ARG; }
  DevBuf* A = state_buf(r, a, false);
  DevBuf* B = state_buf(r, b, false);
  if (!A || !B) { set_error("unknown or empty state array"); return MGD_ERR_ARG; }
  const long nb = (r->N + 255) / 256;
  if (nb * r->ninst > kScratchDoubles) {
    set_error("mgd_inner_product_noalloc: " + std::to_string(nb * r->ninst) +
              " partial sums exceed the fixed scratch of " + std::to_string(kScratchDoubles) +
              " (at most 2^26 grid points over all instances)");
    return MGD_ERR_ARG;
  }
  double* part = g_scratch_ptr[r->device];
  if (!part) { set_error("scratch not set up on the region's device"); return MGD_ERR_STATE; }
  launch_inner_product(r->geo, A->p, B->p, part, out_device, r->stream);
  MGD_HIP(hipGetLastError());
  return MGD_OK;
}

int mgd_terminal_cost(mgd_region* r, int kind, double* J) {
  if (!r || !r->finalized || !J) { set_error("bad argument"); return MGD_ERR_ARG; }
  if (kind != MGD_TERMINAL_KINETIC_ENERGY) { set_error("unknown terminal cost"); return MGD_ERR_ARG; }
  const long nb = (r->N + 255) / 256;
  if (r->ipart.alloc((size_t)nb * r->ninst) || r->iout.alloc(r->ninst)) return MGD_ERR_HIP;
  launch_kinetic_energy(r->geo, r->q.p, r->lam.p, r->ipart.p, r->iout.p, r->stream);
  MGD_HIP(hipGetLastError());
  MGD_HIP(hipMemcpyAsync(J, r->iout.p, sizeof(double) * r->ninst, hipMemcpyDeviceToHost, r->stream));
  MGD_HIP(hipStreamSynchronize(r->stream));
  return MGD_OK;
}

int mgd_compute_rhs(mgd_region* r, int mode, double time, const double* control, double factor) {
  if (!r || !r->finalized) { set_error("region not finalized"); return MGD_ERR_STATE; }
  const int ns = (int)r->sources.size();
  std::vector<double> amp;
  source_table(r, std::vector<double>{time}, amp);
  if (upload(r->src_amp, amp)) return MGD_ERR_HIP;
  if (fused_ok(r) && (mode == MGD_FORWARD || mode == MGD_ADJOINT)) {
    if (mode == MGD_FORWARD) {
      const int na = n_act(r);
      if (control && na > 0) {
        std::vector<double> c(control, control + (size_t)r->ninst * na);
        if (upload(r->control, c)) return MGD_ERR_HIP;
        r->ctrl_stride = na;
      } else if (r->control.p) {
        r->control.alloc(0);
      }
      FwdStage a = fwd_args(r, r->q.p, r->q.p, r->b2.p, r->b1.p, 1, 0.0, 0, 1, false);
      a.rhs_out = r->rhs.p;
      launch_fwd_stage2d(r->geo, dpatches(r), r->tiling, a, r->stream);
    } else {
      AdjStage a;
      std::memset(&a, 0, sizeof a);
      a.lam = r->lam.p; a.l0 = r->lam.p; a.q = r->q.p; a.b2 = r->b2.p; a.lout = r->b1.p;
      a.rhs_out = r->rhs.p; a.stage = 4; a.dt = 0.0; a.forcing_factor = factor;
      a.zero_forcing = r->aforce_zero; a.ip_act = r->ip_act;
      launch_adj_stage2d(r->geo, dpatches(r), r->tiling, a, r->stream);
    }
    MGD_HIP(hipGetLastError());
    MGD_HIP(hipStreamSynchronize(r->stream));
    return MGD_OK;
  }
  if (mode == MGD_FORWARD) {
    const int na = n_act(r);
    const double* ctrl = nullptr;
    if (control && na > 0) {
      std::vector<double> c(control, control + (size_t)r->ninst * na);
      if (upload(r->control, c)) return MGD_ERR_HIP;
      ctrl = r->control.p;
    }
    launch_rhs_forward(r->geo, dpatches(r), r->q.p, r->rhs.p, r->w.p, r->G.p, r->t.p, r->T1.p, ctrl, na,
                       reinterpret_cast<const Source*>(r->d_sources.p), ns, r->src_amp.p,
                       r->src_inst_stride, r->fin, r->stream);
  } else if (mode == MGD_ADJOINT) {
    if (int rc = ensure_adjoint(r)) return rc;
    const int nc = n_cost(r);
    launch_rhs_adjoint(r->geo, dpatches(r), r->q.p, r->lam.p, r->rhs.p, r->w.p, r->G.p, r->t.p,
                       r->T1.p, r->S.p, nc > 0 ? r->aforce.p : nullptr, (long)r->nv * nc, factor,
                       r->fin, r->stream);
  } else {
    set_error("mode must be MGD_FORWARD or MGD_ADJOINT");
    return MGD_ERR_ARG;
  }
  MGD_HIP(hipGetLastError());
  MGD_HIP(hipStreamSynchronize(r->stream));
  return MGD_OK;
}

int mgd_update_adjoint_forcing(mgd_region* r, int zero) {
  if (!r || !r->finalized) { set_error("region not finalized"); return MGD_ERR_STATE; }
  r->aforce_zero = zero != 0;
  if (r->ip_cost < 0) return MGD_OK;
  launch_adj_forcing(r->geo, dpatch(r, r->ip_cost), n_cost(r), r->q.p, r->aforce.p, zero, r->stream);
  MGD_HIP(hipGetLastError());
  MGD_HIP(hipStreamSynchronize(r->stream));
  return MGD_OK;
}

int mgd_functional(mgd_region* r, double* out) {
  if (!r || !r->finalized || !out) { set_error("bad argument"); return MGD_ERR_ARG; }
  if (r->ip_cost < 0) {
    for (int i = 0; i < r->ninst; ++i) out[i] = 0.0;
    return MGD_OK;
  }
  launch_patch_reduce(r->geo, dpatch(r, r->ip_cost), n_cost(r), 0, r->q.p, r->partial.p, 0.0,
                      nullptr, r->acc_J.p, r->stream);
  MGD_HIP(hipGetLastError());
  MGD_HIP(hipMemcpyAsync(out, r->acc_J.p, sizeof(double) * r->ninst, hipMemcpyDeviceToHost, r->stream));
  MGD_HIP(hipStreamSynchronize(r->stream));
  return MGD_OK;
}

int mgd_run_forward(mgd_region* r, double t0, double dt, int nsteps, const double* control, double* J,
                    int use_graph) {
  if (!r || !r->finalized) { set_error("region not finalized"); return MGD_ERR_STATE; }
  if (nsteps < 0 || !(dt > 0.0)) { set_error("bad time stepping"); return MGD_ERR_ARG; }
  if (int rc = maybe_tune(r)) return rc;
  const int na = n_act(r), nst = 4 * nsteps;
  if (control && na > 0) {
    std::vector<double> c(control, control + (size_t)r->ninst * nst * na);
    if (upload(r->control, c)) return MGD_ERR_HIP;
    r->ctrl_stride = (long)nst * na;
  } else if (r->control.p) {
    r->control.alloc(0);
  }
  std::vector<double> amp;
  run_source_table(r, t0, dt, std::max(nsteps, 1), amp);
  if (upload(r->src_amp, amp)) return MGD_ERR_HIP;
  if (fused_ok(r))
    if (int rc = prepare_quadrature(r, std::max(nst, 1), dt)) return rc;
  uint64_t dtb;
  std::memcpy(&dtb, &dt, sizeof dtb);
  const auto key = graph_key(r, 0, {(uintptr_t)nsteps, (uintptr_t)dtb});
  MGD_HIP(hipEventRecord(r->ev0, r->stream));
  if (int rc = run_device_work(r, use_graph != 0, 0, key, [&] { return enqueue_forward_run(r, dt, nsteps); }))
    return rc;
  if (int rc = finish_timed(r)) return rc;
  if (J) MGD_HIP(hipMemcpy(J, r->acc_J.p, sizeof(double) * r->ninst, hipMemcpyDeviceToHost));
  return limits_outcome(r, J, nullptr);
}

int mgd_run_adjoint(mgd_region* r, double t0, double dt, int nsteps, int save_interval,
                    const double* control, double* gradient, double* sensitivity, double* J,
                    int use_graph) {
  if (!r || !r->finalized) { set_error("region not finalized"); return MGD_ERR_STATE; }
  if (nsteps <= 0 || save_interval <= 0 || !(dt > 0.0)) { set_error("bad time stepping"); return MGD_ERR_ARG; }
  if (int rc = prepare_forward_ck(r, t0, dt, nsteps, save_interval, control)) return rc;
  r->ck.valid = true;
  if (int rc = prepare_adjoint_ck(r)) return rc;
  uint64_t dtb;
  std::memcpy(&dtb, &dt, sizeof dtb);
  const bool want_grad = gradient != nullptr;
  const auto key = graph_key(r, 1, {(uintptr_t)dtb, (uintptr_t)want_grad});
  MGD_HIP(hipEventRecord(r->ev0, r->stream));
  if (int rc = run_device_work(r, use_graph != 0, 1, key, [&] {
        if (int e = enqueue_forward_ck(r)) return e;
        return enqueue_adjoint_ck(r, false, want_grad);
      }))
    return rc;
  if (int rc = finish_timed(r)) return rc;
  if (J) MGD_HIP(hipMemcpy(J, r->acc_J.p, sizeof(double) * r->ninst, hipMemcpyDeviceToHost));
  if (int rc = copy_adjoint_outputs(r, gradient, sensitivity)) return rc;
  return limits_outcome(r, J, sensitivity);
}

int mgd_run_forward_checkpointed(mgd_region* r, double t0, double dt, int nsteps, int save_interval,
                                 const double* control, double* J) {
  if (!r || !r->finalized) { set_error("region not finalized"); return MGD_ERR_STATE; }
  if (nsteps <= 0 || save_interval <= 0 || !(dt > 0.0)) { set_error("bad time stepping"); return MGD_ERR_ARG; }
  if (int rc = prepare_forward_ck(r, t0, dt, nsteps, save_interval, control)) return rc;
  MGD_HIP(hipEventRecord(r->ev0, r->stream));
  if (int rc = enqueue_forward_ck(r)) return rc;
  if (int rc = finish_timed(r)) return rc;
  r->ck.valid = true;
  if (J) MGD_HIP(hipMemcpy(J, r->acc_J.p, sizeof(double) * r->ninst, hipMemcpyDeviceToHost));
  return limits_outcome(r, J, nullptr);
}

int mgd_run_adjoint_checkpointed(mgd_region* r, int terminal, double* gradient, double* sensitivity) {
  if (!r || !r->finalized) { set_error("region not finalized"); return MGD_ERR_STATE; }
  if (terminal != 0 && terminal != 1) { set_error("terminal must be 0 or 1"); return MGD_ERR_ARG; }
  if (int rc = prepare_adjoint_ck(r)) return rc;
  MGD_HIP(hipEventRecord(r->ev0, r->stream));
  if (int rc = enqueue_adjoint_ck(r, terminal == 1, gradient != nullptr)) return rc;
  if (int rc = finish_timed(r)) return rc;
  return copy_adjoint_outputs(r, gradient, sensitivity);
}

// ---------------------------------------------------------------- stage-level entry points

int mgd_rk4_substep(mgd_region* r, int mode, int stage, double time, double dt, const double* control) {
  if (!r || !r->finalized) { set_error("region not finalized"); return MGD_ERR_STATE; }
  if (stage < 1 || stage > 4 || !(dt > 0.0)) { set_error("stage must be 1..4 and dt > 0"); return MGD_ERR_ARG; }
  if (mode != MGD_FORWARD && mode != MGD_ADJOINT) {
    set_error("mode must be MGD_FORWARD or MGD_ADJOINT");
    return MGD_ERR_ARG;
  }
  if (int rc = ensure_adjoint(r)) return rc;
  if (int rc = maybe_tune(r)) return rc;
  const long M = (long)r->ninst * M_of(r);
  // the previous stage may still read the per-stage inputs about to be replaced
  MGD_HIP(hipStreamSynchronize(r->stream));
  if (mode == MGD_FORWARD) {
    // this stage's inputs: the source amplitudes at `time`, the actuator forcing
    std::vector<double> amp;
    source_table(r, std::vector<double>{time}, amp);
    if (upload(r->src_amp, amp)) return MGD_ERR_HIP;
    const int na = n_act(r);
    if (control && na > 0) {
      std::vector<double> c(control, control + (size_t)r->ninst * na);
      if (upload(r->control, c)) return MGD_ERR_HIP;
      r->ctrl_stride = na;
    } else if (r->control.p) {
      r->control.alloc(0);
    }
    if (fused_ok(r)) {
      // the stage kernel writes out of place: the level (buffer1) is kept in b1
      if (r->q_alt.alloc(M)) return MGD_ERR_HIP;
      if (stage == 1) launch_copy(M, r->q.p, r->b1.p, r->stream);
      FwdStage a = fwd_args(r, r->q.p, r->b1.p, r->b2.p, r->q_alt.p, stage, dt, 0, 1, false);
      launch_fwd_stage2d(r->geo, dpatches(r), r->tiling, a, r->stream);
      std::swap(r->q.p, r->q_alt.p);
      std::swap(r->q.n, r->q_alt.n);
    } else {
      forward_stage(r, r->q, r->q_alt, r->b1.p, r->b2.p, 0, stage, dt, nullptr);
    }
  } else {
    const double factor = stage == 4 ? 2.0 : (stage == 2 ? 0.5 : 1.0);
    if (fused_ok(r)) {
      if (r->lam_alt.alloc(M)) return MGD_ERR_HIP;
      if (stage == 4) launch_copy(M, r->lam.p, r->b1a.p, r->stream);
      AdjStage a;
      std::memset(&a, 0, sizeof a);
      a.lam = r->lam.p; a.l0 = r->b1a.p; a.q = r->q.p; a.b2 = r->b2a.p; a.lout = r->lam_alt.p;
      a.stage = stage; a.dt = dt; a.forcing_factor = factor; a.zero_forcing = r->aforce_zero;
      a.ip_act = r->ip_act;
      launch_adj_stage2d(r->geo, dpatches(r), r->tiling, a, r->stream);
      std::swap(r->lam.p, r->lam_alt.p);
      std::swap(r->lam.n, r->lam_alt.n);
    } else {
      const int nc = n_cost(r);
      const RkStage rk{stage, dt, r->lam.p, r->b1a.p, r->b2a.p};
      if (!launch_rhs_adjoint(r->geo, dpatches(r), r->q.p, r->lam.p, r->rhs.p, r->w.p, r->G.p, r->t.p,
                              r->T1.p, r->S.p, nc > 0 ? r->aforce.p : nullptr, (long)r->nv * nc,
                              factor, r->fin, r->stream, &rk))
        launch_rk_adjoint(M, stage, dt, r->lam.p, r->b1a.p, r->b2a.p, r->rhs.p, r->stream);
    }
  }
  MGD_HIP(hipGetLastError());
  return MGD_OK;
}

int mgd_checkpoint_migrate(mgd_region* r, int timestep, int stage) {
  if (!r || !r->finalized) { set_error("region not finalized"); return MGD_ERR_STATE; }
  if (!r->ck.valid) { set_error("no checkpointed forward run to migrate in"); return MGD_ERR_STATE; }
  const int nsteps = r->ck.nsteps, K = r->ck.K, nst = 4 * nsteps;
  const double dt = r->ck.dt;
  if (timestep < 0 || timestep > nsteps || stage < 1 || stage > 4 || (timestep == 0 && stage != 4)) {
    set_error("(timestep, stage) outside the checkpointed run");
    return MGD_ERR_ARG;
  }
  const long M = (long)r->ninst * M_of(r);
  const int ts = timestep, st = stage;
  if (fused_ok(r) && r->resident) {
    // state 4 (ts - 1) + st of the resident stage states (4 ts: the state at step ts)
    launch_copy(M, r->wall.p + (4L * (ts - 1) + st) * M, r->q.p, r->stream);
    MGD_HIP(hipGetLastError());
    return MGD_OK;
  }
  if (r->step_mode) {
    // the state at the start of step ts, or the output of stage st of step ts - 1
    const double* src;
    if (st == 4) {
      src = ts < nsteps ? r->sstore.p + (long)ts * M : r->qT.p;
    } else {
      const int t = ts - 1;
      const bool rec = t < r->scache_first && r->rec_step != t;
      if (rec) {
        // the recompute needs the run's actuator forcing and source amplitudes
        MGD_HIP(hipStreamSynchronize(r->stream));
        if (int rc = upload_run_inputs(r)) return rc;
        if (fused_ok(r) && r->b2r.alloc(M)) return MGD_ERR_HIP;
      }
      src = step_stage_state(r, t, st + 1, dt, rec);
    }
    launch_copy(M, src, r->q.p, r->stream);
    MGD_HIP(hipGetLastError());
    return MGD_OK;
  }
  int loaded = r->mig_loaded;
  const bool need = loaded == -1 || ts < loaded || ts > loaded + K || (ts == loaded && st < 4) ||
                    (ts == loaded + K && st == 4);
  if (need) {
    int lt;
    if (ts % K == 0 && st == 4) lt = ts;
    else if (ts % K == 0) lt = ts - K;
    else lt = ts - ts % K;
    loaded = lt;
    // the recompute needs the run's actuator forcing and source amplitudes
    MGD_HIP(hipStreamSynchronize(r->stream));
    if (int rc = upload_run_inputs(r)) return rc;
    double* W = r->window.p;
    launch_copy(M, r->snaps.p + (long)(lt / K) * M, W, r->stream);
    const int t1 = std::min(nsteps, lt + K);
    if (fused_ok(r)) {
      if (r->b2r.alloc(M) || r->qsr.alloc(M)) return MGD_ERR_HIP;
      for (int t = lt; t < t1; ++t) {
        const long base = 4L * (t - lt);
        double* outs[4];
        for (int k = 0; k < 4; ++k) outs[k] = W + (base + 1 + k) * M;
        if (t == lt + K - 1) outs[3] = r->qsr.p;   // the state at the next snapshot: not needed
        fused_forward_step(r, W + base * M, nullptr, nullptr, r->b2r.p, t, dt, nst, false, outs);
      }
    } else {
      launch_copy(M, W, r->qs.p, r->stream);
      int idx = 1;
      for (int s2 = lt + 1; s2 <= t1; ++s2)
        for (int st2 = 1; st2 <= 4; ++st2) {
          const bool last = s2 == lt + K && st2 == 4;
          forward_stage(r, r->qs, r->qs_alt, r->b1.p, r->b2.p, 4 * (s2 - 1) + st2 - 1, st2, dt,
                        last ? nullptr : W + (long)idx * M);
          if (last) break;
          ++idx;
        }
    }
    r->mig_loaded = loaded;
  }
  const int widx = (ts - 1 - loaded) * 4 + st;
  launch_copy(M, r->window.p + (long)widx * M, r->q.p, r->stream);
  MGD_HIP(hipGetLastError());
  return MGD_OK;
}

int mgd_control_gradient(mgd_region* r, double* gradient, double* sensitivity) {
  if (!r || !r->finalized) { set_error("region not finalized"); return MGD_ERR_STATE; }
  const int na = n_act(r);
  if (na == 0) {
    if (sensitivity)
      for (int i = 0; i < r->ninst; ++i) sensitivity[i] = 0.0;
    return MGD_OK;
  }
  if (r->iout.alloc(r->ninst)) return MGD_ERR_HIP;
  if (gradient) {
    if (r->grad.n < (size_t)r->ninst * na && r->grad.alloc((size_t)r->ninst * na)) return MGD_ERR_HIP;
    launch_gradient(r->geo, dpatch(r, r->ip_act), na, r->lam.p, r->grad.p, (long)na, r->stream);
    MGD_HIP(hipMemcpyAsync(gradient, r->grad.p, sizeof(double) * r->ninst * na,
                           hipMemcpyDeviceToHost, r->stream));
  }
  if (sensitivity) {
    launch_patch_reduce(r->geo, dpatch(r, r->ip_act), na, 1, r->lam.p, r->partial.p, 1.0, nullptr,
                        r->iout.p, r->stream);
    MGD_HIP(hipMemcpyAsync(sensitivity, r->iout.p, sizeof(double) * r->ninst,
                           hipMemcpyDeviceToHost, r->stream));
  }
  MGD_HIP(hipGetLastError());
  MGD_HIP(hipStreamSynchronize(r->stream));
  return MGD_OK;
}

int mgd_adjoint_initial_condition(mgd_region* r, int keep) {
  if (!r || !r->finalized) { set_error("region not finalized"); return MGD_ERR_STATE; }
  if (!r->ck.valid) { set_error("no checkpointed forward run to start the adjoint from"); return MGD_ERR_STATE; }
  if (int rc = ensure_adjoint(r)) return rc;
  const long M = (long)r->ninst * M_of(r);
  const int nc = n_cost(r);
  if (!keep) launch_fill(M, 0.0, r->lam.p, r->stream);
  if (nc > 0) {
    launch_adj_forcing(r->geo, dpatch(r, r->ip_cost), nc, r->qT.p, r->aforce.p, 0, r->stream);
    launch_adj_terminal(r->geo, dpatch(r, r->ip_cost), nc, r->aforce.p, -r->ck.dt / 6.0, r->lam.p,
                        r->stream);
  }
  MGD_HIP(hipGetLastError());
  return MGD_OK;
}

int mgd_stencil_apply(mgd_region* r, int op, int dir, double* host, int ncomp) {
  if (!r || !host || ncomp <= 0 || dir < 0 || dir >= r->nd || op < 0 || op >= kNumSlots) {
    set_error("bad argument");
    return MGD_ERR_ARG;
  }
  DevBuf a, b;
  const size_t n = (size_t)ncomp * r->N;
  if (a.alloc(n) || b.alloc(n)) return MGD_ERR_HIP;
  MGD_HIP(hipMemcpy(a.p, host, n * sizeof(double), hipMemcpyHostToDevice));
  launch_apply(r->geo, op, dir, ncomp, a.p, b.p, r->stream);
  MGD_HIP(hipGetLastError());
  MGD_HIP(hipStreamSynchronize(r->stream));
  MGD_HIP(hipMemcpy(host, b.p, n * sizeof(double), hipMemcpyDeviceToHost));
  return MGD_OK;
}

int mgd_region_get_grid(mgd_region* r, int which, double* out) {
  if (!r || !out) { set_error("bad argument"); return MGD_ERR_ARG; }
  const int nd = r->nd;
  const long N = r->N;
  for (long p = 0; p < N; ++p) {
    int ijk[3];
    unflat_host(r, p, ijk);
    double det = r->h[0][ijk[0]];
    if (nd > 1) det = det * r->h[1][ijk[1]];
    if (nd > 2) det = det * r->h[2][ijk[2]];
    switch (which) {
      case 0:
        for (int a = 0; a < nd; ++a)
          for (int b = 0; b < nd; ++b) out[(a * nd + b) * N + p] = a == b ? metric_host(r, a, ijk) : 0.0;
        break;
      case 1: out[p] = 1.0 / det; break;
      case 2: {
        double s = 1.0;
        for (int d = 0; d < nd; ++d) s = r->hw[d][ijk[d]] * s;
        out[p] = s * det;
        break;
      }
      case 3:
        for (int d = 0; d < nd; ++d) out[d * N + p] = std::fabs(metric_host(r, d, ijk));
        break;
      default: set_error("unknown grid array"); return MGD_ERR_ARG;
    }
  }
  return MGD_OK;
}

int mgd_last_timings(mgd_region* r, double* out, int* launches) {
  if (!r) return MGD_ERR_ARG;
  for (int i = 0; i < 3; ++i) {
    if (out) out[i] = r->last_ms[i];
    if (launches) launches[i] = r->last_launch[i];
  }
  return MGD_OK;
}

int mgd_kernel_timer(mgd_region* r, int which, int on) {
  if (!r) { set_error("null region"); return MGD_ERR_ARG; }
  if (which != MGD_KERNEL_FWD_MARCH3D) { set_error("only the z-march kernel is timed per launch"); return MGD_ERR_ARG; }
  r->timing = on != 0;
  if (r->timing) r->tcount = r->tcopy = 0;
  return MGD_OK;
}

int mgd_kernel_timer_read(mgd_region* r, long* launches, long* with_copy, double* ms) {
  if (!r || !launches || !with_copy || !ms) { set_error("null argument"); return MGD_ERR_ARG; }
  *launches = r->tcount;
  *with_copy = r->tcopy;
  *ms = 0.0;
  if (r->tcount == 0) return MGD_OK;
  MGD_HIP(hipEventSynchronize(r->tev[2 * (size_t)r->tcount - 1]));
  double sum = 0.0;
  for (long k = 0; k < r->tcount; ++k) {
    float e = 0.0f;
    MGD_HIP(hipEventElapsedTime(&e, r->tev[2 * k], r->tev[2 * k + 1]));
    sum += e;
  }
  *ms = sum;
  return MGD_OK;
}

int mgd_time_kernel(mgd_region* r, int which, int reps, double* ms_per_launch, double* bytes) {
  if (!r || !r->finalized || reps <= 0) { set_error("bad argument"); return MGD_ERR_ARG; }
  if (int rc = maybe_tune(r)) return rc;
  const long Mi = M_of(r), M = (long)r->ninst * Mi;
  const double pts = (double)r->ninst * r->N;
  const double w = 8.0 * r->nv;   // bytes of one state point
  if (r->qy.alloc(M) || r->ly.alloc(M) || r->l0.alloc(M) || r->qs.alloc(M)) return MGD_ERR_HIP;
  std::vector<double> amp;
  source_table(r, std::vector<double>{0.0, 0.5e-3, 0.5e-3, 1e-3}, amp);
  if (upload(r->src_amp, amp)) return MGD_ERR_HIP;
  if (r->control.p) r->control.alloc(0);
  const double dt = 1e-3;
  int per_rep = 1;
  auto launch = [&]() {
    switch (which) {
      case MGD_KERNEL_FWD_STAGE:
        // one RK4 step: stages 1-4 on the level state qs with stage states in qy / ly
        fused_forward_step(r, r->qs.p, r->qy.p, r->ly.p, r->b2.p, 0, dt, 1, false, nullptr);
        break;
      case MGD_KERNEL_ADJ_STAGE: {
        // one adjoint RK4 step: stages 4-1 on the level lam with stage adjoints in ly / l0,
        // the forward stage state held at q
        double* L0 = r->lam.p;
        double* LX = r->ly.p;
        double* LY = r->l0.p;
        for (int i = 4; i >= 1; --i) {
          AdjStage a;
          std::memset(&a, 0, sizeof a);
          a.lam = i == 4 ? L0 : (i == 3 ? LX : (i == 2 ? LY : LX));
          a.lout = i == 4 ? LX : (i == 3 ? LY : (i == 2 ? LX : L0));
          a.l0 = L0; a.q = r->q.p; a.b2 = r->b2.p;
          a.stage = i; a.dt = dt; a.forcing_factor = i == 4 ? 2.0 : (i == 2 ? 0.5 : 1.0);
          a.ip_act = r->ip_act;
          launch_adj_stage2d(r->geo, dpatches(r), r->tiling, a, r->stream);
        }
        break;
      }
      case MGD_KERNEL_COPY:
        launch_copy(M, r->q.p, r->qs.p, r->stream);
        break;
      case MGD_KERNEL_RHS_FORWARD:
        launch_rhs_forward(r->geo, dpatches(r), r->q.p, r->rhs.p, r->w.p, r->G.p, r->t.p, r->T1.p, nullptr,
                           n_act(r), reinterpret_cast<const Source*>(r->d_sources.p),
                           (int)r->sources.size(), r->src_amp.p, r->src_inst_stride,
                           r->fin, r->stream);
        break;
      case MGD_KERNEL_RHS_ADJOINT:
        launch_rhs_adjoint(r->geo, dpatches(r), r->q.p, r->lam.p, r->rhs.p, r->w.p, r->G.p, r->t.p,
                           r->T1.p, r->S.p, nullptr, 0, 1.0, r->fin, r->stream);
        break;
      case MGD_KERNEL_FWD_MARCH3D: {
        // one RK4 step of march launches as the runs issue them (qs -> qy -> qs_alt2 -> qy
        // -> qs, stages 2 and 3 reading the state at the start of the step from qs), from
        // the stress slots formed once (the timing does not depend on their values)
        double* const in[4] = {r->qs.p, r->qy.p, r->qs_alt2.p, r->qy.p};
        double* const out[4] = {r->qy.p, r->qs_alt2.p, r->qy.p, r->qs.p};
        for (int i = 1; i <= 4; ++i) {
          RkStage rk{i, dt, in[i - 1], i == 1 ? nullptr : r->qs.p, r->b2.p, nullptr, out[i - 1]};
          launch_fwd3d_march(r->geo, rk.y, r->G.p, r->fin.kolmogorov, r->rhs.p, &rk, r->stream);
        }
        break;
      }
    }
  };
  if (which == MGD_KERNEL_FWD_STAGE || which == MGD_KERNEL_ADJ_STAGE) {
    if (!fused_ok(r)) { set_error("no fused stage kernel for this region"); return MGD_ERR_UNSUPPORTED; }
    per_rep = 4;
    // well-defined operands: every buffer holds the current state
    launch_copy(M, r->q.p, r->qs.p, r->stream);
    launch_copy(M, r->q.p, r->qy.p, r->stream);
    launch_copy(M, r->q.p, r->b2.p, r->stream);
    launch_copy(M, r->q.p, r->ly.p, r->stream);
    launch_copy(M, r->q.p, r->l0.p, r->stream);
    launch_copy(M, r->q.p, r->lam.p, r->stream);
  } else if (which == MGD_KERNEL_FWD_MARCH3D) {
    if (!fwd3d_march_path(r->geo, r->fin, (int)r->sources.size())) {
      set_error("the 3-D z-march forward kernel does not serve this region");
      return MGD_ERR_UNSUPPORTED;
    }
    per_rep = 4;
    if (r->qs_alt2.alloc(M)) return MGD_ERR_HIP;
    launch_copy(M, r->q.p, r->qs.p, r->stream);
    launch_copy(M, r->q.p, r->b2.p, r->stream);
    if (r->geo.viscous) launch_stress3d(r->geo, r->q.p, r->G.p, r->stream);
  } else if (which == MGD_KERNEL_RHS_FORWARD || which == MGD_KERNEL_RHS_ADJOINT) {
    if (which == MGD_KERNEL_RHS_ADJOINT) {
      if (int rc = ensure_adjoint(r)) return rc;
      launch_copy(M, r->q.p, r->lam.p, r->stream);
    }
  } else if (which != MGD_KERNEL_COPY) {
    set_error("unknown kernel");
    return MGD_ERR_ARG;
  }
  launch();   // warm-up
  MGD_HIP(hipEventRecord(r->ev0, r->stream));
  for (int i = 0; i < reps; ++i) launch();
  MGD_HIP(hipEventRecord(r->ev1, r->stream));
  MGD_HIP(hipGetLastError());
  MGD_HIP(hipEventSynchronize(r->ev1));
  float ms = 0.f;
  MGD_HIP(hipEventElapsedTime(&ms, r->ev0, r->ev1));
  if (ms_per_launch) *ms_per_launch = ms / (reps * per_rep);
  if (bytes) {
    switch (which) {
      // state arrays per step: forward 3 + 5 + 5 + 3 (stage state, level, buffer2 in;
      // buffer2, next stage state out), adjoint 4 + 6 + 6 + 4 (adjoint stage state, level,
      // forward stage state, buffer2 in; buffer2, next adjoint out)
      case MGD_KERNEL_FWD_STAGE: *bytes = 16.0 / 4.0 * w * pts; break;
      case MGD_KERNEL_ADJ_STAGE: *bytes = 20.0 / 4.0 * w * pts; break;
      // one rhs evaluation: state in, rhs out (forward); state and adjoint in (adjoint)
      case MGD_KERNEL_RHS_ADJOINT: *bytes = 3.0 * w * pts; break;
      // the march kernel's own operands per stage, averaged over the 4 stages: the stage
      // state and the 9 stress / heat-flux slots in, RK buffers in (0, 2, 2, 1 states),
      // RK buffers and the next stage state out (3, 2, 2, 1 states)
      case MGD_KERNEL_FWD_MARCH3D: *bytes = (w + 72.0 + 5.0 / 4.0 * w + 7.0 / 4.0 * w) * pts; break;
      default: *bytes = 2.0 * w * pts; break;
    }
  }
  return MGD_OK;
}

int mgd_sync(mgd_region* r) {
  if (!r) return MGD_ERR_ARG;
  MGD_HIP(hipStreamSynchronize(r->stream));
  return MGD_OK;
}

int mgd_group_create(mgd_region* const* regions, int n, mgd_group** out) {
  if (!out || !regions || n < 1 || n > kMaxGroup) {
    set_error("a group holds 1 .. 64 regions");
    return MGD_ERR_ARG;
  }
  *out = nullptr;
  auto G = std::make_unique<mgd_group>();
  for (int k = 0; k < n; ++k) {
    if (!regions[k] || !regions[k]->finalized) { set_error("group member not finalized"); return MGD_ERR_STATE; }
    G->m.push_back(regions[k]);
  }
  MGD_HIP(hipStreamCreate(&G->st));
  MGD_HIP(hipStreamCreate(&G->st_rec));
  MGD_HIP(hipEventCreateWithFlags(&G->ev_fwd, hipEventDisableTiming));
  for (int i = 0; i < 2; ++i) {
    MGD_HIP(hipEventCreateWithFlags(&G->ev_ready[i], hipEventDisableTiming));
    MGD_HIP(hipEventCreateWithFlags(&G->ev_free[i], hipEventDisableTiming));
  }
  *out = G.release();
  return MGD_OK;
}

int mgd_group_destroy(mgd_group* G) {
  if (!G) return MGD_OK;
  if (G->st) (void)hipStreamSynchronize(G->st);
  if (G->st_rec) (void)hipStreamSynchronize(G->st_rec);
  if (G->ev_fwd) (void)hipEventDestroy(G->ev_fwd);
  for (int i = 0; i < 2; ++i) {
    if (G->ev_ready[i]) (void)hipEventDestroy(G->ev_ready[i]);
    if (G->ev_free[i]) (void)hipEventDestroy(G->ev_free[i]);
  }
  if (G->st) (void)hipStreamDestroy(G->st);
  if (G->st_rec) (void)hipStreamDestroy(G->st_rec);
  delete G;
  return MGD_OK;
}

int mgd_group_run_adjoint(mgd_group* G, double t0, const double* dt, int nsteps, int save_interval,
                          const double* const* controls, double* const* gradients,
                          double* const* sensitivities, double* const* Js) {
  if (!G) { set_error("null group"); return MGD_ERR_ARG; }
  if (nsteps <= 0 || save_interval <= 0 || !dt) { set_error("bad time stepping"); return MGD_ERR_ARG; }
  const int n = (int)G->m.size();
  for (int k = 0; k < n; ++k) {
    mgd_region* r = G->m[k];
    if (!(dt[k] > 0.0)) { set_error("bad time step"); return MGD_ERR_ARG; }
    if (!fused_ok(r)) {
      set_error("group members must run on the fused 2-D path");
      return MGD_ERR_UNSUPPORTED;
    }
  }
  // every member's stage states resident where every member asks for it
  // (MGD_OPT_RESIDENT_STATES = 1), else every member on snapshots and windows
  int nres = 0;
  for (mgd_region* r : G->m) nres += r->opt_resident == 1;
  if (nres != 0 && nres != n) {
    set_error("MGD_OPT_RESIDENT_STATES = 1 on some members of a group but not on all");
    return MGD_ERR_ARG;
  }
  const bool gres = nres == n;
  // every member works on the group's streams for the duration of the run
  std::vector<hipStream_t> streams;
  std::vector<int> shapes;
  for (mgd_region* r : G->m) {
    streams.push_back(r->stream);
    shapes.push_back(r->tiling.shape);
    r->stream = G->st;
  }
  for (mgd_region* r : G->m) r->group_run = true;
  auto restore = [&]() {
    for (int k = 0; k < n; ++k) {
      mgd_region* r = G->m[k];
      r->group_run = false;
      r->group_resident = false;
      r->stream = streams[k];
      if (r->g_swapped) swap_tilings(r);
      if (r->tiling.shape != shapes[k]) (void)install_tiling(r, shapes[k]);
    }
  };
  auto run = [&]() -> int {
    for (int k = 0; k < n; ++k) {
      mgd_region* r = G->m[k];
      r->group_resident = gres;
      if (int rc = prepare_forward_ck(r, t0, dt[k], nsteps, save_interval, controls ? controls[k] : nullptr))
        return rc;
      r->ck.valid = true;
      if (int rc = prepare_adjoint_ck(r, true)) return rc;
      if (r->tiling.shape != kGroupShape) {
        // the group tiling, built once per region, in place of the region's own
        if (!r->g_built && make_tiling(r, kGroupShape, r->g_htile, r->g_tiling, r->g_tile_mask) != MGD_OK) {
          set_error("a group member does not fit the 16 x 16 tiles of a group launch");
          return MGD_ERR_UNSUPPORTED;
        }
        r->g_built = true;
        swap_tilings(r);
      }
    }
    const bool want_grad = gradients != nullptr;
    if (int rc = group_tables(G, dt[0], dt, nsteps, save_interval, want_grad)) return rc;
    if (int rc = group_forward_pass(G, nsteps, save_interval)) return rc;
    if (int rc = group_adjoint_pass(G, dt, nsteps, save_interval)) return rc;
    MGD_HIP(hipGetLastError());
    MGD_HIP(hipStreamSynchronize(G->st));
    for (int k = 0; k < n; ++k) {
      mgd_region* r = G->m[k];
      if (Js && Js[k])
        MGD_HIP(hipMemcpy(Js[k], r->acc_J.p, sizeof(double) * r->ninst, hipMemcpyDeviceToHost));
      if (int rc = copy_adjoint_outputs(r, gradients ? gradients[k] : nullptr,
                                        sensitivities ? sensitivities[k] : nullptr))
        return rc;
    }
    return MGD_OK;
  };
  const int rc = run();
  if (rc != MGD_OK) (void)hipStreamSynchronize(G->st);
  restore();
  return rc;
}

}  // extern "C"
