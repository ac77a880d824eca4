// kernels.h -- launchers of the magudi RHS kernels (kernels.hip, fused2d.hip).
#pragma once
#include "mgd_internal.h"

namespace mgd {

constexpr int kBlock = 256;

// F: flux scratch (nd * nv * N per instance) -- fluxes computed once per point and
// differentiated from memory; null recomputes them at every stencil tap
// What the finishing pass of a forward rhs has to do, as far as the host knows it:
// far-field patches present; whether the patches are at most one Kolmogorov forcing
// over the whole domain (`kolmogorov`: its strength array, null for none) -- then, with
// no dissipation and no sources, the pass rides on the last derivative kernel of the
// periodic 3-D path.
struct FinishInfo {
  int has_farfield = 0;
  int simple = 0;
  const double* kolmogorov = nullptr;
};
// One RK4 stage update that a rhs launcher may apply itself (substepForwardRK4 /
// substepAdjointRK4, src/RK4IntegratorImpl.f90:115-156, 215-261): y the state updated in
// place, b1 / b2 the RK buffers, copy (forward, optional) a second destination of the
// new stage state.
// yout (forward, optional): a state buffer other than y the new stage state may be
// written to instead of y, for kernels that cannot update y in place.
struct RkStage {
  int stage;
  double dt;
  double* y;
  double* b1;
  double* b2;
  double* copy;
  double* yout = nullptr;
  // events recorded on the stream right before and after the z-march kernel, when set
  hipEvent_t ev_begin = nullptr;
  hipEvent_t ev_end = nullptr;
  // adjoint stages on the 3-D march path (stress_reuse_path): G already holds the stress /
  // heat-flux slots of the stage state, so the stress pass is skipped
  bool stress_ready = false;
};
// What launch_rhs_forward did with `rk`.
enum RkDone { RK_NOT_APPLIED = 0, RK_IN_PLACE = 1, RK_TO_YOUT = 2 };
// RK_IN_PLACE / RK_TO_YOUT when it applied `rk` (then rhs is not written; with RK_TO_YOUT
// the new stage state is in rk->yout and y is unchanged); RK_NOT_APPLIED when the caller
// still has to (launch_rk_forward) -- always for rk == nullptr.
int launch_rhs_forward(const Geo& g, const Patches* P, const double* q, double* rhs, double* w,
                       double* G, double* t, double* F, const double* control, long control_stride,
                       const Source* src, int nsrc, const double* src_amp, long src_inst_stride,
                       const FinishInfo& fin, hipStream_t st, const RkStage* rk = nullptr);
// Returns true when it applied `rk` (then rhs is not written); false when the caller
// still has to (launch_rk_adjoint) -- always false for rk == nullptr.
bool launch_rhs_adjoint(const Geo& g, const Patches* P, const double* q, const double* lam,
                        double* rhs, double* w, double* G, double* t, double* T1, double* S,
                        const double* aforce, long aforce_stride, double forcing_factor,
                        const FinishInfo& fin, hipStream_t st, const RkStage* rk = nullptr);
// The 3-D z-march forward path (periodic in every direction, at most one whole-domain
// Kolmogorov forcing, no dissipation or sources): whether it serves a region; the stress /
// heat-flux kernel (9 slots of the 12-slot gradient buffer S); the march itself, which
// writes rhs, or with rk (rk->yout set, not q) applies the RK4 stage out of place.
bool fwd3d_march_path(const Geo& g, const FinishInfo& fin, int nsrc);
// Whether the forward and the adjoint rhs both run on the 3-D marches, each from the stress
// / heat-flux slots of its state that launch_stress3d writes into G: then slots the forward
// stage computed serve the adjoint stage about the same state.
bool stress_reuse_path(const Geo& g, const FinishInfo& fin, int nsrc);
void launch_stress3d(const Geo& g, const double* q, double* S, hipStream_t st);
void launch_fwd3d_march(const Geo& g, const double* q, const double* S, const double* kstr,
                        double* rhs, const RkStage* rk, hipStream_t st);
void launch_rk_forward(long M, int stage, double dt, double* q, double* b1, double* b2,
                       const double* rhs, double* qcopy, hipStream_t st);
void launch_rk_adjoint(long M, int stage, double dt, double* lam, double* b1, double* b2,
                       const double* rhs, hipStream_t st);
void launch_copy(long M, const double* a, double* b, hipStream_t st);
void launch_fill(long M, double v, double* b, hipStream_t st);
int reduce_blocks(int np);
void launch_patch_reduce(const Geo& g, const DPatch* pt, int np, int which, const double* x,
                         double* partial, double weight, double* acc, double* raw, hipStream_t st);
void launch_adj_forcing(const Geo& g, const DPatch* pt, int np, const double* q, double* out,
                        int zero, hipStream_t st);
void launch_adj_terminal(const Geo& g, const DPatch* pt, int np, const double* af, double factor,
                         double* lam, hipStream_t st);
void launch_gradient(const Geo& g, const DPatch* pt, int np, const double* lam, double* out,
                     long inst_stride, hipStream_t st);
// out[inst] = <a, b>_H over the instance's state arrays; partial: ninst * ceil(N/256)
void launch_inner_product(const Geo& g, const double* a, const double* b, double* partial,
                          double* out, hipStream_t st);
// computeTimeStepSize (mode 0: out = min cfl / speed, param = cfl) or computeCfl (mode 1:
// out = max speed * dt, param = dt) per instance; partial: ninst * ceil(N/256)
void launch_wave_speed(const Geo& g, const double* q, int mode, double param, double* partial,
                       double* out, hipStream_t st);
// Solution limits: flag[inst] = 1 where density or temperature of instance inst reaches
// range (rho_min, rho_max, T_min, T_max) or is not a number (flags are not cleared)
void launch_check_limits(const Geo& g, const double* q, const double* range, double* flag,
                         hipStream_t st);
// Terminal kinetic energy: J[inst] = sum_p H_p |rho u|^2 / (2 rho), lam = dJ/dq pointwise;
// partial: ninst * ceil(N/256)
void launch_kinetic_energy(const Geo& g, const double* q, double* lam, double* partial, double* J,
                           hipStream_t st);
void launch_lincomb(long M, const double* a, const double* b, double s, const double* c, double* d,
                    hipStream_t st);
void launch_apply(const Geo& g, int slot, int d, int ncomp, const double* x, double* y,
                  hipStream_t st);

}  // namespace mgd
