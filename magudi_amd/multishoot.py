"""Multiple-shooting drivers over the HIP path: msforward and msadjoint.

Mirrors bin/msforward.f90 (segment loop :242-365, aggregation :381-400) and
bin/msadjoint.f90 (segment loop :287-384, totals :420-470) for the configuration
without a state mollifier.  The reference runs the Nsplit time segments one after
another and exchanges states through files; here the segments are the instances of
one batched region, so every RK4 stage of every segment is one launch:

* instance k holds segment k: its initial condition ic_k (MGD_INITIAL_STATE), its
  start time (per-instance time offset) and its slice of the global control
  (controlTimestepOffset = k * Nts);
* msforward: one checkpointed forward run of all segments, then the matching
  mismatches ic_k - end_(k-1) and their SBP-norm inner products on the device;
* msadjoint: the matching terminal conditions penaltyScale (end_k - ic_(k+1)) of all
  segments at once (they depend only on the forward result), one adjoint run of all
  segments, then the IC gradients lambda_(k+1) + penaltyScale (ic_(k+1) - end_k).

Penalties follow src/MSPenaltyImpl.f90: quadratic (0.5 L2sq, gradient factor 1) and
Huber with a threshold.
"""
from __future__ import annotations

import ctypes as C
import math
from typing import Optional, Sequence

import numpy as np

from . import capi
from .capi import check, dptr
from .cases import Case
from .solver import Region

__all__ = ["MultipleShooting", "penalty_norm", "penalty_gradient_factor", "segment_initial_guesses"]


def segment_initial_guesses(case: Case, nsplit: int, nts: int, perturbation: float = 1e-3,
                            seed: int = 0, dt: Optional[float] = None):
    """Initial guesses of the segments' initial conditions, as msforward takes them from the
    <prefix>-<timestep>.q files of a baseline run: the states of one forward run from the
    case's initial condition at t = k nts dt (on the GPU), each but the first offset by a
    seeded random perturbation of the momenta (so that the matching conditions are not
    met exactly); returns (ics, t0s)."""
    from .solver import Solver
    dt = dt if dt is not None else case.dt
    R = Region(case)
    S = Solver(R, dt=dt, nsteps=nts)
    q = case.initial_state()
    rng = np.random.Generator(np.random.PCG64(seed))
    ics, t0s = [q], [0.0]
    for k in range(1, nsplit):
        R.states[0].conservedVariables = q
        S.runForward(t0=(k - 1) * nts * dt)
        q = R.states[0].conservedVariables
        ic = q.copy()
        for c in range(1, case.nd + 1):
            ic[c] = ic[c] + perturbation * rng.standard_normal(case.npoints)
        ics.append(ic)
        t0s.append(k * nts * dt)
    R.close()
    return ics, t0s


def penalty_norm(kind: str, L2sq: float, threshold: float) -> float:
    """t_MSPenalty%norm (src/MSPenaltyImpl.f90:3-71)."""
    if kind == "quadratic":
        return 0.5 * L2sq
    if kind == "huber":
        n = math.sqrt(L2sq)
        return n - 0.5 * threshold if n >= threshold else 0.5 * L2sq / threshold
    raise ValueError(f"Unknown time_splitting/penalty_type '{kind}': expected 'quadratic' or 'huber'.")


def penalty_gradient_factor(kind: str, L2sq: float, threshold: float) -> float:
    """t_MSPenalty%gradientFactor (src/MSPenaltyImpl.f90:25-97)."""
    if kind == "quadratic":
        return 1.0
    if kind == "huber":
        n = math.sqrt(L2sq)
        return n if n >= threshold else threshold
    raise ValueError(f"Unknown time_splitting/penalty_type '{kind}': expected 'quadratic' or 'huber'.")


class MultipleShooting:
    """time_splitting/* options of magudi.inp: number_of_segments (nsplit),
    segment_length (nts), penalty_weight, penalty_type, huber/threshold; save_interval
    is the checkpoint interval of the adjoint runs."""

    def __init__(self, case: Case, nsplit: int, nts: int, penalty_weight: float,
                 penalty_type: str = "quadratic", huber_threshold: float = 1e-12,
                 save_interval: Optional[int] = None, dt: Optional[float] = None,
                 force_plain: bool = False):
        penalty_norm(penalty_type, 1.0, huber_threshold)   # validates the type
        self.case = case
        self.nsplit, self.nts = nsplit, nts
        self.weight = penalty_weight
        self.kind, self.threshold = penalty_type, huber_threshold
        self.K = save_interval if save_interval is not None else max(1, min(nts, case.save_interval))
        self.dt = dt if dt is not None else case.dt
        self.region = Region(case, n_instances=nsplit, force_plain=force_plain)
        self.L2sq = np.zeros(nsplit)
        self.segment_cost = np.zeros(nsplit)
        self._fwd_done = False

    # -- helpers ---------------------------------------------------------------
    def _lincomb(self, dst, dst_inst, a, a_inst, b, b_inst, s, c=-1, c_inst=0):
        check(capi.lib().mgd_state_lincomb(self.region.handle, dst, dst_inst, a, a_inst, b, b_inst,
                                           s, c, c_inst))

    def _inner(self, a, b) -> np.ndarray:
        out = np.zeros(self.nsplit)
        check(capi.lib().mgd_inner_product(self.region.handle, a, b, dptr(out)))
        return out

    def _segment_control(self, control):
        if control is None or self.region.n_actuator == 0:
            return None
        c = np.ascontiguousarray(control, dtype=np.float64).reshape(
            self.nsplit, 4 * self.nts, self.region.n_actuator)
        return np.ascontiguousarray(c)

    # -- msforward -------------------------------------------------------------
    def msforward(self, ics: Sequence[np.ndarray], t0s: Sequence[float],
                  control: Optional[np.ndarray] = None) -> float:
        """Segment runs from ic_k at times t0s[k]; returns J = sum of segment costs +
        penalty_weight * norm(sum of mismatch L2sq)."""
        r = self.region
        if len(ics) != self.nsplit or len(t0s) != self.nsplit:
            raise ValueError("one initial condition and one start time per segment")
        for k, ic in enumerate(ics):
            r.states[k].initialCondition = ic
        r.restart()
        offs = np.ascontiguousarray(t0s, dtype=np.float64)
        check(capi.lib().mgd_region_set_time_offsets(r.handle, dptr(offs)))
        self._control = self._segment_control(control)
        J = np.zeros(self.nsplit)
        check(capi.lib().mgd_run_forward_checkpointed(r.handle, 0.0, self.dt, self.nts, self.K,
                                                      dptr(self._control), dptr(J)))
        self.segment_cost = J
        # mismatch ic_k - end_(k-1) and its SBP-norm L2sq (bin/msforward.f90:300-318)
        for k in range(1, self.nsplit):
            self._lincomb(capi.SCRATCH, k, capi.INITIAL_STATE, k, capi.FINAL_STATE, k - 1, 1.0)
        self._lincomb(capi.SCRATCH, 0, capi.INITIAL_STATE, 0, capi.INITIAL_STATE, 0, 0.0)
        ip = self._inner(capi.SCRATCH, capi.SCRATCH)
        self.L2sq = np.zeros(self.nsplit)
        self.L2sq[1:] = ip[1:]
        L2sum = 0.0
        for k in range(1, self.nsplit):
            L2sum = L2sum + self.L2sq[k]
        self.L2sum = L2sum
        total = 0.0
        for k in range(self.nsplit):
            total = total + J[k]
        self.J_time_integral = total
        self.J_penalty = self.weight * penalty_norm(self.kind, L2sum, self.threshold)
        self._fwd_done = True
        return total + self.J_penalty

    # -- msadjoint -------------------------------------------------------------
    def msadjoint(self, want_control_gradient: bool = True):
        """Adjoint of the last msforward.  Returns a dict: ic_grad (nsplit, nv, N; index 0
        is the adjoint at the start of segment 0), control_grad (global stages x actuator
        points or None), ctrl_ip, ic_ip (per segment) and total."""
        if not self._fwd_done:
            raise RuntimeError("run msforward first")
        r = self.region
        scale = self.weight / penalty_gradient_factor(self.kind, self.L2sum, self.threshold)
        # matching adjoint terminals (bin/msadjoint.f90:377-400): scale (end_k - ic_(k+1));
        # the last segment starts from zero
        last = self.nsplit - 1
        for k in range(last):
            self._lincomb(capi.ADJOINT_VARS, k, capi.FINAL_STATE, k, capi.INITIAL_STATE, k + 1, scale)
        self._lincomb(capi.ADJOINT_VARS, last, capi.FINAL_STATE, last, capi.FINAL_STATE, last, 0.0)
        sens = np.zeros(self.nsplit)
        grad = None
        na = r.n_actuator
        if want_control_gradient and na > 0:
            grad = np.zeros((self.nsplit, 4 * self.nts, na))
        check(capi.lib().mgd_run_adjoint_checkpointed(r.handle, 1, dptr(grad), dptr(sens)))
        # IC gradients (bin/msadjoint.f90:340-352): lambda_k + scale (ic_k - end_(k-1))
        for k in range(1, self.nsplit):
            self._lincomb(capi.SCRATCH, k, capi.INITIAL_STATE, k, capi.FINAL_STATE, k - 1, scale,
                          capi.ADJOINT_VARS, k)
        ip = self._inner(capi.SCRATCH, capi.SCRATCH)
        ic_ip = np.zeros(self.nsplit)
        ic_ip[1:] = ip[1:]
        ic_grad = np.empty((self.nsplit, r.nv, r.npoints))
        ic_grad[0] = r.states[0].adjointVariables
        for k in range(1, self.nsplit):
            out = np.empty((r.nv, r.npoints))
            check(capi.lib().mgd_state_get(r.handle, capi.SCRATCH, k, dptr(out)))
            ic_grad[k] = out
        total = 0.0
        for k in range(self.nsplit):
            total = total + sens[k] + ic_ip[k]
        return {"ic_grad": ic_grad,
                "control_grad": None if grad is None else grad.reshape(-1, na),
                "ctrl_ip": sens, "ic_ip": ic_ip, "total": total, "scale": scale}
