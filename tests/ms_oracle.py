"""Oracle restatement of the multiple-shooting drivers (test infrastructure only).

msforward (bin/msforward.f90:242-365, 381-400) and msadjoint (bin/msadjoint.f90:287-384,
400-470) without the state mollifier, built on the CPU oracle's per-segment runs
(oracle/orc_solver.c: orc_run_forward, orc_run_adjoint_ic).  Segment k starts from
its initial condition ic_k at time t0s[k] and runs nts steps with its slice of the
global control (controlTimestepOffset = k * nts).  Penalty norms and gradient factors
follow src/MSPenaltyImpl.f90 (quadratic: 0.5 L2sq and 1; Huber with threshold h:
sqrt(L2sq) - h/2 or 0.5 L2sq / h, and max(sqrt(L2sq), h)).
"""
import math

import numpy as np


def penalty_norm(kind, L2sq, threshold):
    if kind == "quadratic":
        return 0.5 * L2sq
    n = math.sqrt(L2sq)
    return n - 0.5 * threshold if n >= threshold else 0.5 * L2sq / threshold


def penalty_gradient_factor(kind, L2sq, threshold):
    if kind == "quadratic":
        return 1.0
    n = math.sqrt(L2sq)
    return n if n >= threshold else threshold


def inner(norm, f, g):
    """Grid%computeInnerProduct (src/GridImpl.f90:1067-1170) of (nv, N) arrays."""
    s = 0.0
    for c in range(f.shape[0]):
        s = s + float(np.sum(f[c] * norm * g[c]))
    return s


def control_slice(control, k, nts):
    if control is None:
        return None
    return np.ascontiguousarray(control[4 * nts * k:4 * nts * (k + 1)])


def ms_forward(O, ics, t0s, dt, nts, weight, control=None, kind="quadratic", threshold=1e-12):
    """Returns (J, segment costs, L2sq per segment, end states)."""
    norm = O.grid(2)[0]
    nsplit = len(ics)
    cost, L2sq, ends = [], [0.0] * nsplit, []
    L2sum = 0.0
    for k in range(nsplit):
        if k > 0:
            d = ics[k] - ends[k - 1]
            L2sq[k] = inner(norm, d, d)
            L2sum = L2sum + L2sq[k]
        Jk, qT = O.run_forward(ics[k], dt, nts, control=control_slice(control, k, nts), t0=t0s[k])
        cost.append(Jk)
        ends.append(qT)
    J = sum(cost) + weight * penalty_norm(kind, L2sum, threshold)
    return J, cost, L2sq, ends


def ms_adjoint(O, ics, t0s, dt, nts, K, weight, control=None, kind="quadratic", threshold=1e-12):
    """Returns dict: ic_grad (list, index 0 = adjoint at the start of segment 0),
    control_grad (global, stages x actuator points), ctrl_ip, ic_ip, total."""
    norm = O.grid(2)[0]
    nsplit = len(ics)
    _, _, L2sq, ends = ms_forward(O, ics, t0s, dt, nts, weight, control, kind, threshold)
    scale = weight / penalty_gradient_factor(kind, sum(L2sq), threshold)
    na = O.n_actuator
    ic_grad = [None] * nsplit
    ctrl_ip = [0.0] * nsplit
    ic_ip = [0.0] * nsplit
    cgrad = np.zeros((4 * nts * nsplit, max(na, 1)))[:, :na]
    lam_next = None
    for k in range(nsplit - 1, -1, -1):
        lamT = None
        if k < nsplit - 1:
            g = lam_next + scale * (ics[k + 1] - ends[k])
            ic_grad[k + 1] = g
            ic_ip[k + 1] = inner(norm, g, g)
            lamT = scale * (ends[k] - ics[k + 1])
        sens, grad, lam0, _, _ = O.run_adjoint_ic(ics[k], dt, nts, K,
                                                  control=control_slice(control, k, nts),
                                                  t0=t0s[k], lamT=lamT)
        ctrl_ip[k] = sens
        cgrad[4 * nts * k:4 * nts * (k + 1)] = grad
        lam_next = lam0
    ic_grad[0] = lam_next
    total = 0.0
    for k in range(nsplit):
        total = total + ctrl_ip[k] + ic_ip[k]
    return {"ic_grad": ic_grad, "control_grad": cgrad, "ctrl_ip": ctrl_ip, "ic_ip": ic_ip,
            "total": total, "L2sq": L2sq, "scale": scale}
