#!/usr/bin/env python3
"""Benchmark of the magudi RHS hot path on one MI355X.

Headline workload (one "step" = one full pass of the hot path): the 3-D viscous
periodic Kolmogorov flow on a 256^3 grid (SBP 4-8, Navier-Stokes with Re 100 at
Mach 0.2, Pr 0.7, forcing f_x = A sin(2 pi 4 y), seeded solenoidal initial velocity),
200 RK4 steps forward with HBM snapshots every 25 steps, J = kinetic energy of the
final state and its gradient dJ/dq(T) on the device, then the discrete adjoint back to
t0 with window recompute, whose result is the gradient dJ/dq0.  Every part of the pass
runs inside the timed region; the initial condition is restored from a device copy.
Synthetic input built in magudi_amd/cases.py.

value = grid-point-steps per second of the forward+adjoint pass (256^3 * 200 / time).
Also reported: the dominant kernel (the 3-D z-march forward kernel) against the HBM
roofline, timed with HIP events on the region stream, and its HBM traffic from two
rocprofv3 counter passes run by this script; the 3-D forward and adjoint rhs; the 2-D
monopole legs (401^2 forward + adjoint with the control gradient, 20-instance batches,
the amplitude sweep, uneven batches and ensembles, the generic per-operator path); the
CPU oracle on a bounded sample of the headline workload.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
METRIC = "FP64 grid-pt-steps/s fwd+adjoint, batched vs plain; % of HBM roofline"


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--c4-n", type=int, default=256, help="grid of the headline 3-D run")
    ap.add_argument("--c4-nsteps", type=int, default=200, help="RK4 steps of the headline run")
    ap.add_argument("--n", type=int, default=401, help="grid of the 2-D monopole legs")
    ap.add_argument("--nsteps", type=int, default=960)
    ap.add_argument("--save-interval", type=int, default=32)
    ap.add_argument("--batch", type=int, default=20)
    ap.add_argument("--no-extras", action="store_true",
                    help="headline only (no 2-D, batched, plain or CPU legs)")
    ap.add_argument("--no-traffic", action="store_true",
                    help="skip the two rocprofv3 counter passes for roofline.traffic")
    ap.add_argument("--fd-sweep-only", action="store_true",
                    help="only the batched finite-difference sweep over control amplitudes")
    ap.add_argument("--kolmogorov-only", action="store_true",
                    help="only the headline 3-D Kolmogorov forward + adjoint run (no extras, "
                         "no counter passes)")
    ap.add_argument("--viscous3d-only", action="store_true",
                    help="only the 3-D viscous rhs leg (its profiling workload)")
    ap.add_argument("--monopole-only", action="store_true",
                    help="only the 401^2 monopole forward + adjoint legs")
    ap.add_argument("--batched-only", action="store_true",
                    help="only the batched 401^2 monopole forward + adjoint (the step store; "
                         "its profiling workload)")
    ap.add_argument("--tile-shape", type=int, default=-1,
                    help="tile shape of the fused 2-D kernels (-1: autotune)")
    ap.add_argument("--kernels-only", action="store_true",
                    help="only the HIP-event timing of the headline's dominant kernel and the "
                         "state copy (the workload of the rocprofv3 counter passes)")
    ap.add_argument("--kernels-only-2d", action="store_true",
                    help="the same for the fused 2-D stage kernels on the 401^2 monopole")
    return ap.parse_args()


def counter_pass(counter, args):
    """One rocprofv3 counter pass (a child process) over `bench.py <args>`; returns
    {kernel name: (dispatches, mean counter value per dispatch)} or None."""
    import csv
    import glob
    import shutil
    import subprocess
    import tempfile
    exe = shutil.which("rocprofv3")
    if not exe:
        return None
    out = tempfile.mkdtemp(prefix="mgd_pmc_", dir=os.environ.get("TMPDIR", "/tmp"))
    cmd = [exe, "--pmc", counter, "--output-format", "csv", "-d", out, "-o", "run", "--",
           sys.executable, os.path.join(ROOT, "bench.py"), *args]
    try:
        subprocess.run(cmd, cwd=out, stdout=subprocess.DEVNULL, stderr=subprocess.DEVNULL,
                       timeout=240, check=False)
    except subprocess.TimeoutExpired:
        return None
    acc = {}
    for path in glob.glob(os.path.join(out, "**", "*counter_collection.csv"), recursive=True):
        with open(path) as f:
            for row in csv.DictReader(f):
                if row.get("Counter_Name") != counter:
                    continue
                k = row["Kernel_Name"]
                n, v = acc.get(k, (0, 0.0))
                acc[k] = (n + 1, v + float(row["Counter_Value"]))
    shutil.rmtree(out, ignore_errors=True)
    return {k: (n, v / n) for k, (n, v) in acc.items()} or None


def measured_traffic(args, kernel, copy_bytes):
    """HBM bytes per launch of `kernel` (a substring of its name) from FETCH_SIZE and
    WRITE_SIZE (KiB per dispatch), each scaled by what the same counter reads for the state
    copy kernel of known size (copy_bytes read and copy_bytes written)."""
    res = {}
    for counter in ("FETCH_SIZE", "WRITE_SIZE"):
        c = counter_pass(counter, args)
        if not c:
            return None
        kv = [v for k, (n, v) in c.items() if kernel in k]
        cv = [v for k, (n, v) in c.items() if "k_copy" in k]
        if not kv or not cv:
            return None
        factor = copy_bytes / (cv[0] * 1024.0)
        res[counter] = {"kib_per_dispatch": kv[0], "copy_kib": cv[0], "factor": factor,
                        "bytes": kv[0] * 1024.0 * factor}
    return {"bytes": res["FETCH_SIZE"]["bytes"] + res["WRITE_SIZE"]["bytes"], "passes": res}


def run_hot_path(R, S, steps, use_graph=False):
    """steps x (restart from the resident initial condition + forward + adjoint with the
    control gradient); use_graph replays the run's device work as one HIP graph."""
    R.sync()
    t0 = time.perf_counter()
    for _ in range(steps):
        R.restart()
        out = S.runAdjoint(want_gradient=True, use_graph=use_graph)
    R.sync()
    return (time.perf_counter() - t0) / steps, out


def cpu_baseline(n=64, nsteps=8, save_interval=4):
    """The CPU oracle (plain C; its per-point and per-pencil loops on every OpenMP thread
    the host gives, OMP_NUM_THREADS) on a bounded sample of the headline workload: the
    same Kolmogorov flow on an n^3 grid for nsteps RK4 steps, forward with snapshots,
    terminal kinetic energy, adjoint back from dJ/dq(T).  The terminal state comes from an
    untimed forward run first (the oracle's adjoint driver redoes the forward pass)."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import oracle_bind as ob
    from magudi_amd import cases
    from magudi_amd.functionals import kinetic_energy
    c = cases.kolmogorov3d(n=n, nsteps=nsteps)
    O = ob.Region(c)
    q0 = c.initial_state()
    norm = O.grid(2)[0]
    threads = ob.lib().orc_threads()
    _, qT = O.run_forward(q0, c.dt, nsteps)
    t0 = time.perf_counter()
    _, lamT = kinetic_energy(qT, norm, 3)
    O.run_adjoint_ic(q0, c.dt, nsteps, save_interval, lamT=lamT)
    dt = time.perf_counter() - t0
    return {"value": c.npoints * nsteps / dt, "unit": "grid-pt-steps/s", "cores": threads,
            "kind": "port",
            "sample": f"Kolmogorov {n}^3 (the headline flow on a smaller grid), {nsteps} RK4 "
                      f"steps forward with snapshots every {save_interval} + terminal kinetic "
                      f"energy + adjoint (oracle/ C restatement, {threads} OpenMP threads, "
                      f"{dt:.1f} s wall, {dt * threads:.0f} s of CPU work)"}


def viscous3d_leg(n):
    """One forward and one adjoint rhs evaluation of the 3-D viscous periodic Kolmogorov
    flow at n^3, HIP-event timed (the stress pass and the z marches, after the march-shape
    autotuner)."""
    from magudi_amd import capi, cases
    from magudi_amd.solver import Region
    c = cases.kolmogorov3d(n=n, nsteps=1)
    R = Region(c)
    R.states[0].conservedVariables = c.initial_state()
    f_ms, f_b = R.time_kernel(capi.KERNEL_RHS_FORWARD, reps=5)
    a_ms, a_b = R.time_kernel(capi.KERNEL_RHS_ADJOINT, reps=5)
    c_ms, c_b = R.time_kernel(capi.KERNEL_COPY, reps=5)   # counter calibration
    del R
    leg = {"workload": f"Kolmogorov {n}^3 viscous periodic rhs (Re_ac 500, Pr 0.7)",
           "path": "3-D kernels (forward: stress pass + z march; adjoint: stress pass, "
                   "Jacobian march, finishing march)",
           "state_copy": {"ms": round(c_ms, 4), "bytes": c_b,
                          "achieved_GBs": round(c_b / (c_ms * 1e-3) / 1e9, 1)}}
    for k, ms, b in (("forward_rhs", f_ms, f_b), ("adjoint_rhs", a_ms, a_b)):
        gbs = b / (ms * 1e-3) / 1e9
        leg[k] = {"ms": round(ms, 4), "algorithmic_bytes": b, "achieved_GBs": round(gbs, 1),
                  "frac": round(gbs / HBM_PEAK_GBS, 4), "pt_per_s": n ** 3 / (ms * 1e-3)}
    return leg


class KolmogorovRun:
    """The headline: the 3-D viscous periodic Kolmogorov flow as the configuration runs it.
    One pass = restart from the resident initial condition, nsteps RK4 steps forward
    keeping the state of every step and the inner stage states of the last steps in HBM
    (the step store; save_interval nsteps / 8), J = kinetic energy of the final state and
    dJ/dq(T) on the device, the discrete adjoint back from dJ/dq(T), recomputing the inner
    stage states of the steps not held (its result, the adjoint state at t0, is dJ/dq0 and
    stays on the device)."""

    def __init__(self, n, nsteps, step_store=-1):
        from magudi_amd import cases
        from magudi_amd.solver import Region, Solver
        self.case = cases.kolmogorov3d(n=n, nsteps=nsteps)
        self.R = Region(self.case)
        self.R.set_step_store(step_store)
        self.R.states[0].initialCondition = self.case.initial_state()
        self.S = Solver(self.R, dt=self.case.dt, nsteps=nsteps,
                        save_interval=self.case.save_interval)
        self.pts = self.case.npoints * nsteps

    def one(self):
        self.R.restart()
        self.S.runForwardCheckpointed()
        J = self.R.terminalKineticEnergy()
        self.S.runAdjointCheckpointed(terminal=True, want_gradient=False)
        return J

    def timed(self, steps):
        """(seconds per pass, J, forward seconds of a pass, z-march launches in the timed
        passes, those of them that also wrote a state copy, their summed ms -- HIP events
        around every launch on the region stream)"""
        self.R.sync()
        self.R.kernel_timer(True)
        t0 = time.perf_counter()
        for _ in range(steps):
            J = self.one()
        self.R.sync()
        t = (time.perf_counter() - t0) / steps
        launches, with_copy, ms = self.R.kernel_timer_read()
        self.R.kernel_timer(False)
        # forward half alone, for the split
        self.R.restart()
        self.R.sync()
        t1 = time.perf_counter()
        self.S.runForwardCheckpointed()
        self.R.sync()
        return t, float(J[0]), time.perf_counter() - t1, launches, with_copy, ms


def fd_sweep_leg(n=401, nsteps=960, save_interval=32, batch=20):
    """The batched finite-difference gradient sweep: J0 and Grad0 from one forward +
    adjoint run, then the forward run re-done for batch control amplitudes
    A_k = 10^(-2 - 0.25 k) along the adjoint gradient, all amplitudes as the instances
    of one region (one launch per RK stage), against the same re-runs one after
    another.  Reports the relative error |(J(A_k) - J0) / A_k - Grad0| / Grad0 per
    amplitude."""
    from magudi_amd import cases
    from magudi_amd.solver import Region, Solver
    c = cases.acoustic_monopole(n=n, nsteps=nsteps, save_interval=save_interval)
    q0 = c.initial_state()
    R = Region(c)
    R.states[0].conservedVariables = q0
    S = Solver(R)
    sens, grad, J = S.runAdjoint()
    J0, g0 = float(J[0]), float(sens[0])
    amps = [10.0 ** (-2.0 - 0.25 * k) for k in range(batch)]
    ctl = np.stack([a * grad[0] for a in amps])
    Rb = Region(c, n_instances=batch)
    for k in range(batch):
        Rb.states[k].initialCondition = q0
    Sb = Solver(Rb)
    Rb.restart()
    Sb.runForward(control=ctl)                      # warm-up (tile autotuning)
    Rb.restart()
    Rb.sync()
    t0 = time.perf_counter()
    Jb = Sb.runForward(control=ctl)
    Rb.sync()
    tb = time.perf_counter() - t0
    # the same batched run replayed as one HIP graph (captured on its first call)
    Rb.restart()
    Sb.runForward(control=ctl, use_graph=True)
    Rb.restart()
    Rb.sync()
    t0 = time.perf_counter()
    Jg = Sb.runForward(control=ctl, use_graph=True)
    Rb.sync()
    tg = time.perf_counter() - t0
    # the same re-runs one after another on the single-instance region
    R.states[0].initialCondition = q0
    Js = []
    R.sync()
    t0 = time.perf_counter()
    for k in range(batch):
        R.restart()
        Js.append(float(S.runForward(control=ctl[k])[0]))
    R.sync()
    ts = time.perf_counter() - t0
    err = [abs((float(Jb[k]) - J0) / amps[k] - g0) / abs(g0) for k in range(batch)]
    return {"workload": f"{batch} forward re-runs of the {n}x{n} monopole ({nsteps} RK4 steps) "
                        f"at control amplitudes 10^(-2 - 0.25 k) along the adjoint gradient",
            "value": batch * c.npoints * nsteps / min(tb, tg), "unit": "grid-pt-steps/s",
            "eager_value": batch * c.npoints * nsteps / tb,
            "graph_value": batch * c.npoints * nsteps / tg,
            "graph_same_J": bool(np.array_equal(Jg, Jb)),
            "sequential_value": batch * c.npoints * nsteps / ts, "speedup": ts / min(tb, tg),
            "J0": J0, "Grad0": g0, "same_J_as_single_runs": bool(all(float(Jb[k]) == Js[k] for k in range(batch))),
            "rel_error": [[amps[k], err[k]] for k in range(batch)]}


def small_grid_leg(n=201, nsteps=96, save_interval=32, batch=20):
    """Batched against plain on a small grid: batched fused forward+adjoint (batch instances
    per launch) against the generic per-instance path, whose many small launches
    make it launch-bound at this size, and against the fused kernels on one instance.
    Every run recomputes its windows from the snapshots (the plain runs with the window
    cache off: on a small grid it would keep every window and skip the recompute the
    fused runs do)."""
    from magudi_amd import cases
    from magudi_amd.solver import Region, Solver
    c = cases.acoustic_monopole(n=n, nsteps=nsteps, save_interval=save_interval)
    q0 = c.initial_state()
    out = {"grid": [n, n], "nsteps": nsteps, "instances": batch}
    for key, ninst, plain, graph in (("batched_fused", batch, False, False),
                                     ("batched_fused_graph", batch, False, True),
                                     ("fused_single", 1, False, False),
                                     ("fused_single_graph", 1, False, True),
                                     ("plain_single", 1, True, False)):
        R = Region(c, n_instances=ninst, force_plain=plain)
        if plain:
            R.set_window_cache(0)   # window recompute, as the fused runs do (see below)
        for i in range(ninst):
            R.states[i].initialCondition = q0
        S = Solver(R)
        run_hot_path(R, S, 1, graph)
        t, _ = run_hot_path(R, S, 1, graph)
        out[key] = ninst * c.npoints * nsteps / t
        del R, S
    best = max(out["batched_fused"], out["batched_fused_graph"])
    # the north-star ratio (batched against the plain per-instance path) and the batching
    # gain alone (against the fused kernels on one instance)
    out["ratio"] = best / out["plain_single"]
    out["ratio_vs_fused_single"] = best / max(out["fused_single"], out["fused_single_graph"])
    return out


def uneven_leg(nsteps=32, save_interval=8):
    """The mixed batch (1-D periodic waves of 2048 points, 2-D monopoles on odd grids
    drawn from [201, 513] with seed 5, 513 x 257 mixing layers): forward + adjoint of
    every member, all regions issued concurrently, against the same regions run one after
    another.  Each mixing layer is a multiple-shooting member: msforward + msadjoint over 4
    time segments of `nsteps` steps (the segments are the instances of its region), the
    segments starting from the states of a baseline run at t = k nsteps dt with a seeded
    momentum perturbation (the matching conditions unmet), quadratic penalty of weight 1."""
    from magudi_amd import capi, cases
    from magudi_amd.group import RegionGroup
    cs = cases.uneven_batch(seed=5, nsteps=nsteps, save_interval=save_interval)
    shooting = {k: dict(nsplit=4, nts=nsteps, penalty_weight=1.0, save_interval=save_interval)
                for k, c in enumerate(cs) if c.name.startswith("mixing")}
    G = RegionGroup(cs, shooting=shooting)
    G.restart()
    G.run_adjoint(want_gradient=True)          # autotune and warm up
    G.restart()
    G.sync()
    t0 = time.perf_counter()
    G.run_adjoint(want_gradient=True)
    G.sync()
    t_group = time.perf_counter() - t0
    G.restart()
    G.sync()
    t0 = time.perf_counter()
    for k, s in enumerate(G.solvers):
        if k in G.shooting:
            G._shooting_run(k, None, True)
        else:
            s.runAdjoint(want_gradient=True)
    G.sync()
    t_seq = time.perf_counter() - t0
    # one launch per RK stage for the 18 fused 2-D members, the 1-D waves beside them
    G.restart()
    G.run_adjoint_merged(want_gradient=True)   # tables and warm-up
    G.restart()
    G.sync()
    t0 = time.perf_counter()
    G.run_adjoint_merged(want_gradient=True)
    G.sync()
    t_merged = time.perf_counter() - t0
    # the same with every merged member's stage states resident (MGD_OPT_RESIDENT_STATES
    # 1 on each): no window recompute on the second stream
    for k in G._merged_members():
        capi.check(capi.lib().mgd_region_set_option(G.regions[k].handle, capi.OPT_RESIDENT_STATES, 1))
    G.restart()
    G.run_adjoint_merged(want_gradient=True)
    G.restart()
    G.sync()
    t0 = time.perf_counter()
    G.run_adjoint_merged(want_gradient=True)
    G.sync()
    t_merged_res = time.perf_counter() - t0
    merged = len(G._merged_members())
    G.close()
    pts = G.grid_point_steps
    return {"workload": f"{len(cs)} regions: 8 x 1-D wave 2048, 16 x 2-D monopole 201..513 "
                        f"(seed 5), 2 x mixing layer 513x257 multiple shooting over 4 segments of "
                        f"{nsteps} steps; fwd+adjoint {nsteps} steps per member or segment, "
                        f"snapshots every {save_interval}",
            "shooting_penalty": [float(G.shooting[k][0].J_penalty) for k in sorted(G.shooting)],
            "value": pts / min(t_group, t_merged, t_merged_res), "unit": "grid-pt-steps/s",
            "concurrent_value": pts / t_group, "merged_value": pts / t_merged,
            "merged_resident_value": pts / t_merged_res,
            "merged_members": merged, "sequential_value": pts / t_seq,
            "concurrency_speedup": t_seq / t_group, "merged_speedup": t_seq / min(t_merged, t_merged_res)}


def small_ensemble_leg(count=32, nsteps=32, save_interval=8, seed=3):
    """An ensemble of small 2-D monopoles (odd sizes drawn from [41, 129], seed 3):
    forward + adjoint of every member with one launch per RK stage for all of them,
    against the members issued concurrently and one after another."""
    import numpy as np
    from magudi_amd import cases
    from magudi_amd.group import RegionGroup
    from magudi_amd.solver import Region, Solver
    rng = np.random.default_rng(seed)
    ns = [int(2 * rng.integers(20, 65) + 1) for _ in range(count)]
    G = RegionGroup([cases.acoustic_monopole(n=n, nsteps=nsteps, save_interval=save_interval)
                     for n in ns])

    def timed(fn):
        G.restart()
        G.sync()
        t0 = time.perf_counter()
        fn()
        G.sync()
        return time.perf_counter() - t0
    G.restart()
    G.run_adjoint_merged(want_gradient=True)
    t_m = timed(lambda: G.run_adjoint_merged(want_gradient=True))
    G.run_adjoint(want_gradient=True)
    t_c = timed(lambda: G.run_adjoint(want_gradient=True))
    t_s = timed(lambda: [s.runAdjoint(want_gradient=True) for s in G.solvers])
    [s.runAdjoint(want_gradient=True, use_graph=True) for s in G.solvers]
    t_g = timed(lambda: [s.runAdjoint(want_gradient=True, use_graph=True) for s in G.solvers])
    G.close()
    # the members one after another on the generic per-operator kernels (the plain
    # per-instance path), fewer steps at the same per-step work
    ns_plain = max(save_interval, nsteps // 4)
    plain_pts = 0
    t_p = 0.0
    for n in ns:
        c = cases.acoustic_monopole(n=n, nsteps=ns_plain, save_interval=save_interval)
        Rp = Region(c, force_plain=True)
        Rp.set_window_cache(0)      # window recompute, as the fused runs do
        Rp.states[0].initialCondition = c.initial_state()
        Sp = Solver(Rp)
        run_hot_path(Rp, Sp, 1)
        t, _ = run_hot_path(Rp, Sp, 1)
        t_p += t
        plain_pts += c.npoints * ns_plain
        del Rp, Sp
    pts = G.grid_point_steps
    return {"workload": f"{count} monopoles, odd sizes in [41, 129] (seed {seed}); fwd+adjoint "
                        f"{nsteps} steps, snapshots every {save_interval}",
            "merged_value": pts / t_m, "concurrent_value": pts / t_c, "sequential_value": pts / t_s,
            "sequential_graph_value": pts / t_g, "plain_sequential_value": plain_pts / t_p,
            "unit": "grid-pt-steps/s", "merged_speedup_vs_sequential": t_s / t_m,
            "merged_speedup_vs_sequential_graph": t_g / t_m,
            "merged_speedup_vs_plain": (pts / t_m) / (plain_pts / t_p)}


def monopole_legs(a, extras):
    """The 2-D acoustic monopole (401^2, 960 RK4 steps, snapshots every 32) forward + adjoint
    with the control gradient on the fused stage kernels, its dominant kernel against the
    HBM roofline, and (extras) the batched, generic, resident-state, sweep, uneven-batch,
    ensemble and small-grid legs."""
    from magudi_amd import capi, cases
    from magudi_amd.solver import Region, Solver
    case = cases.acoustic_monopole(n=a.n, nsteps=a.nsteps, save_interval=a.save_interval)
    q0 = case.initial_state()
    R = Region(case, tile_shape=a.tile_shape)
    R.states[0].initialCondition = q0
    S = Solver(R)
    run_hot_path(R, S, 1)
    t_step, (sens, _, J) = run_hot_path(R, S, 2)
    pts = case.npoints * a.nsteps
    value = pts / t_step
    plan = R.checkpoint_plan
    resident = plan["mode"] == "fused" and plan["held"] > 0
    step_store = plan["mode"] == "step store"
    # compulsory bytes per grid-point step: 4 x 320 with every stage state resident, 4 x 448
    # with windows (each stage recomputed once, 128 B), 4 x 320 + 3 x 128 per step not held
    # on the step store
    job_b = (4 * 320 + 3 * 128 * (a.nsteps - plan["held"]) / a.nsteps if step_store
             else 4 * (320 if resident else 448))
    R.restart()
    fwd_ms, fwd_b = R.time_kernel(capi.KERNEL_FWD_STAGE, reps=50)
    adj_ms, adj_b = R.time_kernel(capi.KERNEL_ADJ_STAGE, reps=50)
    out = {"workload": f"monopole {a.n}x{a.n} forward+adjoint with the control gradient, "
                       f"{a.nsteps} RK4 steps, " +
                       ("every RK stage state resident in HBM" if resident else
                        f"the state of every step and the stage states of {plan['held']} steps "
                        f"in HBM, the others recomputed" if step_store else
                        f"HBM snapshots every {a.save_interval} steps with recompute"),
           "checkpoint_plan": plan,
           "value": value, "unit": "grid-pt-steps/s", "ms_per_run": t_step * 1e3,
           "path": "fused 2-D stage kernels",
           "tile_shape": ["32x16", "16x32", "16x16", "24x16", "16x16/2", "16x12/2"][R.tile_shape]
           if R.tile_shape >= 0 else "generic",
           "kernels": {k: {"avg_launch_us": round(ms * 1e3, 3), "bytes_per_launch": b,
                           "achieved_GBs": round(b / (ms * 1e-3) / 1e9, 1),
                           "frac": round(b / (ms * 1e-3) / 1e9 / HBM_PEAK_GBS, 4)}
                       for k, ms, b in (("k_fwd_stage2d", fwd_ms, fwd_b),
                                        ("k_adj_stage2d", adj_ms, adj_b))},
           # whole run against the compulsory-traffic model of one fwd+adjoint pt-step, per
           # state component: forward RK4 128 B, storing its 4 stage states 32 B, adjoint
           # RK4 128 B, reading the 4 forward stage states 32 B, and with window recompute
           # the recompute's RK4 128 B
           "job_roofline": {"bytes_per_pt_step": job_b,
                            "achieved_GBs": round(value * job_b / 1e9, 1),
                            "frac": round(value * job_b / 1e9 / HBM_PEAK_GBS, 4)},
           "J": float(J[0]), "Grad0": float(sens[0])}
    if not extras:
        return out
    Rb = Region(case, n_instances=a.batch)
    for i in range(a.batch):
        Rb.states[i].initialCondition = q0
    Sb = Solver(Rb)
    run_hot_path(Rb, Sb, 1)
    tb, _ = run_hot_path(Rb, Sb, 1)
    run_hot_path(Rb, Sb, 1, True)
    tbg, (sens_b, _, J_b) = run_hot_path(Rb, Sb, 1, True)
    plan_b = Rb.checkpoint_plan
    Rb.close()
    del Rb, Sb
    # the same batch with snapshots and window recompute in place of the step store
    Rw = Region(case, n_instances=a.batch)
    Rw.set_step_store(0)
    for i in range(a.batch):
        Rw.states[i].initialCondition = q0
    Sw = Solver(Rw)
    run_hot_path(Rw, Sw, 1)
    tbw, (sens_w, _, J_w) = run_hot_path(Rw, Sw, 1)
    Rw.close()
    del Rw, Sw
    # generic per-operator path on one instance, fewer steps, same per-step work
    ns_plain = max(a.save_interval, a.nsteps // 10)
    Rp = Region(case, force_plain=True)
    Rp.set_window_cache(0)          # window recompute, as the fused runs do
    Rp.states[0].initialCondition = q0
    Sp = Solver(Rp, nsteps=ns_plain)
    run_hot_path(Rp, Sp, 1)
    tp, _ = run_hot_path(Rp, Sp, 1)
    del Rp, Sp
    out["batched"] = {"instances": a.batch, "value": a.batch * pts / tb, "ms_per_run": tb * 1e3,
                      "graph_value": a.batch * pts / tbg, "graph_ms_per_run": tbg * 1e3,
                      "speedup_vs_single_fused": (a.batch * pts / min(tb, tbg)) / value,
                      "checkpoint_plan": plan_b,
                      "windows_value": a.batch * pts / tbw, "windows_ms_per_run": tbw * 1e3,
                      "step_store_speedup_vs_windows": tbw / tb,
                      "same_J_and_Grad0_as_windows": bool(np.array_equal(J_b, J_w) and
                                                          np.array_equal(sens_b, sens_w))}
    out["plain"] = {"value": case.npoints * ns_plain / tp, "nsteps": ns_plain,
                    "ms_per_run": tp * 1e3,
                    "batched_speedup_vs_plain": (a.batch * pts / tb) / (case.npoints * ns_plain / tp)}
    # snapshots every save_interval steps + window recompute (each recompute stage sharing
    # a launch with an adjoint stage) instead of every RK stage state resident in HBM
    # (19.8 GB at 401^2 x 960 steps): same results
    Rr = Region(case, tile_shape=R.tile_shape, resident_states=False)
    Rr.set_step_store(0)
    Rr.states[0].initialCondition = q0
    Sr = Solver(Rr)
    run_hot_path(Rr, Sr, 1)
    tr, (sens_r, _, J_r) = run_hot_path(Rr, Sr, 1)
    del Rr, Sr
    out["window_recompute"] = {"value": pts / tr, "ms_per_run": tr * 1e3,
                               "same_J_and_Grad0": bool(J_r[0] == J[0] and sens_r[0] == sens[0])}
    return out


def batched_leg(a):
    """The batched monopole run alone (a.batch instances of 401^2, 960 steps, with the
    control gradient): one warm-up run, one timed; its storage plan."""
    from magudi_amd import cases
    from magudi_amd.solver import Region, Solver
    case = cases.acoustic_monopole(n=a.n, nsteps=a.nsteps, save_interval=a.save_interval)
    q0 = case.initial_state()
    Rb = Region(case, n_instances=a.batch)
    for i in range(a.batch):
        Rb.states[i].initialCondition = q0
    Sb = Solver(Rb)
    run_hot_path(Rb, Sb, 1)
    tb, _ = run_hot_path(Rb, Sb, 1)
    out = {"instances": a.batch, "value": a.batch * case.npoints * a.nsteps / tb,
           "ms_per_run": tb * 1e3, "checkpoint_plan": Rb.checkpoint_plan,
           "tile_shape": Rb.tile_shape}
    Rb.close()
    return out


def checkpoint_plans_leg(n, nsteps, t_step, J):
    """The headline pass with snapshots every nsteps / 8 steps and windows (the window cache
    holding the last windows free memory allows) in place of the step store: same J, the
    time against the step store's."""
    K = KolmogorovRun(n, nsteps, step_store=0)
    K.one()
    t, Jw, _, launches, _, _ = K.timed(1)
    plan = K.R.checkpoint_plan
    K.R.close()
    del K
    return {"step_store_s_per_pass": round(t_step, 3), "windows_s_per_pass": round(t, 3),
            "windows_plan": plan, "windows_march_launches_per_pass": launches,
            "same_J": bool(Jw == J), "speedup": round(t / t_step, 3)}


def kernels_only(a):
    """HIP-event timing of the headline's dominant kernel and of the state copy of the same
    region (the calibration the counter passes scale by)."""
    from magudi_amd import capi
    K = KolmogorovRun(a.c4_n, 8)
    K.R.restart()
    res = {}
    for name, k in (("k_fwd3d_march", capi.KERNEL_FWD_MARCH3D), ("k_copy", capi.KERNEL_COPY)):
        ms, b = K.R.time_kernel(k, reps=3)
        res[name] = {"avg_launch_us": ms * 1e3, "bytes_per_launch": b}
    return res


def main():
    a = parse()
    if a.gpus != 1:
        print(f"bench.py measures one GPU only; --gpus {a.gpus} is not supported "
              "(batched instances and ensemble members would shard across GPUs, one "
              "process per GPU, without communication; a single large grid as z slabs with "
              "a halo exchange per stage)", file=sys.stderr)
        return 2
    from magudi_amd import capi

    if a.kolmogorov_only:
        a.no_extras = a.no_traffic = True
    if a.viscous3d_only:
        print(json.dumps({"viscous3d": viscous3d_leg(a.c4_n)}), flush=True)
        return 0
    if a.fd_sweep_only:
        print(json.dumps({"fd_sweep": fd_sweep_leg(a.n, a.nsteps, a.save_interval, a.batch)}),
              flush=True)
        return 0
    if a.monopole_only:
        print(json.dumps({"monopole401": monopole_legs(a, extras=False)}), flush=True)
        return 0
    if a.batched_only:
        print(json.dumps({"batched": batched_leg(a)}), flush=True)
        return 0
    if a.kernels_only:
        print(json.dumps({"kernels_only": kernels_only(a)}), flush=True)
        return 0
    if a.kernels_only_2d:
        from magudi_amd import cases
        from magudi_amd.solver import Region
        case = cases.acoustic_monopole(n=a.n, nsteps=a.nsteps, save_interval=a.save_interval)
        R = Region(case, tile_shape=a.tile_shape)
        R.states[0].initialCondition = case.initial_state()
        R.restart()
        res = {}
        for name, k in (("k_fwd_stage2d", capi.KERNEL_FWD_STAGE),
                        ("k_adj_stage2d", capi.KERNEL_ADJ_STAGE), ("k_copy", capi.KERNEL_COPY)):
            ms, b = R.time_kernel(k, reps=5)
            res[name] = {"avg_launch_us": ms * 1e3, "bytes_per_launch": b}
        res["tile_shape"] = R.tile_shape
        print(json.dumps({"kernels_only": res}), flush=True)
        return 0

    # ---- headline: the 3-D Kolmogorov forward + adjoint pass
    K = KolmogorovRun(a.c4_n, a.c4_nsteps)
    for _ in range(max(1, a.warmup)):
        K.one()
    t_step, J, t_fwd, launches, with_copy, march_ms = K.timed(a.steps)
    march_shape = K.R.march_shape
    plan = K.R.checkpoint_plan
    value = K.pts / t_step
    print(f"headline: {t_step:.3f} s per pass ({value:.4g} grid-pt-steps/s)", file=sys.stderr,
          flush=True)
    # dominant kernel (by launches x time: 4 nsteps forward launches per pass plus those of
    # the recompute, against 4 nsteps launches of each adjoint kernel), timed with
    # HIP events around each of its launches in the timed passes.  Its algorithmic bytes
    # per launch: the stage state and the 9 stress / heat-flux slots in, the RK buffers
    # and the new stage state (0, 2, 2, 1 states in; 2, 2, 2, 1 out, over the 4 stages:
    # stages 2 and 3 read the state at the start of the step where stage 1 read it) --
    # 232 B per point -- plus 40 B per point for the launches that also write the state
    # into the step store, the stage cache or the recompute buffer.
    K.R.restart()
    cms, cb = K.R.time_kernel(capi.KERNEL_COPY, reps=3)
    f_ms, f_b = K.R.time_kernel(capi.KERNEL_RHS_FORWARD, reps=3)
    a_ms, a_b = K.R.time_kernel(capi.KERNEL_RHS_ADJOINT, reps=3)
    M_bytes = 8.0 * 5 * K.case.npoints
    K.R.close()                 # its step store holds most of the device memory
    del K
    ms = march_ms / max(launches, 1)
    b = (232.0 + 40.0 * with_copy / max(launches, 1)) * M_bytes / 40.0
    achieved = b / (ms * 1e-3) / 1e9
    traffic = None
    if not a.no_traffic:
        tr = measured_traffic(["--kernels-only", "--c4-n", str(a.c4_n)], "k_fwd3d_march", M_bytes)
        traffic = tr["bytes"] if tr else None
    roofline = {"bound": "hbm", "kernel": "k_fwd3d_march", "achieved": round(achieved, 1),
                "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4),
                "traffic": traffic, "avg_launch_us": round(ms * 1e3, 2), "bytes_per_launch": b,
                "launches_timed": launches, "launches_with_state_copy": with_copy}
    n = a.c4_n
    out = {
        "metric": METRIC, "value": value, "unit": "grid-pt-steps/s", "n_gpus": 1,
        "steps": a.steps, "warmup": a.warmup, "ms_per_step": t_step * 1e3,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64",
        "data": "synthetic (Kolmogorov flow set-up and seeded solenoidal field generated in "
                "magudi_amd/cases.py)",
        "config": {"workload": f"Kolmogorov {n}^3 viscous periodic forward + adjoint, "
                               f"{a.c4_nsteps} RK4 steps, the state of every step and the "
                               f"stage states of the last {plan['held']} steps held in HBM, "
                               f"the other steps' stages recomputed, terminal kinetic energy "
                               f"and dJ/dq(T) on the device",
                   "checkpoint_plan": plan,
                   "grid": [n, n, n], "nsteps": a.c4_nsteps,
                   "save_interval": max(1, a.c4_nsteps // 8), "instances": 1,
                   "march_shape": {"forward": list(march_shape[0]),
                                   "adjoint": list(march_shape[1])},
                   "path": "3-D kernels: stress pass + z-march forward; adjoint stress pass + "
                           "Jacobian march + finishing march"},
        "roofline": roofline,
        "J": J, "forward_s": round(t_fwd, 3), "adjoint_s": round(t_step - t_fwd, 3),
        "rhs3d": {"forward_ms": round(f_ms, 4), "adjoint_ms": round(a_ms, 4),
                  "forward_frac_80B": round(f_b / (f_ms * 1e-3) / 1e9 / HBM_PEAK_GBS, 4),
                  "adjoint_frac_120B": round(a_b / (a_ms * 1e-3) / 1e9 / HBM_PEAK_GBS, 4)},
        "state_copy": {"avg_launch_us": round(cms * 1e3, 2),
                       "achieved_GBs": round(cb / (cms * 1e-3) / 1e9, 1)},
    }
    if traffic is not None:
        # the counter passes run the kernel's launches without the state copy (232 B / pt)
        out["roofline"]["traffic_over_algorithmic"] = round(traffic / (232.0 * M_bytes / 40.0), 3)
        out["roofline"]["traffic_workload"] = "launches without the state copy (--kernels-only)"
    if not a.no_extras:
        out["checkpoint_plans"] = checkpoint_plans_leg(a.c4_n, a.c4_nsteps, t_step, J)
        out["monopole401"] = monopole_legs(a, extras=True)
        out["fd_sweep"] = fd_sweep_leg(a.n, a.nsteps, a.save_interval, a.batch)
        out["uneven_batch"] = uneven_leg()
        out["small_ensemble"] = small_ensemble_leg()
        out["small_grid"] = small_grid_leg()
        out["cpu_baseline"] = cpu_baseline()
    else:
        out["cpu_baseline"] = None
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
