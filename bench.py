#!/usr/bin/env python3
"""Benchmark of the magudi RHS hot path on one MI355X.

Headline workload (one "step" = one full pass of the hot path): the 2-D acoustic
monopole on a 401 x 401 grid, inviscid Euler, SBP 4-8 with dissipation, far-field
SAT + sponges, forward run of 960 RK4 steps followed by the discrete adjoint over
the same 960 steps with HBM snapshots every 32 steps and window recompute
(Solver%runAdjoint with a uniform checkpointer).  Synthetic input: the quiescent
state with the Gaussian monopole source and a control mollifier, all built in
magudi_amd/cases.py.

value = grid-point-steps per second of the forward+adjoint pass (N * 960 / time).
Also reported: the 20-instance batched rate (finite-difference sweep shape), the
generic per-operator ("plain") path, the HBM roofline fraction of the dominant
kernel timed with HIP events on its own stream, and the CPU oracle on a bounded
sample.
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
    ap.add_argument("--n", type=int, default=401)
    ap.add_argument("--nsteps", type=int, default=960)
    ap.add_argument("--save-interval", type=int, default=32)
    ap.add_argument("--batch", type=int, default=20)
    ap.add_argument("--no-extras", action="store_true",
                    help="headline only (no batched / plain / CPU legs)")
    ap.add_argument("--cpu-steps", type=int, default=64)
    ap.add_argument("--c4-n", type=int, default=256,
                    help="grid of the 3-D viscous Kolmogorov rhs leg (0: skip)")
    ap.add_argument("--fd-sweep-only", action="store_true",
                    help="only the batched finite-difference sweep over control amplitudes")
    ap.add_argument("--kolmogorov-only", action="store_true",
                    help="only the 3-D Kolmogorov forward + adjoint run (c4-n^3, 200 steps)")
    ap.add_argument("--viscous3d-only", action="store_true",
                    help="only the 3-D viscous rhs leg (its profiling workload)")
    ap.add_argument("--tile-shape", type=int, default=-1,
                    help="tile shape of the fused 2-D kernels (-1: autotune)")
    ap.add_argument("--kernels-only", action="store_true",
                    help="only the HIP-event timing of the stage kernels and the copy kernel "
                         "(the workload of the rocprofv3 counter runs)")
    return ap.parse_args()


def measured_traffic():
    """HBM bytes per launch from the committed rocprofv3 counter summary, or {}.

    Lines of the form `traffic <kernel> <bytes per launch>` in
    profiles/pmc_traffic_monopole401.txt (FETCH_SIZE + WRITE_SIZE, scaled by the
    copy-kernel calibration factor written in the same file)."""
    path = os.path.join(ROOT, "profiles", "pmc_traffic_monopole401.txt")
    out = {}
    if os.path.exists(path):
        for line in open(path):
            f = line.split()
            if len(f) == 3 and f[0] == "traffic":
                out[f[1]] = float(f[2])
    return out


def run_hot_path(R, S, steps):
    """steps x (restart from the resident initial condition + forward + adjoint)."""
    R.sync()
    t0 = time.perf_counter()
    for _ in range(steps):
        R.restart()
        out = S.runAdjoint(want_gradient=False)
    R.sync()
    return (time.perf_counter() - t0) / steps, out


def cpu_baseline(n, nsteps, save_interval):
    """CPU oracle (plain C; its per-point and per-pencil loops on every OpenMP thread the
    host gives, OMP_NUM_THREADS) on a bounded sample of the same workload."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import oracle_bind as ob
    from magudi_amd import cases
    c = cases.acoustic_monopole(n=n, nsteps=nsteps, save_interval=save_interval)
    O = ob.Region(c)
    q0 = c.initial_state()
    threads = ob.lib().orc_threads()
    t0 = time.perf_counter()
    O.run_adjoint(q0, c.dt, nsteps, save_interval)
    dt = time.perf_counter() - t0
    return {"value": c.npoints * nsteps / dt, "unit": "grid-pt-steps/s", "cores": threads,
            "kind": "port",
            "sample": f"{n}x{n} monopole forward+adjoint, {nsteps} RK4 steps, snapshots every "
                      f"{save_interval} (oracle/ C restatement, {threads} OpenMP threads, "
                      f"{dt:.1f} s wall)"}


def viscous3d_leg(n):
    """One forward and one adjoint rhs evaluation of the 3-D viscous periodic Kolmogorov
    flow at n^3, HIP-event timed (per-operator kernels: one-pass periodic gradient,
    LDS-staged flux divergence, line-march derivatives)."""
    from magudi_amd import capi, cases
    from magudi_amd.solver import Region
    c = cases.kolmogorov3d(n=n, nsteps=1)
    R = Region(c)
    R.states[0].conservedVariables = c.initial_state()
    f_ms, f_b = R.time_kernel(capi.KERNEL_RHS_FORWARD, reps=5)
    a_ms, a_b = R.time_kernel(capi.KERNEL_RHS_ADJOINT, reps=5)
    del R
    leg = {"workload": f"Kolmogorov {n}^3 viscous periodic rhs (Re_ac 500, Pr 0.7)",
           "path": "3-D per-operator kernels (periodic gradient / flux-divergence / line derivatives)"}
    for k, ms, b in (("forward_rhs", f_ms, f_b), ("adjoint_rhs", a_ms, a_b)):
        gbs = b / (ms * 1e-3) / 1e9
        leg[k] = {"ms": round(ms, 4), "algorithmic_bytes": b, "achieved_GBs": round(gbs, 1),
                  "frac": round(gbs / HBM_PEAK_GBS, 4), "pt_per_s": n ** 3 / (ms * 1e-3)}
    return leg


def kolmogorov_run_leg(n, nsteps=200):
    """The 3-D viscous periodic Kolmogorov flow as the configuration runs it: nsteps RK4
    steps forward with HBM snapshots (every nsteps / 8), J = kinetic energy of the final
    state, then the discrete adjoint back from dJ/dq(T) with window recompute.  A short
    run of one window first allocates the buffers; then one timed run."""
    from magudi_amd import cases
    from magudi_amd.functionals import kinetic_energy
    from magudi_amd.solver import Region, Solver
    c = cases.kolmogorov3d(n=n, nsteps=nsteps)
    q0 = c.initial_state()
    R = Region(c)
    norm = R.grid(2)[0]

    def run(steps):
        S = Solver(R, dt=c.dt, nsteps=steps, save_interval=c.save_interval)
        R.states[0].conservedVariables = q0
        R.sync()
        t0 = time.perf_counter()
        S.runForwardCheckpointed()
        R.sync()
        t1 = time.perf_counter()
        J, lamT = kinetic_energy(R.states[0].conservedVariables, norm, 3)
        t2 = time.perf_counter()
        S.runAdjointCheckpointed(terminal=lamT, want_gradient=False)
        R.sync()
        t3 = time.perf_counter()
        return J, t1 - t0, t3 - t2, t2 - t1

    run(c.save_interval)
    J, tf, ta, th = run(nsteps)
    lam_max = float(abs(R.states[0].adjointVariables).max())
    del R
    t = tf + ta
    return {"workload": f"Kolmogorov {n}^3 viscous periodic, {nsteps} RK4 steps forward + "
                        f"adjoint from the terminal kinetic energy, snapshots every "
                        f"{c.save_interval} steps with recompute",
            "value": n ** 3 * nsteps / t, "unit": "grid-pt-steps/s", "seconds": round(t, 3),
            "forward_s": round(tf, 3), "adjoint_s": round(ta, 3),
            "host_terminal_s": round(th, 3), "J": J, "max_abs_lambda0": lam_max}


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
            "value": batch * c.npoints * nsteps / tb, "unit": "grid-pt-steps/s",
            "sequential_value": batch * c.npoints * nsteps / ts, "speedup": ts / tb,
            "J0": J0, "Grad0": g0, "same_J_as_single_runs": bool(all(float(Jb[k]) == Js[k] for k in range(batch))),
            "rel_error": [[amps[k], err[k]] for k in range(batch)]}


def small_grid_leg(n=201, nsteps=96, save_interval=32, batch=20):
    """Batched against plain on a small grid: batched fused forward+adjoint (batch instances
    per launch) against the generic per-instance path, whose many small launches
    make it launch-bound at this size."""
    from magudi_amd import cases
    from magudi_amd.solver import Region, Solver
    c = cases.acoustic_monopole(n=n, nsteps=nsteps, save_interval=save_interval)
    q0 = c.initial_state()
    out = {"grid": [n, n], "nsteps": nsteps, "instances": batch}
    for key, ninst, plain in (("batched_fused", batch, False), ("plain_single", 1, True)):
        R = Region(c, n_instances=ninst, force_plain=plain)
        for i in range(ninst):
            R.states[i].initialCondition = q0
        S = Solver(R)
        run_hot_path(R, S, 1)
        t, _ = run_hot_path(R, S, 1)
        out[key] = ninst * c.npoints * nsteps / t
    out["ratio"] = out["batched_fused"] / out["plain_single"]
    return out


def uneven_leg(nsteps=32, save_interval=8):
    """The mixed batch (1-D periodic waves of 2048 points, 2-D monopoles on odd grids
    drawn from [201, 513] with seed 5, 513 x 257 mixing layers with 4 segments each as
    instances): forward + adjoint of every member, all regions issued concurrently,
    against the same regions run one after another."""
    from magudi_amd import cases
    from magudi_amd.group import RegionGroup
    cs = cases.uneven_batch(seed=5, nsteps=nsteps, save_interval=save_interval)
    ninst = [4 if c.name.startswith("mixing") else 1 for c in cs]
    G = RegionGroup(cs, n_instances=ninst)
    G.restart()
    G.run_adjoint(want_gradient=False)          # autotune and warm up
    G.restart()
    G.sync()
    t0 = time.perf_counter()
    G.run_adjoint(want_gradient=False)
    G.sync()
    t_group = time.perf_counter() - t0
    G.restart()
    G.sync()
    t0 = time.perf_counter()
    for s in G.solvers:
        s.runAdjoint(want_gradient=False)
    G.sync()
    t_seq = time.perf_counter() - t0
    # one launch per RK stage for the 18 fused 2-D members, the 1-D waves beside them
    G.restart()
    G.run_adjoint_merged(want_gradient=False)   # tables and warm-up
    G.restart()
    G.sync()
    t0 = time.perf_counter()
    G.run_adjoint_merged(want_gradient=False)
    G.sync()
    t_merged = time.perf_counter() - t0
    merged = len(G._merged_members())
    G.close()
    pts = G.grid_point_steps
    return {"workload": f"{len(cs)} regions: 8 x 1-D wave 2048, 16 x 2-D monopole 201..513 "
                        f"(seed 5), 2 x mixing layer 513x257 x 4 segments; fwd+adjoint "
                        f"{nsteps} steps, snapshots every {save_interval}",
            "value": pts / min(t_group, t_merged), "unit": "grid-pt-steps/s",
            "concurrent_value": pts / t_group, "merged_value": pts / t_merged,
            "merged_members": merged, "sequential_value": pts / t_seq,
            "concurrency_speedup": t_seq / t_group, "merged_speedup": t_seq / t_merged}


def small_ensemble_leg(count=32, nsteps=32, save_interval=8, seed=3):
    """An ensemble of small 2-D monopoles (odd sizes drawn from [41, 129], seed 3):
    forward + adjoint of every member with one launch per RK stage for all of them,
    against the members issued concurrently and one after another."""
    import numpy as np
    from magudi_amd import cases
    from magudi_amd.group import RegionGroup
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
    G.run_adjoint_merged(want_gradient=False)
    t_m = timed(lambda: G.run_adjoint_merged(want_gradient=False))
    G.run_adjoint(want_gradient=False)
    t_c = timed(lambda: G.run_adjoint(want_gradient=False))
    t_s = timed(lambda: [s.runAdjoint(want_gradient=False) for s in G.solvers])
    G.close()
    pts = G.grid_point_steps
    return {"workload": f"{count} monopoles, odd sizes in [41, 129] (seed {seed}); fwd+adjoint "
                        f"{nsteps} steps, snapshots every {save_interval}",
            "merged_value": pts / t_m, "concurrent_value": pts / t_c, "sequential_value": pts / t_s,
            "unit": "grid-pt-steps/s", "merged_speedup_vs_sequential": t_s / t_m}


def main():
    a = parse()
    if a.gpus != 1:
        print(f"bench.py measures one GPU only; --gpus {a.gpus} is not supported "
              "(batched instances and ensemble members would shard across GPUs, one "
              "process per GPU, without communication)", file=sys.stderr)
        return 2
    from magudi_amd import capi, cases
    from magudi_amd.solver import Region, Solver

    if a.viscous3d_only:
        print(json.dumps({"viscous3d": viscous3d_leg(a.c4_n)}), flush=True)
        return 0
    if a.kolmogorov_only:
        print(json.dumps({"kolmogorov_run": kolmogorov_run_leg(a.c4_n)}), flush=True)
        return 0
    if a.fd_sweep_only:
        print(json.dumps({"fd_sweep": fd_sweep_leg(a.n, a.nsteps, a.save_interval, a.batch)}),
              flush=True)
        return 0
    case = cases.acoustic_monopole(n=a.n, nsteps=a.nsteps, save_interval=a.save_interval)
    q0 = case.initial_state()
    R = Region(case, tile_shape=a.tile_shape)
    R.states[0].initialCondition = q0
    if a.kernels_only:
        R.restart()
        res = {}
        for name, k in (("k_fwd_stage2d", capi.KERNEL_FWD_STAGE),
                        ("k_adj_stage2d", capi.KERNEL_ADJ_STAGE), ("k_copy", capi.KERNEL_COPY)):
            ms, b = R.time_kernel(k, reps=5)
            res[name] = {"avg_launch_us": ms * 1e3, "bytes_per_launch": b}
        res["tile_shape"] = R.tile_shape
        # calibration copy larger than the L2 of all XCDs (a.batch states)
        Rb = Region(case, n_instances=a.batch)
        for i in range(a.batch):
            Rb.states[i].initialCondition = q0
        Rb.restart()
        ms, b = Rb.time_kernel(capi.KERNEL_COPY, reps=5)
        res["k_copy_batched"] = {"avg_launch_us": ms * 1e3, "bytes_per_launch": b}
        print(json.dumps({"kernels_only": res}), flush=True)
        return 0
    S = Solver(R)
    for _ in range(a.warmup):
        run_hot_path(R, S, 1)
    t_step, (sens, _, J) = run_hot_path(R, S, a.steps)
    pts = case.npoints * a.nsteps
    value = pts / t_step

    # dominant kernel: HIP-event timing of the fused stage kernels on the region stream
    # (a full RK4 step of 4 stage launches per repetition; bytes = average per launch)
    R.restart()
    fwd_ms, fwd_b = R.time_kernel(capi.KERNEL_FWD_STAGE, reps=50)
    adj_ms, adj_b = R.time_kernel(capi.KERNEL_ADJ_STAGE, reps=50)
    cpy_ms, cpy_b = R.time_kernel(capi.KERNEL_COPY, reps=50)
    launches_fwd = 8 * a.nsteps          # forward pass + window recompute
    launches_adj = 4 * a.nsteps
    kernels = {
        "k_fwd_stage2d": {"ms": fwd_ms, "bytes": fwd_b, "launches": launches_fwd},
        "k_adj_stage2d": {"ms": adj_ms, "bytes": adj_b, "launches": launches_adj},
    }
    dom = max(kernels, key=lambda k: kernels[k]["ms"] * kernels[k]["launches"])
    d = kernels[dom]
    achieved = d["bytes"] / (d["ms"] * 1e-3) / 1e9
    traffic = measured_traffic().get(dom) if a.n == 401 else None
    roofline = {"bound": "hbm", "kernel": dom, "achieved": round(achieved, 1),
                "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4),
                "traffic": traffic, "avg_launch_us": round(d["ms"] * 1e3, 3),
                "bytes_per_launch": d["bytes"]}
    other = {k: {"avg_launch_us": round(v["ms"] * 1e3, 3),
                 "achieved_GBs": round(v["bytes"] / (v["ms"] * 1e-3) / 1e9, 1)}
             for k, v in kernels.items()}
    other["state_copy"] = {"avg_launch_us": round(cpy_ms * 1e3, 3),
                           "achieved_GBs": round(cpy_b / (cpy_ms * 1e-3) / 1e9, 1)}
    kernel_share = sum(v["ms"] * v["launches"] for v in kernels.values()) / (t_step * 1e3)

    out = {
        "metric": METRIC, "value": value, "unit": "grid-pt-steps/s", "n_gpus": 1,
        "steps": a.steps, "warmup": a.warmup, "ms_per_step": t_step * 1e3,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64",
        "data": "synthetic (acoustic monopole set-up generated in magudi_amd/cases.py)",
        "config": {"workload": f"monopole {a.n}x{a.n} forward+adjoint, {a.nsteps} RK4 steps, "
                               f"HBM snapshots every {a.save_interval} steps with recompute",
                   "grid": [a.n, a.n], "nsteps": a.nsteps, "save_interval": a.save_interval,
                   "instances": 1, "path": "fused 2-D stage kernels",
                   "tile_shape": ["32x16", "16x32", "16x16", "24x16", "16x16/2", "16x12/2"][R.tile_shape]
                   if R.tile_shape >= 0 else "generic"},
        "roofline": roofline,
        # whole job against the compulsory-traffic model of one fwd+adjoint pt-step with
        # window recompute, per state component: forward RK4 128 B, adjoint RK4 128 B,
        # reading the 4 forward stage states 32 B, recompute 128 B, storing its stages 32 B
        "job_roofline": {"bytes_per_pt_step": 448 * 4,
                         "achieved_GBs": round(value * 448 * 4 / 1e9, 1),
                         "frac": round(value * 448 * 4 / 1e9 / HBM_PEAK_GBS, 4)},
        "kernels": other,
        "kernel_time_share": round(kernel_share, 3),
        "J": float(J[0]), "Grad0": float(sens[0]),
    }
    if not a.no_extras:
        # batched: a.batch instances of the same grid in every launch
        Rb = Region(case, n_instances=a.batch)
        for i in range(a.batch):
            Rb.states[i].initialCondition = q0
        Sb = Solver(Rb)
        run_hot_path(Rb, Sb, 1)
        tb, _ = run_hot_path(Rb, Sb, 1)
        # plain per-instance path (generic per-operator kernels), fewer steps, same per-step work
        ns_plain = max(a.save_interval, a.nsteps // 10)
        Rp = Region(case, force_plain=True)
        Rp.states[0].initialCondition = q0
        Sp = Solver(Rp, nsteps=ns_plain)
        run_hot_path(Rp, Sp, 1)
        tp, _ = run_hot_path(Rp, Sp, 1)
        out["batched"] = {"instances": a.batch, "value": a.batch * pts / tb,
                          "ms_per_step": tb * 1e3,
                          "speedup_vs_single_fused": (a.batch * pts / tb) / value}
        out["plain"] = {"value": case.npoints * ns_plain / tp, "nsteps": ns_plain,
                        "ms_per_step": tp * 1e3,
                        "batched_speedup_vs_plain": (a.batch * pts / tb) / (case.npoints * ns_plain / tp)}
        # the same run with every RK stage state resident in HBM (19.8 GB at 401^2 x 960
        # steps) instead of snapshots + window recompute: same results, no recompute
        del Rb, Sb
        Rr = Region(case, tile_shape=R.tile_shape, resident_states=True)
        Rr.states[0].initialCondition = q0
        Sr = Solver(Rr)
        run_hot_path(Rr, Sr, 1)
        tr, (sens_r, _, J_r) = run_hot_path(Rr, Sr, a.steps)
        out["resident_states"] = {"value": pts / tr, "ms_per_step": tr * 1e3,
                                  "state_bytes": 8 * 4 * case.npoints * (4 * a.nsteps + 1),
                                  "same_J_and_Grad0": bool(J_r[0] == J[0] and sens_r[0] == sens[0])}
        del Rr, Sr
        if a.c4_n > 0:
            out["viscous3d"] = viscous3d_leg(a.c4_n)
            out["kolmogorov_run"] = kolmogorov_run_leg(a.c4_n)
        out["fd_sweep"] = fd_sweep_leg(a.n, a.nsteps, a.save_interval, a.batch)
        out["uneven_batch"] = uneven_leg()
        out["small_ensemble"] = small_ensemble_leg()
        out["small_grid"] = small_grid_leg()
        out["cpu_baseline"] = cpu_baseline(a.n, a.cpu_steps, max(1, a.cpu_steps // 2))
    else:
        out["cpu_baseline"] = None
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
