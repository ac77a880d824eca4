"""Uneven batches: independent regions of any size and problem type run together.

Two ways, with the same results as each region run alone:

* concurrent runs: each region owns its HIP streams, so the runs of several regions
  overlap on the GPU when they are issued from different host threads (the C ABI calls
  release the Python interpreter lock); each region keeps its own tile shape;
* one launch per RK stage for all members (``run_adjoint_merged``,
  ``mgd_group_run_adjoint``): the fused 2-D members of equal step count march in lock
  step, every launch covering the 16 x 16 tiles of all of them; the other members run
  concurrently afterwards.  Many small grids gain most (launch-bound alone).
"""
from __future__ import annotations

import concurrent.futures as cf
import ctypes as C
from typing import List, Optional, Sequence

import numpy as np

from . import capi
from .cases import Case
from .capi import check, dptr
from .solver import Region, Solver

__all__ = ["RegionGroup"]


class RegionGroup:
    def __init__(self, cases: Sequence[Case], n_instances: Optional[Sequence[int]] = None,
                 force_plain: bool = False, workers: int = 8):
        n_instances = list(n_instances) if n_instances is not None else [1] * len(cases)
        self.cases = list(cases)
        self.regions = [Region(c, n_instances=ni, force_plain=force_plain)
                        for c, ni in zip(self.cases, n_instances)]
        self.solvers = [Solver(r) for r in self.regions]
        for r, c in zip(self.regions, self.cases):
            for i in range(r.n_instances):
                r.states[i].initialCondition = c.initial_state()
        self.workers = max(1, workers)

    @property
    def grid_point_steps(self) -> int:
        return sum(r.npoints * r.n_instances * s.nTimesteps for r, s in zip(self.regions, self.solvers))

    def _map(self, fn):
        with cf.ThreadPoolExecutor(max_workers=min(self.workers, len(self.regions))) as ex:
            return list(ex.map(fn, range(len(self.regions))))

    def restart(self):
        for r in self.regions:
            r.restart()

    def run_adjoint(self, controls: Optional[List[Optional[np.ndarray]]] = None,
                    want_gradient: bool = True):
        """Solver%runAdjoint of every region, concurrently; list of (sens, grad, J)."""
        controls = controls or [None] * len(self.regions)

        def one(k):
            return self.solvers[k].runAdjoint(control=controls[k], want_gradient=want_gradient)
        return self._map(one)

    def _merged_members(self) -> List[int]:
        """Members the one-launch path takes: fused 2-D regions (16 x 16 tiles fit) with
        the step count and snapshot interval of the first of them."""
        idx = [k for k, r in enumerate(self.regions) if r.nd == 2 and r.tile_shape >= 0]
        if not idx:
            return []
        s0 = self.solvers[idx[0]]
        return [k for k in idx if self.solvers[k].nTimesteps == s0.nTimesteps
                and self.solvers[k].saveInterval == s0.saveInterval]

    def _group_handle(self, members: List[int]):
        key = tuple(members)
        if getattr(self, "_group_key", None) != key:
            self.close()
            arr = (C.c_void_p * len(members))(*[self.regions[k].handle.value for k in members])
            h = C.c_void_p()
            check(capi.lib().mgd_group_create(arr, len(members), C.byref(h)))
            self._group, self._group_key = h, key
        return self._group

    def close(self):
        if getattr(self, "_group", None) is not None:
            capi.lib().mgd_group_destroy(self._group)
            self._group = None
            self._group_key = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _run_adjoint_merged(self, members, controls, want_gradient):
        """mgd_group_run_adjoint over `members`: list of (sens, grad, J) per member."""
        s0 = self.solvers[members[0]]
        outs, keep = [], []
        n = len(members)
        dts = np.array([self.solvers[k].dt for k in members], dtype=np.float64)
        ctl = (C.POINTER(C.c_double) * n)()
        grd = (C.POINTER(C.c_double) * n)()
        sen = (C.POINTER(C.c_double) * n)()
        jj = (C.POINTER(C.c_double) * n)()
        for j, k in enumerate(members):
            r = self.regions[k]
            J = np.zeros(r.n_instances)
            sens = np.zeros(r.n_instances)
            grad = None
            if want_gradient and r.n_actuator > 0:
                grad = np.zeros((r.n_instances, 4 * s0.nTimesteps, r.n_actuator))
            c = None if controls[k] is None else np.ascontiguousarray(controls[k], dtype=np.float64)
            keep.append(c)
            ctl[j], grd[j], sen[j], jj[j] = dptr(c), dptr(grad), dptr(sens), dptr(J)
            outs.append((sens, grad, J))
        check(capi.lib().mgd_group_run_adjoint(self._group_handle(members), 0.0, dptr(dts),
                                               s0.nTimesteps, s0.saveInterval, ctl,
                                               grd if want_gradient else None, sen, jj))
        return outs

    def run_adjoint_merged(self, controls: Optional[List[Optional[np.ndarray]]] = None,
                           want_gradient: bool = True):
        """Solver%runAdjoint of every region: the fused 2-D members of equal step count in
        one launch per stage (mgd_group_run_adjoint), the others concurrently beside them;
        list of (sens, grad, J) in member order."""
        controls = controls or [None] * len(self.regions)
        members = self._merged_members()
        rest = [k for k in range(len(self.regions)) if k not in members]
        results = [None] * len(self.regions)

        def one(k):
            return self.solvers[k].runAdjoint(control=controls[k], want_gradient=want_gradient)
        # the merged launches first, then the other members concurrently: issued at the
        # same time, their many small launches would share the hardware queues with the
        # merged ones and hold them up
        if members:
            for k, o in zip(members, self._run_adjoint_merged(members, controls, want_gradient)):
                results[k] = o
        if rest:
            with cf.ThreadPoolExecutor(max_workers=max(1, min(self.workers, len(rest)))) as ex:
                for k, o in zip(rest, ex.map(one, rest)):
                    results[k] = o
        return results

    def run_forward(self, controls: Optional[List[Optional[np.ndarray]]] = None):
        controls = controls or [None] * len(self.regions)

        def one(k):
            return self.solvers[k].runForward(control=controls[k])
        return self._map(one)

    def sync(self):
        for r in self.regions:
            r.sync()
