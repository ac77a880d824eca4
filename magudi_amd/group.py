"""Uneven batches: independent regions of any size and problem type run together.

Each region owns its HIP streams, so the runs of several regions overlap on the GPU
when they are issued from different host threads (the C ABI calls release the
Python interpreter lock).  Each region keeps its own tile shape (autotuned for its
grid) and its own checkpoints; results are those of the region run alone.
"""
from __future__ import annotations

import concurrent.futures as cf
from typing import List, Optional, Sequence

import numpy as np

from .cases import Case
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

    def run_forward(self, controls: Optional[List[Optional[np.ndarray]]] = None):
        controls = controls or [None] * len(self.regions)

        def one(k):
            return self.solvers[k].runForward(control=controls[k])
        return self._map(one)

    def sync(self):
        for r in self.regions:
            r.sync()
