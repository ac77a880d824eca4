"""ctypes binding of the CPU oracle (oracle/liborc.so) -- test infrastructure only.

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg use this
module; it is the checker the HIP path is compared against.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB = os.path.join(ROOT, "oracle", "liborc.so")

MAXB = 20


class OrcOp(C.Structure):
    _fields_ = [("sym", C.c_int), ("iw", C.c_int), ("bw", C.c_int), ("bd", C.c_int),
                ("interior", C.c_double * (2 * MAXB + 1)),
                ("b1", C.c_double * (MAXB * MAXB)), ("b2", C.c_double * (MAXB * MAXB)),
                ("norm", C.c_double * MAXB), ("nnorm", C.c_int)]

    def interior_dict(self):
        hw = self.iw // 2
        return {k: self.interior[k + MAXB] for k in range(-hw, hw + 1)}

    def block(self, right=False):
        """Boundary block as an array B[m, c] (row m, column c), 0-based."""
        src = self.b2 if right else self.b1
        a = np.array(src[:]).reshape(MAXB, MAXB)
        return a[: self.bd, : self.bw].copy()

    def dense(self, n, periodic=False, off=0):
        """Dense n x n matrix of the operator (brute-force check)."""
        D = np.zeros((n, n))
        for j in range(n):
            e = np.zeros(n)
            e[j] = 1.0
            x = e.copy()
            apply(self, 0, (n, 1, 1), 0 if not periodic else (1 if off == 0 else 2), x.reshape(1, n))
            D[:, j] = x
        return D


class OrcDesc(C.Structure):
    _fields_ = [("nd", C.c_int), ("n", C.c_int * 3), ("periodicity", C.c_int * 3),
                ("periodic_length", C.c_double * 3), ("scheme", C.c_int), ("viscous", C.c_int),
                ("dissipation", C.c_int), ("composite_dissipation", C.c_int),
                ("gamma", C.c_double), ("reynolds", C.c_double), ("prandtl", C.c_double),
                ("power_law", C.c_double), ("bulk_ratio", C.c_double),
                ("dissipation_amount", C.c_double)]


class OrcPatchDesc(C.Structure):
    _fields_ = [("type", C.c_int), ("normal_dir", C.c_int), ("extent", C.c_int * 6),
                ("p", C.c_double * 4)]


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            subprocess.check_call(["make", "-s", "-C", os.path.join(ROOT, "oracle")])
        L = C.CDLL(LIB)
        dp = C.POINTER(C.c_double)
        L.orc_threads.restype = C.c_int
        L.orc_threads.argtypes = []
        L.orc_set_threads.argtypes = [C.c_int]
        L.orc_op_setup.argtypes = [C.POINTER(OrcOp), C.c_int, C.c_int]
        L.orc_op_adjoint.argtypes = [C.POINTER(OrcOp), C.POINTER(OrcOp)]
        L.orc_op_apply.argtypes = [C.POINTER(OrcOp), C.c_int, C.POINTER(C.c_int), C.c_int, dp, C.c_int]
        L.orc_op_apply_norm.argtypes = [C.POINTER(OrcOp), C.c_int, C.POINTER(C.c_int), C.c_int, dp,
                                        C.c_int, C.c_int]
        L.orc_region_create.restype = C.c_void_p
        L.orc_region_create.argtypes = [C.POINTER(OrcDesc), dp]
        L.orc_region_destroy.argtypes = [C.c_void_p]
        L.orc_region_add_patch.argtypes = [C.c_void_p, C.POINTER(OrcPatchDesc)]
        L.orc_region_set_field.argtypes = [C.c_void_p, C.c_int, dp]
        L.orc_region_add_source.argtypes = [C.c_void_p, dp, C.c_double, C.c_double, C.c_double,
                                            C.c_double]
        L.orc_region_finalize.argtypes = [C.c_void_p]
        L.orc_region_get_grid.argtypes = [C.c_void_p, C.c_int, dp]
        L.orc_compute_rhs.argtypes = [C.c_void_p, C.c_int, dp, dp, C.c_double, dp, dp, C.c_double, dp]
        L.orc_functional.restype = C.c_double
        L.orc_functional.argtypes = [C.c_void_p, dp]
        L.orc_time_step_size.restype = C.c_double
        L.orc_time_step_size.argtypes = [C.c_void_p, dp, C.c_double]
        L.orc_cfl.restype = C.c_double
        L.orc_cfl.argtypes = [C.c_void_p, dp, C.c_double]
        L.orc_adjoint_forcing.argtypes = [C.c_void_p, dp, dp]
        L.orc_cost_target_npoints.argtypes = [C.c_void_p]
        L.orc_actuator_npoints.argtypes = [C.c_void_p]
        L.orc_gradient.argtypes = [C.c_void_p, dp, dp]
        L.orc_sensitivity.restype = C.c_double
        L.orc_sensitivity.argtypes = [C.c_void_p, dp]
        L.orc_run_forward.restype = C.c_double
        L.orc_run_forward.argtypes = [C.c_void_p, dp, C.c_double, C.c_double, C.c_int, dp]
        L.orc_run_adjoint.restype = C.c_double
        L.orc_run_adjoint.argtypes = [C.c_void_p, dp, C.c_double, C.c_double, C.c_int, C.c_int, dp,
                                      dp, dp, dp]
        L.orc_run_adjoint_ic.restype = C.c_double
        L.orc_run_adjoint_ic.argtypes = [C.c_void_p, dp, C.c_double, C.c_double, C.c_int, C.c_int,
                                         dp, dp, dp, dp, dp, dp]
        for f in ("orc_inviscid_jacobian", "orc_incoming_jacobian", "orc_inviscid_flux"):
            getattr(L, f).restype = None
        L.orc_inviscid_jacobian.argtypes = [C.c_int, dp, dp, C.c_double, dp]
        L.orc_incoming_jacobian.argtypes = [C.c_int, dp, dp, C.c_double, C.c_int, dp]
        L.orc_inviscid_flux.argtypes = [C.c_int, dp, dp, C.c_double, dp]
        _lib = L
    return _lib


def _p(a):
    if a is None:
        return None
    assert a.dtype == np.float64 and a.flags.c_contiguous
    return a.ctypes.data_as(C.POINTER(C.c_double))


FIRST, SECOND, DISS, DISS_T, COMPOSITE = range(5)


def op(kind, scheme=0):
    o = OrcOp()
    rc = lib().orc_op_setup(C.byref(o), scheme, kind)
    assert rc == 0
    return o


def adjoint(o):
    a = OrcOp()
    lib().orc_op_adjoint(C.byref(o), C.byref(a))
    return a


def apply(o, direction, n, periodicity, x):
    x = np.ascontiguousarray(x)
    ncomp = x.shape[0] if x.ndim > 1 else 1
    nn = (C.c_int * 3)(*n)
    lib().orc_op_apply(C.byref(o), direction, nn, periodicity, _p(x), ncomp)
    return x


def apply_norm(o, direction, n, periodic, x, inverse=False):
    nn = (C.c_int * 3)(*n)
    ncomp = x.shape[0] if x.ndim > 1 else 1
    lib().orc_op_apply_norm(C.byref(o), direction, nn, int(periodic), _p(x), ncomp, int(inverse))
    return x


class Region:
    """Oracle region built from a magudi_amd.cases.Case."""

    def __init__(self, case):
        self.case = case
        d = OrcDesc()
        d.nd = case.nd
        for i in range(3):
            d.n[i] = case.n[i]
            d.periodicity[i] = case.periodicity[i]
            d.periodic_length[i] = case.periodic_length[i]
        d.scheme = case.scheme
        d.viscous = int(case.viscous)
        d.dissipation = int(case.dissipation)
        d.composite_dissipation = int(case.composite_dissipation)
        d.gamma = case.gamma
        d.reynolds = case.reynolds
        d.prandtl = case.prandtl
        d.power_law = case.power_law
        d.bulk_ratio = case.bulk_ratio
        d.dissipation_amount = case.dissipation_amount
        coords = np.ascontiguousarray(case.coords, dtype=np.float64)
        self.h = lib().orc_region_create(C.byref(d), _p(coords))
        assert self.h
        for p in case.patches:
            pd = OrcPatchDesc()
            pd.type = p.type
            pd.normal_dir = p.normal_dir
            for i in range(6):
                pd.extent[i] = p.extent[i]
            for i in range(4):
                pd.p[i] = p.params[i] if i < len(p.params) else 0.0
            assert lib().orc_region_add_patch(self.h, C.byref(pd)) == 0
        for which, arr in ((0, case.target), (1, case.control_mollifier), (2, case.target_mollifier)):
            if arr is not None:
                a = np.ascontiguousarray(arr, dtype=np.float64)
                lib().orc_region_set_field(self.h, which, _p(a))
        for s in case.sources:
            loc = np.array(list(s.location) + [0.0] * 3, dtype=np.float64)[:3]
            lib().orc_region_add_source(self.h, _p(loc), s.amplitude, s.frequency, s.radius, s.phase)
        lib().orc_region_finalize(self.h)
        self.N = case.npoints
        self.nv = case.nv

    def __del__(self):
        if getattr(self, "h", None):
            lib().orc_region_destroy(self.h)
            self.h = None

    def grid(self, which):
        nd = self.case.nd
        size = {0: nd * nd, 1: 1, 2: 1, 3: nd}[which] * self.N
        out = np.zeros(size)
        lib().orc_region_get_grid(self.h, which, _p(out))
        return out.reshape(-1, self.N)

    def rhs(self, mode, q, lam=None, time=0.0, control=None, aforcing=None, factor=0.0):
        q = np.ascontiguousarray(q, dtype=np.float64)
        out = np.zeros((self.nv, self.N))
        lam_ = None if lam is None else np.ascontiguousarray(lam, dtype=np.float64)
        ctl = None if control is None else np.ascontiguousarray(control, dtype=np.float64)
        af = None if aforcing is None else np.ascontiguousarray(aforcing, dtype=np.float64)
        lib().orc_compute_rhs(self.h, mode, _p(q), _p(lam_), time, _p(ctl), _p(af), factor, _p(out))
        return out

    def functional(self, q):
        q = np.ascontiguousarray(q, dtype=np.float64)
        return lib().orc_functional(self.h, _p(q))

    def time_step_size(self, q, cfl):
        return lib().orc_time_step_size(self.h, _p(np.ascontiguousarray(q, dtype=np.float64)), cfl)

    def cfl(self, q, dt):
        return lib().orc_cfl(self.h, _p(np.ascontiguousarray(q, dtype=np.float64)), dt)

    def adjoint_forcing(self, q):
        n = lib().orc_cost_target_npoints(self.h)
        out = np.zeros((self.nv, n))
        q = np.ascontiguousarray(q, dtype=np.float64)
        lib().orc_adjoint_forcing(self.h, _p(q), _p(out))
        return out

    @property
    def n_actuator(self):
        return lib().orc_actuator_npoints(self.h)

    def run_forward(self, q0, dt, nsteps, control=None, t0=0.0):
        q = np.ascontiguousarray(q0, dtype=np.float64).copy()
        ctl = None if control is None else np.ascontiguousarray(control, dtype=np.float64)
        J = lib().orc_run_forward(self.h, _p(q), t0, dt, nsteps, _p(ctl))
        return J, q

    def run_adjoint(self, q0, dt, nsteps, save_interval, control=None, t0=0.0):
        q0 = np.ascontiguousarray(q0, dtype=np.float64)
        na = self.n_actuator
        grad = np.zeros((nsteps * 4, max(na, 1)))
        lam0 = np.zeros((self.nv, self.N))
        J = C.c_double(0.0)
        ctl = None if control is None else np.ascontiguousarray(control, dtype=np.float64)
        sens = lib().orc_run_adjoint(self.h, _p(q0), t0, dt, nsteps, save_interval, _p(ctl),
                                     _p(grad), _p(lam0), C.cast(C.pointer(J), C.POINTER(C.c_double)))
        return sens, grad[:, :na], lam0, J.value

    def run_adjoint_ic(self, q0, dt, nsteps, save_interval, control=None, t0=0.0, lamT=None):
        """run_adjoint with a nonzero adjoint terminal state; also returns the final state."""
        q0 = np.ascontiguousarray(q0, dtype=np.float64)
        na = self.n_actuator
        grad = np.zeros((nsteps * 4, max(na, 1)))
        lam0 = np.zeros((self.nv, self.N))
        qT = np.zeros((self.nv, self.N))
        J = C.c_double(0.0)
        ctl = None if control is None else np.ascontiguousarray(control, dtype=np.float64)
        lt = None if lamT is None else np.ascontiguousarray(lamT, dtype=np.float64)
        sens = lib().orc_run_adjoint_ic(self.h, _p(q0), t0, dt, nsteps, save_interval, _p(ctl),
                                        _p(lt), _p(grad), _p(lam0),
                                        C.cast(C.pointer(J), C.POINTER(C.c_double)), _p(qT))
        return sens, grad[:, :na], lam0, J.value, qT
