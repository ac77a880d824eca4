"""Pins the oracle's pointwise physics, RHS and adjoint (oracle/orc_*.c).

Mirrors the reference's relation tests (no golden vectors exist in magudi):
  test/inviscid_flux_jacobian.f90         -- FD order of the flux Jacobian
  test/adjoint_relation/full_rhs.f90      -- <R^+ l, dq>_H + <l, R'(q) dq>_H -> 0 at O(h)
  test/adjoint_relation/time_integrator.f90, README.md:80-157 -- (J(A)-J0)/A -> Grad0
Special cases the mathematics fixes: quiescent state with no source gives rhs = 0,
A+ + A- = A for the characteristic split, R L = I.
"""
import numpy as np
import pytest

import oracle_bind as ob
from magudi_amd import cases


def rand_state(nd, n, rng):
    q = np.zeros((nd + 2, n))
    q[0] = rng.uniform(0.5, 2.0, n)
    for j in range(nd):
        q[j + 1] = q[0] * rng.uniform(-0.5, 0.5, n)
    q[nd + 1] = q[0] * rng.uniform(0.5, 2.0, n) / 1.4 + 0.5 / q[0] * (q[1:nd + 1] ** 2).sum(0)
    return q


@pytest.mark.parametrize("nd", [2, 3])
def test_inviscid_jacobian_fd(nd):
    rng = np.random.default_rng(nd)
    q = rand_state(nd, 1, rng)[:, 0].copy()
    m = rng.uniform(-1, 1, nd)
    A = np.zeros((nd + 2) * (nd + 2))
    qp = np.ascontiguousarray(q)
    ob.lib().orc_inviscid_jacobian(nd, ob._p(qp), ob._p(m), 1.4, ob._p(A))
    A = A.reshape(nd + 2, nd + 2).T            # column-major -> A[i, j]
    F0 = np.zeros(nd + 2)
    ob.lib().orc_inviscid_flux(nd, ob._p(qp), ob._p(m), 1.4, ob._p(F0))
    dq = rng.standard_normal(nd + 2)
    errs = []
    for h in [1e-3, 1e-4, 1e-5]:
        q1 = np.ascontiguousarray(q + h * dq)
        F1 = np.zeros(nd + 2)
        ob.lib().orc_inviscid_flux(nd, ob._p(q1), ob._p(m), 1.4, ob._p(F1))
        errs.append(np.abs((F1 - F0) / h - A @ dq).max())
    assert errs[2] < errs[0] * 2e-2 and errs[2] < 1e-4 * max(1.0, np.abs(A @ dq).max())


@pytest.mark.parametrize("nd", [2, 3])
def test_incoming_jacobian_split(nd):
    rng = np.random.default_rng(10 + nd)
    q = np.ascontiguousarray(rand_state(nd, 1, rng)[:, 0])
    m = rng.uniform(-1, 1, nd)
    n = nd + 2
    Ap, Am, A = np.zeros(n * n), np.zeros(n * n), np.zeros(n * n)
    ob.lib().orc_incoming_jacobian(nd, ob._p(q), ob._p(m), 1.4, 1, ob._p(Ap))
    ob.lib().orc_incoming_jacobian(nd, ob._p(q), ob._p(m), 1.4, -1, ob._p(Am))
    ob.lib().orc_inviscid_jacobian(nd, ob._p(q), ob._p(m), 1.4, ob._p(A))
    assert np.abs(Ap + Am - A).max() < 1e-12 * max(1, np.abs(A).max())
    # A+ has only non-negative eigenvalues
    ev = np.linalg.eigvals(Ap.reshape(n, n).T)
    assert ev.real.min() > -1e-12


def test_quiescent_rhs_is_zero():
    c = cases.acoustic_monopole(n=41, nsteps=1)
    c.sources = []
    r = ob.Region(c)
    q = c.target.copy()
    rhs = r.rhs(1, q)
    assert np.abs(rhs).max() < 1e-13


def test_grid_metrics_uniform():
    c = cases.acoustic_monopole(n=41, nsteps=1)
    r = ob.Region(c)
    dx = 28.0 / 40
    m = r.grid(0)
    assert np.allclose(m[0], dx, rtol=0, atol=1e-12) and np.allclose(m[3], dx, rtol=0, atol=1e-12)
    assert np.abs(m[1]).max() == 0 and np.abs(m[2]).max() == 0
    jac = r.grid(1)[0]
    assert np.allclose(jac, 1 / dx ** 2, rtol=1e-11, atol=0)
    norm = r.grid(2)[0]
    assert abs(norm.sum() - 28.0 ** 2) < 1e-9       # H-quadrature integrates 1 exactly


def _adjoint_relation(case, seed, amp=0.05):
    rng = np.random.default_rng(seed)
    r = ob.Region(case)
    N, nv = case.npoints, case.nv
    q = case.target + amp * rng.standard_normal((nv, N)) * np.array([[1.0]] + [[1.0]] * (nv - 1))
    q[0] = np.abs(q[0])
    lam = rng.standard_normal((nv, N))
    dq = rng.standard_normal((nv, N))
    norm = r.grid(2)[0]
    s1 = np.sum(r.rhs(-1, q, lam) * dq * norm)
    R0 = r.rhs(1, q)
    errs = []
    hs = [1e-3 * 10 ** (-k / 2) for k in range(12)]
    for h in hs:
        s2 = np.sum(lam * (r.rhs(1, q + h * dq) - R0) / h * norm)
        errs.append(abs(s1 + s2) / abs(s1))
    return np.array(errs)


def test_adjoint_relation_inviscid_monopole():
    c = cases.acoustic_monopole(n=41, nsteps=1)
    errs = _adjoint_relation(c, 3)
    # first-order convergence in h, down to well below 1e-6
    assert errs[4] < errs[0] * 0.05 and errs.min() < 1e-6, errs


def test_adjoint_relation_viscous_farfield():
    c = cases.acoustic_monopole(n=41, nsteps=1)
    c.viscous = True
    c.reynolds = 200.0
    c.prandtl = 0.7
    c.patches = [cases.Patch(p.name, p.type, p.normal_dir, p.extent,
                             (1.0, 0.7, 0, 0) if p.type == cases.FAR_FIELD else p.params)
                 for p in c.patches]
    errs = _adjoint_relation(c, 4)
    order = -np.polyfit(np.log10([1e-3 * 10 ** (-k / 2) for k in range(6)]), np.log10(errs[:6]), 1)[0] * -1
    assert abs(order - 1.0) < 0.1 and errs.min() < 1e-5, errs


def test_adjoint_relation_viscous_periodic_3d_powerlaw():
    c = cases.kolmogorov3d(n=12, nsteps=1)
    c.power_law = 0.666
    c.bulk_ratio = 0.6
    c.dissipation = True
    c.dissipation_amount = 0.01
    errs = _adjoint_relation(c, 5)
    assert errs[4] < errs[0] * 0.05 and errs.min() < 1e-5, errs


@pytest.mark.parametrize("viscous", [False, True])
def test_adjoint_relation_walls(viscous):
    """Slip and isothermal wall penalties (src/ImpenetrableWallImpl.f90:60-207,
    src/IsothermalWallImpl.f90:101-330) keep the discrete adjoint relation."""
    c = cases.walled_monopole(n=41, viscous=viscous)
    errs = _adjoint_relation(c, 6)
    assert errs[4] < errs[0] * 0.05 and errs.min() < 1e-5, errs


def test_slip_wall_penalty_vanishes_for_tangential_flow():
    """The slip-wall penalty is proportional to the normal momentum: zero when the flow
    at the wall is tangential."""
    c = cases.walled_monopole(n=41)
    c.patches = [p for p in c.patches if p.type == cases.IMPENETRABLE_WALL]
    c.dissipation = False
    c.sources = []
    r = ob.Region(c)
    q = c.target.copy()
    q[1] = 0.1 * q[0]                         # uniform flow along x: tangential to y walls
    q[3] = q[3] + 0.5 * q[1] ** 2 / q[0]
    base = cases.walled_monopole(n=41)
    base.patches = []
    base.dissipation = False
    base.sources = []
    assert np.array_equal(r.rhs(1, q), ob.Region(base).rhs(1, q))


def test_isothermal_wall_quiescent_at_wall_temperature():
    """At rest at the wall temperature T_w = 1/(gamma-1) the no-slip penalty
    J (0, rho u, rho v, E - rho T_w / gamma) vanishes: the rhs is zero to rounding,
    while a hotter wall drives the energy at the wall."""
    c = cases.walled_monopole(n=41, viscous=True)
    c.sources = []
    c.dissipation = False
    r = ob.Region(c)
    assert np.abs(r.rhs(1, c.target)).max() < 1e-12
    hot = [cases.Patch(p.name, p.type, p.normal_dir, p.extent, (1.0, 1.0, 1.5 * p.params[2], 0))
           if p.type == cases.ISOTHERMAL_WALL else p for p in c.patches]
    c.patches = hot
    assert np.abs(ob.Region(c).rhs(1, c.target)).max() > 1e-3


def test_gradient_fd_relation():
    """(J(A) - J(0)) / A -> Grad0 along the adjoint gradient (README.md:80-157)."""
    c = cases.acoustic_monopole(n=41, nsteps=24, save_interval=5)
    c.dt = 0.2
    r = ob.Region(c)
    q0 = c.initial_state()
    sens, grad, lam0, J0 = r.run_adjoint(q0, c.dt, c.nsteps, c.save_interval)
    J0f, _ = r.run_forward(q0, c.dt, c.nsteps)
    assert J0 == J0f                       # the adjoint's forward pass is the forward run
    assert sens > 0
    errs = []
    for A in [1e-1, 1e-2, 1e-3]:
        J1, _ = r.run_forward(q0, c.dt, c.nsteps, control=A * grad)
        errs.append(abs((J1 - J0) / A - sens) / sens)
    assert errs[2] < 1e-3 and errs[2] < errs[0] * 0.05, errs


def test_initial_condition_gradient():
    """dJ/dq0 along dq equals <lambda(0), dq>_H (discrete adjoint of RK4 + RHS)."""
    c = cases.acoustic_monopole(n=41, nsteps=12, save_interval=4)
    c.dt = 0.2
    r = ob.Region(c)
    q0 = c.initial_state()
    x, y = c.coords
    q0[3] += 1e-2 * np.exp(-((x - 7.5) ** 2 + (y - 1.0) ** 2))   # pressure pulse on the target
    dq = np.zeros_like(q0)
    dq[3] = np.exp(-((x - 6.0) ** 2 + y ** 2))
    dq[1] = 0.3 * np.exp(-((x - 8.0) ** 2 + (y + 1.0) ** 2))
    sens, grad, lam0, J0 = r.run_adjoint(q0, c.dt, c.nsteps, c.save_interval)
    norm = r.grid(2)[0]
    pred = np.sum(lam0 * dq * norm)
    errs = []
    for h in [1e-4, 1e-5, 1e-6]:
        J1, _ = r.run_forward(q0 + h * dq, c.dt, c.nsteps)
        errs.append(abs((J1 - J0) / h - pred) / abs(pred))
    assert errs[-1] < 1e-4 and errs[-1] < errs[0] * 0.05, errs


def test_oracle_threads_do_not_change_results():
    """The oracle's OpenMP loops (per point / per pencil) give the serial values bit for
    bit, whatever the thread count: forward and adjoint rhs of a viscous 3-D periodic
    case and a whole forward + adjoint run of the monopole."""
    from magudi_amd import cases
    L = ob.lib()
    saved = L.orc_threads()
    try:
        c = cases.kolmogorov3d(n=12, nsteps=1)
        c.dissipation = True
        rng = np.random.default_rng(4)
        q = c.initial_state() + 1e-3 * rng.standard_normal((c.nv, c.npoints))
        lam = rng.standard_normal((c.nv, c.npoints))
        m = cases.acoustic_monopole(n=41, nsteps=6, save_interval=3)
        m.dt = 0.2
        outs = []
        for nt in (1, 3):
            L.orc_set_threads(nt)
            O = ob.Region(c)
            Om = ob.Region(m)
            outs.append((O.rhs(1, q, time=0.3), O.rhs(-1, q, lam, factor=0.5),
                         Om.run_adjoint(m.initial_state(), m.dt, m.nsteps, m.save_interval)))
        (f1, a1, r1), (f2, a2, r2) = outs
        assert np.array_equal(f1, f2) and np.array_equal(a1, a2)
        assert r1[0] == r2[0] and r1[3] == r2[3]
        assert np.array_equal(r1[1], r2[1]) and np.array_equal(r1[2], r2[2])
    finally:
        L.orc_set_threads(saved)
