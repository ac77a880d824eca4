"""Pins the oracle's SBP operators (oracle/orc_ops.c) by what the mathematics fixes.

Mirrors the reference's own operator tests:
  test/stencil_coefficients.f90  -- Taylor order of interior and boundary rows
  test/SBP_property.f90          -- H D + (H D)^T = B (boundary terms only)
  test/adjoint_relation/SBP_operators.f90 -- <D f, g>_H = <f, D^+ g>_H
  test/dissipation_self_adjoint.f90
plus brute force on small dense matrices.  No golden vectors exist in the
reference; these identities are what pin the coefficient tables.
"""
import math

import numpy as np
import pytest

import oracle_bind as ob

EPS = np.finfo(float).eps


def taylor_residuals(coefs, offsets, deriv, kmax):
    """max_k |sum c_j j^k / k! - delta_{k,deriv}| for k = 0..kmax."""
    res = []
    for k in range(kmax + 1):
        s = sum(c * (o ** k) / math.factorial(k) for c, o in zip(coefs, offsets))
        res.append(abs(s - (1.0 if (deriv > 0 and k == deriv) else 0.0)))
    return np.array(res)


@pytest.mark.parametrize("kind,deriv,inner,bnd,tol", [
    (ob.FIRST, 1, 8, 4, 300 * EPS),
    (ob.SECOND, 2, 8, 4, 64 * EPS),
    (ob.COMPOSITE, 0, 8, 4, 64 * EPS),
])
def test_taylor_order(kind, deriv, inner, bnd, tol):
    o = ob.op(kind)
    I = o.interior_dict()
    offs = sorted(I)
    # interior: exact for degrees < inner + deriv, fails at inner + deriv
    r = taylor_residuals([I[k] for k in offs], offs, deriv, inner + deriv)
    assert np.all(r[: inner + deriv] < tol * len(offs) * 10), r
    assert r[inner + deriv] > 1e-6
    B = o.block()
    for m in range(o.bd):
        offsets = [c - m for c in range(o.bw)]
        r = taylor_residuals(list(B[m]), offsets, deriv, bnd + deriv - 1)
        scale = max(1.0, np.abs(B[m]).max()) * o.bw
        assert np.all(r < tol * scale * 10), (m, r)


def test_first_derivative_skew_and_sbp_property():
    o = ob.op(ob.FIRST)
    I = o.interior_dict()
    for k in range(1, 5):
        assert I[-k] == -I[k]
    n = 48
    D = o.dense(n)
    h = np.ones(n)
    h[:8] = o.norm[:8]
    h[-8:] = o.norm[:8][::-1]
    H = np.diag(h)
    Bm = np.zeros((n, n))
    Bm[0, 0], Bm[-1, -1] = -1.0, 1.0
    Q = H @ D
    assert np.abs(Q + Q.T - Bm).max() < 50 * EPS
    # H is the norm of the operator: positive weights, sum equals the domain length n-1
    assert np.all(h > 0) and abs(h.sum() - (n - 1)) < 1e-12


def test_adjoint_operator_is_H_transpose():
    o = ob.op(ob.FIRST)
    a = ob.adjoint(o)
    assert (a.bw, a.bd) == (16, 12)
    n = 48
    D = o.dense(n)
    A = a.dense(n)
    h = np.ones(n)
    h[:8] = o.norm[:8]
    h[-8:] = o.norm[:8][::-1]
    Dt = np.diag(1 / h) @ D.T @ np.diag(h)
    assert np.abs(A - Dt).max() < 1e-12
    # SBP_property.f90: H00 (D f + D^+ f) equals -f at the left end and +f at the right end
    rng = np.random.default_rng(0)
    f = rng.standard_normal(n)
    s = D @ f + A @ f
    assert abs(h[0] * s[0] + f[0]) < 1e-12 and abs(h[-1] * s[-1] - f[-1]) < 1e-12
    assert np.abs(s[1:-1]).max() < 1e-12


def test_adjoint_relation_periodic():
    o = ob.op(ob.FIRST)
    a = ob.adjoint(o)
    n = 37
    D = o.dense(n, periodic=True)
    A = a.dense(n, periodic=True)
    assert np.abs(A + D).max() < 1e-15          # skew-symmetric interior
    assert np.abs(D + D.T).max() < 1e-15


def test_dissipation_pair_is_transpose_and_negative_semidefinite():
    d = ob.op(ob.DISS)
    dt = ob.op(ob.DISS_T)
    n = 40
    A = d.dense(n)
    At = dt.dense(n)
    assert np.abs(A.T - At).max() == 0.0
    # the non-composite dissipation -D4^T D4 is symmetric negative semidefinite
    M = -At @ A
    assert np.abs(M - M.T).max() == 0.0
    assert np.linalg.eigvalsh(M).max() < 1e-10
    # composite dissipation is self-adjoint in the H norm
    c = ob.op(ob.COMPOSITE)
    Cm = c.dense(n)
    h = np.ones(n)
    h[:8] = c.norm[:8]
    h[-8:] = c.norm[:8][::-1]
    HC = np.diag(h) @ Cm
    assert np.abs(HC - HC.T).max() < 1e-12


def test_apply_direction_and_norm_inverse():
    """Operator along y of a 2-D array equals the 1-D operator applied to columns."""
    o = ob.op(ob.FIRST)
    nx, ny = 9, 23
    rng = np.random.default_rng(1)
    x = rng.standard_normal((2, ny, nx))
    y = ob.apply(o, 1, (nx, ny, 1), 0, x.reshape(2, -1).copy()).reshape(2, ny, nx)
    D = o.dense(ny)
    assert np.abs(y - np.einsum("ij,cjk->cik", D, x)).max() < 1e-13
    z = ob.apply_norm(o, 1, (nx, ny, 1), False, x.reshape(2, -1).copy(), inverse=True).reshape(2, ny, nx)
    h = np.ones(ny)
    h[:8] = o.norm[:8]
    h[-8:] = o.norm[:8][::-1]
    assert np.abs(z - x / h[None, :, None]).max() < 1e-15


def test_derivative_of_polynomial_exact():
    o = ob.op(ob.FIRST)
    n = 64
    xg = np.linspace(-1.3, 2.1, n)
    hgrid = xg[1] - xg[0]
    f = xg ** 4 - 2 * xg ** 3 + xg
    df = ob.apply(o, 0, (n, 1, 1), 0, f.reshape(1, -1).copy()).ravel() / hgrid
    assert np.abs(df - (4 * xg ** 3 - 6 * xg ** 2 + 1)).max() < 1e-10
