"""Costs on the final state, for runs whose adjoint starts from a nonzero terminal state.

magudi's own functionals are time integrals (SOUND: src/AcousticNoiseImpl.f90:123-206,
accumulated by Solver%runForward, src/SolverImpl.f90:837-841); a cost on the final state
enters the discrete adjoint as its terminal condition instead
(adjoint_nonzero_initial_condition, src/SolverImpl.f90:368-428, 1096-1102).  With
J = sum_p H_p f(q_p) over the SBP norm H, dJ = <f'(q), dq>_H, so the terminal adjoint
state is the pointwise derivative f'(q); the adjoint at t0 is then dJ/dq0 in the same
inner product.
"""
from __future__ import annotations

import numpy as np


def kinetic_energy(q: np.ndarray, norm: np.ndarray, nd: int):
    """Kinetic energy J = sum_p H_p |rho u|^2 / (2 rho) of one state q (nv x N,
    conserved variables) and its terminal adjoint state dJ/dq (pointwise derivative:
    -|rho u|^2 / (2 rho^2) for rho, u_i for the momenta, 0 for the energy)."""
    rho = q[0]
    mom = q[1:1 + nd]
    m2 = np.zeros_like(rho)
    for i in range(nd):
        m2 = m2 + mom[i] * mom[i]
    ke = 0.5 * m2 / rho
    J = float(np.sum(norm * ke))
    lam = np.zeros_like(q)
    lam[0] = -0.5 * m2 / (rho * rho)
    for i in range(nd):
        lam[1 + i] = mom[i] / rho
    return J, lam
