"""Multiple shooting on the GPU (magudi_amd/multishoot.py) against the oracle
restatement of msforward / msadjoint (tests/ms_oracle.py, pinned by finite
differences in test_oracle_multishoot.py).  Tolerances: J 1e-12 relative, gradients
and inner products 1e-10 relative."""
import numpy as np
import pytest

import ms_oracle as ms
from test_oracle_multishoot import ms_case, ms_inputs

pytestmark = pytest.mark.gpu


def rel(a, b):
    a, b = np.asarray(a), np.asarray(b)
    return np.abs(a - b).max() / max(np.abs(b).max(), 1e-300)


@pytest.mark.parametrize("plain", [False, True])
@pytest.mark.parametrize("kind,threshold", [("quadratic", 1e-12), ("huber", 1e-6)])
def test_multishoot_matches_oracle(plain, kind, threshold):
    from magudi_amd.multishoot import MultipleShooting
    c = ms_case()
    O, ics, t0s = ms_inputs(c)
    nts, K, w = 4, 2, 20.0
    na = O.n_actuator
    ctl = 1e-2 * np.random.default_rng(7).standard_normal((3 * 4 * nts, na))
    Jo, cost_o, L2_o, _ = ms.ms_forward(O, ics, t0s, c.dt, nts, w, control=ctl, kind=kind,
                                        threshold=threshold)
    out_o = ms.ms_adjoint(O, ics, t0s, c.dt, nts, K, w, control=ctl, kind=kind, threshold=threshold)
    M = MultipleShooting(c, nsplit=3, nts=nts, penalty_weight=w, penalty_type=kind,
                         huber_threshold=threshold, save_interval=K, dt=c.dt, force_plain=plain)
    J = M.msforward(ics, t0s, control=ctl)
    assert abs(J - Jo) <= 1e-12 * abs(Jo)
    assert rel(M.segment_cost, cost_o) < 1e-12
    assert rel(M.L2sq[1:], L2_o[1:]) < 1e-12
    out = M.msadjoint()
    assert rel(out["ctrl_ip"], out_o["ctrl_ip"]) < 1e-10
    assert rel(out["ic_ip"][1:], out_o["ic_ip"][1:]) < 1e-10
    assert abs(out["total"] - out_o["total"]) <= 1e-10 * abs(out_o["total"])
    for k in range(3):
        assert rel(out["ic_grad"][k], out_o["ic_grad"][k]) < 1e-10, k
    assert rel(out["control_grad"], out_o["control_grad"]) < 1e-10
