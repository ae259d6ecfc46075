"""Memory plan of the budgeted gradient + image job (host arithmetic only)."""
import numpy as np
import pytest

from taichirtm_amd import budget
from taichirtm_amd.batch import ShotSpec


def _shots(n, nt, nr):
    return [ShotSpec(np.array([[50, 1]], np.int32), (np.zeros((1, nt)),) * 3,
                     np.array([[k, 1] for k in range(nr)], np.int32)) for _ in range(n)]


def test_plan_raises_below_one_shot():
    with pytest.raises(MemoryError):
        budget.plan(400, 200, _shots(1, 100, 10), 1 << 20)


@pytest.mark.parametrize("share", [True, False])
def test_salt_plan_fits_and_scales_with_budget(share):
    """4000 x 2000 x 8000, 32 shots of 3000 receivers: checkpoints every
    ~sqrt(nt * 5 / 2.75) steps (5 state fields per checkpoint against 2.75
    history fields per step: the compact history's 2 velocities per step and
    3 stresses per 4 steps); the group grows with the budget and always fits
    it.
    Sharing the forward history, the RTM pass holds only its backward
    propagator next to the FWI pass."""
    shots = _shots(32, 8000, 3000)
    cells = 4000 * 2000
    last = 0
    for gb in (40, 64, 128, 256):
        p = budget.plan(4000, 2000, shots, gb * 1e9, share_forward=share)
        assert 110 <= p.L_fwi <= 130
        if share:
            assert p.L_rtm == p.L_fwi and p.shot_bytes_rtm < 20 * cells * 4 + 6 * 3000 * 8000 * 4
            per = p.shot_bytes_fwi + p.shot_bytes_rtm
        else:
            assert 120 <= p.L_rtm <= 160
            per = max(p.shot_bytes_fwi, p.shot_bytes_rtm)
        assert p.shared_bytes + p.group * per <= gb * 1e9
        assert p.groups * p.group >= 32 and p.group >= last
        # FWI per shot: state, adjoint, gradients, workspace, checkpoints, history
        assert p.shot_bytes_fwi >= (28 + 5 * (8000 // p.L_fwi) + 2 * p.L_fwi +
                                    3 * (p.L_fwi // 4)) * cells * 4
        last = p.group
    assert last == 10
