"""Launch choices of the autotuner change the schedule, not the arithmetic:
every choice gives bit-identical results (per output element the same
products are accumulated in the same order), and autotune() leaves the
state untouched."""

import numpy as np
import pytest

from tests.test_gpu_parity import Case

pytestmark = pytest.mark.gpu


def _step_outputs(c, tune):
    import torch
    c.wg.set_state(*c.state)
    c.wg.tune = list(tune)
    c.wg._sync_struct()
    c.wg.step(1e-17)
    torch.cuda.synchronize()
    return [x.copy() for x in c.wg.get_state()] + [x.copy() for x in c.wg.outputs()]


@pytest.mark.parametrize("prec", ["fp64", "bf16x3"])
def test_every_launch_choice_gives_identical_results(prec):
    c = Case(64, 8, seed=3, prec=prec)
    ref = _step_outputs(c, [0, 0, 0, 0])
    for t1 in (2, 4, 8, 16):
        for t2 in (0, 1):
            got = _step_outputs(c, [0, t1, t2, 0])
            for a, b in zip(got, ref):
                assert np.array_equal(a, b), (t1, t2)
    for t3 in (2, 4, 8, 16):
        got = _step_outputs(c, [0, 0, 0, t3])
        for a, b in zip(got, ref):
            assert np.array_equal(a, b), t3


def test_autotune_keeps_state_and_picks_valid_choices():
    c = Case(64, 8, seed=4, prec="bf16x3")
    before = c.wg.get_state()
    tuned = c.wg.autotune(reps=2)
    after = c.wg.get_state()
    for a, b in zip(before, after):
        assert np.array_equal(a, b)
    assert c.wg.tune[0] == 0 and c.wg.tune[1] in (2, 4, 8, 16) and c.wg.tune[2] in (0, 1)
    assert c.wg.tune[3] in (2, 4, 8, 16)
    assert set(tuned) == {1, 2, 3}
