"""What the step carries from one call to the next (psa_group.wtag): the
p, C, D operand planes that the split-BF16 transport column pass leaves for
the next RHS, and the diagonal sums of C, D that the sources leave for the
next prep.  A group that carries them steps bit for bit like one that
recomputes them (wtag NULL), across level swaps, host writes of the state
(which invalidate the tags), the 8- and 16-line column passes and a column
pass that leaves no planes; and the tags do point at the current level
after a step."""

import numpy as np
import pytest

from tests.helpers import params_pair, random_state, wire_length

pytestmark = pytest.mark.gpu

LEVELS = ("pA", "pB", "CA", "CB", "DA", "DB")


def _pair(prec, N, nw, col_lines, seed):
    from pulsesuite_amd.engine import WireGroup
    L = wire_length(N)
    prm, _ = params_pair(L)
    rr = np.linspace(-1.2 * L, 1.2 * L, 96)
    rng = np.random.default_rng(seed)
    st = random_state(rng, N, nw)
    E = 2e7 * np.cos(2 * np.pi * rr / (3 * L)) + 0j
    out = []
    for carry in (True, False):
        wg = WireGroup(prm, nw, rr, prec=prec, carry=carry)
        wg.tune[1] = 8
        wg.tune[3] = col_lines
        wg._sync_struct()
        wg.set_state(*st)
        wg.set_fields(E, 0.5 * E, 0 * E)
        out.append(wg)
    return out, rng


def _same(a, b, what):
    import torch
    torch.cuda.synchronize()
    for k in LEVELS:
        assert np.array_equal(a.s[k].cpu().numpy(), b.s[k].cpu().numpy()), (what, k)
    assert np.array_equal(a.s["out"].cpu().numpy(), b.s["out"].cpu().numpy()), (what, "out")


@pytest.mark.parametrize("prec,N,col_lines", [("bf16x3", 64, 8), ("bf16x3", 96, 16), ("bf16x3", 128, 4),
                                              ("fp64", 64, 8)])
def test_carried_step_equals_recomputed_step(prec, N, col_lines):
    import torch
    nw = 3
    print(f"[carry] prec={prec} N={N} col_lines={col_lines}", flush=True)
    (a, b), rng = _pair(prec, N, nw, col_lines, seed=N + col_lines)
    dt = 2e-17
    for it in range(4):
        a.step(dt)
        b.step(dt)
        _same(a, b, f"step {it}")
    # the tags name the current level 3 of every wire (the planes only when
    # the column pass wrote them: split-BF16 with 8 or 16 lines per block)
    torch.cuda.synchronize()
    tag = a.s["wtag"].cpu().numpy().view(np.uint64).reshape(nw, 2)
    per = N * N * 16
    for w in range(nw):
        assert tag[w, 1] == a.s["CB"].data_ptr() + w * per, w
        if prec == "bf16x3" and col_lines >= 8:
            assert tag[w, 0] == a.s["pB"].data_ptr() + w * per, w
    # a host write of the state drops what was carried
    st = random_state(rng, N, nw)
    a.set_state(*st)
    b.set_state(*st)
    for it in range(3):
        a.step(dt)
        b.step(dt)
        _same(a, b, f"after set_state, step {it}")
    # so does a write of one wire followed by invalidate()
    for g in (a, b):
        g.s["CB"][1, 2, 5] += 1e-3
        g.s["pB"][2, 3, 1] -= 2e-3
        g.invalidate(1)
        g.invalidate(2)
    for it in range(2):
        a.step(dt)
        b.step(dt)
        _same(a, b, f"after a wire write, step {it}")
