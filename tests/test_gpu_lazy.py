"""The deferred leapfrog average (psa_group.lazy): a group that leaves
X2 = (X1 + X3) / 2 to the next step's RHS steps bit for bit like one that
averages in place, in every precision mode, with the transport's fused
average, with the LF = 0 transport and without transport, per group and in
the multi-group step; the state a host reads (get_state) is the averaged
one, and host writes reset the flags."""

import numpy as np
import pytest

from oracle import sbe_oracle as O
from tests.helpers import random_state, wire_length

pytestmark = pytest.mark.gpu


def _flags(**kw):
    f = list(O.SHIPPED_FLAGS)
    for k, v in kw.items():
        f[O.FLAG_NAMES.index(k)] = v
    return tuple(f)


def _group(prec, N, nw, flags, seed, lazy):
    from pulsesuite_amd.engine import WireGroup
    from pulsesuite_amd.tables import QWParams
    L = wire_length(N)
    rr = np.linspace(-1.2 * L, 1.2 * L, 96)
    wg = WireGroup(QWParams(L=L, me_rel=0.067, flags=flags), nw, rr, prec=prec, lazy=lazy)
    wg.set_state(*random_state(np.random.default_rng(seed), N, nw))
    E = 2e7 * np.cos(2 * np.pi * rr / (3 * L)) + 0j
    wg.set_fields(E, 0.5 * E, 0 * E)
    return wg


def _same(a, b, what):
    for x, y in zip(a.get_state(), b.get_state()):
        assert np.array_equal(x, y), what
    for x, y in zip(a.outputs(), b.outputs()):
        assert np.array_equal(x, y), what
    assert np.array_equal(a.observables()["I0"], b.observables()["I0"]), what


@pytest.mark.parametrize("prec,N,flags", [
    ("bf16x3", 64, O.SHIPPED_FLAGS), ("fp64", 64, O.SHIPPED_FLAGS), ("fp32", 96, O.SHIPPED_FLAGS),
    ("bf16x3", 128, _flags(LF=0)), ("fp64", 64, _flags(DCTrans=0)),
])
def test_deferred_average_equals_in_place_average(prec, N, flags):
    import torch
    print(f"[lazy] prec={prec} N={N} flags={flags}", flush=True)
    dt = 2e-17
    a = _group(prec, N, 3, flags, N + 1, lazy=True)
    b = _group(prec, N, 3, flags, N + 1, lazy=False)
    for it in range(3):
        a.step(dt)
        b.step(dt)
        torch.cuda.synchronize()
        # every step leaves the flags set (level A holds this step's X1)
        assert a.s["lazy"].cpu().numpy().tolist() == [1, 1, 1]
        # the level the next step reads (B) is the same; A differs until averaged
        for k in ("pB", "CB", "DB"):
            assert np.array_equal(a.s[k].cpu().numpy(), b.s[k].cpu().numpy()), (it, k)
    _same(a, b, "after 3 steps")
    assert a.s["lazy"].cpu().numpy().tolist() == [0, 0, 0]  # get_state wrote the average out
    # stepping on from a written-out average, and on a wire subset
    for it in range(2):
        a.step(dt)
        b.step(dt)
    a.step(dt, 1, 2)
    b.step(dt, 1, 2)
    _same(a, b, "after a subset step")
    # a host write resets the flags: both groups continue from the same state
    st = random_state(np.random.default_rng(5), N, 3)
    a.step(dt)
    b.step(dt)
    a.set_state(*st)
    b.set_state(*st)
    assert a.s["lazy"].cpu().numpy().tolist() == [0, 0, 0]
    for it in range(2):
        a.step(dt)
        b.step(dt)
    _same(a, b, "after set_state")


@pytest.mark.parametrize("prec", ["fp64", "bf16x3"])
def test_deferred_average_in_the_multi_group_step(prec):
    import torch
    from pulsesuite_amd.engine import MultiStep
    print(f"[lazy multi] prec={prec}", flush=True)
    dt = 2e-17
    sizes = ((64, 2), (96, 3), (128, 2))
    a = [_group(prec, N, nw, O.SHIPPED_FLAGS, 3 + N, lazy=True) for N, nw in sizes]
    b = [_group(prec, N, nw, O.SHIPPED_FLAGS, 3 + N, lazy=False) for N, nw in sizes]
    ma, mb = MultiStep(a, dt), MultiStep(b, dt)
    for _ in range(4):
        ma()
        mb()
    torch.cuda.synchronize()
    for ga, gb in zip(a, b):
        _same(ga, gb, f"N={ga.N}")
