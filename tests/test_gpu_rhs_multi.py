"""The persistent multi-group RHS (psa_rhs_multi through the C ABI): three
wire groups of different N in one launch over a work queue give exactly the
per-group psa_rhs results, in FP64, FP32 and split-BF16, for the raw RHS
(mode 0) and the leapfrog write (mode 1); repeated launches (the queue
resets itself) and a graph replay give the same bits; the multi-group
step (its groups on streams of their own) equals the per-group steps, also
replayed from a graph.  Uneven Nk
{64, 128, 192, 256} stepped together in the reduced modes stay within
1e-6 of FP64 in density and polarization."""

import numpy as np
import pytest

from tests.helpers import params_pair, random_state, wire_length

pytestmark = pytest.mark.gpu

SIZES = ((64, 3), (96, 2), (128, 4))


def _groups(prec, seed=3, sizes=SIZES):
    from pulsesuite_amd.engine import WireGroup
    out = []
    rng = np.random.default_rng(seed)
    for N, nw in sizes:
        prm, _ = params_pair(wire_length(N))
        rr = np.linspace(-1.2 * wire_length(N), 1.2 * wire_length(N), 96)
        wg = WireGroup(prm, nw, rr, prec=prec)
        wg.set_state(*random_state(rng, N, nw))
        E = 2e7 * np.cos(2 * np.pi * rr / (3 * wire_length(N))) + 0j
        wg.set_fields(E, 0.5 * E, 0 * E)
        wg.prop2qw()
        wg.prep()
        wg.calc_h()
        wg.offdiag()
        out.append(wg)
    return out


def _levels(gs):
    return [[g.s[k].cpu().numpy().copy() for k in ("pX", "CX", "DX")] for g in gs]


@pytest.mark.parametrize("prec", ["fp64", "fp32", "bf16x3"])
@pytest.mark.parametrize("mode", [0, 1])
def test_multi_rhs_equals_per_group(prec, mode):
    import torch
    from pulsesuite_amd.engine import MultiRhs
    print(f"[rhs_multi] prec={prec} mode={mode}", flush=True)
    gs = _groups(prec)
    dt = 10e-18
    for g in gs:
        g.rhs(mode, dt)
    torch.cuda.synchronize()
    ref = _levels(gs)
    for g in gs:
        for k in ("pX", "CX", "DX"):
            g.s[k].zero_()
    multi = MultiRhs(gs, mode=mode, dt=dt)
    for rep in range(3):
        multi()
        torch.cuda.synchronize()
        got = _levels(gs)
        for a, b in zip(got, ref):
            for x, y in zip(a, b):
                assert np.array_equal(x, y), rep
        assert int(multi.queue.abs().sum().item()) == 0
    # captured in a graph and replayed
    for g in gs:
        for k in ("pX", "CX", "DX"):
            g.s[k].zero_()
    torch.cuda.synchronize()
    s = torch.cuda.Stream()
    torch.cuda.synchronize()
    graph = torch.cuda.CUDAGraph()
    with torch.cuda.graph(graph, stream=s):
        multi()
    for _ in range(2):
        graph.replay()
        torch.cuda.synchronize()
        for a, b in zip(_levels(gs), ref):
            for x, y in zip(a, b):
                assert np.array_equal(x, y)


@pytest.mark.parametrize("prec", ["fp64", "fp32", "bf16x3"])
def test_multi_step_equals_per_group_steps(prec):
    import torch
    from pulsesuite_amd.engine import MultiStep
    print(f"[step_multi] prec={prec}", flush=True)
    dt = 10e-18
    a = _groups(prec, seed=5)
    b = _groups(prec, seed=5)
    for _ in range(4):
        for g in a:
            g.step(dt)
    ms = MultiStep(b, dt)
    for _ in range(4):
        ms()
    torch.cuda.synchronize()
    assert len(ms._packs) == 2  # one description per level parity
    for ga, gb in zip(a, b):
        for x, y in zip(ga.get_state(), gb.get_state()):
            assert np.array_equal(x, y)
        for x, y in zip(ga.outputs(), gb.outputs()):
            assert np.array_equal(x, y)


def _noise():
    import torch
    x = torch.randn(1 << 16, device="cuda", dtype=torch.float64)
    return float((x * x).sum())


@pytest.mark.parametrize("prec", ["fp64", "bf16x3"])
def test_multi_step_graph_replay_matches_eager(prec):
    """psa_step_multi runs each group's stages on a stream of its own; a
    graph captured after eager calls (the streams exist, so the capture
    forks into them and joins back) and replayed three times, with eager
    work and synchronisations between the replays, gives the eager run's
    bits."""
    import torch
    from pulsesuite_amd.engine import MultiStep
    print(f"[step_multi graph] prec={prec}", flush=True)
    dt = 10e-18
    a = _groups(prec, seed=7)
    ma = MultiStep(a, dt)
    for _ in range(8):
        ma()
    torch.cuda.synchronize()
    b = _groups(prec, seed=7)
    mb = MultiStep(b, dt)
    s = torch.cuda.Stream()
    with torch.cuda.stream(s):
        for _ in range(2):
            mb()
    torch.cuda.synchronize()
    g = torch.cuda.CUDAGraph()
    with torch.cuda.graph(g, stream=s):  # two steps: back to this level parity
        mb()
        mb()
    for _ in range(3):
        g.replay()
        torch.cuda.synchronize()
        _noise()
    torch.cuda.synchronize()
    for ga, gb in zip(a, b):
        for x, y in zip(ga.get_state(), gb.get_state()):
            assert np.array_equal(x, y)
        for x, y in zip(ga.outputs(), gb.outputs()):
            assert np.array_equal(x, y)


UNEVEN = ((64, 2), (128, 2), (192, 2), (256, 2))


@pytest.mark.parametrize("prec", ["fp32", "bf16x3"])
def test_uneven_nk_reduced_precision_vs_fp64(prec):
    """Wires of Nk 64, 128, 192, 256 stepped together (psa_step_multi, the
    persistent RHS) for 20 leapfrog steps in a reduced mode against the
    same groups in FP64: the densities (diagonals of C, D), the coherences
    p and the polarization Pxx stay within 1e-6 relative per group."""
    import torch
    from pulsesuite_amd.engine import MultiStep
    print(f"[uneven] prec={prec}", flush=True)
    dt = 10e-18
    ref = _groups("fp64", seed=9, sizes=UNEVEN)
    red = _groups(prec, seed=9, sizes=UNEVEN)
    a, b = MultiStep(ref, dt), MultiStep(red, dt)
    for _ in range(20):
        a()
        b()
    torch.cuda.synchronize()
    for ga, gb in zip(ref, red):
        Sa, Sb = ga.get_state(), gb.get_state()
        for k, nm in ((3, "C3"), (4, "D3")):
            na = np.einsum("iiw->iw", Sa[k])
            nb = np.einsum("iiw->iw", Sb[k])
            e = float(np.max(np.abs(nb - na)) / np.max(np.abs(na)))
            print(f"  N={ga.N} {nm} density err {e:.2e}", flush=True)
            assert e < 1e-6
        e = float(np.max(np.abs(Sb[5] - Sa[5])) / np.max(np.abs(Sa[5])))
        ep = float(np.max(np.abs(gb.outputs()[0] - ga.outputs()[0])) / np.max(np.abs(ga.outputs()[0])))
        print(f"  N={ga.N} p err {e:.2e} Pxx err {ep:.2e}", flush=True)
        assert e < 1e-6 and ep < 1e-6
