"""HIP-graph capture of the RK4 wire step and of the coupled wires + field
step: the graph replayed several times, with eager work and
synchronisations between the replays, reproduces the eager run bit for
bit (no memset nodes; the record index and the step counter live on the
device)."""

import numpy as np
import pytest

from tests.helpers import params_pair, random_state, wire_length

pytestmark = pytest.mark.gpu


def _group(N, nw, rr, seed, integrator="leapfrog"):
    from pulsesuite_amd.engine import WireGroup
    prm, _ = params_pair(wire_length(N))
    wg = WireGroup(prm, nw, rr, prec="fp64", integrator=integrator)
    wg.set_state(*random_state(np.random.default_rng(seed), N, nw))
    return wg


def _noise():
    import torch
    x = torch.randn(1 << 16, device="cuda", dtype=torch.float64)
    return float((x * x).sum())


def test_rk4_step_graph_replay_matches_eager():
    import torch
    N, nw, dt = 64, 4, 0.2e-15
    L = wire_length(N)
    rr = np.linspace(-1.2 * L, 1.2 * L, 96)
    E = 2e7 * np.cos(2 * np.pi * rr / (3 * L)) + 0j
    print("[graph rk4] eager", flush=True)
    a = _group(N, nw, rr, 5, "rk4")
    a.set_fields(E, 0.5 * E, 0 * E)
    for _ in range(8):
        a.step(dt)
    torch.cuda.synchronize()
    print("[graph rk4] capture", flush=True)
    b = _group(N, nw, rr, 5, "rk4")
    b.set_fields(E, 0.5 * E, 0 * E)
    torch.cuda.synchronize()
    s = torch.cuda.Stream()
    with torch.cuda.stream(s):
        for _ in range(2):
            b.step(dt)
    torch.cuda.synchronize()
    g = torch.cuda.CUDAGraph()
    with torch.cuda.graph(g, stream=s):
        b.step(dt)
    for _ in range(6):
        g.replay()
        torch.cuda.synchronize()
        _noise()
    torch.cuda.synchronize()
    for x, y in zip(a.get_state()[3:], b.get_state()[3:]):
        assert np.array_equal(x, y)
    for k in range(3):
        assert np.array_equal(a.outputs()[k], b.outputs()[k])


def test_pstd_coupled_step_graph_replay_matches_eager():
    import torch
    from pulsesuite_amd.pstd import Pstd1D, abc_profile
    N, nw, dt = 32, 2, 5e-17
    L = wire_length(N)
    rr = np.linspace(-1.2 * L, 1.2 * L, 64)
    Np, dy = 256, 2e-9
    abc = abc_profile(Np, dy, dt, 3.1, 60e-9, 60e-9)

    def setup():
        wg = _group(N, nw, rr, 9)
        p = Pstd1D(Np, dy, 3.1, dt, wires_y=[200e-9, 230e-9], sheet=1e4, abc=abc,
                   obs_y=[40e-9, 400e-9], nrec=16)
        E, H = p.pulse(1e8, 150e-9, 40e-9, 800e-9, "x")
        p.set_fields(E, H)
        p.attach([(wg, 0)])
        return wg, p

    print("[graph pstd] eager", flush=True)
    wa, pa = setup()
    for _ in range(8):
        pa.couple_step()
    torch.cuda.synchronize()
    print("[graph pstd] capture", flush=True)
    wb, pb = setup()
    torch.cuda.synchronize()
    s = torch.cuda.Stream()
    with torch.cuda.stream(s):
        for _ in range(2):
            pb.couple_step()
    torch.cuda.synchronize()
    g = torch.cuda.CUDAGraph()
    with torch.cuda.graph(g, stream=s):  # two steps: the level swap returns to this parity
        pb.couple_step()
        pb.couple_step()
    for _ in range(3):
        g.replay()
        torch.cuda.synchronize()
        _noise()
    torch.cuda.synchronize()
    assert int(pb.t["count"].item()) == 8
    assert np.array_equal(pa.records(), pb.records())
    assert np.array_equal(pa.t["state"].cpu().numpy(), pb.t["state"].cpu().numpy())
    for x, y in zip(wa.get_state(), wb.get_state()):
        assert np.array_equal(x, y)
