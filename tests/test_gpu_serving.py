"""Serving: continuous batching over slot pools (GPU).

Each request is an independent single-wire simulation, so its result must
not depend on what else is in flight: the same request alone and in a mix
of others (same and different Nk, admitted into other slots at other
ticks) gives the same trace and final state within 1e-12 relative.  The
graph-replayed ticks must reproduce the eager ticks exactly."""

import math

import numpy as np
import pytest

from oracle import sbe_oracle as O
from tests.helpers import params_pair, pulse, rel, wire_length

pytestmark = pytest.mark.gpu

RR = np.linspace(-2e-6, 2e-6, 512)
DT = 10e-18


def _req(rid, nk, e0=3e7, fwhm=2e-15, det=0.02, nsteps=40, arrival=0.0):
    from pulsesuite_amd.serving import SimRequest
    return SimRequest(rid=rid, nk=nk, e0=e0, fwhm=fwhm, detuning_ev=det, nsteps=nsteps,
                      arrival=arrival)


def _serve(reqs, graphs=True, prec="fp64", integrator="leapfrog", budget=8 << 30, on_tick=None, nan_check=64):
    from pulsesuite_amd.serving import Server
    srv = Server(RR, dt=DT, prec=prec, hbm_budget=budget, max_slots=16, maxsteps=200,
                 graphs=graphs, integrator=integrator, nan_check=nan_check)
    done = srv.run(reqs, on_tick=on_tick)
    return {r.rid: r for r in done}, srv


@pytest.mark.parametrize("integrator", ["leapfrog", "rk4"])
def test_graph_replay_matches_eager(integrator):
    mk = lambda: [_req(0, 64, nsteps=30), _req(1, 64, e0=1e7, nsteps=45),  # noqa: E731
                  _req(2, 64, fwhm=3e-15, nsteps=20)]
    print(f"[serve] eager {integrator}", flush=True)
    a, _ = _serve(mk(), graphs=False, integrator=integrator)
    print(f"[serve] graphs {integrator}", flush=True)
    b, srv = _serve(mk(), graphs=True, integrator=integrator)
    assert sum(len(p.graphs) for p in srv.pools.values()) > 0
    for rid in a:
        assert a[rid].trace.shape == b[rid].trace.shape
        assert np.array_equal(a[rid].trace, b[rid].trace), rid
        assert np.array_equal(a[rid].final["p"], b[rid].final["p"]), rid


@pytest.mark.parametrize("prec,integrator", [("fp64", "leapfrog"), ("bf16x3", "leapfrog"), ("bf16x3", "rk4")])
def test_request_same_alone_and_in_mix(prec, integrator):
    print(f"[serve] alone {prec} {integrator}", flush=True)
    alone, _ = _serve([_req(7, 64, nsteps=50)], prec=prec, integrator=integrator)
    mix = [_req(1, 64, e0=5e7, nsteps=80), _req(2, 128, nsteps=40), _req(3, 64, fwhm=3e-15, nsteps=30),
           _req(7, 64, nsteps=50, arrival=0.002), _req(4, 128, e0=1e7, nsteps=60, arrival=0.001)]
    print(f"[serve] mix {prec} {integrator}", flush=True)
    got, srv = _serve(mix, prec=prec, integrator=integrator)
    assert set(got) == {1, 2, 3, 4, 7}
    a, b = alone[7], got[7]
    assert np.isfinite(a.trace).all()
    # the pulse drives a polarisation and the densities move
    assert np.abs(a.trace[:, 0]).max() > 0
    assert rel(b.trace, a.trace) < 1e-12
    for k in ("ne", "nh", "p"):
        assert rel(b.final[k], a.final[k]) < 1e-12, k
    lat = [r.t_done - r.arrival for r in got.values()]
    assert min(lat) > 0


@pytest.mark.parametrize("prec", ["fp64", "bf16x3"])
def test_merged_ticks_equal_per_pool_ticks(prec):
    """Stepping every pool with one persistent multi-group RHS launch
    (Server(merged=True)) gives the per-pool results bit for bit."""
    from pulsesuite_amd.serving import Server
    mk = lambda: [_req(1, 64, e0=5e7, nsteps=40), _req(2, 128, nsteps=30),  # noqa: E731
                  _req(3, 64, fwhm=3e-15, nsteps=25), _req(4, 128, e0=1e7, nsteps=35)]
    print(f"[serve] per-pool {prec}", flush=True)
    a, _ = _serve(mk(), prec=prec)
    print(f"[serve] merged {prec}", flush=True)
    srv = Server(RR, dt=DT, prec=prec, hbm_budget=8 << 30, max_slots=16, maxsteps=200, merged=True)
    b = {r.rid: r for r in srv.run(mk())}
    assert set(a) == set(b) == {1, 2, 3, 4}
    assert len(srv._multi) > 0 and len(srv._mgraphs) > 0
    for rid in a:
        assert np.array_equal(a[rid].trace, b[rid].trace), rid
        for k in ("ne", "nh", "p"):
            assert np.array_equal(a[rid].final[k], b[rid].final[k]), (rid, k)


def test_served_request_matches_oracle():
    """One served Nk = 64 request (FP64) against the oracle solver driven by
    the same sbetestprop pulse (sbetestprop.py:183-198, Ey = 2 Ex, Ez = 0)
    at the request's clock (first step at the wire-off gate's opening):
    the recorded Pxx at the middle point and the carrier total of every
    step within 1e-9.  This pins the device pulse (psa_drive) and the
    per-step record (psa_record)."""
    print("[serve] served request vs oracle", flush=True)
    req = _req(0, 64, e0=3e7, fwhm=2e-15, det=0.02, nsteps=40)
    got, srv = _serve([req])
    tr = got[0].trace
    pool = srv.pools[64]
    prm, oprm = params_pair(wire_length(64))
    E0x, E0y, E0z, w0, tw, tp = req.pulse(prm.gap)
    t0 = req.gate_open_time(prm.gap, DT)
    sv = O.Solver(oprm, 1, RR, math.sqrt(E0x**2 + E0y**2 + E0z**2))
    xmid = len(RR) // 2
    ones = np.ones(len(RR), dtype=complex)
    for n in range(req.nsteps):
        Ex = pulse(t0 + n * DT, E0x, tw, tp, w0)
        out = sv.qw_calculator(Ex * ones, 2.0 * Ex * ones, 0.0 * ones, DT, 1)
        assert out is not None, n
        ref_p = out[0][xmid]
        ref_ne = np.sum(np.abs(np.diag(sv.CC2[:, :, 0])))
        ref_nh = np.sum(np.abs(np.diag(sv.DD2[:, :, 0])))
        assert abs(tr[n, 0] + 1j * tr[n, 1] - ref_p) <= 1e-9 * np.max(np.abs(out[0])), n
        assert abs(tr[n, 2] - ref_ne) <= 1e-9 * ref_ne, n
        assert abs(tr[n, 3] - ref_nh) <= 1e-9 * ref_nh, n
    assert pool.xmid == xmid


def test_request_same_alone_and_in_mix_large_nk():
    """Alone versus in a mix with Nk = 256 and 512 requests (split-BF16):
    the Nk = 256 request's trace and final state agree within 1e-12."""
    print("[serve] alone Nk=256 bf16x3", flush=True)
    alone, _ = _serve([_req(7, 256, nsteps=20)], prec="bf16x3", budget=16 << 30)
    mix = [_req(1, 512, e0=5e7, nsteps=20), _req(2, 256, nsteps=30), _req(7, 256, nsteps=20, arrival=0.002),
           _req(4, 512, e0=1e7, nsteps=15, arrival=0.001), _req(5, 64, nsteps=25)]
    print("[serve] mix Nk=64/256/512 bf16x3", flush=True)
    got, _ = _serve(mix, prec="bf16x3", budget=16 << 30)
    assert set(got) == {1, 2, 4, 5, 7}
    a, b = alone[7], got[7]
    assert np.isfinite(a.trace).all() and np.abs(a.trace[:, 0]).max() > 0
    assert rel(b.trace, a.trace) < 1e-12
    for k in ("ne", "nh", "p"):
        assert rel(b.final[k], a.final[k]) < 1e-12, k


def test_nonfinite_request_is_isolated_on_the_device():
    """A NaN written from the host into one in-flight request's state (as a
    diverged wire would hold) makes that request fail; every other request
    of the same pool gives exactly the results of a run without it."""
    import torch
    mk = lambda: [_req(1, 64, e0=5e7, nsteps=40), _req(2, 64, nsteps=40), _req(3, 64, fwhm=3e-15, nsteps=40),  # noqa: E731
                  _req(4, 128, nsteps=30)]
    print("[serve] clean mix", flush=True)
    clean, _ = _serve(mk(), nan_check=8)
    calls = []

    def inject(srv, t):
        calls.append(t)
        if len(calls) == 10:
            pool = srv.pools[64]
            i = next(i for i, r in enumerate(pool.slots) if r is not None and r.rid == 2)
            with torch.cuda.stream(pool.stream):
                pool.group.s["CB"][i, 3, 3] = float("nan")
                pool.group.s["CA"][i, 3, 3] = float("nan")
                pool.group.invalidate(i)

    print("[serve] mix with a NaN injected into request 2", flush=True)
    got, srv = _serve(mk(), on_tick=inject, nan_check=8)
    assert [r.rid for r in srv.failed] == [2]
    assert srv.failed[0].status == "failed"
    assert set(got) == {1, 3, 4}
    for rid in got:
        assert np.array_equal(got[rid].trace, clean[rid].trace), rid
        for k in ("ne", "nh", "p"):
            assert np.array_equal(got[rid].final[k], clean[rid].final[k]), (rid, k)
