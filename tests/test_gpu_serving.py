"""Serving: continuous batching over slot pools (GPU).

Each request is an independent single-wire simulation, so its result must
not depend on what else is in flight: the same request alone and in a mix
of others (same and different Nk, admitted into other slots at other
ticks) gives the same trace and final state within 1e-12 relative.  The
graph-replayed ticks must reproduce the eager ticks exactly."""

import numpy as np
import pytest

from tests.helpers import rel

pytestmark = pytest.mark.gpu

RR = np.linspace(-2e-6, 2e-6, 512)
DT = 10e-18


def _req(rid, nk, e0=3e7, fwhm=2e-15, det=0.02, nsteps=40, arrival=0.0):
    from pulsesuite_amd.serving import SimRequest
    return SimRequest(rid=rid, nk=nk, e0=e0, fwhm=fwhm, detuning_ev=det, nsteps=nsteps,
                      arrival=arrival)


def _serve(reqs, graphs=True, prec="fp64", integrator="leapfrog"):
    from pulsesuite_amd.serving import Server
    srv = Server(RR, dt=DT, prec=prec, hbm_budget=8 << 30, max_slots=16, maxsteps=200,
                 graphs=graphs, integrator=integrator)
    done = srv.run(reqs)
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
