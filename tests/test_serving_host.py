"""Host-side logic of the serving layer (no GPU): the wire-off gate of a
request's pulse and the footprint arithmetic used for admission."""

import numpy as np


def test_serving_gate_open_time_brute_force():
    """The wire-off gate opens at the first step with |E| >= 1e-3 Emax0
    (SBEs.py:5849-5858) for the sbetestprop pulse (sbetestprop.py:183-198)."""
    from pulsesuite_amd.serving import SimRequest
    r = SimRequest(rid=0, nk=64, e0=2e7, fwhm=20e-15, detuning_ev=0.01, nsteps=10)
    dt = 1e-16
    t_open = r.gate_open_time(1.5, dt)
    E0x, E0y, E0z, w0, tw, tp = r.pulse(1.5)
    emax = np.sqrt(E0x**2 + E0y**2)

    def mag(t):
        x = w0 * (t - tp)
        f = np.exp(-(x**2) / (w0 * tw) ** 2) * np.cos(x) * np.exp(-((x / (2 * tw * w0)) ** 20))
        return abs(f) * emax

    n = int(round(t_open / dt))
    assert mag(n * dt) >= 1e-3 * emax
    assert all(mag(m * dt) < 1e-3 * emax for m in range(n))
    assert 0 < t_open < tp


def test_slot_footprint_grows_with_nk_and_bf16_planes():
    from pulsesuite_amd.serving import slot_bytes
    a = slot_bytes(64, 4096, "fp64", 1000)
    b = slot_bytes(128, 4096, "fp64", 1000)
    c = slot_bytes(128, 4096, "bf16x3", 1000)
    # 12 complex N x N matrices dominate: 4x per doubling of N, plus the
    # planes (the fields and the split-K scratch scale with the propagation grid)
    fixed = 4096 * 7 * 16 + 8 * 2 * 4096 * 5 * 8
    assert 3.5 < (b - fixed) / (a - fixed) < 4.5
    assert c - b == 14 * 4 * 128 * 128 * 2


# ---------------------------------------------------------------------------
# Failure handling of the request loop, on the CPU: a pool with the slot
# interface of SlotPool whose "device" is a per-slot step counter; a request
# with a NaN amplitude records NaN observables, as a diverged wire would.
# ---------------------------------------------------------------------------

class FakePool:
    def __init__(self, nk, cap, rr, dt, prec, maxsteps, graphs=True, prm_kw=None, integrator="leapfrog"):
        if nk % 32:
            raise ValueError(f"k-grid gave N={nk + 1}, wanted {nk}")
        self.nk, self.cap, self.maxsteps = nk, int(cap), int(maxsteps)
        self.slots = [None] * self.cap
        self.done_steps = [0] * self.cap
        self.rows = [[] for _ in range(self.cap)]
        self.ticks = 0
        self.admitted = []

    def footprint(self):
        return 1000 * self.cap

    def free(self):
        return sum(1 for s in self.slots if s is None)

    def active(self):
        return sum(1 for s in self.slots if s is not None)

    def prefix(self):
        return max((i + 1 for i, s in enumerate(self.slots) if s is not None), default=0)

    def admit(self, req, now):
        assert 1 <= req.nsteps <= self.maxsteps
        i = self.slots.index(None)
        self.slots[i], self.done_steps[i], self.rows[i] = req, 0, []
        req.t_admit, req.status = now, "running"
        self.admitted.append((req.rid, i))
        return i

    def release(self, i):
        r, self.slots[i] = self.slots[i], None
        return r

    def tick(self):
        for i, r in enumerate(self.slots):
            if r is not None:
                self.done_steps[i] += 1
                self.rows[i].append([r.e0 * self.done_steps[i], 0.0, 0.1, 0.1])
        self.ticks += 1
        return self.prefix()

    def nonfinite(self):
        return [i for i, r in enumerate(self.slots)
                if r is not None and self.rows[i] and not np.isfinite(self.rows[i][-1]).all()]

    def harvest(self, now_fn):
        out = []
        for i, r in enumerate(self.slots):
            if r is not None and self.done_steps[i] >= r.nsteps:
                r.trace = np.array(self.rows[i])
                r.final = {}
                r.status = "done" if np.isfinite(r.trace).all() else "failed"
                r.t_done = now_fn()
                self.slots[i] = None
                out.append(r)
        return out


def _server(**kw):
    from pulsesuite_amd.serving import Server
    kw.setdefault("hbm_budget", 1 << 40)
    return Server(np.linspace(-1e-6, 1e-6, 64), dt=1e-17, prec="fp64", max_slots=4, maxsteps=50,
                  pool_factory=FakePool, **kw)


def _reqs(specs):
    from pulsesuite_amd.serving import SimRequest
    return [SimRequest(rid=i, nk=nk, e0=e0, fwhm=1e-14, detuning_ev=0.0, nsteps=n, arrival=a)
            for i, (nk, e0, n, a) in enumerate(specs)]


def test_impossible_requests_are_rejected_before_the_clock():
    """A step count beyond the trace buffer (or zero), an unsupported k-grid
    and an Nk with no HBM room are rejected with a reason; the others run;
    the loop ends (no request waits forever)."""
    reqs = _reqs([(64, 1.0, 10, 0.0), (64, 1.0, 51, 0.0), (64, 1.0, 0, 0.0), (100, 1.0, 10, 0.0),
                  (512, 1.0, 10, 0.0)])
    from pulsesuite_amd.serving import pool_bytes, slot_bytes
    # room for the Nk = 64 pool, not for a slot of Nk = 512 after it
    budget = pool_bytes(64) + 4 * slot_bytes(64, 64, "fp64", 50) + 4000 + pool_bytes(512) + 1
    srv = _server(hbm_budget=budget)
    done = srv.run(reqs)
    assert [r.rid for r in done] == [0]
    why = {r.rid: r.error for r in srv.rejected}
    assert set(why) == {1, 2, 3, 4}
    assert "outside [1, 50]" in why[1] and "outside [1, 50]" in why[2]
    assert "not supported" in why[3]
    assert "no HBM room" in why[4]
    assert all(r.status == "rejected" for r in srv.rejected)
    assert done[0].status == "done" and len(done[0].trace) == 10


def test_no_budget_rejects_everything_and_returns():
    srv = _server(hbm_budget=1000)
    reqs = _reqs([(64, 1.0, 10, 0.0), (128, 1.0, 10, 0.0)])
    assert srv.run(reqs) == []
    assert sorted(r.rid for r in srv.rejected) == [0, 1]


def test_cancel_waiting_and_in_flight():
    """Cancelling an in-flight request frees its slot for a waiting one;
    cancelling a waiting request keeps it from ever running."""
    reqs = _reqs([(64, 1.0, 40, 0.0), (64, 1.0, 40, 0.0), (64, 1.0, 40, 0.0), (64, 1.0, 40, 0.0),
                  (64, 2.0, 10, 0.0), (64, 3.0, 10, 0.0)])
    srv = _server()
    calls = []

    def on_tick(s, t):
        calls.append(t)
        if len(calls) == 5:
            s.cancel(1)   # in flight in slot 1
            s.cancel(5)   # still waiting (four slots, all busy)

    done = srv.run(reqs, on_tick=on_tick)
    st = {r.rid: r.status for r in done + srv.cancelled}
    assert st == {0: "done", 2: "done", 3: "done", 4: "done", 1: "cancelled", 5: "cancelled"}
    pool = srv.pools[64]
    assert (4, 1) in pool.admitted            # rid 4 took the freed slot 1
    assert all(rid != 5 for rid, _ in pool.admitted)


def test_nonfinite_request_is_evicted_and_others_finish():
    reqs = _reqs([(64, 1.0, 30, 0.0), (64, float("nan"), 30, 0.0), (64, 2.0, 30, 0.0),
                  (64, 4.0, 5, 0.0), (64, 5.0, 5, 0.0)])
    srv = _server(nan_check=4)
    done = srv.run(reqs)
    assert sorted(r.rid for r in done) == [0, 2, 3, 4]
    assert [r.rid for r in srv.failed] == [1]
    assert srv.failed[0].status == "failed" and "non-finite" in srv.failed[0].error
    # evicted at the first check (tick 4), long before its 30 steps
    assert srv.pools[64].ticks < 30 + 5 + 5
    for r in done:
        assert np.isfinite(r.trace).all() and len(r.trace) == r.nsteps
