"""Continuous-batching server for many independent SBE simulations on one GPU.

A simulation request is one quantum wire driven by the sbetestprop pulse
(``sbetestprop.py:183-198``: Ey = 2 Ex as there, Ez = 0, spatially
uniform) for a given number of active steps, recording Pxx at the middle
point and the carrier totals every step (``sbetestprop.py:210-224``).  Every
simulation is a single-wire solver of its own, so all run with the
reference wire index w = 1.

Design (one process, one GPU):

* One ``SlotPool`` per k-grid size Nk: a ``WireGroup`` with a fixed number
  of wire slots, allocated once.  Its capacity comes from the HBM budget
  (``footprint`` per slot + the pool's shared tables).  The density
  matrices of the requests in flight stay resident in their slots.
* Continuous batching: requests are admitted into free slots between ticks
  (lowest free slot first, so the occupied slots form a short prefix);
  a tick advances every occupied slot of every pool by one step.  The
  RHS and all other stages of all in-flight requests of one Nk run in the
  same kernel launches; pools of different Nk run on their own HIP streams,
  so their kernels overlap on the device.
* A tick of a pool is drive -> QWCalculator step -> record on the device
  (``psa_drive``, ``psa_step`` -- or ``psa_step_rk4`` for a server built
  with ``integrator="rk4"`` -- and ``psa_record``) for the occupied prefix,
  optionally replayed from a captured HIP graph (one per pointer parity of
  the leapfrog's level swap and prefix length; RK4 updates the state in
  place, so one per prefix length).
* Requests whose step count is reached are harvested (trace and final
  populations copied out), their slot is freed and reused.
* ``Server(..., merged=True)`` steps all pools together instead: the drives,
  one ``psa_step_multi`` whose RHS covers the occupied slots of every pool
  in one persistent launch over per-XCD work queues, and the records, on one
  stream, replayed from a graph per (Nk, parity, prefix) combination; the
  results equal the per-pool ticks exactly.

Before its first active step a request waits for the reference's wire-off
gate (``SBEs.py:5849-5858``): steps with |E| < 1e-3 Emax0 before the field
first exceeds that do nothing, so the clock starts at the first open step.

Failure handling.  Every request ends with a ``status``: ``"done"``,
``"rejected"`` (before the clock starts, with the reason in ``error``: a
step count beyond the trace buffer or below one, a k-grid the device path
does not support, or no room for a slot of its Nk in the HBM budget --
nothing waits forever), ``"cancelled"`` (``Server.cancel(rid)``, while
waiting or in flight: its slot is freed at once and reset when reused) or
``"failed"`` (its recorded observables became non-finite: checked every
``nan_check`` ticks and at harvest; the slot is freed).  Wires never share
data inside a step, so a request that fails leaves the others' results
unchanged.
"""

import ctypes
import math
import time
from dataclasses import dataclass, field

import numpy as np
import torch
from scipy.constants import e as e0
from scipy.constants import hbar

from . import capi, tables
from .engine import MultiRhs, WireGroup


@dataclass
class SimRequest:
    """One simulation: wire k-grid size, pulse and step count."""

    rid: int
    nk: int
    e0: float                 # peak Ex (V/m); Ey = 2 e0, Ez = 0
    fwhm: float               # field FWHM (s); tw = fwhm / (2 sqrt(ln 2))
    detuning_ev: float        # hbar omega0 = gap + detuning
    nsteps: int               # active steps
    arrival: float = 0.0      # seconds after the server clock starts
    # filled by the server
    t_admit: float = field(default=None, repr=False)
    t_done: float = field(default=None, repr=False)
    trace: np.ndarray = field(default=None, repr=False)
    final: dict = field(default=None, repr=False)
    status: str = field(default="pending", repr=False)  # done, rejected, cancelled, failed
    error: str = field(default=None, repr=False)

    def pulse(self, gap_ev):
        """(E0x, E0y, E0z, omega0, tw, tp)."""
        tw = self.fwhm / (2.0 * math.sqrt(math.log(2.0)))
        omega0 = (gap_ev + self.detuning_ev) * e0 / hbar
        return self.e0, 2.0 * self.e0, 0.0, omega0, tw, 3.0 * self.fwhm

    def gate_open_time(self, gap_ev, dt):
        """First step time at which |E| >= 1e-3 Emax0 (the wire-off gate)."""
        E0x, E0y, E0z, w0, tw, tp = self.pulse(gap_ev)
        emax = math.sqrt(E0x**2 + E0y**2 + E0z**2)
        n = np.arange(int(math.ceil(2 * tp / dt)) + 2)
        t = n * dt
        x = w0 * (t - tp)
        env = np.exp(-(x**2) / (w0 * tw) ** 2) * np.cos(x) * np.exp(-((x / (2 * tw * w0)) ** 20))
        mag = np.abs(env) * emax
        open_ = np.nonzero(mag >= 1e-3 * emax)[0]
        return float(t[open_[0]]) if open_.size else 0.0


def wire_length(nk, me_rel=0.067, gap_ev=1.5):
    """Wire length giving exactly nk k-points (SBEs.py:4754-4758)."""
    from scipy.constants import m_e
    kmax = math.sqrt(1.2 * gap_ev * e0 * 2.0 * me_rel * m_e / hbar**2)
    return (nk / 2 + 0.5) * math.pi / kmax


class SlotPool:
    """Fixed-capacity slots of one Nk on one stream."""

    def __init__(self, nk, capacity, rr, dt, prec, maxsteps, graphs=True, prm_kw=None, autotune=True,
                 integrator="leapfrog"):
        prm_kw = dict(prm_kw or {})
        prm_kw.setdefault("me_rel", 0.067)
        self.prm = tables.QWParams(L=wire_length(nk, prm_kw["me_rel"]), **prm_kw)
        self.group = WireGroup(self.prm, capacity, rr, wire_ids=np.ones(capacity, np.int32), prec=prec,
                               integrator=integrator)
        self.rk4 = integrator == "rk4"
        if self.group.N != nk:
            raise ValueError(f"k-grid gave N={self.group.N}, wanted {nk}")
        self.nk, self.cap, self.dt = nk, int(capacity), float(dt)
        self.maxsteps = int(maxsteps)
        d = self.group.dev
        self.drive = torch.zeros(self.cap, 8, dtype=torch.float64, device=d)
        self.trace = torch.zeros(self.cap, self.maxsteps, 4, dtype=torch.float64, device=d)
        self.count = torch.zeros(self.cap, dtype=torch.int32, device=d)
        self.slots = [None] * self.cap
        self.done_steps = [0] * self.cap
        self.stream = torch.cuda.Stream(device=d)
        self.use_graphs = graphs
        self.graphs = {}
        self.parity = 0
        self.xmid = len(rr) // 2
        self.lib = self.group.lib
        self.ticks = 0
        if autotune:
            self.tuned = self.group.autotune()

    def footprint(self):
        """Device bytes held by this pool."""
        return self.group.footprint_bytes() + self.trace.numel() * 8

    # -- slots -----------------------------------------------------------
    def free(self):
        return sum(1 for s in self.slots if s is None)

    def active(self):
        return sum(1 for s in self.slots if s is not None)

    def prefix(self):
        hi = 0
        for i, s in enumerate(self.slots):
            if s is not None:
                hi = i + 1
        return hi

    def admit(self, req, now):
        """Put `req` into the lowest free slot (the server has checked that
        its step count fits the trace buffer)."""
        if not 1 <= req.nsteps <= self.maxsteps:
            raise ValueError("request step count outside the pool's trace buffer")
        i = self.slots.index(None)
        with torch.cuda.stream(self.stream):
            self.group.reset_wire(i)
            t0 = req.gate_open_time(self.prm.gap, self.dt)
            row = list(req.pulse(self.prm.gap)) + [t0, 0.0]
            self.drive[i].copy_(torch.tensor(row, dtype=torch.float64))
            self.count[i] = 0
        self.slots[i] = req
        self.done_steps[i] = 0
        req.t_admit = now
        req.status = "running"
        return i

    def release(self, i):
        """Free slot i without harvesting it (cancellation, failure); the
        slot's state is reset when it is next admitted."""
        req = self.slots[i]
        self.slots[i] = None
        return req

    def nonfinite(self):
        """Slots whose last recorded step holds a non-finite observable
        (one small gather and copy, on this pool's stream)."""
        hi = self.prefix()
        if hi == 0:
            return []
        with torch.cuda.stream(self.stream):
            cnt = self.count[:hi].long()
            rows = self.trace[torch.arange(hi, device=cnt.device), (cnt - 1).clamp(min=0)]
            # a slot with nothing recorded yet holds a previous request's rows
            ok = (torch.isfinite(rows).all(dim=1) | (cnt == 0)).cpu().numpy()
        return [i for i in range(hi) if self.slots[i] is not None and not ok[i]]

    # -- one tick ----------------------------------------------------------
    def _launch(self, hi, S, stream_handle):
        c = capi.check
        c(self.lib.psa_drive(S, 0, hi, self.drive.data_ptr(), self.dt, stream_handle))
        step = self.lib.psa_step_rk4 if self.rk4 else self.lib.psa_step
        c(step(S, 0, hi, self.dt, stream_handle))
        c(self.lib.psa_record(S, 0, hi, self.xmid, self.trace.data_ptr(), self.count.data_ptr(),
                              self.maxsteps, stream_handle))

    def tick(self):
        """Advance every occupied slot by one step (asynchronous)."""
        hi = self.prefix()
        if hi == 0:
            return 0
        hi = min(self.cap, -(-hi // 8) * 8)  # prefix rounded up to 8 (fewer graph shapes)
        import ctypes
        S = ctypes.byref(self.group.S)
        with torch.cuda.stream(self.stream):
            key = (self.parity, hi)
            if self.use_graphs and self.ticks >= 2:
                g = self.graphs.get(key)
                if g is None:
                    g = torch.cuda.CUDAGraph()
                    with torch.cuda.graph(g, stream=self.stream):
                        self._launch(hi, S, ctypes.c_void_p(torch.cuda.current_stream().cuda_stream))
                    self.graphs[key] = g
                g.replay()
            else:
                self._launch(hi, S, ctypes.c_void_p(self.stream.cuda_stream))
        if not self.rk4:  # the leapfrog's new level becomes level 3; RK4 updates it in place
            self.group.swap_levels()
            self.parity ^= 1
        self.ticks += 1
        for i, s in enumerate(self.slots):
            if s is not None:
                self.done_steps[i] += 1
        return hi

    def harvest(self, now_fn):
        """Collect finished requests (waits for this pool's stream)."""
        fin = [i for i, s in enumerate(self.slots) if s is not None and self.done_steps[i] >= s.nsteps]
        if not fin:
            return []
        self.stream.synchronize()
        out = []
        g = self.group
        for i in fin:
            req = self.slots[i]
            req.trace = self.trace[i, :req.nsteps].cpu().numpy()
            req.status = "done"
            if not np.isfinite(req.trace).all():
                req.status, req.error = "failed", "non-finite observables in the trace"
            # final level-3 populations and polarization matrix of this slot
            # (after the swap, level 3 is in the B buffers)
            C = g.s["CB"][i].cpu().numpy()
            D = g.s["DB"][i].cpu().numpy()
            p = g.s["pB"][i].cpu().numpy()
            req.final = dict(ne=np.real(np.diag(C)).copy(), nh=np.real(np.diag(D)).copy(), p=p)
            req.t_done = now_fn()
            self.slots[i] = None
            out.append(req)
        return out


def slot_bytes(nk, nx, prec, maxsteps, integrator="leapfrog"):
    """Per-slot device bytes of a pool (state, scratch, split planes, trace)."""
    N = nk
    b = (15 if integrator == "rk4" else 12) * N * N * 16  # pA..Hho (+ the RK4 accumulators)
    b += 2 * N * 16 + (8 * N + 4 * (2 * N - 1) + 5 * N) * 8
    b += 2 * N * (2 * N - 1) * 8 + 11 * N * 8 + 3 * (2 * N - 1) * 16
    b += 3 * nx * 16 + 4 * nx * 16 + (9 * (2 * N + 2) + 8 * (2 * N - 1) + 8) * 8
    if prec == "bf16x3":
        b += 14 * 4 * (-(-N // 128) * 128) * (-(-N // 32) * 32) * 2
    b += 8 * 2 * max((2 * N + 2) * 4, (2 * N - 1) * 3, nx * 5) * 8  # split-K partial sums (bnx <= nx)
    b += maxsteps * 4 * 8 + 8 * 8
    return b


def pool_bytes(nk):
    """Shared per-pool tables: the many-body lists, and the stored Lorentzian
    table of the off-diagonal dephasing where N > 512 needs it."""
    N = nk
    lor = 2 * N * N * 2 * N * 8 if N > 512 else 0
    return lor + 64 * N * N * N // 8


class Server:
    """Admission by HBM footprint, continuous batching over Nk pools."""

    def __init__(self, rr, dt=10e-18, prec="bf16x3", hbm_budget=64 << 30, max_slots=64,
                 maxsteps=20000, graphs=True, prm_kw=None, integrator="leapfrog", merged=False,
                 nan_check=64, pool_factory=None):
        self.rr = np.asarray(rr, dtype=float)
        # the pool class (SlotPool on the device; tests of the host logic
        # substitute one with the same methods)
        self.pool_factory = pool_factory or SlotPool
        self.nan_check = int(nan_check)
        self.rejected, self.cancelled, self.failed = [], [], []
        self._cancel = set()
        self.dt, self.prec = float(dt), prec
        self.integrator = integrator
        # merged: every tick steps all pools together on one stream, their RHS
        # as one persistent launch over the per-XCD work queues
        # (psa_step_multi); results equal the per-pool ticks exactly
        self.merged = bool(merged)
        if self.merged and integrator != "leapfrog":
            raise ValueError("merged ticks step leapfrog pools")
        self._stream = None
        self._multi = {}
        self._mgraphs = {}
        self._mseen = {}
        self.budget, self.max_slots = int(hbm_budget), int(max_slots)
        self.maxsteps, self.graphs = int(maxsteps), graphs
        self.prm_kw = prm_kw
        self.pools = {}
        self.used = 0

    def pool(self, nk):
        """The slot pool of `nk`, built on first use; None when the HBM
        budget left cannot hold its tables and one slot."""
        p = self.pools.get(nk)
        if p is None:
            per = slot_bytes(nk, len(self.rr), self.prec, self.maxsteps, self.integrator)
            room = self.budget - self.used - pool_bytes(nk)
            cap = min(self.max_slots, room // per)
            if cap < 1:
                return None
            p = self.pool_factory(nk, cap, self.rr, self.dt, self.prec, self.maxsteps, self.graphs, self.prm_kw,
                                  integrator=self.integrator)
            if self.merged:
                if self._stream is None:
                    self._stream = torch.cuda.Stream(device=p.group.dev)
                p.stream = self._stream
            self.used += p.footprint()
            self.pools[nk] = p
        return p

    def cancel(self, rid):
        """Cancel request `rid` (waiting or in flight) at the next loop
        iteration of run(); it ends in ``cancelled``."""
        self._cancel.add(rid)

    def _reject(self, r, why):
        r.status, r.error = "rejected", why
        self.rejected.append(r)

    def _admission_check(self, pending):
        """Reject, before the clock starts, every request that could never
        run: a step count outside [1, maxsteps], an Nk whose pool cannot be
        built (unsupported k-grid) or does not fit the HBM budget."""
        ok = []
        bad_nk = {}
        for nk in sorted({r.nk for r in pending}):
            try:
                if self.pool(nk) is None:
                    bad_nk[nk] = (f"no HBM room for a slot of Nk={nk} "
                                  f"(budget {self.budget} B, {self.used} B in use)")
            except (ValueError, capi.PsaError) as exc:
                bad_nk[nk] = f"Nk={nk} not supported: {exc}"
        for r in pending:
            if not 1 <= r.nsteps <= self.maxsteps:
                self._reject(r, f"nsteps={r.nsteps} outside [1, {self.maxsteps}]")
            elif r.nk in bad_nk:
                self._reject(r, bad_nk[r.nk])
            else:
                ok.append(r)
        return ok

    def _tick_merged(self, pools):
        """One step of every occupied prefix of every pool: the drives, one
        psa_step_multi over all pools, the records, on one stream; replayed
        from a captured graph once the same (Nk, parity, prefix) combination
        has run twice."""
        his = [min(p.cap, -(-p.prefix() // 8) * 8) for p in pools]
        key = tuple((p.nk, p.parity, hi) for p, hi in zip(pools, his))
        m = self._multi.get(key)
        if m is None:
            m = MultiRhs([p.group for p in pools], mode=1, dt=self.dt, ranges=[(0, hi) for hi in his])
            self._multi[key] = m
        lib = pools[0].lib

        def launch(h):
            for p, hi in zip(pools, his):
                capi.check(lib.psa_drive(ctypes.byref(p.group.S), 0, hi, p.drive.data_ptr(), p.dt, h))
            capi.check(lib.psa_step_multi(ctypes.cast(m.host, ctypes.c_void_p), m.packed.data_ptr(),
                                          m.queue.data_ptr(), self.dt, h))
            for p, hi in zip(pools, his):
                capi.check(lib.psa_record(ctypes.byref(p.group.S), 0, hi, p.xmid, p.trace.data_ptr(),
                                          p.count.data_ptr(), p.maxsteps, h))

        seen = self._mseen.get(key, 0)
        self._mseen[key] = seen + 1
        with torch.cuda.stream(self._stream):
            if self.graphs and seen >= 2:
                gph = self._mgraphs.get(key)
                if gph is None:
                    gph = torch.cuda.CUDAGraph()
                    with torch.cuda.graph(gph, stream=self._stream):
                        launch(ctypes.c_void_p(torch.cuda.current_stream().cuda_stream))
                    self._mgraphs[key] = gph
                gph.replay()
            else:
                launch(ctypes.c_void_p(self._stream.cuda_stream))
        for p in pools:
            p.group.swap_levels()
            p.parity ^= 1
            p.ticks += 1
            for i, r in enumerate(p.slots):
                if r is not None:
                    p.done_steps[i] += 1

    def _cancel_and_fail(self, waiting, now):
        """Apply cancellations and evict requests with non-finite observables."""
        if self._cancel:
            keep = []
            for r in waiting:
                if r.rid in self._cancel:
                    r.status, r.t_done = "cancelled", now
                    self.cancelled.append(r)
                else:
                    keep.append(r)
            waiting[:] = keep
            for p in self.pools.values():
                for i, r in enumerate(p.slots):
                    if r is not None and r.rid in self._cancel:
                        p.release(i)
                        r.status, r.t_done = "cancelled", now
                        self.cancelled.append(r)
        for p in self.pools.values():
            if self.nan_check > 0 and p.active() and p.ticks % self.nan_check == 0 and p.ticks > 0:
                for i in p.nonfinite():
                    r = p.release(i)
                    r.status, r.error, r.t_done = "failed", f"non-finite observables after tick {p.ticks}", now
                    self.failed.append(r)

    def run(self, requests, progress=None, time_limit=None, on_tick=None):
        """Serve `requests` (sorted by arrival, seconds from start).  Returns
        the finished requests; latencies are t_done - arrival.  Requests that
        end otherwise are in ``rejected``, ``cancelled`` and ``failed``.
        ``on_tick(server, now)`` is called once per loop iteration."""
        pending = self._admission_check(sorted(requests, key=lambda r: r.arrival))
        if torch.cuda.is_available():
            torch.cuda.synchronize()
        t0 = time.perf_counter()
        now = lambda: time.perf_counter() - t0  # noqa: E731
        waiting, done = [], []
        qi = 0
        steps = 0
        while qi < len(pending) or waiting or any(p.active() for p in self.pools.values()):
            t = now()
            while qi < len(pending) and pending[qi].arrival <= t:
                pending[qi].status = "waiting"
                waiting.append(pending[qi])
                qi += 1
            if on_tick is not None:
                on_tick(self, t)
            self._cancel_and_fail(waiting, t)
            # admit in arrival order where a slot of the right Nk is free
            rest = []
            for r in waiting:
                p = self.pools.get(r.nk)
                if p is not None and p.free() > 0:
                    p.admit(r, t)
                else:
                    rest.append(r)
            waiting = rest
            busy = False
            act = [p for p in self.pools.values() if p.active()]
            if act and self.merged:
                steps += sum(p.active() for p in act)
                self._tick_merged(act)
                busy = True
            else:
                for p in act:
                    steps += p.active()
                    p.tick()
                    busy = True
            for p in self.pools.values():
                for r in p.harvest(now):
                    (done if r.status == "done" else self.failed).append(r)
            if not busy and qi < len(pending):
                # idle until the next arrival
                time.sleep(max(0.0, min(pending[qi].arrival - now(), 0.01)))
            if progress and len(done) and len(done) % progress == 0:
                print(f"[serve] done {len(done)}/{len(pending)} t={now():.1f}s", flush=True)
            if time_limit is not None and now() > time_limit:
                break
        if torch.cuda.is_available():
            torch.cuda.synchronize()
        self.wall = now()
        self.steps = steps
        return done
