"""Device engine: a group of quantum wires sharing one k-grid, stepped on
one GPU through the C ABI.

A ``WireGroup`` owns every device buffer as a torch tensor (PyTorch is the
allocator and the stream provider only) and fills the ``psa_group`` struct
the kernels read.  State levels follow the reference names: level 2
(``YY2/CC2/DD2``, the averaged level) and level 3 (``YY3/CC3/DD3``, the
newest level) of ``SBESolver`` (``SBEs.py:4770-4778``); the reference
layout [N, N, Nw] is converted to the device's wire-major [Nw, N, N].
"""

import ctypes

import numpy as np
import torch
from scipy.constants import e as e0
from scipy.constants import hbar

from . import capi, tables

F64 = torch.float64
C128 = torch.complex128


def _prep_stride(N):
    return 8 * N + 4 * (2 * N - 1) + 5 * N


def _src_stride(N, Nr):
    return 9 * Nr + 8 * (2 * N - 1) + 8


class WireGroup:
    """Nw wires of one k-grid size on one device.

    prm    -- tables.QWParams (material, numerics, flags)
    nw     -- number of wires
    rr     -- propagation grid the fields live on (QWCalculator's rr)
    wire_ids -- 1-based reference wire index per wire (sets yw(w), (-1)^w)
    prec   -- "fp64" (reference-equivalent), "fp32" or "bf16x3"
    integrator -- "leapfrog" (the reference's leapfrog + average) or "rk4"
               (classic RK4, psa_step_rk4: the state is level 3 and stays
               there; level 2 is stage scratch)
    """

    def __init__(self, prm, nw, rr, wire_ids=None, prec="fp64", device="cuda", host_mb=False,
                 lor_table=False, integrator="leapfrog", carry=True, lazy=True):
        self.lib = capi.load()
        # carry: let the library reuse, from one step to the next, the
        # operand planes and diagonal sums it left behind (psa_group.wtag)
        self.carry = bool(carry)
        # lazy: the leapfrog average X2 = (X1 + X3) / 2 is formed where the
        # next step reads it (psa_group.lazy) instead of in place; the
        # state reads (get_state) write it out first
        self.lazy = bool(lazy)
        if not torch.cuda.is_available():
            raise capi.PsaError("no GPU visible: the device path has no CPU fallback")
        self.dev = torch.device(device)
        self.prm = prm
        self.nw = int(nw)
        self.rr = np.asarray(rr, dtype=float)
        self.prec = prec
        if integrator not in ("leapfrog", "rk4"):
            raise ValueError(f"unknown integrator {integrator!r}")
        self.integrator = integrator
        if prec not in capi.PREC:
            raise ValueError(f"unknown precision {prec!r}")
        if prm.flags.get("FreePot"):
            # _preparation never fills Wee/Whh (SBEs.py:5173-5226): FreePot
            # adds zero in the reference; nothing to do here either.
            pass
        self.grid, tb = tables.build_tables(prm, self.rr, self._vint, host_mb=host_mb)
        self.tab = tb
        g = self.grid
        N, Nr, Nx = g["N"], g["Nr"], len(self.rr)
        self.N, self.Nr, self.Nx = N, Nr, Nx
        d = self.dev
        T = lambda a, dt=F64: torch.as_tensor(np.ascontiguousarray(a), dtype=dt).to(d)  # noqa: E731
        self.t = {
            "kr": T(g["kr"]), "Ee": T(g["Ee"]), "Eh": T(g["Eh"]),
            "vtab": T(tb["vtab"]), "chi": T(tb["chi"]), "qeh": T(tb["qeh"]),
            "EP": T(tb["EP"]), "HP": T(tb["HP"]),
            "EPT": T(tb["EP"].T), "HPT": T(tb["HP"].T), "xe": T(tb["xe"]), "xh": T(tb["xh"]),
            "didx": T(tb["didx"], torch.int32), "iq": T(tb["iq"], torch.int32),
            "homega": T(tb["homega"]), "hsq": T(tb["hsq"]), "window": T(tb["window"]),
            "rq": T(g["r"]), "ytr": T(tb["ytr"]),
            "gmat": T(tb["gmat"], C128),
            "m2": T(tb["m2"].T if tb["bnx"] else np.zeros((Nr, 1))),
            "phi": T(tb["phi"], C128), "twid": T(tb["twid"], C128),
        }
        # The off-diagonal dephasing kernel makes its Lorentzian operand in
        # registers (N <= 512); the stored table [2][N*N][2N] (built once by
        # psa_lorentz_table, 0.54 GB at N = 256) serves larger N, or on request.
        od = bool(prm.flags.get("OffDiagDph"))
        self.lor_table = od and (bool(lor_table) or N > 512)
        if self.lor_table:
            self.t["lor"] = torch.zeros(2 * N * N * 2 * N, dtype=F64, device=d)
        if tb["mb"] is None:
            self.mb = self._device_mb_lists()
            self.mb_nnz = list(self.mb_sizes)
        else:
            self.mb = []
            for off, idx, val in tb["mb"]:
                self.mb.append((T(off, torch.int32),
                                T(idx.astype(np.int64).astype(np.int32) if idx.size else np.zeros(1, np.int32), torch.int32),
                                T(val if val.size else np.zeros(1))))
            self.mb_nnz = [int(v.size) for _, _, v in tb["mb"]]
        nwz = self.nw
        z = lambda *s, dt=C128: torch.zeros(s, dtype=dt, device=d)  # noqa: E731
        self.s = {}
        for k in ("pA", "pB", "pX", "CA", "CB", "CX", "DA", "DB", "DX", "Heh", "Heo", "Hho"):
            self.s[k] = z(nwz, N, N)
        self.s["hdiag"] = z(nwz, 2, N)
        self.s["prep"] = z(nwz, _prep_stride(N), dt=F64)
        self.s["dedh"] = z(nwz, 2, N, 2 * N - 1, dt=F64)
        # [Nw][4][N] rates; wire fastest: occ [2][N][Nw], V^2 [3][N][Nw], |n| [2][N][Nw]
        self.s["wmb"] = z(11 * nwz * N, dt=F64)
        self.s["exq"] = z(nwz, 3, 2 * N - 1)
        self.s["efield"] = z(3, nwz, Nx)
        self.s["src"] = z(nwz, _src_stride(N, Nr), dt=F64)
        self.s["out"] = z(4, nwz, Nx)
        self.s["obs"] = z(nwz, 8, dt=F64)
        self.s["tphase"] = z(nwz, N)
        self.s["part"] = z(nwz, 256, dt=F64)
        if integrator == "rk4":  # stage accumulators R = k1 + 2 k2 + 2 k3
            for k in ("pR", "CR", "DR"):
                self.s[k] = z(nwz, N, N)
        # split-bf16 operand planes (bf16x3 only): [Nw][14 slots][4 planes] on
        # a zero-padded (128 ceil(N/128)) x (32 ceil(N/32)) grid, in 16 KB
        # chunks of 64 rows x 32 k
        nspl = nwz * 14 * 4 * (-(-N // 128) * 128) * (-(-N // 32) * 32) if prec == "bf16x3" else 1
        self.s["spl"] = torch.zeros(nspl, dtype=torch.int16, device=d)
        # per-wire reuse tags (psa_group.wtag): which level the p, C, D planes
        # and the carried diagonal sums belong to; zeroed on every host write
        # of the state
        self.s["wtag"] = torch.zeros(2 * nwz, dtype=torch.int64, device=d)
        # per-wire flags of the deferred leapfrog average (psa_group.lazy)
        self.s["lazy"] = torch.zeros(nwz, dtype=torch.int32, device=d)
        # split-K partial sums of the source / propagation-grid products
        self._kscr_len = 8 * 2 * max(Nr * 4 * nwz, (2 * N - 1) * 3 * nwz, max(tb["bnx"], 1) * 5 * nwz)
        self.s["kscr"] = torch.empty(self._kscr_len, dtype=F64, device=d)
        if wire_ids is None:
            wire_ids = np.arange(1, nwz + 1)
        self.wire_ids = np.asarray(wire_ids, dtype=np.int32)
        if self.wire_ids.shape != (nwz,):
            raise ValueError("wire_ids must have one entry per wire")
        self.s["wire_id"] = T(self.wire_ids, torch.int32)
        self._wid_host = self.wire_ids.copy()
        self.tune = [0, 0, 0, 0]  # launch choices (psa_group.tune), 0 = defaults
        self.init_state()
        self._sync_struct()
        if self.lor_table:
            capi.check(self.lib.psa_lorentz_table(ctypes.byref(self.S), self._stream()))

    # ------------------------------------------------------------------
    def _device_mb_lists(self):
        """The eight many-body lists built on the device (psa_mb_list):
        count per target, prefix sum on the host, fill."""
        g, prm, d = self.grid, self.prm, self.dev
        N = len(g["Ee"])
        Ee = torch.as_tensor(g["Ee"], dtype=F64, device=d)
        Eh = torch.as_tensor(g["Eh"], dtype=F64, device=d)
        st = self._stream()
        lz = 1 if prm.flags.get("LorentzDelta") else 0
        out, sizes = [], []
        for L in range(8):
            cnt = torch.zeros(N, dtype=torch.int32, device=d)
            capi.check(self.lib.psa_mb_list(L, N, Ee.data_ptr(), Eh.data_ptr(), prm.gam_e, prm.gam_h,
                                            hbar, lz, None, cnt.data_ptr(), None, None, st))
            c = cnt.cpu().numpy().astype(np.int64)
            off = np.zeros(N + 1, dtype=np.int64)
            off[1:] = np.cumsum(c)
            nnz = int(off[-1])
            if nnz >= 2**31:
                raise capi.PsaError("many-body list too long for 32-bit offsets")
            offt = torch.as_tensor(off.astype(np.int32), device=d)
            idx = torch.zeros(max(nnz, 1), dtype=torch.int32, device=d)
            val = torch.zeros(max(nnz, 1), dtype=F64, device=d)
            if nnz:
                capi.check(self.lib.psa_mb_list(L, N, Ee.data_ptr(), Eh.data_ptr(), prm.gam_e, prm.gam_h,
                                                hbar, lz, offt.data_ptr(), None, idx.data_ptr(),
                                                val.data_ptr(), st))
            out.append((offt, idx, val))
            sizes.append(nnz)
        torch.cuda.synchronize()
        self.mb_sizes = sizes
        return out

    def _vint(self, Qr, r, ae, ah, Delta0, pref):
        d = self.dev
        Q = torch.as_tensor(Qr, dtype=F64, device=d)
        y = torch.as_tensor(r, dtype=F64, device=d)
        out = torch.zeros(3, len(Qr), dtype=F64, device=d)
        capi.check(self.lib.psa_vint_table(Q.data_ptr(), len(Qr), y.data_ptr(), len(r),
                                           ae, ah, Delta0, pref, out.data_ptr(),
                                           self._stream()))
        torch.cuda.synchronize()
        return out.cpu().numpy()

    def _stream(self):
        return ctypes.c_void_p(torch.cuda.current_stream(self.dev).cuda_stream)

    def _sync_struct(self):
        g, tb, prm = self.grid, self.tab, self.prm
        S = capi.PsaGroup()
        S.N, S.Nr, S.Nw, S.Nx = self.N, self.Nr, self.nw, self.Nx
        S.Nhw = len(tb["homega"])
        S.flags = prm.flag_bits()
        S.prec = capi.PREC[self.prec]
        S.gx0, S.gnx, S.bx0, S.bnx = tb["gx0"], tb["gnx"], tb["bx0"], tb["bnx"]
        for i, n in enumerate(self.mb_nnz):
            S.nnz_mb[i] = int(n)
        S.hbar, S.e0, S.gap, S.L = hbar, e0, prm.gap, prm.L
        S.ehint, S.dcv, S.area, S.dkr = g["ehint"], g["dcv"], g["area"], g["dkr"]
        S.dr_r = g["r"][1] - g["r"][0]
        S.drr = self.rr[1] - self.rr[0] if len(self.rr) > 1 else 1.0
        S.gam_e, S.gam_h, S.gam_eh, S.Edc = prm.gam_e, prm.gam_h, g["gam_eh"], prm.Edc
        S.Vscale, S.n1d_max, S.dhw = tb["Vscale"], tb["n1d_max"], tb["dhw"]
        for i, v in enumerate(self.tune):
            S.tune[i] = int(v)
        for k, v in self.t.items():
            setattr(S, k, v.data_ptr())
        for i, (off, idx, val) in enumerate(self.mb):
            S.mb_off[i] = off.data_ptr()
            S.mb_idx[i] = idx.data_ptr()
            S.mb_val[i] = val.data_ptr()
        for k, v in self.s.items():
            setattr(S, k, v.data_ptr())
        S.kscr_len = self._kscr_len
        S.wid_host = self._wid_host.ctypes.data
        if not self.carry:
            S.wtag = None
        if not self.lazy:
            S.lazy = None
        self.S = S

    # ------------------------------------------------------------------
    # state
    # ------------------------------------------------------------------
    def init_state(self):
        """Initial populations FermiDistr(Ee + gap/2) for both carriers
        (SBEs.py:4829-4840), coherences zero."""
        N = self.N
        f = tables.fermi(self.grid["Ee"] + self.prm.gap / 2.0)
        diag = torch.as_tensor(np.diag(f).astype(np.complex128), device=self.dev)
        for k in ("pA", "pB", "pX"):
            self.s[k].zero_()
        for k in ("CA", "CB", "DA", "DB"):
            self.s[k].copy_(diag.expand(self.nw, N, N))
        self.s["CX"].zero_()
        self.s["DX"].zero_()
        self.s["lazy"].zero_()
        self.invalidate()

    def invalidate(self, w=None):
        """Forget what the library carries over between steps (the p, C, D
        operand planes and the diagonal sums) for wire w, or every wire:
        needed after any write of p, C or D outside the step entry points."""
        if w is None:
            self.s["wtag"].zero_()
        else:
            self.s["wtag"][2 * w:2 * w + 2].zero_()

    def reset_wire(self, w):
        """Initial state (init_state) for wire w only, levels 2 and 3."""
        N = self.N
        f = tables.fermi(self.grid["Ee"] + self.prm.gap / 2.0)
        diag = torch.as_tensor(np.diag(f).astype(np.complex128), device=self.dev)
        for k in ("pA", "pB"):
            self.s[k][w].zero_()
        for k in ("CA", "CB", "DA", "DB"):
            self.s[k][w].copy_(diag)
        self.s["lazy"][w] = 0
        self.invalidate(w)

    def swap_levels(self):
        """Level 3 <- the new level for every wire of the group (pointer
        swap of the B and X buffers); wires that were not stepped exchange
        stale buffers, which callers reset before use."""
        for a, b in (("pB", "pX"), ("CB", "CX"), ("DB", "DX")):
            self.s[a], self.s[b] = self.s[b], self.s[a]
        self._sync_struct()

    def autotune(self, reps=2, rounds=3):
        """Pick the launch choices of this group by timing them on its own
        buffers: tune[2] GEMM block order (XCD-grouped or linear; the
        split-BF16 RHS has a static schedule of its own), tune[1] / tune[3]
        transport lines per block for the row / column pass (4, 2, 8, 16),
        timed over full steps with the level swap (the column pass carries
        the leapfrog average there, and in the split-BF16 mode writes the
        next step's operand planes).  The candidates of a knob are timed in
        turn, `rounds` times over (reps calls each), and each keeps its
        fastest round, so a clock or thermal drift during the timing does
        not pick a choice.  The state is unchanged: the levels the steps
        overwrite are saved and restored.  Returns {knob: (choice, ms per
        candidate)}."""
        od = bool(self.prm.flags.get("OffDiagDph"))
        self.prop2qw()
        self.prep()
        torch.cuda.synchronize()

        def contractions():
            self.calc_h()
            if od:
                self.offdiag()
            self.rhs(1, 1e-18)

        def full_step():  # a step as the time loop runs it (with the level swap,
            self.step(1e-18)  # so what one step leaves for the next counts)

        knobs = [(2, [0, 1], contractions)]
        saved = None
        if self.prm.flags.get("DCTrans") and self.prm.flags.get("LF"):
            saved = {k: self.s[k].clone() for k in ("pA", "pB", "pX", "CA", "CB", "CX", "DA", "DB", "DX",
                                                      "lazy")}
            knobs.append((1, [4, 2, 8, 16], full_step))
            knobs.append((3, [4, 2, 8, 16], full_step))
        out = {}
        for idx, cands, fn in knobs:
            times = [float("inf")] * len(cands)
            for r in range(rounds):
                for i, c in enumerate(cands):
                    self.tune[idx] = c
                    self._sync_struct()
                    # an untimed call first: the first launch of a choice sets
                    # its attributes, and a step leaves state for the next
                    # (the column pass's planes), so the timed calls start
                    # from what this choice itself leaves
                    fn()
                    torch.cuda.synchronize()
                    a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
                    a.record()
                    for _ in range(reps):
                        fn()
                    b.record()
                    torch.cuda.synchronize()
                    times[i] = min(times[i], a.elapsed_time(b) / reps)
            best = int(np.argmin(times))
            # the XCD-grouped block order (candidate 0 of knob 2) reads about
            # half the HBM bytes of the linear one; it is kept unless the
            # linear order is clearly faster (timing noise aside)
            if idx == 2 and times[0] <= 1.03 * times[best]:
                best = 0
            self.tune[idx] = cands[best]
            self._sync_struct()
            out[idx] = (self.tune[idx], times)
        if saved is not None:  # the levels by role (the buffers may have swapped)
            for k, v in saved.items():
                self.s[k].copy_(v)
            self.invalidate()
            torch.cuda.synchronize()
        return out

    def footprint_bytes(self):
        """Device bytes held by this group (state, scratch and tables)."""
        tot = sum(t.numel() * t.element_size() for t in self.s.values())
        tot += sum(t.numel() * t.element_size() for t in self.t.values())
        tot += sum(sum(x.numel() * x.element_size() for x in m) for m in self.mb)
        return int(tot)

    def set_state(self, C2, D2, p2, C3, D3, p3):
        """Load levels 2 and 3 from arrays in the reference layout [N, N, Nw]."""
        for k, a in (("CA", C2), ("DA", D2), ("pA", p2), ("CB", C3), ("DB", D3), ("pB", p3)):
            a = np.asarray(a, dtype=np.complex128)
            if a.ndim == 2:
                a = a[:, :, None]
            self.s[k].copy_(torch.as_tensor(np.ascontiguousarray(np.moveaxis(a, 2, 0)), device=self.dev))
        self.s["lazy"].zero_()
        self.invalidate()

    def hamiltonians(self):
        """(Heh, Heo, Hho) as complex128 arrays [Nw][N][N] in the device
        order (transposed, [k2][k1]); the split-BF16 mode stores them as
        complex64 packed from the start of each buffer."""
        torch.cuda.synchronize()
        n = self.nw * self.N * self.N
        out = []
        for k in ("Heh", "Heo", "Hho"):
            t = self.s[k]
            if self.prec == "bf16x3":
                t = t.reshape(-1).view(torch.complex64)[:n].reshape(self.nw, self.N, self.N)
            out.append(t.cpu().numpy().astype(np.complex128))
        return tuple(out)

    def average_levels(self, w0=0, nw=None):
        """Write out the deferred leapfrog average of wires [w0, w0+nw)
        (psa_average_levels): level A then holds X2 as the reference keeps
        it.  Host code reads or writes level A only after this."""
        w0, nw = self._range(w0, nw)
        capi.check(self.lib.psa_average_levels(ctypes.byref(self.S), w0, nw, self._stream()))

    def get_state(self):
        """(C2, D2, p2, C3, D3, p3) in the reference layout [N, N, Nw]."""
        self.average_levels()
        torch.cuda.synchronize()
        out = []
        for k in ("CA", "DA", "pA", "CB", "DB", "pB"):
            out.append(np.moveaxis(self.s[k].cpu().numpy(), 0, 2).copy())
        return tuple(out)

    # ------------------------------------------------------------------
    # stages (each one C-ABI call on the current stream)
    # ------------------------------------------------------------------
    def _range(self, w0, nw):
        nw = self.nw - w0 if nw is None else nw
        return int(w0), int(nw)

    def set_fields(self, Exx, Eyy, Ezz, w0=0, nw=None):
        """Fields on rr for wires [w0, w0+nw): arrays of shape [Nx] (same
        for every wire) or [nw, Nx]; numpy or device tensors."""
        w0, nw = self._range(w0, nw)
        for f, E in enumerate((Exx, Eyy, Ezz)):
            t = torch.as_tensor(E, dtype=C128).to(self.dev)
            if t.ndim == 1:
                t = t.expand(nw, self.Nx)
            self.s["efield"][f, w0:w0 + nw].copy_(t)

    def prop2qw(self, w0=0, nw=None):
        w0, nw = self._range(w0, nw)
        capi.check(self.lib.psa_prop2qw(ctypes.byref(self.S), w0, nw, self._stream()))

    def prep(self, w0=0, nw=None):
        w0, nw = self._range(w0, nw)
        capi.check(self.lib.psa_prep(ctypes.byref(self.S), w0, nw, self._stream()))

    def calc_h(self, w0=0, nw=None):
        w0, nw = self._range(w0, nw)
        capi.check(self.lib.psa_calc_h(ctypes.byref(self.S), w0, nw, self._stream()))

    def offdiag(self, w0=0, nw=None):
        w0, nw = self._range(w0, nw)
        capi.check(self.lib.psa_offdiag_tables(ctypes.byref(self.S), w0, nw, self._stream()))

    def rhs(self, mode=0, dt=0.0, w0=0, nw=None):
        w0, nw = self._range(w0, nw)
        capi.check(self.lib.psa_rhs(ctypes.byref(self.S), w0, nw, mode, dt, self._stream()))

    def relaxation(self, dt, w0=0, nw=None):
        w0, nw = self._range(w0, nw)
        capi.check(self.lib.psa_relaxation(ctypes.byref(self.S), w0, nw, dt, self._stream()))

    def transport(self, dt, w0=0, nw=None):
        w0, nw = self._range(w0, nw)
        capi.check(self.lib.psa_transport(ctypes.byref(self.S), w0, nw, dt, self._stream()))

    def finish(self, w0=0, nw=None):
        w0, nw = self._range(w0, nw)
        capi.check(self.lib.psa_finish(ctypes.byref(self.S), w0, nw, self._stream()))

    def sources(self, w0=0, nw=None):
        w0, nw = self._range(w0, nw)
        capi.check(self.lib.psa_sources(ctypes.byref(self.S), w0, nw, self._stream()))

    def qw2prop(self, w0=0, nw=None):
        w0, nw = self._range(w0, nw)
        capi.check(self.lib.psa_qw2prop(ctypes.byref(self.S), w0, nw, self._stream()))

    def rhs_eval(self, w0=0, nw=None):
        """One RHS evaluation on level 3 (the reference's p2/C2/D2 at
        checkout): fills pX/CX/DX with dp/dt, dC/dt, dD/dt."""
        if self.prm.flags.get("Optics"):
            self.prop2qw(w0, nw)
        self.prep(w0, nw)
        self.calc_h(w0, nw)
        if self.prm.flags.get("OffDiagDph"):
            self.offdiag(w0, nw)
        self.rhs(0, 0.0, w0, nw)

    def advance_levels(self, w0=0, nw=None):
        """After a step: level 3 <- the new level (pointer swap for the whole
        group, device copy for a sub-range)."""
        w0, nw = self._range(w0, nw)
        if w0 == 0 and nw == self.nw:
            self.swap_levels()
        else:
            for a, b in (("pB", "pX"), ("CB", "CX"), ("DB", "DX")):
                self.s[a][w0:w0 + nw].copy_(self.s[b][w0:w0 + nw])

    def step(self, dt, w0=0, nw=None):
        """One full QWCalculator step (without the host bookkeeping) for
        wires [w0, w0+nw) with the fields already in ``efield``: the
        reference's leapfrog + average, or RK4 for an "rk4" group."""
        w0, nw = self._range(w0, nw)
        if self.integrator == "rk4":
            capi.check(self.lib.psa_step_rk4(ctypes.byref(self.S), w0, nw, dt, self._stream()))
            return
        capi.check(self.lib.psa_step(ctypes.byref(self.S), w0, nw, dt, self._stream()))
        self.advance_levels(w0, nw)

    def outputs(self, w0=0, nw=None):
        """(Pxx, Pyy, Pzz, Rho) as numpy [nw, Nx]."""
        w0, nw = self._range(w0, nw)
        o = self.s["out"][:, w0:w0 + nw].cpu().numpy()
        return o[0], o[1], o[2], o[3]

    def observables(self):
        """Per-wire I0 (CalcI0), sum ne2, sum nh2 as numpy."""
        ob = self.s["obs"].cpu().numpy()
        return dict(I0=ob[:, 0], sum_ne2=ob[:, 1], sum_nh2=ob[:, 2])


class MultiRhs:
    """The RHS (``psa_rhs`` with ``mode`` and ``dt``) of several wire groups
    of any N and one precision mode, as one persistent launch over work
    queues, one per XCD (``psa_rhs_multi``).  The groups' structs are packed when this is
    built; build one per level parity (``swap_levels`` exchanges the level
    pointers) and again when a group's wire range changes.  Results equal
    the per-group ``WireGroup.rhs`` exactly."""

    def __init__(self, groups, mode=0, dt=0.0, ranges=None):
        self.lib = capi.load()
        self.groups = list(groups)
        n = len(self.groups)
        if not 1 <= n <= capi.PSA_RHS_MULTI_MAX:
            raise ValueError(f"1 to {capi.PSA_RHS_MULTI_MAX} groups")
        ranges = ranges or [(0, g.nw) for g in self.groups]
        self.dev = self.groups[0].dev
        ptrs = (ctypes.c_void_p * n)(*[ctypes.addressof(g.S) for g in self.groups])
        w0 = (ctypes.c_int32 * n)(*[int(r[0]) for r in ranges])
        nw = (ctypes.c_int32 * n)(*[int(r[1]) for r in ranges])
        nb = int(self.lib.psa_rhs_multi_bytes(ctypes.cast(nw, ctypes.c_void_p), n))
        self.host = ctypes.create_string_buffer(nb)
        capi.check(self.lib.psa_rhs_multi_pack(ctypes.cast(ptrs, ctypes.c_void_p), ctypes.cast(w0, ctypes.c_void_p),
                                               ctypes.cast(nw, ctypes.c_void_p), n, int(mode), float(dt),
                                               ctypes.cast(self.host, ctypes.c_void_p)))
        self.packed = torch.frombuffer(bytearray(self.host.raw), dtype=torch.uint8).to(self.dev)
        self.queue = torch.zeros(capi.PSA_RHS_MULTI_QUEUE, dtype=torch.int32, device=self.dev)

    def __call__(self):
        capi.check(self.lib.psa_rhs_multi(ctypes.cast(self.host, ctypes.c_void_p), self.packed.data_ptr(),
                                          self.queue.data_ptr(),
                                          ctypes.c_void_p(torch.cuda.current_stream(self.dev).cuda_stream)))


class MultiStep:
    """One leapfrog step (``WireGroup.step``) of several wire groups of any
    N and one precision mode, with their RHS as one persistent launch over
    work queues (``psa_step_multi``).  A packed description is kept per
    level parity.  Results equal ``WireGroup.step`` per group exactly."""

    def __init__(self, groups, dt):
        self.groups = list(groups)
        if any(g.integrator != "leapfrog" for g in self.groups):
            raise ValueError("MultiStep steps leapfrog groups")
        self.dt = float(dt)
        self.lib = capi.load()
        self._packs = {}

    def _pack(self):
        key = tuple(g.S.pB for g in self.groups)
        if key not in self._packs:
            self._packs[key] = MultiRhs(self.groups, mode=1, dt=self.dt)
        return self._packs[key]

    def __call__(self):
        m = self._pack()
        capi.check(self.lib.psa_step_multi(ctypes.cast(m.host, ctypes.c_void_p), m.packed.data_ptr(),
                                           m.queue.data_ptr(), self.dt,
                                           ctypes.c_void_p(torch.cuda.current_stream(m.dev).cuda_stream)))
        for g in self.groups:
            g.advance_levels()
