"""Drop-in for ``pulsesuite.PSTD3D.SBEs`` on one MI355X.

Keeps the reference API (``SBEs.py:4619-6053``):

* ``SBESolver(q, rr, r0, Emaxxx, lam, Nw, QW, backend="auto")`` reading
  ``params/qw.params`` and ``params/mb.params`` from the working directory
  (``SBEs.py:4978-5022``);
* ``SBESolver.QWCalculator(Exx, Eyy, Ezz, Vrr, rr, q, dt, w, Pxx, Pyy, Pzz,
  Rho, DoQWP, DoQWDl)`` with the same bookkeeping (``SBEs.py:5837-5847``),
  wire-off gate (``5849-5858``), in-place outputs (strided views allowed),
  ``Rho = RhoH - RhoE`` and the P.E energy bookkeeping (``5925-5979``);
* the module shims ``InitializeSBE``, ``QWCalculator``, ``chiqw``, ``getqc``,
  ``QWArea``, ``ShutOffOptics``, ``WriteSBEsData``, ``ReadSBEsData``.

The state of all wires lives on the GPU (``engine.WireGroup``); every call
runs the HIP path.  ``QWCalculatorBatch`` steps all wires in one call (what
a throughput driver should use).  Output files (``dataQW/info.WW.t.dat``,
``dataQW/EP.WW.t.dat``, ``output/PdotE.*.dat``, backups) follow the
reference formats when ``io=True``.
"""

import os

import numpy as np
import torch
from scipy.constants import c as c0
from scipy.constants import e as e0
from scipy.constants import epsilon_0 as eps0
from scipy.constants import hbar
from scipy.constants import k as kB

from . import tables
from .engine import WireGroup

twopi = 2.0 * np.pi
ii = 1j
eV = e0
# printITReal2 (qwoptics.py:1189-1215) writes the coordinate column only in
# the first file of the process
_prop_first = [True]


def _f32(x):
    return np.float32(x)


def _write_it(V, name):
    """WriteIt / WriteIT1D (usefulsubs.py:2339, 1915-1919)."""
    with open(f"dataQW/{name}.dat", "w", encoding="utf-8") as f:
        f.write("".join(f"{v}\n" for v in V))


def _print_it(Dx, z, n, name, width=6):
    """printIT (usefulsubs.py:2068-2100): z, Re, Im as float32."""
    fn = f"dataQW/{name}{n:0{width}d}.dat"
    os.makedirs(os.path.dirname(fn), exist_ok=True)
    Dx = np.asarray(Dx, dtype=complex)
    with open(fn, "w", encoding="utf-8") as f:
        f.write("".join(f"{_f32(a)} {_f32(b.real)} {_f32(b.imag)}\n" for a, b in zip(z, Dx)))


def _print_itr(Dx, z, n, name, width=6):
    """printITR (usefulsubs.py:2103-2133) and printGam (dephasing.py:417-427,
    width 5): z, value as float32."""
    fn = f"dataQW/{name}{n:0{width}d}.dat"
    os.makedirs(os.path.dirname(fn), exist_ok=True)
    with open(fn, "w", encoding="utf-8") as f:
        f.write("".join(f"{_f32(a)} {_f32(b)}\n" for a, b in zip(z, Dx)))


def _print_it2d(Dx, n, name):
    """printIT2D (usefulsubs.py:2136-2168): |Dx| as float32, column by column."""
    fn = f"dataQW/{name}{n:07d}.dat"
    os.makedirs(os.path.dirname(fn), exist_ok=True)
    with open(fn, "w", encoding="utf-8") as f:
        f.write("".join(f"{_f32(v)}\n" for v in np.abs(np.asarray(Dx)).T.reshape(-1)))


def _print_it_real2(Dx, z, n, name):
    """printITReal2 (qwoptics.py:1189-1215): Re Dx (and z in the first file
    of the process) as float32."""
    fn = f"dataQW/{name}{n:06d}.dat"
    os.makedirs(os.path.dirname(fn), exist_ok=True)
    re = np.real(np.asarray(Dx))
    with open(fn, "w", encoding="utf-8") as f:
        if _prop_first[0]:
            f.write("".join(f"{_f32(a)} {_f32(b)}\n" for a, b in zip(re, z)))
            _prop_first[0] = False
        else:
            f.write("".join(f"{_f32(a)}\n" for a in re))


def yw(w):
    """qwoptics.py:1093-1106."""
    return (-1) ** int(np.floor((w - 1) / 2.0))


def QWChi1(lam, dky, Ee, Eh, area, geh, dcv):
    """Linear susceptibility of the wire (qwoptics.py:1109-1146)."""
    ww = twopi * c0 / lam
    d1 = Ee + Eh - ii * hbar * geh - hbar * ww
    d2 = Ee + Eh + ii * hbar * geh + hbar * ww
    return 4 * dcv**2 / eps0 / area * dky / twopi * np.sum((Ee + Eh) / (d1 * d2))


class SBESolver:
    """Class-based SBE solver with the reference constructor signature.

    Extra keywords (all optional): ``params`` (a ``tables.QWParams``
    instead of reading params/*.params), ``prec`` ("fp64" default,
    "fp32"), ``io`` (write the reference's per-step text files),
    ``keep_level1`` (keep the YY1/CC1/DD1 level for backups)."""

    def __init__(self, q, rr, r0, Emaxxx, lam, Nw, QW, backend="auto", params=None,
                 prec="fp64", io=True, keep_level1=True):
        self._backend = backend
        self.Nw = int(Nw)
        self.io = bool(io)
        self.keep_level1 = bool(keep_level1)
        self.start = False
        if not QW:
            if self.io:
                os.makedirs("dataQW/Wire/info", exist_ok=True)
                open("dataQW/Wire/info/ETHz.t.dat", "w").close()
            return
        if params is None:
            params = tables.QWParams.from_files("params/qw.params", "params/mb.params")
        self.prm = params
        self.rr = np.asarray(rr, dtype=float)
        self.Emax0 = float(Emaxxx)
        self.lam = float(lam)
        self.group = WireGroup(params, self.Nw, self.rr, wire_ids=np.arange(1, self.Nw + 1),
                               prec=prec)
        g = self.group.grid
        for k in ("kr", "r", "Qr", "NQ0", "dkr", "dr", "dcv", "alphae", "alphah",
                  "ehint", "qc", "area", "Ee", "Eh"):
            setattr(self, k, g[k])
        self.Nk, self.Nr = g["N"], g["Nr"]
        self.L, self.gap = params.L, params.gap
        self.me, self.mh = params.me, params.mh
        self.gam_e, self.gam_h = params.gam_e, params.gam_h
        self.gam_eh = g["gam_eh"]
        self.jmax = params.jmax
        for n, v in params.flags.items():
            setattr(self, n, v)
        self.t = 0.0
        self.xxx = 1
        self.jjj = 1
        self.wireoff = True
        self.I0 = np.zeros(self.Nw)
        self.chiw = QWChi1(lam, self.dkr, self.Ee + self.gap, self.Eh, self.area,
                           self.gam_eh, self.dcv)
        self.EPEnergy = 0.0
        self.EPEnergyW = 0.0
        self._P0 = self._P1 = None
        self._file_972 = self._file_973 = None
        self._lvl1 = None
        # real-space field at r[Nr//2] for the EP file: window * spline row
        self._m1 = tables.spline_matrix(self.rr, self.r) if self.io else None
        self._m2 = None
        if self.io:
            os.makedirs("dataQW", exist_ok=True)
            for w in range(1, self.Nw + 1):
                for nm in ("info", "EP", "XQ"):
                    open(f"dataQW/{nm}.{w:02d}.t.dat", "w").close()
            os.makedirs("dataQW/Wire/info", exist_ok=True)
            open("dataQW/Wire/info/ETHz.t.dat", "w").close()
            self._write_init_files()
            self.record_xqw(np.zeros(self.Nk), np.zeros(self.Nk), np.full(self.Nk, self.gam_e),
                            np.full(self.Nk, self.gam_h), 0)
        self.start = True

    def _write_init_files(self):
        """Start-up files of SBESolver.__init__ (SBEs.py:4857-4870), the
        window of QWOptics.__init__ (qwoptics.py:650), the empty Fe/Fh files
        of the DC-field module (dcfield.py:792-793) and the MaxOffDiag files
        of InitializeDephasing (dephasing.py:538-539)."""
        _write_it(self.kr, "kr")
        _write_it(self.Qr, "Qr")
        _write_it(self.r, "R")
        _write_it(self.Ee / eV, "Ee.k")
        _write_it(self.Eh / eV, "Eh.k")
        _write_it((self.Ee + self.Eh + self.gap - hbar * c0 * twopi / self.lam) / eV, "Echw.k")
        _write_it((self.Ee + self.Eh + self.gap) / eV, "Etrn.k")
        with open("dataQW/Etr.dat", "w", encoding="utf-8") as f:
            f.write("".join(f"{k} {(a + b + self.gap) / eV}\n"
                            for k, a, b in zip(self.kr, self.Ee, self.Eh)))
        _print_it(self.group.tab["window"] * (ii) ** 0.0, self.r, 0, "Envl.y.")
        if self.DCTrans:
            open("dataQW/Fe.dat", "w").close()
            open("dataQW/Fh.dat", "w").close()
        open("dataQW/Wire/info/MaxOffDiag.e.dat", "w").close()
        open("dataQW/Wire/info/MaxOffDiag.h.dat", "w").close()

    def _write_max_offdiag(self, i):
        """One line per wire step of OffDiagDephasing2 (dephasing.py:733-743):
        t, max and min of De (resp. Dh), from psa_offdiag_extrema."""
        import ctypes
        from . import capi
        g = self.group
        capi.check(g.lib.psa_offdiag_extrema(ctypes.byref(g.S), i, 1, g._stream()))
        ob = g.s["obs"][i].cpu().numpy()
        t = np.real(self.t)
        for fn, a, b in (("e", ob[4], ob[5]), ("h", ob[6], ob[7])):
            with open(f"dataQW/Wire/info/MaxOffDiag.{fn}.dat", "a", encoding="utf-8") as f:
                f.write(f"{t} {a} {b}\n")

    def _write_snapshots(self, Exx, Eyy, Ezz, out, w):
        """The WriteFields files of one wire step (every jmax steps): the
        relaxation rates (SBEs.py:5282-5307), WriteSBESolns and
        WriteDephasing (5712-5715; qwoptics.py:1263-1291, dephasing.py:430-437),
        WriteQWFields in QW2Prop (qwoptics.py:792-793, 1316-1349) and
        WritePropFields (SBEs.py:5919-5922, qwoptics.py:1352-1386)."""
        i = w - 1
        g = self.group
        N, Nr, xxx = self.Nk, self.Nr, self.xxx
        kr, r, rr = self.kr, self.r, self.rr
        ws = f"{w:02d}"
        pr = g.s["prep"][i].cpu().numpy()
        if self.EHs or self.Phonon:
            mb = g.s["wmb"][i * 4 * N:(i + 1) * 4 * N].cpu().numpy().reshape(4, N)
            _print_itr(mb[0], kr, xxx, f"Wire/Win/Win.e.{ws}.k.")
            _print_itr(mb[2], kr, xxx, f"Wire/Win/Win.h.{ws}.k.")
            _print_itr(mb[1], kr, xxx, f"Wire/Wout/Wout.e.{ws}.k.")
            _print_itr(mb[3], kr, xxx, f"Wire/Wout/Wout.h.{ws}.k.")
        # the new level is level 3 (B) once the step has advanced the levels
        C3 = g.s["CB"][i].cpu().numpy()
        D3 = g.s["DB"][i].cpu().numpy()
        p3 = g.s["pB"][i].cpu().numpy()
        hd = g.s["hdiag"][i].cpu().numpy()
        _print_it2d(C3, xxx, f"Wire/C/C.{ws}.k.kp.")
        _print_it2d(D3, xxx, f"Wire/D/D.{ws}.k.kp.")
        _print_it2d(p3, xxx, f"Wire/P/P.{ws}.k.kp.")
        _print_it(np.diag(C3), kr, xxx, f"Wire/ne/ne.{ws}.k.")
        _print_it(np.diag(D3), kr, xxx, f"Wire/nh/nh.{ws}.k.")
        _print_it(hd[0] / eV, kr, xxx, f"Wire/Ee/Ee.{ws}.k.")
        _print_it(hd[1] / eV, kr, xxx, f"Wire/Eh/Eh.{ws}.k.")
        if w == int(np.ceil(self.Nw / 2.0)):
            # the middle wire's statistics on the console (SBEs.py:5755-5767),
            # with CalcVD / CalcPD (dcfield.py:128-178)
            for lab, n, m in (("Elec", np.diag(C3), self.me), ("Hole", np.diag(D3), self.mh)):
                vd = np.sum(np.real(n) * hbar * kr / m) / np.sum(np.real(n) + 1e-200)
                pd = np.sum(np.abs(n) * hbar * kr) / np.sum(np.abs(n) + 1e-200)
                print(f"{xxx:8d} {lab} {np.sum(np.real(n)):18.6e} {np.max(np.real(n)):18.6e} "
                      f"{np.min(np.real(n)):18.6e} {vd:18.6e} {np.real(pd):18.6e}")
        if self.DiagDph:
            gs = 8 * N + 4 * (2 * N - 1)  # PSA_PREP_GAME
            _print_itr(pr[gs:gs + N], kr, xxx, f"Wire/Ge/Ge.{ws}.k.", width=5)
            _print_itr(pr[gs + N:gs + 2 * N], kr, xxx, f"Wire/Gh/Gh.{ws}.k.", width=5)
        # QW2Prop: the QW-grid fields (window x spline of the rr fields),
        # polarizations and charge densities on r
        src = g.s["src"][i].cpu().numpy()
        win = g.tab["window"]
        Eq = [win * np.real(self._m1 @ np.asarray(E)) + 0j for E in (Exx, Eyy, Ezz)]
        P = [src[c * Nr:(c + 1) * Nr] + 0j for c in range(3)]
        re, rh = src[3 * Nr:4 * Nr] + 0j, src[4 * Nr:5 * Nr] + 0j
        zq = np.zeros(Nr, dtype=complex)
        for nm, v in (("Ex", Eq[0]), ("Ey", Eq[1]), ("Ez", Eq[2]), ("Vr", zq), ("Px", P[0]),
                      ("Py", P[1]), ("Pz", P[2]), ("Re", re), ("Rh", rh), ("Rho", rh - re)):
            _print_it(v, r, xxx, f"Wire/{nm}/{nm}.{ws}.r.")
        # propagation grid: RhoE, RhoH as QW2Prop forms them (splines to rr,
        # then |.| renormalised to the QW-grid total, qwoptics.py:795-805)
        if self._m2 is None:
            self._m2 = tables.spline_matrix(self.r, self.rr)
        dr = r[1] - r[0]
        dRR = rr[1] - rr[0] if len(rr) > 1 else 1.0
        RhoE = self._m2 @ re
        RhoH = self._m2 @ rh
        if self.LF:
            total = (np.sum(re) * dr + np.sum(rh) * dr + 1e-100) / 2.0
            RhoE = np.abs(RhoE) + 0j
            RhoH = np.abs(RhoH) + 0j
            RhoE = RhoE * total / (np.sum(RhoE) * dRR + 1e-100)
            RhoH = RhoH * total / (np.sum(RhoH) * dRR + 1e-100)
        zr = np.zeros(len(rr), dtype=complex)
        for nm, v in (("Ex", Exx), ("Ey", Eyy), ("Ez", Ezz), ("Vr", zr), ("Px", out[0]),
                      ("Py", out[1]), ("Pz", out[2]), ("Re", RhoE), ("Rh", RhoH)):
            _print_it_real2(v, rr, xxx, f"Prop/{nm}/{nm}.{ws}.r.")
        _print_it(RhoH - RhoE, rr, xxx, f"Prop/Rho/Rho.{ws}.r.")

    def xqw_table(self, fe, fh, game, gamh):
        """chi(q, w) for q = 0..Nk (units of dkr) and w = iw dw, dw = 2 gap /
        hbar / 2000 (SBEs.py:5395-5407), on the device (psa_xqw_table).
        Returns (w [nw], X [Nk+1][nw] complex)."""
        import ctypes
        from . import capi
        lib = self.group.lib
        wmax = self.gap / hbar * 2.0
        dw = wmax / 2000.0
        nw = int(wmax / dw) + 1
        N = self.Nk
        dev = self.group.dev
        T = lambda a: torch.as_tensor(np.ascontiguousarray(a, dtype=float), device=dev)  # noqa: E731
        args = [T(self.Ee), T(self.Eh), T(fe), T(fh), T(game), T(gamh)]
        out = torch.empty((N + 1, nw), dtype=torch.complex128, device=dev)
        x = 2.0 * self.dkr / twopi / self.area * self.dcv * self.dcv / eps0
        stream = ctypes.c_void_p(torch.cuda.current_stream(dev).cuda_stream)
        capi.check(lib.psa_xqw_table(N, *[a.data_ptr() for a in args], self.gap, x, hbar,
                                     1e-4 * e0, dw, nw, out.data_ptr(), stream))
        return np.arange(nw) * dw, out.cpu().numpy()

    def record_xqw(self, fe, fh, game, gamh, ind):
        """_record_Xqw (SBEs.py:5395-5432): dataQW/Wire/Xqw/Xqw.<ind>.dat,
        lines "w Re Im" (16.6e) with q fastest, and Xqw.params once when the
        Xqwparams flag is set."""
        w, X = self.xqw_table(fe, fh, game, gamh)
        os.makedirs("dataQW/Wire/Xqw", exist_ok=True)
        nq = X.shape[0]
        rows = np.empty((len(w) * nq, 3))
        rows[:, 0] = np.repeat(w, nq)
        rows[:, 1] = X.T.real.reshape(-1)
        rows[:, 2] = X.T.imag.reshape(-1)
        np.savetxt(f"dataQW/Wire/Xqw/Xqw.{ind:06d}.dat", rows, fmt="%16.6e %16.6e %16.6e")
        if getattr(self, "Xqwparams", False):
            wmax = self.gap / hbar * 2.0
            dw = wmax / 2000.0
            with open("dataQW/Wire/Xqw/Xqw.params", "w", encoding="utf-8") as f:
                f.write(f"Nq = {self.Nk + 1}\n")
                f.write(f"Nw = {len(w)}\n")
                f.write(" \n")
                f.write(f"hbar dw (eV) = {dw * hbar / e0}\n")
                f.write(f"dq (rad/m) = {self.dkr}\n")
                f.write(" \n")
                f.write("wmin = 0.0\n")
                f.write(f"wmax = {wmax * hbar / e0}\n")
                f.write("qmin = 0.0\n")
                f.write(f"qmax = {self.Nk * self.dkr}\n")
            self.Xqwparams = False

    # ------------------------------------------------------------------
    # reference-style state access ([N, N, Nw], host copies)
    # ------------------------------------------------------------------
    def _levels(self):
        return self.group.get_state()

    @property
    def CC2(self):
        return self._levels()[0]

    @property
    def DD2(self):
        return self._levels()[1]

    @property
    def YY2(self):
        return self._levels()[2]

    @property
    def CC3(self):
        return self._levels()[3]

    @property
    def DD3(self):
        return self._levels()[4]

    @property
    def YY3(self):
        return self._levels()[5]

    def set_levels(self, C2, D2, p2, C3, D3, p3):
        """Load levels 2 and 3 (reference layout [N, N, Nw])."""
        self.group.set_state(C2, D2, p2, C3, D3, p3)

    # ------------------------------------------------------------------
    def ShutOffOptics(self):
        """Disable optical coupling (SBEs.py:5516-5518)."""
        self.Optics = False
        self.prm.flags["Optics"] = False
        self.group._sync_struct()

    def chiqw(self):
        return self.chiw

    def getqc(self):
        return self.qc

    def QWArea(self):
        return self.area

    # ------------------------------------------------------------------
    def _bookkeeping(self, w, dt):
        """SBEs.py:5837-5847."""
        if w == 1:
            if self.jjj == self.jmax:
                self.jjj = 0
            self.xxx += 1
            self.jjj += 1
            self.t += dt
        return self.jjj == self.jmax

    def _gate(self, Exx, Eyy, Ezz):
        """SBEs.py:5849-5858: True while the wire stays off."""
        mag = np.max(np.sqrt(np.abs(Exx) ** 2 + np.abs(Eyy) ** 2 + np.abs(Ezz) ** 2))
        if self.wireoff and mag < 1e-3 * self.Emax0:
            return True
        self.wireoff = False
        return False

    def _snapshot_level1(self, w0, nw):
        if not self.keep_level1:
            return
        s = self.group.s
        self.group.average_levels(w0, nw)  # level A as the reference keeps it
        if self._lvl1 is None:
            self._lvl1 = {k: torch.zeros_like(s[k]) for k in ("pA", "CA", "DA")}
        for k in ("pA", "CA", "DA"):
            self._lvl1[k][w0:w0 + nw].copy_(s[k][w0:w0 + nw])

    def QWCalculator(self, Exx, Eyy, Ezz, Vrr, rr, q, dt, w, Pxx, Pyy, Pzz, Rho, DoQWP, DoQWDl):
        """One step of wire w (1-based) on the GPU; outputs in place."""
        Pxx[:] = 0.0
        Pyy[:] = 0.0
        Pzz[:] = 0.0
        Rho[:] = 0.0
        write_fields = self._bookkeeping(w, dt)
        if self._gate(Exx, Eyy, Ezz):
            return
        DoQWP[0] = self.Optics
        DoQWDl[0] = self.LF
        i = w - 1
        g = self.group
        g.set_fields(np.asarray(Exx), np.asarray(Eyy), np.asarray(Ezz), w0=i, nw=1)
        self._snapshot_level1(i, 1)
        g.step(dt, w0=i, nw=1)
        P = g.outputs(i, 1)
        out = (P[0][0], P[1][0], P[2][0], P[3][0])
        self.I0[i] = g.observables()["I0"][i]
        # The P.E energy and the Prop/* snapshots see the wire's real sources;
        # ignorewire only zeroes what the caller gets back (SBEs.py:5919-5985).
        self._energy(np.asarray(Exx), np.asarray(Eyy), np.asarray(Ezz), out, i)
        if self.io:
            self._write_files(np.asarray(Exx), np.asarray(Eyy), np.asarray(Ezz), dt, w)
            if self.OffDiagDph:
                self._write_max_offdiag(i)
            if write_fields:
                self._write_snapshots(np.asarray(Exx), np.asarray(Eyy), np.asarray(Ezz), out, w)
        if self.ignorewire:
            out = tuple(np.zeros_like(o) for o in out)
        Pxx[:] = out[0]
        Pyy[:] = out[1]
        Pzz[:] = out[2]
        Rho[:] = out[3]

    def QWCalculatorBatch(self, Exx, Eyy, Ezz, dt):
        """Step every wire at once; fields [Nw, Nx] (or [Nx] shared) as numpy
        or device tensors.  Returns device tensors (Pxx, Pyy, Pzz, Rho) of
        shape [Nw, Nx].  Bookkeeping as Nw successive QWCalculator calls;
        the gate is evaluated on the host when the fields are numpy."""
        self._bookkeeping(1, dt)
        if isinstance(Exx, np.ndarray) and self._gate(Exx, Eyy, Ezz):
            z = torch.zeros(self.Nw, len(self.rr), dtype=torch.complex128, device=self.group.dev)
            return z, z.clone(), z.clone(), z.clone()
        self.wireoff = False
        g = self.group
        g.set_fields(Exx, Eyy, Ezz)
        self._snapshot_level1(0, self.Nw)
        g.step(dt)
        o = g.s["out"]
        return o[0], o[1], o[2], o[3]

    # ------------------------------------------------------------------
    def _energy(self, Exx, Eyy, Ezz, out, i):
        """P.E bookkeeping on the propagation grid (SBEs.py:5925-5979); the
        reference keeps one history for all wires."""
        Pxx, Pyy, Pzz, _ = out
        drr = self.rr[1] - self.rr[0] if len(self.rr) > 1 else 1.0
        if self._P0 is None:
            self._P0 = [np.zeros(len(Pxx)) for _ in range(3)]
            self._P1 = [np.zeros(len(Pxx)) for _ in range(3)]
            if self.io:
                os.makedirs("output", exist_ok=True)
                self._file_972 = open("output/PdotE.p.dat", "w", encoding="utf-8")
                self._file_973 = open("output/PdotE.w.dat", "w", encoding="utf-8")
        for c, (P, E) in enumerate(((Pxx, Exx), (Pyy, Eyy), (Pzz, Ezz))):
            self.EPEnergy += np.sum((np.real(P) - self._P0[c]) * np.real(E) * 0.5) * drr * self.area
        for c, P in enumerate((Pxx, Pyy, Pzz)):
            self._P0[c][:] = self._P1[c]
            self._P1[c][:] = np.real(P)
        if self._m1 is not None:
            # same on the wire grid with the real-space QW fields and sources
            g = self.group
            W = g.tab["window"]
            src = g.s["src"][i].cpu().numpy()
            Nr = self.Nr
            if not hasattr(self, "_PW"):
                self._PW = [[np.zeros(Nr) for _ in range(3)] for _ in range(2)]
            for c, E in enumerate((Exx, Eyy, Ezz)):
                Eq = W * np.real(self._m1 @ E)
                Pq = src[c * Nr:(c + 1) * Nr]
                self.EPEnergyW += np.sum((Pq - self._PW[0][c]) * Eq * 0.5) * self.dr * self.area
                self._PW[0][c][:] = self._PW[1][c]
                self._PW[1][c][:] = Pq
        if self._file_972 is not None:
            self._file_972.write(f"{self.t} {self.EPEnergy / e0}\n")
            self._file_972.flush()
            self._file_973.write(f"{self.t} {self.EPEnergyW / e0}\n")
            self._file_973.flush()

    def _write_files(self, Exx, Eyy, Ezz, dt, w):
        """dataQW/info.WW.t.dat and EP.WW.t.dat lines (SBEs.py:5773-5811)."""
        i = w - 1
        g = self.group
        N, Nr = self.Nk, self.Nr
        g.average_levels(i, 1)  # level 2 as the reference keeps it
        C2 = g.s["CA"][i].diagonal().cpu().numpy()
        D2 = g.s["DA"][i].diagonal().cpu().numpy()
        ne2 = np.abs(C2).astype(complex)
        nh2 = np.abs(D2).astype(complex)
        kr, Ee, Eh, me, mh = self.kr, self.Ee, self.Eh, self.me, self.mh
        small = 1e-200

        def vd(m, n):
            return np.sum(np.real(n) * hbar * kr / m) / np.sum(np.real(n) + small)

        def temp(n, E):
            return 2.0 * np.real(np.sum(n * E) / (np.sum(n) + 1e-100)) / kB

        def i0n(n, m):
            dk = kr[1] - kr[0]
            return -e0 * np.sum(n * kr * hbar / m) * 2.0 * dk

        src = g.s["src"][i].cpu().numpy()
        rc = src[5 * Nr:9 * Nr].view(complex).reshape(2, Nr)
        X3d = (g.s["CB"][i].diagonal().cpu().numpy(), g.s["DB"][i].diagonal().cpu().numpy())
        dr = self.r[1] - self.r[0]
        rmax = []
        for c in range(2):
            re = rc[c].copy()
            re = re - np.real(np.sum(re[0:10]) + np.sum(re[Nr - 10:Nr])) / 20.0
            tot = np.sum(np.abs(X3d[c])) + 1e-50
            re = re * tot / (np.sum(np.abs(re)) * dr + 1e-100)
            Rq = -np.fft.fft(re) / Nr * 2.0 * e0 / self.area * self.ehint
            rmax.append(np.max(np.abs(Rq)))
        cols = [self.xxx * dt, vd(me, ne2), vd(mh, nh2),
                np.real(2 * np.sum(ne2) / self.L), np.real(2 * np.sum(nh2) / self.L),
                np.real(np.sum(ne2 * (Ee + self.gap))) / e0, np.real(np.sum(nh2 * Eh)) / e0,
                temp(ne2, Ee), temp(nh2, Eh), rmax[0], rmax[1], 0.0,
                np.real(hbar * np.sum((ne2 - nh2) * kr)) * 2.0,
                np.real(hbar * np.sum(ne2 * kr)), np.real(-hbar * np.sum(nh2 * kr)),
                self.I0[i], np.real(i0n(ne2, me)), np.real(i0n(1 - nh2, -mh))]
        with open(f"dataQW/info.{w:02d}.t.dat", "a", encoding="utf-8") as f:
            f.write(" ".join(f"{c:18.6e}" for c in cols) + "\n")
        W = g.tab["window"][Nr // 2]
        Ef = [W * np.real(self._m1[Nr // 2] @ E) for E in (Exx, Eyy, Ezz)]
        Pm = [src[c * Nr + Nr // 2] for c in range(3)]
        with open(f"dataQW/EP.{w:02d}.t.dat", "a", encoding="utf-8") as f:
            f.write(" ".join(f"{c:18.6e}" for c in [self.xxx * dt] + Ef + Pm) + "\n")

    # ------------------------------------------------------------------
    # backups (SBEs.py:5524-5569): one "re im" pair per line, order w, k2, k1
    # ------------------------------------------------------------------
    def _all_levels(self):
        C2, D2, p2, C3, D3, p3 = self._levels()
        if self._lvl1 is not None:
            l1 = {k: np.moveaxis(v.cpu().numpy(), 0, 2) for k, v in self._lvl1.items()}
            C1, D1, p1 = l1["CA"], l1["DA"], l1["pA"]
        else:
            C1, D1, p1 = np.zeros_like(C2), np.zeros_like(D2), np.zeros_like(p2)
        return dict(CC1=C1, CC2=C2, CC3=C3, DD1=D1, DD2=D2, DD3=D3, YY1=p1, YY2=p2, YY3=p3)

    def write_sbes_data(self, n):
        d = "dataQW/backup/"
        os.makedirs(d, exist_ok=True)
        for key, arr in self._all_levels().items():
            with open(f"{d}{key}.{n}.dat", "w", encoding="utf-8") as f:
                for w in range(arr.shape[2]):
                    blk = arr[:, :, w].T.reshape(-1)  # k2 outer, k1 inner
                    f.write("".join(f"{np.real(v)} {np.imag(v)}\n" for v in blk))

    def read_sbes_data(self, Nt):
        d = "dataQW/backup/"
        N, Nw = self.Nk, self.Nw
        lv = {}
        for key in ("CC1", "CC2", "CC3", "DD1", "DD2", "DD3", "YY1", "YY2", "YY3"):
            v = np.loadtxt(f"{d}{key}.{Nt}.dat").reshape(Nw, N, N, 2)
            a = v[..., 0] + 1j * v[..., 1]          # [w][k2][k1]
            lv[key] = np.moveaxis(np.swapaxes(a, 1, 2), 0, 2)  # [k1][k2][w]
        self.set_levels(lv["CC2"], lv["DD2"], lv["YY2"], lv["CC3"], lv["DD3"], lv["YY3"])
        if self.keep_level1:
            self._snapshot_level1(0, Nw)
            for k, key in (("CA", "CC1"), ("DA", "DD1"), ("pA", "YY1")):
                self._lvl1[k].copy_(torch.as_tensor(np.ascontiguousarray(
                    np.moveaxis(lv[key], 2, 0)), device=self.group.dev))

    def close(self):
        for f in (self._file_972, self._file_973):
            if f is not None:
                f.close()
        self._file_972 = self._file_973 = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


# ---------------------------------------------------------------------------
# module-level shims (SBEs.py:6008-6053)
# ---------------------------------------------------------------------------
_default_solver = None


def InitializeSBE(q, rr, r0, Emaxxx, lam, Nw, QW):
    global _default_solver
    _default_solver = SBESolver(q, rr, r0, Emaxxx, lam, Nw, QW)


def QWCalculator(Exx, Eyy, Ezz, Vrr, rr, q, dt, w, Pxx, Pyy, Pzz, Rho, DoQWP, DoQWDl):
    _default_solver.QWCalculator(Exx, Eyy, Ezz, Vrr, rr, q, dt, w, Pxx, Pyy, Pzz, Rho,
                                 DoQWP, DoQWDl)


def chiqw():
    return _default_solver.chiqw()


def getqc():
    return _default_solver.getqc()


def QWArea():
    return _default_solver.QWArea()


def ShutOffOptics():
    _default_solver.ShutOffOptics()


def WriteSBEsData(n):
    _default_solver.write_sbes_data(n)


def ReadSBEsData(Nt):
    _default_solver.read_sbes_data(Nt)
