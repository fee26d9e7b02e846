"""Drop-in for the reference's ``qwoptics`` module (``PSTD3D/qwoptics.py``):
the ``QWOptics`` class with its per-step transforms on the GPU, with the
reference's call signatures and in-place outputs.

* ``QWOptics(RR, L, dcv, kr, Qr, Ee, Eh, ehint, area, gap)``
  (``qwoptics.py:565-628``): the window ``_QWWindow``, the phases
  ``_Expikr``/``_Expikrc``, the dipole matrices ``_Xcv0 ... _Zvc0``,
  ``_Vol``, ``_dcv0`` as host attributes (init-time tables).
* per call, on the device: ``Prop2QW`` (spline of the propagation-grid
  fields, window, real part, FFTG), ``QWPolarization3``, ``QWRho5`` and
  ``QW2Prop`` (iFFTG, the Plasmonics renormalisations, spline back to the
  propagation grid) through ``psa_mod_*``.  The cubic spline is applied as
  the linear map it is (``rescale_1D`` is linear in the data): its matrix
  is built once per pair of grids on the host and kept on the device.
* ``yw``, ``QWChi1`` (host scalar), ``GetVn1n2`` (on the device,
  ``psa_mod_vn1n2``), and the file writers ``WriteQWFields``,
  ``WriteSBESolns``, ``WritePLSpectrum``, ``WritePropFields``,
  ``printITReal2``, ``printITReal``, ``printIT3D`` in the reference's formats
  (``qwoptics.py:1093-1386``).  ``printITReal2`` writes the coordinate column
  into the first file of the process only, a state it shares with the SBE
  drop-in's Prop snapshots (the reference's ``qwoptics._firsttime``).

The tables the methods read are the instance attributes at call time (the
reference's methods return early when they are None, and so do these).
Unlike the reference, constructing ``QWOptics`` writes no ``Envl.y`` file;
the SBE drop-in writes its start-up files itself.
"""

import hashlib

import numpy as np
import torch

from . import capi
from .coulomb import C128, F64, device, stream, to_dev, to_host
from .tables import spline_matrix

twopi = 2.0 * np.pi


def yw(w):
    """qwoptics.py:1093-1106: (-1)^floor((w - 1) / 2)."""
    return (-1) ** int(np.floor((w - 1) / 2.0))


def WriteQWFields(QY, Ex, Ey, Ez, Vr, Px, Py, Pz, Re, Rh, sp, w, xxx):
    """qwoptics.py:1316-1349: dataQW/Wire/<name>/<name>.<ww>.<sp>.<xxx>.dat."""
    from .SBEs import _print_it
    ws = f"{w:02d}"
    for nm, v in (("Ex", Ex), ("Ey", Ey), ("Ez", Ez), ("Vr", Vr), ("Px", Px), ("Py", Py), ("Pz", Pz),
                  ("Re", Re), ("Rh", Rh), ("Rho", Rh - Re)):
        _print_it(np.asarray(v), QY, xxx, f"Wire/{nm}/{nm}.{ws}.{sp}.")


def QWChi1(lam, dky, Ee, Eh, area, geh, dcv):
    """qwoptics.py:1109-1146: linear susceptibility of the wire."""
    from .SBEs import QWChi1 as _chi1
    return _chi1(lam, dky, Ee, Eh, area, geh, dcv)


def GetVn1n2(kr, rcv, Hcc, Hhh, Hcv, Vcc, Vvv, Vcv, Vvc):
    """qwoptics.py:1149-1186: the interaction matrices Vcc, Vvv, Vcv and
    Vvc = Vcv^H, written in place; on the device."""
    from scipy.constants import hbar
    N = len(kr)
    for a in (Hcc, Hhh, Hcv, Vcc, Vvv, Vcv, Vvc):
        if np.shape(a) != (N, N):
            raise ValueError("GetVn1n2: matrices need shape (Nk, Nk)")
    if len(rcv) != N:
        raise ValueError("GetVn1n2: rcv needs Nk entries")
    lib, dev = capi.load(), device()
    ins = [to_dev(a, C128, dev) for a in (rcv, Hcc, Hhh, Hcv)]
    outs = [torch.empty(N, N, dtype=C128, device=dev) for _ in range(4)]
    capi.check(lib.psa_mod_vn1n2(N, *[t.data_ptr() for t in ins], hbar, *[t.data_ptr() for t in outs],
                                 stream(dev)))
    for dst, t in zip((Vcc, Vvv, Vcv, Vvc), outs):
        dst[...] = to_host(t)


def printITReal2(Dx, z, n, file):
    """qwoptics.py:1189-1215: Re Dx per line, with z as a second column in
    the first file of the process; dataQW/<file><n:06d>.dat."""
    from .SBEs import _print_it_real2
    _print_it_real2(Dx, z, n, file)


def printITReal(Dx, z, n, file):
    """qwoptics.py:1218-1237: z and Re Dx per line, as float32."""
    import os
    fn = f"dataQW/{file}{n:06d}.dat"
    os.makedirs(os.path.dirname(fn), exist_ok=True)
    with open(fn, "w", encoding="utf-8") as f:
        f.write("".join(f"{np.float32(a)} {np.float32(np.real(b))}\n" for a, b in zip(z, Dx)))


def printIT3D(Dx, z, n, file):
    """qwoptics.py:1240-1260: Re and Im of the flattened array, "%.6e"."""
    import os
    fn = f"dataQW/{file}{n:06d}.dat"
    os.makedirs(os.path.dirname(fn), exist_ok=True)
    Dx = np.asarray(Dx)
    np.savetxt(fn, np.column_stack([np.real(Dx.flatten()), np.imag(Dx.flatten())]), fmt="%.6e")


def WriteSBESolns(ky, ne, nh, C, D, P, Ee, Eh, w, xxx):
    """qwoptics.py:1263-1291: |C|, |D|, |P| (printIT2D) and ne, nh, Ee, Eh
    (printIT) of wire w at index xxx under dataQW/Wire/."""
    from .SBEs import _print_it, _print_it2d
    ws = f"{w:02d}"
    for nm, v in (("C", C), ("D", D), ("P", P)):
        _print_it2d(v, xxx, f"Wire/{nm}/{nm}.{ws}.k.kp.")
    for nm, v in (("ne", ne), ("nh", nh), ("Ee", Ee), ("Eh", Eh)):
        _print_it(v, ky, xxx, f"Wire/{nm}/{nm}.{ws}.k.")


def WritePLSpectrum(hw, PLS, w, xxx):
    """qwoptics.py:1294-1313: dataQW/Wire/PL/pl.<ww>.hw.<xxx>.dat against hw / e."""
    from scipy.constants import e as e0

    from .SBEs import _print_it
    _print_it(np.asarray(PLS, dtype=complex), np.asarray(hw) / e0, xxx, f"Wire/PL/pl.{w:02d}.hw.")


def WritePropFields(y, Ex, Ey, Ez, Vr, Px, Py, Pz, Re, Rh, sp, w, xxx):
    """qwoptics.py:1352-1386: the propagation-grid fields (printITReal2) and
    Rho = Rh - Re (printIT) under dataQW/Prop/."""
    from .SBEs import _print_it, _print_it_real2
    ws = f"{w:02d}"
    for nm, v in (("Ex", Ex), ("Ey", Ey), ("Ez", Ez), ("Vr", Vr), ("Px", Px), ("Py", Py), ("Pz", Pz),
                  ("Re", Re), ("Rh", Rh)):
        _print_it_real2(v, y, xxx, f"Prop/{nm}/{nm}.{ws}.{sp}.")
    _print_it(np.asarray(Rh) - np.asarray(Re), y, xxx, f"Prop/Rho/Rho.{ws}.{sp}.")


def _key(a):
    a = np.ascontiguousarray(np.asarray(a, dtype=float))
    return (len(a), hashlib.sha1(a.tobytes()).hexdigest())


class QWOptics:
    """qwoptics.py:565-1090 with the per-step transforms on the device."""

    def __init__(self, RR, L, dcv, kr, Qr, Ee, Eh, ehint, area, gap):
        self.lib = capi.load()
        self.dev = device()
        self._small = 1e-100
        self._epsr = 9.1
        kr = np.asarray(kr, dtype=float)
        Nk = len(kr)
        self.CalcQWWindow(RR, L)
        self.CalcExpikr(RR, kr)
        self._dcv0 = dcv
        self._Vol = L * area / ehint
        sg = np.where(np.arange(Nk) % 2 == 0, 1.0, -1.0)
        self._Xcv0 = (dcv * np.broadcast_to(sg[None, :], (Nk, Nk))).astype(complex)
        self._Ycv0 = np.full((Nk, Nk), dcv, dtype=complex)
        self._Zcv0 = np.full((Nk, Nk), -dcv, dtype=complex)
        self._Xvc0 = np.conj(self._Xcv0.T)
        self._Yvc0 = np.conj(self._Ycv0.T)
        self._Zvc0 = np.conj(self._Zcv0.T)
        self._maps = {}

    # ---- init-time tables (qwoptics.py:630-663) ----
    def CalcQWWindow(self, YY, L):
        self._QWWindow = np.exp(-((np.asarray(YY, dtype=float) / (L / 2.0)) ** 150))

    def CalcExpikr(self, y, ky):
        self._Expikr = np.exp(1j * np.outer(ky, y))
        self._Expikrc = np.conj(self._Expikr)

    def Xcv(self, k, kp):
        return self._Xcv0[k, kp]

    def Ycv(self, k, kp):
        return self._Ycv0[k, kp]

    def Zcv(self, k, kp):
        return self._Zcv0[k, kp]

    # ---- device helpers ----
    def _map(self, x0, x1):
        """Device spline matrix of rescale_1D(x0, ., x1), built once per
        pair of grids."""
        k = (_key(x0), _key(x1))
        if k not in self._maps:
            self._maps[k] = to_dev(spline_matrix(x0, x1), F64, self.dev)
        return self._maps[k]

    def _dft(self, x, sign):
        """FFTG (sign -1) / iFFTG (sign +1) of the rows of device x [v][n]."""
        n = x.shape[-1]
        y = torch.empty_like(x)
        capi.check(self.lib.psa_mod_dft(n, x.shape[0], sign, -1.0 / n if sign < 0 else -1.0, x.data_ptr(),
                                        y.data_ptr(), stream(self.dev)))
        return y

    def _resample(self, M, x, w=None, real_in=False):
        y = torch.empty(x.shape[0], M.shape[0], dtype=C128, device=self.dev)
        capi.check(self.lib.psa_mod_resample(M.shape[0], M.shape[1], x.shape[0], M.data_ptr(),
                                             None if w is None else w.data_ptr(), int(real_in), x.data_ptr(),
                                             y.data_ptr(), stream(self.dev)))
        return y

    @staticmethod
    def _put(dst, src):
        dst[...] = src

    # ---- per-step transforms ----
    def Prop2QW(self, RR, Exx, Eyy, Ezz, Vrr, Edc, R, Ex, Ey, Ez, Vr, t, xxx):
        """qwoptics.py:667-715: spline RR -> R (zero outside RR), window,
        real part, FFTG; Ex, Ey, Ez, Vr filled in place."""
        for a in (Ex, Ey, Ez, Vr):
            a[:] = 0.0
        M = self._map(RR, R)
        x = to_dev(np.stack([Exx, Eyy, Ezz, Vrr]), C128, self.dev)
        W = to_dev(self._QWWindow, F64, self.dev) if self._QWWindow is not None else None
        y = self._dft(self._resample(M, x, W, real_in=True), -1)
        h = to_host(y)
        for a, v in zip((Ex, Ey, Ez, Vr), h):
            self._put(a, v)

    def QWPolarization3(self, y, ky, p, ehint, area, L, Px, Py, Pz, xxx, w):
        """qwoptics.py:809-940: Px, Py, Pz (FFTG'd) filled in place."""
        Px[:] = 0.0
        Py[:] = 0.0
        Pz[:] = 0.0
        if self._Xvc0 is None or self._Yvc0 is None or self._Zvc0 is None:
            return
        if self._Expikr is None or self._Expikrc is None or self._QWWindow is None:
            return
        N, Nr = len(ky), len(y)
        dev = self.dev
        pd = to_dev(p, C128, dev)
        X, Y, Z = (to_dev(a, C128, dev) for a in (self._Xvc0, self._Yvc0, self._Zvc0))
        E, W = to_dev(self._Expikr, C128, dev), to_dev(self._QWWindow, F64, dev)
        out = torch.empty(3, Nr, dtype=C128, device=dev)
        capi.check(self.lib.psa_mod_pol3(N, Nr, pd.data_ptr(), X.data_ptr(), Y.data_ptr(), Z.data_ptr(),
                                         E.data_ptr(), W.data_ptr(), float(yw(w)), float((-1) ** w),
                                         2 * ehint / area / L, out.data_ptr(), stream(dev)))
        h = to_host(self._dft(out, -1))
        for a, v in zip((Px, Py, Pz), h):
            self._put(a, v)

    def QWRho5(self, Qr, kr, R, L, kkp, p, CC, DD, ne, nh, re, rh, xxx, jjj):
        """qwoptics.py:942-1071: re, rh (FFTG'd) filled in place."""
        re[:] = 0.0
        rh[:] = 0.0
        if self._Expikr is None or self._Expikrc is None or self._QWWindow is None:
            return
        N, Nr = len(kr), len(R)
        dev = self.dev
        Cd, Dd = to_dev(CC, C128, dev), to_dev(DD, C128, dev)
        ned, nhd = to_dev(ne, C128, dev), to_dev(nh, C128, dev)
        E, W = to_dev(self._Expikr, C128, dev), to_dev(self._QWWindow, F64, dev)
        out = torch.empty(2, Nr, dtype=C128, device=dev)
        capi.check(self.lib.psa_mod_rho5(N, Nr, Cd.data_ptr(), Dd.data_ptr(), E.data_ptr(), W.data_ptr(),
                                         float(L), float(R[1] - R[0]), ned.data_ptr(), nhd.data_ptr(),
                                         out.data_ptr(), stream(dev)))
        h = to_host(self._dft(out, -1))
        self._put(re, h[0])
        self._put(rh, h[1])

    def QW2Prop(self, r, Qr, Ex, Ey, Ez, Vr, Px, Py, Pz, re, rh, RR, Pxx, Pyy, Pzz, RhoE, RhoH, w, xxx,
                WriteFields, Plasmonics):
        """qwoptics.py:717-805: iFFTG of the nine QW arrays (in place),
        the Plasmonics renormalisation, WriteQWFields, spline r -> RR into
        Pxx, Pyy, Pzz, RhoE, RhoH, and the renormalisation on RR."""
        dev = self.dev
        dr = r[1] - r[0]
        dRR = RR[1] - RR[0]
        qa = (Ex, Ey, Ez, Vr, Px, Py, Pz, re, rh)
        x = self._dft(to_dev(np.stack(qa), C128, dev), +1)
        total = torch.zeros(1, dtype=F64, device=dev)
        if Plasmonics:
            capi.check(self.lib.psa_mod_plasmonics(x.shape[1], float(dr), 0, x[7].data_ptr(), x[8].data_ptr(),
                                                   total.data_ptr(), stream(dev)))
        h = to_host(x)
        for a, v in zip(qa, h):
            self._put(a, v)
        if WriteFields:
            WriteQWFields(r, Ex, Ey, Ez, Vr, Px, Py, Pz, re, rh, "r", w, xxx)
        M = self._map(r, RR)
        y = self._resample(M, x[4:9].contiguous())
        if Plasmonics:
            capi.check(self.lib.psa_mod_plasmonics(y.shape[1], float(dRR), 1, y[3].data_ptr(), y[4].data_ptr(),
                                                   total.data_ptr(), stream(dev)))
        h = to_host(y)
        for a, v in zip((Pxx, Pyy, Pzz, RhoE, RhoH), h):
            self._put(a, v)
