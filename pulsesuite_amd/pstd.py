"""1-D pseudo-spectral field propagator coupled to the quantum wires.

The reference drives its wires with spatially uniform analytic pulses
(``PSTD3D/sbetestprop.py:183-198``) and has no propagator on that path; it
ships the staggered differences ``dEdx``/``dHdx`` (``usefulsubs.py:612-655``,
``704-747``) and the real-space absorber ``ApplyABC`` (``usefulsubs.py:584-609``).
``Pstd1D`` is a propagator built from them, on the device behind the
``psa_pstd_*`` entry points of the C ABI:

* the pulse travels along y, normal to the wires, which lie along x (the
  reference's 3-D geometry: ``Ystart``/``Ywin`` in ``params/pulse.params``
  set the propagation window, 8.192 um = 4096 points of 2 nm);
* (Ex, Hz) and (Ez, Hx) in a background of index n0, E on y_j = y0 + j dy,
  H on y_j + dy/2, held in k-space;
* y derivatives as k-space multipliers: ``"spectral"`` (i k exp(+-i k dy/2))
  or ``"fd"`` (the reference's dEdx / dHdx differences); each mode advanced
  over dt by the exact solution of its 2x2 system with the current held
  constant over the step, so dt has no stability limit;
* an optional absorbing profile applied in real space once per step
  (ApplyABC), built by :func:`abc_profile`;
* wire slot s on grid point jw[s]: the transverse field there drives the
  wire (Exx, Ezz broadcast along its propagation grid rr), and the wire's
  Pxx, Pzz averaged over rr times ``sheet`` are the grid polarization at
  jw[s].  ``sheet = area / (lz dy)`` for a sheet of wires of cross-section
  ``area`` repeated every ``lz`` along z (:func:`sheet_factor`).

One coupled step (:meth:`Pstd1D.couple_step`): every wire group steps with
the field of the current time, the wires' new polarizations give the
current of the step, the propagator advances, and the new field is
scattered to the wires.  All of it is asynchronous on the current stream.
"""

import ctypes

import numpy as np
import torch
from scipy.constants import epsilon_0 as eps0
from scipy.constants import mu_0 as mu0

from . import capi

F64 = torch.float64
C128 = torch.complex128


def sheet_factor(area, lz, dy):
    """Wire polarization -> grid polarization for wires of cross-section
    ``area`` (m^2) repeated every ``lz`` (m) along z, on a grid of spacing dy."""
    return float(area) / (float(lz) * float(dy))


def abc_profile(N, dy, dt, n0, width_lo, width_hi, atten=10.0, power=3):
    """Real absorbing profile exp(-sigma(y) dt) with sigma = s_max (d/W)^power
    over a layer of width W at each end (d = depth into the layer) and
    s_max set so that a pulse crossing the layer once is damped by
    exp(-atten): s_max W / ((power + 1) v) = atten, v = c/n0."""
    v = 1.0 / np.sqrt(eps0 * n0 * n0 * mu0)
    j = np.arange(N, dtype=float)
    sig = np.zeros(N)
    for W, d in ((width_lo, (width_lo / dy) - j), (width_hi, j - (N - 1 - width_hi / dy))):
        if W and W > 0:
            u = np.clip(d / (W / dy), 0.0, 1.0)
            sig += atten * (power + 1) * v / W * u ** power
    return np.exp(-sig * dt)


def _mode_table(N, dy, eps, mu, dt, deriv):
    k = 2.0 * np.pi * np.fft.fftfreq(N, dy)
    if deriv == "spectral":
        ike = 1j * k * np.exp(0.5j * k * dy)
        ikh = 1j * k * np.exp(-0.5j * k * dy)
    elif deriv == "fd":
        z = np.exp(2j * np.pi * np.arange(N) / N)
        ike = (z - 1.0) / dy
        ikh = (1.0 - np.conj(z)) / dy
    else:
        raise ValueError(f"unknown derivative {deriv!r}")
    w2 = np.maximum(-(ike * ikh).real, 0.0) / (eps * mu)
    w = np.sqrt(w2)
    pos = w > 0
    ws = np.where(pos, w, 1.0)
    C = np.cos(w * dt)
    S = np.where(pos, np.sin(w * dt) / ws, dt)
    Q = np.where(pos, 2.0 * np.sin(0.5 * w * dt) ** 2 / (ws * ws), 0.5 * dt * dt)
    t = np.zeros((N, 8))
    t[:, 0], t[:, 1] = ike.real, ike.imag
    t[:, 2], t[:, 3] = ikh.real, ikh.imag
    t[:, 4], t[:, 5], t[:, 6] = C, S, Q
    return t


def _ptr(t):
    return ctypes.c_void_p(t.data_ptr())


class Pstd1D:
    """The propagator on ``N`` points y_j = y0 + j dy (N a power of two,
    64..8192) in a background of index ``n0``, stepping by ``dt``.

    wires_y -- y of each wire slot (m), snapped to the nearest grid point
    sheet   -- wire polarization -> grid polarization (:func:`sheet_factor`)
    abc     -- None, (width_lo, width_hi) in m for :func:`abc_profile`, or a
               real profile array of N values
    obs_y   -- y of the recorded observation points (m)
    nrec    -- steps the record holds
    deriv   -- "spectral" or "fd"
    """

    def __init__(self, N, dy, n0, dt, wires_y=(), sheet=1.0, abc=None, obs_y=(), nrec=0,
                 y0=0.0, deriv="spectral", device="cuda"):
        self.lib = capi.load()
        if not torch.cuda.is_available():
            raise capi.PsaError("no GPU visible: the device path has no CPU fallback")
        N = int(N)
        if N < 64 or N > 8192 or N & (N - 1):
            raise ValueError("N must be a power of two in [64, 8192]")
        self.N, self.dy, self.n0, self.dt, self.y0 = N, float(dy), float(n0), float(dt), float(y0)
        self.eps, self.mu = eps0 * self.n0 ** 2, mu0
        self.v = 1.0 / np.sqrt(self.eps * self.mu)
        self.eta = np.sqrt(self.mu / self.eps)
        self.y = self.y0 + self.dy * np.arange(N)
        self.dev = torch.device(device)
        self.jw = self.grid_index(wires_y)
        self.obs = self.grid_index(obs_y)
        if len(self.jw) > capi.PSA_PSTD_MAXW:
            raise ValueError(f"at most {capi.PSA_PSTD_MAXW} wire slots")
        if abc is not None and not hasattr(abc, "__len__"):
            raise ValueError("abc: (width_lo, width_hi) or a profile array")
        if abc is not None and len(abc) == 2:
            abc = abc_profile(N, self.dy, self.dt, self.n0, abc[0], abc[1])
        d = self.dev
        T = lambda a, dt=F64: torch.as_tensor(np.ascontiguousarray(a), dtype=dt).to(d)  # noqa: E731
        nw, nobs, nrec = len(self.jw), len(self.obs), int(nrec)
        self.t = {
            "state": torch.zeros(4, N, dtype=C128, device=d),
            "mode": T(_mode_table(N, self.dy, self.eps, self.mu, self.dt, deriv)),
            "abc": T(np.ones(N) if abc is None else np.asarray(abc, dtype=float)),
            "twid": T(np.exp(-2j * np.pi * np.arange(N) / N), C128),
            "jw": T(self.jw if nw else np.zeros(1), torch.int32),
            "obs": T(self.obs if nobs else np.zeros(1), torch.int32),
            "pold": torch.zeros(2, max(nw, 1), dtype=C128, device=d),
            "jsrc": torch.zeros(2, max(nw, 1), dtype=C128, device=d),
            "ew": torch.zeros(2, max(nw, 1), dtype=C128, device=d),
            "rec": torch.zeros(max(nrec, 1), 2, max(nobs, 1), dtype=C128, device=d),
            "count": torch.zeros(1, dtype=torch.int64, device=d),
            "real": torch.zeros(4, N, dtype=C128, device=d),
        }
        if self.t["abc"].shape != (N,):
            raise ValueError("abc profile must have N values")
        S = capi.PsaPstd()
        S.N, S.nw, S.nobs = N, nw, nobs
        S.flags = capi.PSA_PSTD_ABC if abc is not None else 0
        S.dy, S.eps, S.mu, S.dt, S.sheet = self.dy, self.eps, self.mu, self.dt, float(sheet)
        for k in ("state", "mode", "abc", "twid", "jw", "obs", "pold", "jsrc", "ew", "rec", "count"):
            setattr(S, k, _ptr(self.t[k]))
        S.nrec = nrec
        self.S = S
        self.nw, self.nobs, self.nrec = nw, nobs, nrec

    # ------------------------------------------------------------------
    def grid_index(self, ys):
        j = np.rint((np.asarray(ys, dtype=float).reshape(-1) - self.y0) / self.dy).astype(np.int64)
        if j.size and (j.min() < 0 or j.max() >= self.N):
            raise ValueError("a wire or observation point lies outside the grid")
        return j.astype(np.int32)

    def _stream(self):
        return ctypes.c_void_p(torch.cuda.current_stream(self.dev).cuda_stream)

    def set_fields(self, Ex, Hz, Ez=None, Hx=None):
        """Real-space fields: E on y_j, H on y_j + dy/2 (None = zero)."""
        z = np.zeros(self.N)
        for i, f in enumerate((Ex, Hz, Ez, Hx)):
            self.t["state"][i].copy_(torch.as_tensor(z if f is None else f, dtype=C128).to(self.dev))
        capi.check(self.lib.psa_pstd_to_k(ctypes.byref(self.S), self._stream()))

    def pulse(self, E0, yc, width, lam, pol="x", direction=+1):
        """(E, H) of a pulse E0 exp(-((y - yc)/width)^2) cos(k (y - yc)),
        k = 2 pi n0 / lam, travelling along +y (direction +1) or -y, with
        its periodic images; H on the half points.  pol "x": (Ex, Hz),
        "z": (Ez, Hx)."""
        k0 = 2.0 * np.pi * self.n0 / lam
        L = self.N * self.dy
        yc = self.y0 + (yc - self.y0) % L

        def f(y):
            out = np.zeros_like(y)
            for m in (-1, 0, 1):
                u = y - yc - m * L
                out += E0 * np.exp(-(u / width) ** 2) * np.cos(k0 * u)
            return out
        E = f(self.y)
        Hh = f(self.y + 0.5 * self.dy) / self.eta
        # +y: Hz = -Ex/eta, Hx = +Ez/eta
        H = (-Hh if pol == "x" else Hh) * direction
        return E, H

    def fields(self):
        """Real-space (Ex, Hz, Ez, Hx) as numpy [4][N] (synchronises)."""
        capi.check(self.lib.psa_pstd_to_real(ctypes.byref(self.S), _ptr(self.t["real"]), self._stream()))
        return self.t["real"].cpu().numpy()

    def energy(self):
        """sum_j (eps |E|^2 + mu |H|^2) dy over both pairs (J/m^2)."""
        f = self.fields()
        return float((self.eps * (np.sum(np.abs(f[0]) ** 2) + np.sum(np.abs(f[2]) ** 2))
                      + self.mu * (np.sum(np.abs(f[1]) ** 2) + np.sum(np.abs(f[3]) ** 2))) * self.dy)

    def sample(self):
        capi.check(self.lib.psa_pstd_sample(ctypes.byref(self.S), self._stream()))

    def advance(self):
        capi.check(self.lib.psa_pstd_advance(ctypes.byref(self.S), self._stream()))

    def set_currents(self, J):
        """Current densities of the next step per wire slot, [2][nw] (x, z)."""
        self.t["jsrc"].copy_(torch.as_tensor(np.asarray(J), dtype=C128).to(self.dev))

    def gather(self, group, slot0, w0=0, nw=None):
        nw = group.nw - w0 if nw is None else nw
        capi.check(self.lib.psa_pstd_gather(ctypes.byref(self.S), ctypes.byref(group.S), int(w0), int(nw),
                                            int(slot0), self._stream()))

    def scatter(self, group, slot0, w0=0, nw=None):
        nw = group.nw - w0 if nw is None else nw
        capi.check(self.lib.psa_pstd_scatter(ctypes.byref(self.S), ctypes.byref(group.S), int(w0), int(nw),
                                             int(slot0), self._stream()))

    def wire_fields(self):
        """(Ex, Ez) at each wire slot, numpy [2][nw]."""
        return self.t["ew"][:, :self.nw].cpu().numpy()

    def records(self):
        """The observation record so far: numpy [steps][2][nobs] (Ex, Ez)."""
        c = min(int(self.t["count"].item()), self.nrec)
        return self.t["rec"][:c, :, :self.nobs].cpu().numpy()

    # ------------------------------------------------------------------
    # coupled stepping
    # ------------------------------------------------------------------
    def attach(self, groups):
        """Wire groups [(WireGroup, slot0), ...]: group g's wire w sits in
        slot slot0 + w.  Writes the current field into their efield."""
        for g, s0 in groups:
            if s0 < 0 or s0 + g.nw > self.nw:
                raise ValueError("wire group outside the slots")
        self.groups = list(groups)
        self.sample()
        for g, s0 in self.groups:
            self.scatter(g, s0)

    def couple_step(self):
        """One step of dt for the attached groups and the field."""
        for g, _ in self.groups:
            g.step(self.dt)
        for g, s0 in self.groups:
            self.gather(g, s0)
        self.advance()
        for g, s0 in self.groups:
            self.scatter(g, s0)
