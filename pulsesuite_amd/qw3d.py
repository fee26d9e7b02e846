"""3-D coupling of the quantum wires to a Maxwell grid: the per-step part of
``rhoPJ.QuantumWire`` (``PSTD3D/rhoPJ.py:117-490``) on the device, behind
``psa_qw3d_drive`` / ``psa_qw3d_deposit`` of the C ABI.

The wires of one ``WireGroup`` lie along x of a grid [nx][ny][nz] (their
propagation grid ``rr`` is the grid's x axis, as QuantumWire passes ``rx``).
Each step (:meth:`QuantumWire3D.step`):

* drive: the grid fields at the wire point (x = 0, y0, z0), trilinearly
  interpolated (``EAtXYZ``, ``usefulsubs.py:1996-2065``), broadcast along
  every wire;
* the wires step (``psa_step``);
* deposit: J = sum over wires of P / dt / Nw, plus ``CalcJfx`` of the
  summed charge / Nw along x (``rhoPJ.py:491-523``); J placed into Jx, Jy,
  Jz, the mean charge into Rho and the mean polarization into fresh Px, Py,
  Pz with the profiles gy = exp(-(y - y0)^2 / 2 ay^2), gz likewise
  (``QWPlacement``, ``rhoPJ.py:655-700``); then the longitudinal fields of
  ``ElongfromRho`` (``rhoPJ.py:703-880``) by 3-D FFTs.

As in the reference, placement zeroes a grid before it deposits
(``rhoPJ.py:682``), so Jx, Jy, Jz and Rho are overwritten every step (a grid
whose flag is off keeps the caller's values), and the previous polarization
and charge are discarded at the top of every call (``rhoPJ.py:256-270``), so
J = sum P / (dt Nw) and CalcJfx sees a previous charge of zero.
``keep_old=True`` departs from the reference on purpose: it keeps the
previous call's P and Rho, so J is the polarization current
sum (P - P_old) / (dt Nw) and CalcJfx sees the change of the charge.

The 3-D Maxwell propagator that would consume J, Rho and the longitudinal
fields is not part of this package; the caller owns the grid arrays.
"""

import ctypes

import numpy as np
import torch
from scipy.constants import epsilon_0 as eps0

from . import capi

C128 = torch.complex128
F64 = torch.float64

EL_NAMES = ("ExlfromPRho", "EylfromPRho", "EzlfromPRho", "ExlfromP", "EylfromP", "EzlfromP",
            "ExlfromRho", "EylfromRho", "EzlfromRho", "RhoBound")


def _locator(x, x0):
    """usefulsubs.py:1805-1830: i with x[i] <= x0 < x[i+1], clipped to
    [0, len(x) - 2]."""
    i = int(np.searchsorted(x, x0, side="right")) - 1
    return max(0, min(i, len(x) - 2))


def _len_ok(n):
    if n < 2 or n > 512:
        return False
    while n % 2 == 0:
        n //= 2
    while n % 3 == 0:
        n //= 3
    return n == 1


class QuantumWire3D:
    """group -- the wires (``WireGroup`` with Nx = len(rx));
    rx, ry, rz -- grid coordinates; qx, qy, qz -- grid wavenumbers (FFT
    order); y0, z0, ay, az -- wire position and profile widths
    (``params/qwarray.params``); epsr -- background dielectric constant;
    dt -- step; curr, rho, pol -- DoQWCurr, DoQWDl, DoQWP; keep_old -- use
    the previous call's P and Rho in J (not the reference's behaviour)."""

    def __init__(self, group, rx, ry, rz, qx, qy, qz, y0, z0, ay, az, epsr, dt,
                 curr=True, rho=True, pol=True, keep_old=False):
        self.g = group
        self.lib = capi.load()
        rx, ry, rz = (np.asarray(a, dtype=float) for a in (rx, ry, rz))
        nx, ny, nz = len(rx), len(ry), len(rz)
        if not all(_len_ok(n) for n in (nx, ny, nz)):
            raise ValueError("grid sizes must be 2^a 3^b in [2, 512]")
        if group.Nx != nx:
            raise ValueError("the wires' propagation grid must be the grid's x axis")
        self.shape = (nx, ny, nz)
        dev = group.dev
        self.dev = dev
        T = lambda a, dt=F64: torch.as_tensor(np.ascontiguousarray(a), dtype=dt).to(dev)  # noqa: E731
        cell, w = [], []
        for a, a0 in ((rx, 0.0), (ry, y0), (rz, z0)):
            i = _locator(a, a0)
            cell.append(i)
            w.append((a0 - a[i]) / (a[i + 1] - a[i]))
        n = nx * ny * nz
        self.t = {
            "qx": T(qx), "qy": T(qy), "qz": T(qz),
            "gy": T(np.exp(-((ry - y0) ** 2) / (2.0 * ay**2))),
            "gz": T(np.exp(-((rz - z0) ** 2) / (2.0 * az**2))),
            "tw": T(np.concatenate([np.exp(-2j * np.pi * np.arange(m) / m) for m in (nx, ny, nz)]), C128),
            "Fk": torch.zeros(4, n, dtype=C128, device=dev),
            "El": torch.zeros(10, nx, ny, nz, dtype=C128, device=dev),
            "lines": torch.zeros(8, nx, dtype=C128, device=dev),
            "old": torch.zeros(4, group.nw, nx, dtype=C128, device=dev),
        }
        S = capi.PsaQw3d()
        S.nx, S.ny, S.nz = nx, ny, nz
        S.flags = ((capi.PSA_QW3D_CURR if curr else 0) | (capi.PSA_QW3D_RHO if rho else 0)
                   | (capi.PSA_QW3D_POL if pol else 0) | (capi.PSA_QW3D_KEEP_OLD if keep_old else 0))
        S.i0, S.j0, S.k0 = cell
        S.ax, S.ay, S.az = w
        S.dt, S.dx, S.eps = float(dt), float(rx[1] - rx[0]), eps0 * float(epsr)
        for k in ("qx", "qy", "qz", "gy", "gz", "tw", "Fk", "El", "lines", "old"):
            setattr(S, k, self.t[k].data_ptr())
        self.S = S
        self.dt = float(dt)

    def _grid(self, a, name):
        if not (isinstance(a, torch.Tensor) and a.dtype == C128 and a.is_contiguous()
                and tuple(a.shape) == self.shape and a.is_cuda):
            raise ValueError(f"{name}: a contiguous complex128 device tensor of shape {self.shape}")
        return a.data_ptr()

    def _stream(self):
        return ctypes.c_void_p(torch.cuda.current_stream(self.dev).cuda_stream)

    def drive(self, Ex, Ey, Ez):
        self.S.Ex, self.S.Ey, self.S.Ez = (self._grid(a, n) for a, n in ((Ex, "Ex"), (Ey, "Ey"), (Ez, "Ez")))
        capi.check(self.lib.psa_qw3d_drive(ctypes.byref(self.S), ctypes.byref(self.g.S), self._stream()))

    def deposit(self, Jx, Jy, Jz, Rho):
        self.S.Jx, self.S.Jy, self.S.Jz, self.S.Rho = (
            self._grid(a, n) for a, n in ((Jx, "Jx"), (Jy, "Jy"), (Jz, "Jz"), (Rho, "Rho")))
        capi.check(self.lib.psa_qw3d_deposit(ctypes.byref(self.S), ctypes.byref(self.g.S), self._stream()))

    def step(self, Ex, Ey, Ez, Jx, Jy, Jz, Rho):
        """One QuantumWire call: drive, the wires' step, deposit.  Jx, Jy,
        Jz, Rho are overwritten in place when placed; the longitudinal fields and the
        bound charge are in :attr:`El` (device, [10][nx][ny][nz], order
        ``EL_NAMES``)."""
        self.drive(Ex, Ey, Ez)
        self.g.step(self.dt)
        self.deposit(Jx, Jy, Jz, Rho)
        return self.t["El"]

    @property
    def El(self):
        return self.t["El"]
