"""CPU FP64 oracle for the 1-D pseudo-spectral field propagator coupled to
the quantum wires.

TEST INFRASTRUCTURE ONLY.  Nothing in the ``pulsesuite_amd`` package imports
this module; only ``tests/`` and ``__graft_entry__.smoke()`` use it, as the
checker.

The reference has no propagator on the wire path (sbetestprop drives the
wires with spatially uniform analytic pulses, ``PSTD3D/sbetestprop.py:183-198``).
It ships the building blocks of one: the staggered forward / backward
differences ``dEdx`` / ``dHdx`` (``PSTD3D/usefulsubs.py:612-655`` and
``704-747``) and the real-space absorber ``ApplyABC`` (inverse FFT, multiply
by a real profile, FFT; ``usefulsubs.py:584-609``).  This module restates
those three and adds the propagator built from them:

* geometry: the pulse travels along y, normal to the wires (which lie along
  x, as in the reference's 3-D grid where ``Ystart``/``Ywin`` set the
  propagation window, ``params/pulse.params``).  The transverse pairs
  (Ex, Hz) and (Ez, Hx) are propagated; E lives on the grid points y_j,
  H on the half points y_j + dy/2;
* Maxwell in a background of index n0 (eps = eps0 n0^2, mu = mu0):
  dEx/dt = (dHz/dy - Jx)/eps, dHz/dt = (dEx/dy)/mu,
  dEz/dt = (-dHx/dy - Jz)/eps, dHx/dt = -(dEz/dy)/mu;
* space: the y derivatives are multipliers in k-space (numpy.fft
  convention): spectral ``ik exp(+-ik dy/2)`` or the reference's differences
  ``(exp(ik dy) - 1)/dy`` (dEdx) and ``(1 - exp(-ik dy))/dy`` (dHdx);
* time: each k mode is advanced over dt by the exact solution of its 2x2
  system with the current density held constant over the step (the
  pseudo-spectral analytical time domain form: no CFL limit, exact
  vacuum dispersion of the chosen derivative);
* absorber: once per step, ApplyABC's real-space multiplication;
* wires: wire w sits on grid point j_w; the field there drives it
  (broadcast along its own propagation grid rr); its polarization averaged
  over rr, times ``sheet`` (cross-section / (transverse period * dy)), is
  the grid polarization at j_w, and J = dP/dt over the step.

Pinning (tests/test_pstd_oracle.py): the difference multipliers against
the reference's dEdx/dHdx restated below, vacuum translation of a pulse at
c/n0 against the closed form, conservation of the field energy, the field
radiated by a current sheet against the image-sum closed form
E = -(eta/2) K(t - |y - y0|/v), and the uniform (k = 0) field a net sheet
polarization leaves in a periodic box.
"""

import numpy as np
from scipy.constants import epsilon_0 as eps0
from scipy.constants import mu_0 as mu0

# array order of the state: Ex, Hz, Ez, Hx
EX, HZ, EZ, HX = 0, 1, 2, 3


# ---------------------------------------------------------------------------
# reference building blocks (1-D restatements)
# ---------------------------------------------------------------------------

def dEdx(E, dx):
    """Periodic forward difference (E[i+1] - E[i]) / dx
    (usefulsubs.py:612-655, one column)."""
    iu = np.roll(np.arange(E.shape[0]), -1)
    return (E[iu] - E) / dx


def dHdx(H, dx):
    """Periodic backward difference (H[i] - H[i-1]) / dx
    (usefulsubs.py:704-747, one column)."""
    idn = np.roll(np.arange(H.shape[0]), 1)
    return (H - H[idn]) / dx


def apply_abc(F, abc):
    """ApplyABC (usefulsubs.py:584-609) on a k-space line: IFFT, multiply by
    the real profile, FFT."""
    return np.fft.fft(np.fft.ifft(F) * abc)


# ---------------------------------------------------------------------------
# propagator
# ---------------------------------------------------------------------------

def multipliers(N, dy, deriv="spectral"):
    """(ike, ikh): k-space multipliers of d/dy taking E (grid points) to the
    half points and H (half points) to the grid points."""
    k = 2.0 * np.pi * np.fft.fftfreq(N, dy)
    if deriv == "spectral":
        return 1j * k * np.exp(0.5j * k * dy), 1j * k * np.exp(-0.5j * k * dy)
    if deriv == "fd":
        z = np.exp(2j * np.pi * np.arange(N) / N)  # exp(i k dy)
        return (z - 1.0) / dy, (1.0 - 1.0 / z) / dy
    raise ValueError(deriv)


def mode_coeffs(ike, ikh, eps, mu, dt):
    """cos(w dt), sin(w dt)/w, (1 - cos(w dt))/w^2 with w^2 = -ike ikh/(eps mu)."""
    w2 = np.maximum(-(ike * ikh).real / (eps * mu), 0.0)
    w = np.sqrt(w2)
    C = np.cos(w * dt)
    S = np.where(w > 0, np.sin(w * dt) / np.where(w > 0, w, 1.0), dt)
    h = 0.5 * w * dt
    Q = np.where(w > 0, 2.0 * np.sin(h) ** 2 / np.where(w > 0, w2, 1.0), 0.5 * dt * dt)
    return C, S, Q


class Pstd1D:
    """The propagator on N grid points of spacing dy (k-space state)."""

    def __init__(self, N, dy, n0, dt, deriv="spectral", abc=None, jw=(), sheet=1.0):
        self.N, self.dy, self.dt = int(N), float(dy), float(dt)
        self.eps, self.mu = eps0 * n0 * n0, mu0
        self.ike, self.ikh = multipliers(self.N, self.dy, deriv)
        self.C, self.S, self.Q = mode_coeffs(self.ike, self.ikh, self.eps, self.mu, self.dt)
        self.abc = None if abc is None else np.asarray(abc, dtype=float)
        self.jw = np.asarray(jw, dtype=int)
        self.sheet = float(sheet)
        self.F = np.zeros((4, self.N), dtype=complex)
        self.pold = np.zeros((2, len(self.jw)), dtype=complex)
        self.J = np.zeros((2, len(self.jw)), dtype=complex)

    def set_real(self, fields):
        self.F = np.fft.fft(np.asarray(fields, dtype=complex), axis=1)

    def real(self):
        return np.fft.ifft(self.F, axis=1)

    def energy(self):
        """sum_j (eps |E|^2 + mu |H|^2) dy over both pairs (Parseval)."""
        f = self.F
        return (self.eps * (np.sum(np.abs(f[EX]) ** 2) + np.sum(np.abs(f[EZ]) ** 2))
                + self.mu * (np.sum(np.abs(f[HZ]) ** 2) + np.sum(np.abs(f[HX]) ** 2))) * self.dy / self.N

    def gather(self, pbar, dt):
        """Wire polarizations pbar[2][nw] (x, z; averaged over rr) at the end
        of the step -> sheet current densities of the step."""
        p = np.asarray(pbar, dtype=complex) * self.sheet
        self.J = (p - self.pold) / dt
        self.pold = p

    def source_k(self, f):
        """k-space current density of the point sheets: sum_w J_w exp(-2 pi i k j_w / N)."""
        m = np.arange(self.N)
        out = np.zeros(self.N, dtype=complex)
        for w, j in enumerate(self.jw):
            out += self.J[f, w] * np.exp(-2j * np.pi * ((m * j) % self.N) / self.N)
        return out

    def advance(self):
        F = self.F
        for f, (e, h, s) in enumerate(((EX, HZ, 1.0), (EZ, HX, -1.0))):
            Jk = self.source_k(f) if len(self.jw) else 0.0
            a = s * self.ikh / self.eps
            b = s * self.ike / self.mu
            d = a * F[h] - Jk / self.eps
            E1 = self.C * F[e] + self.S * d
            H1 = F[h] + b * (self.S * F[e] + self.Q * d)
            F[e], F[h] = E1, H1
        if self.abc is not None:
            for i in range(4):
                F[i] = apply_abc(F[i], self.abc)

    def at_wires(self):
        """(Ex, Ez) at each wire's grid point, [2][nw] complex."""
        r = np.fft.ifft(self.F[[EX, EZ]], axis=1)
        return r[:, self.jw]


# ---------------------------------------------------------------------------
# closed forms used to pin the propagator
# ---------------------------------------------------------------------------

def gauss_pulse(y, E0, yc, width, k0):
    """E0 exp(-((y - yc)/width)^2) cos(k0 (y - yc))."""
    u = y - yc
    return E0 * np.exp(-(u / width) ** 2) * np.cos(k0 * u)


def periodic_images(y, y0, L, fn, nimg=2):
    """sum over periodic images m of fn(y - y0 - m L)."""
    out = 0.0
    for m in range(-nimg, nimg + 1):
        out = out + fn(y - y0 - m * L)
    return out


def sheet_field(y, t, y0, L, v, eta, K):
    """E of a current sheet K(t) (A/m, K = 0 for t < 0) at y0 in a periodic
    box of length L: -(eta/2) sum_images K(t - |y - y0 - m L| / v)."""
    return periodic_images(y, y0, L, lambda u: -0.5 * eta * K(t - np.abs(u) / v))
