"""CPU FP64 oracle for the PulseSuite quantum-wire SBE time step.

TEST INFRASTRUCTURE ONLY.  Nothing in the ``pulsesuite_amd`` package imports
this module; only ``tests/``, ``__graft_entry__.smoke()`` and ``bench.py``'s
CPU-baseline leg use it, as the checker.

This is a NumPy restatement of the reference algorithm, written from the
reference sources (file:line cited per function, paths relative to
``src/pulsesuite/``).  It reproduces the reference's quirks on purpose:

* leapfrog step + average (``PSTD3D/SBEs.py:5673-5706``), relaxation with
  ``dt`` (not ``2 dt``) (``SBEs.py:5312-5321``);
* ``FermiDistr(Ee + gap/2)`` for both carriers (``SBEs.py:4829-4840``);
* the second ``area`` assignment (``SBEs.py:4811-4813``);
* the ``Eps1D`` denominator without pi (``coulomb.py:1186-1192``);
* ``_K03_jit`` for the Coulomb tables versus an accurate K0 for Chi1D;
* lower-triangular unscreened tables: negative ``kkp`` entries are dropped
  (``coulomb.py:567-573``);
* the one-index shift of the stored ``Ceh/Cee/Chh`` tables as read by
  ``_MBCE_core``/``_MBCH_core`` (``coulomb.py:956-1071`` reads
  ``Ceh[k-1, kp-1, k1-1]`` from a table filled at 1-based indices);
* the ``g = [gam_eh, gam_e, gam_h]`` order for off-diagonal dephasing
  (``SBEs.py:5205``, ``dephasing.py:726-727``);
* ``CalcGammaE/H`` replacing the constant rates (``SBEs.py:5216-5219``);
* near-tie ``rint`` in ``Vxx2`` (``dephasing.py:362-378``), reproduced by
  computing ``kr`` and the ratios in the reference's operation order.

Pinning (see tests/test_oracle_*.py): every function is checked against
the golden vectors written from the reference (``tests/golden/*.npz``), and
further against closed forms / library routines where the mathematics fixes
them (scipy.special.k0, numpy.fft, brute-force loops at tiny N, exactness
of the cubic spline on cubics, conservation identities).
"""

import math

import numpy as np
from scipy.constants import c as c0
from scipy.constants import e as e0
from scipy.constants import epsilon_0 as eps0
from scipy.constants import hbar
from scipy.constants import k as kB
from scipy.constants import m_e as me0

pi = np.pi
twopi = 2.0 * np.pi
eV = 1.602176634e-19
EULER_GAMMA = 0.5772156649015329
TEMP = 77.0  # phonons.py:22, emission.py:29
SMALL = 1e-200  # SBEs.py:4703


# ---------------------------------------------------------------------------
# Parameters (SBEs.py:4978-5022; params/qw.params, params/mb.params)
# ---------------------------------------------------------------------------

FLAG_NAMES = ("Optics", "Excitons", "EHs", "Screened", "Phonon", "DCTrans",
              "LF", "FreePot", "DiagDph", "OffDiagDph", "Recomb", "PLSpec",
              "ignorewire", "Xqwparams", "LorentzDelta")
SHIPPED_FLAGS = (1, 1, 1, 1, 1, 1, 1, 0, 1, 1, 1, 0, 0, 1, 0)


class Params:
    """Material/numerics parameters as read by ``_read_qw_params``."""

    def __init__(self, L=100e-9, Delta0=5e-9, gap_eV=1.5, me_rel=0.07,
                 mh_rel=0.45, HO_eV=100e-3, gam_e=1e12, gam_h=1e12,
                 gam_eh=1e12, epsr=9.1, Oph_eV=36e-3, Gph_eV=3e-3, Edc=0.0,
                 jmax=500, ntmax=20000, flags=SHIPPED_FLAGS):
        self.L = float(L)
        self.Delta0 = float(Delta0)
        self.gap = gap_eV * e0
        self.me = me_rel * me0
        self.mh = mh_rel * me0
        self.HO = HO_eV * e0
        self.gam_e = float(gam_e)
        self.gam_h = float(gam_h)
        self.gam_eh = float(gam_eh)
        self.epsr = float(epsr)
        self.Oph = Oph_eV * e0 / hbar
        self.Gph = Gph_eV * e0 / hbar
        self.Edc = float(Edc)
        self.jmax = int(jmax)
        self.ntmax = int(ntmax)
        for name, v in zip(FLAG_NAMES, flags):
            setattr(self, name, bool(int(v)))


# ---------------------------------------------------------------------------
# Grid and derived constants (SBEs.py:4754-4823, 5028-5051; typespace.py)
# ---------------------------------------------------------------------------

def get_space_array(n, length):
    """typespace.py:815-849."""
    if n == 1:
        return np.array([0.0])
    return np.linspace(-length / 2.0, length / 2.0, n)


def get_k_array(n, length):
    """typespace.py:852-886 (centred, NOT FFT order)."""
    if n == 1:
        return np.array([0.0])
    dk = twopi / length if length > 0 else 1.0
    k = np.arange(n, dtype=float) * dk
    return k - k[n // 2]


def get_array0_index(x):
    """usefulsubs.py:1012-1042."""
    dx = x[2] - x[1]
    xi = np.round((x + 0.1 * dx) / dx).astype(np.int32)
    for i in range(len(x)):
        if xi[i] == 0:
            return i
    raise ValueError("Zero index not found in array")


class Grid:
    """Everything ``SBESolver.__init__`` derives before the module tables."""

    def __init__(self, prm):
        self.prm = prm
        kmax = np.sqrt(1.2 * prm.gap * 2.0 * prm.me / (hbar**2))
        self.dkr = twopi / (2.0 * prm.L)
        nk = int(np.floor(kmax / self.dkr) * 2 + 1)
        self.Nr = nk * 2
        self.Nk = nk - 1
        N = self.Nk
        dnk = (N - 1) // 2
        self.NK0 = dnk + 1
        self.kr = np.zeros(N)
        for k in range(N):
            self.kr[k] = self.dkr * (-dnk + k - 0.5)
        self.r = get_space_array(self.Nr, 2 * prm.L)
        qx = get_k_array(self.Nr, 2 * prm.L)
        qx = np.roll(qx, self.Nr // 2)
        qx[0] = -qx[0]
        self.Qr = qx
        self.NQ0 = get_array0_index(qx)
        self.dr = (self.r[2] - self.r[1]) * (self.Nr - 1) / float(self.Nr)
        self.kkp = np.zeros((N, N), dtype=np.int64)
        for k in range(N):
            for kp in range(N):
                q = self.kr[k] - self.kr[kp]
                self.kkp[k, kp] = int(np.round(q / self.dkr)) + self.NQ0
        me, mh = prm.me, prm.mh
        self.dcv = np.sqrt((e0 * hbar) ** 2 / (6.0 * me0 * prm.gap)
                           * (me0 / me - 1.0))
        self.alphae = np.sqrt(me * prm.HO) / hbar
        self.alphah = np.sqrt(mh * prm.HO) / hbar
        ae, ah = self.alphae, self.alphah
        self.ehint = np.sqrt(2.0 * ae * ah / (ae**2 + ah**2))
        self.gam_eh = (prm.gam_e + prm.gam_h) / 2.0
        self.qc = 2.0 * ae * ah / (ae + ah)
        self.area = np.sqrt(2.0 * pi) / np.sqrt(ae**2 + ah**2) * prm.Delta0
        self.Ee = hbar**2 * self.kr**2 / (2.0 * me)
        self.Eh = hbar**2 * self.kr**2 / (2.0 * mh)


def fermi_distr(En):
    """phonons.py:317-340."""
    return 1.0 / (np.exp(En / kB / TEMP) + 1.0)


# ---------------------------------------------------------------------------
# Bessel K0: the crude _K03_jit (coulomb.py:38-66) and an accurate K0
# restating the Zhang-Jin series/asymptotics behind usefulsubs.K03
# (usefulsubs.py:1649-1800), real positive argument only.
# ---------------------------------------------------------------------------

def k03_jit(x):
    """Vectorised ``_K03_jit`` (coulomb.py:38-66), including its
    non-standard series (term*(1/k - gamma)) and its constants."""
    x = np.asarray(x, dtype=float)
    out = np.zeros_like(x)
    big = x > 100.0
    tiny = (~big) & (x < 1e-10)
    ser = (~big) & (~tiny) & (x < 2.0)
    asy = (~big) & (~tiny) & (x >= 2.0)
    out[tiny] = 23.0
    xs = x[ser]
    x2 = xs * xs
    res = -np.log(0.5 * xs) - EULER_GAMMA
    term = np.ones_like(xs)
    active = np.ones(xs.shape, dtype=bool)
    for k in range(1, 20):
        nt = term * 0.25 * x2 / (k * k)
        term = np.where(active, nt, term)
        res = np.where(active, res + term * (1.0 / k - EULER_GAMMA), res)
        active &= ~(np.abs(term) < 1e-15)
    out[ser] = res
    xa = x[asy]
    out[asy] = np.sqrt(np.pi / (2.0 * xa)) * np.exp(-xa)
    return out


def _asym_i0_coeffs(n):
    # I0(z) ~ e^z/sqrt(2 pi z) sum_k ((2k-1)!!)^2 / (k! 8^k z^k)
    c = []
    num = 1.0
    for k in range(1, n + 1):
        num *= (2 * k - 1) ** 2
        c.append(num / (math.factorial(k) * 8.0**k))
    return np.array(c)


def _asym_i0k0_coeffs(n):
    # I0(z) K0(z) ~ 1/(2z) (1 + sum_k c_k z^(-2k)) with (nu = 0 in the
    # product expansion of I_nu K_nu)
    # c_k = (2k-1)!!/(2k)!! * ((2k-1)!!)^2 / 4^k, i.e. the recurrence
    # c_k = c_{k-1} (2k-1)^3 / (8k).
    c = []
    v = 1.0
    for k in range(1, n + 1):
        v *= (2 * k - 1) ** 3 / (8.0 * k)
        c.append(v)
    return np.array(c)


_A_I0 = _asym_i0_coeffs(12)
_A_I0K0 = _asym_i0k0_coeffs(10)


def k0_accurate(x):
    """Accurate K0(x), x > 0: power series for x <= 9, otherwise the
    asymptotic expansion of I0*K0 divided by I0 (Zhang & Jin, Computation
    of Special Functions, 1996, routine CIK01), as used by usefulsubs.K03;
    zero above 100 as K03 returns."""
    x = np.atleast_1d(np.asarray(x, dtype=float))
    out = np.zeros_like(x)
    for idx, xv in np.ndenumerate(x):
        if xv > 100.0:
            out[idx] = 0.0
            continue
        z2 = xv * xv
        if xv <= 18.0:
            i0 = 1.0
            cr = 1.0
            for k in range(1, 51):
                cr = 0.25 * cr * z2 / (k * k)
                i0 += cr
                if abs(cr / i0) < 1e-15:
                    break
        else:
            k0n = 12 if xv < 35.0 else (9 if xv < 50.0 else 7)
            s = 1.0
            for k in range(1, k0n + 1):
                s += _A_I0[k - 1] * xv ** (-k)
            i0 = np.exp(xv) / np.sqrt(2.0 * pi * xv) * s
        if xv <= 9.0:
            cs = 0.0
            ct = -np.log(0.5 * xv) - EULER_GAMMA
            w0 = 0.0
            cr = 1.0
            cw = 0.0
            for k in range(1, 51):
                w0 += 1.0 / k
                cr = 0.25 * cr / (k * k) * z2
                cs += cr * (w0 + ct)
                if abs((cs - cw) / cs) < 1e-15:
                    break
                cw = cs
            out[idx] = ct + cs
        else:
            s = 1.0
            for k in range(1, 11):
                s += _A_I0K0[k - 1] * z2 ** (-k)
            out[idx] = 0.5 / xv * s / i0
    return out


# ---------------------------------------------------------------------------
# Init tables
# ---------------------------------------------------------------------------

def vint(Qyk, y, ae, ah, Delta0):
    """coulomb.py:69-104 / 107-177 (JIT path, _K03_jit)."""
    Ny = len(y)
    N1 = Ny // 4
    N2 = 3 * Ny // 4
    aey2 = (ae * y) ** 2
    ahy2 = (ah * y) ** 2
    kmin = (ae + ah) / 4.0
    dk = max(abs(Qyk), kmin)
    mult = ae * ah / np.pi * (y[1] - y[0]) ** 2
    yi = y[N1:N2 + 1]
    rr = np.sqrt((yi[:, None] - yi[None, :]) ** 2 + Delta0**2)
    ex = np.exp(-aey2[N1:N2 + 1][:, None] - ahy2[N1:N2 + 1][None, :])
    return float(np.sum(ex * mult * k03_jit(dk * rr)))


class Tables:
    """Module-level tables of coulomb/phonons/dephasing/emission/qwoptics."""

    def __init__(self, g, vtab=None):
        """``vtab`` (3, Nr) may supply precomputed unscreened Coulomb
        integrals (only the CPU-baseline timing uses it, to skip the slow
        table build that is not part of a step)."""
        prm = g.prm
        N, Nr = g.Nk, g.Nr
        ky = g.kr
        ae, ah = g.alphae, g.alphah
        # --- coulomb.py:502-577 CalcCoulombArrays
        pref = e0**2 / (twopi * eps0 * prm.epsr * prm.L)
        NQ = len(g.Qr)
        if vtab is None:
            eh = np.array([pref * vint(g.Qr[k], g.r, ae, ah, prm.Delta0)
                           for k in range(NQ)])
            ee = np.array([pref * vint(g.Qr[k], g.r, ae, ae, prm.Delta0)
                           for k in range(NQ)])
            hh = np.array([pref * vint(g.Qr[k], g.r, ah, ah, prm.Delta0)
                           for k in range(NQ)])
            vtab = np.stack([eh, ee, hh])
        self.vtab = np.asarray(vtab)  # (3, Nr), Qr order
        self.V0 = np.zeros((3, N, N))
        for k in range(N):
            for q in range(N):
                i = g.kkp[k, q]
                if 0 <= i < NQ:
                    self.V0[:, k, q] = self.vtab[:, i]
        # --- coulomb.py:302-333 MakeQs, 580-630 CalcChi1D
        d = np.abs(ky[None, :] - ky[:, None])  # [k1,k2] = |ky[k2]-ky[k1]|
        self.qe = np.maximum(d, ae / 2.0)
        self.qh = np.maximum(d, ah / 2.0)
        Re = np.sqrt((2.0 / ae) ** 2 + prm.Delta0**2)
        Rh = np.sqrt((2.0 / ah) ** 2 + prm.Delta0**2)
        scale = e0**2 / (twopi * eps0 * prm.epsr * hbar**2)
        self.Chi1De = prm.me * k0_accurate(self.qe * Re) / self.qe * scale
        self.Chi1Dh = prm.mh * k0_accurate(self.qh * Rh) / self.qh * scale
        # --- phonons.py:55-136
        self.NO = 1.0 / (np.exp(hbar * prm.Oph / kB / TEMP) - 1.0)
        hO, hG = hbar * prm.Oph, hbar * prm.Gph
        idel = 1.0 - np.eye(N)
        Ee, Eh = g.Ee, g.Eh
        dEe = Ee[:, None] - Ee[None, :]
        dEh = Eh[:, None] - Eh[None, :]
        EP = (self.NO / ((dEe - hO) ** 2 + hG**2)
              + (self.NO + 1.0) / ((dEe + hO) ** 2 + hG**2))
        HP = (self.NO / ((dEh - hO) ** 2 + hG**2)
              + (self.NO + 1.0) / ((dEh + hO) ** 2 + hG**2))
        self.EP = EP * 2.0 * prm.Gph * idel
        self.HP = HP * 2.0 * prm.Gph * idel
        self.Vscale = hbar * prm.Oph * prm.epsr * (1.0 / 8.2 - 1.0 / 10.0)
        # --- dephasing.py:472-539
        dk = ky[1] - ky[0]
        kmaxd = ky[N - 1] + dk
        kmind = ky[0] - dk
        Nk0 = (N - 1) // 2 + 1

        def idx(v):
            return int(np.round(np.clip(v, kmind, kmaxd) / dk)) + Nk0 - 1

        self.xe = prm.me / hbar**2 * np.abs(ky) / (np.abs(ky) + dk * 1e-5) ** 2 / dk
        self.xh = prm.mh / hbar**2 * np.abs(ky) / (np.abs(ky) + dk * 1e-5) ** 2 / dk
        me, mh = prm.me, prm.mh
        z = lambda: np.zeros((N, N), dtype=np.int64)  # noqa: E731
        self.k_p_q, self.k_m_q = z(), z()
        self.k1p, self.k1p_m_q, self.k1, self.k1_m_q = z(), z(), z(), z()
        for q in range(N):
            for k in range(N):
                self.k_p_q[k, q] = idx(ky[k] + ky[q])
                self.k_m_q[k, q] = idx(ky[k] - ky[q])
                k1p0 = ((me + mh) * ky[q] - 2 * mh * ky[k]) / 2.0 / me
                self.k1p[k, q] = idx(k1p0)
                self.k1p_m_q[k, q] = idx(k1p0 - ky[q])
                k10 = ((me + mh) * ky[q] - 2 * me * ky[k]) / 2.0 / mh
                self.k1[k, q] = idx(k10)
                self.k1_m_q[k, q] = idx(k10 - ky[q])
        # Vxx2 index (dephasing.py:362-378)
        self.iq = np.round(np.abs(ky / dk)).astype(np.int64)
        # --- emission.py:35-91, 220-255
        kBT = kB * TEMP
        hg = hbar * g.gam_eh
        hwmax = (kBT + hg) * 4.0
        dhw = min(kBT, hg) / 20.0
        Nw = int(np.ceil(hwmax / dhw))
        self.HOmega = (np.arange(Nw) + 0.5) * dhw
        dcv = abs(g.dcv)
        self.square = (3.0 * dcv**2 * g.ehint / eps0 / np.sqrt(prm.epsr)
                       * g.ehint / hbar * lrtz(self.HOmega, hg)
                       * np.exp(-self.HOmega / kB / TEMP))
        # --- qwoptics.py:575-663
        self.window = np.exp(-((g.r / (prm.L / 2.0)) ** 150))
        self.Expikr = np.exp(1j * np.outer(ky, g.r))
        kh = np.arange(N)
        self.Xvc0 = np.tile((g.dcv * (-1.0) ** kh)[:, None], (1, N)).astype(complex)
        self.Yvc0 = np.full((N, N), g.dcv, dtype=complex)
        self.Zvc0 = np.full((N, N), -g.dcv, dtype=complex)
        # --- dcfield.py:767-796 (Transport's phase ramp)
        dky = ky[2] - ky[1]
        self.Ytr = get_k_array(N, (N - 1) * dky)


def lrtz(a, b):
    """usefulsubs.py:1256-1279."""
    return (b / pi) / (a**2 + b**2)


def gauss_delta(a, b):
    """coulomb.py:236-268."""
    return 1.0 / (np.sqrt(pi) * b) * np.exp(-((a / b) ** 2))


# ---------------------------------------------------------------------------
# Per-step physics
# ---------------------------------------------------------------------------

def screened_arrays(g, t, screened, ne, nh):
    """coulomb.py:1172-1230: returns VC (N,N,3) and E1D (N,N)."""
    N = g.Nk
    VC = np.zeros((N, N, 3))
    VC[:, :, 0] = t.V0[0]
    VC[:, :, 1] = t.V0[1]
    VC[:, :, 2] = t.V0[2]
    E1D = np.ones((N, N))
    if screened:
        dens = np.sum(np.real(ne) + np.real(nh)) / 2.0 / g.prm.L
        dmax = min(t.qe[1, 1], t.qh[1, 1]) / pi * 0.99
        n1 = min(dens, dmax)
        E1D = (np.ones_like(t.Chi1De)
               - t.Chi1De * 2 * np.log(np.abs((t.qe - pi * n1) / (t.qe + n1)))
               - t.Chi1Dh * 2 * np.log(np.abs((t.qh - pi * n1) / (t.qh + n1))))
        for i in range(3):
            VC[:, :, i] = VC[:, :, i] / E1D
    return VC, E1D


def yw(w):
    """qwoptics.py:1093-1106."""
    return (-1) ** int(np.floor((w - 1) / 2.0))


def calc_meh(g, t, Ex, Ey, Ez, w):
    """SBEs.py:5153-5163 (negative kkp wraps Python-style)."""
    q = g.kkp  # [ke, kh]
    X = t.Xvc0.conj().T  # Xcv0[ke,kh]
    Y = t.Yvc0.conj().T
    Z = t.Zvc0.conj().T
    return (-g.ehint * X * Ex[q] - g.ehint * Y * Ey[q] * yw(w)
            - g.ehint * Z * Ez[q] * (-1) ** w)


def calc_h(g, Meh, C, D, p, VC, excitons, lf):
    """SBEs.py:808-924 (_CalcH_jit), Wee = Whh = 0 (FreePot never adds)."""
    N = g.Nk
    V = np.zeros((2 * N + 1, 3))
    for q in range(1, N):
        V[N + q] = VC[q, 0, :]
        V[N - q] = VC[q, 0, :]
    Heh = Meh.astype(complex).copy()
    Hee = np.zeros((N, N), dtype=complex)
    Hhh = np.zeros((N, N), dtype=complex)
    for k in range(N):
        Hee[k, k] = g.Ee[k] + g.prm.gap
        Hhh[k, k] = g.Eh[k]
    pt, Ct, Dt = p.T, C.T, D.T
    if not excitons:
        return Heh, Hee, Hhh
    k1 = np.arange(N)[:, None]
    k2 = np.arange(N)[None, :]
    # exchange sums over q != 0 with both k1+q and k2+q in range
    for q in range(-(N - 1), N):
        if q == 0:
            continue
        lo = max(0, -q)
        hi = min(N, N - q)
        sl = slice(lo, hi)
        slq = slice(lo + q, hi + q)
        Heh[sl, sl] += V[N + q, 0] * pt[slq, slq]
        if lf:
            Hee[sl, sl] -= V[N + q, 1] * Ct[slq, slq]
            Hhh[sl, sl] -= V[N + q, 2] * Dt[slq, slq]
        else:
            idx = np.arange(lo, hi)
            Hee[idx, idx] -= V[N + q, 1] * Ct[idx + q, idx + q]
            Hhh[idx, idx] -= V[N + q, 2] * Dt[idx + q, idx + q]
    if lf:
        # direct terms: q_d = k1 - k2 != 0
        qd = (k1 - k2)
        dC = np.zeros(2 * N + 1, dtype=complex)  # sum_k C[k, k+s], s in [-N,N]
        dD = np.zeros(2 * N + 1, dtype=complex)  # sum_k D[k+s, k]
        for s in range(-(N - 1), N):
            dC[N + s] = np.trace(C, offset=s)
            dD[N + s] = np.trace(D, offset=-s)
        m = qd != 0
        Hee[m] += (V[N + qd, 1] * dC[N + qd] - V[N + qd, 0] * dD[N + qd])[m]
        Hhh[m] += (V[N + qd, 2] * np.array([np.trace(D, offset=s) for s in range(-N, N + 1)])[N + qd]
                   - V[N + qd, 0] * np.array([np.trace(C, offset=-s) for s in range(-N, N + 1)])[N + qd])[m]
    return Heh, Hee, Hhh


def dpdt(C, D, p, Heh, Hee, Hhh, GamE, GamH, OffP):
    """SBEs.py:1941-1965 (_dpdt_jit); p is p[kh, ke], Heh is Heh[ke, kh]."""
    t1 = Hhh @ p
    t2 = p @ Hee.T
    t3 = Heh.T @ C
    t4 = (Heh @ D).T
    r = (t1 + t2 - t3 - t4 + Heh.T
         - 1j * hbar * (GamE[None, :] + GamH[:, None]) * p + OffP)
    return r / (1j * hbar)


def dCdt(C, D, p, Heh, Hee, Hhh):
    """SBEs.py:2308-2327 (_dCdt_jit)."""
    Peh = np.conj(p.T)
    Hhe = np.conj(Heh.T)
    r = C @ Hee.T - Hee.T @ C + Peh @ Heh.T - Hhe.T @ p
    return r / (1j * hbar)


def dDdt(C, D, p, Heh, Hee, Hhh):
    """SBEs.py:2606-2625 (_dDdt_jit)."""
    Peh = np.conj(p.T)
    Hhe = np.conj(Heh.T)
    r = D @ Hhh.T - Hhh.T @ D + Peh.T @ Heh - Hhe @ p.T
    return r / (1j * hbar)


def vxx2(t, Vmat):
    """dephasing.py:362-378: V[rint(|ky/dq|), 0]**2."""
    return Vmat[t.iq, 0] ** 2


def gamma_e(g, t, ne0, nh0, VC):
    """dephasing.py:28-64, 541-582."""
    N = g.Nk
    Veh2 = vxx2(t, VC[:, :, 0])
    Vee2 = vxx2(t, VC[:, :, 1])
    ne = np.zeros(N + 2)
    nh = np.zeros(N + 2)
    se = np.zeros(N + 2)
    sh = np.zeros(N + 2)
    ne[1:N + 1] = np.real(ne0)
    nh[1:N + 1] = np.real(nh0)
    se[1:N + 1] = 1.0 - ne[1:N + 1]
    sh[1:N + 1] = 1.0 - nh[1:N + 1]
    G = np.zeros(N)
    for q in range(N):
        for k in range(N):
            i = t.k_p_q[k, q] + 1
            if 1 <= i < N + 1:
                G[k] += pi / hbar * Vee2[q] * ne[i] * se[i] * np.abs(t.xe[q])
    for q in range(N):
        for k in range(N):
            a = t.k1p_m_q[k, q] + 1
            b = t.k1p[k, q] + 1
            c = t.k_m_q[k, q] + 1
            if 1 <= a < N + 1 and 1 <= b < N + 1 and 1 <= c < N + 1:
                G[k] += pi / hbar * Veh2[q] * (
                    nh[a] * sh[b] * ne[c] + sh[a] * nh[b] * se[c]) * np.abs(t.xh[q])
    return G


def gamma_h(g, t, ne0, nh0, VC):
    """dephasing.py:67-103, 584-625."""
    N = g.Nk
    Veh2 = vxx2(t, VC[:, :, 0])
    Vhh2 = vxx2(t, VC[:, :, 2])
    ne = np.zeros(N + 2)
    nh = np.zeros(N + 2)
    se = np.zeros(N + 2)
    sh = np.zeros(N + 2)
    ne[1:N + 1] = np.real(ne0)
    nh[1:N + 1] = np.real(nh0)
    se[1:N + 1] = 1.0 - ne[1:N + 1]
    sh[1:N + 1] = 1.0 - nh[1:N + 1]
    G = np.zeros(N)
    for q in range(N):
        for kp in range(N):
            i = t.k_p_q[kp, q] + 1
            if 1 <= i < N + 1:
                G[kp] += pi / hbar * Vhh2[q] * nh[i] * sh[i] * np.abs(t.xh[q])
    for q in range(N):
        for kp in range(N):
            a = t.k1_m_q[kp, q] + 1
            b = t.k1[kp, q] + 1
            c = t.k_m_q[kp, q] + 1
            if 1 <= a < N + 1 and 1 <= b < N + 1 and 1 <= c < N + 1:
                G[kp] += pi / hbar * Veh2[q] * (
                    ne[a] * se[b] * nh[c] + se[a] * ne[b] * sh[c]) * np.abs(t.xe[q])
    return G


def _vsq(VC, N):
    Vsq = np.zeros((2 * N + 1, 3))
    for q in range(1, N):
        Vsq[q + N] = VC[q, 0, :] ** 2
        Vsq[-q + N] = VC[q, 0, :] ** 2
    return Vsq


def _offdiag_deph2(g, Ea, Eb, na, nb, Vsq, ch, gaa, gab):
    """The two sums of CalcOffDiagDeph_E2/H2 for carrier a (na, Ea) against
    itself (channel ch, broadening gaa) and against the other carrier b
    (channel 0, broadening gab): for every q, all k with k + q in range at
    once (a row per k), the p-sums over the reference's index sets; each row
    is reduced on its own, in the order the loop over k would."""
    N = g.Nk
    D = np.zeros((2 * N + 1, N))
    P = np.arange(N)
    for q in range(-N, N + 1):
        ks = P[(P + q >= 0) & (P + q < N)]
        if not ks.size:
            continue
        kc = ks[:, None]
        # loop 1: p - q in [1, N-1]
        pm = P[(P - q >= 1) & (P - q < N)]
        if pm.size:
            Ed = Ea[kc + q] + Ea[pm - q][None, :] - Ea[pm][None, :] - Ea[kc]
            D[q + N, ks] += np.sum(
                Vsq[q + N, ch] * lrtz(Ed, hbar * gaa)
                * (na[pm - q][None, :] * (1.0 - na[pm])[None, :] * (1.0 - na[kc])
                   + (1.0 - na[pm - q])[None, :] * na[pm][None, :] * na[kc]), axis=1)
        # loop 2: p + q in [0, N-1]
        pp = P[(P + q >= 0) & (P + q < N)]
        if pp.size:
            Ed = Ea[kc + q] + Eb[pp + q][None, :] - Eb[pp][None, :] - Ea[kc]
            D[q + N, ks] += np.sum(
                Vsq[q + N, 0] * lrtz(Ed, hbar * gab)
                * (nb[pp + q][None, :] * (1.0 - nb[pp])[None, :] * (1.0 - na[kc])
                   + (1.0 - nb[pp + q])[None, :] * nb[pp][None, :] * na[kc]), axis=1)
    return D * pi / hbar


def offdiag_deph_e2(g, ne, nh, VC, gee, geh):
    """dephasing.py:226-270, 381-397 (vectorised over k and p, same index
    sets and per-element operation order)."""
    return _offdiag_deph2(g, g.Ee, g.Eh, np.abs(ne), np.abs(nh), _vsq(VC, g.Nk), 1, gee, geh)


def offdiag_deph_h2(g, ne, nh, VC, ghh, geh):
    """dephasing.py:273-317, 400-414."""
    return _offdiag_deph2(g, g.Eh, g.Ee, np.abs(nh), np.abs(ne), _vsq(VC, g.Nk), 2, ghh, geh)


def offdiag_dephasing2(g, ne, nh, p, VC):
    """dephasing.py:320-356, 714-743; g = [gam_eh, gam_e, gam_h]."""
    N = g.Nk
    gg = (g.gam_eh, g.prm.gam_e, g.prm.gam_h)
    De = offdiag_deph_e2(g, ne, nh, VC, gg[0], gg[2])
    Dh = offdiag_deph_h2(g, ne, nh, VC, gg[1], gg[2])
    kh = np.arange(N)[:, None]
    m = np.arange(N)[None, :]
    A = Dh[m - kh + N, kh.repeat(N, 1)]  # A[kh, m] = Dh[m-kh+N, kh]
    A = np.where(m == kh, 0.0, A)
    ke = np.arange(N)[None, :]
    mm = np.arange(N)[:, None]
    B = De[mm - ke + N, ke.repeat(N, 0)]  # B[m, ke] = De[m-ke+N, ke]
    B = np.where(mm == ke, 0.0, B)
    x = A @ p + p @ B
    return x * 1j * hbar, De, Dh


def ec(ne, nh, VC):
    """emission.py:133-171."""
    N = len(ne)
    idel = 1.0 - np.eye(N)
    Veh, Vee, Vhh = VC[:, :, 0], VC[:, :, 1], VC[:, :, 2]
    out = np.zeros(N)
    for k in range(N):
        out[k] = (np.sum(ne * (Vee[k, k] - Vee[:, k]) + nh * (Vhh[k, k] - Vhh[:, k]))
                  + np.sum(ne * Veh[k, k] * idel[:, k] - nh * Veh[k, k] * idel[:, k])
                  - Veh[k, k])
    return out


def rho0(hw):
    """emission.py:201-217."""
    return hw**2 / (c0**3 * pi**2 * hbar**3)


def spont_emission(g, t, ne, nh, VC):
    """emission.py:94-130, 174-198."""
    Ek = g.prm.gap + g.Ee + g.Eh + ec(np.real(ne), np.real(nh), VC)
    dhw = t.HOmega[1] - t.HOmega[0]
    H = t.HOmega[None, :] + Ek[:, None]
    return np.sum(H * rho0(H) * t.square[None, :], axis=1) * dhw


def mb_tables(g, kind, ge, gh, lorentz=False):
    """Stored (N+1)^3 tables of CalcMBArrays (coulomb.py:367-499; Gaussian
    branch 448-497, ``lorentz``: the LorentzDelta branch 405-446), index 0
    planes zero, 1-based storage."""
    N = g.Nk
    Ee, Eh = g.Ee, g.Eh
    geh = (ge + gh) / 2.0
    S = np.zeros((N + 1, N + 1, N + 1))
    a = np.arange(N)[:, None, None]
    b = np.arange(N)[None, :, None]
    c = np.arange(N)[None, None, :]
    if kind == "eh":
        k30 = c + b - a
        ok = (k30 >= 0) & (k30 < N)
        kk = np.clip(k30, 0, N - 1)
        Ed = Ee[a] + Eh[b] - Eh[kk] - Ee[c]
        if lorentz:
            val = 2.0 * geh * (a != c) * (b != k30) / (Ed**2 + (hbar * geh) ** 2)
        else:
            val = twopi / hbar * (a != c) * (b != k30) * gauss_delta(Ed, hbar * geh)
    else:
        E = Ee if kind == "ee" else Eh
        gam = ge if kind == "ee" else gh
        k30 = a + b - c
        ok = (k30 >= 0) & (k30 < N)
        kk = np.clip(k30, 0, N - 1)
        Ed = E[a] + E[b] - E[kk] - E[c]
        if lorentz:
            val = 2.0 * gam * (a != c) * (b != k30) / (Ed**2 + (hbar * gam) ** 2)
        else:
            val = twopi / hbar * (a != c) * (b != k30) * gauss_delta(Ed, hbar * gam)
    S[1:, 1:, 1:] = np.where(ok, val, 0.0)
    return S


# The many-body sums below run over the non-zero entries of each term only:
# the 3-D weight of a term (the stored table read at the loop indices, zero
# where the shifted index leaves [0, N)) does not change during a run, so its
# non-zero entries are found once per table and kept.  Each entry's product
# is formed in the reference's factor order; the per-k sums are taken in
# entry order (np.bincount) instead of numpy's pairwise order.
_MB_TERMS = {}


def _mb_term(S, key, build):
    """Entries (out, i1, i2, kk, w) of one term of table S, cached per table
    object (the table itself is held so that its id stays valid)."""
    hit = _MB_TERMS.get((id(S), key))
    if hit is None or hit[0] is not S:
        W, kk, out_axis = build()
        nz = np.nonzero(W)
        idx = [np.broadcast_to(a, W.shape)[nz].astype(np.int64) for a in np.ix_(*(np.arange(n) for n in W.shape))]
        hit = (S, (idx[out_axis], idx[1], idx[2], np.broadcast_to(kk, W.shape)[nz].astype(np.int64), W[nz]))
        _MB_TERMS[(id(S), key)] = hit
    return hit[1]


def _term_weight(S, sel, shifted, N):
    ok = (shifted >= 0) & (shifted < N)
    return np.where(ok, S[sel], 0.0), np.clip(shifted, 0, N - 1)


def mbce(g, S_eh, S_ee, ne0, nh0, VC):
    """coulomb.py:956-1012, 1298-1311 (reads the stored tables with the
    0-based loop indices, i.e. shifted by one)."""
    N = g.Nk
    ne = np.abs(ne0)
    nh = np.abs(nh0)
    Veh2 = VC[:, :, 0] ** 2
    Vee2 = VC[:, :, 1] ** 2
    K = np.arange(N)[:, None, None]
    P = np.arange(N)[None, :, None]
    Q = np.arange(N)[None, None, :]
    # term A: k1p = P+Q-K
    k, p, q, kk, w = _mb_term(S_eh, "eA", lambda: _term_weight(S_eh, (K, P, Q), P + Q - K, N) + (0,))
    Win = np.bincount(k, Veh2[k, q] * (1.0 - nh[p]) * nh[kk] * ne[q] * w, minlength=N)
    # term B: k1p = P+K-Q
    k, p, q, kk, w = _mb_term(S_eh, "eB", lambda: _term_weight(S_eh, (Q, P, K), P + K - Q, N) + (0,))
    Wout = np.bincount(k, Veh2[q, k] * (1.0 - ne[q]) * (1.0 - nh[p]) * nh[kk] * w, minlength=N)
    # term C (P plays k2, Q plays k4): k30 = K+P-Q
    k, p, q, kk, w = _mb_term(S_ee, "eC", lambda: _term_weight(S_ee, (K, P, Q), K + P - Q, N) + (0,))
    Win += np.bincount(k, Vee2[k, q] * (1.0 - ne[p]) * ne[kk] * ne[q] * w, minlength=N)
    # term D: k30 = Q+P-K
    k, p, q, kk, w = _mb_term(S_ee, "eD", lambda: _term_weight(S_ee, (Q, P, K), Q + P - K, N) + (0,))
    Wout += np.bincount(k, Vee2[q, k] * (1.0 - ne[q]) * (1.0 - ne[p]) * ne[kk] * w, minlength=N)
    return Win, Wout


def mbch(g, S_eh, S_hh, ne0, nh0, VC):
    """coulomb.py:1015-1071, 1313-1326 (sums over axes 1 and 2 into the
    axis-0 index KP)."""
    N = g.Nk
    ne = np.abs(ne0)
    nh = np.abs(nh0)
    Veh2 = VC[:, :, 0] ** 2
    Vhh2 = VC[:, :, 2] ** 2
    KP = np.arange(N)[:, None, None]
    K = np.arange(N)[None, :, None]
    Q = np.arange(N)[None, None, :]
    # k1p = KP+Q-K
    o, k, q, kk, w = _mb_term(S_eh, "hA", lambda: _term_weight(S_eh, (K, KP, Q), KP + Q - K, N) + (0,))
    Win = np.bincount(o, Veh2[k, q] * (1.0 - ne[k]) * nh[kk] * ne[q] * w, minlength=N)
    # k1p = KP+K-Q, the table read at (K, k1p, Q)
    def build_b():
        k1p = KP + K - Q
        ok = (k1p >= 0) & (k1p < N)
        kk = np.clip(k1p, 0, N - 1)
        return np.where(ok, S_eh[K, kk, Q], 0.0), kk, 0
    o, k, q, kk, w = _mb_term(S_eh, "hB", build_b)
    Wout = np.bincount(o, Veh2[q, k] * (1.0 - ne[k]) * (1.0 - nh[kk]) * ne[q] * w, minlength=N)
    # B2 = K, F = Q: k3p = KP+B2-F
    o, b2, f, kk, w = _mb_term(S_hh, "hC", lambda: _term_weight(S_hh, (KP, K, Q), KP + K - Q, N) + (0,))
    Win += np.bincount(o, Vhh2[o, f] * (1.0 - nh[b2]) * nh[kk] * nh[f] * w, minlength=N)
    # k3p = F+B2-KP
    o, b2, f, kk, w = _mb_term(S_hh, "hD", lambda: _term_weight(S_hh, (Q, K, KP), Q + K - KP, N) + (0,))
    Wout += np.bincount(o, Vhh2[f, o] * (1.0 - nh[f]) * (1.0 - nh[b2]) * nh[kk] * w, minlength=N)
    return Win, Wout


def mbpe(t, ne, VC, E1D):
    """phonons.py:139-202."""
    Vep = VC[:, :, 1] / E1D * t.Vscale
    Win = np.sum(Vep * ne[None, :] * t.EP, axis=1)        # EPT[i,k] = EP[k,i]
    Wout = np.sum(Vep * (1.0 - ne[None, :]) * t.EP.T, axis=1)
    return Win, Wout


def mbph(t, nh, VC, E1D):
    """phonons.py:205-268."""
    Vhp = VC[:, :, 2] / E1D * t.Vscale
    Win = np.sum(Vhp * nh[:, None] * t.HP.T, axis=0)      # HPT[i,kp] = HP[kp,i]
    Wout = np.sum(Vhp * (1.0 - nh[:, None]) * t.HP, axis=0)
    return Win, Wout


def transport(t, X, Edc, Eac, dt, lf=True):
    """dcfield.py:1176-1236: ShiftN2D of the whole matrix (k1nek2 = LF
    True, 1188-1199), else ShiftN1D of its diagonal only (1154-1174,
    1229-1232), the off-diagonal entries kept."""
    dk = -e0 * (Edc + Eac) / hbar * dt
    Y = t.Ytr
    if not lf:
        out = np.array(X, dtype=complex, copy=True)
        nk = np.fft.ifft(np.fft.fft(np.diag(X)) * np.exp(-1j * Y * dk))
        out[np.diag_indices(len(nk))] = nk
        return out
    F = np.fft.fft2(X)
    F = F * np.exp(-1j * (Y[:, None] + Y[None, :]) * dk)
    return np.fft.ifft2(F)


def fftg(F):
    """usefulsubs.py:2245-2268."""
    return -np.fft.fft(F) / (len(F) * 1.0)


def ifftg(F):
    """usefulsubs.py:2271-2297."""
    return -np.fft.ifft(F) * len(F)


def qw_polarization3(g, t, p, w):
    """qwoptics.py:49-76, 809-940: returns FFTG'd Px, Py, Pz."""
    E = t.Expikr
    Ec = np.conj(E)
    W = t.window
    # prod[r] = sum_{ke,kh} p[kh,ke] * Xvc0[kh,ke] * E[ke,r] * Ec[kh,r]
    Px = np.real(np.einsum("hk,kr,hr->r", p * t.Xvc0, E, Ec)) * W
    Py = np.real(np.einsum("hk,kr,hr->r", p * t.Yvc0, E, Ec)) * W * yw(w)
    Pz = np.real(np.einsum("hk,kr,hr->r", p * t.Zvc0, E, Ec)) * W * ((-1) ** w)
    s = 2 * g.ehint / g.area / g.prm.L
    return fftg(Px * s + 0j), fftg(Py * s + 0j), fftg(Pz * s + 0j)


def qw_rho5(g, t, C, D, ne, nh):
    """qwoptics.py:360-374, 942-1071: returns FFTG'd re, rh."""
    E = t.Expikr
    Ec = np.conj(E)
    W = t.window
    L = g.prm.L
    Nr = g.Nr
    dr = g.r[1] - g.r[0]
    NeT = np.sum(np.abs(ne)) + 1e-50
    NhT = np.sum(np.abs(nh)) + 1e-50
    re = np.einsum("ab,ar,br->r", C, Ec, E) * W / (2.0 * L)
    rh = np.einsum("ab,ar,br->r", np.conj(D), Ec, E) * W / (2.0 * L)
    re = re - np.real(np.sum(re[0:10]) + np.sum(re[Nr - 10:Nr])) / 20.0
    rh = rh - np.real(np.sum(rh[0:10]) + np.sum(rh[Nr - 10:Nr])) / 20.0
    re = re * NeT / (np.sum(np.abs(re)) * dr + 1e-100)
    rh = rh * NhT / (np.sum(np.abs(rh)) * dr + 1e-100)
    return fftg(re), fftg(rh)


def ek_renorm(n, En, V):
    """dcfield.py:338-345."""
    return En + np.sum(n[None, :] * (np.diag(V)[:, None] - V), axis=1) / 2.0


def drift_vt(n, Ec, dkk):
    """dcfield.py:347-360."""
    N = len(n)
    dEdk = np.zeros(N, dtype=complex)
    dEdk[1:N - 1] = (Ec[2:] - Ec[:-2]) / 2.0 / dkk
    dEdk[0] = 2.0 * dEdk[1] - dEdk[2]
    dEdk[N - 1] = 2.0 * dEdk[N - 2] - dEdk[N - 3]
    v = np.sum(dEdk * n) / (1e-100 + np.sum(n)) / hbar
    return np.real(v)


def calc_i0(g, ne, nh, VC):
    """dcfield.py:1110-1152."""
    Ec = ek_renorm(np.real(ne), g.Ee, VC[:, :, 1])
    ve = drift_vt(np.real(ne), Ec, g.dkr)
    Ec = ek_renorm(np.real(nh), g.Eh, VC[:, :, 2])
    vh = drift_vt(np.real(nh), Ec, g.dkr)
    return -e0 * (ve + vh) * np.sum(ne) * g.dkr * 2.0


# ---------------------------------------------------------------------------
# FMM cubic spline (spliner.py:153-188, 1136-1171, 1287-1396, 1579-1636)
# ---------------------------------------------------------------------------

def spline_coeffs(x, y):
    """spline_dpc: coefficients b, c, d (columns of y handled together)."""
    y = np.asarray(y)
    n = len(x)
    shape = y.shape
    b = np.zeros(shape, dtype=y.dtype)
    c = np.zeros(shape, dtype=y.dtype)
    d = np.zeros(shape, dtype=y.dtype)
    dd = np.zeros(n)  # d holds the interval widths first
    if n < 3:
        b[0] = (y[1] - y[0]) / (x[1] - x[0])
        b[1] = b[0]
        return b, c, d
    bb = np.zeros(n)
    dd[0] = x[1] - x[0]
    c[1] = (y[1] - y[0]) / dd[0]
    for i in range(1, n - 1):
        dd[i] = x[i + 1] - x[i]
        bb[i] = 2.0 * (dd[i - 1] + dd[i])
        c[i + 1] = (y[i + 1] - y[i]) / dd[i]
        c[i] = c[i + 1] - c[i]
    bb[0] = -dd[0]
    bb[n - 1] = -dd[n - 2]
    c[0] = 0.0
    c[n - 1] = 0.0
    if n != 3:
        c[0] = c[2] / (x[3] - x[1]) - c[1] / (x[2] - x[0])
        c[n - 1] = c[n - 2] / (x[n - 1] - x[n - 3]) - c[n - 3] / (x[n - 2] - x[n - 4])
        c[0] = c[0] * dd[0] ** 2 / (x[3] - x[0])
        c[n - 1] = -c[n - 1] * dd[n - 2] ** 2 / (x[n - 1] - x[n - 4])
    for i in range(1, n):
        tt = dd[i - 1] / bb[i - 1]
        bb[i] = bb[i] - tt * dd[i - 1]
        c[i] = c[i] - tt * c[i - 1]
    c[n - 1] = c[n - 1] / bb[n - 1]
    for i in range(n - 2, -1, -1):
        c[i] = (c[i] - dd[i] * c[i + 1]) / bb[i]
    b[n - 1] = (y[n - 1] - y[n - 2]) / dd[n - 2] + dd[n - 2] * (c[n - 2] + 2.0 * c[n - 1])
    for i in range(n - 1):
        b[i] = (y[i + 1] - y[i]) / dd[i] - dd[i] * (c[i + 1] + 2.0 * c[i])
        d[i] = (c[i + 1] - c[i]) / dd[i]
        c[i] = 3.0 * c[i]
    c[n - 1] = 3.0 * c[n - 1]
    d[n - 1] = d[n - 2]
    return b, c, d


def locate(xx, x):
    """spliner.py:1136-1171."""
    n = len(xx)
    if x < xx[0]:
        return 0
    if x >= xx[n - 1]:
        return n - 1
    jl, ju = 0, n
    while ju - jl > 1:
        jm = (ju + jl) // 2
        if x >= xx[jm]:
            jl = jm
        else:
            ju = jm
    return jl


def rescale_1d(x0, z0, x1):
    """rescale_1D_dpc (spliner.py:153-188): zero outside [min x0, max x0]."""
    b, c, d = spline_coeffs(x0, z0)
    out = np.zeros(len(x1), dtype=complex)
    lo, hi = np.min(x0), np.max(x0)
    for i, u in enumerate(x1):
        if lo <= u <= hi:
            j = locate(x0, u)
            dx = u - x0[j]
            out[i] = z0[j] + dx * (b[j] + dx * (c[j] + dx * d[j]))
    return out


def prop2qw(g, t, rr, Exx):
    """qwoptics.py:667-715 for one field: spline, window, real, FFTG."""
    E = rescale_1d(rr, Exx, g.r)
    E = E * t.window
    E = np.real(E) + 0j
    return fftg(E)


# ---------------------------------------------------------------------------
# Solver (SBEs.py:4619-5985): state, one step, QWCalculator
# ---------------------------------------------------------------------------

def calc_xqw(iq, w, dkr, fe, fh, Ee, Eh, gap, area, game, gamh, dcv):
    """Linear susceptibility chi(q, w) of the wire (_CalcXqw_jit,
    SBEs.py:3165-3243, called through SBESolver.calc_Xqw 5331-5357):
    dephasing floored at 0.1 meV, Ec = gap + Ee, Ev = -Eh, fc = fe,
    fv = 1 - fh, x = 2 dkr / (2 pi area) dcv^2 / eps0, and over
    k in [max(0, -iq), min(N, N - iq)) the absorption sum
    x (fc[k] - fv[k+q]) / (Ev[k+q] - Ec[k] - hbar w - i (hgh[k+q] + hge[k]))
    followed by the emission sum
    x (fv[k] - fc[k+q]) / (Ec[k+q] - Ev[k] - hbar w - i (hge[k+q] + hgh[k]))."""
    N = len(Ee)
    floor = 1e-4 * e0
    hge = np.maximum(floor, hbar * np.asarray(game))
    hgh = np.maximum(floor, hbar * np.asarray(gamh))
    Ec, Ev = gap + Ee, -Eh
    fc, fv = fe, 1.0 - fh
    x = 2.0 * dkr / twopi / area * dcv * dcv / eps0
    k0, k1 = (0, N - iq) if iq >= 0 else (-iq, N)
    X = 0j
    for k in range(k0, k1):
        X += x * (fc[k] - fv[k + iq]) / (Ev[k + iq] - Ec[k] - hbar * w - 1j * (hgh[k + iq] + hge[k]))
    for k in range(k0, k1):
        X += x * (fv[k] - fc[k + iq]) / (Ec[k + iq] - Ev[k] - hbar * w - 1j * (hge[k + iq] + hgh[k]))
    return X


def rk4(y, dydt, h, D):
    """Classic RK4 step (_rk4, libpulsesuite/integrator.py:143-157):
    returns y + h/6 (k1 + k4 + 2 (k2 + k3)) with dydt = k1 and D(y) the
    derivative (autonomous here: the wire step holds the field)."""
    ytmp = y + (h / 2.0) * dydt
    dyt = D(ytmp)
    ytmp = y + (h / 2.0) * dyt
    dym = D(ytmp)
    ytmp = y + h * dym
    dym = dym + dyt
    dyt = D(ytmp)
    return y + (h / 6.0) * (dydt + dyt + 2.0 * dym)


class Solver:
    """Mirror of SBESolver for parity: one N, Nw wires, state in the
    reference's [N, N, Nw] layout."""

    def __init__(self, prm, Nw, rr, Emax0, tables=None, grid=None, integrator="leapfrog"):
        self.prm = prm
        self.integrator = integrator
        self.g = grid if grid is not None else Grid(prm)
        self.t = tables if tables is not None else Tables(self.g)
        g = self.g
        N = g.Nk
        self.Nw = Nw
        self.rr = np.asarray(rr, dtype=float)
        self.Emax0 = float(Emax0)
        f = fermi_distr(g.Ee + prm.gap / 2.0)
        z = lambda: np.zeros((N, N, Nw), dtype=complex)  # noqa: E731
        self.YY1, self.YY2, self.YY3 = z(), z(), z()
        self.CC1, self.CC2, self.CC3 = z(), z(), z()
        self.DD1, self.DD2, self.DD3 = z(), z(), z()
        for w in range(Nw):
            for k in range(N):
                self.CC1[k, k, w] = f[k]
                self.DD1[k, k, w] = f[k]
        self.CC2[:] = self.CC1
        self.CC3[:] = self.CC1
        self.DD2[:] = self.DD1
        self.DD3[:] = self.DD1
        self.I0 = np.zeros(Nw)
        self.xxx = 1
        self.jjj = 1
        self.tt = 0.0
        self.wireoff = True
        self.S_eh = self.S_ee = self.S_hh = None
        self.last = {}

    def _mb(self):
        if self.S_eh is None:
            g, prm = self.g, self.prm
            lz = bool(prm.LorentzDelta)  # mb.params line 15 (SBEs.py:5015-5022)
            self.S_eh = mb_tables(g, "eh", prm.gam_e, prm.gam_h, lz)
            self.S_ee = mb_tables(g, "ee", prm.gam_e, prm.gam_h, lz)
            self.S_hh = mb_tables(g, "hh", prm.gam_e, prm.gam_h, lz)

    def preparation(self, p2, C2, D2, Ex, Ey, Ez, w):
        """SBEs.py:5173-5226."""
        g, t, prm = self.g, self.t, self.prm
        N = g.Nk
        ne = np.diag(C2).copy()
        nh = np.diag(D2).copy()
        Meh = np.zeros((N, N), dtype=complex)
        if prm.Optics:
            Meh = calc_meh(g, t, Ex, Ey, Ez, w)
        VC, E1D = screened_arrays(g, t, prm.Screened, ne, nh)
        Heh, Hee, Hhh = calc_h(g, Meh, C2, D2, p2, VC, prm.Excitons, prm.LF)
        GamE = np.full(N, prm.gam_e)
        GamH = np.full(N, prm.gam_h)
        if prm.DiagDph:
            GamE = gamma_e(g, t, ne, nh, VC)
            GamH = gamma_h(g, t, ne, nh, VC)
        OffP = np.zeros((N, N), dtype=complex)
        if prm.OffDiagDph:
            OffP = offdiag_dephasing2(g, ne, nh, p2, VC)[0]
        Rsp = np.zeros(N)
        if prm.Recomb:
            Rsp = spont_emission(g, t, ne, nh, VC)
        return dict(Heh=Heh, Hee=Hee, Hhh=Hhh, VC=VC, E1D=E1D, GamE=GamE,
                    GamH=GamH, OffP=OffP, Rsp=Rsp, Meh=Meh)

    def rhs(self, p2, C2, D2, Ex, Ey, Ez, w):
        pr = self.preparation(p2, C2, D2, Ex, Ey, Ez, w)
        dp = dpdt(C2, D2, p2, pr["Heh"], pr["Hee"], pr["Hhh"], pr["GamE"],
                  pr["GamH"], pr["OffP"])
        dC = dCdt(C2, D2, p2, pr["Heh"], pr["Hee"], pr["Hhh"])
        dD = dDdt(C2, D2, p2, pr["Heh"], pr["Hee"], pr["Hhh"])
        return dp, dC, dD, pr

    def relaxation(self, ne, nh, VC, E1D, Rsp, dt):
        """SBEs.py:5232-5325."""
        prm, g, t = self.prm, self.g, self.t
        N = g.Nk
        if prm.EHs or prm.Phonon:
            WinE = np.zeros(N)
            WoutE = np.zeros(N)
            WinH = np.zeros(N)
            WoutH = np.zeros(N)
            if prm.EHs:
                self._mb()
                a, b = mbce(g, self.S_eh, self.S_ee, np.real(ne), np.real(nh), VC)
                WinE += a
                WoutE += b
                a, b = mbch(g, self.S_eh, self.S_hh, np.real(ne), np.real(nh), VC)
                WinH += a
                WoutH += b
            if prm.Phonon:
                a, b = mbpe(t, np.real(ne), VC, E1D)
                WinE += a
                WoutE += b
                a, b = mbph(t, np.real(nh), VC, E1D)
                WinH += a
                WoutH += b
            We = WinE + WoutE
            Wh = WinH + WoutH
            ne = (np.exp(-We * dt) * ((-1 + np.exp(We * dt) + ne) * WinE + ne * WoutE)
                  / (We + SMALL))
            nh = (np.exp(-Wh * dt) * ((-1 + np.exp(Wh * dt) + nh) * WinH + nh * WoutH)
                  / (Wh + SMALL))
        if prm.Recomb:
            ne = ne * np.exp(-Rsp * nh * dt)
            nh = nh * np.exp(-Rsp * ne * dt)
        return ne, nh

    def sbe_step(self, Ex, Ey, Ez, dt, w):
        """SBEs.py:5612-5820 for wire w (1-based); returns the QW-space
        (FFTG'd) sources Px, Py, Pz, re, rh."""
        g, t, prm = self.g, self.t, self.prm
        N = g.Nk
        i = w - 1
        p1 = self.YY2[:, :, i].copy()
        p2 = self.YY3[:, :, i].copy()
        C1 = self.CC2[:, :, i].copy()
        C2 = self.CC3[:, :, i].copy()
        D1 = self.DD2[:, :, i].copy()
        D2 = self.DD3[:, :, i].copy()
        dp, dC, dD, pr = self.rhs(p2, C2, D2, Ex, Ey, Ez, w)
        self.last = dict(dp=dp, dC=dC, dD=dD, **pr)
        C3 = C1 + dC * dt * 2.0
        D3 = D1 + dD * dt * 2.0
        p3 = p1 + dp * dt * 2.0
        if prm.EHs or prm.Phonon:
            ne3 = np.diag(C3).copy()
            nh3 = np.diag(D3).copy()
            ne3, nh3 = self.relaxation(ne3, nh3, pr["VC"], pr["E1D"], pr["Rsp"], dt)
            C3[np.diag_indices(N)] = ne3
            D3[np.diag_indices(N)] = nh3
        if prm.DCTrans:
            Eac = np.real(Ex[g.NQ0])
            C3 = transport(t, C3, prm.Edc, Eac, dt, prm.LF)
            D3 = transport(t, D3, prm.Edc, Eac, dt, prm.LF)
            p3 = transport(t, p3, prm.Edc, Eac, dt, prm.LF)
        ne3 = np.abs(np.diag(C3)).astype(complex)
        nh3 = np.abs(np.diag(D3)).astype(complex)
        total = (np.sum(np.abs(ne3)) + np.sum(np.abs(nh3))) / 2.0
        ne3 = ne3 * total / (np.sum(np.abs(ne3)) + SMALL)
        nh3 = nh3 * total / (np.sum(np.abs(nh3)) + SMALL)
        C3[np.diag_indices(N)] = ne3
        D3[np.diag_indices(N)] = nh3
        p2 = (p1 + p3) / 2.0
        C2 = (C1 + C3) / 2.0
        D2 = (D1 + D3) / 2.0
        ne2 = np.abs(np.diag(C2)).astype(complex)
        nh2 = np.abs(np.diag(D2)).astype(complex)
        Nr = g.Nr
        z = np.zeros(Nr, dtype=complex)
        Px, Py, Pz = z, z, z
        if prm.Optics:
            Px, Py, Pz = qw_polarization3(g, t, p3, w)
        if prm.LF:
            re, rh = qw_rho5(g, t, C3, D3, ne3, nh3)
            re = 2.0 * e0 * re / g.area * g.ehint
            rh = 2.0 * e0 * rh / g.area * g.ehint
        else:
            re, rh = z, z
        self.I0[i] = np.real(calc_i0(g, ne2, nh2, pr["VC"]))
        self.YY1[:, :, i] = p1
        self.YY2[:, :, i] = p2
        self.YY3[:, :, i] = p3
        self.CC1[:, :, i] = C1
        self.CC2[:, :, i] = C2
        self.CC3[:, :, i] = C3
        self.DD1[:, :, i] = D1
        self.DD2[:, :, i] = D2
        self.DD3[:, :, i] = D3
        self.last.update(ne2=ne2, nh2=nh2, ne3=ne3, nh3=nh3, re=re, rh=rh,
                         Px=Px, Py=Py, Pz=Pz)
        return Px, Py, Pz, re, rh

    def rk4_step(self, Ex, Ey, Ez, dt, w):
        """The wire step with classic RK4 (rk4, the restatement of _rk4 at
        integrator.py:143-157) for the coherent part in place of the leapfrog
        + average of SBEs.py:5673-5706: four RHS evaluations of the state
        (level 3) with the field of the step start; relaxation with the
        screening of the last stage, transport and the diagonal
        renormalisation on the new state, which is also what CalcI0 and the
        sources read.  The new state goes to levels 2 and 3."""
        g, t, prm = self.g, self.t, self.prm
        N = g.Nk
        i = w - 1
        y = np.array([self.YY3[:, :, i], self.CC3[:, :, i], self.DD3[:, :, i]])
        last = {}

        def D(ys):
            dp, dC, dD, pr = self.rhs(ys[0], ys[1], ys[2], Ex, Ey, Ez, w)
            last["pr"] = pr
            return np.array([dp, dC, dD])

        y = rk4(y, D(y), dt, D)
        pr = last["pr"]
        p3, C3, D3 = y[0].copy(), y[1].copy(), y[2].copy()
        if prm.EHs or prm.Phonon:
            ne3, nh3 = self.relaxation(np.diag(C3).copy(), np.diag(D3).copy(),
                                       pr["VC"], pr["E1D"], pr["Rsp"], dt)
            C3[np.diag_indices(N)] = ne3
            D3[np.diag_indices(N)] = nh3
        if prm.DCTrans:
            Eac = np.real(Ex[g.NQ0])
            C3 = transport(t, C3, prm.Edc, Eac, dt, prm.LF)
            D3 = transport(t, D3, prm.Edc, Eac, dt, prm.LF)
            p3 = transport(t, p3, prm.Edc, Eac, dt, prm.LF)
        ne3 = np.abs(np.diag(C3)).astype(complex)
        nh3 = np.abs(np.diag(D3)).astype(complex)
        total = (np.sum(np.abs(ne3)) + np.sum(np.abs(nh3))) / 2.0
        ne3 = ne3 * total / (np.sum(np.abs(ne3)) + SMALL)
        nh3 = nh3 * total / (np.sum(np.abs(nh3)) + SMALL)
        C3[np.diag_indices(N)] = ne3
        D3[np.diag_indices(N)] = nh3
        Nr = g.Nr
        z = np.zeros(Nr, dtype=complex)
        Px, Py, Pz = z, z, z
        if prm.Optics:
            Px, Py, Pz = qw_polarization3(g, t, p3, w)
        if prm.LF:
            re, rh = qw_rho5(g, t, C3, D3, ne3, nh3)
            re = 2.0 * e0 * re / g.area * g.ehint
            rh = 2.0 * e0 * rh / g.area * g.ehint
        else:
            re, rh = z, z
        self.I0[i] = np.real(calc_i0(g, ne3, nh3, pr["VC"]))
        for lv in (self.YY2, self.YY3):
            lv[:, :, i] = p3
        for lv in (self.CC2, self.CC3):
            lv[:, :, i] = C3
        for lv in (self.DD2, self.DD3):
            lv[:, :, i] = D3
        self.last = dict(pr, ne3=ne3, nh3=nh3, re=re, rh=rh, Px=Px, Py=Py, Pz=Pz)
        return Px, Py, Pz, re, rh

    def qw2prop(self, Px, Py, Pz, re, rh):
        """qwoptics.py:717-805 (LF branch)."""
        g = self.g
        rr = self.rr
        dr = g.r[1] - g.r[0]
        dRR = rr[1] - rr[0]
        Px, Py, Pz = ifftg(Px), ifftg(Py), ifftg(Pz)
        re, rh = ifftg(re), ifftg(rh)
        LF = self.prm.LF
        if LF:
            re = np.abs(re) + 0j
            rh = np.abs(rh) + 0j
            total = (np.sum(re) * dr + np.sum(rh) * dr + 1e-100) / 2.0
            re = re * total / (np.sum(re) * dr + 1e-100)
            rh = rh * total / (np.sum(rh) * dr + 1e-100)
        Pxx = rescale_1d(g.r, Px, rr)
        Pyy = rescale_1d(g.r, Py, rr)
        Pzz = rescale_1d(g.r, Pz, rr)
        RhoE = rescale_1d(g.r, re, rr)
        RhoH = rescale_1d(g.r, rh, rr)
        if LF:
            RhoE = np.abs(RhoE) + 0j
            RhoH = np.abs(RhoH) + 0j
            RhoE = RhoE * total / (np.sum(RhoE) * dRR + 1e-100)
            RhoH = RhoH * total / (np.sum(RhoH) * dRR + 1e-100)
        return Pxx, Pyy, Pzz, RhoE, RhoH

    def qw_calculator(self, Exx, Eyy, Ezz, dt, w):
        """SBEs.py:5826-5985 minus file output; returns Pxx, Pyy, Pzz, Rho
        (None while the wire-off gate holds)."""
        prm = self.prm
        if w == 1:
            if self.jjj == prm.jmax:
                self.jjj = 0
            self.xxx += 1
            self.jjj += 1
            self.tt += dt
        mag = np.max(np.sqrt(np.abs(Exx) ** 2 + np.abs(Eyy) ** 2 + np.abs(Ezz) ** 2))
        if self.wireoff and mag < 1e-3 * self.Emax0:
            return None
        self.wireoff = False
        Ex = prop2qw(self.g, self.t, self.rr, Exx)
        Ey = prop2qw(self.g, self.t, self.rr, Eyy)
        Ez = prop2qw(self.g, self.t, self.rr, Ezz)
        step = self.rk4_step if self.integrator == "rk4" else self.sbe_step
        Px, Py, Pz, re, rh = step(Ex, Ey, Ez, dt, w)
        Pxx, Pyy, Pzz, RhoE, RhoH = self.qw2prop(Px, Py, Pz, re, rh)
        Rho = RhoH - RhoE
        if prm.ignorewire:
            z = np.zeros_like(Pxx)
            return z, z.copy(), z.copy(), z.copy()
        return Pxx, Pyy, Pzz, Rho
