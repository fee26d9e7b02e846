"""Drop-in for the reference's ``coulomb`` module (``PSTD3D/coulomb.py``):
the functions the SBE step calls, on the GPU, with the reference's call
signatures, in-place outputs and module state.

* ``CoulombModule(y, ky, L, Delta0, me, mh, Ee, Eh, ge, gh, alphae, alphah,
  er, Qy, kkp, screened, LorentzDelta=False)`` (``coulomb.py:1079-1164``):
  the (N+1)^3 many-body tables and k3 are built on the device
  (``psa_mod_mb_arrays``), the unscreened Coulomb integrals too
  (``psa_vint_table``); qe/qh, UnDel and Chi1D are O(N^2) host tables.
* per call, on the device: ``CalcScreenedArrays`` / ``Eps1D``
  (``psa_mod_screen``), ``MBCE`` / ``MBCE2`` / ``MBCH`` (``psa_mod_mbc``).
* module level (``coulomb.py:1337-1447``): the singleton made by
  ``InitializeCoulomb``, ``SetLorentzDelta`` before or after it, the
  wrappers, and ``coulomb._k3``-style reads of the tables (host copies).

Inputs are the caller's numpy arrays; outputs are written into the caller's
arrays with ``X[...] = ...`` (strided views work).  The tables stay on the
device; the attributes ``k3``, ``Ceh``, ``Cee``, ``Chi1De`` ... return host
arrays.  There is no CPU fallback: without the HIP library or a GPU the
constructor raises.

Off the SBE step, also on the device: ``CalcMVeh`` (``psa_mod_mveh``),
``BGRenorm`` / ``EeRenorm`` / ``EhRenorm`` (``psa_mod_renorm``), ``Vint``
(``psa_vint_table`` for one momentum) and ``Vehint`` (``psa_mod_vehint``,
with the accurate K0 as the reference's Vehint uses).  ``GetChi1Dqw`` and
``GetEps1Dqw`` are scalar diagnostics of one (q, w) and are evaluated on the
host, as ``CalcChi1D`` is (the chi(q, w) record of the SBE drop-in runs on
the device through ``psa_xqw_table``).
"""

import ctypes
from math import pi

import numpy as np
import torch
from scipy.constants import e as e0
from scipy.constants import epsilon_0 as eps0
from scipy.constants import hbar

from . import capi
from .tables import k0_zhang_jin

twopi = 2.0 * pi
F64 = torch.float64
C128 = torch.complex128


# ---------------------------------------------------------------------------
# device plumbing shared by the module drop-ins
# ---------------------------------------------------------------------------

def device():
    if not torch.cuda.is_available():
        raise capi.PsaError("no GPU visible: the module drop-ins have no CPU fallback")
    return torch.device("cuda")


def stream(dev):
    return ctypes.c_void_p(torch.cuda.current_stream(dev).cuda_stream)


def to_dev(a, dtype, dev):
    """Contiguous device copy of a host array-like."""
    np_dt = np.complex128 if dtype == C128 else (np.int32 if dtype == torch.int32 else np.float64)
    return torch.as_tensor(np.ascontiguousarray(np.asarray(a, dtype=np_dt))).to(dev)


def to_host(t):
    torch.cuda.current_stream(t.device).synchronize()
    return t.cpu().numpy()


# ---------------------------------------------------------------------------
# pure builders (coulomb.py:245-630)
# ---------------------------------------------------------------------------

def GaussDelta(a, b):
    """coulomb.py:245-268."""
    if abs(b) < 1e-300:
        return 0.0
    return 1.0 / (np.sqrt(pi) * b) * np.exp(-((a / b) ** 2))


def MakeK3(ky):
    """coulomb.py:271-299: k3[k1, k2, k4] = k1 + k2 - k4 (1-based), 0
    outside [1, N]."""
    N = len(ky)
    i = np.arange(1, N + 1)
    k3 = i[:, None, None] + i[None, :, None] - i[None, None, :]
    return np.where((k3 >= 1) & (k3 <= N), k3, 0).astype(np.int32)


def MakeQs(ky, ae, ah):
    """coulomb.py:302-333: q[k1, k2] = max(|ky[k2] - ky[k1]|, a / 2)."""
    ky = np.asarray(ky, dtype=float)
    d = np.abs(ky[None, :] - ky[:, None])
    return np.maximum(d, ae / 2.0), np.maximum(d, ah / 2.0)


def MakeUnDel(ky):
    """coulomb.py:336-364: (N+1) x (N+1) ones, zero row/column 0 and
    diagonal."""
    N = len(ky)
    u = np.ones((N + 1, N + 1))
    u[0, :] = 0.0
    u[:, 0] = 0.0
    u[np.arange(1, N + 1), np.arange(1, N + 1)] = 0.0
    return u


def _mb_arrays_dev(Ee, Eh, ge, gh, lorentz, dev):
    lib = capi.load()
    N = len(Ee)
    if N > 1024:
        raise ValueError("the many-body tables need N <= 1024")
    k3 = torch.zeros((N, N, N), dtype=torch.int32, device=dev)
    C = [torch.zeros((N + 1,) * 3, dtype=F64, device=dev) for _ in range(3)]
    # inputs held in locals until the launch is queued (a temporary's block
    # could otherwise be handed to the next allocation before the kernel runs)
    Ed, Hd = to_dev(Ee, F64, dev), to_dev(Eh, F64, dev)
    capi.check(lib.psa_mod_mb_arrays(N, Ed.data_ptr(), Hd.data_ptr(), float(ge), float(gh), hbar,
                                     int(bool(lorentz)), k3.data_ptr(), C[0].data_ptr(), C[1].data_ptr(),
                                     C[2].data_ptr(), stream(dev)))
    return k3, C


def CalcMBArrays(ky, Ee, Eh, ge, gh, k3, UnDel, LorentzDelta=False):
    """coulomb.py:367-499 on the device.  ``k3`` and ``UnDel`` must be
    ``MakeK3(ky)`` and ``MakeUnDel(ky)`` (as every caller in the reference
    passes them); returns host (Ceh, Cee, Chh)."""
    N = len(ky)
    if np.shape(k3) != (N, N, N) or np.shape(UnDel) != (N + 1, N + 1):
        raise ValueError("k3 / UnDel shapes do not match ky")
    _, C = _mb_arrays_dev(Ee, Eh, ge, gh, LorentzDelta, device())
    return tuple(to_host(c) for c in C)


def CalcCoulombArrays(y, ky, er, alphae, alphah, L, Delta0, Qy, kkp, ReadArrays=False, ScrewThis=False):
    """coulomb.py:502-577: the Vint integrals (crude K0, middle half of y)
    on the device for every Qy, then V0[k, q] = table[kkp[k, q]] for
    0 <= kkp < NQ."""
    N, NQ = len(ky), len(Qy)
    V = [np.zeros((N, N)) for _ in range(3)]
    if ReadArrays or ScrewThis:
        return tuple(V)
    dev = device()
    lib = capi.load()
    out = torch.zeros(3, NQ, dtype=F64, device=dev)
    pref = e0**2 / (twopi * eps0 * er * L)
    yd, Qd = to_dev(y, F64, dev), to_dev(Qy, F64, dev)
    capi.check(lib.psa_vint_table(Qd.data_ptr(), NQ, yd.data_ptr(), len(y), float(alphae), float(alphah),
                                  float(Delta0), pref, out.data_ptr(), stream(dev)))
    tab = to_host(out)
    kkp = np.asarray(kkp)
    ok = (kkp >= 0) & (kkp < NQ)
    idx = np.where(ok, kkp, 0)
    for c in range(3):
        V[c] = np.where(ok, tab[c][idx], 0.0)
    return tuple(V)


def Vint(Qyk, y, alphae, alphah, Delta0):
    """coulomb.py:107-177: the interaction integral of one momentum Qyk (the
    reference's JIT path, crude K0), on the device (psa_vint_table's eh
    channel with unit prefactor)."""
    dev = device()
    lib = capi.load()
    out = torch.zeros(3, 1, dtype=F64, device=dev)
    yd, Qd = to_dev(y, F64, dev), to_dev([float(Qyk)], F64, dev)
    capi.check(lib.psa_vint_table(Qd.data_ptr(), 1, yd.data_ptr(), len(y), float(alphae), float(alphah),
                                  float(Delta0), 1.0, out.data_ptr(), stream(dev)))
    return float(to_host(out)[0, 0])


def Vehint(k, q, y, ky, alphae, alphah, Delta0):
    """coulomb.py:180-242: the integral at momentum ky[k-1] - ky[q-1]
    (1-based k, q; 0 reads index 0) with the accurate K0, on the device."""
    if len(y) < 4:
        raise ValueError("Vehint: y needs at least 4 points")
    ki = k - 1 if k > 0 else 0
    qi = q - 1 if q > 0 else 0
    dev = device()
    lib = capi.load()
    out = torch.zeros(1, dtype=F64, device=dev)
    yd = to_dev(y, F64, dev)
    capi.check(lib.psa_mod_vehint(float(ky[ki] - ky[qi]), yd.data_ptr(), len(y), float(alphae), float(alphah),
                                  float(Delta0), out.data_ptr(), stream(dev)))
    return float(to_host(out)[0])


def _k0(x):
    return float(k0_zhang_jin(np.array([x]))[0])


def GetChi1Dqw(alphae, alphah, Delta0, epsr, game, gamh, ky, Ee, Eh, ne, nh, qq, w):
    """coulomb.py:633-709: chi(q, w) of the wire at one (q, w); returns
    (Re chi, Im chi).  Host scalar diagnostic (one pass over Nk)."""
    ky = np.asarray(ky, dtype=float)
    Nk = len(ky)
    dk = ky[1] - ky[0]
    hge = np.maximum(hbar * np.asarray(game), 1e-4 * e0)
    hgh = np.maximum(hbar * np.asarray(gamh), 1e-4 * e0)
    Re = np.sqrt((2.0 / alphae) ** 2 + Delta0**2)
    Rh = np.sqrt((2.0 / alphah) ** 2 + Delta0**2)
    beta = e0**2 / (4 * pi * eps0 * epsr)
    ql = abs(qq)
    Ke = _k0(max(ql, alphah / 2.0) * Re) * beta * 2
    Kh = _k0(max(ql, alphae / 2.0) * Rh) * beta * 2
    q = int(round(qq / dk))
    k = np.arange(max(0, -q), min(Nk, Nk - q))
    kq = k + q
    chi = 0.0 + 0.0j
    for a, b in zip(k, kq):
        de = hbar * w - Ee[b] + Ee[a] + 1j * (hge[b] + hge[a])
        if abs(de) > 1e-300:
            chi -= Ke / pi * (ne[a] - ne[b]) / de * dk
        dh = hbar * w - Eh[b] + Eh[a] + 1j * (hgh[b] + hgh[a])
        if abs(dh) > 1e-300:
            chi -= Kh / pi * (nh[a] - nh[b]) / dh * dk
    return np.real(chi), np.imag(chi)


def GetEps1Dqw(alphae, alphah, Delta0, epsr, me, mh, n1D, q, w):
    """coulomb.py:712-776: the plasmon-pole dielectric function at one
    (q, w); returns (Re eps, Im eps).  Host scalar diagnostic."""
    Re = np.sqrt((2.0 / alphae) ** 2 + Delta0**2)
    Rh = np.sqrt((2.0 / alphah) ** 2 + Delta0**2)
    beta = e0**2 / (4 * pi * eps0 * epsr)
    if abs(q) < 1.0:
        q = 1.0
    ql = abs(q)
    Ce = 2 * beta * me / pi / hbar**2 / ql
    Ch = 2 * beta * mh / pi / hbar**2 / ql
    Ke = _k0(max(ql, alphah / 2.0) * Re)
    Kh = _k0(max(ql, alphae / 2.0) * Rh)
    oep = hbar * ql / 2.0 / me * abs(ql + pi * n1D)
    oem = hbar * ql / 2.0 / me * abs(ql - pi * n1D)
    ohp = hbar * ql / 2.0 / mh * abs(ql + pi * n1D)
    ohm = hbar * ql / 2.0 / mh * abs(ql - pi * n1D)
    epr = (1.0 - Ce * Ke * np.log(abs((w**2 - oem**2) / (w**2 - oep**2)))
           - Ch * Kh * np.log(abs((w**2 - ohm**2) / (w**2 - ohp**2))))
    epi = 0.0
    if min(oem, oep) < w < max(oem, oep) and min(ohm, ohp) < w < max(ohm, ohp):
        epi = -Ce * Ke - Ch * Kh
    if np.isnan(epr) or np.isnan(epi):
        print(f"NAN in EpsL(q,w) at (q,w) = {q}, {w}")
        print(f"Re[EpsL(q,w)] = {epr}")
        print(f"Im[EpsL(q,w)] = {epi}")
    return epr, epi


def CalcChi1D(ky, alphae, alphah, Delta0, epsr, me, mh, qe, qh):
    """coulomb.py:580-630 with the accurate K0 (usefulsubs.K03)."""
    Re = np.sqrt((2.0 / alphae) ** 2 + Delta0**2)
    Rh = np.sqrt((2.0 / alphah) ** 2 + Delta0**2)
    scale = e0**2 / (twopi * eps0 * epsr * hbar**2)

    def chi(m, q, R):
        u, inv = np.unique(q, return_inverse=True)
        return (m * k0_zhang_jin(u * R) / u)[inv].reshape(q.shape) * scale

    return chi(me, np.asarray(qe), Re), chi(mh, np.asarray(qh), Rh)


# ---------------------------------------------------------------------------
# CoulombModule
# ---------------------------------------------------------------------------

class CoulombModule:
    """coulomb.py:1079-1326 with the tables on the device."""

    _HOST = ("k3", "Ceh", "Cee", "Chh")

    def __init__(self, y, ky, L, Delta0, me, mh, Ee, Eh, ge, gh, alphae, alphah, er, Qy, kkp,
                 screened, LorentzDelta=False):
        self.lib = capi.load()
        self.dev = dev = device()
        ky = np.asarray(ky, dtype=float)
        N = len(ky)
        if N < 2 or N > 1024:
            raise ValueError("CoulombModule: need 2 <= N <= 1024")
        self.N = N
        self.LorentzDelta = bool(LorentzDelta)
        self.UnDel = MakeUnDel(ky)
        self.qe, self.qh = MakeQs(ky, alphae, alphah)
        k3, C = _mb_arrays_dev(Ee, Eh, ge, gh, self.LorentzDelta, dev)
        self.Veh0, self.Vee0, self.Vhh0 = CalcCoulombArrays(y, ky, er, alphae, alphah, L, Delta0, Qy, kkp)
        self.Chi1De, self.Chi1Dh = CalcChi1D(ky, alphae, alphah, Delta0, er, me, mh, self.qe, self.qh)
        self._d = {
            "k3": k3, "Ceh": C[0], "Cee": C[1], "Chh": C[2],
            "V0": to_dev(np.stack([self.Veh0, self.Vee0, self.Vhh0]), F64, dev),
            "chi": to_dev(np.stack([self.Chi1De, self.Chi1Dh]), F64, dev),
            "q": to_dev(np.stack([self.qe, self.qh]), F64, dev),
            "ud": to_dev(self.UnDel, F64, dev),
        }
        self._nmax = min(self.qe[1, 1], self.qh[1, 1]) / pi * 0.99
        self._host = {}

    def __getattr__(self, name):
        if name in CoulombModule._HOST:
            h = self.__dict__["_host"]
            if name not in h:
                h[name] = to_host(self.__dict__["_d"][name])
            return h[name]
        raise AttributeError(name)

    def SetLorentzDelta(self, boolean):
        """coulomb.py:1166-1168: the flag only (the tables keep the
        broadening they were built with, as in the reference)."""
        self.LorentzDelta = bool(boolean)

    def Eps1D(self, n1D, Nk=None):
        """coulomb.py:1172-1193."""
        E1D = torch.empty(self.N, self.N, dtype=F64, device=self.dev)
        d = self._d
        capi.check(self.lib.psa_mod_screen(self.N, 1, 0.0, 0.0, None, None, float(n1D), None,
                                           d["chi"].data_ptr(), d["q"].data_ptr(), None, E1D.data_ptr(),
                                           stream(self.dev)))
        return to_host(E1D)

    def CalcScreenedArrays(self, screened, L, ne, nh, VC, E1D):
        """coulomb.py:1195-1230: fills VC (N, N, 3) and E1D (N, N) in place."""
        N, d, dev = self.N, self._d, self.dev
        if len(ne) != N or len(nh) != N:
            raise ValueError("ne, nh must have N entries")
        ned, nhd = to_dev(ne, C128, dev), to_dev(nh, C128, dev)
        VCd = torch.empty(N, N, 3, dtype=F64, device=dev)
        Ed = torch.empty(N, N, dtype=F64, device=dev)
        capi.check(self.lib.psa_mod_screen(N, int(bool(screened)), float(L), self._nmax, ned.data_ptr(),
                                           nhd.data_ptr(), 0.0, d["V0"].data_ptr(), d["chi"].data_ptr(),
                                           d["q"].data_ptr(), VCd.data_ptr(), Ed.data_ptr(), stream(dev)))
        VC[...] = to_host(VCd)
        E1D[...] = to_host(Ed)

    def undell(self, k, q):
        """coulomb.py:1244-1246: UnDel(k, q) with 1-based indices."""
        return self.UnDel[k, q]

    def CalcMVeh(self, p, VC, MVeh):
        """coulomb.py:1234-1241: MVeh[k, kp, f] = sum_q p[q, qp, f] Veh[k, q]
        UnDel[k+1, q+1] UnDel[kp+1, qp+1] over the k3 momentum conservation;
        p, MVeh (N, N, Nf)."""
        N, dev = self.N, self.dev
        p = np.asarray(p)
        if p.ndim != 3 or p.shape[:2] != (N, N) or np.shape(MVeh) != p.shape or np.shape(VC) != (N, N, 3):
            raise ValueError("p and MVeh need shape (N, N, Nf) and VC shape (N, N, 3)")
        pd, vcd = to_dev(p, C128, dev), to_dev(VC, F64, dev)
        M = torch.empty(p.shape, dtype=C128, device=dev)
        capi.check(self.lib.psa_mod_mveh(N, p.shape[2], pd.data_ptr(), vcd.data_ptr(), self._d["k3"].data_ptr(),
                                         self._d["ud"].data_ptr(), M.data_ptr(), stream(dev)))
        MVeh[...] = to_host(M)

    def _renorm(self, which, ne, nh, VC, BGR):
        N, dev = self.N, self.dev
        if np.shape(VC) != (N, N, 3) or np.shape(BGR) != (N, N):
            raise ValueError("VC needs shape (N, N, 3) and BGR (N, N)")
        for v in (ne, nh):
            if v is not None and len(v) != N:
                raise ValueError("occupations need N entries")
        ned = None if ne is None else to_dev(ne, C128, dev)
        nhd = None if nh is None else to_dev(nh, C128, dev)
        vcd = to_dev(VC, F64, dev)
        B = torch.empty(N, N, dtype=C128, device=dev)
        capi.check(self.lib.psa_mod_renorm(N, which, None if ned is None else ned.data_ptr(),
                                           None if nhd is None else nhd.data_ptr(), vcd.data_ptr(),
                                           self._d["ud"].data_ptr(), B.data_ptr(), stream(dev)))
        BGR[...] = to_host(B)

    def BGRenorm(self, C, D, VC, BGR):
        """coulomb.py:1248-1259: band-gap renormalisation from diag C, diag D."""
        self._renorm(0, np.diag(C), np.diag(D), VC, BGR)

    def EeRenorm(self, ne, VC, BGR):
        """coulomb.py:1261-1269: electron energy renormalisation."""
        self._renorm(1, ne, None, VC, BGR)

    def EhRenorm(self, nh, VC, BGR):
        """coulomb.py:1271-1280: hole energy renormalisation."""
        self._renorm(2, None, nh, VC, BGR)

    def _mbc(self, hole, ne0, nh0, VC, Win, Wout):
        N, d, dev = self.N, self._d, self.dev
        if len(ne0) != N or len(nh0) != N or np.shape(VC) != (N, N, 3):
            raise ValueError("ne0, nh0 must have N entries and VC shape (N, N, 3)")
        wi, wo = to_dev(Win, F64, dev), to_dev(Wout, F64, dev)
        ned, nhd, vcd = to_dev(ne0, C128, dev), to_dev(nh0, C128, dev), to_dev(VC, F64, dev)
        capi.check(self.lib.psa_mod_mbc(N, hole, ned.data_ptr(), nhd.data_ptr(), vcd.data_ptr(),
                                        d["k3"].data_ptr(), d["Ceh"].data_ptr(),
                                        d["Chh" if hole else "Cee"].data_ptr(), wi.data_ptr(), wo.data_ptr(),
                                        stream(dev)))
        Win[...] = to_host(wi)
        Wout[...] = to_host(wo)

    def MBCE2(self, ne0, nh0, ky, Ee, Eh, VC, geh, ge, Win, Wout):
        """coulomb.py:1283-1296 (the same sums as MBCE)."""
        self._mbc(0, ne0, nh0, VC, Win, Wout)

    def MBCE(self, ne0, nh0, ky, Ee, Eh, VC, geh, ge, Win, Wout):
        """coulomb.py:1298-1311: electron in/out rates added to Win, Wout."""
        self._mbc(0, ne0, nh0, VC, Win, Wout)

    def MBCH(self, ne0, nh0, ky, Ee, Eh, VC, geh, gh, Win, Wout):
        """coulomb.py:1313-1326: hole in/out rates added to Win, Wout."""
        self._mbc(1, ne0, nh0, VC, Win, Wout)


# ---------------------------------------------------------------------------
# module singleton and wrappers (coulomb.py:1337-1447)
# ---------------------------------------------------------------------------

_instance = None
_LorentzDelta = False


def SetLorentzDelta(boolean):
    global _LorentzDelta
    _LorentzDelta = bool(boolean)
    if _instance is not None:
        _instance.SetLorentzDelta(boolean)


def InitializeCoulomb(y, ky, L, Delta0, me, mh, Ee, Eh, ge, gh, alphae, alphah, er, Qy, kkp, screened):
    global _instance
    _instance = CoulombModule(y, ky, L, Delta0, me, mh, Ee, Eh, ge, gh, alphae, alphah, er, Qy, kkp,
                              screened, LorentzDelta=_LorentzDelta)


def _require_instance():
    if _instance is None:
        raise ValueError("CoulombModule not initialized. Call InitializeCoulomb first.")
    return _instance


def CalcScreenedArrays(screened, L, ne, nh, VC, E1D):
    _require_instance().CalcScreenedArrays(screened, L, ne, nh, VC, E1D)


def Eps1D(n1D, Nk):
    return _require_instance().Eps1D(n1D, Nk)


def undell(k, q):
    return _require_instance().undell(k, q)


def CalcMVeh(p, VC, MVeh, k3=None, UnDel=None):
    _require_instance().CalcMVeh(p, VC, MVeh)


def BGRenorm(C, D, VC, BGR, UnDel=None):
    _require_instance().BGRenorm(C, D, VC, BGR)


def EeRenorm(ne, VC, BGR, UnDel=None):
    _require_instance().EeRenorm(ne, VC, BGR)


def EhRenorm(nh, VC, BGR, UnDel=None):
    _require_instance().EhRenorm(nh, VC, BGR)


def MBCE2(ne0, nh0, ky, Ee, Eh, VC, geh, ge, Win, Wout, k3=None, Ceh=None, Cee=None):
    _require_instance().MBCE2(ne0, nh0, ky, Ee, Eh, VC, geh, ge, Win, Wout)


def MBCE(ne0, nh0, ky, Ee, Eh, VC, geh, ge, Win, Wout, k3=None, Ceh=None, Cee=None):
    _require_instance().MBCE(ne0, nh0, ky, Ee, Eh, VC, geh, ge, Win, Wout)


def MBCH(ne0, nh0, ky, Ee, Eh, VC, geh, gh, Win, Wout, k3=None, Ceh=None, Chh=None):
    _require_instance().MBCH(ne0, nh0, ky, Ee, Eh, VC, geh, gh, Win, Wout)


_ATTRS = {"_k3": "k3", "_UnDel": "UnDel", "_qe": "qe", "_qh": "qh", "_Ceh": "Ceh", "_Cee": "Cee",
          "_Chh": "Chh", "_Veh0": "Veh0", "_Vee0": "Vee0", "_Vhh0": "Vhh0", "_Chi1De": "Chi1De",
          "_Chi1Dh": "Chi1Dh"}


def __getattr__(name):
    """coulomb.py:1427-1447: ``coulomb._k3`` etc. read the singleton."""
    if name in _ATTRS:
        return None if _instance is None else getattr(_instance, _ATTRS[name])
    raise AttributeError(f"module 'coulomb' has no attribute {name!r}")
