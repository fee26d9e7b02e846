"""Drop-in for the reference's ``phonons`` module (``PSTD3D/phonons.py``)
with the per-step rates on the GPU.

* ``InitializePhonons(ky, Ee, Eh, L, epsr, Gph, Oph)`` (``phonons.py:55-136``)
  fills the module state ``_EP, _EPT, _HP, _HPT, _idel, _NO, _Vscale``
  (O(N^2) host tables, init-time).
* ``MBPE(ne, VC, E1D, Win, Wout)`` / ``MBPH(nh, VC, E1D, Win, Wout)``
  (``phonons.py:139-268``) run on the device (``psa_mod_mbp``) and add the
  in/out rates into the caller's ``Win``, ``Wout``.  They read the module
  state at call time, so tables assigned directly to ``phonons._EP`` etc.
  are used, as in the reference.
* ``Cq2``, ``FermiDistr``, ``BoseDistr``, ``N00`` (``phonons.py:271-380``):
  the scalar helpers, on the host.

Without the HIP library or a GPU, ``MBPE``/``MBPH`` raise (no CPU fallback).
"""

import numpy as np
from scipy.constants import hbar
from scipy.constants import k as kB

from . import capi
from .coulomb import C128, F64, device, stream, to_dev, to_host

_small = 1e-200
_Temp = 77.0
_epsr0 = 10.0
_epsrINF = 8.2
_Vscale = 0.0
_NO = 0.0
_EP = None
_EPT = None
_HP = None
_HPT = None
_idel = None


def InitializePhonons(ky, Ee, Eh, L, epsr, Gph, Oph):
    """phonons.py:55-136."""
    global _NO, _EP, _EPT, _HP, _HPT, _idel, _Vscale
    _NO = 1.0 / (np.exp(hbar * Oph / kB / _Temp) - 1.0)
    Nk = len(ky)
    Ee = np.asarray(Ee, dtype=float)
    Eh = np.asarray(Eh, dtype=float)
    _idel = np.ones((Nk, Nk))
    _idel[np.arange(Nk), np.arange(Nk)] = 0.0
    hO, hG2 = hbar * Oph, (hbar * Gph) ** 2

    def table(E):
        dE = E[:, None] - E[None, :]
        return _NO / ((dE - hO) ** 2 + hG2) + (_NO + 1.0) / ((dE + hO) ** 2 + hG2)

    _EP = table(Ee) * 2.0 * Gph * _idel
    _HP = table(Eh) * 2.0 * Gph * _idel
    _EPT = _EP.T
    _HPT = _HP.T
    _Vscale = hbar * Oph * epsr * (1.0 / _epsrINF - 1.0 / _epsr0)


def _mbp(hole, n, VC, E1D, Win, Wout):
    P, PT = (_HP, _HPT) if hole else (_EP, _EPT)
    if P is None or PT is None:
        raise ValueError("phonon tables not initialized. Call InitializePhonons first.")
    N = len(n)
    if np.shape(VC) != (N, N, 3) or np.shape(E1D) != (N, N) or np.shape(P) != (N, N):
        raise ValueError("VC must be (N, N, 3), E1D and the phonon tables (N, N)")
    lib = capi.load()
    dev = device()
    wi, wo = to_dev(Win, F64, dev), to_dev(Wout, F64, dev)
    # inputs held in locals until the launch is queued
    nd, vcd, ed = to_dev(n, C128, dev), to_dev(VC, F64, dev), to_dev(E1D, F64, dev)
    pd, ptd = to_dev(P, F64, dev), to_dev(PT, F64, dev)
    capi.check(lib.psa_mod_mbp(N, hole, nd.data_ptr(), vcd.data_ptr(), ed.data_ptr(), pd.data_ptr(),
                               ptd.data_ptr(), float(_Vscale), wi.data_ptr(), wo.data_ptr(), stream(dev)))
    Win[...] = to_host(wi)
    Wout[...] = to_host(wo)


def MBPE(ne, VC, E1D, Win, Wout):
    """phonons.py:139-202: Win[k] += sum_i Vep[k,i] ne[i] EPT[i,k],
    Wout[k] += sum_i Vep[k,i] (1 - ne[i]) EP[i,k], Vep = VC[:,:,1] / E1D
    Vscale."""
    _mbp(0, ne, VC, E1D, Win, Wout)


def MBPH(nh, VC, E1D, Win, Wout):
    """phonons.py:205-268: Win[kp] += sum_i Vhp[i,kp] nh[i] HPT[i,kp],
    Wout[kp] += sum_i Vhp[i,kp] (1 - nh[i]) HP[i,kp], Vhp = VC[:,:,2] / E1D
    Vscale."""
    _mbp(1, nh, VC, E1D, Win, Wout)


def Cq2(q, V, E1D):
    """phonons.py:271-314."""
    dq = 1.0 if len(q) < 2 else q[1] - q[0]
    iq = np.round(np.abs(np.asarray(q) / dq)).astype(int)
    out = np.zeros(len(q))
    for i, idx in enumerate(iq):
        if 0 <= idx < V.shape[0] and 0 <= idx < E1D.shape[0]:
            out[i] = V[idx, 0] / E1D[idx, 0] * _Vscale
    return out


def FermiDistr(En):
    """phonons.py:317-340."""
    return 1.0 / (np.exp(En / kB / _Temp) + 1.0)


def BoseDistr(En):
    """phonons.py:343-362."""
    return 1.0 / (np.exp(En / kB / _Temp) - 1.0)


def N00():
    """phonons.py:365-380."""
    return _NO
