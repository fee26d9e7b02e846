"""Drop-in for the reference's ``emission`` module (``PSTD3D/emission.py``)
with ``SpontEmission`` on the GPU.

* ``InitializeEmission(ky, Ee, Eh, dcv, epsr, geh, ehint)``
  (``emission.py:35-91``): the module state ``_RScale, _HOmega, _square,
  _idel`` (host tables, ``CalcHOmega`` ``emission.py:220-255``).
* ``SpontEmission(ne, nh, Ee, Eh, gap, geh, VC, Rsp)`` (``emission.py:94-130``)
  and ``Ec(ne, nh, VC)`` (``emission.py:133-171``) on the device
  (``psa_mod_spont``).
* ``SpontIntegral``, ``rho0`` (``emission.py:174-217``): scalar helpers on
  the host.
* ``Calchw(hw, PLS, Estart, Emax)`` (``emission.py:258-288``): the photon
  energy grid, and ``PLSpectrum(ne, nh, Ee, Eh, gap, geh, VC, hw, t, PLS)``
  (``emission.py:291-361``) on the device (``psa_mod_plspectrum``).
"""

import numpy as np
import torch
from scipy.constants import c as c0
from scipy.constants import epsilon_0 as eps0
from scipy.constants import hbar
from scipy.constants import k as kB

from . import capi
from .coulomb import C128, F64, device, stream, to_dev, to_host

pi = np.pi
_RScale = 0.0
_Temp = 77.0
_HOmega = None
_square = None
_idel = None


def _lrtz(a, b):
    return (b / pi) / (a**2 + b**2)


def CalcHOmega(kBT, hg):
    """emission.py:220-255."""
    global _HOmega
    hwmax = (kBT + hg) * 4.0
    dhw = min(kBT, hg) / 20.0
    Nw = int(np.ceil(hwmax / dhw))
    if Nw < 10:
        raise ValueError("Error: temperature is too low in emission.py")
    _HOmega = 0.0 + (np.arange(Nw) + 0.5) * dhw


def InitializeEmission(ky, Ee, Eh, dcv, epsr, geh, ehint):
    """emission.py:35-91."""
    global _RScale, _square, _idel
    Nk = len(ky)
    _idel = np.ones((Nk, Nk))
    np.fill_diagonal(_idel, 0.0)
    _RScale = 3.0 * dcv**2 / eps0 / np.sqrt(epsr) * ehint**2
    CalcHOmega(kB * _Temp, hbar * geh)
    _square = (3.0 * dcv**2 * ehint / eps0 / np.sqrt(epsr) * ehint / hbar
               * _lrtz(_HOmega, hbar * geh) * np.exp(-_HOmega / kB / _Temp))


def rho0(hw):
    """emission.py:201-217: photon density of states."""
    return hw**2 / (c0**3 * pi**2 * hbar**3)


def SpontIntegral(Ek):
    """emission.py:174-198."""
    if _HOmega is None or len(_HOmega) < 2:
        raise ValueError("HOmega not initialized. Call InitializeEmission first.")
    dhw = _HOmega[1] - _HOmega[0]
    return np.sum((_HOmega + Ek) * rho0(_HOmega + Ek) * _square) * dhw


def _spont(ne, nh, Ee, Eh, gap, VC, Rsp, want_ec):
    if _HOmega is None or len(_HOmega) < 2 or _square is None:
        raise ValueError("HOmega not initialized. Call InitializeEmission first.")
    N = len(ne)
    if np.shape(VC) != (N, N, 3) or len(nh) != N:
        raise ValueError("ne, nh need N entries and VC shape (N, N, 3)")
    lib, dev = capi.load(), device()
    ned, nhd, vcd = to_dev(ne, C128, dev), to_dev(nh, C128, dev), to_dev(VC, F64, dev)
    Ed, Hd = to_dev(Ee, F64, dev), to_dev(Eh, F64, dev)
    hd, sd = to_dev(_HOmega, F64, dev), to_dev(_square, F64, dev)
    r = torch.empty(N, dtype=F64, device=dev)
    ec = torch.empty(N, dtype=F64, device=dev) if want_ec else None
    capi.check(lib.psa_mod_spont(N, len(_HOmega), ned.data_ptr(), nhd.data_ptr(), vcd.data_ptr(), Ed.data_ptr(),
                                 Hd.data_ptr(), float(gap), hd.data_ptr(), sd.data_ptr(),
                                 float(_HOmega[1] - _HOmega[0]), c0**3 * pi**2 * hbar**3, r.data_ptr(),
                                 None if ec is None else ec.data_ptr(), stream(dev)))
    if Rsp is not None:
        Rsp[...] = to_host(r)
    return to_host(ec) if want_ec else None


def SpontEmission(ne, nh, Ee, Eh, gap, geh, VC, Rsp):
    """emission.py:94-130: Rsp[k] = SpontIntegral(gap + Ee + Eh + Ec)."""
    _spont(ne, nh, Ee, Eh, gap, VC, Rsp, False)


def Ec(ne, nh, VC):
    """emission.py:133-171: the many-body Coulomb energy correction."""
    N = len(ne)
    z = np.zeros(N)
    if _HOmega is None:
        raise ValueError("HOmega not initialized. Call InitializeEmission first.")
    return _spont(ne, nh, z, z, 0.0, VC, None, True)


def Calchw(hw, PLS, Estart, Emax):
    """emission.py:258-288: hw[w] = Estart + w (Emax - Estart) / Nw; PLS = 0."""
    Nw = len(hw)
    hw[:] = 0.0
    PLS[:] = 0.0
    dhw = (Emax - Estart) / float(Nw)
    for w in range(Nw):
        hw[w] = Estart + w * dhw


def PLSpectrum(ne, nh, Ee, Eh, gap, geh, VC, hw, t, PLS):
    """emission.py:291-361: adds the photoluminescence of the occupations ne,
    nh (interpolated 100x in k) at the photon energies hw to PLS, then
    scales PLS by softtheta(t, geh); on the device."""
    N = len(ne)
    if N < 2 or len(nh) != N or np.shape(VC) != (N, N, 3) or len(Ee) != N or len(Eh) != N:
        raise ValueError("ne, nh, Ee, Eh need N >= 2 entries and VC shape (N, N, 3)")
    if len(PLS) != len(hw):
        raise ValueError("PLS and hw need the same length")
    lib, dev = capi.load(), device()
    ned, nhd = to_dev(ne, C128, dev), to_dev(nh, C128, dev)
    Ed, Hd, vcd = to_dev(Ee, F64, dev), to_dev(Eh, F64, dev), to_dev(VC, F64, dev)
    hwd = to_dev(hw, F64, dev)
    pls = to_dev(np.real(np.asarray(PLS)), F64, dev)
    work = torch.empty(N + 1, dtype=F64, device=dev)
    tfac = 0.5 * (1.0 + 2.0 / pi * np.arctan(t / geh))
    capi.check(lib.psa_mod_plspectrum(N, len(hw), ned.data_ptr(), nhd.data_ptr(), Ed.data_ptr(), Hd.data_ptr(),
                                      float(gap), hbar * geh, vcd.data_ptr(), hwd.data_ptr(),
                                      c0**3 * pi**2 * hbar**3, kB, float(_RScale), float(tfac),
                                      work.data_ptr(), pls.data_ptr(), stream(dev)))
    PLS[...] = to_host(pls)
