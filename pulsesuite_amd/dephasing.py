"""Drop-in for the reference's ``dephasing`` module (``PSTD3D/dephasing.py``)
with the per-step functions on the GPU.

* ``InitializeDephasing(ky, me, mh)`` (``dephasing.py:472-539``): the index
  tables ``_k_p_q, _k_m_q, _k1_m_q, _k1p_m_q, _k1, _k1p`` and the weights
  ``_xe, _xh`` (O(N^2) host tables, module globals that may be read or
  assigned directly), and the ``MaxOffDiag.e/h`` files it opens.
* per call, on the device: ``CalcGammaE`` / ``CalcGammaH``
  (``psa_mod_gamma``), ``CalcOffDiagDeph_E2`` / ``_H2`` and
  ``OffDiagDephasing2`` (``psa_mod_offdiag2``; the latter also appends the
  max/min line of each table, as the reference does).
* ``Vxx2``, ``printGam``, ``WriteDephasing`` (``dephasing.py:362-378, 417-437``).
* off the SBE step (which uses the version-2 functions), also on the
  device: ``CalcOffDiagDeph_E`` / ``_H`` and ``OffDiagDephasing``
  (``dephasing.py:107-218, 627-712``; ``psa_mod_offdiag1``,
  ``psa_mod_offdiag1_apply``), with the reference's one-index shift of the
  second index of p in the contraction.
"""

import os

import numpy as np
import torch
from scipy.constants import hbar

from . import capi
from .coulomb import C128, F64, device, stream, to_dev, to_host

_k_p_q = None
_k_m_q = None
_k1_m_q = None
_k1p_m_q = None
_k1 = None
_k1p = None
_xe = None
_xh = None
_fe_file = None
_fh_file = None


def InitializeDephasing(ky, me, mh):
    """dephasing.py:472-539."""
    global _k_p_q, _k_m_q, _k1_m_q, _k1p_m_q, _k1, _k1p, _xe, _xh, _fe_file, _fh_file
    ky = np.asarray(ky, dtype=float)
    Nk = len(ky)
    dk = ky[1] - ky[0] if Nk > 1 else 1.0
    kmax, kmin = ky[Nk - 1] + dk, ky[0] - dk
    Nk0 = (Nk - 1) // 2 + 1

    def idx(v):
        return (np.round(np.clip(v, kmin, kmax) / dk)).astype(np.int64).astype(np.int32) + np.int32(Nk0 - 1)

    _xe = me / hbar**2 * np.abs(ky) / (np.abs(ky) + dk * 1e-5) ** 2 / dk
    _xh = mh / hbar**2 * np.abs(ky) / (np.abs(ky) + dk * 1e-5) ** 2 / dk
    K, Q = ky[:, None], ky[None, :]
    _k_p_q = idx(K + Q)
    _k_m_q = idx(K - Q)
    k1p0 = ((me + mh) * Q - 2 * mh * K) / 2.0 / me
    _k1p = idx(k1p0)
    _k1p_m_q = idx(k1p0 - Q)
    k10 = ((me + mh) * Q - 2 * me * K) / 2.0 / mh
    _k1 = idx(k10)
    _k1_m_q = idx(k10 - Q)
    os.makedirs("dataQW/Wire/info", exist_ok=True)
    _fe_file = open("dataQW/Wire/info/MaxOffDiag.e.dat", "w", encoding="utf-8")
    _fh_file = open("dataQW/Wire/info/MaxOffDiag.h.dat", "w", encoding="utf-8")


def Vxx2(q, V):
    """dephasing.py:362-378: V[rint(|q/dq|), 0]^2."""
    dq = q[1] - q[0] if len(q) > 1 else 1.0
    iq = np.round(np.abs(np.asarray(q) / dq)).astype(np.int32)
    return np.asarray(V)[iq, 0] ** 2


def _gamma(hole, ky, ne0, nh0, VC, out):
    if _k_p_q is None:
        raise RuntimeError("InitializeDephasing must be called first")
    N = len(ky)
    if np.shape(VC) != (N, N, 3) or len(ne0) != N or len(nh0) != N:
        raise ValueError("ne0, nh0 need N entries and VC shape (N, N, 3)")
    ky = np.asarray(ky, dtype=float)
    dq = ky[1] - ky[0] if N > 1 else 1.0
    iq = np.round(np.abs(ky / dq)).astype(np.int32)
    if iq.max() >= N:
        raise ValueError("Vxx2: rint(|ky/dq|) outside VC")
    tabs = np.stack([_k_p_q, _k_m_q, _k1_m_q if hole else _k1p_m_q, _k1 if hole else _k1p])
    lib, dev = capi.load(), device()
    ned, nhd, vcd = to_dev(ne0, C128, dev), to_dev(nh0, C128, dev), to_dev(VC, F64, dev)
    iqd, td = to_dev(iq, torch.int32, dev), to_dev(tabs, torch.int32, dev)
    xed, xhd = to_dev(_xe, F64, dev), to_dev(_xh, F64, dev)
    res = torch.empty(N, dtype=F64, device=dev)
    capi.check(lib.psa_mod_gamma(N, hole, ned.data_ptr(), nhd.data_ptr(), vcd.data_ptr(), iqd.data_ptr(),
                                 td.data_ptr(), xed.data_ptr(), xhd.data_ptr(), hbar, res.data_ptr(), stream(dev)))
    out[...] = to_host(res)


def CalcGammaE(ky, ne0, nh0, VC, GammaE):
    """dephasing.py:541-580: electron dephasing rates into GammaE."""
    _gamma(0, ky, ne0, nh0, VC, GammaE)


def CalcGammaH(ky, ne0, nh0, VC, GammaH):
    """dephasing.py:582-625: hole dephasing rates into GammaH."""
    _gamma(1, ky, ne0, nh0, VC, GammaH)


def _offdiag2(ne, nh, p, Ee, Eh, gee, ghh, geh, VC, N):
    if np.shape(VC) != (N, N, 3) or len(ne) != N or len(nh) != N:
        raise ValueError("ne, nh need N entries and VC shape (N, N, 3)")
    lib, dev = capi.load(), device()
    ned, nhd, vcd = to_dev(ne, C128, dev), to_dev(nh, C128, dev), to_dev(VC, F64, dev)
    Ed, Hd = to_dev(Ee, F64, dev), to_dev(Eh, F64, dev)
    pd = to_dev(p, C128, dev) if p is not None else None
    D = torch.empty(2, 2 * N + 1, N, dtype=F64, device=dev)
    x = torch.empty(N, N, dtype=C128, device=dev) if p is not None else None
    capi.check(lib.psa_mod_offdiag2(N, ned.data_ptr(), nhd.data_ptr(), None if pd is None else pd.data_ptr(),
                                    vcd.data_ptr(), Ed.data_ptr(), Hd.data_ptr(), float(gee), float(ghh), float(geh),
                                    hbar, D.data_ptr(), None if x is None else x.data_ptr(), stream(dev)))
    return to_host(D), (None if x is None else to_host(x))


def CalcOffDiagDeph_E2(ne, nh, ky, Ee, Eh, gee, geh, VC, Nk):
    """dephasing.py:381-396: De [2Nk+1][Nk]."""
    return _offdiag2(ne, nh, None, Ee, Eh, gee, gee, geh, VC, Nk)[0][0]


def CalcOffDiagDeph_H2(ne, nh, ky, Ee, Eh, ghh, geh, VC, Nk):
    """dephasing.py:399-414: Dh [2Nk+1][Nk]."""
    return _offdiag2(ne, nh, None, Ee, Eh, ghh, ghh, geh, VC, Nk)[0][1]


def OffDiagDephasing2(ne, nh, p, ky, Ee, Eh, g, VC, t, x):
    """dephasing.py:714-743: x = i hbar (Dh p + p De) with g = (gee, ghh,
    geh) as the reference passes it; appends t, max, min of each table to
    the MaxOffDiag files opened by InitializeDephasing."""
    N = len(ky)
    D, xd = _offdiag2(ne, nh, p, Ee, Eh, g[0], g[1], g[2], VC, N)
    x[...] = xd
    for f, tab in ((_fe_file, D[0]), (_fh_file, D[1])):
        if f is not None:
            f.write(f"{np.real(t)} {np.max(np.real(tab))} {np.min(np.real(tab))}\n")
            f.flush()


def printGam(Dx, z, n, file):
    """dephasing.py:417-427: z and Dx per line as float32,
    dataQW/<file><n:05d>.dat."""
    fn = f"dataQW/{file}{n:05d}.dat"
    os.makedirs(os.path.dirname(fn), exist_ok=True)
    with open(fn, "w", encoding="utf-8") as f:
        for zi, di in zip(z, Dx):
            f.write(f"{np.float32(zi)} {np.float32(di)}\n")


def WriteDephasing(ky, gamE, gamH, w, xxx):
    """dephasing.py:417-437: dataQW/Wire/Ge|Gh/G?.<ww>.k.<xxxxx>.dat."""
    wire = f"{w:02d}"
    for name, g in ((f"Wire/Ge/Ge.{wire}.k.", gamE), (f"Wire/Gh/Gh.{wire}.k.", gamH)):
        printGam(g, ky, xxx, name)


def _od1_tables(hole, ne0, nh0, Ee, Eh, gss, gso, VC, dev):
    N = len(ne0)
    if _k_p_q is None:
        raise RuntimeError("InitializeDephasing must be called first")
    if np.shape(VC) != (N, N, 3) or len(nh0) != N or np.shape(_k_p_q) != (N, N):
        raise ValueError("ne0, nh0 need N entries, VC shape (N, N, 3), the index tables (N, N)")
    lib = capi.load()
    ned, nhd = to_dev(ne0, C128, dev), to_dev(nh0, C128, dev)
    vcd, Ed, Hd = to_dev(VC, F64, dev), to_dev(Ee, F64, dev), to_dev(Eh, F64, dev)
    kp = torch.as_tensor(np.ascontiguousarray(_k_p_q, dtype=np.int32), device=dev)
    km = torch.as_tensor(np.ascontiguousarray(_k_m_q, dtype=np.int32), device=dev)
    D = torch.empty(N, N, dtype=F64, device=dev)
    capi.check(lib.psa_mod_offdiag1(N, int(hole), ned.data_ptr(), nhd.data_ptr(), vcd.data_ptr(), Ed.data_ptr(),
                                    Hd.data_ptr(), kp.data_ptr(), km.data_ptr(), float(gss), float(gso), hbar,
                                    D.data_ptr(), stream(dev)))
    return D, kp, (ned, nhd, vcd, Ed, Hd, km)


def CalcOffDiagDeph_E(ne0, nh0, ky, Ee0, Eh0, gee, geh, VC):
    """dephasing.py:627-655 (_CalcOffDiagDeph_E_jit 107-147): the version-1
    electron table D[k, q], returned as a new (Nk, Nk) array."""
    return to_host(_od1_tables(0, ne0, nh0, Ee0, Eh0, gee, geh, VC, device())[0])


def CalcOffDiagDeph_H(ne0, nh0, ky, Ee0, Eh0, ghh, geh, VC):
    """dephasing.py:657-685 (_CalcOffDiagDeph_H_jit 150-190): the version-1
    hole table D[k, q], returned as a new (Nk, Nk) array."""
    return to_host(_od1_tables(1, ne0, nh0, Ee0, Eh0, ghh, geh, VC, device())[0])


def OffDiagDephasing(ne, nh, p, ky, Ee, Eh, g, VC, x):
    """dephasing.py:687-712 (_OffDiagDephasing_jit 193-218): x = i hbar (the
    Dh contraction over qp + the De contraction over q) with g = (gee, ghh,
    geh), written into x."""
    N = len(ky)
    if np.shape(p) != (N, N) or np.shape(x) != (N, N):
        raise ValueError("p and x need shape (Nk, Nk)")
    dev = device()
    DE, kp, keep = _od1_tables(0, ne, nh, Ee, Eh, g[0], g[2], VC, dev)
    DH, _, keep2 = _od1_tables(1, ne, nh, Ee, Eh, g[1], g[2], VC, dev)
    pd, kyd = to_dev(p, C128, dev), to_dev(ky, F64, dev)
    xd = torch.empty(N, N, dtype=C128, device=dev)
    capi.check(capi.load().psa_mod_offdiag1_apply(N, DE.data_ptr(), DH.data_ptr(), pd.data_ptr(), kp.data_ptr(),
                                                  kyd.data_ptr(), hbar, xd.data_ptr(), stream(dev)))
    x[...] = to_host(xd)
    del keep, keep2  # the inputs of the queued launches lived until here
