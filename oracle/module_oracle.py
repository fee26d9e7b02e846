"""CPU FP64 oracle for the reference module functions that sit off the
QWCalculator path (coulomb, emission, qwoptics free functions).

TEST INFRASTRUCTURE ONLY.  Nothing in the ``pulsesuite_amd`` package imports
this module; only ``tests/`` use it, as the checker.

Each function restates the reference loop it cites (paths relative to
``src/pulsesuite/``) with the reference's summation order where the sum is
sequential, so the restatement agrees with the reference's pure-Python
loops bit for bit; NumPy's pairwise ``np.sum`` (PLSpectrum's photon sums
and temperatures) is followed by calling ``np.sum`` the same way.

Pinning (``tests/test_module_oracle.py``): every function is checked
against ``tests/golden/ref_modx.npz`` (written from the reference by
``tests/golden/make_golden.py extras``) and against what the mathematics
fixes: the renormalisations' algebraic reduction to matrix-vector
products, GetVn1n2 as diagonal-matrix products, Vehint against a
scipy.special.k0 double sum, CalcMVeh against an independent einsum over
the momentum-conservation delta.
"""

import numpy as np
from scipy.constants import c as c0
from scipy.constants import e as e0
from scipy.constants import epsilon_0 as eps0
from scipy.constants import hbar
from scipy.constants import k as kB

from .sbe_oracle import ec, k0_accurate, rho0

pi = np.pi
ii = 1j
eV = e0
SMALL = 1e-100


def calc_mveh(p, VC, k3, UnDel):
    """CalcMVeh (coulomb.py:1234-1241) -> _CalcMVeh_core (786-810):
    MVeh[k, kp, f] = sum_q p[q, qp, f] Veh[k, q] UnDel[k+1, q+1] UnDel[kp+1, qp+1]
    with qp + 1 = k3[kp, q, k] (terms with k3 = 0 skipped), q ascending."""
    N, _, Nf = p.shape
    Veh = VC[:, :, 0]
    M = np.zeros((N, N, Nf), dtype=complex)
    k = np.arange(N)[:, None]
    kp = np.arange(N)[None, :]
    for q in range(N):
        qp1 = k3[kp, q, k]  # [k, kp]
        ok = qp1 > 0
        qp = np.where(ok, qp1 - 1, 0)
        term = (p[q, qp, :] * Veh[k, q][..., None] * UnDel[k + 1, q + 1][..., None]
                * UnDel[kp + 1, qp + 1][..., None])
        M = np.where(ok[..., None], M + term, M)
    return M


def bg_renorm(C, D, VC, UnDel):
    """BGRenorm (coulomb.py:1248-1259) -> _BGRenorm_core (813-831):
    BGR[k, kp] = -sum_i nh_i Vhh[i, kp] U[i+1, kp+1] - sum_i ne_i Vee[i, k] U[i+1, k+1]."""
    ne, nh = np.diag(C), np.diag(D)
    Vee, Vhh = VC[:, :, 1], VC[:, :, 2]
    N = len(ne)
    shh = np.zeros(N, dtype=complex)
    see = np.zeros(N, dtype=complex)
    for i in range(N):
        shh = shh + nh[i] * Vhh[i, :] * UnDel[i + 1, 1:]
        see = see + ne[i] * Vee[i, :] * UnDel[i + 1, 1:]
    return -shh[None, :] - see[:, None]


def carrier_renorm(n, V, UnDel):
    """EeRenorm / EhRenorm (coulomb.py:1261-1280) -> _EeRenorm_core,
    _EhRenorm_core (834-897): with a = sum_i n_i V[kp, kp], b = sum_i n_i V[i, kp] U[i+1, kp+1],
    c = sum_i n_i V[k, k], d = sum_i n_i V[i, k] U[i+1, k+1]:
    BGR[k, kp] = 2a - b + 2c - d - 2a - c, evaluated left to right."""
    N = len(n)
    dg = np.diag(V)
    a = np.zeros(N, dtype=complex)
    b = np.zeros(N, dtype=complex)
    for i in range(N):
        a = a + n[i] * dg
        b = b + n[i] * V[i, :] * UnDel[i + 1, 1:]
    # a, b index kp; the same sums index k for c, d
    A, B = a[None, :], b[None, :]
    Cc, Dd = a[:, None], b[:, None]
    return 2.0 * A - B + 2.0 * Cc - Dd - 2.0 * A - Cc


def vn1n2(rcv, Hcc, Hhh, Hcv):
    """GetVn1n2 (qwoptics.py:1149-1186) -> _GetVn1n2_jit (522-542):
    Hvv = -Hhh^T, Hvc = Hcv^H, rvc = conj(rcv); returns Vcc, Vvv, Vcv, Vvc with
    Vvc = Vcv^H."""
    Hvv = -Hhh.T
    Hvc = np.conj(Hcv.T)
    rvc = np.conj(rcv)
    m = -ii / hbar
    Vcv = m * (rcv[:, None] * Hvv - Hcc * rcv[None, :])
    Vcc = m * (rcv[:, None] * Hvc - Hcv * rvc[None, :])
    Vvv = m * (rvc[:, None] * Hcv - Hvc * rcv[None, :])
    return Vcc, Vvv, Vcv, np.conj(Vcv.T)


def vehint(k, q, y, ky, ae, ah, Delta0):
    """Vehint (coulomb.py:180-242): the Vint double sum over the middle half
    of y with dk = max(|ky[k-1] - ky[q-1]|, (ae + ah)/4) (index 0 for k or
    q = 0) and the accurate K0 (usefulsubs.K03), summed row by row."""
    Ny = len(y)
    aey2 = (ae * y) ** 2
    ahy2 = (ah * y) ** 2
    ki = k - 1 if k > 0 else 0
    qi = q - 1 if q > 0 else 0
    kmin = (ae + ah) / 4.0
    dk = max(abs(ky[ki] - ky[qi]), kmin)
    mult = ae * ah / pi * (y[1] - y[0]) ** 2
    N1, N2 = Ny // 4, 3 * Ny // 4
    v = 0.0
    for i in range(N1, N2 + 1):
        r = np.sqrt((y[i] - y[N1:N2 + 1]) ** 2 + Delta0**2)
        ex = np.exp(-aey2[i] - ahy2[N1:N2 + 1])
        for j, val in enumerate(ex * mult * k0_accurate(dk * r)):
            v += val
    return v


def chi1dqw(ae, ah, Delta0, epsr, game, gamh, ky, Ee, Eh, ne, nh, qq, w):
    """GetChi1Dqw (coulomb.py:633-709): returns (Re chi, Im chi)."""
    Nk = len(ky)
    dk = ky[1] - ky[0]
    hge = np.maximum(hbar * game, 1e-4 * eV)
    hgh = np.maximum(hbar * gamh, 1e-4 * eV)
    qmine, qminh = ae / 2.0, ah / 2.0
    Re = np.sqrt((2.0 / ae) ** 2 + Delta0**2)
    Rh = np.sqrt((2.0 / ah) ** 2 + Delta0**2)
    beta = e0**2 / (4 * pi * eps0 * epsr)
    ql = abs(qq)
    Ke = k0_accurate(max(ql, qminh) * Re)[0] * beta * 2
    Kh = k0_accurate(max(ql, qmine) * Rh)[0] * beta * 2
    chi = 0.0 + 0.0j
    q = int(round(qq / dk))
    for k in range(max(0, -q), min(Nk, Nk - q)):
        kq = k + q
        de = hbar * w - Ee[kq] + Ee[k] + ii * (hge[kq] + hge[k])
        if abs(de) > 1e-300:
            chi -= Ke / pi * (ne[k] - ne[kq]) / de * dk
        dh = hbar * w - Eh[kq] + Eh[k] + ii * (hgh[kq] + hgh[k])
        if abs(dh) > 1e-300:
            chi -= Kh / pi * (nh[k] - nh[kq]) / dh * dk
    return np.real(chi), np.imag(chi)


def eps1dqw(ae, ah, Delta0, epsr, me, mh, n1D, q, w):
    """GetEps1Dqw (coulomb.py:712-776): returns (Re eps, Im eps)."""
    qmine, qminh = ae / 2.0, ah / 2.0
    Re = np.sqrt((2.0 / ae) ** 2 + Delta0**2)
    Rh = np.sqrt((2.0 / ah) ** 2 + Delta0**2)
    beta = e0**2 / (4 * pi * eps0 * epsr)
    if abs(q) < 1.0:
        q = 1.0
    ql = abs(q)
    Ce = 2 * beta * me / pi / hbar**2 / ql
    Ch = 2 * beta * mh / pi / hbar**2 / ql
    Ke = k0_accurate(max(ql, qminh) * Re)[0]
    Kh = k0_accurate(max(ql, qmine) * Rh)[0]
    oep = hbar * ql / 2.0 / me * abs(ql + pi * n1D)
    oem = hbar * ql / 2.0 / me * abs(ql - pi * n1D)
    ohp = hbar * ql / 2.0 / mh * abs(ql + pi * n1D)
    ohm = hbar * ql / 2.0 / mh * abs(ql - pi * n1D)
    epr = (1.0 - Ce * Ke * np.log(abs((w**2 - oem**2) / (w**2 - oep**2)))
           - Ch * Kh * np.log(abs((w**2 - ohm**2) / (w**2 - ohp**2))))
    epi = 0.0
    if min(oem, oep) < w < max(oem, oep) and min(ohm, ohp) < w < max(ohm, ohp):
        epi = -Ce * Ke - Ch * Kh
    return epr, epi


def calchw(Nw, Estart, Emax):
    """Calchw (emission.py:258-288): hw[w] = Estart + w (Emax - Estart)/Nw."""
    dhw = (Emax - Estart) / float(Nw)
    return np.array([Estart + w * dhw for w in range(Nw)])


def _lrtz(a, b):
    return (b / pi) / (a**2 + b**2)


def _softtheta(x, g):
    return 0.5 * (1.0 + 2.0 / pi * np.arctan(x / g))


def _interp(f, x0):
    """LinearInterp_dp/_dpc (libpulsesuite/helpers.py:951-995) on the grid
    ky = 0, 1, ..., N-1 with locator's clamp to [0, N-2] (usefulsubs.py:1803-1830)."""
    N = len(f)
    i = np.clip(np.searchsorted(np.arange(N, dtype=float), x0, side="right") - 1, 0, N - 2)
    xi, xi1 = i.astype(float), (i + 1).astype(float)
    return (f[i] * (xi1 - x0) + f[i + 1] * (x0 - xi)) / (xi1 - xi)


def pl_spectrum(ne, nh, Ee, Eh, gap, geh, VC, hw, t, PLS, rscale):
    """PLSpectrum (emission.py:291-361): returns the updated PLS."""
    Nk = len(ne)
    X = 100
    Ek = gap + Ee + Eh + ec(np.abs(ne), np.abs(nh), VC)
    Te = 2.0 * np.real(np.sum(ne * Ee) / (np.sum(ne) + SMALL)) / kB
    Th = 2.0 * np.real(np.sum(nh * Eh) / (np.sum(nh) + SMALL)) / kB
    T = (Te + Th) / 2.0
    qy = np.arange(X * Nk) / float(X)
    E = _interp(Ek, qy)
    nenh = np.abs(_interp(ne * nh, qy))
    out = np.array(PLS, dtype=float).copy()
    for w in range(len(hw)):
        out[w] = out[w] + rscale * np.sum(hw[w] * rho0(hw[w]) * nenh * np.exp(-np.abs(hw[w] - E) / kB / T)
                                          * _lrtz(hw[w] - E, hbar * geh)
                                          * _softtheta(hw[w] - E[X * Nk // 2], hbar * geh))
    return out * _softtheta(t, geh)


def offdiag_deph1(ne0, nh0, VC, Ee, Eh, kpq, kmq, gss, gso, hole):
    """CalcOffDiagDeph_E (hole False) / CalcOffDiagDeph_H (hole True)
    (dephasing.py:107-190, 627-685) on 0-based index tables (an entry
    outside [0, N) drops the term, as the reference's extended arrays do):
    D[k, q] = pi/hbar [sum_k1 Vss(k1, a)^2 L(Es[a] + Es[k] - Es[k1] - Es[b], hbar gss)
              (ns[a] ns[k] (1 - ns[k1]) + (1 - ns[a]) (1 - ns[k]) ns[k1])
            + sum_k1 Vso^2 L(Eo[c] + Es[k] - Eo[k1] - Es[b], hbar gso)
              (no[c] (1 - no[k1]) ns[k] + (1 - no[c]) no[k1] (1 - ns[k]))]
    with b = kpq[k, q], a = kpq[k1, q], c = kmq[k1, q]; Vso = Veh at (k, b)
    for electrons and at (k1, c) for holes; the same-carrier sum runs first,
    both over k1 ascending."""
    N = len(ne0)
    ns, no = (np.abs(nh0), np.abs(ne0)) if hole else (np.abs(ne0), np.abs(nh0))
    Es, Eo = (Eh, Ee) if hole else (Ee, Eh)
    Vss2 = VC[:, :, 2 if hole else 1] ** 2
    Veh2 = VC[:, :, 0] ** 2
    hs, ho = hbar * gss, hbar * gso
    D = np.zeros((N, N))
    k = np.arange(N)[:, None]
    b = kpq                        # [k, q]
    okb = (b >= 0) & (b < N)
    bb = np.where(okb, b, 0)
    for k1 in range(N):
        a = kpq[k1][None, :]       # [1, q]
        ok = okb & (a >= 0) & (a < N)
        aa = np.where(ok, a, 0)
        Ed = Es[aa] + Es[k] - Es[k1] - Es[bb]
        occ = ns[aa] * ns[k] * (1.0 - ns[k1]) + (1.0 - ns[aa]) * (1.0 - ns[k]) * ns[k1]
        D = np.where(ok, D + Vss2[k1, aa] * _lrtz(Ed, hs) * occ, D)
    for k1 in range(N):
        c = kmq[k1][None, :]
        ok = okb & (c >= 0) & (c < N)
        cc = np.where(ok, c, 0)
        Ed = Eo[cc] + Es[k] - Eo[k1] - Es[bb]
        occ = no[cc] * (1.0 - no[k1]) * ns[k] + (1.0 - no[cc]) * no[k1] * (1.0 - ns[k])
        v2 = Veh2[k1, cc] if hole else Veh2[k, bb]
        D = np.where(ok, D + v2 * _lrtz(Ed, ho) * occ, D)
    return D * pi / hbar


def offdiag_dephasing1(ne, nh, p, ky, Ee, Eh, g, VC, kpq, kmq):
    """OffDiagDephasing (dephasing.py:193-218, 687-712):
    x[k, kp] = i hbar [sum_qp DH[kp, qp] pt[kpq[qp, kp] + 1, k] u[qp]
                       + sum_q DE[k, q] pp[kpq[q, k] + 1, kp] u[q]],
    u = |ky| / (|ky| + 1e-10), DE, DH the tables with (gee, geh), (ghh, geh).
    pt and pp are p and p^T placed at [1:N+1, 1:N+1] of (N+2)^2 arrays and
    read with the 0-based column index, as the reference does: the second
    index is shifted by one, pt[i + 1, k] = p[i, k - 1] (0 for k = 0) and
    pp[i + 1, kp] = p[kp - 1, i] (0 for kp = 0)."""
    N = len(ky)
    DE = offdiag_deph1(ne, nh, VC, Ee, Eh, kpq, kmq, g[0], g[2], False)
    DH = offdiag_deph1(ne, nh, VC, Ee, Eh, kpq, kmq, g[1], g[2], True)
    u = np.abs(ky) / (np.abs(ky) + 1e-10)
    # p padded with a zero row and column in front: ps[i + 1, j + 1] = p[i, j]
    ps = np.zeros((N + 1, N + 1), dtype=complex)
    ps[1:, 1:] = p
    x = np.zeros((N, N), dtype=complex)
    k = np.arange(N)[:, None]
    kp = np.arange(N)[None, :]
    for qp in range(N):
        r = kpq[qp][None, :]       # kpq[qp, kp]
        ok = (r >= 0) & (r < N)
        rr = np.where(ok, r, 0)
        x = np.where(ok, x + DH[kp, qp] * ps[rr + 1, k] * u[qp], x)
    for q in range(N):
        r = kpq[q][:, None]        # kpq[q, k]
        ok = (r >= 0) & (r < N)
        rr = np.where(ok, r, 0)
        x = np.where(ok, x + DE[k, q] * ps[kp, rr + 1] * u[q], x)
    return x * ii * hbar


__all__ = ["calc_mveh", "bg_renorm", "carrier_renorm", "vn1n2", "vehint", "chi1dqw", "eps1dqw",
           "calchw", "pl_spectrum", "offdiag_deph1", "offdiag_dephasing1", "c0"]
