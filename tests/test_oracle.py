"""Pins the CPU oracle: golden vectors from the reference, closed forms,
library routines and brute force at tiny N.  CPU only."""

import os

import numpy as np
import pytest
from scipy.special import k0 as scipy_k0

from oracle import sbe_oracle as O

TAGS = ["n16_mb", "n16_rhs", "n32_mb"]


def rel(a, b):
    a = np.asarray(a)
    b = np.asarray(b)
    return np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300)


def load(golden_dir, tag):
    return np.load(os.path.join(golden_dir, f"ref_{tag}.npz"))


def setup(z):
    prm = O.Params(L=float(z["L"]), me_rel=0.067, flags=tuple(z["mb"]))
    g = O.Grid(prm)
    return prm, g, O.Tables(g)


@pytest.mark.parametrize("tag", TAGS)
def test_grid_and_tables_match_reference(golden_dir, tag):
    z = load(golden_dir, tag)
    prm, g, t = setup(z)
    assert g.Nk == int(z["nk"]) and g.Nr == int(z["nr"])
    for k in ("kr", "r", "Qr", "Ee", "Eh"):
        assert rel(getattr(g, k), z[k]) == 0.0, k
    assert np.array_equal(g.kkp, z["kkp"])
    assert g.NQ0 == int(z["NQ0"])
    for k in ("dcv", "area", "ehint", "dr", "alphae", "alphah", "qc", "dkr"):
        assert rel(getattr(g, k), z[k]) == 0.0, k
    # Vint with the crude _K03_jit, summation order differs: ulp level
    for i, k in enumerate(("Veh0", "Vee0", "Vhh0")):
        assert rel(t.V0[i], z[k]) < 1e-14, k
    # lower-triangular: negative kkp entries are dropped (coulomb.py:567-573)
    assert np.all(np.triu(t.V0[0], 1) == 0.0)
    assert rel(t.qe, z["qe"]) == 0 and rel(t.qh, z["qh"]) == 0
    assert rel(t.Chi1De, z["Chi1De"]) < 1e-12
    assert rel(t.Chi1Dh, z["Chi1Dh"]) < 1e-12
    for k, v in (("EP", t.EP), ("HP", t.HP), ("xe", t.xe), ("xh", t.xh),
                 ("Expikr", t.Expikr), ("QWWindow", t.window)):
        assert rel(v, z[k]) < 1e-15, k
    for k in ("k_p_q", "k_m_q", "k1_m_q", "k1p_m_q", "k1", "k1p"):
        assert np.array_equal(getattr(t, k), z[k]), k
    assert rel(t.Vscale, z["Vscale"]) == 0 and rel(t.NO, z["NO"]) == 0
    if "HOmega" in z.files:
        assert rel(t.HOmega, z["HOmega"]) == 0
        assert rel(t.square, z["square"]) < 1e-15
    if "Ceh" in z.files:
        assert rel(O.mb_tables(g, "eh", prm.gam_e, prm.gam_h), z["Ceh"]) < 1e-15
        assert rel(O.mb_tables(g, "ee", prm.gam_e, prm.gam_h), z["Cee"]) < 1e-15
        assert rel(O.mb_tables(g, "hh", prm.gam_e, prm.gam_h), z["Chh"]) < 1e-15


@pytest.mark.parametrize("tag", TAGS)
def test_step_functions_match_reference(golden_dir, tag):
    z = load(golden_dir, tag)
    prm, g, t = setup(z)
    C, D, p = z["fC"], z["fD"], z["fp"]
    ne = np.diag(C).copy()
    nh = np.diag(D).copy()
    VC, E1D = O.screened_arrays(g, t, True, ne, nh)
    assert rel(VC, z["fVC"]) < 1e-14
    assert rel(E1D, z["fE1D"]) < 1e-14
    VC = z["fVC"]
    Heh, Hee, Hhh = O.calc_h(g, z["fMeh"], C, D, p, VC, True, True)
    assert rel(Heh, z["fHeh"]) < 1e-14
    # Hee/Hhh carry the band energies on the diagonal: compare the
    # Coulomb (off-diagonal) part relative to its own size as well
    off = lambda m: m - np.diag(np.diag(m))  # noqa: E731
    assert rel(Hee, z["fHee"]) < 1e-14 and rel(off(Hee), off(z["fHee"])) < 1e-14
    assert rel(Hhh, z["fHhh"]) < 1e-14 and rel(off(Hhh), off(z["fHhh"])) < 1e-14
    assert rel(O.gamma_e(g, t, ne, nh, VC), z["fGamE"]) < 1e-14
    assert rel(O.gamma_h(g, t, ne, nh, VC), z["fGamH"]) < 1e-14
    assert rel(O.offdiag_dephasing2(g, ne, nh, p, VC)[0], z["fOffP"]) < 1e-14
    S_eh = O.mb_tables(g, "eh", prm.gam_e, prm.gam_h)
    S_ee = O.mb_tables(g, "ee", prm.gam_e, prm.gam_h)
    S_hh = O.mb_tables(g, "hh", prm.gam_e, prm.gam_h)
    a, b = O.mbce(g, S_eh, S_ee, np.real(ne), np.real(nh), VC)
    assert rel(a, z["fWinE"]) < 1e-14 and rel(b, z["fWoutE"]) < 1e-14
    a, b = O.mbch(g, S_eh, S_hh, np.real(ne), np.real(nh), VC)
    assert rel(a, z["fWinH"]) < 1e-14 and rel(b, z["fWoutH"]) < 1e-14
    re, rh = O.qw_rho5(g, t, C, D, ne, nh)
    assert rel(re, z["fre"]) < 1e-14 and rel(rh, z["frh"]) < 1e-14


def run_golden_trajectory(z):
    prm, g, t = setup(z)
    s = O.Solver(prm, 2, z["rr"], float(z["Emax0"]), tables=t, grid=g)
    for nm, dst in (("C2", s.CC2), ("D2", s.DD2), ("p2", s.YY2),
                    ("C3", s.CC3), ("D3", s.DD3), ("p3", s.YY3)):
        dst[:] = z["s0_" + nm]
    dt = float(z["dt"])
    outs = []
    states = []
    for it in range(int(z["nsteps"])):
        Exx = z["t_Exx"][it]
        Eyy = z["t_Eyy"][it]
        Ezz = np.zeros_like(Exx)
        for w in (1, 2):
            outs.append(s.qw_calculator(Exx, Eyy, Ezz, dt, w))
        states.append((s.CC2.copy(), s.DD2.copy(), s.YY2.copy(), s.CC3.copy(),
                       s.DD3.copy(), s.YY3.copy(), s.I0.copy()))
    return outs, states


# trajectories also with LF = 0 (diagonal-only transport, no QWRho5, no
# Hartree terms) and LorentzDelta = 1 (Lorentzian many-body tables)
@pytest.mark.parametrize("tag", TAGS + ["n16_nolf", "n16_lztraj"])
def test_qwcalculator_trajectory_matches_reference(golden_dir, tag):
    z = load(golden_dir, tag)
    outs, states = run_golden_trajectory(z)
    for j, (Pxx, Pyy, Pzz, Rho) in enumerate(outs):
        assert rel(Pxx, z["t_Pxx"][j]) < 1e-13
        assert rel(Pyy, z["t_Pyy"][j]) < 1e-13
        assert rel(Pzz, z["t_Pzz"][j]) < 1e-13
        # Rho = RhoH - RhoE cancels: 1e-12 relative to max|Rho|
        assert rel(Rho, z["t_Rho"][j]) < 1e-12
    for it, st in enumerate(states):
        for v, nm in zip(st, ("C2", "D2", "p2", "C3", "D3", "p3", "I0")):
            assert rel(v, z["t_" + nm][it]) < 1e-13, (it, nm)


# ---------------------------------------------------------------------------
# Closed forms, library routines, brute force
# ---------------------------------------------------------------------------

def test_transport_lf0_diagonal_shift(golden_dir):
    """LF = 0 transport (ShiftN1D of the diagonal, dcfield.py:1154-1174,
    1229-1232) against the 2-D shift the golden trajectories pin: the 2-D
    shift of a rank-one matrix u v^T is S(u) S(v)^T with S the 1-D shift, so
    a wrong sign or bin order in S fails; off-diagonal entries are kept, a
    zero field is the identity and S is unitary."""
    z = load(golden_dir, "n16_mb")
    prm, g, t = setup(z)
    N = g.Nk
    rng = np.random.default_rng(5)
    Edc, Eac, dt = 1e6, 3e7, 2e-15

    def S(u):
        return np.diag(O.transport(t, np.diag(u), Edc, Eac, dt, lf=False))

    u = rng.standard_normal(N) + 1j * rng.standard_normal(N)
    v = rng.standard_normal(N) + 1j * rng.standard_normal(N)
    two_d = O.transport(t, np.outer(u, v), Edc, Eac, dt, lf=True)
    assert rel(two_d, np.outer(S(u), S(v))) < 1e-13
    assert abs(np.linalg.norm(S(u)) / np.linalg.norm(u) - 1.0) < 1e-13
    X = rng.standard_normal((N, N)) + 1j * rng.standard_normal((N, N))
    Y = O.transport(t, X, Edc, Eac, dt, lf=False)
    off = ~np.eye(N, dtype=bool)
    assert np.array_equal(Y[off], X[off])
    assert rel(np.diag(Y), S(np.diag(X))) < 1e-15
    assert rel(O.transport(t, X, 0.0, 0.0, dt, lf=False), X) < 1e-14
    # the shift moves weight: not the identity at a finite field
    assert rel(np.diag(Y), np.diag(X)) > 1e-3


def test_k0_accurate_vs_scipy():
    # Zhang-Jin CIK01: series for x <= 9 loses ~1e-9 to cancellation near
    # x = 9 and the I0K0 asymptotic is truncated at 10 terms just above 9;
    # elsewhere it is at the 1e-13 level.
    x = np.concatenate([np.linspace(0.05, 8.0, 50), np.linspace(12, 99, 40)])
    assert rel(O.k0_accurate(x), scipy_k0(x)) < 1e-12
    x = np.linspace(8.0, 12.0, 41)
    assert np.max(np.abs(O.k0_accurate(x) / scipy_k0(x) - 1)) < 1e-8
    assert O.k0_accurate(np.array([150.0]))[0] == 0.0


def test_k03_jit_branches():
    # the crude _K03_jit: constants and the large-x asymptote are exact
    assert O.k03_jit(np.array([1e-12]))[0] == 23.0
    assert O.k03_jit(np.array([101.0]))[0] == 0.0
    x = np.array([2.0, 5.0, 40.0])
    assert rel(O.k03_jit(x), np.sqrt(np.pi / (2 * x)) * np.exp(-x)) == 0
    # its series is -ln(x/2) - gamma + sum_k (x^2/4)^k/(k!)^2 (1/k - gamma):
    # brute force with exact factorials
    from math import factorial, log
    for xv in (0.3, 1.0, 1.9):
        s = -log(0.5 * xv) - O.EULER_GAMMA
        for k in range(1, 20):
            term = (0.25 * xv * xv) ** k / factorial(k) ** 2
            s += term * (1.0 / k - O.EULER_GAMMA)
            if abs(term) < 1e-15:
                break
        assert abs(O.k03_jit(np.array([xv]))[0] - s) < 1e-14 * abs(s)


def test_fftg_is_scaled_numpy_fft():
    rng = np.random.default_rng(0)
    f = rng.standard_normal(34) + 1j * rng.standard_normal(34)
    assert rel(O.fftg(f), -np.fft.fft(f) / 34) < 1e-15
    assert rel(O.ifftg(O.fftg(f)), f) < 1e-14


def test_spline_exact_on_cubics():
    # the FMM end conditions match third derivatives: exact on cubics
    x0 = np.sort(np.random.default_rng(1).uniform(-1, 1, 40))
    x0[0], x0[-1] = -1.0, 1.0
    f = lambda x: 1 - 2 * x + 0.5 * x**2 - 0.7 * x**3  # noqa: E731
    x1 = np.linspace(-1.2, 1.2, 77)
    out = O.rescale_1d(x0, f(x0) + 0j, x1)
    inside = (x1 >= -1) & (x1 <= 1)
    assert np.max(np.abs(out[inside] - f(x1[inside]))) < 1e-12
    assert np.all(out[~inside] == 0)


def test_grid_size_pinned_by_docs_fixture():
    # docs/source/examples/dataQW/Wire/Xqw/Xqw.params:1 records Nq = Nk+1 =
    # 115 for the shipped L = 100 nm, me = 0.07
    g = O.Grid(O.Params())
    assert g.Nk == 114 and g.Nr == 230


def test_rhs_restated_with_loops_at_tiny_n():
    """dpdt/dCdt/dDdt (SBEs.py:1941-2625) element by element."""
    rng = np.random.default_rng(3)
    n = 5
    r = lambda: rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n))  # noqa: E731
    C, D, p, Heh, Hee, Hhh, OffP = r(), r(), r(), r(), r(), r(), r()
    GE, GH = rng.uniform(size=n), rng.uniform(size=n)
    ih = 1j * O.hbar
    dp = np.zeros((n, n), complex)
    dC = np.zeros((n, n), complex)
    dD = np.zeros((n, n), complex)
    for ke in range(n):
        for kh in range(n):
            s = sum(Hhh[kh, k] * p[k, ke] + Hee[ke, k] * p[kh, k]
                    - Heh[k, kh] * C[k, ke] - Heh[ke, k] * D[k, kh]
                    for k in range(n))
            dp[kh, ke] = (s + Heh[ke, kh] - ih * (GE[ke] + GH[kh]) * p[kh, ke]
                          + OffP[kh, ke]) / ih
    for k1 in range(n):
        for k2 in range(n):
            dC[k1, k2] = sum(Hee[k2, k] * C[k1, k] - Hee[k, k1] * C[k, k2]
                             + Heh[k2, k] * np.conj(p[k, k1])
                             - np.conj(Heh[k1, k]) * p[k, k2] for k in range(n)) / ih
            dD[k1, k2] = sum(Hhh[k2, k] * D[k1, k] - Hhh[k, k1] * D[k, k2]
                             + Heh[k, k2] * np.conj(p[k1, k])
                             - np.conj(Heh[k, k1]) * p[k2, k] for k in range(n)) / ih
    assert rel(O.dpdt(C, D, p, Heh, Hee, Hhh, GE, GH, OffP), dp) < 1e-14
    assert rel(O.dCdt(C, D, p, Heh, Hee, Hhh), dC) < 1e-14
    assert rel(O.dDdt(C, D, p, Heh, Hee, Hhh), dD) < 1e-14


def test_rhs_invariants():
    """Hermitian C, D, H => dC/dt, dD/dt Hermitian; charge neutrality of
    the optical generation: d(tr C)/dt = d(tr D)/dt."""
    z = None
    rng = np.random.default_rng(4)
    n = 12

    def herm():
        a = rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n))
        return a + a.conj().T

    C, D, Hee, Hhh = herm(), herm(), herm(), herm()
    p = rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n))
    Heh = rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n))
    dC = O.dCdt(C, D, p, Heh, Hee, Hhh)
    dD = O.dDdt(C, D, p, Heh, Hee, Hhh)
    assert rel(dC, dC.conj().T) < 1e-13
    assert rel(dD, dD.conj().T) < 1e-13
    assert abs(np.trace(dC) - np.trace(dD)) < 1e-12 * np.max(np.abs(dC))
    del z


def test_relaxation_update_is_exact_solution():
    """SBEs.py:5312-5321 is the exact solution of dn/dt = Win(1-n) - Wout n
    over dt."""
    Win, Wout, n0, dt = 3e12, 5e12, 0.2, 1e-14
    W = Win + Wout
    n = np.exp(-W * dt) * ((-1 + np.exp(W * dt) + n0) * Win + n0 * Wout) / (W + 1e-200)
    exact = Win / W + (n0 - Win / W) * np.exp(-W * dt)
    assert abs(n - exact) < 1e-14


def test_unscreened_eps_is_one():
    """coulomb.py:1215-1222: screened = False leaves E1D = 1 and VC = V0."""
    g = O.Grid(O.Params(L=O.Params().L / 3))
    t = O.Tables(g)
    ne = np.full(g.Nk, 0.3 + 0j)
    VC, E1D = O.screened_arrays(g, t, False, ne, ne)
    assert np.all(E1D == 1.0)
    assert np.array_equal(VC[:, :, 0], t.V0[0])


def test_mbce_brute_force_loops_tiny_n(golden_dir):
    """_MBCE_core restated as literal loops over 1-based padded arrays."""
    z = load(golden_dir, "n16_mb")
    prm, g, t = setup(z)
    N = g.Nk
    S_eh = O.mb_tables(g, "eh", prm.gam_e, prm.gam_h)
    S_ee = O.mb_tables(g, "ee", prm.gam_e, prm.gam_h)
    ne0 = np.real(np.diag(z["fC"]))
    nh0 = np.real(np.diag(z["fD"]))
    VC = z["fVC"]
    ne = np.zeros(N + 1)
    nh = np.zeros(N + 1)
    ne[1:] = np.abs(ne0)
    nh[1:] = np.abs(nh0)
    k3 = z["k3"]
    Veh2 = VC[:, :, 0] ** 2
    Vee2 = VC[:, :, 1] ** 2
    Win = np.zeros(N)
    Wout = np.zeros(N)
    for k in range(1, N + 1):
        for q1 in range(1, N + 1):
            for q2 in range(1, N + 1):
                kp, k1 = q1, q2
                v = k3[kp - 1, k1 - 1, k - 1]
                if v > 0:
                    Win[k - 1] += (Veh2[k - 1, k1 - 1] * (1 - nh[kp]) * nh[v]
                                   * ne[k1] * S_eh[k - 1, kp - 1, k1 - 1])
                v = k3[kp - 1, k - 1, k1 - 1]
                if v > 0:
                    Wout[k - 1] += (Veh2[k1 - 1, k - 1] * (1 - ne[k1]) * (1 - nh[kp])
                                    * nh[v] * S_eh[k1 - 1, kp - 1, k - 1])
                k2, k4 = q1, q2
                v = k3[k - 1, k2 - 1, k4 - 1]
                if v > 0:
                    Win[k - 1] += (Vee2[k - 1, k4 - 1] * (1 - ne[k2]) * ne[v]
                                   * ne[k4] * S_ee[k - 1, k2 - 1, k4 - 1])
                v = k3[k4 - 1, k2 - 1, k - 1]
                if v > 0:
                    Wout[k - 1] += (Vee2[k4 - 1, k - 1] * (1 - ne[k4]) * (1 - ne[k2])
                                    * ne[v] * S_ee[k4 - 1, k2 - 1, k - 1])
    a, b = O.mbce(g, S_eh, S_ee, ne0, nh0, VC)
    assert rel(a, Win) < 1e-14 and rel(b, Wout) < 1e-14



# ---------------------------------------------------------------------------
# RK4 (the restatement of _rk4, libpulsesuite/integrator.py:143-157), checked
# as the reference's own tests check it (tests/libpulsesuite/
# test_integrator.py:153-201) and against the exact amplification factor
# ---------------------------------------------------------------------------

def _t4(z):
    return 1.0 + z + z**2 / 2.0 + z**3 / 6.0 + z**4 / 24.0


def test_rk4_restatement():
    decay = lambda y: -y  # noqa: E731
    y = O.rk4(np.array([1.0]), np.array([-1.0]), 0.01, decay)
    assert np.isclose(y[0], np.exp(-0.01), rtol=1e-10)
    rot = lambda y: 1j * y  # noqa: E731
    y = O.rk4(np.array([1.0 + 0j]), np.array([1j]), 0.001, rot)
    assert np.isclose(abs(y[0]), 1.0, rtol=1e-8)
    e1 = abs(O.rk4(np.array([1.0]), np.array([-1.0]), 0.1, decay)[0] - np.exp(-0.1))
    e2 = abs(O.rk4(np.array([1.0]), np.array([-1.0]), 0.05, decay)[0] - np.exp(-0.05))
    assert e1 / e2 > 10
    ho = lambda y: np.array([y[1], -y[0]])  # noqa: E731
    y = np.array([1.0, 0.0])
    for _ in range(1000):
        y = O.rk4(y, ho(y), 0.01, ho)
    assert np.isclose(y[0] ** 2 + y[1] ** 2, 1.0, rtol=1e-6)
    # linear y' = lam y: one step multiplies by the degree-4 Taylor polynomial
    lam = np.array([-0.3 + 2.1j, -1.7 - 0.4j, 2.9j])
    y0 = np.array([1.0 + 0.5j, -0.2j, 3.0])
    y = O.rk4(y0, lam * y0, 0.7, lambda v: lam * v)
    assert np.max(np.abs(y - _t4(0.7 * lam) * y0)) < 1e-14 * np.max(np.abs(y0))


def test_rk4_wire_step_free_rotation():
    """All mb.params terms off: dp = (-i (Ee[ke] + gap + Eh[kh]) / hbar -
    gam_e - gam_h) p, dC = -i (Ee[k2] - Ee[k1]) / hbar C, dD likewise with
    Eh (the free-carrier Hamiltonian and the constant dephasing), so three
    RK4 wire steps multiply every element by T4(dt lam)^3."""
    from scipy.constants import hbar
    from tests.helpers import params_pair, random_state, wire_length
    N = 16
    L = wire_length(N)
    _, oprm = params_pair(L, (0,) * 15)
    s = O.Solver(oprm, 2, np.linspace(-L, L, 32), 0.0, integrator="rk4")
    st = random_state(np.random.default_rng(4), N, 2)
    # equal electron and hole totals: the diagonal renormalisation
    # (SBEs.py:5692-5702) is then the identity
    for c, d in ((st[0], st[1]), (st[3], st[4])):
        for w in range(2):
            np.fill_diagonal(d[:, :, w], np.diag(c[:, :, w]))
    for dst, src in zip((s.CC2, s.DD2, s.YY2, s.CC3, s.DD3, s.YY3), st):
        dst[:] = src
    g = s.g
    lp = -1j * (g.Ee[None, :] + oprm.gap + g.Eh[:, None]) / hbar - (oprm.gam_e + oprm.gam_h)
    lc = -1j * (g.Ee[None, :] - g.Ee[:, None]) / hbar
    ld = -1j * (g.Eh[None, :] - g.Eh[:, None]) / hbar
    dt = 1e-15
    z = np.zeros(32, dtype=complex)
    for _ in range(3):
        for w in (1, 2):
            s.qw_calculator(z, z, z, dt, w)
    for lam, x0, x in ((lp, st[5], s.YY3), (lc, st[3], s.CC3), (ld, st[4], s.DD3)):
        ref = _t4(dt * lam)[:, :, None] ** 3 * x0
        assert np.max(np.abs(x - ref)) < 1e-13 * np.max(np.abs(ref))


def test_calc_xqw_matches_reference(golden_dir):
    """chi(q, w) against SBESolver.calc_Xqw of the reference at random
    occupations and dephasing (one rate below the 0.1 meV floor), q of both
    signs and the empty q = N, w from 0 to twice the gap."""
    z = np.load(os.path.join(golden_dir, "ref_xqw.npz"))
    for a, iq in enumerate(z["iqs"]):
        for b, w in enumerate(z["ws"]):
            got = O.calc_xqw(int(iq), float(w), float(z["dkr"]), z["fe"], z["fh"], z["Ee"], z["Eh"],
                             float(z["gap"]), float(z["area"]), z["game"], z["gamh"], float(z["dcv"]))
            ref = z["X"][a, b]
            assert abs(got - ref) <= 1e-13 * max(abs(ref), 1e-300) + 1e-300, (iq, w)
