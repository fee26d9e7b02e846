"""GPU QWOptics drop-in (pulsesuite_amd.qwoptics) through the C ABI
(psa_mod_dft, psa_mod_resample, psa_mod_pol3, psa_mod_rho5,
psa_mod_plasmonics), on the SBE grids of the golden runs:

* the init-time tables against the ones the reference built (Expikr,
  QWWindow);
* QWRho5 against the densities the reference computed (fre, frh);
* QWPolarization3, Prop2QW and QW2Prop (with and without Plasmonics)
  against the oracle, whose versions the golden QWCalculator trajectories
  pin.

Tolerance 1e-12 relative to the largest entry (direct DFTs and a different
summation order against numpy's FFT and einsum)."""

import types

import numpy as np
import pytest

from oracle import sbe_oracle as O
from tests.helpers import load_golden, rel

pytestmark = pytest.mark.gpu

TOL = 1e-12


def _setup(tag):
    from pulsesuite_amd.qwoptics import QWOptics
    z = load_golden(tag)
    prm = O.Params(L=float(z["L"]), me_rel=0.067, flags=tuple(z["mb"]))
    g = O.Grid(prm)
    t = O.Tables(g)
    qw = QWOptics(g.r, prm.L, g.dcv, g.kr, g.Qr, g.Ee, g.Eh, g.ehint, g.area, prm.gap)
    return z, prm, g, t, qw


@pytest.mark.parametrize("tag", ["n16_mb", "n32_mb"])
def test_qwoptics_tables_and_rho5_match_reference(tag):
    print(f"[qwoptics] rho5 {tag}", flush=True)
    z, prm, g, t, qw = _setup(tag)
    assert rel(qw._Expikr, z["Expikr"]) < 1e-15 and rel(qw._QWWindow, z["QWWindow"]) < 1e-15
    assert qw.Xcv(3, 1) == -g.dcv and qw.Ycv(0, 2) == g.dcv and qw.Zcv(1, 1) == -g.dcv
    C, D, p = z["fC"], z["fD"], z["fp"]
    ne, nh = np.diag(C).copy(), np.diag(D).copy()
    re = np.zeros(g.Nr, dtype=complex)
    rh = np.zeros(g.Nr, dtype=complex)
    qw.QWRho5(g.Qr, g.kr, g.r, prm.L, g.kkp, p, C, D, ne, nh, re, rh, 1, 1)
    assert rel(re, z["fre"]) < TOL and rel(rh, z["frh"]) < TOL


@pytest.mark.parametrize("tag", ["n16_mb", "n32_mb"])
def test_qwpolarization3_and_prop2qw_match_oracle(tag):
    print(f"[qwoptics] pol3/prop2qw {tag}", flush=True)
    z, prm, g, t, qw = _setup(tag)
    p = z["fp"]
    for w in (1, 2, 3, 4):
        P = [np.zeros(g.Nr, dtype=complex) for _ in range(3)]
        qw.QWPolarization3(g.r, g.kr, p, g.ehint, g.area, prm.L, *P, 1, w)
        ref = O.qw_polarization3(g, t, p, w)
        for a, b in zip(P, ref):
            assert rel(a, b) < TOL, w
    rng = np.random.default_rng(3)
    rr = z["rr"]
    F = [rng.standard_normal(len(rr)) * 1e7 + 1j * rng.standard_normal(len(rr)) * 1e7 for _ in range(4)]
    Q = [np.zeros(g.Nr, dtype=complex) for _ in range(4)]
    qw.Prop2QW(rr, *F, 0.0, g.r, *Q, 0.0, 1)
    for a, f in zip(Q, F):
        assert rel(a, O.prop2qw(g, t, rr, f)) < TOL


@pytest.mark.parametrize("plasmonics", [True, False])
def test_qw2prop_matches_oracle(plasmonics):
    print(f"[qwoptics] qw2prop plasmonics={plasmonics}", flush=True)
    z, prm, g, t, qw = _setup("n32_mb")
    rr = z["rr"]
    rng = np.random.default_rng(8)
    Nr, Nx = g.Nr, len(rr)
    arrs = [rng.standard_normal(Nr) + 1j * rng.standard_normal(Nr) for _ in range(9)]
    keep = [a.copy() for a in arrs]
    outs = [np.zeros(Nx, dtype=complex) for _ in range(5)]
    qw.QW2Prop(g.r, g.Qr, *arrs, rr, *outs, 1, 1, False, plasmonics)
    # the nine QW arrays are inverse-transformed in place
    for a, b in zip(arrs[:7], keep[:7]):
        assert rel(a, O.ifftg(b)) < TOL
    solver = types.SimpleNamespace(g=g, rr=rr, prm=types.SimpleNamespace(LF=plasmonics))
    ref = O.Solver.qw2prop(solver, *keep[4:9])
    for a, b in zip(outs, ref):
        assert rel(a, b) < TOL
