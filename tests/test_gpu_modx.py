"""The module functions off the QWCalculator path on the device, through
the C ABI (psa_mod_mveh, psa_mod_renorm, psa_mod_vn1n2, psa_mod_plspectrum,
psa_mod_vehint, psa_vint_table), on the golden run's own grid:

* against the vectors the reference wrote (tests/golden/ref_modx.npz):
  CalcMVeh, BGRenorm, EeRenorm, EhRenorm bit for bit (sequential sums in
  the reference's order, no contraction into FMAs); GetVn1n2 within 1e-15
  and Vvc = Vcv^H exactly; PLSpectrum within 1e-12 (tree sums against
  numpy's pairwise sums, device exp/atan); Vint within 1e-13 and Vehint
  within 1e-12 (the reference's complex CIK01 K0);
* against the oracle at N = 48 on random inputs, the same bars;
* outputs written into strided views;
* the version-1 off-diagonal dephasing functions."""

import numpy as np
import pytest

from oracle import module_oracle as M
from oracle import sbe_oracle as O
from tests.helpers import load_golden, rel

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def z():
    return load_golden("modx")


def _init(L, N=None):
    from pulsesuite_amd import coulomb, emission
    prm = O.Params(L=L, me_rel=0.067)
    g = O.Grid(prm)
    coulomb.SetLorentzDelta(False)
    coulomb.InitializeCoulomb(g.r, g.kr, prm.L, prm.Delta0, prm.me, prm.mh, g.Ee, g.Eh, prm.gam_e,
                              prm.gam_h, g.alphae, g.alphah, prm.epsr, g.Qr, g.kkp, True)
    emission.InitializeEmission(g.kr, g.Ee, g.Eh, abs(g.dcv), prm.epsr, g.gam_eh, g.ehint)
    return prm, g


def test_coulomb_extras_match_reference(z):
    from pulsesuite_amd import coulomb
    print("[modx] coulomb extras N=16", flush=True)
    prm, g = _init(float(z["L"]))
    assert np.allclose(g.kr, z["kr"], rtol=1e-15, atol=0)
    N = g.Nk
    MV = np.zeros((N, N, 3), dtype=complex)
    coulomb.CalcMVeh(z["p3"], z["fVC"], MV)
    assert np.array_equal(MV, z["MVeh"])
    big = np.zeros((N + 3, 2 * N), dtype=complex)
    B = big[1:N + 1, ::2]  # a strided output view
    coulomb.BGRenorm(z["fC"], z["fD"], z["fVC"], B)
    assert np.array_equal(B, z["BGR"]) and not big[0].any() and not big[:, 1::2].any()
    coulomb.EeRenorm(np.diag(z["fC"]).copy(), z["fVC"], B)
    assert np.array_equal(B, z["EeR"])
    coulomb.EhRenorm(np.diag(z["fD"]).copy(), z["fVC"], B)
    assert np.array_equal(B, z["EhR"])
    for q, ref in zip(z["vint_q"], z["vint"]):
        assert abs(coulomb.Vint(q, z["r"], z["alphae"], z["alphah"], z["Delta0"]) - ref) <= 1e-13 * abs(ref), q
    for (a, b), ref in zip(z["vehint_kq"], z["vehint"]):
        got = coulomb.Vehint(int(a), int(b), z["r"], z["kr"], z["alphae"], z["alphah"], z["Delta0"])
        assert abs(got - ref) <= 1e-12 * abs(ref), (a, b)


def test_qwoptics_emission_extras_match_reference(z):
    from pulsesuite_amd import emission, qwoptics
    print("[modx] GetVn1n2, PLSpectrum N=16", flush=True)
    _init(float(z["L"]))
    N = int(z["nk"])
    V = [np.zeros((N, N), dtype=complex) for _ in range(4)]
    qwoptics.GetVn1n2(z["kr"], z["rcv"], z["Hcc"], z["Hhh"], z["Hcv"], *V)
    for v, k in zip(V, ("Vcc", "Vvv", "Vcv", "Vvc")):
        assert rel(v, z[k]) < 1e-15, k
    assert np.array_equal(V[3], np.conj(V[2].T))
    assert abs(emission._RScale - float(z["rscale"])) <= 1e-15 * float(z["rscale"])
    PLS = z["pls0"].copy()
    emission.PLSpectrum(z["pl_ne"], z["pl_nh"], z["Ee"], z["Eh"], float(z["gap"]), float(z["gam_eh"]), z["fVC"],
                        z["hw"], float(z["pl_t"]), PLS)
    assert rel(PLS, z["pls"]) < 1e-12


def test_extras_against_oracle_n48():
    from pulsesuite_amd import coulomb, emission, qwoptics
    from tests.helpers import wire_length
    print("[modx] extras vs oracle N=48", flush=True)
    prm, g = _init(wire_length(48))
    N = g.Nk
    assert N == 48
    rng = np.random.default_rng(48)
    k3 = coulomb._k3
    U = coulomb._UnDel
    p = rng.normal(size=(N, N, 2)) + 1j * rng.normal(size=(N, N, 2))
    VC = rng.uniform(0.1, 1.0, size=(N, N, 3)) * 1e-21
    MV = np.zeros((N, N, 2), dtype=complex)
    coulomb.CalcMVeh(p, VC, MV)
    assert np.array_equal(MV, M.calc_mveh(p, VC, k3, U))
    C = rng.normal(size=(N, N)) + 1j * rng.normal(size=(N, N))
    D = rng.normal(size=(N, N)) + 1j * rng.normal(size=(N, N))
    B = np.zeros((N, N), dtype=complex)
    coulomb.BGRenorm(C, D, VC, B)
    assert np.array_equal(B, M.bg_renorm(C, D, VC, U))
    ne = np.diag(C).copy()
    coulomb.EeRenorm(ne, VC, B)
    assert np.array_equal(B, M.carrier_renorm(ne, VC[:, :, 1], U))
    rcv = rng.normal(size=N) + 1j * rng.normal(size=N)
    H = [rng.normal(size=(N, N)) + 1j * rng.normal(size=(N, N)) for _ in range(3)]
    V = [np.zeros((N, N), dtype=complex) for _ in range(4)]
    qwoptics.GetVn1n2(g.kr, rcv, *H, *V)
    for a, b in zip(V, M.vn1n2(rcv, *H)):
        assert rel(a, b) < 1e-15
    hw = np.zeros(64)
    PLS = np.zeros(64)
    emission.Calchw(hw, PLS, prm.gap - 0.02 * O.e0, prm.gap + 0.3 * O.e0)
    nep = rng.uniform(0.0, 0.5, N) * np.exp(1j * rng.uniform(0, 6, N))
    nhp = rng.uniform(0.0, 0.5, N) * np.exp(1j * rng.uniform(0, 6, N))
    VCs = np.zeros((N, N, 3))
    E1D = np.zeros((N, N))
    coulomb.CalcScreenedArrays(True, prm.L, np.abs(nep) + 0j, np.abs(nhp) + 0j, VCs, E1D)
    PLS[:] = rng.uniform(0.0, 1e-8, 64)
    want = M.pl_spectrum(nep, nhp, g.Ee, g.Eh, prm.gap, g.gam_eh, VCs, hw, 2e11, PLS, emission._RScale)
    emission.PLSpectrum(nep, nhp, g.Ee, g.Eh, prm.gap, g.gam_eh, VCs, hw, 2e11, PLS)
    assert np.abs(want).max() > 10 * 1e-8
    assert rel(PLS, want) < 1e-12


def test_offdiag_dephasing_v1_matches_reference_and_oracle(z):
    """CalcOffDiagDeph_E/H and OffDiagDephasing (the version-1 functions,
    psa_mod_offdiag1 / _apply) on the golden run's index tables: the tables
    and the contraction within 1e-14 of the reference's values (sequential
    sums in the reference's order; |n| by the device's hypot against
    libm's); again against the oracle on random occupations at N = 48."""
    from pulsesuite_amd import dephasing
    from tests.helpers import wire_length
    print("[modx] off-diagonal dephasing v1", flush=True)
    prm, g = _init(float(z["L"]))
    dephasing.InitializeDephasing(g.kr, prm.me, prm.mh)
    assert np.array_equal(dephasing._k_p_q, z["od_kpq"]) and np.array_equal(dephasing._k_m_q, z["od_kmq"])
    gg = z["od_g"]
    DE = dephasing.CalcOffDiagDeph_E(z["od_ne"], z["od_nh"], z["kr"], z["Ee"], z["Eh"], gg[0], gg[2], z["fVC"])
    DH = dephasing.CalcOffDiagDeph_H(z["od_ne"], z["od_nh"], z["kr"], z["Ee"], z["Eh"], gg[1], gg[2], z["fVC"])
    assert rel(DE, z["od_DE"]) < 1e-14
    assert rel(DH, z["od_DH"]) < 1e-14
    x = np.zeros((len(z["kr"]), len(z["kr"])), dtype=complex)
    dephasing.OffDiagDephasing(z["od_ne"], z["od_nh"], z["od_p"], z["kr"], z["Ee"], z["Eh"], list(gg), z["fVC"], x)
    assert rel(x, z["od_x"]) < 1e-14
    prm, g = _init(wire_length(48))
    dephasing.InitializeDephasing(g.kr, prm.me, prm.mh)
    N = g.Nk
    rng = np.random.default_rng(11)
    ne = rng.uniform(0, 0.6, N) * np.exp(1j * rng.uniform(0, 6, N))
    nh = rng.uniform(0, 0.6, N) * np.exp(1j * rng.uniform(0, 6, N))
    VC = rng.uniform(0.1, 1.0, (N, N, 3)) * 1e-21
    p = (rng.normal(size=(N, N)) + 1j * rng.normal(size=(N, N))) * 0.05
    gg = (3e12, 4e12, 2e12)
    kpq, kmq = dephasing._k_p_q, dephasing._k_m_q
    for hole, fn, gs in ((False, dephasing.CalcOffDiagDeph_E, gg[0]), (True, dephasing.CalcOffDiagDeph_H, gg[1])):
        got = fn(ne, nh, g.kr, g.Ee, g.Eh, gs, gg[2], VC)
        assert rel(got, M.offdiag_deph1(ne, nh, VC, g.Ee, g.Eh, kpq, kmq, gs, gg[2], hole)) < 1e-14, hole
    x = np.zeros((N, N), dtype=complex)
    dephasing.OffDiagDephasing(ne, nh, p, g.kr, g.Ee, g.Eh, gg, VC, x)
    assert rel(x, M.offdiag_dephasing1(ne, nh, p, g.kr, g.Ee, g.Eh, gg, VC, kpq, kmq)) < 1e-14
