"""GPU module drop-ins (pulsesuite_amd.coulomb, pulsesuite_amd.phonons)
through the C ABI (psa_mod_*, psa_vint_table):

* on the SBE grids of the golden runs: the module tables (unscreened
  Coulomb integrals, k3, Ceh/Cee/Chh) and the per-call functions
  (CalcScreenedArrays, MBCE, MBCH) against the vectors the reference wrote;
  MBPE/MBPH against the oracle (whose phonon rates the golden trajectories
  pin);
* on the reference test-suite's own fixture (test_coulomb.py:709-742: an
  energy table and Qy/kkp that are not the SBE grid's, random VC that is
  not a function of k - k'): MBCE == MBCE2, the
  rates against the oracle, both broadening branches of the tables,
  screening on and off, accumulation into Win/Wout and strided outputs.

Tolerances (relative to the largest entry): 1e-12 for every FP64 output
(device exp/log and a different summation order against numpy)."""

import types

import numpy as np
import pytest

from oracle import sbe_oracle as O
from tests.helpers import load_golden, rel

pytestmark = pytest.mark.gpu

TOL = 1e-12


def _init_sbe_grid(tag):
    from pulsesuite_amd import coulomb, phonons
    z = load_golden(tag)
    prm = O.Params(L=float(z["L"]), me_rel=0.067, flags=tuple(z["mb"]))
    g = O.Grid(prm)
    coulomb.SetLorentzDelta(False)
    coulomb.InitializeCoulomb(g.r, g.kr, prm.L, prm.Delta0, prm.me, prm.mh, g.Ee, g.Eh, prm.gam_e,
                              prm.gam_h, g.alphae, g.alphah, prm.epsr, g.Qr, g.kkp, True)
    phonons.InitializePhonons(g.kr, g.Ee, g.Eh, prm.L, prm.epsr, prm.Gph, prm.Oph)
    return z, prm, g


@pytest.mark.parametrize("tag", ["n16_mb", "n32_mb"])
def test_coulomb_module_matches_reference_vectors(tag):
    from pulsesuite_amd import coulomb
    print(f"[modules] coulomb {tag}", flush=True)
    z, prm, g = _init_sbe_grid(tag)
    N = g.Nk
    for k in ("Veh0", "Vee0", "Vhh0", "qe", "qh", "Chi1De", "Chi1Dh"):
        assert rel(getattr(coulomb, "_" + k), z[k]) < TOL, k
    if "k3" in z.files:
        assert np.array_equal(coulomb._k3, z["k3"])
        for k in ("Ceh", "Cee", "Chh"):
            assert rel(getattr(coulomb, "_" + k), z[k]) < TOL, k
    ne, nh = np.diag(z["fC"]).copy(), np.diag(z["fD"]).copy()
    VC = np.zeros((N, N, 3))
    E1D = np.zeros((N, N))
    coulomb.CalcScreenedArrays(True, prm.L, ne, nh, VC, E1D)
    assert rel(VC, z["fVC"]) < TOL and rel(E1D, z["fE1D"]) < TOL
    VC = z["fVC"]
    for fn, tags, gam in ((coulomb.MBCE, ("fWinE", "fWoutE"), prm.gam_e),
                          (coulomb.MBCH, ("fWinH", "fWoutH"), prm.gam_h)):
        Win, Wout = np.zeros(N), np.zeros(N)
        fn(np.real(ne), np.real(nh), g.kr, g.Ee, g.Eh, VC, g.gam_eh, gam, Win, Wout)
        assert rel(Win, z[tags[0]]) < TOL and rel(Wout, z[tags[1]]) < TOL, fn.__name__
        # the rates are added to what the arrays hold
        fn(np.real(ne), np.real(nh), g.kr, g.Ee, g.Eh, VC, g.gam_eh, gam, Win, Wout)
        assert rel(Win, 2 * z[tags[0]]) < TOL and rel(Wout, 2 * z[tags[1]]) < TOL


def test_coulomb_module_lorentz_tables_match_reference():
    """The LorentzDelta branch of the device many-body tables (psa_mod
    mb_arrays) and MBCE/MBCH on them against the reference's own vectors
    (tests/golden/ref_n16_lorentz.npz)."""
    from pulsesuite_amd import coulomb
    print("[modules] coulomb LorentzDelta = 1", flush=True)
    z = load_golden("n16_lorentz")
    prm = O.Params(L=float(z["L"]), me_rel=0.067, flags=O.SHIPPED_FLAGS)
    g = O.Grid(prm)
    coulomb.SetLorentzDelta(True)
    try:
        coulomb.InitializeCoulomb(g.r, g.kr, prm.L, prm.Delta0, prm.me, prm.mh, g.Ee, g.Eh, prm.gam_e,
                                  prm.gam_h, g.alphae, g.alphah, prm.epsr, g.Qr, g.kkp, True)
        assert np.array_equal(coulomb._k3, z["k3"])
        for k in ("Ceh", "Cee", "Chh"):
            assert rel(getattr(coulomb, "_" + k), z[k]) < TOL, k
        N = g.Nk
        ne, nh = np.real(np.diag(z["fC"])), np.real(np.diag(z["fD"]))
        for fn, tags, gam in ((coulomb.MBCE, ("fWinE", "fWoutE"), prm.gam_e),
                              (coulomb.MBCH, ("fWinH", "fWoutH"), prm.gam_h)):
            Win, Wout = np.zeros(N), np.zeros(N)
            fn(ne, nh, g.kr, g.Ee, g.Eh, z["fVC"], g.gam_eh, gam, Win, Wout)
            assert rel(Win, z[tags[0]]) < TOL and rel(Wout, z[tags[1]]) < TOL, fn.__name__
    finally:
        coulomb.SetLorentzDelta(False)


@pytest.mark.parametrize("tag", ["n16_mb", "n32_mb"])
def test_phonon_rates_match_oracle(tag):
    from pulsesuite_amd import phonons
    print(f"[modules] phonons {tag}", flush=True)
    z, prm, g = _init_sbe_grid(tag)
    t = O.Tables(g)
    ne, nh = np.diag(z["fC"]).copy(), np.diag(z["fD"]).copy()
    VC, E1D = z["fVC"], z["fE1D"]
    N = g.Nk
    for fn, ofn, n in ((phonons.MBPE, O.mbpe, ne), (phonons.MBPH, O.mbph, nh)):
        Win, Wout = np.full(N, 0.5), np.full(N, 0.25)
        fn(n, VC, E1D, Win, Wout)
        a, b = ofn(t, np.real(n), VC, E1D)
        assert rel(Win - 0.5, a) < TOL and rel(Wout - 0.25, b) < TOL, fn.__name__
    # tables assigned to the module state directly are the ones used
    old = phonons._EP, phonons._EPT
    try:
        phonons._EP = 2.0 * old[0]
        phonons._EPT = phonons._EP.T
        Win, Wout = np.zeros(N), np.zeros(N)
        phonons.MBPE(ne, VC, E1D, Win, Wout)
        a, b = O.mbpe(t, np.real(ne), VC, E1D)
        assert rel(Win, 2 * a) < TOL and rel(Wout, 2 * b) < TOL
    finally:
        phonons._EP, phonons._EPT = old


def _fixture_module(N, lorentz):
    """test_coulomb.py:709-742 (CoulombModule fixture)."""
    from pulsesuite_amd.coulomb import CoulombModule
    y = np.linspace(-10e-6, 10e-6, 2 * N)
    ky = np.linspace(-1e7, 1e7, N)
    Ee = 0.5 * 9.109e-31 * (ky / 1e7) ** 2 * (1.6e-19)
    Eh = 0.5 * 1.5 * 9.109e-31 * (ky / 1e7) ** 2 * (1.6e-19)
    NQ = 2 * N - 1
    Qy = np.linspace(-2e7, 2e7, NQ)
    kk = np.arange(N)[:, None] - np.arange(N)[None, :] + NQ // 2
    kkp = np.where((kk >= 0) & (kk < NQ), kk, -1).astype(np.int32)
    cm = CoulombModule(y, ky, 100e-6, 10e-9, 0.067 * 9.109e-31, 0.5 * 9.109e-31, Ee, Eh, 1e12, 1e12, 1e6,
                       1.2e6, 12.0, Qy, kkp, False, LorentzDelta=lorentz)
    return cm, types.SimpleNamespace(Nk=N, Ee=Ee, Eh=Eh, ky=ky)


@pytest.mark.parametrize("lorentz", [False, True])
def test_coulomb_module_on_the_reference_fixture(lorentz):
    N = 32
    print(f"[modules] fixture lorentz={lorentz}", flush=True)
    cm, g = _fixture_module(N, lorentz)
    assert cm.k3.shape == (N, N, N) and cm.UnDel.shape == (N + 1, N + 1)
    assert cm.Ceh.shape == (N + 1,) * 3 and cm.Veh0.shape == (N, N) and cm.Chi1De.shape == (N, N)
    assert cm.LorentzDelta is lorentz
    S = {k: O.mb_tables(g, k, 1e12, 1e12, lorentz=lorentz) for k in ("eh", "ee", "hh")}
    assert rel(cm.Ceh, S["eh"]) < TOL and rel(cm.Cee, S["ee"]) < TOL and rel(cm.Chh, S["hh"]) < TOL
    rng = np.random.default_rng(11)
    VC = 1e-20 * rng.random((N, N, 3))
    ne0 = rng.uniform(0.0, 0.9, N) * np.exp(1j * rng.uniform(0, 6.3, N))
    nh0 = rng.uniform(0.0, 0.9, N)
    ga = types.SimpleNamespace(Nk=N)
    a = [np.zeros(N) for _ in range(4)]
    cm.MBCE(ne0, nh0, g.ky, g.Ee, g.Eh, VC, 1e12, 1e12, a[0], a[1])
    cm.MBCE2(ne0, nh0, g.ky, g.Ee, g.Eh, VC, 1e12, 1e12, a[2], a[3])
    assert np.array_equal(a[0], a[2]) and np.array_equal(a[1], a[3])
    assert np.all(a[0] >= 0) and np.all(a[1] >= 0)
    wi, wo = O.mbce(ga, S["eh"], S["ee"], ne0, nh0, VC)
    assert rel(a[0], wi) < TOL and rel(a[1], wo) < TOL
    Win, Wout = np.zeros(N), np.zeros(N)
    cm.MBCH(ne0, nh0, g.ky, g.Ee, g.Eh, VC, 1e12, 1e12, Win, Wout)
    wi, wo = O.mbch(ga, S["eh"], S["hh"], ne0, nh0, VC)
    assert rel(Win, wi) < TOL and rel(Wout, wo) < TOL


def test_screening_on_the_reference_fixture():
    N = 32
    print("[modules] screening", flush=True)
    cm, g = _fixture_module(N, False)
    ne = 0.5 * np.ones(N, dtype=complex)
    nh = 0.5 * np.ones(N, dtype=complex)
    big = np.zeros((N, N, 5))
    VC = big[:, :, 1:4]  # strided output, as the reference's callers may pass
    E1D = np.zeros((N, N))
    cm.CalcScreenedArrays(False, 100e-6, ne, nh, VC, E1D)
    assert np.all(E1D == 1.0)
    assert np.array_equal(VC, np.stack([cm.Veh0, cm.Vee0, cm.Vhh0], axis=2))
    assert np.all(big[:, :, 0] == 0) and np.all(big[:, :, 4] == 0)
    cm.CalcScreenedArrays(True, 100e-6, ne, nh, VC, E1D)
    gs = types.SimpleNamespace(Nk=N, prm=types.SimpleNamespace(L=100e-6))
    t = types.SimpleNamespace(V0=[cm.Veh0, cm.Vee0, cm.Vhh0], qe=cm.qe, qh=cm.qh, Chi1De=cm.Chi1De,
                              Chi1Dh=cm.Chi1Dh)
    VCo, Eo = O.screened_arrays(gs, t, True, ne, nh)
    assert not np.allclose(E1D, 1.0)
    assert rel(E1D, Eo) < TOL and rel(VC, VCo) < TOL
    # Eps1D at the clipped density equals the screening's dielectric function
    n1 = min(np.sum(np.real(ne) + np.real(nh)) / 2.0 / 100e-6, min(cm.qe[1, 1], cm.qh[1, 1]) / np.pi * 0.99)
    assert rel(cm.Eps1D(n1, N), E1D) < TOL


def test_module_rejects_bad_shapes():
    from pulsesuite_amd import coulomb, phonons
    _init_sbe_grid("n16_mb")
    N = 16
    with pytest.raises(ValueError):
        coulomb.MBCE(np.zeros(N), np.zeros(N), 0, 0, 0, np.zeros((N, N, 2)), 0, 0, np.zeros(N), np.zeros(N))
    with pytest.raises(ValueError):
        phonons.MBPH(np.zeros(N), np.zeros((N, N, 3)), np.ones((N, N + 1)), np.zeros(N), np.zeros(N))
