"""Host side of the coulomb / phonons module drop-ins (no GPU): the
init-time builders against the tables the reference wrote
(tests/golden/ref_n16_mb.npz, ref_n32_mb.npz), the module state and its
errors, and the oracle's LorentzDelta many-body branch (which the
reference's vectors do not cover) pinned by its support and its value at
zero energy mismatch."""

import types

import numpy as np
import pytest

from oracle import sbe_oracle as O
from tests.helpers import load_golden, rel


def _setup(tag):
    z = load_golden(tag)
    prm = O.Params(L=float(z["L"]), me_rel=0.067, flags=tuple(z["mb"]))
    return z, prm, O.Grid(prm)


@pytest.mark.parametrize("tag", ["n16_mb", "n32_mb"])
def test_coulomb_builders_match_reference(tag):
    from pulsesuite_amd import coulomb
    z, prm, g = _setup(tag)
    N = g.Nk
    qe, qh = coulomb.MakeQs(g.kr, g.alphae, g.alphah)
    assert rel(qe, z["qe"]) == 0 and rel(qh, z["qh"]) == 0
    ce, ch = coulomb.CalcChi1D(g.kr, g.alphae, g.alphah, prm.Delta0, prm.epsr, prm.me, prm.mh, qe, qh)
    assert rel(ce, z["Chi1De"]) < 1e-12 and rel(ch, z["Chi1Dh"]) < 1e-12
    u = coulomb.MakeUnDel(g.kr)
    assert u.shape == (N + 1, N + 1) and u[0].sum() == 0 and u[:, 0].sum() == 0
    assert u.sum() == N * N - N
    if "k3" in z.files:
        assert np.array_equal(coulomb.MakeK3(g.kr), z["k3"])


@pytest.mark.parametrize("tag", ["n16_mb", "n32_mb"])
def test_phonon_tables_match_reference(tag):
    from pulsesuite_amd import phonons
    z, prm, g = _setup(tag)
    phonons.InitializePhonons(g.kr, g.Ee, g.Eh, prm.L, prm.epsr, prm.Gph, prm.Oph)
    assert rel(phonons._EP, z["EP"]) < 1e-15 and rel(phonons._HP, z["HP"]) < 1e-15
    assert np.array_equal(phonons._EPT, phonons._EP.T)
    assert phonons._Vscale == z["Vscale"] and phonons._NO == z["NO"] == phonons.N00()
    assert np.all(np.diag(phonons._idel) == 0) and phonons._idel.sum() == g.Nk * (g.Nk - 1)


def test_module_errors_without_state():
    from pulsesuite_amd import capi, coulomb, phonons
    n = np.zeros(4)
    with pytest.raises(ValueError, match="not initialized"):
        coulomb.MBCE(n, n, n, n, n, np.zeros((4, 4, 3)), 0.0, 0.0, n.copy(), n.copy())
    old = phonons._EP
    phonons._EP = None
    try:
        with pytest.raises(ValueError, match="InitializePhonons"):
            phonons.MBPE(n, np.zeros((4, 4, 3)), np.ones((4, 4)), n.copy(), n.copy())
    finally:
        phonons._EP = old
    assert coulomb._k3 is None or coulomb._instance is not None
    import torch
    if not torch.cuda.is_available():
        z, prm, g = _setup("n16_mb")
        with pytest.raises(capi.PsaError):
            coulomb.CoulombModule(g.r, g.kr, prm.L, prm.Delta0, prm.me, prm.mh, g.Ee, g.Eh, prm.gam_e,
                                  prm.gam_h, g.alphae, g.alphah, prm.epsr, g.Qr, g.kkp, True)


def test_oracle_lorentz_tables_support_and_zero_mismatch():
    """The LorentzDelta tables (coulomb.py:405-446) are nonzero exactly
    where the Gaussian ones are, and where the energy mismatch vanishes
    their ratio to the Gaussian is 1/sqrt(pi) (2/(hbar^2 g) against
    2 pi/hbar / (sqrt(pi) hbar g)), with the width of each table."""
    N = 12
    ky = np.linspace(-1e8, 1e8, N)
    g = types.SimpleNamespace(Nk=N, Ee=3e-20 * (ky / 1e8) ** 2, Eh=1e-20 * (ky / 1e8) ** 2)
    # widths comparable to the band energies, so no Gaussian underflows
    ge, gh = 1.3e14, 0.7e14
    for kind in ("eh", "ee", "hh"):
        G = O.mb_tables(g, kind, ge, gh)
        L = O.mb_tables(g, kind, ge, gh, lorentz=True)
        assert np.array_equal(G != 0, L != 0), kind
        gam = {"eh": (ge + gh) / 2, "ee": ge, "hh": gh}[kind]
        peak_g = 2 * np.pi / O.hbar / (np.sqrt(np.pi) * O.hbar * gam)
        at_zero = np.isclose(G, peak_g, rtol=1e-12, atol=0)
        assert at_zero.sum() > 0, kind
        assert np.allclose(L[at_zero] / G[at_zero], 1 / np.sqrt(np.pi), rtol=1e-12)


def test_oracle_lorentz_tables_match_reference():
    """CalcMBArrays with LorentzDelta = 1 (coulomb.py:405-446), as the
    reference's SBESolver builds it from mb.params line 15, and MBCE/MBCH on
    those tables (tests/golden/ref_n16_lorentz.npz)."""
    from tests.helpers import load_golden
    z = load_golden("n16_lorentz")
    prm = O.Params(L=float(z["L"]), me_rel=0.067, flags=O.SHIPPED_FLAGS)
    g = O.Grid(prm)
    assert rel(g.Ee, z["Ee"]) == 0 and rel(g.Eh, z["Eh"]) == 0
    S = {k: O.mb_tables(g, k, prm.gam_e, prm.gam_h, lorentz=True) for k in ("eh", "ee", "hh")}
    for k, nm in (("eh", "Ceh"), ("ee", "Cee"), ("hh", "Chh")):
        assert rel(S[k], z[nm]) < 1e-15, nm
        # a wrong dependence on the energy mismatch would not survive this:
        # the Lorentzian is nowhere zero where the index conditions hold
        assert np.count_nonzero(z[nm]) > 0.5 * np.count_nonzero(O.mb_tables(g, k, 1e15, 1e15))
    ne, nh = np.real(np.diag(z["fC"])), np.real(np.diag(z["fD"]))
    a, b = O.mbce(g, S["eh"], S["ee"], ne, nh, z["fVC"])
    assert rel(a, z["fWinE"]) < 1e-14 and rel(b, z["fWoutE"]) < 1e-14
    a, b = O.mbch(g, S["eh"], S["hh"], ne, nh, z["fVC"])
    assert rel(a, z["fWinH"]) < 1e-14 and rel(b, z["fWoutH"]) < 1e-14


def test_qwoptics_host_helpers():
    from pulsesuite_amd import qwoptics
    assert [qwoptics.yw(w) for w in range(1, 9)] == [O.yw(w) for w in range(1, 9)] == [1, 1, -1, -1, 1, 1, -1, -1]
    import torch
    if not torch.cuda.is_available():
        from pulsesuite_amd import capi
        z, prm, g = _setup("n16_mb")
        with pytest.raises(capi.PsaError):
            qwoptics.QWOptics(g.r, prm.L, g.dcv, g.kr, g.Qr, g.Ee, g.Eh, g.ehint, g.area, prm.gap)


@pytest.mark.parametrize("tag", ["n16_mb", "n32_mb"])
def test_dephasing_and_emission_tables_match_reference(tag, tmp_path, monkeypatch):
    from pulsesuite_amd import dephasing, emission
    z, prm, g = _setup(tag)
    monkeypatch.chdir(tmp_path)  # InitializeDephasing opens dataQW/Wire/info/MaxOffDiag.*
    dephasing.InitializeDephasing(g.kr, prm.me, prm.mh)
    for k in ("k_p_q", "k_m_q", "k1_m_q", "k1p_m_q", "k1", "k1p"):
        assert np.array_equal(getattr(dephasing, "_" + k), z[k]), k
    assert rel(dephasing._xe, z["xe"]) < 1e-15 and rel(dephasing._xh, z["xh"]) < 1e-15
    assert (tmp_path / "dataQW/Wire/info/MaxOffDiag.e.dat").exists()
    dephasing._fe_file.close()
    dephasing._fh_file.close()
    dephasing._fe_file = dephasing._fh_file = None
    emission.InitializeEmission(g.kr, g.Ee, g.Eh, abs(g.dcv), prm.epsr, g.gam_eh, g.ehint)
    assert rel(emission._HOmega, z["HOmega"]) == 0 and rel(emission._square, z["square"]) < 1e-15
    assert emission._RScale == z["RScale"]
