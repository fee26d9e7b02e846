"""GPU dephasing / emission drop-ins (pulsesuite_amd.dephasing,
pulsesuite_amd.emission) through the C ABI (psa_mod_gamma,
psa_mod_offdiag2, psa_mod_spont), on the SBE grids of the golden runs:
CalcGammaE/H and OffDiagDephasing2 against the vectors the reference wrote
(fGamE, fGamH, fOffP), the off-diagonal tables, the MaxOffDiag lines and
SpontEmission/Ec against the oracle (whose versions the golden trajectories
with recombination pin).  Tolerance 1e-12 relative to the largest entry."""

import numpy as np
import pytest

from oracle import sbe_oracle as O
from tests.helpers import load_golden, rel

pytestmark = pytest.mark.gpu

TOL = 1e-12


@pytest.mark.parametrize("tag", ["n16_mb", "n32_mb"])
def test_dephasing_matches_reference_vectors(tag, tmp_path, monkeypatch):
    from pulsesuite_amd import dephasing
    print(f"[dephasing] {tag}", flush=True)
    z = load_golden(tag)
    prm = O.Params(L=float(z["L"]), me_rel=0.067, flags=tuple(z["mb"]))
    g = O.Grid(prm)
    monkeypatch.chdir(tmp_path)
    dephasing.InitializeDephasing(g.kr, prm.me, prm.mh)
    N = g.Nk
    C, D, p, VC = z["fC"], z["fD"], z["fp"], z["fVC"]
    ne, nh = np.diag(C).copy(), np.diag(D).copy()
    GE, GH = np.zeros(N), np.zeros(N)
    dephasing.CalcGammaE(g.kr, ne, nh, VC, GE)
    dephasing.CalcGammaH(g.kr, ne, nh, VC, GH)
    assert rel(GE, z["fGamE"]) < TOL and rel(GH, z["fGamH"]) < TOL
    x = np.zeros((N, N), dtype=complex)
    gg = np.array([g.gam_eh, prm.gam_e, prm.gam_h])
    dephasing.OffDiagDephasing2(ne, nh, p, g.kr, g.Ee, g.Eh, gg, VC, 1.5e-15, x)
    assert rel(x, z["fOffP"]) < TOL
    De = O.offdiag_deph_e2(g, ne, nh, VC, gg[0], gg[2])
    Dh = O.offdiag_deph_h2(g, ne, nh, VC, gg[1], gg[2])
    assert rel(dephasing.CalcOffDiagDeph_E2(ne, nh, g.kr, g.Ee, g.Eh, gg[0], gg[2], VC, N), De) < TOL
    assert rel(dephasing.CalcOffDiagDeph_H2(ne, nh, g.kr, g.Ee, g.Eh, gg[1], gg[2], VC, N), Dh) < TOL
    dephasing._fe_file.close()
    dephasing._fh_file.close()
    dephasing._fe_file = dephasing._fh_file = None
    for fn, tab in (("e", De), ("h", Dh)):
        t, mx, mn = (float(v) for v in (tmp_path / f"dataQW/Wire/info/MaxOffDiag.{fn}.dat").read_text().split())
        assert t == 1.5e-15
        assert abs(mx - tab.max()) <= TOL * abs(tab).max() and abs(mn - tab.min()) <= TOL * abs(tab).max()


@pytest.mark.parametrize("tag", ["n16_mb", "n32_mb"])
def test_spont_emission_matches_oracle(tag):
    from pulsesuite_amd import emission
    print(f"[emission] {tag}", flush=True)
    z = load_golden(tag)
    prm = O.Params(L=float(z["L"]), me_rel=0.067, flags=tuple(z["mb"]))
    g = O.Grid(prm)
    t = O.Tables(g)
    emission.InitializeEmission(g.kr, g.Ee, g.Eh, abs(g.dcv), prm.epsr, g.gam_eh, g.ehint)
    N = g.Nk
    ne, nh, VC = np.diag(z["fC"]).copy(), np.diag(z["fD"]).copy(), z["fVC"]
    R = np.zeros(N)
    emission.SpontEmission(ne, nh, g.Ee, g.Eh, prm.gap, g.gam_eh, VC, R)
    assert rel(R, O.spont_emission(g, t, ne, nh, VC)) < TOL
    assert rel(emission.Ec(np.real(ne), np.real(nh), VC), O.ec(np.real(ne), np.real(nh), VC)) < TOL
    # the scalar helper on the host agrees with the device integral
    Ek = prm.gap + g.Ee[3] + g.Eh[3] + O.ec(np.real(ne), np.real(nh), VC)[3]
    assert abs(emission.SpontIntegral(Ek) - R[3]) <= TOL * abs(R).max()
