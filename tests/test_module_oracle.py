"""The oracle of the module functions off the QWCalculator path
(oracle/module_oracle.py) and the drop-ins' host-side pieces, on the CPU.

* every oracle function against the vectors the reference wrote
  (tests/golden/ref_modx.npz, make_golden.py extras): bit for bit where
  the reference sums sequentially (CalcMVeh, BGRenorm, EeRenorm, EhRenorm,
  GetEps1Dqw, PLSpectrum), 1e-14 relative where it does not (GetVn1n2:
  numpy's SIMD complex products; GetChi1Dqw) and 1e-12 for Vehint (the
  reference evaluates K0 through the complex CIK01 routine, 4e-13 apart);
* against what the mathematics fixes: the renormalisations reduce to
  matrix-vector products (2a - 2a cancels), CalcMVeh is a contraction
  over the delta of k + qp = kp + q, GetVn1n2 on diagonal Hamiltonians
  has a closed form and Vvc = Vcv^H, Vehint against scipy.special.k0,
  PLSpectrum with ne nh = 0 only rescales PLS;
* the version-1 off-diagonal dephasing (CalcOffDiagDeph_E/H,
  OffDiagDephasing) against the reference's values, the contraction with
  its one-index shift of p's second index, and against a direct loop over
  the reference's own formula at a second size;
* the host-side drop-in functions (GetChi1Dqw, GetEps1Dqw, QWChi1,
  Calchw) and the file writers (printITReal2, printITReal, printIT3D,
  WriteSBESolns, WritePLSpectrum, WritePropFields) against the
  reference's values and files."""

import os

import numpy as np
import pytest
from scipy.special import k0 as k0_scipy

from oracle import module_oracle as M
from oracle import sbe_oracle as O
from tests.helpers import load_golden, rel


@pytest.fixture(scope="module")
def z():
    return load_golden("modx")


def _k3_undel(n):
    i = np.arange(1, n + 1)
    k3 = i[:, None, None] + i[None, :, None] - i[None, None, :]
    k3 = np.where((k3 >= 1) & (k3 <= n), k3, 0).astype(np.int32)
    U = np.ones((n + 1, n + 1))
    U[0, :] = 0.0
    U[:, 0] = 0.0
    U[np.arange(1, n + 1), np.arange(1, n + 1)] = 0.0
    return k3, U


def test_sequential_sums_bit_exact(z):
    n = int(z["nk"])
    k3, U = _k3_undel(n)
    assert np.array_equal(M.calc_mveh(z["p3"], z["fVC"], k3, U), z["MVeh"])
    assert np.array_equal(M.bg_renorm(z["fC"], z["fD"], z["fVC"], U), z["BGR"])
    assert np.array_equal(M.carrier_renorm(np.diag(z["fC"]), z["fVC"][:, :, 1], U), z["EeR"])
    assert np.array_equal(M.carrier_renorm(np.diag(z["fD"]), z["fVC"][:, :, 2], U), z["EhR"])
    eps = [M.eps1dqw(z["alphae"], z["alphah"], z["Delta0"], z["epsr"], z["me"], z["mh"], float(z["eps_n1d"]),
                     q, w) for q, w in z["eps_qw"]]
    assert np.array_equal(np.array(eps), z["eps"])
    pls = M.pl_spectrum(z["pl_ne"], z["pl_nh"], z["Ee"], z["Eh"], float(z["gap"]), float(z["gam_eh"]), z["fVC"],
                        z["hw"], float(z["pl_t"]), z["pls0"], float(z["rscale"]))
    assert np.array_equal(pls, z["pls"])


def test_other_functions_against_reference(z):
    V = M.vn1n2(z["rcv"], z["Hcc"], z["Hhh"], z["Hcv"])
    for v, k in zip(V, ("Vcc", "Vvv", "Vcv", "Vvc")):
        assert rel(v, z[k]) < 1e-15, k
    for (a, b), ref in zip(z["vehint_kq"], z["vehint"]):
        got = M.vehint(int(a), int(b), z["r"], z["kr"], z["alphae"], z["alphah"], z["Delta0"])
        assert abs(got - ref) <= 1e-12 * abs(ref), (a, b)
    for q, ref in zip(z["vint_q"], z["vint"]):
        assert abs(O.vint(q, z["r"], z["alphae"], z["alphah"], z["Delta0"]) - ref) <= 1e-14 * abs(ref)
    chi = [M.chi1dqw(z["alphae"], z["alphah"], z["Delta0"], z["epsr"], z["game"], z["gamh"], z["kr"], z["Ee"],
                     z["Eh"], z["fe"], z["fh"], q, w) for q, w in z["chi_qw"]]
    assert rel(chi, z["chi"]) < 1e-14
    # the PL spectrum's photon energy grid (Calchw) on the reference's end points
    from scipy.constants import e
    hw = M.calchw(len(z["hw"]), z["gap"] - 0.05 * e, z["gap"] + 0.25 * e)
    assert np.array_equal(hw, z["hw"])


def test_renormalisations_reduce_to_products(z):
    """EeRenorm = c - b - d and BGRenorm = -(nh Vhh') - (ne Vee') with
    V' = V masked by UnDel: matrix-vector products, no loop over i."""
    n = int(z["nk"])
    _, U = _k3_undel(n)
    Um = U[1:, 1:]
    VC = z["fVC"]
    ne, nh = np.diag(z["fC"]), np.diag(z["fD"])
    for n_, V, got in ((ne, VC[:, :, 1], M.carrier_renorm(ne, VC[:, :, 1], U)),
                       (nh, VC[:, :, 2], M.carrier_renorm(nh, VC[:, :, 2], U))):
        b = n_ @ (V * Um)
        c = np.sum(n_) * np.diag(V)
        want = c[:, None] - b[None, :] - b[:, None]
        scale = np.max(np.abs(np.sum(n_) * np.diag(V)))
        assert np.max(np.abs(got - want)) < 1e-13 * scale
    want = -(nh @ (VC[:, :, 2] * Um))[None, :] - (ne @ (VC[:, :, 1] * Um))[:, None]
    assert rel(M.bg_renorm(z["fC"], z["fD"], VC, U), want) < 1e-14
    # a plausible slip (the k and kp roles swapped) is caught
    assert rel(want.T, want) > 1e-3


def test_mveh_is_a_momentum_conserving_contraction():
    rng = np.random.default_rng(3)
    for n, nf in ((5, 1), (9, 3)):
        k3, U = _k3_undel(n)
        p = rng.normal(size=(n, n, nf)) + 1j * rng.normal(size=(n, n, nf))
        VC = rng.normal(size=(n, n, 3))
        k = np.arange(n)
        # delta[k, kp, q, qp] = 1 where k + qp = kp + q
        delta = (k[:, None, None, None] + k[None, None, None, :]
                 == k[None, :, None, None] + k[None, None, :, None]).astype(float)
        Um = U[1:, 1:]
        want = np.einsum("abqr,qrf,aq,aq,br->abf", delta, p, VC[:, :, 0], Um, Um)
        assert rel(M.calc_mveh(p, VC, k3, U), want) < 1e-14


def test_vn1n2_closed_form_and_adjoint():
    rng = np.random.default_rng(4)
    n = 7
    rcv = rng.normal(size=n) + 1j * rng.normal(size=n)
    a, b = 1.3e-20, -0.4e-20
    Vcc, Vvv, Vcv, Vvc = M.vn1n2(rcv, a * np.eye(n), b * np.eye(n), np.zeros((n, n), complex))
    from scipy.constants import hbar
    assert rel(Vcv, np.diag(1j / hbar * (a + b) * rcv)) < 1e-15
    assert np.max(np.abs(Vcc)) == 0.0 and np.max(np.abs(Vvv)) == 0.0
    H = [rng.normal(size=(n, n)) + 1j * rng.normal(size=(n, n)) for _ in range(3)]
    Vcc, Vvv, Vcv, Vvc = M.vn1n2(rcv, *H)
    assert np.array_equal(Vvc, np.conj(Vcv.T))
    # Vcc with Hcv = diag(h): (-i/hbar) (rcv_k conj(h_k) - h_k conj(rcv_k)) on the diagonal
    h = rng.normal(size=n) + 1j * rng.normal(size=n)
    Vcc = M.vn1n2(rcv, H[0], H[1], np.diag(h))[0]
    assert rel(np.diag(Vcc), -1j / hbar * (rcv * np.conj(h) - h * np.conj(rcv))) < 1e-15


def test_vehint_against_scipy_k0(z):
    y, kr = z["r"], z["kr"]
    ae, ah, D0 = float(z["alphae"]), float(z["alphah"]), float(z["Delta0"])
    Ny = len(y)
    N1, N2 = Ny // 4, 3 * Ny // 4
    yi = y[N1:N2 + 1]
    for k, q in ((1, 5), (7, 2)):
        dk = max(abs(kr[k - 1] - kr[q - 1]), (ae + ah) / 4.0)
        r = np.sqrt((yi[:, None] - yi[None, :]) ** 2 + D0**2)
        want = np.sum(np.exp(-(ae * yi[:, None]) ** 2 - (ah * yi[None, :]) ** 2) * ae * ah / np.pi
                      * (y[1] - y[0]) ** 2 * k0_scipy(dk * r))
        # CIK01's K0 is within 1e-8 of the exact K0 (its series and
        # asymptotic branches meet at x = 9 with ~8e-9 error there)
        assert abs(M.vehint(k, q, y, kr, ae, ah, D0) - want) < 2e-8 * abs(want)


def test_pl_spectrum_without_carriers_only_rescales(z):
    n = int(z["nk"])
    pls0 = z["pls0"]
    zero = np.zeros(n, dtype=complex)
    ne = z["pl_ne"]
    out = M.pl_spectrum(ne, zero, z["Ee"], z["Eh"], float(z["gap"]), float(z["gam_eh"]), z["fVC"], z["hw"],
                        float(z["pl_t"]), pls0, float(z["rscale"]))
    tf = 0.5 * (1.0 + 2.0 / np.pi * np.arctan(float(z["pl_t"]) / float(z["gam_eh"])))
    assert np.array_equal(out, pls0 * tf)


def test_host_dropin_functions(z):
    from pulsesuite_amd import coulomb, qwoptics
    chi = [coulomb.GetChi1Dqw(z["alphae"], z["alphah"], z["Delta0"], z["epsr"], z["game"], z["gamh"], z["kr"],
                              z["Ee"], z["Eh"], z["fe"], z["fh"], q, w) for q, w in z["chi_qw"]]
    assert rel(chi, z["chi"]) < 1e-14
    eps = [coulomb.GetEps1Dqw(z["alphae"], z["alphah"], z["Delta0"], z["epsr"], z["me"], z["mh"],
                              float(z["eps_n1d"]), q, w) for q, w in z["eps_qw"]]
    assert rel(eps, z["eps"]) < 1e-14
    dkr = z["kr"][1] - z["kr"][0]
    chi1 = qwoptics.QWChi1(800e-9, dkr, z["Ee"], z["Eh"], z["area"], z["gam_eh"], z["dcv"])
    assert abs(chi1 - z["qwchi1"]) <= 1e-14 * abs(z["qwchi1"])
    from pulsesuite_amd import emission
    hw = np.zeros(len(z["hw"]))
    P = np.ones(len(z["hw"]))
    from scipy.constants import e
    emission.Calchw(hw, P, z["gap"] - 0.05 * e, z["gap"] + 0.25 * e)
    assert np.array_equal(hw, z["hw"]) and not P.any()


def test_file_writers_match_reference_files(z, tmp_path):
    from pulsesuite_amd import SBEs, qwoptics
    cwd = os.getcwd()
    os.chdir(tmp_path)
    try:
        f1, y8, F = z["f1"], z["y8"], list(z["F"])
        SBEs._prop_first[0] = True
        qwoptics.printITReal2(f1, y8, 3, "t/r2a.")
        qwoptics.printITReal2(f1, y8, 4, "t/r2b.")
        qwoptics.printITReal(f1, y8, 5, "t/r1.")
        qwoptics.printIT3D(z["c3"], y8, 6, "t/c3.")
        SBEs._prop_first[0] = True
        qwoptics.WritePropFields(y8, *F, "r", 2, 7)
        C, D = z["fC"], z["fD"]
        qwoptics.WriteSBESolns(z["kr"], np.diag(C).copy(), np.diag(D).copy(), C, D, z["fp"], z["Ee"] + 0j,
                               z["Eh"] + 0j, 2, 8)
        qwoptics.WritePLSpectrum(z["hw"], z["pls"], 2, 9)
        names = [str(s) for s in z["file_names"]]
        assert len(names) == 22
        for nm, txt in zip(names, z["file_texts"]):
            with open(os.path.join("dataQW", nm)) as f:
                assert f.read() == str(txt), nm
    finally:
        os.chdir(cwd)
        SBEs._prop_first[0] = True


def test_offdiag_dephasing_v1_against_reference(z):
    g = z["od_g"]
    DE = M.offdiag_deph1(z["od_ne"], z["od_nh"], z["fVC"], z["Ee"], z["Eh"], z["od_kpq"], z["od_kmq"], g[0], g[2],
                         False)
    DH = M.offdiag_deph1(z["od_ne"], z["od_nh"], z["fVC"], z["Ee"], z["Eh"], z["od_kpq"], z["od_kmq"], g[1], g[2],
                         True)
    assert np.array_equal(DE, z["od_DE"])
    assert rel(DH, z["od_DH"]) < 1e-15  # two entries one ulp apart
    x = M.offdiag_dephasing1(z["od_ne"], z["od_nh"], z["od_p"], z["kr"], z["Ee"], z["Eh"], g, z["fVC"], z["od_kpq"],
                             z["od_kmq"])
    assert np.array_equal(x, z["od_x"])
    # the shift is what the reference does: the unshifted contraction is far off
    ps = z["od_p"]
    assert np.abs(x[:, 0]).max() > 0 and np.abs(x[0, :]).max() > 0
    assert rel(M.offdiag_dephasing1(z["od_ne"], z["od_nh"], np.roll(ps, 1, axis=1), z["kr"], z["Ee"], z["Eh"], g,
                                    z["fVC"], z["od_kpq"], z["od_kmq"]), x) > 1e-3


def test_offdiag_tables_v1_brute_force():
    """The table restatement against a plain loop over the reference's
    formula (dephasing.py:107-190) on random inputs with tables that have
    out-of-range entries."""
    from scipy.constants import hbar
    rng = np.random.default_rng(8)
    N = 7
    ne = rng.uniform(0, 0.6, N) * np.exp(1j * rng.uniform(0, 6, N))
    nh = rng.uniform(0, 0.6, N) * np.exp(1j * rng.uniform(0, 6, N))
    VC = rng.uniform(0.1, 1.0, (N, N, 3)) * 1e-21
    Ee = np.sort(rng.uniform(0, 1e-20, N))
    Eh = np.sort(rng.uniform(0, 1e-21, N))
    kpq = rng.integers(-2, N + 2, (N, N)).astype(np.int32)
    kmq = rng.integers(-2, N + 2, (N, N)).astype(np.int32)
    for hole in (False, True):
        gss, gso = (4e12, 2e12) if hole else (3e12, 2e12)
        ns, no = (np.abs(nh), np.abs(ne)) if hole else (np.abs(ne), np.abs(nh))
        Es, Eo = (Eh, Ee) if hole else (Ee, Eh)
        Vs = VC[:, :, 2 if hole else 1]
        want = np.zeros((N, N))
        L = lambda a, b: (b / np.pi) / (a * a + b * b)  # noqa: E731
        for k in range(N):
            for q in range(N):
                b = kpq[k, q]
                if not 0 <= b < N:
                    continue
                acc = 0.0
                for k1 in range(N):
                    a = kpq[k1, q]
                    if 0 <= a < N:
                        acc += Vs[k1, a] ** 2 * L(Es[a] + Es[k] - Es[k1] - Es[b], hbar * gss) * (
                            ns[a] * ns[k] * (1 - ns[k1]) + (1 - ns[a]) * (1 - ns[k]) * ns[k1])
                for k1 in range(N):
                    c = kmq[k1, q]
                    if 0 <= c < N:
                        v = VC[k1, c, 0] if hole else VC[k, b, 0]
                        acc += v ** 2 * L(Eo[c] + Es[k] - Eo[k1] - Es[b], hbar * gso) * (
                            no[c] * (1 - no[k1]) * ns[k] + (1 - no[c]) * no[k1] * (1 - ns[k]))
                want[k, q] = acc * np.pi / hbar
        got = M.offdiag_deph1(ne, nh, VC, Ee, Eh, kpq, kmq, gss, gso, hole)
        assert rel(got, want) < 1e-14, hole


def test_write_dephasing_reproduces_reference_files(tmp_path):
    """WriteDephasing / printGam (dephasing.py:417-437): the Ge and Gh files
    of the reference's snapshot run, rewritten from the values they hold
    (float32 text round-trips), are the same text."""
    from pulsesuite_amd import dephasing
    z = load_golden("n16_snap")
    names = [str(n) for n in z["file_names"]]
    texts = {n: str(t) for n, t in zip(names, z["file_texts"])}
    ge = "dataQW/Wire/Ge/Ge.01.k.00002.dat"
    gh = "dataQW/Wire/Gh/Gh.01.k.00002.dat"
    a = np.loadtxt(texts[ge].splitlines())
    b = np.loadtxt(texts[gh].splitlines())
    assert np.array_equal(a[:, 0], b[:, 0])
    cwd = os.getcwd()
    os.chdir(tmp_path)
    try:
        dephasing.WriteDephasing(a[:, 0], a[:, 1], b[:, 1], 1, 2)
        for n in (ge, gh):
            with open(n) as f:
                assert f.read() == texts[n], n
    finally:
        os.chdir(cwd)
