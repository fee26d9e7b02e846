"""The SBEs drop-in (reference call signatures, params files from the CWD,
per-wire QWCalculator calls) against the reference's golden trajectories,
and the backup text format round trip."""

import os

import numpy as np
import pytest

from tests.helpers import load_golden, rel

pytestmark = pytest.mark.gpu


def write_params(d, z):
    os.makedirs(os.path.join(d, "params"), exist_ok=True)
    qw = [float(z["L"]), 5e-9, 1.5, 0.067, 0.45, 100e-3, 1e12, 1e12, 1e12, 9.1,
          36e-3, 3e-3, 0.0, 500, 20000]
    with open(os.path.join(d, "params", "qw.params"), "w") as f:
        for v in qw:
            f.write(f"{v!r}    # value\n")
    with open(os.path.join(d, "params", "mb.params"), "w") as f:
        for v in z["mb"]:
            f.write(f"{int(v)}          # flag\n")


@pytest.mark.parametrize("tag", ["n16_mb", "n32_mb"])
def test_dropin_qwcalculator_matches_reference(tag, tmp_path, monkeypatch):
    z = load_golden(tag)
    write_params(str(tmp_path), z)
    monkeypatch.chdir(tmp_path)
    from pulsesuite_amd import SBEs
    print(f"[dropin] {tag}", flush=True)
    rr, qrr = z["rr"], z["qrr"]
    s = SBEs.SBESolver(qrr, rr, 0.0, float(z["Emax0"]), float(z["lam"]), 2, True)
    assert s.Nk == int(z["nk"])
    s.set_levels(z["s0_C2"], z["s0_D2"], z["s0_p2"], z["s0_C3"], z["s0_D3"], z["s0_p3"])
    dt = float(z["dt"])
    nx = len(rr)
    # outputs into strided views, as rhoPJ passes Pxx[:, w]
    P = np.zeros((nx, 2, 4), dtype=complex)
    for it in range(int(z["nsteps"])):
        Exx, Eyy = z["t_Exx"][it], z["t_Eyy"][it]
        Ezz = np.zeros_like(Exx)
        for w in (1, 2):
            s.QWCalculator(Exx, Eyy, Ezz, np.zeros_like(Exx), rr, qrr, dt, w,
                           P[:, w - 1, 0], P[:, w - 1, 1], P[:, w - 1, 2], P[:, w - 1, 3],
                           [True], [False])
            j = 2 * it + w - 1
            for c, nm in enumerate(("Pxx", "Pyy", "Pzz", "Rho")):
                assert rel(P[:, w - 1, c], z["t_" + nm][j]) < 1e-9, (it, w, nm)
        assert rel(s.CC3, z["t_C3"][it]) < 1e-9
        assert rel(s.YY2, z["t_p2"][it]) < 1e-9
    # info file columns against the reference's own info file
    info = np.loadtxt("dataQW/info.01.t.dat", ndmin=2)
    ref = np.real(z["info1"])
    for col in (0, 1, 2, 3, 4, 5, 6, 12, 13, 14, 15):
        assert np.allclose(info[:, col], ref[:, col], rtol=1e-5, atol=1e-30), col
    s.close()


def test_backup_round_trip(tmp_path, monkeypatch):
    z = load_golden("n16_mb")
    write_params(str(tmp_path), z)
    monkeypatch.chdir(tmp_path)
    from pulsesuite_amd import SBEs
    s = SBEs.SBESolver(z["qrr"], z["rr"], 0.0, 1.0, 8e-7, 2, True)
    s.set_levels(z["s0_C2"], z["s0_D2"], z["s0_p2"], z["s0_C3"], z["s0_D3"], z["s0_p3"])
    s.write_sbes_data(7)
    a = s._levels()
    t = SBEs.SBESolver(z["qrr"], z["rr"], 0.0, 1.0, 8e-7, 2, True)
    t.read_sbes_data(7)
    b = t._levels()
    for x, y in zip(a, b):
        assert rel(x, y) < 1e-15
