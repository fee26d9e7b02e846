"""Host-side logic of the serving layer (no GPU): the wire-off gate of a
request's pulse and the footprint arithmetic used for admission."""

import numpy as np


def test_serving_gate_open_time_brute_force():
    """The wire-off gate opens at the first step with |E| >= 1e-3 Emax0
    (SBEs.py:5849-5858) for the sbetestprop pulse (sbetestprop.py:183-198)."""
    from pulsesuite_amd.serving import SimRequest
    r = SimRequest(rid=0, nk=64, e0=2e7, fwhm=20e-15, detuning_ev=0.01, nsteps=10)
    dt = 1e-16
    t_open = r.gate_open_time(1.5, dt)
    E0x, E0y, E0z, w0, tw, tp = r.pulse(1.5)
    emax = np.sqrt(E0x**2 + E0y**2)

    def mag(t):
        x = w0 * (t - tp)
        f = np.exp(-(x**2) / (w0 * tw) ** 2) * np.cos(x) * np.exp(-((x / (2 * tw * w0)) ** 20))
        return abs(f) * emax

    n = int(round(t_open / dt))
    assert mag(n * dt) >= 1e-3 * emax
    assert all(mag(m * dt) < 1e-3 * emax for m in range(n))
    assert 0 < t_open < tp


def test_slot_footprint_grows_with_nk_and_bf16_planes():
    from pulsesuite_amd.serving import slot_bytes
    a = slot_bytes(64, 4096, "fp64", 1000)
    b = slot_bytes(128, 4096, "fp64", 1000)
    c = slot_bytes(128, 4096, "bf16x3", 1000)
    # 12 complex N x N matrices dominate: 4x per doubling of N, plus the planes
    # (the fields and the split-K scratch scale with the propagation grid)
    fixed = 4096 * 7 * 16 + 8 * 2 * 4096 * 5 * 8
    assert 3.5 < (b - fixed) / (a - fixed) < 4.5
    assert c - b == 12 * 4 * 128 * 128 * 2
