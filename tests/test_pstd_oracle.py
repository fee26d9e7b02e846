"""The propagator oracle (oracle/pstd_oracle.py) pinned on the CPU: the
difference operators against the reference's dEdx / dHdx semantics and
closed forms, vacuum translation, energy conservation, the field of a
current sheet, the uniform mode, and the absorbing profile's reflection."""

import numpy as np
import pytest
from scipy.special import erf

from oracle import pstd_oracle as P

N0 = 3.1
LAM = 800e-9


def grid(N, dy):
    return np.arange(N) * dy


def test_dEdx_dHdx_closed_forms():
    """usefulsubs.py:612-655 / 704-747 on a ramp (interior slope, wrap point)
    and on a sine (exact difference identities)."""
    N, dy = 64, 3e-9
    y = grid(N, dy)
    r = 2.0 * np.arange(N) + 1.0
    d = P.dEdx(r, dy)
    assert np.allclose(d[:-1], 2.0 / dy, rtol=1e-14)
    assert np.isclose(d[-1], -2.0 * (N - 1) / dy, rtol=1e-14)
    h = P.dHdx(r, dy)
    assert np.allclose(h[1:], 2.0 / dy, rtol=1e-14)
    assert np.isclose(h[0], -2.0 * (N - 1) / dy, rtol=1e-14)
    k = 2 * np.pi * 5 / (N * dy)
    s = np.sin(k * y)
    # sin(k(y+dy)) - sin(ky) = 2 sin(k dy/2) cos(k(y + dy/2))
    assert np.allclose(P.dEdx(s, dy), 2 * np.sin(k * dy / 2) * np.cos(k * (y + dy / 2)) / dy,
                       rtol=0, atol=1e-9 / dy)
    assert np.allclose(P.dHdx(s, dy), 2 * np.sin(k * dy / 2) * np.cos(k * (y - dy / 2)) / dy,
                       rtol=0, atol=1e-9 / dy)


@pytest.mark.parametrize("N", [64, 128, 1000])
def test_difference_multipliers_match_dEdx_dHdx(N):
    """The "fd" k-space multipliers reproduce the reference's differences."""
    dy = 2e-9
    rng = np.random.default_rng(N)
    f = rng.standard_normal(N) + 1j * rng.standard_normal(N)
    ike, ikh = P.multipliers(N, dy, "fd")
    a = np.fft.ifft(ike * np.fft.fft(f))
    b = np.fft.ifft(ikh * np.fft.fft(f))
    ref_a, ref_b = P.dEdx(f, dy), P.dHdx(f, dy)
    assert np.max(np.abs(a - ref_a)) < 1e-12 * np.max(np.abs(ref_a))
    assert np.max(np.abs(b - ref_b)) < 1e-12 * np.max(np.abs(ref_b))


def test_spectral_multipliers_are_staggered_derivatives():
    N, dy = 256, 2e-9
    y = grid(N, dy)
    ike, ikh = P.multipliers(N, dy, "spectral")
    for m in (1, 7, 100):
        k = 2 * np.pi * m / (N * dy)
        s = np.sin(k * y)
        a = np.fft.ifft(ike * np.fft.fft(s))
        b = np.fft.ifft(ikh * np.fft.fft(s))
        assert np.max(np.abs(a - k * np.cos(k * (y + dy / 2)))) < 1e-11 * k
        assert np.max(np.abs(b - k * np.cos(k * (y - dy / 2)))) < 1e-11 * k


def _pulse_pair(o, yc, width, pol, direction, t=0.0):
    L = o.N * o.dy
    y = grid(o.N, o.dy)
    v = 1 / np.sqrt(o.eps * o.mu)
    eta = np.sqrt(o.mu / o.eps)
    k0 = 2 * np.pi * N0 / LAM
    c = yc + direction * v * t
    E = P.periodic_images(y, 0.0, L, lambda u: P.gauss_pulse(u, 1e7, c, width, k0))
    Hh = P.periodic_images(y + o.dy / 2, 0.0, L, lambda u: P.gauss_pulse(u, 1e7, c, width, k0)) / eta
    H = (-Hh if pol == "x" else Hh) * direction
    return E, H


@pytest.mark.parametrize("pol,direction", [("x", 1), ("x", -1), ("z", 1), ("z", -1)])
def test_vacuum_translation_at_c_over_n0(pol, direction):
    """Spectral derivatives + exact mode update: a band-limited pulse moves
    rigidly at c/n0 in either direction (closed form)."""
    N, dy, dt = 1024, 4e-9, 0.5e-15
    o = P.Pstd1D(N, dy, N0, dt)
    E, H = _pulse_pair(o, 2.0e-6, 0.3e-6, pol, direction)
    f = np.zeros((4, N))
    e, h = (P.EX, P.HZ) if pol == "x" else (P.EZ, P.HX)
    f[e], f[h] = E, H
    o.set_real(f)
    W0 = o.energy()
    for _ in range(40):
        o.advance()
    Ea, Ha = _pulse_pair(o, 2.0e-6, 0.3e-6, pol, direction, t=40 * dt)
    r = o.real()
    assert np.max(np.abs(r[e] - Ea)) < 1e-9 * 1e7
    assert np.max(np.abs(r[h] - Ha)) < 1e-9 * 1e7 / np.sqrt(o.mu / o.eps)
    assert abs(o.energy() - W0) < 1e-13 * W0


def test_fd_single_mode_dispersion():
    """The reference's differences in space: mode k moves at the discrete
    phase velocity (2 / (k dy)) sin(k dy / 2) v (closed form)."""
    N, dy, dt = 256, 4e-9, 2e-15
    o = P.Pstd1D(N, dy, N0, dt, deriv="fd")
    y = grid(N, dy)
    v = 1 / np.sqrt(o.eps * o.mu)
    m = 20
    k = 2 * np.pi * m / (N * dy)
    om = 2 * v / dy * np.sin(k * dy / 2)
    # travelling wave E = cos(k y - om t); dHz/dt = (dE/dy)/mu gives Hz on the
    # half points with amplitude -(2 sin(k dy/2)/dy) / (mu om)
    hamp = -(2 * np.sin(k * dy / 2) / dy) / (o.mu * om)
    f = np.zeros((4, N))
    f[P.EX] = np.cos(k * y)
    f[P.HZ] = hamp * np.cos(k * (y + dy / 2))
    o.set_real(f)
    n = 50
    for _ in range(n):
        o.advance()
    t = n * dt
    r = o.real()
    assert np.max(np.abs(r[P.EX] - np.cos(k * y - om * t))) < 1e-11
    assert np.max(np.abs(r[P.HZ] - hamp * np.cos(k * (y + dy / 2) - om * t))) < 1e-11 * abs(hamp)


@pytest.mark.parametrize("deriv", ["spectral", "fd"])
def test_energy_conserved_without_sources(deriv):
    N, dy, dt = 512, 2e-9, 1e-15
    o = P.Pstd1D(N, dy, N0, dt, deriv=deriv)
    rng = np.random.default_rng(3)
    o.set_real(rng.standard_normal((4, N)) * np.array([[1e6], [1e4], [1e6], [1e4]]))
    W0 = o.energy()
    for _ in range(200):
        o.advance()
    assert abs(o.energy() - W0) < 1e-13 * W0


def test_current_sheet_radiates_minus_eta_half_K():
    """A sheet current K(t) at y0 gives E = -(eta/2) K(t - |y - y0|/v) on
    both sides (image sum over the periodic box); the gather takes the
    sheet polarization P(t) = int K dt / dy on the grid."""
    N, dy, dt = 1024, 4e-9, 0.05e-15
    L = N * dy
    j0 = N // 2
    tau, t0, K0 = 5e-15, 25e-15, 1.0
    o = P.Pstd1D(N, dy, N0, dt, jw=[j0])
    v = 1 / np.sqrt(o.eps * o.mu)
    eta = np.sqrt(o.mu / o.eps)

    def Pint(t):
        return K0 * tau * np.sqrt(np.pi) / 2 * (erf((t - t0) / tau) - erf(-t0 / tau))

    n = 800
    for i in range(n):
        o.gather(np.array([[Pint((i + 1) * dt) / dy], [0.0]]), dt)
        o.advance()
    y = grid(N, dy)
    Ea = P.sheet_field(y, n * dt, y[j0], L, v, eta,
                       lambda t: np.where(t > 0, K0 * np.exp(-((t - t0) / tau) ** 2), 0.0))
    r = o.real()[P.EX]
    assert np.max(np.abs(r - Ea)) < 1e-4 * eta / 2 * K0
    # symmetric about the sheet's grid point
    assert np.max(np.abs(r[j0 + 1:j0 + 300] - r[j0 - 1:j0 - 300:-1])) < 1e-12 * eta / 2 * K0


def test_uniform_mode_from_net_current():
    """k = 0: a uniform current density J0 for one step lowers the mean
    field by J0 dt / eps; the curl terms leave it alone."""
    N, dy, dt = 128, 2e-9, 1e-16
    o = P.Pstd1D(N, dy, N0, dt, jw=list(range(N)))
    o.gather(np.full((2, N), 3.0 * dt), dt)  # J = 3 A/m^2 on every point
    o.advance()
    r = o.real()
    assert np.allclose(r[P.EX], -3.0 * dt / o.eps, rtol=1e-12)
    assert np.allclose(r[P.EZ], -3.0 * dt / o.eps, rtol=1e-12)
    hs = 3.0 * dt / o.eps / np.sqrt(o.mu / o.eps)
    assert np.max(np.abs(r[P.HZ])) < 1e-12 * hs and np.max(np.abs(r[P.HX])) < 1e-12 * hs


def test_abc_profile_absorbs_outgoing_pulse():
    """The product's absorbing profile: a pulse run into each layer leaves
    less than 1e-6 of its energy behind."""
    from pulsesuite_amd.pstd import abc_profile
    N, dy, dt = 2048, 4e-9, 0.5e-15
    abc = abc_profile(N, dy, dt, N0, 0.5e-6, 0.5e-6)
    assert abc[N // 2] == 1.0 and abc[0] < 1.0 and abc[-1] < 1.0
    for direction in (1, -1):
        o = P.Pstd1D(N, dy, N0, dt, abc=abc)
        E, H = _pulse_pair(o, N * dy / 2, 0.3e-6, "x", direction)
        f = np.zeros((4, N))
        f[P.EX], f[P.HZ] = E, H
        o.set_real(f)
        W0 = o.energy()
        v = 1 / np.sqrt(o.eps * o.mu)
        for _ in range(int((N * dy / 2 + 2e-6) / (v * dt))):
            o.advance()
        assert o.energy() < 1e-6 * W0
