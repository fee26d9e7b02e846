"""The 3-D coupling oracle (oracle/qw3d_oracle.py) pinned on the CPU by the
reference's own QuantumWire (tests/golden/ref_qw3d.npz: four wires, three
calls, non-zero incoming grids) and by closed forms: trilinear
interpolation, the discrete continuity equation of CalcJfx, the placement
(which overwrites the grid), and Gauss's law / curl-free longitudinal fields
of ElongfromRho."""

import numpy as np
from scipy.constants import epsilon_0

from oracle import qw3d_oracle as Q
from oracle import sbe_oracle as O
from tests.helpers import load_golden, rel


def _grid(nx=16, ny=12, nz=10, d=2e-9):
    rx = (np.arange(nx) - nx // 2) * d
    ry = (np.arange(ny) - ny // 2) * d
    rz = (np.arange(nz) - nz // 2) * d
    q = [2 * np.pi * np.fft.fftfreq(n, d) for n in (nx, ny, nz)]
    return rx, ry, rz, q


def test_eat_xyz_exact_on_trilinear_functions():
    rx, ry, rz, _ = _grid()
    X, Y, Z = np.meshgrid(rx, ry, rz, indexing="ij")
    a, b, c, d, e = 1.5, -2e8, 3e8, 7e8, 4e16
    f = (a + b * X + c * Y + d * Z + e * X * Y + 1j * e * Y * Z).astype(complex)
    for x0, y0, z0 in ((0.3e-9, -1.1e-9, 2.7e-9), (rx[3], ry[5], rz[2]), (-7.9e-9, 5.5e-9, -3.3e-9)):
        ref = a + b * x0 + c * y0 + d * z0 + e * x0 * y0 + 1j * e * y0 * z0
        assert abs(Q.eat_xyz(f, rx, ry, rz, x0, y0, z0) - ref) < 1e-12 * abs(ref)
    assert Q.eat_xyz(f, rx, ry, rz, rx[4], ry[7], rz[1]) == f[4, 7, 1]


def test_calc_jfx_continuity():
    rng = np.random.default_rng(1)
    n, dt, dx = 40, 1e-17, 2e-9
    new = rng.standard_normal(n) + 1j * rng.standard_normal(n)
    old = rng.standard_normal(n) + 1j * rng.standard_normal(n)
    J = Q.calc_jfx(new, old, dt, dx)
    assert J[0] == 0
    # (J[i] - J[i-1]) / dx = -(rho_new - rho_old)[i] / dt
    lhs = (J[1:] - J[:-1]) / dx
    rhs = -(new - old)[1:] / dt
    assert np.max(np.abs(lhs - rhs)) < 1e-9 * np.max(np.abs(rhs))


def test_placement_overwrites_and_is_separable():
    """QWPlacement zeroes the grid (rhoPJ.py:682): whatever the grid held,
    it ends as f gy gz, and placing twice gives the same grid."""
    rng = np.random.default_rng(2)
    nx, ny, nz = 8, 6, 5
    f = rng.standard_normal(nx) + 1j * rng.standard_normal(nx)
    gy, gz = rng.uniform(0, 1, ny), rng.uniform(0, 1, nz)
    F = 1e3 * (rng.standard_normal((nx, ny, nz)) + 1j)
    Q.qw_placement(f, np.ones(nx), gy, gz, F)
    assert np.allclose(F.sum(axis=(1, 2)), f * gy.sum() * gz.sum(), rtol=1e-13)
    assert np.allclose(F[:, 2, 3], f * gy[2] * gz[3], rtol=1e-14)
    F2 = F.copy()
    Q.qw_placement(f, np.ones(nx), gy, gz, F2)
    assert np.array_equal(F, F2)


def qw3d_from_golden(z, keep_old=False):
    """The oracle solver and coupling set up as the reference run was:
    InitializeSBE on the grid's x axis, seeded levels for every wire."""
    prm = O.Params(L=float(z["L"]), me_rel=0.067, flags=O.SHIPPED_FLAGS)
    g = O.Grid(prm)
    nw = int(z["nw"])
    sv = O.Solver(prm, nw, z["rx"], float(z["Emax0"]), tables=O.Tables(g), grid=g)
    for nm, dst in (("C2", sv.CC2), ("D2", sv.DD2), ("p2", sv.YY2),
                    ("C3", sv.CC3), ("D3", sv.DD3), ("p3", sv.YY3)):
        dst[:] = z["s0_" + nm]
    oq = Q.QuantumWire3D(sv, z["rx"], z["ry"], z["rz"], z["qx"], z["qy"], z["qz"],
                         float(z["y0"]), float(z["z0"]), float(z["ay"]), float(z["az"]),
                         float(z["epsr"]), keep_old=keep_old)
    return sv, oq


def test_quantum_wire_matches_reference():
    """Three QuantumWire calls (rhoPJ.py:1180-1240 -> QWArray.QuantumWire)
    from random non-zero Jx, Jy, Jz, Rho grids: the placed grids are
    overwritten and J = sum P / (dt Nw) (the previous P is discarded), so
    the grids and the ten longitudinal outputs match the reference."""
    z = load_golden("qw3d")
    sv, oq = qw3d_from_golden(z)
    grids = [z[k].copy() for k in ("J0x", "J0y", "J0z", "Rho0")]
    dt = float(z["dt"])
    for it in range(z["q_Jx"].shape[0]):
        res = oq.step(z["q_Ex"][it], z["q_Ey"][it], z["q_Ez"][it], dt, *grids)
        for a, k in zip(grids, ("Jx", "Jy", "Jz", "Rho")):
            assert rel(a, z["q_" + k][it]) < 1e-13, (it, k)
        els = list(res[0]) + list(res[1]) + list(res[2]) + [res[3]]
        for c in range(10):
            assert rel(els[c], z["q_El"][it][c]) < 1e-13, (it, c)


def test_keep_old_differs_from_the_reference():
    """keep_old (not the reference) makes J the polarization current: from
    the second call on, J differs from the reference's sum P / dt."""
    z = load_golden("qw3d")
    sv, oq = qw3d_from_golden(z, keep_old=True)
    grids = [z[k].copy() for k in ("J0x", "J0y", "J0z", "Rho0")]
    dt = float(z["dt"])
    oq.step(z["q_Ex"][0], z["q_Ey"][0], z["q_Ez"][0], dt, *grids)
    assert rel(grids[0], z["q_Jx"][0]) < 1e-12
    oq.step(z["q_Ex"][1], z["q_Ey"][1], z["q_Ez"][1], dt, *grids)
    assert rel(grids[1], z["q_Jy"][1]) > 1e-3


def test_elong_gauss_law_and_curl_free():
    """Longitudinal fields: q x E = 0, i q.E_rho = rho / eps (q^2 + s)/q^2,
    i q.(eps E_P) = -i q.P (q^2 + s)/q^2, bound charge = -i q.P."""
    rx, ry, rz, (qx, qy, qz) = _grid(12, 10, 8)
    rng = np.random.default_rng(3)
    shp = (12, 10, 8)
    rho = rng.standard_normal(shp) + 1j * rng.standard_normal(shp)
    P = [rng.standard_normal(shp) + 1j * rng.standard_normal(shp) for _ in range(3)]
    epsr = 9.1
    e_pr, e_p, e_r, rb = Q.elong_from_rho(rho, *P, qx, qy, qz, epsr)
    QX, QY, QZ = np.meshgrid(qx, qy, qz, indexing="ij")
    small = (qx[1] - qx[0]) / 1e4
    q2 = QX**2 + QY**2 + QZ**2
    fac = np.where(q2 > 0, (q2 + small) / np.where(q2 > 0, q2, 1.0), 0.0)
    eps = epsilon_0 * epsr
    k = lambda a: np.fft.fftn(a)  # noqa: E731
    er = [k(a) for a in e_r]
    ep = [k(a) for a in e_p]
    rk = k(rho)
    pk = [k(a) for a in P]
    qdp = QX * pk[0] + QY * pk[1] + QZ * pk[2]
    div_r = 1j * (QX * er[0] + QY * er[1] + QZ * er[2]) * fac
    assert np.max(np.abs(div_r - np.where(q2 > 0, rk / eps, 0))) < 1e-10 * np.max(np.abs(rk / eps))
    div_p = 1j * (QX * ep[0] + QY * ep[1] + QZ * ep[2]) * eps * fac
    assert np.max(np.abs(div_p - np.where(q2 > 0, -1j * qdp, 0))) < 1e-10 * np.max(np.abs(qdp))
    for E in (er, ep):
        curl = (QY * E[2] - QZ * E[1], QZ * E[0] - QX * E[2], QX * E[1] - QY * E[0])
        assert max(np.max(np.abs(c)) for c in curl) < 1e-12 * np.max(np.abs(QX * E[0])) + 1e-300
    for a, b, c in zip(e_pr, e_p, e_r):
        assert np.max(np.abs(a - (b + c))) < 1e-12 * np.max(np.abs(a))
    assert np.max(np.abs(k(rb) + 1j * qdp)) < 1e-12 * np.max(np.abs(qdp))
