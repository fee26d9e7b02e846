"""GPU 3-D coupling (psa_qw3d_drive / psa_qw3d_deposit through the C ABI):
against the reference's own QuantumWire (tests/golden/ref_qw3d.npz: four
wires, three calls, random non-zero incoming J and Rho grids that placement
overwrites), against the oracle on another grid with and without the
keep-old mode, and a repeated deposit giving the same grids."""

import numpy as np
import pytest

from oracle import qw3d_oracle as Q3
from oracle import sbe_oracle as O
from tests.helpers import load_golden, params_pair, random_state, rel, wire_length

pytestmark = pytest.mark.gpu

TRAJ_TOL = 1e-9


def _torch():
    import torch
    return torch


def test_quantum_wire_3d_matches_reference():
    torch = _torch()
    from pulsesuite_amd.engine import WireGroup
    from pulsesuite_amd.qw3d import QuantumWire3D
    z = load_golden("qw3d")
    nw, N = int(z["nw"]), int(z["nk"])
    rx, ry, rz = z["rx"], z["ry"], z["rz"]
    print(f"[qw3d] reference vectors: grid {len(rx)}x{len(ry)}x{len(rz)} N={N} nw={nw}", flush=True)
    prm, _ = params_pair(float(z["L"]))
    wg = WireGroup(prm, nw, rx, prec="fp64")
    wg.set_state(*(z["s0_" + k] for k in ("C2", "D2", "p2", "C3", "D3", "p3")))
    dev = torch.device("cuda")
    G = [torch.as_tensor(z[k].copy(), dtype=torch.complex128, device=dev)
         for k in ("J0x", "J0y", "J0z", "Rho0")]
    dt = float(z["dt"])
    q3 = QuantumWire3D(wg, rx, ry, rz, z["qx"], z["qy"], z["qz"], float(z["y0"]), float(z["z0"]),
                       float(z["ay"]), float(z["az"]), float(z["epsr"]), dt)
    for it in range(z["q_Jx"].shape[0]):
        Ed = [torch.as_tensor(z["q_" + k][it], dtype=torch.complex128, device=dev) for k in ("Ex", "Ey", "Ez")]
        El = q3.step(*Ed, *G)
        torch.cuda.synchronize()
        for a, k in zip(G, ("Jx", "Jy", "Jz", "Rho")):
            assert rel(a.cpu().numpy(), z["q_" + k][it]) < TRAJ_TOL, (it, k)
        got = El.cpu().numpy()
        for c in range(10):
            assert rel(got[c], z["q_El"][it][c]) < TRAJ_TOL, (it, c)


def _setup(keep_old):
    torch = _torch()
    from pulsesuite_amd.engine import WireGroup
    from pulsesuite_amd.qw3d import QuantumWire3D
    nx, ny, nz, d = 64, 12, 8, 2e-9
    rx = (np.arange(nx) - nx // 2) * d
    ry = (np.arange(ny) - ny // 2) * d
    rz = (np.arange(nz) - nz // 2) * d
    qx, qy, qz = (2 * np.pi * np.fft.fftfreq(n, d) for n in (nx, ny, nz))
    y0, z0, ay, az, epsr, dt = 0.7e-9, -0.4e-9, 5e-9, 5e-9, 9.1, 10e-18
    N, nw = 16, 2
    L = wire_length(N)
    prm, oprm = params_pair(L)
    wg = WireGroup(prm, nw, rx, prec="fp64")
    sv = O.Solver(oprm, nw, rx, 0.0)
    rng = np.random.default_rng(17)
    st = random_state(rng, N, nw)
    wg.set_state(*st)
    for dst, src in zip((sv.CC2, sv.DD2, sv.YY2, sv.CC3, sv.DD3, sv.YY3), st):
        dst[:] = src
    X, Y, Z = np.meshgrid(rx, ry, rz, indexing="ij")
    env = np.exp(-(Y**2 + Z**2) / (8e-9) ** 2)
    shp = (nx, ny, nz)
    grids0 = [(rng.standard_normal(shp) + 1j * rng.standard_normal(shp)) * s
              for s in (1e10, 1e10, 1e10, 1e3)]
    dev = torch.device("cuda")
    G = [torch.as_tensor(a.copy(), dtype=torch.complex128, device=dev) for a in grids0]
    q3 = QuantumWire3D(wg, rx, ry, rz, qx, qy, qz, y0, z0, ay, az, epsr, dt, keep_old=keep_old)
    oq = Q3.QuantumWire3D(sv, rx, ry, rz, qx, qy, qz, y0, z0, ay, az, epsr, keep_old=keep_old)
    return wg, q3, oq, G, [a.copy() for a in grids0], (X, env, nx, d, dt)


@pytest.mark.parametrize("keep_old", [False, True])
def test_quantum_wire_3d_matches_oracle(keep_old):
    torch = _torch()
    print(f"[qw3d] oracle, grid 64x12x8 N=16 nw=2 keep_old={keep_old}", flush=True)
    wg, q3, oq, G, O_grids, (X, env, nx, d, dt) = _setup(keep_old)
    dev = torch.device("cuda")
    w0 = 2 * np.pi * 299792458.0 / 800e-9
    for it in range(3):
        t = 40e-15 + it * dt
        amp = [3e7, 1e7, 2e6]
        E = [(a * np.cos(w0 * t + ph) * env * (1 + 0.3 * np.cos(2 * np.pi * X / (nx * d)))).astype(complex)
             for a, ph in zip(amp, (0.0, 0.4, 1.1))]
        Ed = [torch.as_tensor(e, dtype=torch.complex128, device=dev) for e in E]
        El = q3.step(*Ed, *G)
        torch.cuda.synchronize()
        ref = oq.step(*E, dt, *O_grids)
        for a, b in zip(G, O_grids):
            assert rel(a.cpu().numpy(), b) < TRAJ_TOL, it
        got = El.cpu().numpy()
        e_pr, e_p, e_r, rb = ref
        for c, r in enumerate(list(e_pr) + list(e_p) + list(e_r) + [rb]):
            assert rel(got[c], r) < TRAJ_TOL, (it, c)
        P = wg.outputs()
        assert rel(P[0], np.moveaxis(oq.old[0], 1, 0)) < TRAJ_TOL


def test_repeated_deposit_gives_the_same_grids():
    """Depositing the same wire outputs twice leaves the same J and Rho
    (placement overwrites; with the reference's behaviour nothing carries
    over from the first deposit)."""
    torch = _torch()
    print("[qw3d] repeated deposit", flush=True)
    wg, q3, oq, G, _, (X, env, nx, d, dt) = _setup(False)
    dev = torch.device("cuda")
    E = [torch.as_tensor((a * env).astype(complex), dtype=torch.complex128, device=dev)
         for a in (3e7, 1e7, 2e6)]
    q3.drive(*E)
    wg.step(dt)
    q3.deposit(*G)
    first = [a.clone() for a in G] + [q3.El.clone()]
    q3.deposit(*G)
    torch.cuda.synchronize()
    for a, b in zip([*G, q3.El], first):
        assert torch.equal(a, b)
