"""GPU parity: every stage of the HIP path through the C ABI against the
oracle (FP64, tolerance stated per check), and full QWCalculator
trajectories against the reference's golden vectors."""

import numpy as np
import pytest

from oracle import sbe_oracle as O
from tests.helpers import (load_golden, params_pair, pulse, random_state, rel,
                           wire_length)

pytestmark = pytest.mark.gpu

RHS_TOL = 1e-12     # per RHS evaluation, normwise (max |diff| / max |ref|)
TRAJ_TOL = 1e-9     # trajectories


def _engine(prm, nw, rr, prec="fp64", wire_ids=None):
    from pulsesuite_amd.engine import WireGroup
    return WireGroup(prm, nw, rr, wire_ids=wire_ids, prec=prec)


class Case:
    """Matched device group and oracle solver on the same random state."""

    def __init__(self, N, nw, seed, flags=O.SHIPPED_FLAGS, prec="fp64", nx=96):
        import torch
        torch.cuda.synchronize()
        self.L = wire_length(N)
        prm, oprm = params_pair(self.L, flags)
        self.rr = np.linspace(-1.2 * self.L, 1.2 * self.L, nx)
        self.nw = nw
        print(f"[case] N={N} nw={nw} seed={seed} prec={prec}", flush=True)
        self.wg = _engine(prm, nw, self.rr, prec=prec)
        self.og = O.Grid(oprm)
        self.ot = O.Tables(self.og)
        self.os = O.Solver(oprm, nw, self.rr, 1.0, tables=self.ot, grid=self.og)
        rng = np.random.default_rng(seed)
        st = random_state(rng, N, nw)
        self.state = st
        self.wg.set_state(*st)
        for dst, src in zip((self.os.CC2, self.os.DD2, self.os.YY2,
                             self.os.CC3, self.os.DD3, self.os.YY3), st):
            dst[:] = src
        w0 = 2 * np.pi * 299792458.0 / 800e-9
        env = np.cos(2 * np.pi * self.rr / (3 * self.L)) + 0.4
        self.Exx = pulse(45e-15, 3e7, 10e-15, 50e-15, w0) * env + 0j
        self.Eyy = 2.0 * self.Exx
        self.Ezz = 0.3 * self.Exx
        self.wg.set_fields(self.Exx, self.Eyy, self.Ezz)

    def oracle_fields(self):
        g, t = self.og, self.ot
        return tuple(O.prop2qw(g, t, self.rr, E) for E in (self.Exx, self.Eyy, self.Ezz))


@pytest.fixture(scope="module")
def case64():
    return Case(64, 3, seed=11)


def test_library_loads_on_gpu():
    from pulsesuite_amd import capi
    lib = capi.load()
    assert lib.psa_abi_version() == capi.PSA_ABI_VERSION


def test_vint_table(case64):
    c = case64
    for i in range(3):
        assert rel(c.wg.tab["vtab"][i], c.ot.V0[i][:, 0]) < 1e-13


def test_prop2qw_fields(case64):
    import torch
    c = case64
    c.wg.prop2qw()
    torch.cuda.synchronize()
    N, Nr = c.wg.N, c.wg.Nr
    qb = np.arange(-(N - 1), N) % Nr
    exq = c.wg.s["exq"].cpu().numpy()
    for f, E in enumerate(c.oracle_fields()):
        assert rel(exq[0, f], E[qb]) < 1e-13


def test_prep_stage(case64):
    import torch
    c = case64
    c.wg.prep()
    torch.cuda.synchronize()
    N = c.wg.N
    pr = c.wg.s["prep"].cpu().numpy()
    for w in range(c.nw):
        C3, D3 = c.state[3][:, :, w], c.state[4][:, :, w]
        ne, nh = np.diag(C3).copy(), np.diag(D3).copy()
        VC, E1D = O.screened_arrays(c.og, c.ot, True, ne, nh)
        assert rel(pr[w, 4 * N:5 * N], E1D[:, 0]) < 1e-13
        for ch in range(3):
            assert rel(pr[w, (5 + ch) * N:(6 + ch) * N], VC[:, 0, ch]) < 1e-13
        g0 = 8 * N + 4 * (2 * N - 1)
        assert rel(pr[w, g0:g0 + N], O.gamma_e(c.og, c.ot, ne, nh, VC)) < 1e-12
        assert rel(pr[w, g0 + N:g0 + 2 * N], O.gamma_h(c.og, c.ot, ne, nh, VC)) < 1e-12
        assert rel(pr[w, g0 + 2 * N:g0 + 3 * N], O.spont_emission(c.og, c.ot, ne, nh, VC)) < 1e-12


def test_calc_h_stage(case64):
    import torch
    c = case64
    c.wg.prop2qw()
    c.wg.prep()
    c.wg.calc_h()
    torch.cuda.synchronize()
    Ex, Ey, Ez = c.oracle_fields()
    Heh_g, Heo, Hho = c.wg.hamiltonians()
    hd = c.wg.s["hdiag"].cpu().numpy()
    for w in range(c.nw):
        C3, D3, p3 = c.state[3][:, :, w], c.state[4][:, :, w], c.state[5][:, :, w]
        VC, _ = O.screened_arrays(c.og, c.ot, True, np.diag(C3), np.diag(D3))
        Meh = O.calc_meh(c.og, c.ot, Ex, Ey, Ez, w + 1)
        Heh, Hee, Hhh = O.calc_h(c.og, Meh, C3, D3, p3, VC, True, True)
        # the device stores the Hamiltonians transposed ([k2][k1])
        assert rel(Heh_g[w].T, Heh) < 1e-12
        off = lambda m: m - np.diag(np.diag(m))  # noqa: E731
        assert rel(Heo[w].T, off(Hee)) < 1e-12
        assert rel(Hho[w].T, off(Hhh)) < 1e-12
        assert rel(hd[w, 0], np.diag(Hee)) < 1e-14
        assert rel(hd[w, 1], np.diag(Hhh)) < 1e-14


@pytest.mark.parametrize("N", [256, 384, 512])
def test_calc_h_large_n(N):
    """Exchange sums at the serving sizes against the oracle: FP64 (GEMM
    engine) within 1e-12; FP32 (shifted-Toeplitz kernel on FP32 MFMA) and
    split-BF16 (the same blocks on bf16 MFMA, operands split hi/lo) within
    1e-5 of the largest entry; N = 384 has two row groups and a partial
    column block, N = 512 the largest Toeplitz row."""
    import torch
    from pulsesuite_amd.engine import WireGroup
    L = wire_length(N)
    prm, oprm = params_pair(L)
    nx = 96
    rr = np.linspace(-1.2 * L, 1.2 * L, nx)
    og = O.Grid(oprm)
    ot = O.Tables(og)
    nw = 2
    st = random_state(np.random.default_rng(N), N, nw)
    E = 3e7 * np.cos(2 * np.pi * rr / (3 * L)) + 0j
    fields = (E, 2.0 * E, 0.3 * E)
    oq = tuple(O.prop2qw(og, ot, rr, F) for F in fields)
    ref = []
    for w in range(nw):
        C3, D3, p3 = st[3][:, :, w], st[4][:, :, w], st[5][:, :, w]
        VC, _ = O.screened_arrays(og, ot, True, np.diag(C3), np.diag(D3))
        Meh = O.calc_meh(og, ot, *oq, w + 1)
        ref.append(O.calc_h(og, Meh, C3, D3, p3, VC, True, True))
    off = lambda m: m - np.diag(np.diag(m))  # noqa: E731
    for prec, tol in (("fp64", 1e-12), ("fp32", 1e-5), ("bf16x3", 1e-5)):
        print(f"[calc_h] N={N} prec={prec}", flush=True)
        wg = WireGroup(prm, nw, rr, prec=prec)
        wg.set_state(*st)
        wg.set_fields(*fields)
        wg.prop2qw()
        wg.prep()
        wg.calc_h()
        torch.cuda.synchronize()
        Heh_g, Heo, Hho = wg.hamiltonians()  # complex64 storage in the split-BF16 mode
        hd = wg.s["hdiag"].cpu().numpy()
        for w in range(nw):
            Heh, Hee, Hhh = ref[w]
            assert rel(Heh_g[w].T, Heh) < tol, (prec, w)
            assert rel(Heo[w].T, off(Hee)) < tol, (prec, w)
            assert rel(Hho[w].T, off(Hhh)) < tol, (prec, w)
            assert rel(hd[w, 0], np.diag(Hee)) < tol, (prec, w)
            assert rel(hd[w, 1], np.diag(Hhh)) < tol, (prec, w)
        del wg


def test_offdiag_tables(case64):
    import torch
    c = case64
    c.wg.prep()
    c.wg.offdiag()
    torch.cuda.synchronize()
    N = c.wg.N
    dd = c.wg.s["dedh"].cpu().numpy()
    g = c.og
    for w in range(c.nw):
        C3, D3 = c.state[3][:, :, w], c.state[4][:, :, w]
        ne, nh = np.diag(C3).copy(), np.diag(D3).copy()
        VC, _ = O.screened_arrays(c.og, c.ot, True, ne, nh)
        De = O.offdiag_deph_e2(g, ne, nh, VC, g.gam_eh, g.prm.gam_h)
        Dh = O.offdiag_deph_h2(g, ne, nh, VC, g.prm.gam_e, g.prm.gam_h)
        assert rel(dd[w, 0], De[1:2 * N, :].T) < 1e-12
        assert rel(dd[w, 1], Dh[1:2 * N, :].T) < 1e-12


@pytest.mark.parametrize("N,nw", [(256, 40), (384, 5)])
def test_offdiag_register_lorentzian_vs_table(N, nw):
    """The off-diagonal dephasing kernel that makes its Lorentzian operand in
    registers against the GEMM over the stored Lorentzian table (the form
    pinned against the oracle at N = 64 above, there on the same kernel
    family), on one random state: FP64 within 1e-12; FP32 (the reduced
    modes) within 1e-5 of FP64.  N = 384 covers two row groups, 40 wires a
    partial column tile."""
    import torch
    from pulsesuite_amd.engine import WireGroup
    L = wire_length(N)
    prm, _ = params_pair(L)
    rr = np.linspace(-1.2 * L, 1.2 * L, 64)
    st = random_state(np.random.default_rng(N + nw), N, nw)
    print(f"[offdiag] N={N} nw={nw}", flush=True)
    out = {}
    for tag, kw in (("table", dict(lor_table=True)), ("reg", {}), ("reg32", dict(prec="fp32"))):
        wg = WireGroup(prm, nw, rr, **kw)
        assert ("lor" in wg.t) == (tag == "table")
        wg.set_state(*st)
        wg.prep()
        wg.offdiag()
        torch.cuda.synchronize()
        out[tag] = wg.s["dedh"].cpu().numpy()
        del wg
    ref = out["table"]
    assert np.abs(ref).max() > 0
    assert rel(out["reg"], ref) < 1e-12
    for w in range(nw):
        for kind in range(2):
            assert rel(out["reg"][w, kind], ref[w, kind]) < 1e-12, (w, kind)
            assert rel(out["reg32"][w, kind], ref[w, kind]) < 1e-5, (w, kind)


def test_rhs_fp64(case64):
    import torch
    c = case64
    c.wg.rhs_eval()
    torch.cuda.synchronize()
    Ex, Ey, Ez = c.oracle_fields()
    dp_g = c.wg.s["pX"].cpu().numpy()
    dC_g = c.wg.s["CX"].cpu().numpy()
    dD_g = c.wg.s["DX"].cpu().numpy()
    for w in range(c.nw):
        C3, D3, p3 = c.state[3][:, :, w], c.state[4][:, :, w], c.state[5][:, :, w]
        dp, dC, dD, _ = c.os.rhs(p3, C3, D3, Ex, Ey, Ez, w + 1)
        assert rel(dp_g[w], dp) < RHS_TOL
        assert rel(dC_g[w], dC) < RHS_TOL
        assert rel(dD_g[w], dD) < RHS_TOL


@pytest.mark.parametrize("tag", ["n16_mb", "n16_rhs", "n32_mb", "n16_nolf", "n16_lztraj"])
def test_trajectory_vs_reference_golden(tag):
    """Full QWCalculator steps from the reference's own seeded state and
    fields (golden vectors written from the reference)."""
    import torch
    z = load_golden(tag)
    from pulsesuite_amd.tables import QWParams
    prm = QWParams(L=float(z["L"]), me_rel=0.067, flags=tuple(z["mb"]))
    print(f"[golden] {tag}", flush=True)
    wg = _engine(prm, 2, z["rr"])
    wg.set_state(z["s0_C2"], z["s0_D2"], z["s0_p2"], z["s0_C3"], z["s0_D3"], z["s0_p3"])
    dt = float(z["dt"])
    for it in range(int(z["nsteps"])):
        Exx = z["t_Exx"][it]
        wg.set_fields(Exx, z["t_Eyy"][it], np.zeros_like(Exx))
        wg.step(dt)
        torch.cuda.synchronize()
        Pxx, Pyy, Pzz, Rho = wg.outputs()
        for w in range(2):
            j = 2 * it + w
            assert rel(Pxx[w], z["t_Pxx"][j]) < TRAJ_TOL, (it, w)
            assert rel(Pyy[w], z["t_Pyy"][j]) < TRAJ_TOL, (it, w)
            assert rel(Pzz[w], z["t_Pzz"][j]) < TRAJ_TOL, (it, w)
            assert rel(Rho[w], z["t_Rho"][j]) < TRAJ_TOL, (it, w)
        C2, D2, p2, C3, D3, p3 = wg.get_state()
        for v, nm in ((C2, "C2"), (D2, "D2"), (p2, "p2"), (C3, "C3"), (D3, "D3"), (p3, "p3")):
            assert rel(v, z["t_" + nm][it]) < TRAJ_TOL, (it, nm)
        assert rel(wg.observables()["I0"], z["t_I0"][it]) < TRAJ_TOL


def test_step_vs_oracle_n64():
    """Five full steps, three wires (signs yw(w), (-1)^w differ per wire),
    N = 64, all mb.params terms on."""
    import torch
    c = Case(64, 3, seed=5)
    dt = 10e-18
    for it in range(5):
        c.wg.step(dt)
        torch.cuda.synchronize()
        Pxx, Pyy, Pzz, Rho = c.wg.outputs()
        for w in range(c.nw):
            ref = c.os.qw_calculator(c.Exx, c.Eyy, c.Ezz, dt, w + 1)
            assert ref is not None
            assert rel(Pxx[w], ref[0]) < TRAJ_TOL, (it, w)
            assert rel(Pyy[w], ref[1]) < TRAJ_TOL, (it, w)
            assert rel(Pzz[w], ref[2]) < TRAJ_TOL, (it, w)
            assert rel(Rho[w], ref[3]) < TRAJ_TOL, (it, w)
        got = c.wg.get_state()
        for v, r in zip(got, (c.os.CC2, c.os.DD2, c.os.YY2, c.os.CC3, c.os.DD3, c.os.YY3)):
            assert rel(v, r) < TRAJ_TOL, it


def _flags(**kw):
    names = O.FLAG_NAMES
    f = list(O.SHIPPED_FLAGS)
    for k, v in kw.items():
        f[names.index(k)] = v
    return tuple(f)


@pytest.mark.parametrize("what,flags", [("LF=0", _flags(LF=0)), ("LorentzDelta=1", _flags(LorentzDelta=1))])
def test_step_vs_oracle_mb_options(what, flags):
    """Three full steps of two wires at N = 64 with the mb.params options
    that change the fused step: LF = 0 (diagonal-only transport by direct
    DFTs, no QWRho5, no Hartree terms) and LorentzDelta = 1 (dense
    Lorentzian many-body lists); the oracle is pinned for both by the
    reference's trajectories (n16_nolf, n16_lztraj)."""
    import torch
    print(f"[mb options] {what}", flush=True)
    c = Case(64, 2, seed=17, flags=flags)
    dt = 10e-18
    for it in range(3):
        c.wg.step(dt)
        torch.cuda.synchronize()
        Pxx, Pyy, Pzz, Rho = c.wg.outputs()
        for w in range(c.nw):
            ref = c.os.qw_calculator(c.Exx, c.Eyy, c.Ezz, dt, w + 1)
            assert rel(Pxx[w], ref[0]) < TRAJ_TOL, (it, w)
            assert rel(Pyy[w], ref[1]) < TRAJ_TOL, (it, w)
            assert rel(Pzz[w], ref[2]) < TRAJ_TOL, (it, w)
            if np.max(np.abs(ref[3])) > 0:
                assert rel(Rho[w], ref[3]) < TRAJ_TOL, (it, w)
            else:
                assert np.max(np.abs(Rho[w])) == 0.0, (it, w)
        got = c.wg.get_state()
        for v, r in zip(got, (c.os.CC2, c.os.DD2, c.os.YY2, c.os.CC3, c.os.DD3, c.os.YY3)):
            assert rel(v, r) < TRAJ_TOL, it
        assert rel(c.wg.observables()["I0"], c.os.I0) < TRAJ_TOL


@pytest.mark.parametrize("N", [48, 64, 96, 128, 192, 256])
def test_transport_stage(N):
    """Transport alone (row and column FFT passes, radix 4/2/3 stages) on
    dense random p, C, D against fft2 -> phase -> ifft2."""
    import torch
    c = Case(N, 2, seed=31 + N)
    c.wg.prop2qw()
    rng = np.random.default_rng(N)
    X = {k: rng.standard_normal((2, N, N)) + 1j * rng.standard_normal((2, N, N))
         for k in ("CX", "DX", "pX")}
    for k, v in X.items():
        c.wg.s[k].copy_(torch.as_tensor(v))
    dt = 1e-14
    c.wg.transport(dt)
    torch.cuda.synchronize()
    exq = c.wg.s["exq"].cpu().numpy()
    for w in range(2):
        eac = exq[w, 0, N - 1].real
        assert abs(eac) > 0
        for k, v in X.items():
            ref = O.transport(c.ot, v[w], c.og.prm.Edc, eac, dt)
            assert rel(c.wg.s[k][w].cpu().numpy(), ref) < 1e-13, (k, w)


def test_step_vs_oracle_n48_radix3():
    """Three full steps at N = 48 (transport with a radix-3 stage, the
    leapfrog average fused into the transport column pass)."""
    import torch
    c = Case(48, 2, seed=9)
    dt = 10e-18
    for it in range(3):
        c.wg.step(dt)
        torch.cuda.synchronize()
        Pxx, Pyy, Pzz, Rho = c.wg.outputs()
        for w in range(c.nw):
            ref = c.os.qw_calculator(c.Exx, c.Eyy, c.Ezz, dt, w + 1)
            assert rel(Pxx[w], ref[0]) < TRAJ_TOL, (it, w)
            assert rel(Rho[w], ref[3]) < TRAJ_TOL, (it, w)
        got = c.wg.get_state()
        for v, r in zip(got, (c.os.CC2, c.os.DD2, c.os.YY2, c.os.CC3, c.os.DD3, c.os.YY3)):
            assert rel(v, r) < TRAJ_TOL, it


def test_mixed_precision_fp32_error_bounded():
    """FP32-MFMA contractions (operands rounded to FP32, FP32 accumulation;
    diagonals, epilogues and state in FP64) against the FP64 path after
    five steps: the state stays within 1e-6 relative."""
    import torch
    a = Case(64, 2, seed=21)
    b = Case(64, 2, seed=21, prec="fp32")
    for _ in range(5):
        a.wg.step(1e-17)
        b.wg.step(1e-17)
    torch.cuda.synchronize()
    Sa, Sb = a.wg.get_state(), b.wg.get_state()
    errs = [rel(x, y) for x, y in zip(Sb, Sa)]
    pa, pb = a.wg.outputs()[0], b.wg.outputs()[0]
    print("fp32 state errors", errs, "Pxx", rel(pb, pa))
    assert max(errs) < 1e-6
    assert rel(pb, pa) < 1e-6


def test_bf16x3_rhs_and_state_error_bounded():
    """Split-bf16 contractions (hi/lo planes, three products, FP32
    accumulation; diagonals, epilogue and state in FP64): one RHS evaluation
    against the FP64 path (normwise), then five steps of the state.  The
    emulated per-product error of the split is ~2^-17 relative, ten times
    FP32's, so the bounds here are ten times those of the FP32 test."""
    import torch
    a = Case(64, 2, seed=21)
    b = Case(64, 2, seed=21, prec="bf16x3")
    a.wg.rhs_eval()
    b.wg.rhs_eval()
    torch.cuda.synchronize()
    rerr = [rel(b.wg.s[k].cpu().numpy(), a.wg.s[k].cpu().numpy()) for k in ("pX", "CX", "DX")]
    print("bf16x3 rhs errors", rerr)
    assert max(rerr) < 1e-4
    a = Case(64, 2, seed=21)
    b = Case(64, 2, seed=21, prec="bf16x3")
    for _ in range(5):
        a.wg.step(1e-17)
        b.wg.step(1e-17)
    torch.cuda.synchronize()
    Sa, Sb = a.wg.get_state(), b.wg.get_state()
    errs = [rel(x, y) for x, y in zip(Sb, Sa)]
    pa, pb = a.wg.outputs()[0], b.wg.outputs()[0]
    print("bf16x3 state errors", errs, "Pxx", rel(pb, pa))
    assert max(errs) < 1e-5
    assert rel(pb, pa) < 1e-5


@pytest.mark.parametrize("N", [32, 64])
def test_many_body_lists_device_vs_host(N):
    """psa_mb_list (device) against the host builder: same entries in the
    same order, weights within 1e-13 (exp evaluated on the device); an entry
    present in only one list may only be an underflow-edge weight < 1e-300."""
    from pulsesuite_amd import tables
    from pulsesuite_amd.engine import WireGroup
    L = wire_length(N)
    prm, _ = params_pair(L)
    rr = np.linspace(-1.2 * L, 1.2 * L, 64)
    dev = WireGroup(prm, 1, rr, host_mb=False)
    g = dev.grid
    host = tables.many_body_lists(N, g["Ee"], g["Eh"], prm.gam_e, prm.gam_h)
    for li in range(8):
        off_d = dev.mb[li][0].cpu().numpy()
        n = dev.mb_nnz[li]
        idx_d = dev.mb[li][1].cpu().numpy()[:n].astype(np.uint32)
        val_d = dev.mb[li][2].cpu().numpy()[:n]
        off_h, idx_h, val_h = host[li]
        if np.array_equal(off_d, off_h) and np.array_equal(idx_d, idx_h):
            assert n == 0 or rel(val_d, val_h) < 1e-13, li
            continue
        dd = dict(zip(idx_d.tolist(), val_d.tolist()))
        hh = dict(zip(idx_h.tolist(), val_h.tolist()))
        for k in set(dd) ^ set(hh):
            assert max(dd.get(k, 0.0), hh.get(k, 0.0)) < 1e-300, (li, k)
        common = sorted(set(dd) & set(hh))
        a = np.array([dd[k] for k in common])
        b = np.array([hh[k] for k in common])
        assert rel(a, b) < 1e-13, li


@pytest.mark.parametrize("N,nw", [(64, 20), (256, 18)])
def test_many_body_rates_reduced_vs_fp64(N, nw):
    """The in/out rates of the reduced modes (FP32 many-body tables for 16
    wires per block, two wires per lane) against the FP64 kernel (pinned
    against the oracle through the step tests) on one random state, as
    psa_relaxation leaves them (many-body + phonon rates times the
    occupation factors): within 1e-5 per wire and rate; 18 and 20 wires
    cover a partial wire chunk."""
    import torch
    from pulsesuite_amd.engine import WireGroup
    L = wire_length(N)
    prm, _ = params_pair(L)
    rr = np.linspace(-1.2 * L, 1.2 * L, 64)
    st = random_state(np.random.default_rng(7 * N + nw), N, nw)
    print(f"[mb rates] N={N} nw={nw}", flush=True)
    out = {}
    for prec in ("fp64", "fp32"):
        wg = WireGroup(prm, nw, rr, prec=prec)
        wg.set_state(*st)
        for a, b in (("pX", "pB"), ("CX", "CB"), ("DX", "DB")):
            wg.s[a].copy_(wg.s[b])
        wg.prep()
        wg.relaxation(1e-17)
        torch.cuda.synchronize()
        out[prec] = wg.s["wmb"][:4 * N * nw].cpu().numpy().reshape(nw, 4, N)
        del wg
    ref = out["fp64"]
    assert np.abs(ref).max() > 0
    for w in range(nw):
        for r in range(4):
            assert rel(out["fp32"][w, r], ref[w, r]) < 1e-5, (w, r)


@pytest.mark.parametrize("N", [384, 512])
def test_transport_large_n(N):
    """Transport at the serving sizes (radix 4/2/3 stages up to 512) against
    fft2 -> phase -> ifft2; the ramp Y is the group's table (pinned against
    the oracle at N = 48..192 above)."""
    import torch
    from types import SimpleNamespace
    from pulsesuite_amd.engine import WireGroup
    L = wire_length(N)
    prm, _ = params_pair(L)
    rr = np.linspace(-1.2 * L, 1.2 * L, 128)
    print(f"[transport] N={N}", flush=True)
    wg = WireGroup(prm, 2, rr)
    E = 3e7 * np.cos(2 * np.pi * rr / (3 * L)) + 0j
    wg.set_fields(E, 2 * E, 0 * E)
    wg.prop2qw()
    rng = np.random.default_rng(N)
    X = {k: rng.standard_normal((2, N, N)) + 1j * rng.standard_normal((2, N, N)) for k in ("CX", "DX", "pX")}
    for k, v in X.items():
        wg.s[k].copy_(torch.as_tensor(v))
    dt = 1e-14
    wg.transport(dt)
    torch.cuda.synchronize()
    exq = wg.s["exq"].cpu().numpy()
    t = SimpleNamespace(Ytr=wg.tab["ytr"])
    for w in range(2):
        eac = exq[w, 0, N - 1].real
        for k, v in X.items():
            ref = O.transport(t, v[w], prm.Edc, eac, dt)
            assert rel(wg.s[k][w].cpu().numpy(), ref) < 1e-13, (k, w)


# ---------------------------------------------------------------------------
# The benchmarked sizes: multi-tile RHS products (tile offsets m0, n0 > 0),
# the compile-time N = 256 transport, the dephasing tables and the
# many-body sums at N = 256, and the RHS at N = 512, against the oracle.
# ---------------------------------------------------------------------------

# The oracle's tables are the slow part at N = 512 (Vint with the crude K0,
# about a minute), so one case per N serves every test below.
@pytest.fixture(scope="module")
def case256():
    return Case(256, 2, seed=256)


@pytest.fixture(scope="module")
def case512():
    return Case(512, 1, seed=512)


def _rhs_check(c, precs):
    import torch
    from pulsesuite_amd.engine import WireGroup
    Ex, Ey, Ez = c.oracle_fields()
    ref = []
    for w in range(c.nw):
        C3, D3, p3 = c.state[3][:, :, w], c.state[4][:, :, w], c.state[5][:, :, w]
        ref.append(c.os.rhs(p3, C3, D3, Ex, Ey, Ez, w + 1)[:3])
    for prec, tol in precs:
        print(f"[rhs] N={c.wg.N} nw={c.nw} prec={prec}", flush=True)
        wg = c.wg if prec == "fp64" else WireGroup(c.wg.prm, c.nw, c.rr, prec=prec)
        wg.set_state(*c.state)
        wg.set_fields(c.Exx, c.Eyy, c.Ezz)
        wg.rhs_eval()
        torch.cuda.synchronize()
        got = [wg.s[k].cpu().numpy() for k in ("pX", "CX", "DX")]
        for w in range(c.nw):
            for k in range(3):
                e = rel(got[k][w], ref[w][k])
                print(f"  wire {w} eq {k} rel err {e:.2e}", flush=True)
                assert e < tol, (prec, w, k)
        if wg is not c.wg:
            del wg


def test_rhs_vs_oracle_n256(case256):
    """One RHS evaluation (dp, dC, dD with every mb.params term) of two
    wires at N = 256 (4 x 4 output tiles per equation): FP64 within 1e-12
    normwise of the oracle; FP32 and split-BF16 within 1e-5 (the reduced
    modes' contractions, diagonals exact in FP64)."""
    _rhs_check(case256, (("fp64", RHS_TOL), ("fp32", 1e-5), ("bf16x3", 1e-5)))


def test_rhs_vs_oracle_n512(case512):
    """The same at N = 512 (8 x 8 tiles), one wire."""
    _rhs_check(case512, (("fp64", RHS_TOL), ("bf16x3", 1e-5)))


def test_offdiag_tables_n256(case256):
    """De, Dh (the FP64 register-Lorentzian kernel) against the oracle at
    N = 256."""
    import torch
    c = case256
    c.wg.set_state(*c.state)
    c.wg.prep()
    c.wg.offdiag()
    torch.cuda.synchronize()
    N = c.wg.N
    dd = c.wg.s["dedh"].cpu().numpy()
    g = c.og
    for w in range(c.nw):
        C3, D3 = c.state[3][:, :, w], c.state[4][:, :, w]
        ne, nh = np.diag(C3).copy(), np.diag(D3).copy()
        VC, _ = O.screened_arrays(c.og, c.ot, True, ne, nh)
        De = O.offdiag_deph_e2(g, ne, nh, VC, g.gam_eh, g.prm.gam_h)
        Dh = O.offdiag_deph_h2(g, ne, nh, VC, g.prm.gam_e, g.prm.gam_h)
        assert rel(dd[w, 0], De[1:2 * N, :].T) < 1e-12
        assert rel(dd[w, 1], Dh[1:2 * N, :].T) < 1e-12


def test_step_vs_oracle_n256(case256):
    """One full leapfrog step of two wires at N = 256 with every
    mb.params term (multi-tile RHS, N = 256 transport with the fused
    average, many-body and phonon relaxation, sources): outputs and every
    level within 1e-9 of the oracle."""
    import torch
    c = case256
    c.wg.set_state(*c.state)
    c.wg.set_fields(c.Exx, c.Eyy, c.Ezz)
    for dst, src in zip((c.os.CC2, c.os.DD2, c.os.YY2, c.os.CC3, c.os.DD3, c.os.YY3), c.state):
        dst[:] = src
    dt = 10e-18
    print("[step] N=256 nw=2", flush=True)
    c.wg.step(dt)
    torch.cuda.synchronize()
    Pxx, Pyy, Pzz, Rho = c.wg.outputs()
    for w in range(c.nw):
        ref = c.os.qw_calculator(c.Exx, c.Eyy, c.Ezz, dt, w + 1)
        assert ref is not None
        for got, r in zip((Pxx[w], Pyy[w], Pzz[w], Rho[w]), ref):
            assert rel(got, r) < TRAJ_TOL, w
    got = c.wg.get_state()
    for v, r in zip(got, (c.os.CC2, c.os.DD2, c.os.YY2, c.os.CC3, c.os.DD3, c.os.YY3)):
        assert rel(v, r) < TRAJ_TOL
