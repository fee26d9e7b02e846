"""Write golden input/output vectors from the reference PulseSuite SBE path.

Runs ONLY in the build container, where the reference tree is readable at
``/root/reference``.  It is never imported by the tests: the tests read the
``.npz`` fixtures this script writes.

The reference imports ``numba`` and ``pyfftw`` unguarded (``coulomb.py:18``,
``usefulsubs.py:18``); neither is installed here.  They are stubbed before the
import: ``numba.jit`` becomes the identity decorator and ``prange`` is
``range`` (so the JIT kernels run as plain Python, IEEE arithmetic, no
fastmath), and ``pyfftw.interfaces.numpy_fft`` is ``numpy.fft`` (pinned equal
at rtol 1e-12 by ``tests/PSTD3D/test_usefulsubs.py:1104-1141``).
``numba.cuda`` is left absent, which the reference catches
(``SBEs.py:18-23``).

Only functions that no closed form, library routine or invariant pins are
recorded (composed and index-heavy behaviour):

* the grid and every init table of ``SBESolver.__init__`` (``SBEs.py:4623``),
  including ``CalcCoulombArrays``/``Vint``/``_K03_jit`` (``coulomb.py:38-177``)
  and ``CalcChi1D``;
* ``SBESolver._calc_H`` -> ``_CalcH_jit`` (``SBEs.py:808-924``);
* ``CalcGammaE``/``CalcGammaH`` (``dephasing.py:541-625``);
* ``OffDiagDephasing2`` (``dephasing.py:714-743``);
* ``MBCE``/``MBCH`` (``coulomb.py:1298-1326``);
* ``QWOptics.QWRho5`` (``qwoptics.py:942-1071``);
* ``SBESolver.QWCalculator`` trajectories (``SBEs.py:5826-5985``), with the
  shipped ``mb.params`` and with the many-body relaxation/transport switched
  off so the leapfrog RHS can be read back from the state;
* the snapshot files of the ``WriteFields`` steps (``jmax`` = 2: ``Wire/*``
  via ``WriteSBESolns``, ``WriteDephasing``, the relaxation rates,
  ``WriteQWFields``; ``Prop/*`` via ``WritePropFields``) as text;
* ``SBESolver.calc_Xqw`` -> ``_CalcXqw_jit`` (``SBEs.py:5331-5357``,
  ``3165-3243``) and the chi(q, w) files of
  ``_record_Xqw`` (``SBEs.py:5395-5432``);
* in the snapshot run, also ``output/PdotE.p.dat`` / ``PdotE.w.dat``
  (``SBEs.py:5925-5979``) and the text checkpoint of ``write_sbes_data``
  (``SBEs.py:5524-5547``); a run with ``ignorewire`` = 1 for the PdotE lines
  of that mode (``SBEs.py:5981-5985``);
* ``CalcMBArrays`` with ``LorentzDelta`` = 1 (``coulomb.py:367-499``) and
  ``MBCE``/``MBCH`` on those tables;
* ``rhoPJ.QuantumWire`` (``rhoPJ.py:117-449``): two wires on a small 3-D
  grid, three calls with non-zero incoming J and Rho grids;
* the module functions off the QWCalculator path (``extras_case``).

Usage: ``python tests/golden/make_golden.py`` (writes ``tests/golden/*.npz``;
``python tests/golden/make_golden.py xqw`` writes only ``ref_xqw.npz``,
``... snap`` only ``ref_n16_snap.npz`` and ``ref_n16_ignore.npz``,
``... lorentz`` only ``ref_n16_lorentz.npz``, ``... qw3d`` only
``ref_qw3d.npz``, ``... flags`` the LF = 0 and LorentzDelta = 1
trajectories, ``... extras`` only ``ref_modx.npz``).
"""

import contextlib
import io
import os
import sys
import tempfile
import types

import numpy as np

REF_SRC = "/root/reference/src"
OUT_DIR = os.path.dirname(os.path.abspath(__file__))

MB_SHIPPED = [1, 1, 1, 1, 1, 1, 1, 0, 1, 1, 1, 0, 0, 1, 0]
# Optics, Excitons, EHs, Screened, Phonon, DCTrans, LF, FreePot, DiagDph,
# OffDiagDph, Recomb, PLSpec, ignorewire, Xqwparams, LorentzDelta
MB_RHS_ONLY = [1, 1, 0, 1, 0, 0, 1, 0, 1, 1, 0, 0, 0, 0, 0]
MB_IGNORE = [1, 1, 1, 1, 1, 1, 1, 0, 1, 1, 1, 0, 1, 1, 0]
MB_LORENTZ = [1, 1, 1, 1, 1, 1, 1, 0, 1, 1, 1, 0, 0, 1, 1]
MB_NOLF = [1, 1, 1, 1, 1, 1, 0, 0, 1, 1, 1, 0, 0, 1, 0]


def _install_shims():
    numba = types.ModuleType("numba")

    def jit(*args, **kwargs):
        if args and callable(args[0]) and len(args) == 1 and not kwargs:
            return args[0]

        def deco(f):
            return f

        return deco

    numba.jit = jit
    numba.njit = jit
    numba.prange = range
    sys.modules["numba"] = numba

    pyfftw = types.ModuleType("pyfftw")
    interfaces = types.ModuleType("pyfftw.interfaces")
    cache = types.ModuleType("pyfftw.interfaces.cache")
    cache.enable = lambda *a, **k: None
    interfaces.cache = cache
    interfaces.numpy_fft = np.fft
    pyfftw.interfaces = interfaces
    sys.modules["pyfftw"] = pyfftw
    sys.modules["pyfftw.interfaces"] = interfaces
    sys.modules["pyfftw.interfaces.cache"] = cache
    sys.modules["pyfftw.interfaces.numpy_fft"] = np.fft


def wire_length(nk, me_rel=0.067, gap_ev=1.5):
    """Wire length L whose derived grid has exactly ``nk`` points."""
    from scipy.constants import e, hbar, m_e

    kmax = np.sqrt(1.2 * gap_ev * e * 2.0 * me_rel * m_e / hbar**2)
    return (nk / 2 + 0.5) * np.pi / kmax


def write_params(d, L, mb, me_rel=0.067, jmax=500):
    os.makedirs(os.path.join(d, "params"), exist_ok=True)
    qw = [L, 5e-9, 1.5, me_rel, 0.45, 100e-3, 1e12, 1e12, 1e12, 9.1,
          36e-3, 3e-3, 0.0, jmax, 20000]
    with open(os.path.join(d, "params", "qw.params"), "w") as f:
        for v in qw:
            f.write(f"{v if isinstance(v, int) else float(v)!r}\n")
    with open(os.path.join(d, "params", "mb.params"), "w") as f:
        for v in mb:
            f.write(f"{v}\n")


def random_state(rng, n, occ=0.3, coh=0.05, pol=0.02):
    """Hermitian C, D with diagonal occupations and off-diagonal coherences,
    and a dense complex p."""
    def herm(scale_d, scale_o):
        a = (rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n)))
        a = (a + a.conj().T) * 0.5 * scale_o
        np.fill_diagonal(a, rng.uniform(0.0, scale_d, n))
        return a

    C = herm(occ, coh)
    D = herm(occ, coh)
    p = pol * (rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n)))
    return C, D, p


def pulse(t, E0, tw, tp, w0):
    """sbetestprop.py:187-192 analytic pulse."""
    x = w0 * (t - tp)
    return (E0 * np.exp(-(x**2) / (w0 * tw) ** 2) * np.cos(x)
            * np.exp(-(x**20) / (2 * tw * w0) ** 20))


def run_case(nk, seed, mb, nsteps, tag, nx=64, drr=10e-9, jmax=500, files=False):
    rng = np.random.default_rng(seed)
    L = wire_length(nk)
    out = {}
    with tempfile.TemporaryDirectory() as d:
        cwd = os.getcwd()
        os.chdir(d)
        try:
            write_params(d, L, mb, jmax=jmax)
            buf = io.StringIO()
            with contextlib.redirect_stdout(buf):
                from pulsesuite.PSTD3D import (SBEs, coulomb, dephasing,
                                               emission, phonons)
                from pulsesuite.PSTD3D.typespace import (GetKArray,
                                                         GetSpaceArray)

                rr = GetSpaceArray(nx, (nx - 1) * drr)
                qrr = GetKArray(nx, nx * drr)
                lam = 800e-9
                E0x, E0y = 2e7, 4e7
                s = SBEs.SBESolver(qrr, rr, 0.0, np.hypot(E0x, E0y), lam, 2, True)
            n = s.Nk
            assert n == nk, (n, nk)
            ci = coulomb._instance
            out.update(
                L=L, mb=np.array(mb), nk=n, nr=s.Nr, kr=s.kr, r=s.r, Qr=s.Qr,
                NQ0=s.NQ0, kkp=s.kkp, Ee=s.Ee, Eh=s.Eh, dcv=s.dcv,
                ehint=s.ehint, area=s.area, alphae=s.alphae, alphah=s.alphah,
                qc=s.qc, dkr=s.dkr, dr=s.dr, rr=rr, qrr=qrr,
                Veh0=ci.Veh0, Vee0=ci.Vee0, Vhh0=ci.Vhh0, qe=ci.qe, qh=ci.qh,
                Chi1De=ci.Chi1De, Chi1Dh=ci.Chi1Dh,
                EP=phonons._EP, HP=phonons._HP, Vscale=phonons._Vscale,
                NO=phonons._NO,
                k_p_q=dephasing._k_p_q, k_m_q=dephasing._k_m_q,
                k1_m_q=dephasing._k1_m_q, k1p_m_q=dephasing._k1p_m_q,
                k1=dephasing._k1, k1p=dephasing._k1p,
                xe=dephasing._xe, xh=dephasing._xh,
                Expikr=s.qw._Expikr, QWWindow=s.qw._QWWindow,
            )
            if s.Recomb:
                out.update(HOmega=emission._HOmega, square=emission._square,
                           RScale=emission._RScale)
            if nk <= 16:
                out.update(Ceh=ci.Ceh, Cee=ci.Cee, Chh=ci.Chh, k3=ci.k3)

            # ---- per-function vectors on a random many-body state ----
            C, D, p = random_state(rng, n)
            ne = np.diag(C).copy()
            nh = np.diag(D).copy()
            out.update(fC=C, fD=D, fp=p)
            VC = np.zeros((n, n, 3))
            E1D = np.zeros((n, n))
            with contextlib.redirect_stdout(buf):
                coulomb.CalcScreenedArrays(True, s.L, ne, nh, VC, E1D)
            out.update(fVC=VC.copy(), fE1D=E1D.copy())
            # Hamiltonians with a random dipole-coupling matrix
            Meh = 1e-22 * (rng.standard_normal((n, n))
                           + 1j * rng.standard_normal((n, n)))
            Z = np.zeros((n, n), dtype=complex)
            Heh = np.zeros((n, n), dtype=complex)
            Hee = np.zeros((n, n), dtype=complex)
            Hhh = np.zeros((n, n), dtype=complex)
            s._calc_H(Meh, Z, Z, C, D, p, VC, Heh, Hee, Hhh)
            out.update(fMeh=Meh, fHeh=Heh, fHee=Hee, fHhh=Hhh)
            GamE = np.zeros(n)
            GamH = np.zeros(n)
            dephasing.CalcGammaE(s.kr, ne, nh, VC, GamE)
            dephasing.CalcGammaH(s.kr, ne, nh, VC, GamH)
            out.update(fGamE=GamE, fGamH=GamH)
            g = np.array([s.gam_eh, s.gam_e, s.gam_h])
            OffG = np.zeros((n, n), dtype=complex)
            dephasing.OffDiagDephasing2(ne, nh, p, s.kr, s.Ee, s.Eh, g, VC,
                                        0.0, OffG)
            out.update(fOffP=OffG)
            WinE = np.zeros(n)
            WoutE = np.zeros(n)
            WinH = np.zeros(n)
            WoutH = np.zeros(n)
            coulomb.MBCE(np.real(ne), np.real(nh), s.kr, s.Ee, s.Eh, VC,
                         s.gam_eh, s.gam_e, WinE, WoutE)
            coulomb.MBCH(np.real(ne), np.real(nh), s.kr, s.Ee, s.Eh, VC,
                         s.gam_eh, s.gam_h, WinH, WoutH)
            out.update(fWinE=WinE, fWoutE=WoutE, fWinH=WinH, fWoutH=WoutH)
            re = np.zeros(s.Nr, dtype=complex)
            rh = np.zeros(s.Nr, dtype=complex)
            s.qw.QWRho5(s.Qr, s.kr, s.r, s.L, s.kkp, p, C, D, ne, nh, re, rh,
                        1, 1)
            out.update(fre=re, frh=rh)

            # ---- QWCalculator trajectory from a seeded many-body state ----
            for w in range(2):
                C0, D0, p0 = random_state(rng, n, occ=0.2, coh=0.02, pol=0.01)
                C1, D1, p1 = random_state(rng, n, occ=0.2, coh=0.02, pol=0.01)
                s.CC2[:, :, w] = C0
                s.DD2[:, :, w] = D0
                s.YY2[:, :, w] = p0
                s.CC3[:, :, w] = C1
                s.DD3[:, :, w] = D1
                s.YY3[:, :, w] = p1
            out.update(s0_C2=s.CC2.copy(), s0_D2=s.DD2.copy(),
                       s0_p2=s.YY2.copy(), s0_C3=s.CC3.copy(),
                       s0_D3=s.DD3.copy(), s0_p3=s.YY3.copy())
            dt = 10e-18
            w0 = 2 * np.pi * 299792458.0 / lam
            Exx = np.zeros(nx, dtype=complex)
            Eyy = np.zeros(nx, dtype=complex)
            Ezz = np.zeros(nx, dtype=complex)
            Vrr = np.zeros(nx, dtype=complex)
            traj = {k: [] for k in ("Pxx", "Pyy", "Pzz", "Rho", "Exx", "Eyy",
                                    "C2", "D2", "p2", "C3", "D3", "p3", "I0")}
            t = 40e-15
            for it in range(nsteps):
                # a spatially varying field so Prop2QW's spline is exercised
                env = np.cos(2 * np.pi * rr / (nx * drr)) + 0.3
                Exx[:] = pulse(t, E0x, 10e-15, 50e-15, w0) * env
                Eyy[:] = pulse(t, E0y, 10e-15, 50e-15, w0) * env
                for w in (1, 2):
                    Pxx = np.zeros(nx, dtype=complex)
                    Pyy = np.zeros(nx, dtype=complex)
                    Pzz = np.zeros(nx, dtype=complex)
                    Rho = np.zeros(nx, dtype=complex)
                    with contextlib.redirect_stdout(buf):
                        s.QWCalculator(Exx, Eyy, Ezz, Vrr, rr, qrr, dt, w, Pxx,
                                       Pyy, Pzz, Rho, [True], [False])
                    traj["Pxx"].append(Pxx)
                    traj["Pyy"].append(Pyy)
                    traj["Pzz"].append(Pzz)
                    traj["Rho"].append(Rho)
                traj["Exx"].append(Exx.copy())
                traj["Eyy"].append(Eyy.copy())
                traj["C2"].append(s.CC2.copy())
                traj["D2"].append(s.DD2.copy())
                traj["p2"].append(s.YY2.copy())
                traj["C3"].append(s.CC3.copy())
                traj["D3"].append(s.DD3.copy())
                traj["p3"].append(s.YY3.copy())
                traj["I0"].append(s.I0.copy())
                t += dt
            for k, v in traj.items():
                out["t_" + k] = np.array(v)
            out.update(dt=dt, t0=40e-15, nsteps=nsteps, E0x=E0x, E0y=E0y,
                       lam=lam, Emax0=np.hypot(E0x, E0y))
            info = {}
            for w in (1, 2):
                fn = os.path.join("dataQW", f"info.{w:02d}.t.dat")
                with open(fn) as f:
                    info[w] = np.array([[complex(tok) for tok in line.split()]
                                        for line in f if line.strip()])
            out["info1"] = info[1]
            out["info2"] = info[2]
            out["jmax"] = jmax
            if files:
                # the text checkpoint of the final levels (SBEs.py:5524-5547)
                s.write_sbes_data(7)
                s.close()
                # every file the run wrote (the chi(q, w) record excepted):
                # the Wire/* and Prop/* snapshots of the WriteFields steps,
                # the window file, info/EP files, the backups, PdotE
                names, texts = [], []
                walk = list(os.walk("dataQW")) + list(os.walk("output"))
                for root, _, fs in walk:
                    for fn in sorted(fs):
                        path = os.path.join(root, fn)
                        if "Xqw" in path:
                            continue
                        names.append(path)
                        with open(path) as f:
                            texts.append(f.read())
                out["file_names"] = np.array(names)
                out["file_texts"] = np.array(texts)
            s.close()
        finally:
            os.chdir(cwd)
    path = os.path.join(OUT_DIR, f"ref_{tag}.npz")
    np.savez_compressed(path, **out)
    print("wrote", path, "N =", nk)


def lorentz_case(nk=16, seed=11, nx=64, drr=10e-9):
    """CalcMBArrays with LorentzDelta = 1 (coulomb.py:405-446) as
    SBESolver.__init__ builds it from mb.params line 15, and MBCE/MBCH
    (coulomb.py:1298-1326) on those tables at a random state."""
    rng = np.random.default_rng(seed)
    L = wire_length(nk)
    out = {}
    with tempfile.TemporaryDirectory() as d:
        cwd = os.getcwd()
        os.chdir(d)
        try:
            write_params(d, L, MB_LORENTZ)
            buf = io.StringIO()
            with contextlib.redirect_stdout(buf):
                from pulsesuite.PSTD3D import SBEs, coulomb
                from pulsesuite.PSTD3D.typespace import GetKArray, GetSpaceArray
                rr = GetSpaceArray(nx, (nx - 1) * drr)
                qrr = GetKArray(nx, nx * drr)
                s = SBEs.SBESolver(qrr, rr, 0.0, 1e7, 800e-9, 1, True)
            n = s.Nk
            assert n == nk
            ci = coulomb._instance
            assert ci.LorentzDelta
            C, D, p = random_state(rng, n)
            ne = np.diag(C).copy()
            nh = np.diag(D).copy()
            VC = np.zeros((n, n, 3))
            E1D = np.zeros((n, n))
            with contextlib.redirect_stdout(buf):
                coulomb.CalcScreenedArrays(True, s.L, ne, nh, VC, E1D)
            W = [np.zeros(n) for _ in range(4)]
            coulomb.MBCE(np.real(ne), np.real(nh), s.kr, s.Ee, s.Eh, VC,
                         s.gam_eh, s.gam_e, W[0], W[1])
            coulomb.MBCH(np.real(ne), np.real(nh), s.kr, s.Ee, s.Eh, VC,
                         s.gam_eh, s.gam_h, W[2], W[3])
            out.update(nk=n, L=L, kr=s.kr, Ee=s.Ee, Eh=s.Eh, gam_e=s.gam_e,
                       gam_h=s.gam_h, gam_eh=s.gam_eh, Ceh=ci.Ceh, Cee=ci.Cee,
                       Chh=ci.Chh, k3=ci.k3, fC=C, fD=D, fVC=VC,
                       fWinE=W[0], fWoutE=W[1], fWinH=W[2], fWoutH=W[3])
            s.close()
        finally:
            os.chdir(cwd)
    path = os.path.join(OUT_DIR, "ref_n16_lorentz.npz")
    np.savez_compressed(path, **out)
    print("wrote", path, "N =", nk)


def qw3d_case(nk=16, seed=13, nx=32, ny=8, nz=6, d=2e-9, nsteps=3, nw=4):
    """rhoPJ.QuantumWire, i.e. the shim over QWArray.QuantumWire
    (rhoPJ.py:1180-1240, 941-1160; same steps as rhoPJ.py:117-449), for the
    wires of params/qwarray.params on a [nx][ny][nz] grid built from
    typespace.ss, after InitializeSBE on the grid's x axis: the wires start
    from seeded random levels, the incoming Jx, Jy, Jz and Rho grids are
    random and non-zero, the fields are smooth and strong enough to open the
    wire-off gate.  Records every grid after each call.  QWArray allocates
    its per-wire arrays with its default Nw = 4 before it reads the file, so
    only a file with Nw = 4 runs."""
    rng = np.random.default_rng(seed)
    L = wire_length(nk)
    out = {}
    with tempfile.TemporaryDirectory() as dd:
        cwd = os.getcwd()
        os.chdir(dd)
        try:
            write_params(dd, L, MB_SHIPPED)
            with open(os.path.join(dd, "params", "qwarray.params"), "w") as f:
                f.write(f"1\n{nw}\n0.0e-9\n0.0e-9\n5.0e-9\n4.0e-9\n")
            buf = io.StringIO()
            with contextlib.redirect_stdout(buf):
                from pulsesuite.PSTD3D import SBEs, rhoPJ
                from pulsesuite.PSTD3D import typespace as TS
                space = TS.ss(Dims=3, Nx=nx, Ny=ny, Nz=nz, dx=d, dy=d, dz=d, epsr=9.1)
                rx, ry, rz = TS.GetXArray(space), TS.GetYArray(space), TS.GetZArray(space)
                qx, qy, qz = TS.GetKxArray(space), TS.GetKyArray(space), TS.GetKzArray(space)
                lam = 800e-9
                SBEs.InitializeSBE(qx, rx, 0.0, 3e7, lam, nw, True)
            s = SBEs._default_solver
            n = s.Nk
            assert n == nk
            for w in range(nw):
                C0, D0, p0 = random_state(rng, n, occ=0.2, coh=0.02, pol=0.01)
                C1, D1, p1 = random_state(rng, n, occ=0.2, coh=0.02, pol=0.01)
                s.CC2[:, :, w] = C0
                s.DD2[:, :, w] = D0
                s.YY2[:, :, w] = p0
                s.CC3[:, :, w] = C1
                s.DD3[:, :, w] = D1
                s.YY3[:, :, w] = p1
            out.update(s0_C2=s.CC2.copy(), s0_D2=s.DD2.copy(), s0_p2=s.YY2.copy(),
                       s0_C3=s.CC3.copy(), s0_D3=s.DD3.copy(), s0_p3=s.YY3.copy())
            shp = (nx, ny, nz)
            X, Y, Z = np.meshgrid(rx, ry, rz, indexing="ij")
            env = np.exp(-(Y**2 + Z**2) / (8e-9) ** 2)
            grids = [(rng.standard_normal(shp) + 1j * rng.standard_normal(shp)) * sc
                     for sc in (1e10, 1e10, 1e10, 1e3)]
            out.update(J0x=grids[0].copy(), J0y=grids[1].copy(), J0z=grids[2].copy(),
                       Rho0=grids[3].copy())
            El = [np.zeros(shp, dtype=complex) for _ in range(10)]
            dt = 10e-18
            w0 = 2 * np.pi * 299792458.0 / lam
            rec = {k: [] for k in ("Ex", "Ey", "Ez", "Jx", "Jy", "Jz", "Rho", "El")}
            for it in range(nsteps):
                t = 40e-15 + it * dt
                E = [(a * np.cos(w0 * t + ph) * env * (1 + 0.3 * np.cos(2 * np.pi * X / (nx * d))))
                     .astype(complex) for a, ph in zip((3e7, 1e7, 2e6), (0.0, 0.4, 1.1))]
                with contextlib.redirect_stdout(buf):
                    rhoPJ.QuantumWire(space, dt, it + 1, *E, *grids, *El)
                for k, v in zip(("Ex", "Ey", "Ez"), E):
                    rec[k].append(v)
                for k, v in zip(("Jx", "Jy", "Jz", "Rho"), grids):
                    rec[k].append(v.copy())
                rec["El"].append(np.array([a.copy() for a in El]))
            for k, v in rec.items():
                out["q_" + k] = np.array(v)
            qa = rhoPJ._default_qw
            out.update(nk=n, L=L, rx=rx, ry=ry, rz=rz, qx=qx, qy=qy, qz=qz, dt=dt,
                       y0=qa.y0, z0=qa.z0, ay=qa.ay, az=qa.az, epsr=9.1,
                       nw=qa.Nw, lam=lam, Emax0=3e7)
            s.close()
        finally:
            os.chdir(cwd)
    path = os.path.join(OUT_DIR, "ref_qw3d.npz")
    np.savez_compressed(path, **out)
    print("wrote", path, "N =", nk)


def xqw_case(nk=16, seed=5, nx=64, drr=10e-9):
    """SBESolver.calc_Xqw (SBEs.py:5331-5357 -> 3165-3243) at random occupations and dephasing for
    a spread of (iq, w), and the head and tail of the chi(q, w) file and the
    Xqw.params file that SBESolver.__init__ writes (_record_Xqw,
    SBEs.py:5395-5432)."""
    rng = np.random.default_rng(seed)
    L = wire_length(nk)
    out = {}
    with tempfile.TemporaryDirectory() as d:
        cwd = os.getcwd()
        os.chdir(d)
        try:
            write_params(d, L, MB_SHIPPED)
            buf = io.StringIO()
            with contextlib.redirect_stdout(buf):
                from pulsesuite.PSTD3D import SBEs
                from pulsesuite.PSTD3D.typespace import GetKArray, GetSpaceArray
                rr = GetSpaceArray(nx, (nx - 1) * drr)
                qrr = GetKArray(nx, nx * drr)
                s = SBEs.SBESolver(qrr, rr, 0.0, 1e7, 800e-9, 1, True)
            n = s.Nk
            assert n == nk
            fe = rng.uniform(0.0, 0.6, n)
            fh = rng.uniform(0.0, 0.6, n)
            game = rng.uniform(1e11, 1e13, n)
            gamh = rng.uniform(1e11, 1e13, n)
            game[0] = 1.0  # below the 0.1 meV floor
            wmax = s.gap / SBEs.hbar * 2.0
            iqs = np.array([0, 1, 2, n // 2, n - 1, n, -1, -(n // 2)])
            ws = np.array([0.0, 0.37, 0.5, 0.93, 1.0]) * wmax
            # the method (SBEs.py:5331-5357) passes the solver's Nk; the module
            # function reads a module global that only the legacy API sets
            X = np.array([[s.calc_Xqw(int(iq), float(w), s.kr, fe, fh, s.Ee, s.Eh, s.gap,
                                      s.area, game, gamh, s.dcv) for w in ws] for iq in iqs])
            out.update(nk=n, L=L, kr=s.kr, Ee=s.Ee, Eh=s.Eh, gap=s.gap, area=s.area,
                       dcv=s.dcv, dkr=s.dkr, gam_e=s.gam_e, gam_h=s.gam_h, fe=fe, fh=fh,
                       game=game, gamh=gamh, iqs=iqs, ws=ws, X=X)
            with open(os.path.join("dataQW", "Wire", "Xqw", "Xqw.000000.dat")) as f:
                lines = f.readlines()
            out["xqw_nlines"] = len(lines)
            out["xqw_head"] = np.array(lines[:300])
            out["xqw_tail"] = np.array(lines[-300:])
            with open(os.path.join("dataQW", "Wire", "Xqw", "Xqw.params")) as f:
                out["xqw_params"] = np.array(f.readlines())
            s.close()
        finally:
            os.chdir(cwd)
    path = os.path.join(OUT_DIR, "ref_xqw.npz")
    np.savez_compressed(path, **out)
    print("wrote", path, "N =", nk)


def extras_case(nk=16, seed=17, nx=64, drr=10e-9):
    """The module functions off the QWCalculator path, on a solver's own
    grid and tables: CalcMVeh, BGRenorm, EeRenorm, EhRenorm
    (coulomb.py:1234-1280), Vint and Vehint (coulomb.py:107-242),
    GetChi1Dqw and GetEps1Dqw (coulomb.py:633-776), Calchw and PLSpectrum
    (emission.py:258-357), QWChi1 and GetVn1n2 (qwoptics.py:1109-1186), and
    the files of printITReal2, printITReal, printIT3D, WriteSBESolns,
    WritePLSpectrum and WritePropFields (qwoptics.py:1189-1386) as text."""
    rng = np.random.default_rng(seed)
    L = wire_length(nk)
    out = {}

    def crand(*shape):
        return rng.normal(size=shape) + 1j * rng.normal(size=shape)

    with tempfile.TemporaryDirectory() as d:
        cwd = os.getcwd()
        os.chdir(d)
        try:
            write_params(d, L, MB_SHIPPED)
            buf = io.StringIO()
            with contextlib.redirect_stdout(buf):
                from pulsesuite.PSTD3D import SBEs, coulomb, emission, qwoptics
                from pulsesuite.PSTD3D.typespace import GetKArray, GetSpaceArray
                rr = GetSpaceArray(nx, (nx - 1) * drr)
                qrr = GetKArray(nx, nx * drr)
                s = SBEs.SBESolver(qrr, rr, 0.0, 1e7, 800e-9, 1, True)
            n = s.Nk
            assert n == nk
            C, D, p = random_state(rng, n)
            ne = np.diag(C).copy()
            nh = np.diag(D).copy()
            VC = np.zeros((n, n, 3))
            E1D = np.zeros((n, n))
            with contextlib.redirect_stdout(buf):
                coulomb.CalcScreenedArrays(True, s.L, ne, nh, VC, E1D)
            out.update(nk=n, L=L, kr=s.kr, r=s.r, Ee=s.Ee, Eh=s.Eh, gap=s.gap, area=s.area,
                       dcv=s.dcv, ehint=s.ehint, epsr=s.epsr, me=s.me, mh=s.mh, Delta0=s.Delta0,
                       alphae=s.alphae, alphah=s.alphah, gam_e=s.gam_e, gam_h=s.gam_h,
                       gam_eh=s.gam_eh, fC=C, fD=D, fp=p, fVC=VC)
            # CalcMVeh on a three-frequency p[q, qp, f]
            p3 = crand(n, n, 3)
            MV = np.zeros((n, n, 3), dtype=complex)
            coulomb.CalcMVeh(p3, VC, MV)
            out.update(p3=p3, MVeh=MV)
            B = np.zeros((n, n), dtype=complex)
            coulomb.BGRenorm(C, D, VC, B)
            out["BGR"] = B.copy()
            coulomb.EeRenorm(ne, VC, B)
            out["EeR"] = B.copy()
            coulomb.EhRenorm(nh, VC, B)
            out["EhR"] = B.copy()
            # Vint / Vehint on the solver's r grid
            qs = np.array([0.0, 1e5, 0.37 * s.kr[-1], -2.0 * s.kr[3], 5e9])
            out["vint_q"] = qs
            out["vint"] = np.array([coulomb.Vint(q, s.r, s.alphae, s.alphah, s.Delta0) for q in qs])
            kq = np.array([[0, 0], [1, 1], [1, 5], [7, 2], [n, 1], [3, n]])
            out["vehint_kq"] = kq
            out["vehint"] = np.array([coulomb.Vehint(int(a), int(b), s.r, s.kr, s.alphae, s.alphah,
                                                     s.Delta0) for a, b in kq])
            # chi(q, w) and eps(q, w) diagnostics
            game = rng.uniform(1e11, 1e13, n)
            gamh = rng.uniform(1e11, 1e13, n)
            game[0] = 1.0
            fe = rng.uniform(0.0, 0.6, n)
            fh = rng.uniform(0.0, 0.6, n)
            dkr = s.kr[1] - s.kr[0]
            wg = s.gap / SBEs.hbar
            qw = [(0.0, 0.3 * wg), (dkr, 0.01 * wg), (3.2 * dkr, 0.2 * wg), (-4.0 * dkr, 0.05 * wg),
                  (0.5, 1.0 * wg), (n * dkr, 0.7 * wg)]
            chi = [coulomb.GetChi1Dqw(s.alphae, s.alphah, s.Delta0, s.epsr, game, gamh, s.kr, s.Ee,
                                      s.Eh, fe, fh, q, w) for q, w in qw]
            n1d = 2.0e7
            qw2 = [(0.5, 1e12), (1.0e6, 1e12), (3.0e7, 2e13), (8.0e7, 5e12), (2.0e8, 1e14), (6.3e7, 4.0e12),
                   (1.0e8, 9.0e13)]
            eps = [coulomb.GetEps1Dqw(s.alphae, s.alphah, s.Delta0, s.epsr, s.me, s.mh, n1d, q, w)
                   for q, w in qw2]
            out.update(game=game, gamh=gamh, fe=fe, fh=fh, chi_qw=np.array(qw), chi=np.array(chi),
                       eps_n1d=n1d, eps_qw=np.array(qw2), eps=np.array(eps))
            # photoluminescence
            hw = np.zeros(40)
            PLS = np.zeros(40)
            emission.Calchw(hw, PLS, s.gap - 0.05 * SBEs.eV, s.gap + 0.25 * SBEs.eV)
            out["hw"] = hw.copy()
            PLS[:] = rng.uniform(0.0, 1e-7, 40)
            out["pls0"] = PLS.copy()
            nep = np.abs(ne) * np.exp(0.3j * np.arange(n))
            nhp = np.abs(nh) * np.exp(-0.2j * np.arange(n))
            # t enters only through softtheta(t, geh) (emission.py:361)
            t_pl = 0.37 * s.gam_eh
            emission.PLSpectrum(nep, nhp, s.Ee, s.Eh, s.gap, s.gam_eh, VC, hw, t_pl, PLS)
            out.update(pl_ne=nep, pl_nh=nhp, pl_t=t_pl, pls=PLS.copy(), rscale=emission._RScale)
            out["qwchi1"] = qwoptics.QWChi1(800e-9, dkr, s.Ee, s.Eh, s.area, s.gam_eh, s.dcv)
            # GetVn1n2
            rcv = crand(n)
            Hcc, Hhh, Hcv = crand(n, n), crand(n, n), crand(n, n)
            V = [np.zeros((n, n), dtype=complex) for _ in range(4)]
            qwoptics.GetVn1n2(s.kr, rcv, Hcc, Hhh, Hcv, *V)
            out.update(rcv=rcv, Hcc=Hcc, Hhh=Hhh, Hcv=Hcv, Vcc=V[0], Vvv=V[1], Vcv=V[2], Vvc=V[3])
            # version-1 off-diagonal dephasing (dephasing.py:107-225, 627-712) on
            # the solver's index tables, with complex occupations and a random p
            from pulsesuite.PSTD3D import dephasing
            ge1, gh1, geh1 = 3.0e12, 4.5e12, 2.2e12
            od_ne = nep
            od_nh = nhp
            od_p = crand(n, n) * 0.05
            DE = dephasing.CalcOffDiagDeph_E(od_ne, od_nh, s.kr, s.Ee, s.Eh, ge1, geh1, VC)
            DH = dephasing.CalcOffDiagDeph_H(od_ne, od_nh, s.kr, s.Ee, s.Eh, gh1, geh1, VC)
            X1 = np.zeros((n, n), dtype=complex)
            dephasing.OffDiagDephasing(od_ne, od_nh, od_p, s.kr, s.Ee, s.Eh, [ge1, gh1, geh1], VC, X1)
            out.update(od_g=np.array([ge1, gh1, geh1]), od_ne=od_ne, od_nh=od_nh, od_p=od_p, od_DE=DE,
                       od_DH=DH, od_x=X1, od_kpq=dephasing._instance._k_p_q, od_kmq=dephasing._instance._k_m_q)
            # file writers (qwoptics._firsttime starts True: printITReal2's first file has two columns)
            f1 = crand(8)
            y8 = np.linspace(-1e-7, 1e-7, 8)
            qwoptics._firsttime = True
            qwoptics.printITReal2(f1, y8, 3, "t/r2a.")
            qwoptics.printITReal2(f1, y8, 4, "t/r2b.")
            qwoptics.printITReal(f1, y8, 5, "t/r1.")
            c3 = crand(3, 4, 2)
            qwoptics.printIT3D(c3, y8, 6, "t/c3.")
            qwoptics._firsttime = True
            F = [crand(8) for _ in range(9)]
            qwoptics.WritePropFields(y8, *F, "r", 2, 7)
            qwoptics.WriteSBESolns(s.kr, ne, nh, C, D, p, s.Ee + 0j, s.Eh + 0j, 2, 8)
            qwoptics.WritePLSpectrum(hw, PLS, 2, 9)
            out.update(f1=f1, y8=y8, F=np.array(F), c3=c3)
            names, texts = [], []
            for root, _, files in os.walk("dataQW"):
                for fn in sorted(files):
                    pth = os.path.join(root, fn)
                    if not (pth.startswith(os.path.join("dataQW", "t")) or fn.endswith(("0000007.dat",
                            "000007.dat", "000008.dat", "0000008.dat", "000009.dat"))):
                        continue
                    if "Wire" in pth and not any(t in pth for t in ("/C/", "/D/", "/P/", "/ne/", "/nh/",
                                                                     "/Ee/", "/Eh/", "/PL/")):
                        continue
                    names.append(os.path.relpath(pth, "dataQW"))
                    with open(pth) as fh_:
                        texts.append(fh_.read())
            order = np.argsort(names)
            out["file_names"] = np.array(names)[order]
            out["file_texts"] = np.array(texts)[order]
            s.close()
        finally:
            os.chdir(cwd)
    path = os.path.join(OUT_DIR, "ref_modx.npz")
    np.savez_compressed(path, **out)
    print("wrote", path, "N =", nk, "files", len(out["file_names"]))


def main():
    if not os.path.isdir(REF_SRC):
        raise SystemExit("reference tree not found; run in the build container")
    # the reference tree is read only: no bytecode caches written next to it
    sys.dont_write_bytecode = True
    _install_shims()
    sys.path.insert(0, REF_SRC)
    if sys.argv[1:] == ["xqw"]:
        xqw_case()
        return
    if sys.argv[1:] == ["snap"]:
        run_case(16, seed=6, mb=MB_SHIPPED, nsteps=3, tag="n16_snap", jmax=2, files=True)
        run_case(16, seed=9, mb=MB_IGNORE, nsteps=3, tag="n16_ignore", files=True)
        return
    if sys.argv[1:] == ["lorentz"]:
        lorentz_case()
        return
    if sys.argv[1:] == ["qw3d"]:
        qw3d_case()
        return
    if sys.argv[1:] == ["flags"]:
        flag_cases()
        return
    if sys.argv[1:] == ["extras"]:
        extras_case()
        return
    run_case(16, seed=1, mb=MB_SHIPPED, nsteps=4, tag="n16_mb")
    run_case(16, seed=2, mb=MB_RHS_ONLY, nsteps=3, tag="n16_rhs")
    run_case(32, seed=3, mb=MB_SHIPPED, nsteps=3, tag="n32_mb")
    run_case(16, seed=6, mb=MB_SHIPPED, nsteps=3, tag="n16_snap", jmax=2, files=True)
    run_case(16, seed=9, mb=MB_IGNORE, nsteps=3, tag="n16_ignore", files=True)
    xqw_case()
    lorentz_case()
    qw3d_case()
    flag_cases()
    extras_case()


def flag_cases():
    """QWCalculator trajectories with the mb.params options that change the
    fused step: LF = 0 (diagonal-only DC transport, dcfield.py:1229-1232, no
    QWRho5, Hartree terms off) and LorentzDelta = 1 (Lorentzian many-body
    tables, coulomb.py:405-446)."""
    run_case(16, seed=21, mb=MB_NOLF, nsteps=3, tag="n16_nolf")
    run_case(16, seed=22, mb=MB_LORENTZ, nsteps=3, tag="n16_lztraj")


if __name__ == "__main__":
    main()
