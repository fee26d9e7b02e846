/*
 * pulsesuite_amd: MI355X (gfx950) quantum-wire SBE time step, C ABI.
 *
 * The reference (rahulranjansah/pulsesuite) is a Python module API with no
 * FFI; each entry point below replaces one reference function (or one
 * composed step) and is what a ctypes/cffi binding of that function binds.
 * Citations are relative to src/pulsesuite/PSTD3D/ of the reference.
 *
 * Conventions
 *  - Every entry point returns 0 on success, a negative PSA_E* code on
 *    failure, never throws or aborts; psa_last_error() gives the message
 *    (thread-local).
 *  - Pointers inside psa_group are DEVICE pointers (HBM), complex values are
 *    interleaved (re, im) doubles, matrices row-major [N][N], per-wire
 *    blocks wire-major ([Nw][...]).  `stream` is a hipStream_t (NULL = the
 *    default stream).  Calls are asynchronous on that stream.
 *  - A psa_group describes Nw wires that share one k-grid size N (one
 *    SBESolver of the reference).  It is a plain struct the caller fills;
 *    the library keeps no state between calls except lazily-built kernels.
 *  - Calls that write the same group's buffers must not overlap.
 */
#ifndef PULSESUITE_AMD_H
#define PULSESUITE_AMD_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define PSA_ABI_VERSION 5

#define PSA_OK 0
#define PSA_E_INVALID (-1)   /* bad argument / shape */
#define PSA_E_HIP (-2)       /* HIP runtime error */
#define PSA_E_UNSUPPORTED (-3)

/* precision modes of the dense contractions */
#define PSA_PREC_FP64 0      /* FP64 MFMA, reference-equivalent */
#define PSA_PREC_FP32 1      /* FP32 MFMA operands, FP32 accumulate */
#define PSA_PREC_BF16X3 2    /* BF16 hi/lo split, 3 products, FP32 accumulate */

/* mb.params flags (SBEs.py:5004-5022), bit i = line i+1 */
#define PSA_F_OPTICS (1 << 0)
#define PSA_F_EXCITONS (1 << 1)
#define PSA_F_EHS (1 << 2)
#define PSA_F_SCREENED (1 << 3)
#define PSA_F_PHONON (1 << 4)
#define PSA_F_DCTRANS (1 << 5)
#define PSA_F_LF (1 << 6)
#define PSA_F_FREEPOT (1 << 7)
#define PSA_F_DIAGDPH (1 << 8)
#define PSA_F_OFFDIAGDPH (1 << 9)
#define PSA_F_RECOMB (1 << 10)

typedef struct psa_group {
  /* ---- sizes ---- */
  int32_t N;        /* k points per wire (Nk) */
  int32_t Nr;       /* QW real-space points, 2N+2 */
  int32_t Nw;       /* wires in the group */
  int32_t Nx;       /* propagation-grid points (rr) */
  int32_t Nhw;      /* emission photon-energy points */
  int32_t flags;    /* PSA_F_* */
  int32_t prec;     /* PSA_PREC_* */
  int32_t gx0, gnx; /* support of the Prop2QW transform in rr: [gx0, gx0+gnx) */
  int32_t bx0, bnx; /* rows of rr inside the wire for QW2Prop: [bx0, bx0+bnx) */
  int32_t nnz_mb[8];/* entries in each many-body scattering list */
  int32_t tune[4];  /* launch choices, 0 = default: [0] reserved (0),
                       [1] transport lines per block, row pass (0: 4, else 2, 4, 8 or 16),
                       [2] GEMM block order (0: batches grouped per XCD, 1: linear; the
                       split-BF16 RHS has a static schedule of its own),
                       [3] transport lines per block, column pass (0: 4, else 2, 4, 8 or 16;
                       16 and the column choice only for N known at compile time: 64, 96,
                       128, 192, 256, 384, 512; other N take 4 for 16) */
  int32_t pad0;
  /* ---- scalars (SI units) ---- */
  double hbar, e0, gap, L, ehint, dcv, area, dkr, dr_r, drr;
  double gam_e, gam_h, gam_eh, Edc, Vscale, n1d_max, dhw;
  /* ---- per-N tables (read only) ---- */
  const double *kr, *Ee, *Eh;         /* [N] */
  const double *vtab;                 /* [3][N] unscreened V0(d), d = k-k' >= 0 (eh, ee, hh) */
  const double *chi;                  /* [2][N] Chi1De(d), Chi1Dh(d) */
  const double *qeh;                  /* [2][N] qe(d), qh(d) */
  const double *EP, *HP;              /* [N][N] phonon tables */
  const double *EPT, *HPT;            /* [N][N] their transposes */
  const double *xe, *xh;              /* [N] dephasing weights */
  const int32_t *didx;                /* [6][N k][N q] k_p_q, k_m_q, k1p_m_q, k1p, k1_m_q, k1 */
  const int32_t *iq;                  /* [N] Vxx2 index rint(|ky/dq|) */
  const double *homega, *hsq;         /* [Nhw] emission grid and weights */
  const double *window;               /* [Nr] QW window */
  const double *rq;                   /* [Nr] QW real-space grid r */
  const double *ytr;                  /* [N] Transport phase ramp Y */
  const double *gmat;                 /* [2N-1][gnx] complex: Prop2QW (spline, window, FFTG) */
  const double *m2;                   /* [Nr][bnx] real: QW2Prop spline r -> rr (transposed) */
  const double *phi;                  /* [2N-1][Nr] complex exp(i d dkr r) */
  const double *twid;                 /* [N] complex exp(-2 pi i j / N) */
  double *lor;                        /* [2][N*N][2N] Lorentzian kernels of CalcOffDiagDeph_E2/H2 for
                                         De, Dh: rows (q, k) packed per q over the N-|q| rows with
                                         k+q in [0, N); filled once by psa_lorentz_table */
  const int32_t *mb_off[8];           /* CSR offsets [N+1] per many-body list */
  const uint32_t *mb_idx[8];          /* packed (i, j, l) per entry */
  const double *mb_val[8];            /* table value per entry */
  /* ---- per-wire data ---- */
  const int32_t *wire_id;             /* [Nw] 1-based reference wire index w (host-visible copy in wid_host) */
  double *pA, *pB, *pX;               /* [Nw][N][N] complex: YY2, YY3, new level */
  double *CA, *CB, *CX;
  double *DA, *DB, *DX;
  double *Heh, *Heo, *Hho;            /* [Nw][N][N] complex, TRANSPOSED: element [k2][k1] holds
                                         Heh[k1][k2] (resp. Hee, Hhh without their diagonal) */
  double *hdiag;                      /* [Nw][2][N] complex: diagonals of Hee, Hhh */
  double *prep;                       /* [Nw][PSA_PREP_STRIDE(N)] doubles, see PSA_PREP_* */
  double *dedh;                       /* [Nw][2][N][2N-1] off-diagonal dephasing tables D[q][k] stored [k][q + N - 1];
                                         allocate zeroed: entries with k + q outside [0, N) and q = 0 are
                                         zero and the split-BF16 table kernel does not write them */
  double *wmb;                        /* 11 N Nw doubles: [Nw][4][N] many-body Win_e, Wout_e, Win_h, Wout_h
                                         (after psa_relaxation: the total rates times the occupation
                                         factors Win_e (1 - ne), Wout_e ne, Win_h (1 - nh), Wout_h nh);
                                         then, wire fastest: [2][N][Nw] |Re ne3|, |Re nh3| of the new level,
                                         [3][N][Nw] VC(d)^2 (eh, ee, hh), [2][N][Nw] |ne|, |nh| of the checkout level */
  double *exq;                        /* [Nw][3][2N-1] complex: FFTG'd Ex, Ey, Ez at q = -(N-1)..N-1 */
  double *efield;                     /* [3][Nw][Nx] complex fields on rr (Exx, Eyy, Ezz); set by the caller or psa_drive */
  double *src;                        /* [Nw][PSA_SRC_STRIDE] real-space Px, Py, Pz, re, rh + diagonal sums */
  double *out;                        /* [4][Nw][Nx] complex: Pxx, Pyy, Pzz, Rho */
  double *obs;                        /* [Nw][PSA_OBS_N] observables (I0, totals, ...) */
  double *tphase;                     /* [Nw][N] complex Transport phase exp(-i Y dk) */
  double *part;                       /* [Nw][256] scratch (reserved) */
  void *spl;                          /* bf16 split operands, PSA_PREC_BF16X3 only, zero-initialised:
                                         [Nw][14 slots][NR/64][NK/32][4 planes][64][32] with
                                         NR = 128 ceil(N/128), NK = 32 ceil(N/32); slot = 2 matrix +
                                         transposed (p, C_o, D_o, Heh, Hee_o, Hhh_o), then the
                                         dephasing bands Dh, De; planes re_hi, re_lo, im_hi, im_lo;
                                         the 8-k groups of a row XOR-swizzled by (row >> 1) & 3.
                                         Slots 6-11 (the Hamiltonians) stay unused: the RHS kernel
                                         splits the complex64 Hamiltonians as it stages them */
  double *kscr;                       /* optional scratch for split-K partial sums of the source and
                                         propagation-grid products (NULL: no split); kscr_len doubles,
                                         8 x 2 x max(Nr x 4 Nw, (2N-1) x 3 Nw, bnx x 5 Nw) uses all splits */
  int64_t kscr_len;
  const int32_t *wid_host;            /* [Nw] host copy of wire_id */
  double *pR, *CR, *DR;               /* [Nw][N][N] complex RK4 stage accumulators (psa_step_rk4,
                                         psa_rhs modes 2-5 only; NULL otherwise) */
  uint64_t *wtag;                     /* [Nw][2] per-wire reuse tags, zero-initialised, owned by the
                                         library (NULL: nothing is reused): [0] the device address of
                                         the p buffer (+ the wire's offset) whose p, C, D planes spl
                                         holds, [1] that of the C buffer whose diagonal sums the src
                                         block holds.  A caller that writes p, C or D of a wire other
                                         than through the step entry points zeroes both of its tags. */
  int32_t *lazy;                      /* [Nw] deferred leapfrog average, zero-initialised (NULL: the
                                         steps average in place, as the reference does).  With it set,
                                         psa_step, psa_step_multi and psa_step_noalloc leave level A
                                         (pA, CA, DA) holding the level the step added to (X1) and set
                                         lazy[w] = 1 instead of writing X2 = (X1 + X3) / 2 there: the
                                         averaged level is (A + B) / 2 once the caller has made the new
                                         level B.  The next leapfrog RHS forms it where it reads X1 and
                                         writes it back, and psa_average_levels writes it for host
                                         reads; psa_finish and the RK4 step leave lazy[w] = 0.  A
                                         caller that writes level A or B of a wire sets lazy[w] = 0
                                         after calling psa_average_levels (or overwriting both). */
} psa_group;

/* layout of the per-wire prep block (doubles) */
#define PSA_PREP_NE 0                         /* [N] complex */
#define PSA_PREP_NH(N) (2 * (N))              /* [N] complex */
#define PSA_PREP_E1D(N) (4 * (N))             /* [N] */
#define PSA_PREP_VC(N) (5 * (N))              /* [3][N] screened V(d) */
#define PSA_PREP_DSC(N) (8 * (N))             /* [2N-1] complex diagonal sums of C */
#define PSA_PREP_DSD(N) (8 * (N) + 2 * (2 * (N) - 1))
#define PSA_PREP_GAME(N) (8 * (N) + 4 * (2 * (N) - 1))
#define PSA_PREP_GAMH(N) (PSA_PREP_GAME(N) + (N))
#define PSA_PREP_RSP(N) (PSA_PREP_GAME(N) + 2 * (N))
#define PSA_PREP_ANE(N) (PSA_PREP_GAME(N) + 3 * (N))  /* [N] |ne| */
#define PSA_PREP_ANH(N) (PSA_PREP_GAME(N) + 4 * (N))  /* [N] |nh| */
#define PSA_PREP_STRIDE(N) (PSA_PREP_GAME(N) + 5 * (N))

#define PSA_OBS_N 8 /* I0, sum ne2, sum nh2, total(rho); max/min De, max/min Dh (psa_offdiag_extrema) */
/* per-wire src block: [5][Nr] Px, Py, Pz, re, rh (real space, final);
 * [2][Nr] complex re, rh before the LF normalisation; [4][2N-1] complex
 * diagonal sums; 8 scalars */
#define PSA_SRC_RC(Nr) (5 * (Nr))
#define PSA_SRC_DS(Nr) (9 * (Nr))
#define PSA_SRC_STRIDE(N, Nr) (9 * (Nr) + 8 * (2 * (N) - 1) + 8)

/* ---- library ---- */
int psa_abi_version(void);
const char *psa_last_error(void);
/* sizeof(psa_group), so bindings can check their struct mirror */
int psa_group_size(void);

/* Unscreened Coulomb table: out[c][k] = pref * Vint(Qy[k], y, a1_c, a2_c,
 * Delta0) for c in (eh, ee, hh) with the reference's crude K0.
 * Replaces coulomb.CalcCoulombArrays's Vint loop (coulomb.py:502-565,
 * Vint coulomb.py:107-177, _Vint_jit 69-104, _K03_jit 38-66). */
int psa_vint_table(const double *Qy, int32_t NQ, const double *y, int32_t Ny,
                   double alphae, double alphah, double Delta0, double pref,
                   double *out, void *stream);

/* Per-wire preparation: diag of C2/D2, CalcScreenedArrays (coulomb.py:1195-1230),
 * diagonal sums for the Hartree terms, CalcGammaE/H (dephasing.py:541-625),
 * SpontEmission (emission.py:94-130).  SBEs.py:5173-5226 minus CalcH. */
int psa_prep(const psa_group *g, int32_t w0, int32_t nw, void *stream);

/* Hartree-Fock Hamiltonians incl. the dipole coupling Meh:
 * _CalcH_jit (SBEs.py:808-924) + _calc_Meh (SBEs.py:5153-5163). */
int psa_calc_h(const psa_group *g, int32_t w0, int32_t nw, void *stream);

/* Off-diagonal dephasing tables De, Dh: CalcOffDiagDeph_E2/H2
 * (dephasing.py:226-317, 381-414) as used by OffDiagDephasing2 (714-743). */
int psa_offdiag_tables(const psa_group *g, int32_t w0, int32_t nw, void *stream);

/* Many-body scattering list `list` (0..7) on the device: the non-zero
 * entries (i, j, l) of the stored tables of CalcMBArrays (coulomb.py:367-499;
 * the Gaussian branch 448-497, or with lorentz != 0 the LorentzDelta branch
 * 405-446, whose entries are all non-zero) in the order
 * _MBCE_core/_MBCH_core consume them (coulomb.py:956-1071), grouped by
 * target index (the slot mapping of psa_group.mb_*).  Two calls: off ==
 * NULL counts the entries per target into counts[N]; then, with off[N+1]
 * the exclusive prefix sum, fills idx (packed i | j << 10 | l << 20) and
 * val.  Init-time. */
int psa_mb_list(int32_t list, int32_t N, const double *Ee, const double *Eh, double ge, double gh,
                double hbar, int32_t lorentz, const int32_t *off, int32_t *counts, uint32_t *idx,
                double *val, void *stream);

/* The state-independent Lorentzian factors Lrtz(dE, hbar g) of
 * CalcOffDiagDeph_E2/H2 (dephasing.py:226-317; Lrtz dephasing.py:220-224),
 * for every (q, k, p), into g->lor.  Init-time: call once per group after
 * the energy tables are set (psa_offdiag_tables reads it). */
int psa_lorentz_table(const psa_group *g, void *stream);

/* dpdt/dCdt/dDdt (SBEs.py:1941-1965, 2308-2327, 2606-2625) incl. the
 * off-diagonal dephasing OffP (_OffDiagDephasing2_jit, dephasing.py:320-356).
 * mode 0: write dp/dt, dC/dt, dD/dt into pX, CX, DX.
 * mode 1: leapfrog X3 = X1 + 2 dt dX/dt(X2) into pX, CX, DX (SBEs.py:5673-5675).
 * modes 2-5: classic RK4 stage s = mode - 1 (the stage combination of _rk4,
 * integrator.py:143-157) with k_s = dX/dt of the stage input in B and the
 * step start in A: s = 1: R = k1, X = A + dt/2 k1; s = 2: R += 2 k2,
 * X = A + dt/2 k2; s = 3: R += 2 k3, X = A + dt k3; s = 4:
 * X = A + dt/6 (R + k4) (X may alias A).  R = pR, CR, DR. */
int psa_rhs(const psa_group *g, int32_t w0, int32_t nw, int32_t mode,
            double dt, void *stream);

/* Relaxation of the new level's diagonal: MBCE/MBCH (coulomb.py:1298-1326),
 * MBPE/MBPH (phonons.py:157-268), the exponential in/out update and the
 * Recomb factor (SBEs.py:5276-5325). */
int psa_relaxation(const psa_group *g, int32_t w0, int32_t nw, double dt,
                   void *stream);

/* Transport of CX, DX, pX (dcfield.py:1176-1236, DCTrans and LF set). */
int psa_transport(const psa_group *g, int32_t w0, int32_t nw, double dt,
                  void *stream);

/* Diagonal renormalisation and averaging X2 = (X1+X3)/2 (SBEs.py:5692-5706);
 * leaves the new YY2 in *A and the new level YY3 in *X (the caller swaps
 * B and X afterwards). */
int psa_finish(const psa_group *g, int32_t w0, int32_t nw, void *stream);

/* QWPolarization3 (qwoptics.py:809-940), QWRho5 (942-1071) with the
 * 2 e0 ehint/area scaling (SBEs.py:5749-5750), QW2Prop's LF renormalisation
 * (qwoptics.py:785-790) and CalcI0 (dcfield.py:1110-1152), in real space. */
int psa_sources(const psa_group *g, int32_t w0, int32_t nw, void *stream);

/* Prop2QW (qwoptics.py:667-715) for Ex, Ey, Ez of every wire. */
int psa_prop2qw(const psa_group *g, int32_t w0, int32_t nw, void *stream);

/* QW2Prop splines (qwoptics.py:792-805) and Rho = RhoH - RhoE (SBEs.py:5917). */
int psa_qw2prop(const psa_group *g, int32_t w0, int32_t nw, void *stream);

/* Driver side of a served simulation, on the device so a whole step can be
 * graph-captured.  psa_drive: the sbetestprop pulse (sbetestprop.py:183-198),
 * spatially uniform, E_f = E0_f exp(-(w0 (t-tp))^2 / (w0 tw)^2) cos(w0 (t-tp))
 * exp(-((w0 (t-tp)) / (2 tw w0))^20), written to efield for wires
 * [w0, w0+nw) at each wire's own time; drive[w][8] = E0x, E0y, E0z, omega0,
 * tw, tp, t, unused; t advances by dt.  psa_record: the per-step record of
 * sbetestprop.py:210-224 (Pxx at the middle point) plus the carrier totals:
 * trace[w][count[w]][4] = Re Pxx[xmid], Im Pxx[xmid], sum ne2, sum nh2, then
 * count[w]++ (nothing once count[w] == maxn). */
int psa_drive(const psa_group *g, int32_t w0, int32_t nw, double *drive, double dt, void *stream);
int psa_record(const psa_group *g, int32_t w0, int32_t nw, int32_t xmid, double *trace,
               int32_t *count, int32_t maxn, void *stream);

/* One full SBESolver.QWCalculator step (SBEs.py:5826-5985 minus file I/O)
 * for wires [w0, w0+nw): prop2qw, prep, calc_h, offdiag, rhs (leapfrog),
 * relaxation, transport, finish, sources, qw2prop. */
int psa_step(const psa_group *g, int32_t w0, int32_t nw, double dt, void *stream);

/* psa_step_noalloc: NO COUNTERPART IN THE REFERENCE.  For callers that
 * cannot allocate device memory (code under HIP graph capture, tight
 * serving loops, harnesses that forbid hipMalloc): psa_step with the
 * split-K partial sums of its Prop2QW, source and QW2Prop products in a
 * device array owned by the library (PSA_NOALLOC_SCRATCH_DOUBLES doubles,
 * 32 MiB, held on every GPU the library runs on until the process ends)
 * instead of g->kscr, which it ignores.  Calls no hipMalloc / hipFree.
 * Results equal psa_step's with a full-size kscr (the same splits).
 * psa_step_noalloc_scratch gives the doubles a call with nw wires needs
 * (16 max(Nr 4 nw, (2N-1) 3 nw, bnx 5 nw)); larger calls are refused with
 * PSA_E_INVALID.  The array is shared: calls that overlap on one GPU race. */
#define PSA_NOALLOC_SCRATCH_DOUBLES (4L << 20)
int64_t psa_step_noalloc_scratch(const psa_group *g, int32_t nw);
int psa_step_noalloc(const psa_group *g, int32_t w0, int32_t nw, double dt, void *stream);

/* The same step with classic RK4 (the stage combination of _rk4,
 * integrator.py:143-157) in place of the leapfrog + average.  The reference
 * has no RK4 on the wire path; it integrates with the leapfrog of psa_step.
 * The state at the step start is level B (pB, CB, DB) and the new state is
 * written back there; the four RHS evaluations (prep, calc_h, offdiag, rhs on
 * the stage input) use the field of the step start; then relaxation (dt),
 * transport (dt), the diagonal renormalisation (no average), sources and
 * qw2prop on the new state, with the screening of the last stage.  Levels A
 * and X are stage scratch afterwards (a later leapfrog step needs A = B);
 * pR, CR, DR are required. */
int psa_step_rk4(const psa_group *g, int32_t w0, int32_t nw, double dt, void *stream);

/* max and min of the off-diagonal dephasing tables De, Dh of the last
 * psa_offdiag_tables call (the MaxOffDiag.e/h.dat lines written by
 * OffDiagDephasing2, dephasing.py:733-743) into obs[w][4..7]:
 * max De, min De, max Dh, min Dh. */
int psa_offdiag_extrema(const psa_group *g, int32_t w0, int32_t nw, void *stream);

/* chi(q, w) table that SBESolver.__init__ records (_record_Xqw,
 * SBEs.py:5395-5432) through SBESolver.calc_Xqw (SBEs.py:5331-5357 ->
 * _CalcXqw_jit 3165-3243), on the device: out[q][iw] complex for q in
 * [0, N], w = iw dw, iw < nw; x = 2 dkr / (2 pi area) dcv^2 / eps0,
 * hfloor = the 0.1 meV dephasing floor (J).  Ee, Eh, fe, fh, game, gamh:
 * [N] device arrays; N <= 1023. */
int psa_xqw_table(int32_t N, const double *Ee, const double *Eh, const double *fe, const double *fh,
                  const double *game, const double *gamh, double gap, double x, double hbar,
                  double hfloor, double dw, int32_t nw, double *out, void *stream);

/* ---- 3-D coupling of the wires to a Maxwell grid (rhoPJ.QuantumWire) ----
 *
 * The per-step part of QuantumWire (rhoPJ.py:281-449) around the wires'
 * QWCalculator calls, for the Nw wires of one group lying along x of a
 * [nx][ny][nz] grid (nx = the group's Nx, the grid's x line being the
 * wires' propagation grid):
 *  - psa_qw3d_drive: EAtXYZ (usefulsubs.py:1996-2065) of Ex, Ey, Ez at the
 *    wire point (trilinear cell i0, j0, k0 and weights from the caller's
 *    locator), broadcast along x into every wire's efield;
 *  - psa_qw3d_deposit (after psa_step): J = sum_w (P - P_old) / dt / Nw plus
 *    CalcJfx (rhoPJ.py:491-523) of (sum Rho - sum Rho_old) / Nw for x;
 *    placement (QWPlacement, rhoPJ.py:655-700, which zeroes the grid first)
 *    of J into Jx, Jy, Jz, of sum Rho / Nw into Rho and of sum P / Nw into
 *    fresh Px, Py, Pz, with profiles gy, gz: placed grids are OVERWRITTEN,
 *    grids whose flag is off are left as the caller passed them; then
 *    ElongfromRho (rhoPJ.py:703-880): 3-D FFTs of Rho and P, the k-space
 *    loop (rhoPJ.py:526-619), inverse FFTs of the nine longitudinal fields
 *    and the bound charge.
 *    P_old and Rho_old are zero, as in the reference, which discards its
 *    previous values at the top of every call (rhoPJ.py:256-270): J is then
 *    sum P / (dt Nw).  With PSA_QW3D_KEEP_OLD they are the previous call's
 *    values instead (kept in `old`), so J is the polarization current.
 * Grid arrays are complex [nx][ny][nz], z fastest, device pointers. */
#define PSA_QW3D_CURR 1     /* place J (DoQWCurr) */
#define PSA_QW3D_RHO 2      /* place Rho (DoQWDl) */
#define PSA_QW3D_POL 4      /* place P and compute the longitudinal fields (DoQWP) */
#define PSA_QW3D_KEEP_OLD 8 /* J from P - P_old, CalcJfx from Rho - Rho_old (not the reference) */

typedef struct psa_qw3d {
  int32_t nx, ny, nz;           /* each 2^a 3^b in [2, 512] */
  int32_t flags;                /* PSA_QW3D_* */
  int32_t i0, j0, k0;           /* EAtXYZ cell: lower corner */
  int32_t pad;
  double ax, ay, az;            /* EAtXYZ weights of the upper corner (alpha, beta, gamma) */
  double dt, dx, eps;           /* step, x spacing, background permittivity eps0 epsr (F/m) */
  const double *qx, *qy, *qz;   /* [nx], [ny], [nz] wavenumbers (FFT order) */
  const double *gy, *gz;        /* [ny], [nz] real placement profiles */
  const double *tw;             /* complex exp(-2 pi i m / n): [nx], then [ny], then [nz] */
  const double *Ex, *Ey, *Ez;   /* grid fields (read) */
  double *Jx, *Jy, *Jz, *Rho;   /* grid sources, overwritten when placed */
  double *Fk;                   /* [4] grids scratch: Rho, Px, Py, Pz, then their FFTs */
  double *El;                   /* [10] grids: E_l from P+Rho (x, y, z), from P (x, y, z),
                                   from Rho (x, y, z), bound charge */
  double *lines;                /* [8][nx] complex scratch */
  double *old;                  /* [4][Nw][nx] complex: the wires' previous Pxx, Pyy, Pzz, Rho
                                   (read and written only with PSA_QW3D_KEEP_OLD) */
} psa_qw3d;

int psa_qw3d_size(void);
int psa_qw3d_drive(const psa_qw3d *q, const psa_group *g, void *stream);
int psa_qw3d_deposit(const psa_qw3d *q, const psa_group *g, void *stream);

/* ---- 1-D pseudo-spectral field propagator coupled to the wires ----
 *
 * The reference drives the wires with spatially uniform analytic pulses
 * (sbetestprop.py:183-198) and has no propagator on this path; it ships the
 * pieces of one: the staggered differences dEdx / dHdx (usefulsubs.py:612-655,
 * 704-747) and the real-space absorber ApplyABC (usefulsubs.py:584-609).
 * This propagator carries a pulse along y, normal to the wires (which lie
 * along x), through an array of wires, so the transmitted field can be
 * recorded:
 *  - fields (Ex, Hz) and (Ez, Hx) on N points y_j = j dy (E) and
 *    y_j + dy/2 (H), kept in k-space (numpy.fft convention);
 *  - dEx/dt = (dHz/dy - Jx)/eps, dHz/dt = (dEx/dy)/mu,
 *    dEz/dt = (-dHx/dy - Jz)/eps, dHx/dt = -(dEz/dy)/mu;
 *  - d/dy as per-k multipliers (spectral i k exp(+-i k dy/2), or the
 *    reference's differences); each mode advanced over dt by the exact
 *    solution of its 2x2 system with J constant over the step;
 *  - after the step, ApplyABC's real-space multiplication (optional);
 *  - wire slot s sits on grid point jw[s]: the field there drives the wire
 *    (broadcast along its propagation grid rr), and its polarization,
 *    averaged over rr and multiplied by `sheet`, is the grid polarization
 *    at jw[s]; J = (P_new - P_old) / dt.
 * Pointers are device pointers; complex values interleaved (re, im). */
#define PSA_PSTD_MAXW 256
#define PSA_PSTD_ABC 1 /* flags: apply the absorbing profile after each step */

typedef struct psa_pstd {
  int32_t N;          /* grid points, a power of two in [64, 8192] */
  int32_t nw;         /* wire slots, [0, PSA_PSTD_MAXW] */
  int32_t flags;      /* PSA_PSTD_* */
  int32_t nobs;       /* observation points recorded per step */
  double dy;          /* grid spacing (m) */
  double eps, mu;     /* background permittivity eps0 n0^2 and mu0 (SI) */
  double dt;          /* step the mode table was built for (s) */
  double sheet;       /* wire polarization -> grid polarization factor */
  double *state;      /* [4][N] complex k-space lines: Ex, Hz, Ez, Hx */
  const double *mode; /* [N][8] per k: ike (re, im), ikh (re, im), cos(w dt),
                         sin(w dt)/w, (1 - cos(w dt))/w^2, 0; w^2 = -ike ikh/(eps mu) */
  const double *abc;  /* [N] real absorbing profile (PSA_PSTD_ABC) */
  const double *twid; /* [N] complex exp(-2 pi i m / N) */
  const int32_t *jw;  /* [nw] grid point of each wire slot */
  const int32_t *obs; /* [nobs] recorded grid points */
  double *pold;       /* [2][nw] complex sheet polarization (x, z) at the step start */
  double *jsrc;       /* [2][nw] complex current density of the step (x, z) */
  double *ew;         /* [2][nw] complex Ex, Ez at each wire slot */
  double *rec;        /* [nrec][2][nobs] complex Ex, Ez at the observation points */
  int64_t nrec;
  int64_t *count;     /* [1] steps taken by psa_pstd_advance (device) */
} psa_pstd;

/* sizeof(psa_pstd), so bindings can check their struct mirror */
int psa_pstd_size(void);
/* In place: the four lines of `state`, holding real-space values, become
 * their forward FFTs. */
int psa_pstd_to_k(const psa_pstd *p, void *stream);
/* out[4][N] complex = inverse FFT of each line of `state` (state unchanged). */
int psa_pstd_to_real(const psa_pstd *p, double *out, void *stream);
/* ew = Ex, Ez at the wire slots from the current state. */
int psa_pstd_sample(const psa_pstd *p, void *stream);
/* One step of dt: sources from jsrc, mode update, absorber, ew, and the
 * observation record rec[count] (if count < nrec); count++. */
int psa_pstd_advance(const psa_pstd *p, void *stream);
/* Wires [w0, w0+nw) of g -> slots [slot0, slot0+nw): sheet polarization
 * from Pxx, Pzz of g->out averaged over rr; jsrc = (P - pold)/dt; pold = P. */
int psa_pstd_gather(const psa_pstd *p, const psa_group *g, int32_t w0, int32_t nw,
                    int32_t slot0, void *stream);
/* Slots [slot0, slot0+nw) -> g->efield of wires [w0, w0+nw): Exx = Ex and
 * Ezz = Ez of the slot on every rr point (Eyy untouched). */
int psa_pstd_scatter(const psa_pstd *p, const psa_group *g, int32_t w0, int32_t nw,
                     int32_t slot0, void *stream);

/* ---- The RHS of several wire groups in one persistent launch ----
 *
 * psa_rhs (dpdt/dCdt/dDdt, SBEs.py:1941-2625) for up to PSA_RHS_MULTI_MAX
 * groups of any N and one precision mode, as one launch whose workgroups
 * take 64 x 64 output tiles from a work queue (an atomic counter pair in
 * device memory; the launch leaves it at zero).  Every tile is computed as
 * psa_rhs computes it, so the results are identical.  The wires are dealt
 * to one list per XCD (a wire's tiles share that XCD's L2); a workgroup
 * drains its own list, then takes from the others.  The reference has no
 * batched form: it evaluates the RHS wire by wire (SBEs.py:5612-5675). */
#define PSA_RHS_MULTI_MAX 16
#define PSA_RHS_MULTI_QUEUE 16 /* uint32 words of the queue buffer */
/* bytes of the packed description of n groups of nw[i] wires each (-1 if
 * n or an nw[i] is out of range) */
int64_t psa_rhs_multi_bytes(const int32_t *nw, int32_t n);
/* Validate and pack groups[i] (wires [w0[i], w0[i]+nw[i]), psa_rhs mode and
 * dt) into host memory `packed` (psa_rhs_multi_bytes(n) bytes).  The caller
 * keeps a device copy of the same bytes for psa_rhs_multi; repack when a
 * group's buffers, wire range or mode change. */
int psa_rhs_multi_pack(const psa_group *const *groups, const int32_t *w0, const int32_t *nw,
                       int32_t n, int32_t mode, double dt, void *packed);
/* One launch (plus the per-group bf16 operand split in PSA_PREC_BF16X3):
 * packed_host and packed_dev hold the same bytes; queue is a device
 * uint32[PSA_RHS_MULTI_QUEUE] that is zero before the first call (each
 * launch leaves it zero).  Calls sharing a queue must not overlap. */
int psa_rhs_multi(const void *packed_host, const void *packed_dev, uint32_t *queue, void *stream);
/* psa_step (one leapfrog step, SBEs.py:5612-5820) of every packed group,
 * with the RHS of all of them as the one psa_rhs_multi launch; the
 * description must be packed with mode 1 and this dt (and repacked after
 * the caller exchanges the levels: one description per level parity).
 * Results equal psa_step's per group.  With more than one group, each
 * group's stages before and after the RHS run on a library-owned stream
 * per group (default priority, created on the first call per device),
 * forked from `stream` and joined back into it by events; under graph
 * capture they do so once an eager call has created the streams, and run
 * on `stream` in group order otherwise. */
int psa_step_multi(const void *packed_host, const void *packed_dev, uint32_t *queue, double dt,
                   void *stream);

/* The deferred leapfrog average of psa_group.lazy written out: for every
 * wire w in [w0, w0 + nw) with lazy[w] set, A <- (A + B) / 2 for p, C, D
 * (SBEs.py:5704-5706), then lazy[w] = 0.  No counterpart in the reference,
 * which averages in place every step; a no-op when lazy is NULL. */
int psa_average_levels(const psa_group *g, int32_t w0, int32_t nw, void *stream);

/* ---- Module API on the caller's dense arrays (coulomb.py, phonons.py) ----
 *
 * The per-call functions of the reference's coulomb and phonons modules for
 * any k-grid: arrays in the reference's own layouts (row-major): VC
 * [N][N][3], E1D [N][N], k3 [N][N][N] (1-based entries, 0 = invalid),
 * Ceh/Cee/Chh [(N+1)^3] filled at 1-based positions (the zero planes are
 * left as allocated), complex occupations interleaved (re, im).  All device
 * pointers; N <= 1024.  Win/Wout receive += like the reference. */

/* MakeK3 (coulomb.py:271-299) into k3 and CalcMBArrays (coulomb.py:367-499,
 * GaussDelta 245-268; lorentz != 0: the LorentzDelta branch) into Ceh, Cee,
 * Chh.  Replaces CoulombModule.__init__'s table loops (coulomb.py:1124-1164). */
int psa_mod_mb_arrays(int32_t N, const double *Ee, const double *Eh, double ge, double gh,
                      double hbar, int32_t lorentz, int32_t *k3, double *Ceh, double *Cee,
                      double *Chh, void *stream);

/* CoulombModule.CalcScreenedArrays (coulomb.py:1195-1230) with ne, nh given:
 * n1D = min(sum(Re ne + Re nh) / 2 / L, nmax); E1D = Eps1D(n1D) (coulomb.py:
 * 1172-1193) when screened, else 1; VC[.][.][c] = V0[c] / E1D (or V0).
 * ne = nh = NULL: Eps1D(n1d) alone into E1D (VC may be NULL).
 * V0 [3][N][N] = Veh0, Vee0, Vhh0; chi [2][N][N] = Chi1De, Chi1Dh;
 * q [2][N][N] = qe, qh. */
int psa_mod_screen(int32_t N, int32_t screened, double L, double nmax, const double *ne,
                   const double *nh, double n1d, const double *V0, const double *chi,
                   const double *q, double *VC, double *E1D, void *stream);

/* MBCE / MBCE2 (hole = 0; coulomb.py:897-1012, 1283-1311; Cx = Cee) and
 * MBCH (hole = 1; coulomb.py:1015-1071, 1313-1326; Cx = Chh) on the
 * occupations |ne0|, |nh0| (complex [N]). */
int psa_mod_mbc(int32_t N, int32_t hole, const double *ne0, const double *nh0, const double *VC,
                const int32_t *k3, const double *Ceh, const double *Cx, double *Win, double *Wout,
                void *stream);

/* MBPE (hole = 0; phonons.py:139-202; P = EP, PT = EPT) and MBPH (hole = 1;
 * phonons.py:205-268; P = HP, PT = HPT) on the complex occupations n [N]
 * (their real part enters the real rates). */
int psa_mod_mbp(int32_t N, int32_t hole, const double *n, const double *VC, const double *E1D,
                const double *P, const double *PT, double Vscale, double *Win, double *Wout,
                void *stream);

/* ---- Module API: QWOptics methods (qwoptics.py:565-1090) ----
 * Complex arrays interleaved (re, im), row-major; all device pointers. */

/* FFTG (sign -1, scale -1/n; usefulsubs.py:2245-2268) and iFFTG (sign +1,
 * scale -1; usefulsubs.py:2271-2297) of nvec vectors of length n as direct
 * DFTs: out[v][k] = scale sum_j in[v][j] exp(sign 2 pi i j k / n); in != out. */
int psa_mod_dft(int32_t n, int32_t nvec, int32_t sign, double scale, const double *in, double *out,
                void *stream);

/* rescale_1D (spliner.py:153-188, the cubic spline being linear in the
 * data) as the real matrix M [nout][nin]: out[v][r] = w[r] sum_j M[r][j]
 * f(in[v][j]), f = Re if real_in (Prop2QW's window-then-real-part,
 * qwoptics.py:692-713), w = NULL for 1.  in, out complex; in != out. */
int psa_mod_resample(int32_t nout, int32_t nin, int32_t nvec, const double *M, const double *w,
                     int32_t real_in, const double *in, double *out, void *stream);

/* QWOptics.QWPolarization3 before its FFTG (qwoptics.py:49-76, 809-940):
 * out[c][r] (complex [3][Nr], imaginary parts 0) = scale s_c W[r]
 * sum_{ke,kh} Re(p[kh][ke] V_c[kh][ke] E[ke][r] conj(E[kh][r])) with
 * V = Xvc0, Yvc0, Zvc0 (complex [N][N]), E = Expikr (complex [N][Nr]),
 * s = (1, syw, sz), scale = 2 ehint / area / L. */
int psa_mod_pol3(int32_t N, int32_t Nr, const double *p, const double *Xvc, const double *Yvc,
                 const double *Zvc, const double *E, const double *W, double syw, double sz,
                 double scale, double *out, void *stream);

/* QWOptics.QWRho5 before its FFTG (qwoptics.py:360-374, 942-1071): out
 * (complex [2][Nr]) = re, rh summed from C, conj(D) (complex [N][N]) with
 * E = Expikr and W, minus the mean of the real parts of the first and last
 * 10 points, scaled to sum|ne| + 1e-50 (sum|nh| + 1e-50) over
 * (sum |.| dr + 1e-100).  Nr >= 20. */
int psa_mod_rho5(int32_t N, int32_t Nr, const double *C, const double *D, const double *E,
                 const double *W, double L, double dr, const double *ne, const double *nh,
                 double *out, void *stream);

/* QWOptics.QW2Prop's Plasmonics step (qwoptics.py:785-790, 801-805) on
 * complex re, rh [n] in place: absolute values, then each scaled to
 * total / (sum dx + 1e-100); mode 0 first sets *total (device double) =
 * (sum re dx + sum rh dx + 1e-100) / 2, mode 1 reuses it. */
int psa_mod_plasmonics(int32_t n, double dx, int32_t mode, double *re, double *rh, double *total,
                       void *stream);

/* ---- Module API: dephasing.py, emission.py ---- */

/* CalcGammaE (hole = 0; dephasing.py:28-66, 541-580) / CalcGammaH (hole = 1;
 * dephasing.py:68-103, 582-625) into out [N]: occupations Re ne0, Re nh0
 * (complex [N]); V^2(q) = VC[iq[q]][0][c]^2 (Vxx2, dephasing.py:362-378);
 * tabs [4][N][N] int = k_p_q, k_m_q, k1p_m_q, k1p (electrons) or k_p_q,
 * k_m_q, k1_m_q, k1 (holes) of InitializeDephasing (dephasing.py:472-539);
 * xe, xh [N]. */
int psa_mod_gamma(int32_t N, int32_t hole, const double *ne0, const double *nh0, const double *VC,
                  const int32_t *iq, const int32_t *tabs, const double *xe, const double *xh,
                  double hbar, double *out, void *stream);

/* CalcOffDiagDeph_E2 / _H2 (dephasing.py:226-317, 381-414) into D [2][2N+1][N]
 * (De, Dh; row q + N) on |ne0|, |nh0|, widths hbar gee, hbar ghh, hbar geh;
 * with x != NULL also OffDiagDephasing2's contraction (dephasing.py:320-356,
 * 714-743) of p (complex [N][N]) into x (complex [N][N]). */
int psa_mod_offdiag2(int32_t N, const double *ne0, const double *nh0, const double *p,
                     const double *VC, const double *Ee, const double *Eh, double gee, double ghh,
                     double geh, double hbar, double *D, double *x, void *stream);

/* SpontEmission (emission.py:94-130) with Ec (133-171) and SpontIntegral
 * (174-198): Rsp[k] = dhw sum_n (H_n + Ek) (H_n + Ek)^2 / rden square_n,
 * Ek = gap + Ee[k] + Eh[k] + Ec[k] on Re ne0, Re nh0; rden = c0^3 pi^2
 * hbar^3 (rho0, emission.py:201-217); ec != NULL receives Ec. */
int psa_mod_spont(int32_t N, int32_t Nhw, const double *ne0, const double *nh0, const double *VC,
                  const double *Ee, const double *Eh, double gap, const double *homega,
                  const double *square, double dhw, double rden, double *Rsp, double *ec,
                  void *stream);

/* CalcMVeh (coulomb.py:1234-1241, _CalcMVeh_core 786-810):
 * MVeh[k][kp][f] = sum_q p[q][qp][f] VC[k][q][0] UnDel[k+1][q+1] UnDel[kp+1][qp+1]
 * with qp + 1 = k3[kp][q][k] (0: no term).  p, MVeh complex [N][N][Nf];
 * k3 int32 [N][N][N]; UnDel [N+1][N+1]. */
int psa_mod_mveh(int32_t N, int32_t Nf, const double *p, const double *VC, const int32_t *k3,
                 const double *UnDel, double *MVeh, void *stream);

/* BGRenorm (which 0, on ne = diag C, nh = diag D), EeRenorm (1, ne) and
 * EhRenorm (2, nh) (coulomb.py:1248-1280; cores 813-897) into BGR complex
 * [N][N].  ne, nh complex [N] (the one a mode does not read may be NULL). */
int psa_mod_renorm(int32_t N, int32_t which, const double *ne, const double *nh, const double *VC,
                   const double *UnDel, double *BGR, void *stream);

/* GetVn1n2 (qwoptics.py:1149-1186, _GetVn1n2_jit 522-542): Vcc, Vvv, Vcv and
 * Vvc = Vcv^H from rcv complex [N] and Hcc, Hhh, Hcv complex [N][N]. */
int psa_mod_vn1n2(int32_t N, const double *rcv, const double *Hcc, const double *Hhh,
                  const double *Hcv, double hbar, double *Vcc, double *Vvv, double *Vcv,
                  double *Vvc, void *stream);

/* PLSpectrum (emission.py:291-361): PLS[w] = (PLS[w] + rscale sum_j ...) tfac
 * over Nhw photon energies hw, with hgeh = hbar geh, rden = c0^3 pi^2 hbar^3
 * (rho0), tfac = softtheta(t, geh); ne, nh complex [N]; work: device scratch
 * of N + 1 doubles (transition energies and the carrier temperature). */
int psa_mod_plspectrum(int32_t N, int32_t Nhw, const double *ne, const double *nh,
                       const double *Ee, const double *Eh, double gap, double hgeh,
                       const double *VC, const double *hw, double rden, double kB,
                       double rscale, double tfac, double *work, double *PLS, void *stream);

/* CalcOffDiagDeph_E (hole 0) / CalcOffDiagDeph_H (hole 1) (dephasing.py:
 * 107-190, 627-685), the version-1 tables D [N][N] (k, q) from the complex
 * occupations ne0, nh0 [N], VC, the band energies and the index tables kpq,
 * kmq (int32 [N][N], 0-based: the reference's _k_p_q, _k_m_q; an index
 * outside [0, N) drops the term), widths hbar gss (same carrier) and
 * hbar gso (electron-hole). */
int psa_mod_offdiag1(int32_t N, int32_t hole, const double *ne0, const double *nh0, const double *VC,
                     const double *Ee, const double *Eh, const int32_t *kpq, const int32_t *kmq,
                     double gss, double gso, double hbar, double *D, void *stream);

/* OffDiagDephasing (dephasing.py:193-218, 687-712), the version-1
 * contraction of p (complex [N][N]) with the tables DE, DH of
 * psa_mod_offdiag1 into x (complex [N][N]); reads p with the reference's
 * one-index shift of the second index. */
int psa_mod_offdiag1_apply(int32_t N, const double *DE, const double *DH, const double *p,
                           const int32_t *kpq, const double *ky, double hbar, double *x, void *stream);

/* Vehint (coulomb.py:180-242) with dky = ky[k-1] - ky[q-1]: the Vint double
 * sum with the accurate K0 (usefulsubs.K03) into out[0]. */
int psa_mod_vehint(double dky, const double *y, int32_t Ny, double alphae, double alphah,
                   double Delta0, double *out, void *stream);

#ifdef __cplusplus
}
#endif
#endif
