// MFMA GEMM engine for the SBE contractions (gfx950, wave64).
//
// One workgroup (256 threads = 4 waves in a 2x2 arrangement) computes a
// 64x64 output tile of a batched "sum of products" C = sum_t op(A_t) op(B_t).
// Operands are produced tile by tile by the problem's loaders (dense
// matrices with transpose/conjugate, or generated on the fly: Toeplitz
// Coulomb matrices, skewed diagonals, Lorentzian kernels, occupation
// vectors) into LDS as separate real/imaginary planes laid out [k][x].
// Each wave owns a 32x32 sub-tile = 2x2 MFMA 16x16 blocks; a complex
// product costs four real 16x16x4 MFMAs.  The epilogue hands every output
// element with its real/imaginary accumulator to the problem's store().
// A problem supplies nterms(b), krange(t, b, m0, n0, kb, ke) (the K slice
// that can be non-zero for this tile; loaders return zero outside it) and
// active(b, m0, n0) (false when every product in the tile is zero).
//
// MFMA 16x16x4 lane layout: lane l supplies A[l%16][l/16] and B[l/16][l%16]
// for both f32 and f64.  Accumulator element v of lane l is D[4*(l/16)+v][l%16]
// for f32 but D[l/16+4*v][l%16] for f64 (checked on gfx950).
#pragma once

#include "psa_dev.h"

namespace psa {

typedef double dv4 __attribute__((ext_vector_type(4)));
typedef float fv4 __attribute__((ext_vector_type(4)));

template <typename T>
struct Mfma;
template <>
struct Mfma<double> {
  typedef dv4 acc;
  static __device__ __forceinline__ int row(int kl, int v) { return kl + 4 * v; }
  static __device__ __forceinline__ acc mma(double a, double b, acc c) {
    return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
  }
};
template <>
struct Mfma<float> {
  typedef fv4 acc;
  static __device__ __forceinline__ int row(int kl, int v) { return 4 * kl + v; }
  static __device__ __forceinline__ acc mma(float a, float b, acc c) {
    return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
  }
};

constexpr int GBM = 64, GBN = 64, GBK = 16, GLD = GBM + 2;

template <typename T>
struct Tile {
  T v[2][GBK][GLD];  // [re/im][k][x]
};

// Register fragment of one operand slab: each of the 256 threads holds 4
// elements (value and LDS coordinates) between the global load and the LDS
// store, so the next slab's loads are in flight during the MFMAs.
template <typename T>
struct Frag {
  T re[4], im[4];
  uint8_t k[4], x[4];
  __device__ __forceinline__ void set(int e, int kk, int xx, T r, T i) {
    re[e] = r;
    im[e] = i;
    k[e] = (uint8_t)kk;
    x[e] = (uint8_t)xx;
  }
  __device__ __forceinline__ void commit(Tile<T> &t) const {
#pragma unroll
    for (int e = 0; e < 4; ++e) {
      t.v[0][k[e]][x[e]] = re[e];
      t.v[1][k[e]][x[e]] = im[e];
    }
  }
};

// s[k][x] = sgn * op(X)(x0+x, k0+k); op(X)(x, k) = X[x][k] (xmajor) or
// X[k][x]; optionally conjugated; zero outside [0,xext) x [0,kext).
// Thread mapping follows the memory layout so global reads coalesce.
template <typename T>
__device__ __forceinline__ void load_dense(Frag<T> &s, const double *X, int ldx,
                                           int xext, int kext, bool xmajor,
                                           bool conj, double sgn, int x0, int k0,
                                           bool zero_diag = false) {
  const int tid = threadIdx.x;
#pragma unroll
  for (int e = 0; e < 4; ++e) {
    const int idx = tid + 256 * e;
    int x, k;
    if (xmajor) {
      x = idx >> 4;
      k = idx & 15;
    } else {
      k = idx >> 6;
      x = idx & 63;
    }
    const int gx = x0 + x, gk = k0 + k;
    // unconditional load from a clamped in-range address, masked after, so
    // the loads of a slab are in flight together
    const int cx = min(gx, xext - 1), ck = min(gk, kext - 1);
    cd v = ld(X, xmajor ? (long)cx * ldx + ck : (long)ck * ldx + cx);
    if (!(gx < xext && gk < kext && !(zero_diag && gx == gk))) v = cd{0.0, 0.0};
    s.set(e, k, x, (T)(sgn * v.re), (T)(conj ? -sgn * v.im : sgn * v.im));
  }
}

// First term t >= t0 whose K range [kb, ke) is non-empty (nt if none).
template <class P>
__device__ __forceinline__ int next_term(const P &p, int t0, int nt, int b, int m0, int n0,
                                         int &kb, int &ke) {
  for (int t = t0; t < nt; ++t) {
    p.krange(t, b, m0, n0, kb, ke);
    if (kb < ke) return t;
  }
  return nt;
}

// One 64 x 64 output tile (batch b, rows m0.., columns n0..) of problem p;
// lds_raw holds the double-buffered operand tiles (gemm_lds_bytes<T>()).
template <typename T, class P>
__device__ __forceinline__ void gemm_tile(const P &p, int b, int m0, int n0, char *lds_raw) {
  Tile<T> *tiles = reinterpret_cast<Tile<T> *>(lds_raw);  // [2 buffers][A, B]
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int wm = wave >> 1, wn = wave & 1;
  using acc_t = typename Mfma<T>::acc;
  acc_t aR[2][2], aI[2][2];
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int j = 0; j < 2; ++j) {
      aR[i][j] = acc_t{0, 0, 0, 0};
      aI[i][j] = acc_t{0, 0, 0, 0};
    }
  const int kl = lane >> 4, xl = lane & 15;
  const int nt = p.nterms(b);
  // Terms run over their own K ranges [kb, ke) (block-uniform); a tile the
  // problem marks inactive, or one whose ranges are all empty, skips the
  // products and stores zero accumulators.
  int t = nt, k0 = 0, ke = 0;
  if (p.active(b, m0, n0)) t = next_term(p, 0, nt, b, m0, n0, k0, ke);
  if (t < nt) {
    // flattened (term, k-slab) iteration with a one-slab register prefetch
    Frag<T> fa, fb;
    p.fetchA(fa, t, b, m0, k0);
    p.fetchB(fb, t, b, k0, n0);
    fa.commit(tiles[0]);
    fb.commit(tiles[1]);
    __syncthreads();
    int cur = 0;
    while (true) {
      int tn = t, kn = k0 + GBK, ken = ke;
      if (kn >= ke) tn = next_term(p, t + 1, nt, b, m0, n0, kn, ken);
      const bool more = tn < nt;
      if (more) {
        p.fetchA(fa, tn, b, m0, kn);
        p.fetchB(fb, tn, b, kn, n0);
      }
      const Tile<T> &sA = tiles[2 * cur], &sB = tiles[2 * cur + 1];
#pragma unroll
      for (int kq = 0; kq < GBK; kq += 4) {
        const int kk = kq + kl;
        T ar[2], ai[2], br[2], bi[2];
#pragma unroll
        for (int s = 0; s < 2; ++s) {
          ar[s] = sA.v[0][kk][wm * 32 + s * 16 + xl];
          br[s] = sB.v[0][kk][wn * 32 + s * 16 + xl];
          if (!P::A_REAL) ai[s] = sA.v[1][kk][wm * 32 + s * 16 + xl];
          if (!P::B_REAL) bi[s] = sB.v[1][kk][wn * 32 + s * 16 + xl];
        }
        // four phases over the 2x2 sub-blocks, so that two MFMAs on the
        // same accumulator are eight issues apart (no dependency stall)
#pragma unroll
        for (int si = 0; si < 2; ++si)
#pragma unroll
          for (int sj = 0; sj < 2; ++sj) aR[si][sj] = Mfma<T>::mma(ar[si], br[sj], aR[si][sj]);
        if (!P::B_REAL) {
#pragma unroll
          for (int si = 0; si < 2; ++si)
#pragma unroll
            for (int sj = 0; sj < 2; ++sj) aI[si][sj] = Mfma<T>::mma(ar[si], bi[sj], aI[si][sj]);
        }
        if (!P::A_REAL && !P::B_REAL) {
#pragma unroll
          for (int si = 0; si < 2; ++si)
#pragma unroll
            for (int sj = 0; sj < 2; ++sj) aR[si][sj] = Mfma<T>::mma(-ai[si], bi[sj], aR[si][sj]);
        }
        if (!P::A_REAL) {
#pragma unroll
          for (int si = 0; si < 2; ++si)
#pragma unroll
            for (int sj = 0; sj < 2; ++sj) aI[si][sj] = Mfma<T>::mma(ai[si], br[sj], aI[si][sj]);
        }
      }
      if (!more) break;
      fa.commit(tiles[2 * (cur ^ 1)]);
      fb.commit(tiles[2 * (cur ^ 1) + 1]);
      __syncthreads();
      cur ^= 1;
      t = tn;
      k0 = kn;
      ke = ken;
    }
  }
#pragma unroll
  for (int si = 0; si < 2; ++si)
#pragma unroll
    for (int sj = 0; sj < 2; ++sj)
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        const int row = m0 + wm * 32 + si * 16 + Mfma<T>::row(kl, v);
        const int col = n0 + wn * 32 + sj * 16 + xl;
        p.store(b, row, col, (double)aR[si][sj][v], (double)aI[si][sj][v]);
      }
}

template <typename T, class P>
__global__ __launch_bounds__(256) void gemm_kernel(P p) {
  extern __shared__ char lds_raw[];
  // XCD-aware tile order: consecutive workgroups go to different XCDs
  // (round robin over 8).  When the batch count is a multiple of 8, batch b
  // runs entirely on XCD b % 8 (its tiles share that XCD's L2 for their
  // operands) and neighbouring batches still spread over all XCDs
  // (tune[2] = 1 selects the plain linear order).
  const unsigned gx = gridDim.x, gy = gridDim.y, tpb = gx * gy;
  unsigned lin = blockIdx.x + gx * (blockIdx.y + gy * blockIdx.z);
  int b;
  if (gridDim.z % 8 == 0 && p.g.tune[2] == 0) {
    const unsigned xcd = lin % 8, q = lin / 8;
    b = xcd + 8 * (q / tpb);
    lin = q % tpb;
  } else {
    b = lin / tpb;
    lin %= tpb;
  }
  gemm_tile<T, P>(p, b, (lin / gx) * GBM, (lin % gx) * GBN, lds_raw);
}

// Split-K view of a single-batch problem: the batch index is the K split
// (K values [b ks, (b+1) ks)); store() writes the split's raw accumulators
// to scr[b][M][ncol][2], summed in a fixed order by splitk_reduce, which
// applies the problem's own epilogue.  Every (split, row, column) is
// written, zero where the split has no products.
template <class P>
struct SplitK {
  static constexpr bool A_REAL = P::A_REAL, B_REAL = P::B_REAL;
  P p;
  psa_group g;
  int ks, M, ncol;
  double *scr;

  __device__ int nterms(int) const { return p.nterms(0); }
  __device__ bool active(int, int m0, int n0) const { return p.active(0, m0, n0); }
  __device__ void krange(int t, int b, int m0, int n0, int &kb, int &ke) const {
    p.krange(t, 0, m0, n0, kb, ke);
    kb = max(kb, b * ks);
    ke = min(ke, (b + 1) * ks);
  }
  template <typename T>
  __device__ void fetchA(Frag<T> &s, int t, int, int m0, int k0) const {
    p.fetchA(s, t, 0, m0, k0);
  }
  template <typename T>
  __device__ void fetchB(Frag<T> &s, int t, int, int k0, int n0) const {
    p.fetchB(s, t, 0, k0, n0);
  }
  __device__ void store(int b, int r, int c, double re, double im) const {
    if (r >= M || c >= ncol) return;
    double *o = scr + (((long)b * M + r) * ncol + c) * 2;
    o[0] = re;
    o[1] = im;
  }
};

// S <= SPLITK_MAX: every split's partial is loaded before the first add
// (one round of memory latency), then they are added in split order.
constexpr int SPLITK_MAX = 8;
template <class P>
__global__ __launch_bounds__(256) void splitk_reduce(P p, const double *scr, int S, int M, int ncol) {
  const long i = blockIdx.x * (long)blockDim.x + threadIdx.x;
  if (i >= (long)M * ncol) return;
  const int r = (int)(i / ncol), c = (int)(i % ncol);
  double2 v[SPLITK_MAX];
#pragma unroll
  for (int s = 0; s < SPLITK_MAX; ++s)
    if (s < S) v[s] = *reinterpret_cast<const double2 *>(scr + (((long)s * M + r) * ncol + c) * 2);
  double re = 0.0, im = 0.0;
#pragma unroll
  for (int s = 0; s < SPLITK_MAX; ++s)
    if (s < S) {
      re += v[s].x;
      im += v[s].y;
    }
  p.store(0, r, c, re, im);
}

template <typename T>
constexpr size_t gemm_lds_bytes() {
  return 4 * sizeof(Tile<T>);
}

// ===========================================================================
// Problem: the SBE right-hand sides (SBEs.py:1941-1965, 2308-2327,
// 2606-2625).  The diagonals of Hee, Hhh (band energies) and of C, D
// (occupations) are taken out of the products and applied exactly in the
// FP64 epilogue, so the products only carry the Coulomb/coherence
// corrections (this is what lets them run in reduced precision).  For dp
// the off-diagonal dephasing (dephasing.py:320-356) enters as two banded
// products.
// batch b -> equation e = b % 3 (p, C, D), wire w = w0 + b / 3.
// ===========================================================================
struct RhsProb {
  static constexpr bool A_REAL = false, B_REAL = false;
  psa_group g;
  int w0, mode;
  double dt;

  __device__ int nterms(int b) const {
    return (b % 3 == 0 && (g.flags & PSA_F_OFFDIAGDPH)) ? 6 : 4;
  }
  __device__ bool active(int, int, int) const { return true; }
  __device__ void krange(int, int, int, int, int &kb, int &ke) const {
    kb = 0;
    ke = g.N;
  }

  // Operand descriptor: which matrix, how it is read (op(X)(x, k) = X[x][k]
  // when xm, else X[k][x]), conjugation, sign, zeroed diagonal, or a band.
  // The Hamiltonians are stored transposed (buffer [k2][k1] = H[k1][k2]),
  // so their descriptors read the transpose of what the comment names.
  struct Op {
    const double *X;
    int band;  // -1 dense, else band table kind (0 De, 1 Dh)
    bool xm, cj, zd;
    double sg;
  };

  __device__ Op opA(int e, int t, long o) const {
    const double *P = g.pB + o, *C = g.CB + o, *D = g.DB + o;
    const double *He = g.Heh + o, *Ho = g.Heo + o, *Hh = g.Hho + o;
    switch (e * 8 + t) {
      case 0: return {Hh, -1, false, false, false, 1.0};    // dp: Hhh_o
      case 1: return {P, -1, true, false, false, 1.0};      //     p
      case 2: return {He, -1, true, false, false, -1.0};    //     -Heh^T
      case 3: return {D, -1, false, false, true, -1.0};     //     -D_o^T
      case 4: return {nullptr, 1, true, false, false, 1.0}; //     i hbar Dh band
      case 5: return {P, -1, true, false, false, 1.0};      //     p
      case 8: return {C, -1, true, false, true, 1.0};       // dC: C_o
      case 9: return {Ho, -1, true, false, false, -1.0};    //     -Hee_o^T
      case 10: return {P, -1, false, true, false, 1.0};     //     p^H
      case 11: return {He, -1, false, true, false, -1.0};   //     -conj(Heh)
      case 16: return {D, -1, true, false, true, 1.0};      // dD: D_o
      case 17: return {Hh, -1, true, false, false, -1.0};   //     -Hhh_o^T
      case 18: return {P, -1, true, true, false, 1.0};      //     conj(p)
      default: return {He, -1, true, true, false, -1.0};    //     -Heh^H
    }
  }

  __device__ Op opB(int e, int t, long o) const {
    const double *P = g.pB + o, *C = g.CB + o, *D = g.DB + o;
    const double *He = g.Heh + o, *Ho = g.Heo + o, *Hh = g.Hho + o;
    // op(B)[k][x] = X[x][k] when xm
    switch (e * 8 + t) {
      case 0: return {P, -1, false, false, false, 1.0};     // dp: p
      case 1: return {Ho, -1, false, false, false, 1.0};    //     Hee_o^T
      case 2: return {C, -1, false, false, true, 1.0};      //     C_o
      case 3: return {He, -1, false, false, false, 1.0};    //     Heh^T
      case 4: return {P, -1, false, false, false, 1.0};     //     p
      case 5: return {nullptr, 0, true, false, false, 1.0}; //     i hbar De band
      case 8: return {Ho, -1, false, false, false, 1.0};    // dC: Hee_o^T
      case 9: return {C, -1, false, false, true, 1.0};      //     C_o
      case 10: return {He, -1, false, false, false, 1.0};   //     Heh^T
      case 11: return {P, -1, false, false, false, 1.0};    //     p
      case 16: return {Hh, -1, false, false, false, 1.0};   // dD: Hhh_o^T
      case 17: return {D, -1, false, false, true, 1.0};     //     D_o
      case 18: return {He, -1, true, false, false, 1.0};    //     Heh
      default: return {P, -1, true, false, false, 1.0};     //     p^T
    }
  }

  template <typename T>
  __device__ void fetchA(Frag<T> &s, int t, int b, int m0, int k0) const {
    const int e = b % 3, w = w0 + b / 3, N = g.N;
    const Op op = opA(e, t, mat_off(g, w));
    if (op.band < 0)
      load_dense(s, op.X, N, N, N, op.xm, op.cj, op.sg, m0, k0, op.zd);
    else
      load_band(s, w, op.band, m0, k0, true);
  }

  template <typename T>
  __device__ void fetchB(Frag<T> &s, int t, int b, int k0, int n0) const {
    const int e = b % 3, w = w0 + b / 3, N = g.N;
    const Op op = opB(e, t, mat_off(g, w));
    if (op.band < 0)
      load_dense(s, op.X, N, N, N, op.xm, op.cj, op.sg, n0, k0, op.zd);
    else
      load_band(s, w, op.band, n0, k0, false);
  }

  // Banded off-diagonal dephasing operand, value i*hbar*table.
  //  kind 1 (Dh, A side): A[kh][m] = Dh[m-kh][kh],  x = kh, k = m
  //  kind 0 (De, B side): B[m][ke] = De[m-ke][ke],  x = ke, k = m
  // tables stored [w][kind][k][q + N - 1]; q = 0 excluded (undel(0) = 0).
  template <typename T>
  __device__ void load_band(Frag<T> &s, int w, int kind, int x0, int k0, bool) const {
    const int N = g.N, Q = 2 * N - 1;
    const double *tab = g.dedh + ((long)w * 2 + kind) * (long)N * Q;
    const int tid = threadIdx.x;
#pragma unroll
    for (int e = 0; e < 4; ++e) {
      const int idx = tid + 256 * e;
      const int x = idx >> 4, k = idx & 15;
      const int gx = x0 + x, gk = k0 + k;
      const double raw = tab[(long)min(gx, N - 1) * Q + min(max(gk - gx + N - 1, 0), Q - 1)];
      const double v = (gx < N && gk < N && gk != gx) ? raw : 0.0;
      s.set(e, k, x, (T)0, (T)(g.hbar * v));
    }
  }

  // The per-element FP64 operands of the epilogue (loaded apart from the
  // arithmetic, so that a caller can have many elements' loads in flight):
  // x = the checkout level X (p, C or D) at the element, h = Heh[ke][kh]
  // (dp) or H_o[k2][k1] (dC, dD; 0 on the diagonal), x1 = the level the
  // leapfrog / RK4 stage adds to.
  struct Pre {
    cd x, h, x1;
  };

  __device__ Pre preload(int b, int row, int col) const {
    const int N = g.N, e = b % 3, w = w0 + b / 3;
    const long o = mat_off(g, w), ij = (long)row * N + col;
    Pre q;
    if (e == 0) {
      q.x = ld(g.pB + o, ij);
      q.h = ld_h(g, g.Heh, o, ij);  // Heh[ke][kh], stored transposed
    } else {
      q.x = ld((e == 1 ? g.CB : g.DB) + o, ij);
      q.h = (row == col) ? cd{0.0, 0.0} : ld_h(g, e == 1 ? g.Heo : g.Hho, o, ij);  // H_o[k2][k1]
    }
    q.x1 = mode == 0 ? cd{0.0, 0.0} : ld((e == 0 ? g.pA : (e == 1 ? g.CA : g.DA)) + o, ij);
    // deferred leapfrog average (psa_group.lazy): level A holds X1 of the
    // step that made level B; the step adds to their average
    if (mode == 1 && g.lazy && g.lazy[w]) q.x1 = cd{(q.x1.re + q.x.re) / 2.0, (q.x1.im + q.x.im) / 2.0};
    return q;
  }

  // The diagonal quantities the epilogue needs at one index k, as a column
  // (ke for dp, k2 for dC, dD) or a row (kh, k1): h = the band energy of
  // that index (hdiag), n = the occupation, g = the diagonal dephasing
  // rate (dp only).
  struct Side {
    cd h, n;
    double g;
  };

  __device__ Side side(int e, int w, int k, bool col) const {
    const int N = g.N;
    const double *hd = g.hdiag + (long)w * 4 * N;  // [2][N] complex
    const double *pr = g.prep + (long)w * PSA_PREP_STRIDE(N);
    if (e == 0)
      return col ? Side{ld(hd, k), ld(pr + PSA_PREP_NE, k), pr[PSA_PREP_GAME(N) + k]}
                 : Side{ld(hd, N + k), ld(pr + PSA_PREP_NH(N), k), pr[PSA_PREP_GAMH(N) + k]};
    return Side{ld(hd, (e == 1 ? 0 : N) + k), ld(pr + (e == 1 ? PSA_PREP_NE : PSA_PREP_NH(N)), k), 0.0};
  }

  // dX/dt at one element from the contraction sum acc = (re, im):
  // dp: row = kh, col = ke; the diagonal parts (he[ke] + hh[kh]) p, the
  // Pauli factor Heh[ke,kh] (1 - ne[ke] - nh[kh]) from -Heh^T C - (Heh D)^T,
  // and the dephasing -(GamE[ke] + GamH[kh]) p.  dC, dD:
  // (h[k2] - h[k1]) X[k1,k2] + (n[k1] - n[k2]) H_o[k2,k1].
  __device__ cd deriv(int e, double re, double im, const Pre &q, const Side &c, const Side &r) const {
    const double ihb = 1.0 / g.hbar;  // z / (i hbar) = -i z / hbar
    const cd acc = {re, im};
    if (e == 0) {
      const cd pauli = {1.0 - c.n.re - r.n.re, -c.n.im - r.n.im};
      cd z = cadd(acc, cmul(cadd(c.h, r.h), q.x));
      z = cadd(z, cmul(q.h, pauli));
      const cd d = {z.im * ihb, -z.re * ihb};
      return csub(d, cscale(q.x, c.g + r.g));
    }
    cd z = cadd(acc, cmul(csub(c.h, r.h), q.x));
    z = cadd(z, cmul(csub(r.n, c.n), q.h));
    return cd{z.im * ihb, -z.re * ihb};
  }

  __device__ void finish(int b, int row, int col, double re, double im, const Pre &q) const {
    const int N = g.N;
    const int e = b % 3, w = w0 + b / 3;
    const long o = mat_off(g, w);
    const long ij = (long)row * N + col;
    const cd d = deriv(e, re, im, q, side(e, w, col, true), side(e, w, row, false));
    double *out = (e == 0 ? g.pX : (e == 1 ? g.CX : g.DX)) + o;
    if (mode == 0) {
      st(out, ij, d);
    } else if (mode == 1) {
      const cd x1 = q.x1;
      st(out, ij, cd{x1.re + d.re * dt * 2.0, x1.im + d.im * dt * 2.0});
      // the average formed in preload becomes level A (the X1 of this step)
      if (g.lazy && g.lazy[w]) st((e == 0 ? g.pA : (e == 1 ? g.CA : g.DA)) + o, ij, x1);
    } else {
      // RK4 stage mode - 1 (_rk4, integrator.py:143-157): x1 = step start,
      // R = k1 + 2 k2 + 2 k3 so far; out may alias the step start (same
      // element only)
      const cd x0 = q.x1;
      double *R = (e == 0 ? g.pR : (e == 1 ? g.CR : g.DR)) + o;
      if (mode == 5) {
        const cd r = ld(R, ij);
        const double h6 = dt / 6.0;
        st(out, ij, cd{x0.re + h6 * (r.re + d.re), x0.im + h6 * (r.im + d.im)});
      } else {
        const double c = mode == 4 ? dt : 0.5 * dt;
        if (mode == 2) {
          st(R, ij, d);
        } else {
          const cd r = ld(R, ij);
          st(R, ij, cd{r.re + 2.0 * d.re, r.im + 2.0 * d.im});
        }
        st(out, ij, cd{x0.re + c * d.re, x0.im + c * d.im});
      }
    }
  }

  __device__ void store(int b, int row, int col, double re, double im) const {
    if (row >= g.N || col >= g.N) return;
    finish(b, row, col, re, im, preload(b, row, col));
  }
};

// Epilogue of the exchange sums for output (a, dd) of Hamiltonian `which`
// (0 = Heh, 1 = Hee, 2 = Hhh) of wire w: adds Meh, the Hartree (direct)
// terms and the band energies and scatters to (k1, k2) = (a + d, a),
// d = dd - (N - 1), in the transposed Hamiltonian buffers
// (_CalcH_jit, SBEs.py:808-924).
__device__ inline void calch_store(const psa_group &g, int w, int which, int a, int dd, double re,
                                   double im) {
  const int N = g.N;
  if (a >= N || dd >= 2 * N - 1) return;
  const int d = dd - (N - 1);
  const int k1 = a + d, k2 = a;
  if (k1 < 0 || k1 >= N) return;
  const long o = mat_off(g, w);
  const long ij = (long)k2 * N + k1;  // stored transposed: consecutive lanes, consecutive k1
  const double *pr = g.prep + (long)w * PSA_PREP_STRIDE(N);
  const bool lf = (g.flags & PSA_F_LF) != 0;
  double *hd = g.hdiag + (long)w * 4 * N;
  if (which == 0) {
    cd meh = {0.0, 0.0};
    if (g.flags & PSA_F_OPTICS) {
      const int wid = g.wire_id[w];
      const double *eq = g.exq + (long)w * 3 * 2 * (2 * N - 1);
      const cd ex = ld(eq, dd), ey = ld(eq, (2 * N - 1) + dd), ez = ld(eq, 2 * (2 * N - 1) + dd);
      const double xcv = (k2 & 1) ? -g.dcv : g.dcv;  // Xcv0[ke,kh] = dcv (-1)^kh
      const double yv = yw_of(wid), zs = sgn_pow(wid);
      // -ehint*Xcv*Ex - ehint*Ycv*Ey*yw - ehint*Zcv*Ez*(-1)^w, Ycv = dcv, Zcv = -dcv
      const double a1 = -g.ehint * xcv, a2 = -g.ehint * g.dcv, a3 = -g.ehint * (-g.dcv);
      meh.re = a1 * ex.re + (a2 * ey.re) * yv + (a3 * ez.re) * zs;
      meh.im = a1 * ex.im + (a2 * ey.im) * yv + (a3 * ez.im) * zs;
    }
    st_h(g, g.Heh, o, ij, cd{meh.re + re, meh.im + im});
  } else {
    const bool ee = (which == 1);
    if (d == 0) {
      const double e0 = ee ? g.Ee[a] + g.gap : g.Eh[a];
      st(hd, (ee ? 0 : N) + a, cd{e0 - re, -im});
      return;
    }
    double *Hout = ee ? g.Heo : g.Hho;
    if (!lf) {
      st_h(g, Hout, o, ij, cd{0.0, 0.0});
      return;
    }
    const int ad = d > 0 ? d : -d;
    const double *vc = pr + PSA_PREP_VC(N);
    const double vself = vc[(ee ? 1 : 2) * N + ad], veh = vc[ad];
    const cd dsC = ld(pr + PSA_PREP_DSC(N), d + N - 1);
    const cd dsD = ld(pr + PSA_PREP_DSD(N), d + N - 1);
    const cd dsCm = ld(pr + PSA_PREP_DSC(N), -d + N - 1);
    const cd dsDm = ld(pr + PSA_PREP_DSD(N), -d + N - 1);
    cd h;
    if (ee)  // - exchange + Vee(qd) sum_k C[k,k+qd] - Veh(qd) sum_k D[k+qd,k]
      h = cd{-re + vself * dsC.re - veh * dsDm.re, -im + vself * dsC.im - veh * dsDm.im};
    else     // - exchange + Vhh(qd) sum_k D[k,k+qd] - Veh(qd) sum_k C[k+qd,k]
      h = cd{-re + vself * dsD.re - veh * dsCm.re, -im + vself * dsD.im - veh * dsCm.im};
    st_h(g, Hout, o, ij, h);
  }
}

// ===========================================================================
// Problem: Hartree-Fock exchange as a skewed Toeplitz product
// (_CalcH_jit, SBEs.py:808-924).  For diagonal offset d = k1 - k2 and
// a = k2:  O[a][d] = sum_{m != a} V(|m-a|) X[m][m+d]  with X = p (eh),
// C (ee), D (hh); the epilogue adds Meh, the Hartree (direct) terms and
// the band energies and scatters to (k1, k2) = (a+d, a).
// batch b -> which = b % 3 (eh, ee, hh), wire w = w0 + b / 3.
// ===========================================================================
struct CalcHProb {
  static constexpr bool A_REAL = true, B_REAL = false;
  psa_group g;
  int w0;

  __device__ int nterms(int) const { return 1; }
  // output (a, dd) is used only when k1 = a + dd - (N-1) lies in [0, N)
  __device__ bool active(int, int m0, int n0) const {
    const int N = g.N;
    return m0 + n0 <= 2 * N - 2 && m0 + n0 + (GBM - 1) + (GBN - 1) >= N - 1;
  }
  // B[m][dd] = X[m][m + dd - (N-1)] is non-zero only for m + dd in [N-1, 2N-1)
  __device__ void krange(int, int, int, int n0, int &kb, int &ke) const {
    const int N = g.N;
    kb = max(0, N - 1 - (n0 + GBN - 1));
    ke = min(N, 2 * N - 1 - n0);
  }

  template <typename T>
  __device__ void fetchA(Frag<T> &s, int, int b, int m0, int k0) const {
    const int which = b % 3, w = w0 + b / 3, N = g.N;
    const double *vc = g.prep + (long)w * PSA_PREP_STRIDE(N) + PSA_PREP_VC(N) + which * N;
    const int tid = threadIdx.x;
#pragma unroll
    for (int e = 0; e < 4; ++e) {
      const int idx = tid + 256 * e;
      const int k = idx >> 6, x = idx & 63;
      const int a = m0 + x, m = k0 + k;
      double v = 0.0;
      if (a < N && m < N && m != a) v = vc[m > a ? m - a : a - m];
      s.set(e, k, x, (T)v, (T)0);
    }
  }

  template <typename T>
  __device__ void fetchB(Frag<T> &s, int, int b, int k0, int n0) const {
    const int which = b % 3, w = w0 + b / 3, N = g.N;
    const long o = mat_off(g, w);
    const double *X = (which == 0 ? g.pB : (which == 1 ? g.CB : g.DB)) + o;
    const int tid = threadIdx.x;
#pragma unroll
    for (int e = 0; e < 4; ++e) {
      const int idx = tid + 256 * e;
      const int k = idx >> 6, x = idx & 63;
      const int m = k0 + k, dd = n0 + x;
      const int c = m + dd - (N - 1);  // column m + d
      cd v = {0.0, 0.0};
      // about half of this operand is structurally zero: load only the used part
      if (m < N && dd < 2 * N - 1 && c >= 0 && c < N) v = ld(X, (long)m * N + c);
      s.set(e, k, x, (T)v.re, (T)v.im);
    }
  }

  __device__ void store(int b, int a, int dd, double re, double im) const {
    calch_store(g, w0 + b / 3, b % 3, a, dd, re, im);
  }
};

// ===========================================================================
// Problem: off-diagonal dephasing tables (CalcOffDiagDeph_E2/H2,
// dephasing.py:226-317) batched over wires.  For fixed q the sums over p
// are a product of a wire-independent Lorentzian kernel L_q[k][p] (built
// once, psa_lorentz_table) with occupation vectors of every wire:
//   De[q][k] = pi/hbar ( (1-ne_k) R_A + ne_k R_B ),
//   R_s = sum_p Lee(q,k,p) Vee2(q) a_s(p) + sum_p Leh(q,k,p) Veh2(q) b_s(p).
// K runs over [same-carrier p | other-carrier p] (2N); columns are
// (wire, s) pairs, s in {A, B}.  batch b -> qi = b / 2, kind = b % 2
// (0 = De for electrons, 1 = Dh for holes).
// ===========================================================================
// Packed row of (q index qi = q + N - 1, k) in the Lorentzian table: rows
// of q hold k in [max(0,-q), min(N,N-q)), N - |q| of them; this is the
// first row of qi.
__host__ __device__ inline long lorentz_row0(int N, int qi) {
  if (qi <= N - 1) return (long)qi * (qi + 1) / 2;
  const long m = qi - N;
  return (long)N * (N + 1) / 2 + m * (N - 1) - m * (m - 1) / 2;
}

// Lrtz(Ed, hbar g) = (hbar g / pi) / (Ed^2 + (hbar g)^2) of
// CalcOffDiagDeph_E2/H2 (dephasing.py:226-317) for kind 0 (De: rows are
// electrons, gee = gam_eh, geh = gam_h) or 1 (Dh: rows are holes,
// ghh = gam_e, geh = gam_h), K index pp: same carrier p = pp (needs p-q in
// [1, N)), other carrier p = pp - N (needs p+q in [0, N)); 0 elsewhere.
__device__ inline double lorentz_value(const psa_group &g, int kind, int q, int kr, int pp) {
  const int N = g.N, kq = kr + q;
  if (kr < 0 || kr >= N || kq < 0 || kq >= N || pp < 0 || pp >= 2 * N) return 0.0;
  const double *Es = kind == 0 ? g.Ee : g.Eh;  // carrier of row k
  const double *Eo = kind == 0 ? g.Eh : g.Ee;  // the other carrier
  double Ed, hg;
  if (pp < N) {
    const int p = pp, pm = p - q;
    if (pm < 1 || pm >= N) return 0.0;
    Ed = Es[kq] + Es[pm] - Es[p] - Es[kr];
    hg = g.hbar * (kind == 0 ? g.gam_eh : g.gam_e);
  } else {
    const int p = pp - N, pq = p + q;
    if (pq < 0 || pq >= N) return 0.0;
    Ed = Es[kq] + Eo[pq] - Eo[p] - Es[kr];
    hg = g.hbar * g.gam_h;
  }
  return (hg / PI) / (Ed * Ed + hg * hg);
}

struct OffdiagProb {
  static constexpr bool A_REAL = true, B_REAL = true;
  psa_group g;
  int w0, nw;

  // term 0: same-carrier p (K index p), valid for p - q in [1, N);
  // term 1: other-carrier p (K index N + p), valid for p + q in [0, N).
  // Rows k are used only when k + q lies in [0, N).
  __device__ int nterms(int) const { return 2; }
  __device__ bool active(int b, int m0, int) const {
    const int N = g.N, q = b / 2 - (N - 1);
    return m0 < min(N, N - q) && m0 + GBM > max(0, -q);
  }
  __device__ void krange(int t, int b, int, int, int &kb, int &ke) const {
    const int N = g.N, q = b / 2 - (N - 1);
    if (t == 0) {
      kb = max(0, q + 1);
      ke = min(N, N + q);
    } else {
      kb = N + max(0, -q);
      ke = N + min(N, N - q);
    }
  }

  // A[k][pp] from the packed Lorentzian table (lorentz_row), read 16
  // consecutive pp per row.
  template <typename T>
  __device__ void fetchA(Frag<T> &s, int t, int b, int m0, int k0) const {
    const int N = g.N, qi = b / 2, q = qi - (N - 1), kind = b % 2;
    const int kb = max(0, -q), ke = min(N, N - q);
    const double *tab = g.lor + (long)kind * N * N * 2 * N;
    const long r0 = lorentz_row0(N, qi) - kb;
    const int tid = threadIdx.x;
#pragma unroll
    for (int e = 0; e < 4; ++e) {
      const int idx = tid + 256 * e;
      const int x = idx >> 4, k = idx & 15;
      const int kr = m0 + x, pp = k0 + k;
      // clamped in-range address, masked after: the slab's loads are in flight together
      const double raw = tab[(r0 + min(max(kr, kb), ke - 1)) * 2 * N + min(pp, 2 * N - 1)];
      const double v = (kr >= kb && kr < ke && pp < 2 * N && (t == 0) == (pp < N)) ? raw : 0.0;
      s.set(e, k, x, (T)v, (T)0);
    }
  }

  // B[p][(w, s)]: V^2(q) times the occupation factor, from the wire-fastest
  // tables (|n| of the checkout level, VC^2), so lanes (columns) read
  // consecutive wires.
  template <typename T>
  __device__ void fetchB(Frag<T> &s, int t, int b, int k0, int n0) const {
    const int N = g.N, W = g.Nw, q = b / 2 - (N - 1), kind = b % 2;
    const int aq = q > 0 ? q : -q;
    const double *vsq = g.wmb + 6L * N * W, *aocc = g.wmb + 9L * N * W;
    const int tid = threadIdx.x;
#pragma unroll
    for (int e = 0; e < 4; ++e) {
      const int idx = tid + 256 * e;
      const int k = idx >> 6, x = idx & 63;
      const int pp = k0 + k, col = n0 + x;
      // same-carrier occupations: n[p-q] (1-n[p]) | (1-n[p-q]) n[p]
      // other-carrier:            o[p+q] (1-o[p]) | (1-o[p+q]) o[p]
      // loads from clamped in-range indices, masked after
      const int w = w0 + min(col >> 1, nw - 1), sA = col & 1;
      const bool same = pp < N;
      const int p = same ? pp : pp - N;
      const int j = same ? p - q : p + q;  // p-q (same) or p+q (other)
      const double *nn = aocc + ((kind == 0) == same ? 0L : (long)N * W) + w;
      const int jc = min(max(j, 0), N - 1), pc = min(max(p, 0), N - 1);
      const double a = nn[(long)jc * W], c = nn[(long)pc * W];
      const long vrow = same ? (long)(kind == 0 ? 1 : 2) * N : 0L;
      const double V2 = vsq[(vrow + min(aq, N - 1)) * W + w];
      const bool ok = pp < 2 * N && col < 2 * nw && aq > 0 && (t == 0) == same &&
                      (same ? (j >= 1 && j < N) : (j >= 0 && j < N));
      // V^2 ~ 1e-42 J^2 is below the FP32 range: the reduced-precision
      // product runs on B scaled by 2^100 (exact), undone in store()
      const double sc = sizeof(T) == 4 ? 0x1p100 : 1.0;
      const double v = ok ? sc * V2 * (sA == 0 ? a * (1.0 - c) : (1.0 - a) * c) : 0.0;
      s.set(e, k, x, (T)v, (T)0);
    }
  }

  int scaled;  // 1 when the products ran on the 2^100-scaled B operand

  __device__ void store(int b, int kr, int col, double re, double) const {
    if (scaled) re *= 0x1p-100;
    const double other = __shfl_xor(re, 1, 64);
    const int N = g.N, kind = b % 2, qi = b / 2;
    if ((col & 1) || kr >= N || col >= 2 * nw) return;
    const int w = w0 + (col >> 1);
    const double *pr = g.prep + (long)w * PSA_PREP_STRIDE(N);
    const double nk = pr[(kind == 0 ? PSA_PREP_ANE(N) : PSA_PREP_ANH(N)) + kr];
    const double D = ((1.0 - nk) * re + nk * other) * PI / g.hbar;
    g.dedh[(((long)w * 2 + kind) * N + kr) * (2 * N - 1) + qi] = D;
  }
};

// ===========================================================================
// Problem: Prop2QW for Ex, Ey, Ez (qwoptics.py:667-715): spline rr -> r,
// window, real part and FFTG are linear, so the FFT bins the wire needs
// (q = -(N-1)..N-1) are one complex matrix G [2N-1][gnx] applied to the
// real parts of the fields.  Columns = (wire, component).
// ===========================================================================
struct Prop2qwProb {
  static constexpr bool A_REAL = false, B_REAL = true;
  psa_group g;
  int w0, nw;

  __device__ int nterms(int) const { return 1; }
  __device__ bool active(int, int, int) const { return true; }
  __device__ void krange(int, int, int, int, int &kb, int &ke) const {
    kb = 0;
    ke = g.gnx;
  }

  template <typename T>
  __device__ void fetchA(Frag<T> &s, int, int, int m0, int k0) const {
    load_dense(s, g.gmat, g.gnx, 2 * g.N - 1, g.gnx, true, false, 1.0, m0, k0);
  }

  template <typename T>
  __device__ void fetchB(Frag<T> &s, int, int, int k0, int n0) const {
    const int tid = threadIdx.x;
#pragma unroll
    for (int e = 0; e < 4; ++e) {
      const int idx = tid + 256 * e;
      const int x = idx >> 4, k = idx & 15;
      const int col = n0 + x, xi = k0 + k;
      double v = 0.0;
      if (col < 3 * nw && xi < g.gnx) {
        const int w = w0 + col / 3, f = col % 3;
        v = g.efield[(((long)f * g.Nw + w) * g.Nx + g.gx0 + xi) * 2];
      }
      s.set(e, k, x, (T)v, (T)0);
    }
  }

  __device__ void store(int, int j, int col, double re, double im) const {
    const int N = g.N;
    if (j >= 2 * N - 1 || col >= 3 * nw) return;
    const int w = w0 + col / 3, f = col % 3;
    st(g.exq + (long)w * 3 * 2 * (2 * N - 1), (long)f * (2 * N - 1) + j, cd{re, im});
  }
};

// ===========================================================================
// Problem: the source sums of QWPolarization3 / QWRho5 (qwoptics.py:809-1071)
// in real space, val[r][(w, f)] = sum_s phi[s][r] S_f(s) over the diagonal
// sums S_f of the new level (k_diagsums mode 1): f = 0 dcv sum (-1)^kh p,
// 1 dcv sum p, 2 sum C3, 3 sum conj(D3).  The epilogue applies the window,
// the polarisation prefactors and the yw(w), (-1)^w signs.
// ===========================================================================
struct SrcProb {
  static constexpr bool A_REAL = false, B_REAL = false;
  psa_group g;
  int w0, nw;

  __device__ int nterms(int) const { return 1; }
  __device__ bool active(int, int, int) const { return true; }
  __device__ void krange(int, int, int, int, int &kb, int &ke) const {
    kb = 0;
    ke = 2 * g.N - 1;
  }

  template <typename T>
  __device__ void fetchA(Frag<T> &s, int, int, int m0, int k0) const {
    load_dense(s, g.phi, g.Nr, g.Nr, 2 * g.N - 1, false, false, 1.0, m0, k0);
  }

  template <typename T>
  __device__ void fetchB(Frag<T> &s, int, int, int k0, int n0) const {
    const int N = g.N, Q = 2 * N - 1, tid = threadIdx.x;
#pragma unroll
    for (int e = 0; e < 4; ++e) {
      const int idx = tid + 256 * e;
      const int x = idx >> 4, k = idx & 15;
      const int col = n0 + x, si = k0 + k;
      cd v = {0.0, 0.0};
      if (col < 4 * nw && si < Q) {
        const int w = w0 + (col >> 2), f = col & 3;
        v = ld(g.src + (long)w * PSA_SRC_STRIDE(N, g.Nr) + PSA_SRC_DS(g.Nr), (long)f * Q + si);
      }
      s.set(e, k, x, (T)v.re, (T)v.im);
    }
  }

  __device__ void store(int, int r, int col, double re, double im) const {
    const int N = g.N, Nr = g.Nr;
    if (r >= Nr || col >= 4 * nw) return;
    const int w = w0 + (col >> 2), f = col & 3;
    double *src = g.src + (long)w * PSA_SRC_STRIDE(N, Nr);
    const double W = g.window[r];
    const double sP = 2.0 * g.ehint / g.area / g.L;
    const bool optics = (g.flags & PSA_F_OPTICS) != 0;
    if (f == 0) {
      src[r] = optics ? re * W * sP : 0.0;
    } else if (f == 1) {
      const int wid = g.wire_id[w];
      src[Nr + r] = optics ? re * W * yw_of(wid) * sP : 0.0;
      src[2 * Nr + r] = optics ? -re * W * sgn_pow(wid) * sP : 0.0;
    } else {
      const double fl = (g.flags & PSA_F_LF) ? W / (2.0 * g.L) : 0.0;
      st(src + PSA_SRC_RC(Nr), (f == 2 ? 0 : Nr) + r, cd{re * fl, im * fl});
    }
  }
};

// ===========================================================================
// Problem: QW2Prop splines r -> rr (qwoptics.py:792-805) for Px, Py, Pz, re,
// rh of every wire as one real product with the spline matrix m2
// (stored [Nr][bnx]); rows are the rr points bx0 .. bx0+bnx-1 inside the
// wire's r range (the others stay zero).  RhoE/RhoH go to the real and
// imaginary slots of Rho (|.| with LF) for k_qw2prop_b.
// ===========================================================================
struct Qw2pProb {
  static constexpr bool A_REAL = true, B_REAL = true;
  psa_group g;
  int w0, nw;

  __device__ int nterms(int) const { return 1; }
  __device__ bool active(int, int, int) const { return true; }
  __device__ void krange(int, int, int, int, int &kb, int &ke) const {
    kb = 0;
    ke = g.Nr;
  }

  template <typename T>
  __device__ void fetchA(Frag<T> &s, int, int, int m0, int k0) const {
    const int tid = threadIdx.x;
#pragma unroll
    for (int e = 0; e < 4; ++e) {
      const int idx = tid + 256 * e;
      const int k = idx >> 6, x = idx & 63;
      const int b = m0 + x, r = k0 + k;
      double v = 0.0;
      if (b < g.bnx && r < g.Nr) v = g.m2[(long)r * g.bnx + b];
      s.set(e, k, x, (T)v, (T)0);
    }
  }

  template <typename T>
  __device__ void fetchB(Frag<T> &s, int, int, int k0, int n0) const {
    const int N = g.N, Nr = g.Nr, tid = threadIdx.x;
#pragma unroll
    for (int e = 0; e < 4; ++e) {
      const int idx = tid + 256 * e;
      const int x = idx >> 4, k = idx & 15;
      const int col = n0 + x, r = k0 + k;
      double v = 0.0;
      if (col < 5 * nw && r < Nr) {
        const int w = w0 + col / 5, f = col % 5;
        v = g.src[(long)w * PSA_SRC_STRIDE(N, Nr) + f * Nr + r];
      }
      s.set(e, k, x, (T)v, (T)0);
    }
  }

  __device__ void store(int, int b, int col, double re, double) const {
    if (b >= g.bnx || col >= 5 * nw) return;
    const int w = w0 + col / 5, f = col % 5, x = g.bx0 + b;
    const long NW = g.Nw, Nx = g.Nx;
    if (f < 3) {
      st(g.out + ((f * NW + w) * Nx) * 2, x, cd{re, 0.0});
    } else {
      const bool lf = (g.flags & PSA_F_LF) != 0;
      double *Rho = g.out + ((3 * NW + w) * Nx) * 2;
      Rho[2 * x + (f - 3)] = lf ? fabs(re) : re;  // RhoE (re slot), RhoH (im slot)
    }
  }
};

}  // namespace psa
