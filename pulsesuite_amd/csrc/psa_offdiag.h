// Off-diagonal dephasing tables (CalcOffDiagDeph_E2/H2, dephasing.py:226-317)
// as one MFMA kernel that generates its Lorentzian operand in registers.
//
// For fixed q and kind (0 = De, rows are electrons; 1 = Dh, rows are holes)
//   D[q][k] = pi/hbar ( (1-n_k) R_A(k) + n_k R_B(k) ),
//   R_s(k)  = sum_p L(q,k,p) V2(q) f_s(p),
// with K running over the same-carrier p (p - q in [1, N)) and then the
// other-carrier p (p + q in [0, N)), L the Lorentzian of the energy mismatch
// and f_A = n[j](1-n[p]), f_B = (1-n[j]) n[p] (j = p -/+ q) the occupation
// columns of every wire.  Columns are (wire, s) pairs.
//
// Block = 8 waves of 32 rows (FP32) or 4 of 64 (FP64) over 256 rows k x 64
// columns of one (q, kind).  The
// occupation operand B [K][64] is built cooperatively, 32 K indices at a
// time, into LDS (16-byte units [k / E][col][E], E = 16 / sizeof(T), so a
// lane's four K values of one MFMA group are one or two ds_read_b128), with
// the next chunk's global loads in flight during the current chunk's MFMAs.
// The Lorentzian operand A is computed per lane from the band energies: the
// row energies stay in registers, the K-side energies sit in LDS.  Lane l
// supplies A[k = l%16][K = 4(l/16) + j] to MFMA j of a 16-deep slab, and B
// with the same K order, so the four MFMAs of a slab cover its 16 K values.
//
// FP64 computes the Lorentzian exactly as the reference expression; FP32
// (the reduced-precision modes) forms the mismatch in FP64 and the rest in
// FP32 on x = Ed / (hbar g), and runs on the 2^100-scaled occupation columns
// (V^2 ~ 1e-42 J^2 is below the FP32 range), unscaled in the epilogue.
#pragma once

#include "psa_gemm.h"

namespace psa {
namespace odg {

constexpr int ROWS = 256;  // rows per block
// waves per block and 16-row blocks per wave: FP32 8 x 32 rows (4 waves per
// SIMD fit), FP64 4 x 64 rows (its accumulators take twice the registers)
template <typename T>
struct Cfg {
  static constexpr int WAVES = sizeof(T) == 4 ? 8 : 4, RB = ROWS / (16 * WAVES), THREADS = 64 * WAVES;
};
constexpr int COLS = 64;    // (wire, s) columns per block
constexpr int KC = 32;      // K indices per LDS chunk
constexpr int KMAX = 1024;  // K <= 2N - 1 for N <= 512
constexpr int NMAX = 512;

template <typename T>
struct Unit;
template <>
struct Unit<float> {
  typedef fv4 t;
};
template <>
struct Unit<double> {
  typedef double t __attribute__((ext_vector_type(2)));
};

template <typename T>
__global__ __launch_bounds__(Cfg<T>::THREADS, 2) void k_offdiag(psa_group g, int w0, int nw, int scaled) {
  constexpr int WAVES = Cfg<T>::WAVES, RB = Cfg<T>::RB, THREADS = Cfg<T>::THREADS;
  constexpr int E = 16 / sizeof(T);       // elements per 16-byte unit
  constexpr int PER = KC * COLS / THREADS;  // B elements per thread per chunk
  typedef typename Unit<T>::t unit_t;
  __shared__ double e1[KMAX], e2[KMAX];  // K-side energies: E[p -/+ q], E[p]
  __shared__ unit_t bs[2][KC / E * COLS];
  // per-term Lorentzian constants, read by term index in the inner loop (a
  // register select of two constants would let the compiler move their
  // divisions into the loop)
  __shared__ double tcst[2][2];

  const int N = g.N, W = g.Nw, Q = 2 * N - 1;
  const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
  // heavy q first: z = 0, 1 -> q = 0; 2, 3 -> q = 1; 4, 5 -> q = -1; ...
  const int pr = blockIdx.z >> 1, kind = blockIdx.z & 1;
  const int q = (pr & 1) ? (pr + 1) >> 1 : -(pr >> 1), qi = q + N - 1;
  const int aq = q < 0 ? -q : q;
  const int col0 = blockIdx.y * COLS, row0 = blockIdx.x * ROWS;
  const int rb = max(0, -q), re = min(N, N - q);  // rows with k + q in [0, N)
  const int kb0 = max(0, q + 1), K0 = max(0, min(N, N + q) - kb0);
  const int kb1 = rb;
  // q = 0 is excluded (the occupation columns are zero there)
  const int K = (q == 0 || row0 >= re || row0 + ROWS <= rb) ? 0 : K0 + (re - rb);
  const int nch = (K + KC - 1) / KC;
  const double *Es = kind == 0 ? g.Ee : g.Eh;  // carrier of the rows
  const double *Eo = kind == 0 ? g.Eh : g.Ee;  // the other carrier

  for (int kk = tid; kk < nch * KC; kk += THREADS) {
    double a = 0.0, b = 0.0;
    if (kk < K0) {
      const int p = kb0 + kk;
      a = Es[p - q];
      b = Es[p];
    } else if (kk < K) {
      const int p = kb1 + kk - K0;
      a = Eo[p + q];
      b = Eo[p];
    }
    e1[kk] = a;
    e2[kk] = b;
  }

  // column owned by this thread when building B
  const int cB = tid & (COLS - 1), colg = col0 + cB;
  const bool colok = colg < 2 * nw;
  const int wB = w0 + min(colg >> 1, nw - 1), sB = colg & 1;
  const double *occ = g.wmb + 9L * N * W;  // |n| of the checkout level [2][N][W]
  const double *vsq = g.wmb + 6L * N * W;  // VC(d)^2 [3][N][W]
  const double *nnS = occ + (kind == 0 ? 0L : (long)N * W) + wB;
  const double *nnO = occ + (kind == 0 ? (long)N * W : 0L) + wB;
  const double sc = scaled ? 0x1p100 : 1.0;
  const int aqc = min(aq, N - 1);
  const double v2S = colok ? sc * vsq[((long)(kind == 0 ? 1 : 2) * N + aqc) * W + wB] : 0.0;
  const double v2O = colok ? sc * vsq[(long)aqc * W + wB] : 0.0;

  double ra[PER], rc[PER];
  // element e of this thread: unit (tid >> 6) + WAVES (e / E), position e % E
  auto issue = [&](int ch) {
#pragma unroll
    for (int e = 0; e < PER; ++e) {
      const int kk = ch * KC + ((tid >> 6) + WAVES * (e / E)) * E + e % E;
      int p = 0, j = 0;
      const double *nn = nnS;
      if (kk < K0) {
        p = kb0 + kk;
        j = p - q;
      } else if (kk < K) {
        p = kb1 + kk - K0;
        j = p + q;
        nn = nnO;
      }
      ra[e] = nn[(long)j * W];
      rc[e] = nn[(long)p * W];
    }
  };
  auto commit = [&](int ch, unit_t *dst) {
#pragma unroll
    for (int u = 0; u < PER / E; ++u) {
      unit_t val;
#pragma unroll
      for (int s = 0; s < E; ++s) {
        const int e = u * E + s;
        const int kk = ch * KC + ((tid >> 6) + WAVES * u) * E + s;
        const double a = ra[e], c = rc[e];
        const double f = sB ? (1.0 - a) * c : a * (1.0 - c);
        val[s] = (T)(kk < K ? (kk < K0 ? v2S : v2O) * f : 0.0);
      }
      dst[((tid >> 6) + WAVES * u) * COLS + cB] = val;
    }
  };

  using acc_t = typename Mfma<T>::acc;
  acc_t acc[RB][4];
#pragma unroll
  for (int i = 0; i < RB; ++i)
#pragma unroll
    for (int c = 0; c < 4; ++c) acc[i][c] = acc_t{0, 0, 0, 0};

  const int g4 = lane >> 4, r16 = lane & 15;
  const int wrow0 = row0 + 16 * RB * wv;
  const bool wact = K > 0 && wrow0 < re && wrow0 + 16 * RB > rb;  // wave-uniform
  double ekq[RB], ek[RB];
#pragma unroll
  for (int i = 0; i < RB; ++i) {
    const int k = wrow0 + 16 * i + r16;
    ek[i] = Es[min(k, N - 1)];
    ekq[i] = Es[min(max(k + q, 0), N - 1)];
  }
  const double hgS = g.hbar * (kind == 0 ? g.gam_eh : g.gam_e), hgO = g.hbar * g.gam_h;
  // Lrtz(Ed, hg) = (hg / pi) / (Ed^2 + hg^2); FP32: (1 / (pi hg)) / (x^2 + 1), x = Ed / hg
  if (tid < 2) {
    const double hg = tid == 0 ? hgS : hgO;
    if constexpr (sizeof(T) == 8) {
      tcst[tid][0] = hg / PI;
      tcst[tid][1] = hg * hg;
    } else {
      tcst[tid][0] = 1.0 / hg;
      tcst[tid][1] = 1.0 / (PI * hg);
    }
  }
  double ud[RB];
#pragma unroll
  for (int i = 0; i < RB; ++i) ud[i] = ekq[i] - ek[i];

  if (K > 0) {
    issue(0);
    commit(0, bs[0]);
    __syncthreads();
    for (int ch = 0; ch < nch; ++ch) {
      const bool more = ch + 1 < nch;
      if (more) issue(ch + 1);
      if (wact) {
        const unit_t *B = bs[ch & 1];
#pragma unroll
        for (int sl = 0; sl < KC / 16; ++sl) {
          const int kl = 16 * sl + 4 * g4;  // this lane's first K index in the chunk
          const int kk = ch * KC + kl;
          // one 16-byte unit per column block holds E of the lane's K values;
          // the Lorentzian column for each K value is made just before its MFMAs
#pragma unroll
          for (int u = 0; u < 4 / E; ++u) {
            unit_t bu[4];
#pragma unroll
            for (int c = 0; c < 4; ++c) bu[c] = B[(kl / E + u) * COLS + 16 * c + r16];
#pragma unroll
            for (int s = 0; s < E; ++s) {
              const int j = u * E + s;
              const double d1 = e1[kk + j], d2 = e2[kk + j];
              const double *tc = tcst[kk + j < K0 ? 0 : 1];
              T a[RB];
              if constexpr (sizeof(T) == 8) {
                const double c1 = tc[0], c2 = tc[1];
#pragma unroll
                for (int i = 0; i < RB; ++i) {
                  const double Ed = ekq[i] + d1 - d2 - ek[i];
                  a[i] = c1 / (Ed * Ed + c2);
                }
              } else {
                const float ih = (float)tc[0], f1 = (float)tc[1];
                const double dv = d1 - d2;
#pragma unroll
                for (int i = 0; i < RB; ++i) {
                  const float x = (float)(ud[i] + dv) * ih;
                  a[i] = f1 * __builtin_amdgcn_rcpf(fmaf(x, x, 1.0f));
                }
              }
#pragma unroll
              for (int i = 0; i < RB; ++i)
#pragma unroll
                for (int c = 0; c < 4; ++c) acc[i][c] = Mfma<T>::mma(a[i], (T)bu[c][s], acc[i][c]);
            }
          }
        }
      }
      if (more) commit(ch + 1, bs[(ch + 1) & 1]);
      __syncthreads();
    }
  }

  // epilogue: lanes l, l^1 hold the A and B columns of one wire
  const double *pbase = g.prep;
  const double unscale = scaled ? 0x1p-100 : 1.0;
  const double pih = PI / g.hbar;
#pragma unroll
  for (int i = 0; i < RB; ++i)
#pragma unroll
    for (int c = 0; c < 4; ++c)
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        const double rA = (double)acc[i][c][v] * unscale;
        const double rBv = __shfl_xor(rA, 1, 64);
        const int kr = wrow0 + 16 * i + Mfma<T>::row(g4, v);
        const int cg = col0 + 16 * c + r16;
        if ((cg & 1) || cg >= 2 * nw || kr >= N) continue;
        const int w = w0 + (cg >> 1);
        double D = 0.0;
        if (K > 0 && kr >= rb && kr < re) {
          const double *pw = pbase + (long)w * PSA_PREP_STRIDE(N);
          const double nk = pw[(kind == 0 ? PSA_PREP_ANE(N) : PSA_PREP_ANH(N)) + kr];
          D = ((1.0 - nk) * rA + nk * rBv) * pih;
        }
        g.dedh[(((long)w * 2 + kind) * N + kr) * Q + qi] = D;
      }
}

// ---------------------------------------------------------------------------
// Split-BF16 tables (PSA_PREC_BF16X3): the same sums on
// v_mfma_f32_16x16x32_bf16.  Both operands are split as
// hi = bf16(f), lo = bf16(f - hi) of their FP32 value f, and every product
// is hi hi + hi lo + lo hi with FP32 accumulation (relative error of a
// product ~2^-16; the terms are positive, so the sums keep that bound).
//
// Block = 8 waves of 32 rows over 256 rows starting at the first row rb
// with k + q in [0, N), x 128 columns (64 wires x the A and B sums).  Rows
// past re are skipped wave by wave; blocks past re exit at once, so the
// shrinking (q, k) triangle costs no idle blocks.  Entries with k + q
// outside [0, N) and q = 0 are not written (the tables are allocated
// zeroed and no kernel writes them).
//
// The K index runs over the same-carrier p, zero padding to a multiple of
// the 32-K chunk, then the other-carrier p, so every chunk is of one
// carrier.  B [K][128] (the occupation columns times V^2 / (pi hg)) is built
// per chunk by all threads (thread = wire, 4 consecutive K: one pair of
// occupation loads gives both columns of the wire), split, and stored as
// [hi/lo][column][32 K] bf16 with the 16-byte group of 8 K at
// (K >> 3) ^ ((column >> 1) & 3) (the MFMA reads, 16 columns x 4 groups per
// instruction, fall on distinct banks), double-buffered: chunk ch + 1 is
// built between chunk ch's Lorentzian and its products.  The Lorentzian
// 1 / (x^2 + 1) is made per lane in FP32 (8 consecutive K of one row per
// MFMA) from x = (Es[k + q] - Es[k]) / hg (registers) + (e1 - e2) / hg
// (LDS), summed in FP64, then split.  The occupation columns are the MFMA A
// operand and the Lorentzian B, so that each lane ends with both sums (A,
// B) of two wires for one row k.
// ---------------------------------------------------------------------------
namespace bfx {
constexpr int WAVES = 8, THREADS = 64 * WAVES, ROWS = 32 * WAVES, COLS = 128, KC = 32;
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef __bf16 bf16x4 __attribute__((ext_vector_type(4)));
typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));
typedef uint32_t u32x2 __attribute__((ext_vector_type(2)));
}  // namespace bfx

__global__ __launch_bounds__(bfx::THREADS, 4) void k_offdiag_bf(psa_group g, int w0, int nw) {
  constexpr int THREADS = bfx::THREADS, ROWS = bfx::ROWS, COLS = bfx::COLS, KC = bfx::KC;
  using bfx::bf16x4;
  using bfx::bf16x8;
  using bfx::u32x2;
  using bfx::u32x4;
  __shared__ double dvk[KMAX];                                              // (e1 - e2) / hg per K index
  __shared__ __attribute__((aligned(16))) uint16_t bsp[2][2][COLS * KC];  // [buffer][hi, lo][column][K]

  const int N = g.N, W = g.Nw, Q = 2 * N - 1;
  const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
  const int pr = blockIdx.z >> 1, kind = blockIdx.z & 1;
  const int q = (pr & 1) ? (pr + 1) >> 1 : -(pr >> 1), qi = q + N - 1;
  const int aq = q < 0 ? -q : q;
  const int rb = max(0, -q), re = min(N, N - q);
  const int row0 = rb + ROWS * blockIdx.x;
  // K layout: same-carrier p in [0, K0), zero padding to K0p (a multiple
  // of the chunk, so that every chunk is of one carrier), other-carrier p
  // in [K0p, K)
  const int kb0 = max(0, q + 1), K0 = max(0, min(N, N + q) - kb0), kb1 = rb;
  const int K0p = (K0 + KC - 1) / KC * KC;
  const int K = q == 0 ? 0 : K0p + (re - rb);
  if (K == 0 || row0 >= re) return;  // block-uniform, before any barrier
  const int nch = (K + KC - 1) / KC;
  const double *Es = kind == 0 ? g.Ee : g.Eh;
  const double *Eo = kind == 0 ? g.Eh : g.Ee;
  // Lrtz(Ed, hg) = (1 / (pi hg)) / (x^2 + 1), x = Ed / hg, hg of the
  // same-carrier (hgS) or other-carrier (hgO) K value:
  // x = (Es[k + q] - Es[k]) / hg + (e1 - e2) / hg in FP64; the factor
  // 1 / (pi hg) goes into the occupation operand
  const double hgS = g.hbar * (kind == 0 ? g.gam_eh : g.gam_e), hgO = g.hbar * g.gam_h;
  for (int kk = tid; kk < nch * KC; kk += THREADS) {
    double d = 0.0;
    if (kk < K0) {
      const int p = kb0 + kk;
      d = (Es[p - q] - Es[p]) / hgS;
    } else if (kk >= K0p && kk < K) {
      const int p = kb1 + kk - K0p;
      d = (Eo[p + q] - Eo[p]) / hgO;
    }
    dvk[kk] = d;
  }

  // B builder: wire wl of the block's 64, K values 4 kq .. 4 kq + 3 of each chunk
  const int wl = tid & 63, kq = __builtin_amdgcn_readfirstlane(tid >> 6);  // wave-uniform
  const int wcol = blockIdx.y * 64 + wl;
  const bool wok = wcol < nw;
  const int wB = w0 + min(wcol, nw - 1);
  const double *occ = g.wmb + 9L * N * W;  // |n| of the checkout level [2][N][W]
  const double *vsq = g.wmb + 6L * N * W;  // VC(d)^2 [3][N][W]
  const double *nnS = occ + (kind == 0 ? 0L : (long)N * W);  // wire wB at offset wB
  const double *nnO = occ + (kind == 0 ? (long)N * W : 0L);
  const int aqc = min(aq, N - 1);
  // V^2 / (pi hg), 2^100-scaled, as FP32: its rounding is far below the split's
  const float v2S = wok ? (float)(0x1p100 * vsq[((long)(kind == 0 ? 1 : 2) * N + aqc) * W + wB] / (PI * hgS)) : 0.0f;
  const float v2O = wok ? (float)(0x1p100 * vsq[(long)aqc * W + wB] / (PI * hgO)) : 0.0f;
  // the occupations are read as the upper 32 bits of their FP64 values
  // (21 significant bits; the split keeps 16), one register each
  uint32_t ra[4], rc[4];
  auto issue = [&](int ch) {
#pragma unroll
    for (int e = 0; e < 4; ++e) {
      const int kk = ch * KC + 4 * kq + e;
      int p = 0, j = 0;
      const double *nn = nnS;
      if (kk < K0) {
        p = kb0 + kk;
        j = p - q;
      } else if (kk >= K0p && kk < K) {
        p = kb1 + kk - K0p;
        j = p + q;
        nn = nnO;
      }
      ra[e] = reinterpret_cast<const uint32_t *>(nn + (long)j * W + wB)[1];
      rc[e] = reinterpret_cast<const uint32_t *>(nn + (long)p * W + wB)[1];
    }
  };
  // columns 2 wl (A: n[j] (1 - n[p])) and 2 wl + 1 (B: (1 - n[j]) n[p])
  const int cA = 2 * wl, sw = (wl & 3);
  const int boff = cA * KC + (((kq >> 1) ^ sw) << 3) + 4 * (kq & 1);
  auto commit = [&](int ch, int buf) {
    bf16x4 hA, lA, hB, lB;
#pragma unroll
    for (int e = 0; e < 4; ++e) {
      const int kk = ch * KC + 4 * kq + e;
      const double v = kk < K0 ? (double)v2S : ((kk >= K0p && kk < K) ? (double)v2O : 0.0);
      const double na = __builtin_bit_cast(double, (uint64_t)ra[e] << 32);
      const double nc = __builtin_bit_cast(double, (uint64_t)rc[e] << 32);
      const float fa = (float)(v * (na * (1.0 - nc))), fb = (float)(v * ((1.0 - na) * nc));
      hA[e] = (__bf16)fa;
      lA[e] = (__bf16)(fa - (float)hA[e]);
      hB[e] = (__bf16)fb;
      lB[e] = (__bf16)(fb - (float)hB[e]);
    }
    uint16_t *h = bsp[buf][0], *l = bsp[buf][1];
    *reinterpret_cast<u32x2 *>(&h[boff]) = __builtin_bit_cast(u32x2, hA);
    *reinterpret_cast<u32x2 *>(&l[boff]) = __builtin_bit_cast(u32x2, lA);
    *reinterpret_cast<u32x2 *>(&h[boff + KC]) = __builtin_bit_cast(u32x2, hB);
    *reinterpret_cast<u32x2 *>(&l[boff + KC]) = __builtin_bit_cast(u32x2, lB);
  };

  fv4 acc[2][8];
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int c = 0; c < 8; ++c) acc[i][c] = fv4{0.f, 0.f, 0.f, 0.f};

  const int g4 = lane >> 4, r16 = lane & 15;
  const int wrow0 = row0 + 32 * wv;
  const bool wact = wrow0 < re;  // wave-uniform
  double udS[2], udO[2];  // (Es[k + q] - Es[k]) / hg of this lane's rows
#pragma unroll
  for (int i = 0; i < 2; ++i) {
    const int k = min(wrow0 + 16 * i + r16, N - 1);
    const double ud = Es[min(max(k + q, 0), N - 1)] - Es[k];
    udS[i] = ud / hgS;
    udO[i] = ud / hgO;
  }

  issue(0);
  commit(0, 0);
  __syncthreads();
  // per chunk: this chunk's Lorentzian, the next chunk's B into the other
  // buffer (last read before the previous barrier), then the products
  for (int ch = 0; ch < nch; ++ch) {
    const bool more = ch + 1 < nch;
    bf16x8 ah[2], al[2];
    if (wact) {
      const bool same = ch * KC < K0p;  // chunk-uniform
      const int kk0 = ch * KC + 8 * g4;  // this lane's 8 K values
#pragma unroll
      for (int i = 0; i < 2; ++i) {
        const double ud = same ? udS[i] : udO[i];
#pragma unroll
        for (int e = 0; e < 8; ++e) {
          const float x = (float)(ud + dvk[kk0 + e]);
          const float a = __builtin_amdgcn_rcpf(fmaf(x, x, 1.0f));
          ah[i][e] = (__bf16)a;
          al[i][e] = (__bf16)(a - (float)ah[i][e]);
        }
      }
    }
    if (more) {
      issue(ch + 1);
      commit(ch + 1, (ch + 1) & 1);
    }
    if (wact) {
      // occupation columns as the A operand (rows m = columns), Lorentzian
      // as B (columns n = rows k): D[m][n] lands in lane n % 16, m = 4 (l / 16) + v
      const uint16_t *Bh = bsp[ch & 1][0], *Bl = bsp[ch & 1][1];
#pragma unroll
      for (int c = 0; c < 8; ++c) {
        const int col = 16 * c + r16;
        const int off = col * KC + ((g4 ^ ((col >> 1) & 3)) << 3);
        const bf16x8 bh = __builtin_bit_cast(bf16x8, *reinterpret_cast<const u32x4 *>(&Bh[off]));
        const bf16x8 bl = __builtin_bit_cast(bf16x8, *reinterpret_cast<const u32x4 *>(&Bl[off]));
#pragma unroll
        for (int i = 0; i < 2; ++i) acc[i][c] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(bh, ah[i], acc[i][c], 0, 0, 0);
#pragma unroll
        for (int i = 0; i < 2; ++i) acc[i][c] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(bl, ah[i], acc[i][c], 0, 0, 0);
#pragma unroll
        for (int i = 0; i < 2; ++i) acc[i][c] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(bh, al[i], acc[i][c], 0, 0, 0);
        __builtin_amdgcn_sched_barrier(0);
      }
    }
    __syncthreads();
  }
  if (!wact) return;

  // epilogue: acc[i][c][v] is row k = wrow0 + 16 i + lane % 16 of column
  // 16 c + 4 (lane / 16) + v, i.e. the A (v even) and B (v odd) sums of
  // wire 8 c + 2 (lane / 16) + v / 2: both sums of a wire in one lane
  const double pih = PI / g.hbar * 0x1p-100;
  const long aoff = kind == 0 ? PSA_PREP_ANE(N) : PSA_PREP_ANH(N);
#pragma unroll
  for (int i = 0; i < 2; ++i) {
    const int kr = wrow0 + 16 * i + r16;
    if (kr >= re) continue;
#pragma unroll
    for (int c = 0; c < 8; ++c)
#pragma unroll
      for (int h = 0; h < 2; ++h) {
        const int wc = blockIdx.y * 64 + 8 * c + 2 * g4 + h;
        if (wc >= nw) continue;
        const int w = w0 + wc;
        const double nk = g.prep[(long)w * PSA_PREP_STRIDE(N) + aoff + kr];
        const double rA = (double)acc[i][c][2 * h], rB = (double)acc[i][c][2 * h + 1];
        g.dedh[(((long)w * 2 + kind) * N + kr) * Q + qi] = ((1.0 - nk) * rA + nk * rB) * pih;
        if (h) __builtin_amdgcn_sched_barrier(0);
      }
  }
}

}  // namespace odg
}  // namespace psa
