// Hartree-Fock exchange sums of CalcH (_CalcH_jit, SBEs.py:808-924) as one
// MFMA kernel per (wire, Hamiltonian, column block, row group).
//
// For the diagonal offset d = k1 - k2 the exchange term is
//   O_d[a] = sum_{m != a} V(|m - a|) X[m][m + d],   X = p (Heh), C (Hee), D (Hhh),
// with a and m both running over R_d = [max(0, -d), min(N, N - d)).  The
// Toeplitz matrix V(|m - a|) (zero diagonal) is shift-invariant, so a block
// of 64 columns d can run on rows a' = a - s and K m' = m - s with one shift
// s per half: columns 0..31 are d = 32 cb + 1 .. 32 cb + 32 (s = 0) and
// columns 32..63 are d = -32 cb .. -(32 cb + 31) (s = 32 cb).  Both halves
// then need rows and K only below N - 32 cb, which skips the unused half of
// the (a, d) square without idle waves, and the lanes of one MFMA column
// block hold consecutive d of one sign, so the transposed stores
// H[k2 = a][k1 = a + d] are contiguous.
//
// FP32 (the reduced-precision modes); FP64 stays on the GEMM engine.
// Block = 8 waves of 32 rows.  Y[m'][col] =
// X[m' + s][m' + s + d] is staged in LDS 32 K values at a time (16-byte units
// [k / E][col][E], separate real and imaginary planes) with the next chunk's
// loads in flight; the Toeplitz operand comes from a per-block LDS table of
// V.  The epilogue loads each column's Meh or Hartree constants once, then
// adds them to the exchange sums (_CalcH_jit, SBEs.py:808-924).
#pragma once

#include "psa_gemm.h"

namespace psa {
namespace chx {

constexpr int KC = 32;     // K values per LDS chunk
constexpr int COLS = 64;   // columns per block
constexpr int NMAX = 512;

template <typename T>
struct Cfg;
template <>
struct Cfg<float> {
  static constexpr int WAVES = 8, RB = 2;  // waves per block, 16-row blocks per wave
  typedef fv4 unit;
};

template <typename T>
constexpr int rows_per_block() {
  return Cfg<T>::WAVES * 16 * Cfg<T>::RB;
}
template <typename T>
constexpr int threads_per_block() {
  return 64 * Cfg<T>::WAVES;
}

// diagonal offset d and row/K shift s of column x (0..63) of column block cb
__device__ __forceinline__ void col_info(int cb, int x, int &d, int &s) {
  if (x < 32) {
    d = 32 * cb + 1 + x;
    s = 0;
  } else {
    d = -(32 * cb + x - 32);
    s = 32 * cb;
  }
}

// Epilogue of the exchange kernels: rows a' = wrow0 + 16 i + 4 (lane / 16) + v
// and columns 16 c + lane % 16 of the accumulators (the 16 x 16 MFMA output
// layout of both kernels); the constants of all four column blocks first
// (their loads in flight together), then the stores.
__device__ __forceinline__ void calch_epilogue(const psa_group &g, int w, int which, int cb, int wrow0, int g4,
                                               int r16, const fv4 (&aR)[2][4], const fv4 (&aI)[2][4]) {
  constexpr int RB = 2;
  const int N = g.N;
  const double *pr = g.prep + (long)w * PSA_PREP_STRIDE(N);
  const long ow = mat_off(g, w);
  const bool lf = (g.flags & PSA_F_LF) != 0, ee = which == 1;
  cd k0[4], k1c[4];  // which 0: Meh for even / odd k2; else the Hartree term (in k0)
#pragma unroll
  for (int c = 0; c < 4; ++c) {
    int d, sh;
    col_info(cb, 16 * c + r16, d, sh);
    const int ad = min(d < 0 ? -d : d, N - 1), dd = min(d + N - 1, 2 * N - 2);
    k0[c] = cd{0.0, 0.0};
    k1c[c] = cd{0.0, 0.0};
    if (which == 0) {
      if (g.flags & PSA_F_OPTICS) {
        const int wid = g.wire_id[w];
        const double *eq = g.exq + (long)w * 3 * 2 * (2 * N - 1);
        const cd ex = ld(eq, dd), ey = ld(eq, (2 * N - 1) + dd), ez = ld(eq, 2 * (2 * N - 1) + dd);
        const double yv = yw_of(wid), zs = sgn_pow(wid);
        // -ehint*Xcv*Ex - ehint*Ycv*Ey*yw - ehint*Zcv*Ez*(-1)^w, Xcv = dcv (-1)^k2, Ycv = dcv, Zcv = -dcv
        const double a2 = -g.ehint * g.dcv, a3 = -g.ehint * (-g.dcv);
        const double ae = -g.ehint * g.dcv, ao = -g.ehint * -g.dcv;
        k0[c] = cd{ae * ex.re + (a2 * ey.re) * yv + (a3 * ez.re) * zs,
                   ae * ex.im + (a2 * ey.im) * yv + (a3 * ez.im) * zs};
        k1c[c] = cd{ao * ex.re + (a2 * ey.re) * yv + (a3 * ez.re) * zs,
                    ao * ex.im + (a2 * ey.im) * yv + (a3 * ez.im) * zs};
      }
    } else if (lf) {
      const double *vcp = pr + PSA_PREP_VC(N);
      const double vself = vcp[(ee ? 1 : 2) * N + ad], veh = vcp[ad];
      // Hee: Vee(d) sum_k C[k,k+d] - Veh(d) sum_k D[k+d,k]; Hhh: C <-> D
      const cd dsS = ld(pr + (ee ? PSA_PREP_DSC(N) : PSA_PREP_DSD(N)), dd);
      const cd dsO = ld(pr + (ee ? PSA_PREP_DSD(N) : PSA_PREP_DSC(N)), 2 * N - 2 - dd);
      k0[c] = cd{vself * dsS.re - veh * dsO.re, vself * dsS.im - veh * dsO.im};
    }
  }
#pragma unroll
  for (int c = 0; c < 4; ++c) {
    int d, sh;
    col_info(cb, 16 * c + r16, d, sh);
    const int ad = d < 0 ? -d : d;
    const int lo = max(0, -(sh + d)), hi = min(N - sh, N - sh - d);  // valid rows a'
#pragma unroll
    for (int i = 0; i < RB; ++i)
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        const int ar = wrow0 + 16 * i + 4 * g4 + v;
        if (ad > N - 1 || ar < lo || ar >= hi) continue;
        const int a = ar + sh;
        const long ij = (long)a * N + a + d;  // H[k1][k2] stored at [k2 = a][k1 = a + d]
        const double re = (double)aR[i][c][v], im = (double)aI[i][c][v];
        if (which == 0) {
          const cd m = (a & 1) ? k1c[c] : k0[c];
          st_h(g, g.Heh, ow, ij, cd{m.re + re, m.im + im});
        } else if (d == 0) {
          const double e0 = ee ? g.Ee[a] + g.gap : g.Eh[a];
          st(g.hdiag + (long)w * 4 * N, (ee ? 0 : N) + a, cd{e0 - re, -im});
        } else {
          st_h(g, ee ? g.Heo : g.Hho, ow, ij, lf ? cd{-re + k0[c].re, -im + k0[c].im} : cd{0.0, 0.0});
        }
      }
  }
}

template <typename T>
__global__ __launch_bounds__(threads_per_block<T>(), 2) void k_calch(psa_group g, int w0) {
  constexpr int RB = Cfg<T>::RB, ROWS = rows_per_block<T>(), E = 16 / sizeof(T);
  constexpr int THREADS = threads_per_block<T>();
  constexpr int PER = KC * COLS / THREADS;  // Y elements per thread per chunk (consecutive K)
  typedef typename Cfg<T>::unit unit_t;
  __shared__ T vt[NMAX];                              // V(|m' - a'|), V(0) -> 0
  __shared__ unit_t bs[2][2][KC / E * COLS];          // [buffer][re, im]

  const int N = g.N;
  const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
  // grid (3 Nw, row groups, column blocks): the column block varies slowest,
  // so the long column blocks (cb = 0 spans all N rows and K values, the
  // last ones 32) are dispatched first and the short ones fill the tail
  const int which = blockIdx.x % 3, w = w0 + blockIdx.x / 3;
  const int cb = blockIdx.z, row0 = ROWS * blockIdx.y;
  const int Lmax = N - 32 * cb;  // rows a' and K m' below this
  if (row0 >= Lmax) return;      // block-uniform, before any barrier

  const double *vc = g.prep + (long)w * PSA_PREP_STRIDE(N) + PSA_PREP_VC(N) + (long)which * N;
  for (int i = tid; i < N; i += THREADS) vt[i] = (T)(i == 0 ? 0.0 : vc[i]);

  // this thread's Y column and K rows PER (tid >> 6) + e of each chunk
  const long o = mat_off(g, w);
  const double *X = (which == 0 ? g.pB : (which == 1 ? g.CB : g.DB)) + o;
  const int xB = tid & (COLS - 1);
  int dB, sB;
  col_info(cb, xB, dB, sB);
  // K values m' in [loB, hiB) have m = m' + s and m + d in [0, N)
  const int loB = max(0, -(sB + dB)), hiB = min(N - sB, N - sB - dB);
  const int kq = PER * (tid >> 6);
  const int nch = (Lmax + KC - 1) / KC;

  cd rv[PER];
  auto issue = [&](int ch) {
#pragma unroll
    for (int e = 0; e < PER; ++e) {
      const int m = ch * KC + kq + e;
      const bool ok = m >= loB && m < hiB;
      const int r = ok ? m + sB : 0, c = ok ? m + sB + dB : 0;
      rv[e] = ld(X, (long)r * N + c);
    }
  };
  auto commit = [&](int ch, int buf) {
#pragma unroll
    for (int u = 0; u < PER / E; ++u) {
      unit_t vr, vi;
#pragma unroll
      for (int s = 0; s < E; ++s) {
        const int e = u * E + s;
        const int m = ch * KC + kq + e;
        const bool ok = m >= loB && m < hiB;
        vr[s] = (T)(ok ? rv[e].re : 0.0);
        vi[s] = (T)(ok ? rv[e].im : 0.0);
      }
      const int idx = (kq / E + u) * COLS + xB;
      bs[buf][0][idx] = vr;
      bs[buf][1][idx] = vi;
    }
  };

  using acc_t = typename Mfma<T>::acc;
  acc_t aR[RB][4], aI[RB][4];
#pragma unroll
  for (int i = 0; i < RB; ++i)
#pragma unroll
    for (int c = 0; c < 4; ++c) {
      aR[i][c] = acc_t{0, 0, 0, 0};
      aI[i][c] = acc_t{0, 0, 0, 0};
    }
  const int g4 = lane >> 4, r16 = lane & 15;
  const int wrow0 = row0 + 16 * RB * wv;
  const bool wact = wrow0 < Lmax;  // wave-uniform

  issue(0);
  commit(0, 0);
  __syncthreads();
  for (int ch = 0; ch < nch; ++ch) {
    const bool more = ch + 1 < nch;
    if (more) issue(ch + 1);
    if (wact) {
      const unit_t *Br = bs[ch & 1][0], *Bi = bs[ch & 1][1];
#pragma unroll
      for (int sl = 0; sl < KC / 16; ++sl) {
        const int kl = 16 * sl + 4 * g4;  // this lane's first K value in the chunk
#pragma unroll
        for (int u = 0; u < 4 / E; ++u) {
          unit_t br[4], bi[4];
#pragma unroll
          for (int c = 0; c < 4; ++c) {
            br[c] = Br[(kl / E + u) * COLS + 16 * c + r16];
            bi[c] = Bi[(kl / E + u) * COLS + 16 * c + r16];
          }
#pragma unroll
          for (int s = 0; s < E; ++s) {
            const int m = ch * KC + kl + u * E + s;
            T a[RB];
#pragma unroll
            for (int i = 0; i < RB; ++i) {
              const int dlt = m - (wrow0 + 16 * i + r16);
              a[i] = vt[min(dlt < 0 ? -dlt : dlt, N - 1)];
            }
#pragma unroll
            for (int i = 0; i < RB; ++i)
#pragma unroll
              for (int c = 0; c < 4; ++c) {
                aR[i][c] = Mfma<T>::mma(a[i], br[c][s], aR[i][c]);
                aI[i][c] = Mfma<T>::mma(a[i], bi[c][s], aI[i][c]);
              }
          }
        }
      }
    }
    if (more) commit(ch + 1, (ch + 1) & 1);
    __syncthreads();
  }
  if (!wact) return;

  calch_epilogue(g, w, which, cb, wrow0, g4, r16, aR, aI);
}

// ---------------------------------------------------------------------------
// Split-BF16 exchange sums (PSA_PREC_BF16X3): the same shifted-Toeplitz
// products on v_mfma_f32_16x16x32_bf16 with both operands split as
// hi = bf16(f), lo = bf16(f - hi) of their FP32 value f (three products per
// real product, FP32 accumulation; relative error of a product ~2^-16).
//
// Block = 8 waves of 32 rows x the 64 columns of a column block.  The
// Toeplitz operand A[a'][m'] = V(|m' - a'|) needs, per lane and MFMA, the
// 8 consecutive entries T[i0 .. i0 + 7] of the row T[i] = V(|i - OFF|)
// (OFF = TOFF + N - 1, V(0) and |i - OFF| >= N zero): T is kept split in
// LDS as eight copies shifted by 0..7 entries, so that copy i0 % 8 holds
// them as one aligned 16-byte group (one ds_read_b128 per plane; the copy
// stride of 1104 entries puts the 16 lanes of a row group on distinct
// banks).  Y[m'][col] (complex) is split into four planes (Re hi, Re lo,
// Im hi, Im lo) per 32-K chunk as [column][32 K] with the 16-byte group of
// 8 K at (K >> 3) ^ ((column >> 1) & 3), double-buffered, the next chunk's
// loads in flight during the current chunk's products.  The epilogue is
// the FP32 kernel's (same output lane layout).
// ---------------------------------------------------------------------------
namespace bfx {
constexpr int WAVES = 8, THREADS = 64 * WAVES, ROWS = 32 * WAVES, KC = 32;
constexpr int TOFF = 32;  // zero entries in front of T (rows up to 31 past the last valid row)
constexpr int TL = 1104;  // entries per shifted copy: >= 2 NMAX + 72 and TL / 2 = 40 (mod 64)
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef __bf16 bf16x4 __attribute__((ext_vector_type(4)));
typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));
typedef uint32_t u32x2 __attribute__((ext_vector_type(2)));
}  // namespace bfx

__global__ __launch_bounds__(bfx::THREADS, 4) void k_calch_bf(psa_group g, int w0) {
  constexpr int THREADS = bfx::THREADS, ROWS = bfx::ROWS, KCB = bfx::KC, TL = bfx::TL;
  using bfx::bf16x4;
  using bfx::bf16x8;
  using bfx::u32x2;
  using bfx::u32x4;
  __shared__ __attribute__((aligned(16))) uint16_t vtab[2][8][TL];        // [hi, lo][shift c][j] = T[j + c]
  __shared__ __attribute__((aligned(16))) uint16_t ys[2][4][COLS * KCB];  // [buffer][plane][column][K]

  const int N = g.N;
  const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
  const int which = blockIdx.x % 3, w = w0 + blockIdx.x / 3;
  const int cb = blockIdx.z, row0 = ROWS * blockIdx.y;
  const int Lmax = N - 32 * cb;  // rows a' and K m' below this
  if (row0 >= Lmax) return;      // block-uniform, before any barrier
  const int nch = (Lmax + KCB - 1) / KCB;

  const double *vc = g.prep + (long)w * PSA_PREP_STRIDE(N) + PSA_PREP_VC(N) + (long)which * N;
  const int OFF = bfx::TOFF + N - 1;
  for (int idx = tid; idx < 8 * TL; idx += THREADS) {
    const int c = idx / TL, j = idx % TL, dl = j + c - OFF, ad = dl < 0 ? -dl : dl;
    const float v = (float)((ad > 0 && ad < N) ? vc[ad] : 0.0);
    const __bf16 h = (__bf16)v;
    vtab[0][c][j] = __builtin_bit_cast(uint16_t, h);
    vtab[1][c][j] = __builtin_bit_cast(uint16_t, (__bf16)(v - (float)h));
  }

  // Y builder: column xB, K values 4 kq .. 4 kq + 3 of each chunk
  const long o = mat_off(g, w);
  const double *X = (which == 0 ? g.pB : (which == 1 ? g.CB : g.DB)) + o;
  const int xB = tid & (COLS - 1), kq = __builtin_amdgcn_readfirstlane(tid >> 6);  // wave-uniform
  int dB, sB;
  col_info(cb, xB, dB, sB);
  // K values m' in [loB, hiB) have m = m' + s and m + d in [0, N)
  const int loB = max(0, -(sB + dB)), hiB = min(N - sB, N - sB - dB);
  cd rv[4];
  auto issue = [&](int ch) {
#pragma unroll
    for (int e = 0; e < 4; ++e) {
      const int m = ch * KCB + 4 * kq + e;
      const bool ok = m >= loB && m < hiB;
      const int r = ok ? m + sB : 0, c = ok ? m + sB + dB : 0;
      rv[e] = ld(X, (long)r * N + c);
      if (!ok) rv[e] = cd{0.0, 0.0};
    }
  };
  const int yoff = xB * KCB + (((kq >> 1) ^ ((xB >> 1) & 3)) << 3) + 4 * (kq & 1);
  auto commit = [&](int buf) {
    bf16x4 rh, rl, ih, il;
#pragma unroll
    for (int e = 0; e < 4; ++e) {
      const float fr = (float)rv[e].re, fi = (float)rv[e].im;
      rh[e] = (__bf16)fr;
      rl[e] = (__bf16)(fr - (float)rh[e]);
      ih[e] = (__bf16)fi;
      il[e] = (__bf16)(fi - (float)ih[e]);
    }
    *reinterpret_cast<u32x2 *>(&ys[buf][0][yoff]) = __builtin_bit_cast(u32x2, rh);
    *reinterpret_cast<u32x2 *>(&ys[buf][1][yoff]) = __builtin_bit_cast(u32x2, rl);
    *reinterpret_cast<u32x2 *>(&ys[buf][2][yoff]) = __builtin_bit_cast(u32x2, ih);
    *reinterpret_cast<u32x2 *>(&ys[buf][3][yoff]) = __builtin_bit_cast(u32x2, il);
  };

  fv4 aR[2][4], aI[2][4];
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int c = 0; c < 4; ++c) {
      aR[i][c] = fv4{0.f, 0.f, 0.f, 0.f};
      aI[i][c] = fv4{0.f, 0.f, 0.f, 0.f};
    }
  const int g4 = lane >> 4, r16 = lane & 15;
  const int wrow0 = row0 + 32 * wv;
  const bool wact = wrow0 < Lmax;  // wave-uniform

  issue(0);
  commit(0);
  __syncthreads();
  for (int ch = 0; ch < nch; ++ch) {
    const bool more = ch + 1 < nch;
    if (more) issue(ch + 1);
    if (wact) {
      bf16x8 ah[2], al[2];
#pragma unroll
      for (int i = 0; i < 2; ++i) {
        const int i0 = ch * KCB + 8 * g4 - (wrow0 + 16 * i + r16) + OFF;  // >= 1
        const int c = i0 & 7, j = i0 - c;
        ah[i] = __builtin_bit_cast(bf16x8, *reinterpret_cast<const u32x4 *>(&vtab[0][c][j]));
        al[i] = __builtin_bit_cast(bf16x8, *reinterpret_cast<const u32x4 *>(&vtab[1][c][j]));
      }
      const int buf = ch & 1;
#pragma unroll
      for (int c = 0; c < 4; ++c) {
        const int col = 16 * c + r16;
        const int off = col * KCB + ((g4 ^ ((col >> 1) & 3)) << 3);
        const bf16x8 rh = __builtin_bit_cast(bf16x8, *reinterpret_cast<const u32x4 *>(&ys[buf][0][off]));
        const bf16x8 rl = __builtin_bit_cast(bf16x8, *reinterpret_cast<const u32x4 *>(&ys[buf][1][off]));
        const bf16x8 ih = __builtin_bit_cast(bf16x8, *reinterpret_cast<const u32x4 *>(&ys[buf][2][off]));
        const bf16x8 il = __builtin_bit_cast(bf16x8, *reinterpret_cast<const u32x4 *>(&ys[buf][3][off]));
#pragma unroll
        for (int i = 0; i < 2; ++i) {
          aR[i][c] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah[i], rh, aR[i][c], 0, 0, 0);
          aI[i][c] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah[i], ih, aI[i][c], 0, 0, 0);
        }
#pragma unroll
        for (int i = 0; i < 2; ++i) {
          aR[i][c] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah[i], rl, aR[i][c], 0, 0, 0);
          aI[i][c] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah[i], il, aI[i][c], 0, 0, 0);
        }
#pragma unroll
        for (int i = 0; i < 2; ++i) {
          aR[i][c] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(al[i], rh, aR[i][c], 0, 0, 0);
          aI[i][c] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(al[i], ih, aI[i][c], 0, 0, 0);
        }
        __builtin_amdgcn_sched_barrier(0);
      }
    }
    if (more) commit((ch + 1) & 1);
    __syncthreads();
  }
  if (!wact) return;
  calch_epilogue(g, w, which, cb, wrow0, g4, r16, aR, aI);
}

}  // namespace chx
}  // namespace psa
