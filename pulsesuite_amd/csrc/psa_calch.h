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

  // epilogue: the constants of all four column blocks first (their loads in
  // flight together), then the stores
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
        const int ar = wrow0 + 16 * i + Mfma<T>::row(g4, v);
        if (ad > N - 1 || ar < lo || ar >= hi) continue;
        const int a = ar + sh;
        const long ij = (long)a * N + a + d;  // H[k1][k2] stored at [k2 = a][k1 = a + d]
        const double re = (double)aR[i][c][v], im = (double)aI[i][c][v];
        if (which == 0) {
          const cd m = (a & 1) ? k1c[c] : k0[c];
          st(g.Heh + ow, ij, cd{m.re + re, m.im + im});
        } else if (d == 0) {
          const double e0 = ee ? g.Ee[a] + g.gap : g.Eh[a];
          st(g.hdiag + (long)w * 4 * N, (ee ? 0 : N) + a, cd{e0 - re, -im});
        } else {
          double *Hout = (ee ? g.Heo : g.Hho) + ow;
          st(Hout, ij, lf ? cd{-re + k0[c].re, -im + k0[c].im} : cd{0.0, 0.0});
        }
      }
  }
}

}  // namespace chx
}  // namespace psa
