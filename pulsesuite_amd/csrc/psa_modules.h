// Module-API kernels: the coulomb.py / phonons.py functions the SBE step
// calls, on the caller's own dense arrays (any k-grid, any VC), for the
// drop-in modules pulsesuite_amd.coulomb and pulsesuite_amd.phonons.  The
// fused step kernels work on the SBE grid's Toeplitz V(|q|) and compressed
// many-body lists instead; these take the reference's layouts as they are:
//   VC [N][N][3] (channel fastest), E1D [N][N], k3 [N][N][N] (1-based or 0),
//   Ceh/Cee/Chh [(N+1)^3] filled at 1-based positions.
// Every kernel sums in FP64.  Citations are relative to src/pulsesuite/PSTD3D/.
#pragma once

#include "psa_dev.h"

namespace psa {
namespace mod {

constexpr int RED_T = 256;

// Fixed-order block sum of two values (tree over the 256 threads).
__device__ __forceinline__ void block_sum2(double &a, double &b, double (*sh)[RED_T]) {
  const int t = threadIdx.x;
  sh[0][t] = a;
  sh[1][t] = b;
  __syncthreads();
  for (int o = RED_T / 2; o > 0; o >>= 1) {
    if (t < o) {
      sh[0][t] += sh[0][t + o];
      sh[1][t] += sh[1][t + o];
    }
    __syncthreads();
  }
  a = sh[0][0];
  b = sh[1][0];
}

// GaussDelta (coulomb.py:245-268)
__device__ __forceinline__ double gauss_delta(double a, double b) {
  if (fabs(b) < 1e-300) return 0.0;
  const double x = a / b;
  return 1.0 / (sqrt(PI) * b) * exp(-(x * x));
}

// MakeK3 (coulomb.py:271-299) and CalcMBArrays (coulomb.py:367-499).
// Block (k1, k2), threads over k4.  The zero planes (index 0) of the
// (N+1)^3 tables are left as the caller allocated them (zero).
__global__ void k_mb_arrays(int N, const double *Ee, const double *Eh, double ge, double gh, double hbar,
                            int lorentz, int32_t *k3, double *Ceh, double *Cee, double *Chh) {
  const int k1 = blockIdx.x, k2 = blockIdx.y;
  const double geh = (ge + gh) / 2.0;
  const double hge2 = (hbar * ge) * (hbar * ge), hgh2 = (hbar * gh) * (hbar * gh);
  const double hgeh2 = (hbar * geh) * (hbar * geh);
  const long M = N + 1;
  for (int k4 = threadIdx.x; k4 < N; k4 += blockDim.x) {
    // k3[k1][k2][k4] = (k1 + 1) + (k2 + 1) - (k4 + 1), 0 outside [1, N]
    int v = k1 + k2 - k4 + 1;
    if (v < 1 || v > N) v = 0;
    k3[((long)k1 * N + k2) * N + k4] = v;
    const long o = ((long)(k1 + 1) * M + (k2 + 1)) * M + (k4 + 1);
    const double u14 = (k1 == k4) ? 0.0 : 1.0;
    // Ceh: k30 = k3[k4][k2][k1]
    const int ceh = k4 + k2 - k1 + 1;
    if (ceh >= 1 && ceh <= N) {
      const int k30 = ceh - 1;
      const double u23 = (k2 == k30) ? 0.0 : 1.0;
      const double Ed = Ee[k1] + Eh[k2] - Eh[k30] - Ee[k4];
      Ceh[o] = lorentz ? 2.0 * geh * u14 * u23 / (Ed * Ed + hgeh2)
                       : 2.0 * PI / hbar * u14 * u23 * gauss_delta(Ed, hbar * geh);
    }
    // Cee, Chh: k30 = k3[k1][k2][k4]
    if (v >= 1) {
      const int k30 = v - 1;
      const double u23 = (k2 == k30) ? 0.0 : 1.0;
      const double Ede = Ee[k1] + Ee[k2] - Ee[k30] - Ee[k4];
      const double Edh = Eh[k1] + Eh[k2] - Eh[k30] - Eh[k4];
      Cee[o] = lorentz ? 2.0 * ge * u14 * u23 / (Ede * Ede + hge2)
                       : 2.0 * PI / hbar * u14 * u23 * gauss_delta(Ede, hbar * ge);
      Chh[o] = lorentz ? 2.0 * gh * u14 * u23 / (Edh * Edh + hgh2)
                       : 2.0 * PI / hbar * u14 * u23 * gauss_delta(Edh, hbar * gh);
    }
  }
}

// CalcScreenedArrays / Eps1D (coulomb.py:1172-1230).  Block = 256 elements
// of the N x N grid; every block first forms the clipped 1-D density (a
// fixed-order sum) when the occupations are given, else uses n1d.
__global__ __launch_bounds__(RED_T) void k_screen(int N, int screened, double L, double nmax, const cd *ne,
                                                  const cd *nh, double n1d, const double *V0, const double *chi,
                                                  const double *q, double *VC, double *E1D) {
  __shared__ double sh[2][RED_T];
  double n1 = n1d;
  if (screened && ne) {
    double a = 0.0, b = 0.0;
    for (int k = threadIdx.x; k < N; k += RED_T) a += ne[k].re + nh[k].re;
    block_sum2(a, b, sh);
    n1 = fmin(a / 2.0 / L, nmax);
  }
  const long NN = (long)N * N, i = (long)blockIdx.x * RED_T + threadIdx.x;
  if (i >= NN) return;
  double e = 1.0;
  if (screened) {
    const double qe = q[i], qh = q[NN + i];
    e = 1.0 - chi[i] * 2.0 * log(fabs((qe - PI * n1) / (qe + n1)))
        - chi[NN + i] * 2.0 * log(fabs((qh - PI * n1) / (qh + n1)));
  }
  E1D[i] = e;
  if (VC) {
#pragma unroll
    for (int c = 0; c < 3; ++c) VC[3 * i + c] = screened ? V0[c * NN + i] / e : V0[c * NN + i];
  }
}

// MBCE / MBCE2 (coulomb.py:897-1012, 1283-1311) and MBCH (coulomb.py:
// 1015-1071, 1313-1326).  Block = one outer index (k for electrons, kp for
// holes), threads over the N^2 inner pairs (q1, q2); n[0] = 0 pads the
// 1-based occupations |n0|.  The sums are added to Win/Wout in place.
__global__ __launch_bounds__(RED_T) void k_mbc(int N, int hole, const cd *ne0, const cd *nh0, const double *VC,
                                               const int32_t *k3, const double *Ceh, const double *Cx,
                                               double *Win, double *Wout) {
  __shared__ double s_ne[1025], s_nh[1025];
  __shared__ double sh[2][RED_T];
  const int o = blockIdx.x + 1;  // 1-based outer index
  if (threadIdx.x == 0) {
    s_ne[0] = 0.0;
    s_nh[0] = 0.0;
  }
  for (int k = threadIdx.x; k < N; k += RED_T) {
    s_ne[k + 1] = hypot(ne0[k].re, ne0[k].im);
    s_nh[k + 1] = hypot(nh0[k].re, nh0[k].im);
  }
  __syncthreads();
  const long M = N + 1;
  auto K3 = [&](int a, int b, int c) { return k3[((long)(a - 1) * N + (b - 1)) * N + (c - 1)]; };
  auto T = [&](const double *t, int a, int b, int c) { return t[((long)(a - 1) * M + (b - 1)) * M + (c - 1)]; };
  auto V2 = [&](int a, int b, int ch) {
    const double v = VC[((long)(a - 1) * N + (b - 1)) * 3 + ch];
    return v * v;
  };
  const double *ne = s_ne, *nh = s_nh;
  double win = 0.0, wout = 0.0;
  const long NN = (long)N * N;
  for (long pq = threadIdx.x; pq < NN; pq += RED_T) {
    const int q1 = (int)(pq / N) + 1, q2 = (int)(pq % N) + 1;
    if (!hole) {
      const int k = o, kp = q1, k1 = q2;
      int x = K3(kp, k1, k);
      if (x > 0) win += V2(k, k1, 0) * (1.0 - nh[kp]) * nh[x] * ne[k1] * T(Ceh, k, kp, k1);
      x = K3(kp, k, k1);
      if (x > 0) wout += V2(k1, k, 0) * (1.0 - ne[k1]) * (1.0 - nh[kp]) * nh[x] * T(Ceh, k1, kp, k);
      const int k2 = q1, k4 = q2;
      x = K3(k, k2, k4);
      if (x > 0) win += V2(k, k4, 1) * (1.0 - ne[k2]) * ne[x] * ne[k4] * T(Cx, k, k2, k4);
      x = K3(k4, k2, k);
      if (x > 0) wout += V2(k4, k, 1) * (1.0 - ne[k4]) * (1.0 - ne[k2]) * ne[x] * T(Cx, k4, k2, k);
    } else {
      const int kp = o, k = q1, k1 = q2;
      int x = K3(kp, k1, k);
      if (x > 0) win += V2(k, k1, 0) * (1.0 - ne[k]) * nh[x] * ne[k1] * T(Ceh, k, kp, k1);
      x = K3(kp, k, k1);
      if (x > 0) wout += V2(k1, k, 0) * (1.0 - ne[k]) * (1.0 - nh[x]) * ne[k1] * T(Ceh, k, x, k1);
      const int k2p = q1, k4p = q2;
      x = K3(kp, k2p, k4p);
      if (x > 0) win += V2(kp, k4p, 2) * (1.0 - nh[k2p]) * nh[x] * nh[k4p] * T(Cx, kp, k2p, k4p);
      x = K3(k4p, k2p, kp);
      if (x > 0) wout += V2(k4p, kp, 2) * (1.0 - nh[k4p]) * (1.0 - nh[k2p]) * nh[x] * T(Cx, k4p, k2p, kp);
    }
  }
  block_sum2(win, wout, sh);
  if (threadIdx.x == 0) {
    Win[o - 1] += win;
    Wout[o - 1] += wout;
  }
}

// MBPE (phonons.py:139-202) and MBPH (phonons.py:205-268): one block per k,
// threads over i.  Vep[k][i] = VC[k][i][1] / E1D[k][i] Vscale (electrons),
// Vhp[i][k] = VC[i][k][2] / E1D[i][k] Vscale (holes); only the real part
// of the occupations enters the real rates.
__global__ __launch_bounds__(RED_T) void k_mbp(int N, int hole, const cd *n, const double *VC, const double *E1D,
                                               const double *P, const double *PT, double Vscale, double *Win,
                                               double *Wout) {
  __shared__ double sh[2][RED_T];
  const int k = blockIdx.x;
  double a = 0.0, b = 0.0;
  for (int i = threadIdx.x; i < N; i += RED_T) {
    const long ki = hole ? (long)i * N + k : (long)k * N + i;
    const double v = VC[3 * ki + (hole ? 2 : 1)] / E1D[ki] * Vscale;
    const long ik = (long)i * N + k;
    const double nr = n[i].re;
    a += v * nr * PT[ik];
    b += v * (1.0 - nr) * P[ik];
  }
  block_sum2(a, b, sh);
  if (threadIdx.x == 0) {
    Win[k] += a;
    Wout[k] += b;
  }
}

}  // namespace mod
}  // namespace psa
