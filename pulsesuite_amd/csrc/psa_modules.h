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

// Fixed-order block sum of two values (tree over the 256 threads); every
// thread gets the sums.  The closing barrier lets the next call reuse sh.
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
  __syncthreads();
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

// ---------------------------------------------------------------------------
// qwoptics.py module API (QWOptics methods) on the caller's arrays
// ---------------------------------------------------------------------------

// Fixed-order block sum of a complex value.
__device__ __forceinline__ cd block_sum_c(cd a, double (*sh)[RED_T]) {
  double re = a.re, im = a.im;
  block_sum2(re, im, sh);
  return cd{re, im};
}

// Direct DFT of nvec complex vectors of length n: block (k, v), threads
// over j.  out[v][k] = scale sum_j in[v][j] exp(sign 2 pi i (j k mod n) / n).
__global__ __launch_bounds__(RED_T) void k_dft(int n, int sign, double scale, const cd *in, cd *out) {
  __shared__ double sh[2][RED_T];
  const int k = blockIdx.x, v = blockIdx.y;
  const cd *x = in + (long)v * n;
  cd a{0.0, 0.0};
  for (int j = threadIdx.x; j < n; j += RED_T) {
    const long m = ((long)j * k) % n;
    double sn, cs;
    sincos(2.0 * PI * (double)m / (double)n, &sn, &cs);
    sn *= (double)sign;
    a.re += x[j].re * cs - x[j].im * sn;
    a.im += x[j].re * sn + x[j].im * cs;
  }
  a = block_sum_c(a, sh);
  if (threadIdx.x == 0) out[(long)v * n + k] = cd{scale * a.re, scale * a.im};
}

// rescale_1D as a linear map: out[v][r] = w[r] sum_j M[r][j] f(in[v][j]),
// f = Re (real_in) or the identity; block (r, v), threads over j.
__global__ __launch_bounds__(RED_T) void k_resample(int nout, int nin, const double *M, const double *w, int real_in,
                                                    const cd *in, cd *out) {
  __shared__ double sh[2][RED_T];
  const int r = blockIdx.x, v = blockIdx.y;
  const cd *x = in + (long)v * nin;
  const double *row = M + (long)r * nin;
  cd a{0.0, 0.0};
  for (int j = threadIdx.x; j < nin; j += RED_T) {
    a.re += row[j] * x[j].re;
    if (!real_in) a.im += row[j] * x[j].im;
  }
  a = block_sum_c(a, sh);
  if (threadIdx.x == 0) {
    const double s = w ? w[r] : 1.0;
    out[(long)v * nout + r] = cd{a.re * s, a.im * s};
  }
}

// QWPolarization3 sums (qwoptics.py:49-76): block = one r, threads over the
// (ke, kh) pairs; prod = p[kh][ke] V[kh][ke] E[ke][r] conj(E[kh][r]) in
// the reference's order, its real part summed for V = Xvc0, Yvc0, Zvc0.
__global__ __launch_bounds__(RED_T) void k_pol3(int N, int Nr, const cd *p, const cd *X, const cd *Y, const cd *Z,
                                                const cd *E, const double *W, double syw, double sz, double scale,
                                                double *out) {
  __shared__ double sh[2][RED_T];
  const int r = blockIdx.x;
  double ax = 0.0, ay = 0.0, az = 0.0;
  const long NN = (long)N * N;
  for (long i = threadIdx.x; i < NN; i += RED_T) {
    const int ke = (int)(i / N), kh = (int)(i % N);
    const long hk = (long)kh * N + ke;
    const cd e1 = E[(long)ke * Nr + r], e2 = E[(long)kh * Nr + r];
    const cd e2c{e2.re, -e2.im};
    const cd pv = p[hk];
    ax += cmul(cmul(cmul(pv, X[hk]), e1), e2c).re;
    ay += cmul(cmul(cmul(pv, Y[hk]), e1), e2c).re;
    az += cmul(cmul(cmul(pv, Z[hk]), e1), e2c).re;
  }
  double dummy = 0.0;
  block_sum2(ax, ay, sh);
  block_sum2(az, dummy, sh);
  if (threadIdx.x == 0) {
    const double wr = W[r];
    // P = sum * W (* yw, * (-1)^w) * 2 ehint / area / L, as complex values
    out[2 * r] = ax * wr * scale;
    out[2 * r + 1] = 0.0;
    out[2 * ((long)Nr + r)] = ay * wr * syw * scale;
    out[2 * ((long)Nr + r) + 1] = 0.0;
    out[2 * (2L * Nr + r)] = az * wr * sz * scale;
    out[2 * (2L * Nr + r) + 1] = 0.0;
  }
}

// QWRho5 sums (qwoptics.py:360-374): block = one r, threads over (k1, k2):
// re[r] = sum C[k1][k2] conj(E[k1][r]) E[k2][r] W[r] / (2 L), rh with conj(D).
__global__ __launch_bounds__(RED_T) void k_rho5(int N, int Nr, const cd *C, const cd *D, const cd *E, const double *W,
                                                double L, cd *out) {
  __shared__ double sh[2][RED_T];
  const int r = blockIdx.x;
  cd ae{0.0, 0.0}, ah{0.0, 0.0};
  const double f = W[r] / (2.0 * L);
  const long NN = (long)N * N;
  for (long i = threadIdx.x; i < NN; i += RED_T) {
    const int k1 = (int)(i / N), k2 = (int)(i % N);
    const cd e1 = E[(long)k1 * Nr + r], e2 = E[(long)k2 * Nr + r];
    const cd e1c{e1.re, -e1.im};
    const cd c = C[i], d = D[i];
    const cd t = cmul(e1c, e2);
    ae = cadd(ae, cmul(c, t));
    ah = cadd(ah, cmul(cd{d.re, -d.im}, t));
  }
  ae = block_sum_c(ae, sh);
  ah = block_sum_c(ah, sh);
  if (threadIdx.x == 0) {
    out[r] = cd{ae.re * f, ae.im * f};
    out[Nr + r] = cd{ah.re * f, ah.im * f};
  }
}

// QWRho5 tail (qwoptics.py:1062-1068), one block: subtract the mean of the
// real parts of the first and last 10 points, then scale each density to
// sum |n| + 1e-50 over (sum |x| dr + 1e-100).
__global__ __launch_bounds__(RED_T) void k_rho5_finish(int N, int Nr, double dr, const cd *ne, const cd *nh, cd *x) {
  __shared__ double sh[2][RED_T];
  const int t = threadIdx.x;
  double tn = 0.0, th = 0.0;
  for (int k = t; k < N; k += RED_T) {
    tn += hypot(ne[k].re, ne[k].im);
    th += hypot(nh[k].re, nh[k].im);
  }
  block_sum2(tn, th, sh);
  const double NeT = tn + 1e-50, NhT = th + 1e-50;
  double me = 0.0, mh = 0.0;
  for (int k = t; k < Nr; k += RED_T)
    if (k < 10 || k >= Nr - 10) {
      me += x[k].re;
      mh += x[Nr + k].re;
    }
  block_sum2(me, mh, sh);
  me /= 20.0;
  mh /= 20.0;
  double se = 0.0, shh = 0.0;
  for (int k = t; k < Nr; k += RED_T) {
    x[k].re -= me;
    x[Nr + k].re -= mh;
    se += hypot(x[k].re, x[k].im);
    shh += hypot(x[Nr + k].re, x[Nr + k].im);
  }
  block_sum2(se, shh, sh);
  const double fe = NeT / (se * dr + 1e-100), fh = NhT / (shh * dr + 1e-100);
  for (int k = t; k < Nr; k += RED_T) {
    x[k] = cd{x[k].re * fe, x[k].im * fe};
    x[Nr + k] = cd{x[Nr + k].re * fh, x[Nr + k].im * fh};
  }
}

// QW2Prop's Plasmonics renormalisation (qwoptics.py:785-790, 801-805), one
// block: re = |re|, rh = |rh|; mode 0 sets total = (sum re dx + sum rh dx +
// 1e-100) / 2; then re *= total / (sum re dx + 1e-100), rh likewise.
__global__ __launch_bounds__(RED_T) void k_plasmonics(int n, double dx, int mode, cd *re, cd *rh, double *total) {
  __shared__ double sh[2][RED_T];
  const int t = threadIdx.x;
  double se = 0.0, sr = 0.0;
  for (int k = t; k < n; k += RED_T) {
    const double a = hypot(re[k].re, re[k].im), b = hypot(rh[k].re, rh[k].im);
    re[k] = cd{a, 0.0};
    rh[k] = cd{b, 0.0};
    se += a;
    sr += b;
  }
  block_sum2(se, sr, sh);
  const double tot = mode == 0 ? (se * dx + sr * dx + 1e-100) / 2.0 : *total;
  if (mode == 0 && t == 0) *total = tot;
  const double fe = tot / (se * dx + 1e-100), fh = tot / (sr * dx + 1e-100);
  for (int k = t; k < n; k += RED_T) {
    re[k].re *= fe;
    rh[k].re *= fh;
  }
}

// ---------------------------------------------------------------------------
// dephasing.py / emission.py module API
// ---------------------------------------------------------------------------

// CalcGammaE / CalcGammaH (dephasing.py:28-103, 541-625): block = one k,
// threads over q.  t0..t3 = k_p_q, k_m_q, k1p_m_q, k1p (electrons) or
// k_p_q, k_m_q, k1_m_q, k1 (holes), [N][N] row k; an index outside [0, N)
// drops the term as the reference's bounds test does.  V^2(q) =
// VC[iq[q]][0][c]^2 (Vxx2, dephasing.py:362-378).
__global__ __launch_bounds__(RED_T) void k_gamma(int N, int hole, const cd *ne0, const cd *nh0, const double *VC,
                                                 const int32_t *iq, const int32_t *t0, const int32_t *t1,
                                                 const int32_t *t2, const int32_t *t3, const double *xe,
                                                 const double *xh, double pih, double *out) {
  __shared__ double sh[2][RED_T];
  const int k = blockIdx.x;
  auto ok = [&](int i) { return i >= 0 && i < N; };
  auto n_e = [&](int i) { return ne0[i].re; };
  auto n_h = [&](int i) { return nh0[i].re; };
  double acc = 0.0, dummy = 0.0;
  for (int q = threadIdx.x; q < N; q += RED_T) {
    const long r = (long)k * N + q;
    const long vq = (long)iq[q] * N * 3;
    const double veh = VC[vq] * VC[vq];
    const int a = t0[r], d = t1[r], b = t2[r], c = t3[r];
    if (!hole) {
      const double vee = VC[vq + 1] * VC[vq + 1];
      if (ok(a)) acc += pih * vee * n_e(a) * (1.0 - n_e(a)) * fabs(xe[q]);
      if (ok(b) && ok(c) && ok(d))
        acc += pih * veh * (n_h(b) * (1.0 - n_h(c)) * n_e(d) + (1.0 - n_h(b)) * n_h(c) * (1.0 - n_e(d))) * fabs(xh[q]);
    } else {
      const double vhh = VC[vq + 2] * VC[vq + 2];
      if (ok(a)) acc += pih * vhh * n_h(a) * (1.0 - n_h(a)) * fabs(xh[q]);
      if (ok(b) && ok(c) && ok(d))
        acc += pih * veh * (n_e(b) * (1.0 - n_e(c)) * n_h(d) + (1.0 - n_e(b)) * n_e(c) * (1.0 - n_h(d))) * fabs(xe[q]);
    }
  }
  block_sum2(acc, dummy, sh);
  if (threadIdx.x == 0) out[k] = acc;
}

// CalcOffDiagDeph_E2 / _H2 (dephasing.py:226-317, 381-414): block (q + N,
// k, kind), threads over p; D[kind][q + N][k] for q in [-N, N].  Kind 0
// (De): same-carrier electrons (V_ee, width hgS = hbar gee) and electron-
// hole (V_eh, hbar geh); kind 1 (Dh) the mirror with hbar ghh.  Lrtz(a, b)
// = (b / pi) / (a^2 + b^2) (dephasing.py:220-224); V^2(q) = VC[|q|][0][c]^2
// for 0 < |q| < N, else 0.
__global__ __launch_bounds__(RED_T) void k_od2_tables(int N, const cd *ne0, const cd *nh0, const double *VC,
                                                      const double *Ee, const double *Eh, double hgee, double hghh,
                                                      double hgeh, double hbar, double *D) {
  __shared__ double sh[2][RED_T];
  const int q = (int)blockIdx.x - N, k = blockIdx.y, kind = blockIdx.z;
  const int aq = q < 0 ? -q : q;
  auto vsq = [&](int c) {
    if (aq < 1 || aq > N - 1) return 0.0;
    const double v = VC[(long)aq * N * 3 + c];
    return v * v;
  };
  const double vS = vsq(kind == 0 ? 1 : 2), vO = vsq(0);
  const double *ES = kind == 0 ? Ee : Eh, *EO = kind == 0 ? Eh : Ee;
  const cd *nS = kind == 0 ? ne0 : nh0, *nO = kind == 0 ? nh0 : ne0;
  const double hgS = kind == 0 ? hgee : hghh;
  auto lrtz = [](double a, double b) { return (b / PI) / (a * a + b * b); };
  auto mag = [](cd z) { return hypot(z.re, z.im); };
  double acc = 0.0, dummy = 0.0;
  const bool kq = k + q >= 0 && k + q < N;
  if (kq) {
    const double nk = mag(nS[k]);
    for (int p = threadIdx.x; p < N; p += RED_T) {
      const int pm = p - q, pp = p + q;
      if (pm >= 1 && pm < N) {  // q <= p - 1 and p - q < N
        const double Ed = ES[k + q] + ES[pm] - ES[p] - ES[k];
        const double a = mag(nS[pm]), b = mag(nS[p]);
        acc += vS * lrtz(Ed, hgS) * (a * (1.0 - b) * (1.0 - nk) + (1.0 - a) * b * nk);
      }
      if (pp >= 0 && pp < N) {
        const double Ed = ES[k + q] + EO[pp] - EO[p] - ES[k];
        const double a = mag(nO[pp]), b = mag(nO[p]);
        acc += vO * lrtz(Ed, hgeh) * (a * (1.0 - b) * (1.0 - nk) + (1.0 - a) * b * nk);
      }
    }
  }
  block_sum2(acc, dummy, sh);
  if (threadIdx.x == 0) D[((long)kind * (2 * N + 1) + (q + N)) * N + k] = acc * PI / hbar;
}

// OffDiagDephasing2's contraction (dephasing.py:320-356): x[kh][ke] =
// i hbar (sum_{q != 0} Dh[q + N][kh] p[kh + q][ke] + sum_{q != 0} De[q + N][ke]
// p[kh][ke + q]); one thread per (kh, ke), lanes along ke.
__global__ __launch_bounds__(RED_T) void k_od2_apply(int N, const double *D, const cd *p, double hbar, cd *x) {
  const long i = (long)blockIdx.x * RED_T + threadIdx.x;
  if (i >= (long)N * N) return;
  const int kh = (int)(i / N), ke = (int)(i % N);
  const double *De = D, *Dh = D + (long)(2 * N + 1) * N;
  cd s{0.0, 0.0};
  for (int q = -kh; q < N - kh; ++q) {
    if (q == 0) continue;
    const double d = Dh[(long)(q + N) * N + kh];
    const cd v = p[(long)(kh + q) * N + ke];
    s = cd{s.re + d * v.re, s.im + d * v.im};
  }
  for (int q = -ke; q < N - ke; ++q) {
    if (q == 0) continue;
    const double d = De[(long)(q + N) * N + ke];
    const cd v = p[(long)kh * N + ke + q];
    s = cd{s.re + d * v.re, s.im + d * v.im};
  }
  x[i] = cd{-s.im * hbar, s.re * hbar};
}

// SpontEmission (emission.py:94-130) with Ec (133-171) and SpontIntegral
// (174-198): block = one k.  Ek = gap + Ee[k] + Eh[k] + Ec[k] on the real
// occupations; Rsp[k] = dhw sum_n (H_n + Ek) rho0(H_n + Ek) sq_n with
// rho0(h) = h^2 / rden (rden = c0^3 pi^2 hbar^3).  ec (optional) gets Ec.
__global__ __launch_bounds__(RED_T) void k_spont(int N, int Nhw, const cd *ne0, const cd *nh0, const double *VC,
                                                 const double *Ee, const double *Eh, double gap, const double *hom,
                                                 const double *sq, double dhw, double rden, double *Rsp,
                                                 double *ec) {
  __shared__ double sh[2][RED_T];
  const int k = blockIdx.x;
  auto V = [&](int a, int b, int c) { return VC[((long)a * N + b) * 3 + c]; };
  const double vkk_eh = V(k, k, 0), vkk_ee = V(k, k, 1), vkk_hh = V(k, k, 2);
  double s1 = 0.0, s2 = 0.0;
  for (int i = threadIdx.x; i < N; i += RED_T) {
    const double a = ne0[i].re, b = nh0[i].re, id = (i == k) ? 0.0 : 1.0;
    s1 += a * (vkk_ee - V(i, k, 1)) + b * (vkk_hh - V(i, k, 2));
    s2 += a * vkk_eh * id - b * vkk_eh * id;
  }
  block_sum2(s1, s2, sh);
  const double Ec = s1 + s2 - vkk_eh;
  const double Ek = gap + Ee[k] + Eh[k] + Ec;
  double acc = 0.0, dummy = 0.0;
  for (int n = threadIdx.x; n < Nhw; n += RED_T) {
    const double h = hom[n] + Ek;
    acc += h * (h * h / rden) * sq[n];
  }
  block_sum2(acc, dummy, sh);
  if (threadIdx.x == 0) {
    Rsp[k] = acc * dhw;
    if (ec) ec[k] = Ec;
  }
}


// ---------------------------------------------------------------------------
// Module functions off the SBE step (the coulomb, emission and qwoptics
// drop-ins).  Sequential sums run in the reference's order, and contraction
// into FMAs is off, so these give the reference's pure-Python results.
// ---------------------------------------------------------------------------

// CalcMVeh (coulomb.py:1234-1241, _CalcMVeh_core 786-810): thread per
// (k, kp, f), t = (k N + kp) Nf + f;
//   MVeh[k][kp][f] = sum_q p[q][qp][f] Veh[k][q] U[k+1][q+1] U[kp+1][qp+1],
// qp + 1 = k3[kp][q][k] (terms with 0 skipped), q ascending.
__global__ __launch_bounds__(RED_T) void k_mveh(int N, int Nf, const cd *p, const double *VC, const int32_t *k3,
                                                const double *U, cd *M) {
#pragma clang fp contract(off)
  const long t = (long)blockIdx.x * RED_T + threadIdx.x;
  if (t >= (long)N * N * Nf) return;
  const int f = (int)(t % Nf);
  const long r = t / Nf;
  const int kp = (int)(r % N), k = (int)(r / N);
  const long UN = N + 1;
  cd acc{0.0, 0.0};
  for (int q = 0; q < N; ++q) {
    const int qp1 = k3[((long)kp * N + q) * N + k];
    if (qp1 > 0 && qp1 <= N) {
      const cd v = p[((long)q * N + (qp1 - 1)) * Nf + f];
      const double veh = VC[((long)k * N + q) * 3], u1 = U[(k + 1) * UN + q + 1], u2 = U[(kp + 1) * UN + qp1];
      acc.re = acc.re + ((v.re * veh) * u1) * u2;
      acc.im = acc.im + ((v.im * veh) * u1) * u2;
    }
  }
  M[t] = acc;
}

// BGRenorm / EeRenorm / EhRenorm (coulomb.py:1248-1280; _BGRenorm_core
// 813-831, _EeRenorm_core 834-863, _EhRenorm_core 866-897): thread per
// (k, kp), t = k N + kp, sums over i ascending.
//   which 0: BGR = -sum_i nh_i Vhh[i][kp] U[i+1][kp+1] - sum_i ne_i Vee[i][k] U[i+1][k+1]
//   which 1, 2 (n = ne, V = Vee / n = nh, V = Vhh): with a = sum_i n_i V[kp][kp],
//   b = sum_i n_i V[i][kp] U[i+1][kp+1], c = sum_i n_i V[k][k],
//   d = sum_i n_i V[i][k] U[i+1][k+1]: BGR = 2a - b + 2c - d - 2a - c, left to right.
__global__ __launch_bounds__(RED_T) void k_renorm(int N, int which, const cd *ne, const cd *nh, const double *VC,
                                                  const double *U, cd *B) {
#pragma clang fp contract(off)
  const long t = (long)blockIdx.x * RED_T + threadIdx.x;
  if (t >= (long)N * N) return;
  const int kp = (int)(t % N), k = (int)(t / N);
  const long UN = N + 1;
  auto V = [&](int a, int b, int c) { return VC[((long)a * N + b) * 3 + c]; };
  if (which == 0) {
    cd sh{0.0, 0.0}, se{0.0, 0.0};
    for (int i = 0; i < N; ++i) {
      const double wh = V(i, kp, 2), we = V(i, k, 1);
      const double uh = U[(i + 1) * UN + kp + 1], ue = U[(i + 1) * UN + k + 1];
      sh.re = sh.re + (nh[i].re * wh) * uh;
      sh.im = sh.im + (nh[i].im * wh) * uh;
      se.re = se.re + (ne[i].re * we) * ue;
      se.im = se.im + (ne[i].im * we) * ue;
    }
    B[t] = cd{-sh.re - se.re, -sh.im - se.im};
    return;
  }
  const cd *n = (which == 1) ? ne : nh;
  const int c = (which == 1) ? 1 : 2;
  const double vkpkp = V(kp, kp, c), vkk = V(k, k, c);
  cd a{0.0, 0.0}, b{0.0, 0.0}, cc{0.0, 0.0}, d{0.0, 0.0};
  for (int i = 0; i < N; ++i) {
    const cd x = n[i];
    const double wkp = V(i, kp, c), wk = V(i, k, c);
    const double ukp = U[(i + 1) * UN + kp + 1], uk = U[(i + 1) * UN + k + 1];
    a.re = a.re + x.re * vkpkp;
    a.im = a.im + x.im * vkpkp;
    b.re = b.re + (x.re * wkp) * ukp;
    b.im = b.im + (x.im * wkp) * ukp;
    cc.re = cc.re + x.re * vkk;
    cc.im = cc.im + x.im * vkk;
    d.re = d.re + (x.re * wk) * uk;
    d.im = d.im + (x.im * wk) * uk;
  }
  cd r;
  r.re = ((((2.0 * a.re - b.re) + 2.0 * cc.re) - d.re) - 2.0 * a.re) - cc.re;
  r.im = ((((2.0 * a.im - b.im) + 2.0 * cc.im) - d.im) - 2.0 * a.im) - cc.im;
  B[t] = r;
}

// complex product as (ar br - ai bi, ar bi + ai br), no contraction
__device__ __forceinline__ cd cmul_rn(cd a, cd b) {
#pragma clang fp contract(off)
  return cd{a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re};
}

// GetVn1n2 (qwoptics.py:1149-1186, _GetVn1n2_jit 522-542): thread per
// (k1, k2), t = k1 N + k2; Hvv = -Hhh^T, Hvc = Hcv^H, rvc = conj(rcv),
// m = -i / hbar = (-0, -1/hbar):
//   Vcv = m (rcv[k1] Hvv[k1][k2] - Hcc[k1][k2] rcv[k2])
//   Vcc = m (rcv[k1] Hvc[k1][k2] - Hcv[k1][k2] rvc[k2])
//   Vvv = m (rvc[k1] Hcv[k1][k2] - Hvc[k1][k2] rcv[k2]),  Vvc = Vcv^H.
__global__ __launch_bounds__(RED_T) void k_vn1n2(int N, const cd *rcv, const cd *Hcc, const cd *Hhh, const cd *Hcv,
                                                 double hbar, cd *Vcc, cd *Vvv, cd *Vcv, cd *Vvc) {
#pragma clang fp contract(off)
  const long t = (long)blockIdx.x * RED_T + threadIdx.x;
  if (t >= (long)N * N) return;
  const int k2 = (int)(t % N), k1 = (int)(t / N);
  const long tt = (long)k2 * N + k1;
  const cd r1 = rcv[k1], r2 = rcv[k2];
  const cd v1 = cconj(r1), v2 = cconj(r2);
  const cd hh = Hhh[tt], hvv{-hh.re, -hh.im};
  const cd hcvT = Hcv[tt], hvc{hcvT.re, -hcvT.im};
  const cd hcc = Hcc[t], hcv = Hcv[t];
  const double mi = -1.0 / hbar;
  // m z = (-0 zr - mi zi, -0 zi + mi zr) = (-(mi zi), mi zr) exactly
  auto mz = [&](cd z) { return cd{-(mi * z.im), mi * z.re}; };
  const cd zcv = csub(cmul_rn(r1, hvv), cmul_rn(hcc, r2));
  const cd zcc = csub(cmul_rn(r1, hvc), cmul_rn(hcv, v2));
  const cd zvv = csub(cmul_rn(v1, hcv), cmul_rn(hvc, r2));
  const cd vcv = mz(zcv);
  Vcv[t] = vcv;
  Vcc[t] = mz(zcc);
  Vvv[t] = mz(zvv);
  Vvc[tt] = cconj(vcv);
}

// PLSpectrum (emission.py:291-361), part 1, grid N + 1: block k < N writes
// work[k] = gap + Ee[k] + Eh[k] + Ec[k] on |ne|, |nh| (Ec, emission.py:133-171);
// block N writes work[N] = (Te + Th) / 2 with T = 2 Re(sum n E / (sum n + 1e-100)) / kB
// (Temperature, AvgEnergy: usefulsubs.py:1203-1253).
__global__ __launch_bounds__(RED_T) void k_pl_prep(int N, const cd *ne, const cd *nh, const double *Ee,
                                                   const double *Eh, double gap, const double *VC, double kB,
                                                   double *work) {
  __shared__ double sh[2][RED_T];
  const int k = blockIdx.x;
  if (k == N) {
    double nr = 0.0, ni = 0.0, dr = 0.0, di = 0.0, hr = 0.0, hi = 0.0, gr = 0.0, gi = 0.0;
    for (int i = threadIdx.x; i < N; i += RED_T) {
      nr += ne[i].re * Ee[i];
      ni += ne[i].im * Ee[i];
      dr += ne[i].re;
      di += ne[i].im;
      hr += nh[i].re * Eh[i];
      hi += nh[i].im * Eh[i];
      gr += nh[i].re;
      gi += nh[i].im;
    }
    block_sum2(nr, ni, sh);
    block_sum2(dr, di, sh);
    block_sum2(hr, hi, sh);
    block_sum2(gr, gi, sh);
    if (threadIdx.x == 0) {
      dr += 1e-100;
      gr += 1e-100;
      const double ae = (nr * dr + ni * di) / (dr * dr + di * di);
      const double ah = (hr * gr + hi * gi) / (gr * gr + gi * gi);
      work[N] = (2.0 * ae / kB + 2.0 * ah / kB) / 2.0;
    }
    return;
  }
  auto V = [&](int a, int b, int c) { return VC[((long)a * N + b) * 3 + c]; };
  const double vkk_eh = V(k, k, 0), vkk_ee = V(k, k, 1), vkk_hh = V(k, k, 2);
  double s1 = 0.0, s2 = 0.0;
  for (int i = threadIdx.x; i < N; i += RED_T) {
    const double a = cabs_(ne[i]), b = cabs_(nh[i]), id = (i == k) ? 0.0 : 1.0;
    s1 += a * (vkk_ee - V(i, k, 1)) + b * (vkk_hh - V(i, k, 2));
    s2 += a * vkk_eh * id - b * vkk_eh * id;
  }
  block_sum2(s1, s2, sh);
  if (threadIdx.x == 0) work[k] = ((gap + Ee[k]) + Eh[k]) + ((s1 + s2) - vkk_eh);
}

// LinearInterp on ky = 0, 1, ..., N-1 (libpulsesuite/helpers.py:951-995,
// locator usefulsubs.py:1803-1830): x0 = j / 100, i = min(floor(x0), N - 2)
__device__ __forceinline__ void pl_interp(int N, int j, int &i, double &x0) {
  x0 = (double)j / 100.0;
  i = min(j / 100, N - 2);
}

// PLSpectrum part 2: block per photon energy w;
//   PLS[w] = (PLS[w] + RScale sum_j hw rho0(hw) nenh_j exp(-|hw - E_j| / kB / T)
//             Lrtz(hw - E_j, hbar geh) softtheta(hw - E_mid, hbar geh)) softtheta(t, geh)
// over the 100 N interpolated points j (E_j from work, nenh_j = |(ne nh)(q_j)|).
__global__ __launch_bounds__(RED_T) void k_pl(int N, const cd *ne, const cd *nh, const double *hw,
                                              const double *work, double hg, double rden, double kB,
                                              double rscale, double tfac, double *PLS) {
  __shared__ double sh[2][RED_T];
  const int w = blockIdx.x, X = 100, M = X * N;
  const double h = hw[w];
  const double hr = h * ((h * h) / rden);
  const double T = work[N];
  auto E_at = [&](int j) {
    int i;
    double x0;
    pl_interp(N, j, i, x0);
    return work[i] * ((double)(i + 1) - x0) + work[i + 1] * (x0 - (double)i);
  };
  const double st = 0.5 * (1.0 + 2.0 / PI * atan((h - E_at(M / 2)) / hg));
  const double hg2 = hg * hg;
  double acc = 0.0, dummy = 0.0;
  for (int j = threadIdx.x; j < M; j += RED_T) {
    int i;
    double x0;
    pl_interp(N, j, i, x0);
    const double a = (double)(i + 1) - x0, b = x0 - (double)i;
    const double E = work[i] * a + work[i + 1] * b;
    const cd f0 = cmul(ne[i], nh[i]), f1 = cmul(ne[i + 1], nh[i + 1]);
    const double nn = hypot(f0.re * a + f1.re * b, f0.im * a + f1.im * b);
    const double d = h - E;
    const double lz = (hg / PI) / (d * d + hg2);
    acc += (((hr * nn) * exp(-fabs(d) / kB / T)) * lz) * st;
  }
  block_sum2(acc, dummy, sh);
  if (threadIdx.x == 0) PLS[w] = (PLS[w] + rscale * acc) * tfac;
}

// Accurate K0(x) for x > 0 (the reference's usefulsubs.K03 through CIK01):
// Zhang & Jin, Computation of Special Functions (1996), routine CIK01 --
// power series of I0 for x <= 18 (asymptotic series above), K0 by its power
// series for x <= 9 and from the asymptotic expansion of I0 K0 above; zero
// for x > 100 as K03 returns.  The asymptotic coefficients are generated
// from their closed forms: I0: ((2k-1)!!)^2 / (k! 8^k);
// I0 K0: c_k = c_{k-1} (2k-1)^3 / (8k).
__device__ double k0_accurate(double x) {
  if (x > 100.0) return 0.0;
  const double z2 = x * x;
  double i0;
  if (x <= 18.0) {
    i0 = 1.0;
    double cr = 1.0;
    for (int k = 1; k <= 50; ++k) {
      cr = 0.25 * cr * z2 / (double)(k * k);
      i0 += cr;
      if (fabs(cr / i0) < 1e-15) break;
    }
  } else {
    const int kn = x < 35.0 ? 12 : (x < 50.0 ? 9 : 7);
    double s = 1.0, num = 1.0, fact = 1.0, p8 = 1.0;
    for (int k = 1; k <= kn; ++k) {
      num *= (double)((2 * k - 1) * (2 * k - 1));
      fact *= (double)k;
      p8 *= 8.0;
      s += num / (fact * p8) * pow(x, (double)(-k));
    }
    i0 = exp(x) / sqrt(2.0 * PI * x) * s;
  }
  if (x <= 9.0) {
    double cs = 0.0, cw = 0.0, w0 = 0.0, cr = 1.0;
    const double ct = -log(0.5 * x) - 0.5772156649015329;
    for (int k = 1; k <= 50; ++k) {
      w0 += 1.0 / k;
      cr = 0.25 * cr / (double)(k * k) * z2;
      cs += cr * (w0 + ct);
      if (fabs((cs - cw) / cs) < 1e-15) break;
      cw = cs;
    }
    return ct + cs;
  }
  double s = 1.0, v = 1.0;
  for (int k = 1; k <= 10; ++k) {
    const double m = (double)(2 * k - 1);
    v *= m * m * m / (8.0 * k);
    s += v * pow(z2, (double)(-k));
  }
  return 0.5 / x * s / i0;
}

// Vehint (coulomb.py:180-242): one block; the double sum over the middle
// half of y (N1 = Ny/4 .. N2 = 3Ny/4) of exp(-(ae y_i)^2 - (ah y_j)^2)
// (ae ah / pi dy^2) K0(dk r_ij), r_ij = sqrt((y_i - y_j)^2 + Delta0^2),
// dk = max(|dky|, (ae + ah)/4), with the accurate K0.
__global__ __launch_bounds__(RED_T) void k_vehint(double dky, const double *y, int Ny, double ae, double ah,
                                                  double Delta0, double *out) {
  __shared__ double sh[2][RED_T];
  const int N1 = Ny / 4, N2 = 3 * Ny / 4, M = N2 - N1 + 1;
  const double dk = fmax(fabs(dky), (ae + ah) / 4.0);
  const double dy = y[1] - y[0];
  const double mult = ae * ah / PI * (dy * dy);
  double acc = 0.0, dummy = 0.0;
  for (long t = threadIdx.x; t < (long)M * M; t += RED_T) {
    const int i = N1 + (int)(t / M), j = N1 + (int)(t % M);
    const double yi = y[i], yj = y[j];
    const double r = sqrt((yi - yj) * (yi - yj) + Delta0 * Delta0);
    const double ex = exp(-((ae * yi) * (ae * yi)) - (ah * yj) * (ah * yj));
    acc += ex * mult * k0_accurate(dk * r);
  }
  block_sum2(acc, dummy, sh);
  if (threadIdx.x == 0) out[0] = acc;
}


// CalcOffDiagDeph_E (hole 0) / CalcOffDiagDeph_H (hole 1) (dephasing.py:
// 107-190, 627-685), the version-1 tables: thread per (k, q), t = k N + q,
//   D[k][q] = pi/hbar (sum_k1 Vss(k1, a)^2 L(Es[a] + Es[k] - Es[k1] - Es[b], hbar gss)
//             (ns[a] ns[k] (1 - ns[k1]) + (1 - ns[a]) (1 - ns[k]) ns[k1])
//           + sum_k1 Vso^2 L(Eo[c] + Es[k] - Eo[k1] - Es[b], hbar gso)
//             (no[c] (1 - no[k1]) ns[k] + (1 - no[c]) no[k1] (1 - ns[k])))
// with b = kpq[k][q], a = kpq[k1][q], c = kmq[k1][q] (0-based; an index
// outside [0, N) drops the term), n = |n0|, Vso = Veh at (k, b) for electrons
// and (k1, c) for holes; the same-carrier sum first, both over k1 ascending.
__global__ __launch_bounds__(RED_T) void k_od1_tables(int N, int hole, const cd *ne0, const cd *nh0, const double *VC,
                                                      const double *Ee, const double *Eh, const int32_t *kpq,
                                                      const int32_t *kmq, double gss, double gso, double hbar,
                                                      double *D) {
#pragma clang fp contract(off)
  const long t = (long)blockIdx.x * RED_T + threadIdx.x;
  if (t >= (long)N * N) return;
  const int q = (int)(t % N), k = (int)(t / N);
  const cd *ns = hole ? nh0 : ne0, *no = hole ? ne0 : nh0;
  const double *Es = hole ? Eh : Ee, *Eo = hole ? Ee : Eh;
  const int css = hole ? 2 : 1;
  auto V = [&](int a, int b, int c) { return VC[((long)a * N + b) * 3 + c]; };
  auto lrtz = [](double a, double b) { return (b / PI) / (a * a + b * b); };
  double d = 0.0;
  const int b = kpq[(long)k * N + q];
  if (b >= 0 && b < N) {
    const double nk = cabs_(ns[k]), hs = hbar * gss, ho = hbar * gso;
    for (int k1 = 0; k1 < N; ++k1) {
      const int a = kpq[(long)k1 * N + q];
      if (a < 0 || a >= N) continue;
      const double Ed = ((Es[a] + Es[k]) - Es[k1]) - Es[b];
      const double v = V(k1, a, css), na = cabs_(ns[a]), n1 = cabs_(ns[k1]);
      const double occ = (na * nk) * (1.0 - n1) + ((1.0 - na) * (1.0 - nk)) * n1;
      d = d + ((v * v) * lrtz(Ed, hs)) * occ;
    }
    for (int k1 = 0; k1 < N; ++k1) {
      const int c = kmq[(long)k1 * N + q];
      if (c < 0 || c >= N) continue;
      const double Ed = ((Eo[c] + Es[k]) - Eo[k1]) - Es[b];
      const double v = hole ? V(k1, c, 0) : V(k, b, 0), nc = cabs_(no[c]), n1 = cabs_(no[k1]);
      const double occ = (nc * (1.0 - n1)) * nk + ((1.0 - nc) * n1) * (1.0 - nk);
      d = d + ((v * v) * lrtz(Ed, ho)) * occ;
    }
  }
  D[t] = (d * PI) / hbar;
}

// OffDiagDephasing (dephasing.py:193-218, 687-712), the version-1
// contraction: thread per (k, kp), t = k N + kp,
//   x[k][kp] = i hbar (sum_qp DH[kp][qp] P[kpq[qp][kp]][k - 1] u[qp]
//                      + sum_q DE[k][q] P[kp - 1][kpq[q][k]] u[q]),
// u = |ky| / (|ky| + 1e-10), P[i][-1] = P[-1][j] = 0: the reference reads its
// one-padded copies of p and p^T with the unshifted column index.
__global__ __launch_bounds__(RED_T) void k_od1_apply(int N, const double *DE, const double *DH, const cd *p,
                                                     const int32_t *kpq, const double *ky, double hbar, cd *x) {
#pragma clang fp contract(off)
  const long t = (long)blockIdx.x * RED_T + threadIdx.x;
  if (t >= (long)N * N) return;
  const int kp = (int)(t % N), k = (int)(t / N);
  cd acc{0.0, 0.0};
  for (int qp = 0; qp < N; ++qp) {
    const int r = kpq[(long)qp * N + kp];
    if (r < 0 || r >= N) continue;
    const cd v = k >= 1 ? p[(long)r * N + (k - 1)] : cd{0.0, 0.0};
    const double dh = DH[(long)kp * N + qp], u = fabs(ky[qp]) / (fabs(ky[qp]) + 1e-10);
    acc.re = acc.re + (dh * v.re) * u;
    acc.im = acc.im + (dh * v.im) * u;
  }
  for (int q = 0; q < N; ++q) {
    const int r = kpq[(long)q * N + k];
    if (r < 0 || r >= N) continue;
    const cd v = kp >= 1 ? p[(long)(kp - 1) * N + r] : cd{0.0, 0.0};
    const double de = DE[(long)k * N + q], u = fabs(ky[q]) / (fabs(ky[q]) + 1e-10);
    acc.re = acc.re + (de * v.re) * u;
    acc.im = acc.im + (de * v.im) * u;
  }
  x[t] = cd{-acc.im * hbar, acc.re * hbar};
}

}  // namespace mod
}  // namespace psa
