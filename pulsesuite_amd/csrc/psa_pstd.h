// 1-D pseudo-spectral field propagator coupled to the wires (psa_pstd in
// the C ABI).  The reference has no propagator on the wire path; it ships
// the staggered differences dEdx / dHdx (usefulsubs.py:612-655, 704-747) and
// the real-space absorber ApplyABC (usefulsubs.py:584-609), which this
// propagator is built from:
//
//   k_pstd_update   one thread per k: the current density of the point
//                   sheets in k-space (sum over wires of J_w w^(k j_w)) and
//                   the exact 2x2 update of (E, H) over dt for both pairs;
//   k_pstd_line     one workgroup per line (N <= 8192 complex values in
//                   LDS): inverse FFT, absorber, samples at the wires and
//                   observation points, forward FFT -- ApplyABC plus the
//                   reads the wires need, in one pass over HBM;
//   k_pstd_gather   one workgroup per (wire, component): the wire's
//                   polarization averaged over its propagation grid;
//   k_pstd_scatter  the field at each wire onto its propagation grid.
//
// The FFTs are block-level Stockham autosort transforms, radix 4 with a
// final radix 2 when log2 N is odd, in place in LDS (each stage reads its
// inputs into registers, barrier, writes), twiddles from the table twid.
#pragma once

#include "psa_dev.h"

namespace psa {
namespace pstd {

// LDS slot of element i: bits 0-3 XOR-ed by 5 ((i >> 4) & 3), a permutation
// within each 16-element block that puts the stride-4 writes of the first
// radix-4 stage (and the runs of four at stride 16 of the second) on 16
// different 16-byte bank groups per 16 lanes.
__device__ __forceinline__ int slot(int i) { return i ^ (5 * ((i >> 4) & 3)); }

template <int LOGN>
struct Fft {
  static constexpr int N = 1 << LOGN;
  static constexpr int T = N >= 4096 ? 1024 : (N / 4 < 64 ? 64 : N / 4);  // threads
};

// Twiddles w^m = exp(-2 pi i m / N): from the global table, or from an LDS
// quarter table q[r] = w^r (r < N/4) rotated by (-i)^(m / (N/4)) -- exact
// sign flips and swaps, so both give the same values.
struct TwGlobal {
  const double *t;
  __device__ __forceinline__ cd operator()(int m) const { return ld(t, m); }
};
template <int LOGN>
struct TwQuarter {
  const cd *q;
  __device__ __forceinline__ cd operator()(int m) const {
    const cd a = q[m & ((1 << (LOGN - 2)) - 1)];
    switch (m >> (LOGN - 2)) {
      case 0: return a;
      case 1: return cd{a.im, -a.re};
      case 2: return cd{-a.re, -a.im};
      default: return cd{-a.im, a.re};
    }
  }
};

// x[j + r M] w^(r k N/(4 Ns)) -> DFT4 -> x[(j - k) 4 + k + r Ns], k = j mod Ns;
// SIGN -1 forward (numpy.fft.fft), +1 inverse without the 1/N.
template <int LOGN, int SIGN, class TW>
__device__ __forceinline__ void block_fft(cd *x, const TW &twid, int tid) {
  constexpr int N = Fft<LOGN>::N, T = Fft<LOGN>::T;
  constexpr int M4 = N / 4, IT4 = (M4 + T - 1) / T;
  // stages not unrolled: unrolling lets the compiler hoist later stages'
  // twiddle loads and costs registers for nothing (the barriers serialise)
#pragma unroll 1
  for (int s = 0; s < LOGN / 2; ++s) {
    const int Ns = 1 << (2 * s), step = N / (4 * Ns);
    cd o[IT4][4];
    int base[IT4];
#pragma unroll
    for (int it = 0; it < IT4; ++it) {
      const int j = tid + T * it;
      if (M4 % T == 0 || j < M4) {
        const int k = j & (Ns - 1);
        cd a[4];
        a[0] = x[slot(j)];
#pragma unroll
        for (int r = 1; r < 4; ++r) {
          cd w = twid(r * k * step);
          if (SIGN > 0) w.im = -w.im;
          a[r] = s == 0 ? x[slot(j + r * M4)] : cmul(x[slot(j + r * M4)], w);
        }
        const cd s02 = cadd(a[0], a[2]), d02 = csub(a[0], a[2]);
        const cd s13 = cadd(a[1], a[3]), t13 = csub(a[1], a[3]);
        // (a1 - a3) * (SIGN i)
        const cd d13 = SIGN > 0 ? cd{-t13.im, t13.re} : cd{t13.im, -t13.re};
        o[it][0] = cadd(s02, s13);
        o[it][1] = cadd(d02, d13);
        o[it][2] = csub(s02, s13);
        o[it][3] = csub(d02, d13);
        base[it] = (j - k) * 4 + k;
      }
    }
    __syncthreads();
#pragma unroll
    for (int it = 0; it < IT4; ++it) {
      const int j = tid + T * it;
      if (M4 % T == 0 || j < M4) {
#pragma unroll
        for (int r = 0; r < 4; ++r) x[slot(base[it] + r * Ns)] = o[it][r];
      }
    }
    __syncthreads();
  }
  if constexpr (LOGN & 1) {
    // final radix 2, Ns = N/2: k = j, outputs at j and j + N/2 (in place)
    constexpr int M2 = N / 2, IT2 = (M2 + T - 1) / T;
#pragma unroll
    for (int it = 0; it < IT2; ++it) {
      const int j = tid + T * it;
      if (M2 % T == 0 || j < M2) {
        cd w = twid(j);
        if (SIGN > 0) w.im = -w.im;
        const cd a0 = x[slot(j)], a1 = cmul(x[slot(j + M2)], w);
        x[slot(j)] = cadd(a0, a1);
        x[slot(j + M2)] = csub(a0, a1);
      }
    }
    __syncthreads();
  }
}

enum LineMode { TO_K = 0, TO_REAL = 1, STEP = 2, SAMPLE = 3 };

// sc * x[j] of an LDS line of N values; NaN for a point outside the grid
template <int N>
__device__ __forceinline__ cd at(const cd *x, int j, double sc) {
  if ((unsigned)j >= (unsigned)N) return cd{__builtin_nan(""), __builtin_nan("")};
  return cscale(x[slot(j)], sc);
}

// One line per workgroup.  TO_K / TO_REAL / STEP with the absorber: lines
// 0..3 = blockIdx.x; SAMPLE and STEP without the absorber: the E lines
// 0 and 2 (blockIdx.x 0, 1).
template <int LOGN>
__global__ __launch_bounds__(Fft<LOGN>::T) void k_pstd_line(psa_pstd p, int mode, double *out) {
  constexpr int N = Fft<LOGN>::N, T = Fft<LOGN>::T;
  // the quarter twiddle table in LDS up to N = 4096 (80 KB with the line);
  // N = 8192 reads the global table (the line alone takes 128 KB)
  constexpr bool QTW = LOGN <= 12;
  __shared__ cd x[N];
  __shared__ cd tq[QTW ? N / 4 : 1];
  const int tid = threadIdx.x;
  if constexpr (QTW)
    for (int i = tid; i < N / 4; i += T) tq[i] = ld(p.twid, i);
  const auto tw = [&]() {
    if constexpr (QTW) return TwQuarter<LOGN>{tq};
    else return TwGlobal{p.twid};
  }();
  const bool abc = mode == STEP && (p.flags & PSA_PSTD_ABC);
  const int line = (mode == SAMPLE || (mode == STEP && !abc)) ? 2 * blockIdx.x : blockIdx.x;
  double *st_line = p.state + 2L * N * line;
  for (int i = tid; i < N; i += T) x[slot(i)] = ld(st_line, i);
  __syncthreads();
  if (mode == TO_K) {
    block_fft<LOGN, -1>(x, tw, tid);
    for (int i = tid; i < N; i += T) st(st_line, i, x[slot(i)]);
    return;
  }
  block_fft<LOGN, +1>(x, tw, tid);
  const double inv = 1.0 / (double)N;
  if (mode == TO_REAL) {
    for (int i = tid; i < N; i += T) st(out + 2L * N * line, i, cscale(x[slot(i)], inv));
    return;
  }
  if (abc) {
    for (int i = tid; i < N; i += T) x[slot(i)] = cscale(x[slot(i)], p.abc[i] * inv);
    __syncthreads();
  }
  const double sc = abc ? 1.0 : inv;
  if ((line & 1) == 0) {  // E lines: the wires' drive and the record
    const int f = line >> 1;
    for (int s = tid; s < p.nw; s += T) st(p.ew, (long)f * p.nw + s, at<N>(x, p.jw[s], sc));
    if (mode == STEP && p.nobs > 0) {
      const long c = p.count[0] - 1;
      if (c >= 0 && c < p.nrec)
        for (int o = tid; o < p.nobs; o += T) st(p.rec, (c * 2 + f) * p.nobs + o, at<N>(x, p.obs[o], sc));
    }
  }
  if (abc) {
    __syncthreads();
    block_fft<LOGN, -1>(x, tw, tid);
    for (int i = tid; i < N; i += T) st(st_line, i, x[slot(i)]);
  }
}

// One thread per k: J^(k) = sum_w J_w exp(-2 pi i k j_w / N) for x and z,
// then for each pair (E, H, s) = (Ex, Hz, +1), (Ez, Hx, -1), with
// a = s ikh / eps, b = s ike / mu, d = a H - J^/eps:
//   E' = C E + S d,   H' = H + b (S E + Q d).
constexpr int UPD_T = 64;
__global__ __launch_bounds__(UPD_T) void k_pstd_update(psa_pstd p) {
  __shared__ cd js[2][PSA_PSTD_MAXW];
  __shared__ int jws[PSA_PSTD_MAXW];
  const int tid = threadIdx.x, nw = p.nw, N = p.N;
  for (int s = tid; s < nw; s += UPD_T) {
    js[0][s] = ld(p.jsrc, s);
    js[1][s] = ld(p.jsrc, nw + s);
    jws[s] = p.jw[s];
  }
  if (blockIdx.x == 0 && tid == 0) p.count[0] += 1;
  __syncthreads();
  const int k = blockIdx.x * UPD_T + tid;
  if (k >= N) return;
  cd jx{0.0, 0.0}, jz{0.0, 0.0};
  // eight wires at a time, their twiddle loads in flight together
  for (int s0 = 0; s0 < nw; s0 += 8) {
    cd w[8];
#pragma unroll
    for (int u = 0; u < 8; ++u) {
      const int s = min(s0 + u, nw - 1);
      w[u] = ld(p.twid, (int)(((long)k * jws[s]) & (N - 1)));
    }
#pragma unroll
    for (int u = 0; u < 8; ++u) {
      if (s0 + u < nw) {
        jx = cadd(jx, cmul(js[0][s0 + u], w[u]));
        jz = cadd(jz, cmul(js[1][s0 + u], w[u]));
      }
    }
  }
  const double *m = p.mode + 8L * k;
  const cd ike{m[0], m[1]}, ikh{m[2], m[3]};
  const double C = m[4], S = m[5], Q = m[6];
  const double ie = 1.0 / p.eps, im = 1.0 / p.mu;
#pragma unroll
  for (int f = 0; f < 2; ++f) {
    const double sg = f == 0 ? 1.0 : -1.0;
    const cd J = f == 0 ? jx : jz;
    double *le = p.state + 2L * N * (2 * f), *lh = le + 2L * N;
    const cd E = ld(le, k), H = ld(lh, k);
    const cd a = cscale(ikh, sg * ie), b = cscale(ike, sg * im);
    const cd d = csub(cmul(a, H), cscale(J, ie));
    const cd E1 = cadd(cscale(E, C), cscale(d, S));
    const cd H1 = cadd(H, cmul(b, cadd(cscale(E, S), cscale(d, Q))));
    st(le, k, E1);
    st(lh, k, H1);
  }
}

// Wire w0 + blockIdx.x, component blockIdx.y (0: Pxx -> x, 1: Pzz -> z):
// P = sheet * mean over rr of out; jsrc = (P - pold) / dt; pold = P.
__global__ __launch_bounds__(256) void k_pstd_gather(psa_pstd p, psa_group g, int w0, int slot0) {
  __shared__ double sh[8];
  const int f = blockIdx.y, w = w0 + blockIdx.x, s = slot0 + blockIdx.x;
  const double *o = g.out + 2L * ((long)(f == 0 ? 0 : 2) * g.Nw + w) * g.Nx;
  // four independent partial sums per thread (loads in flight together),
  // combined in a fixed order
  double re[4] = {0.0, 0.0, 0.0, 0.0}, im[4] = {0.0, 0.0, 0.0, 0.0};
  for (int i0 = threadIdx.x; i0 < g.Nx; i0 += 1024) {
    cd v[4];
#pragma unroll
    for (int u = 0; u < 4; ++u) v[u] = i0 + 256 * u < g.Nx ? ld(o, i0 + 256 * u) : cd{0.0, 0.0};
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      re[u] += v[u].re;
      im[u] += v[u].im;
    }
  }
  const double rs = block_sum((re[0] + re[1]) + (re[2] + re[3]), sh);
  const double is = block_sum((im[0] + im[1]) + (im[2] + im[3]), sh);
  if (threadIdx.x == 0) {
    const double sc = p.sheet / (double)g.Nx, idt = 1.0 / p.dt;
    const cd P{rs * sc, is * sc};
    const long at = (long)f * p.nw + s;
    const cd P0 = ld(p.pold, at);
    st(p.jsrc, at, cscale(csub(P, P0), idt));
    st(p.pold, at, P);
  }
}

// efield[0 / 2][w][:] = ew[0 / 1][slot] for wires w0 + blockIdx.y.
__global__ __launch_bounds__(256) void k_pstd_scatter(psa_pstd p, psa_group g, int w0, int slot0) {
  const int f = blockIdx.z, w = w0 + blockIdx.y, s = slot0 + blockIdx.y;
  const int i = blockIdx.x * 256 + threadIdx.x;
  if (i >= g.Nx) return;
  const cd e = ld(p.ew, (long)f * p.nw + s);
  st(g.efield, ((long)(f == 0 ? 0 : 2) * g.Nw + w) * g.Nx + i, e);
}

}  // namespace pstd
}  // namespace psa
