// Device-side helpers shared by the pulsesuite_amd kernels (gfx950).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/pulsesuite_amd.h"

namespace psa {

struct cd {
  double re, im;
};

__device__ __forceinline__ cd cmk(double r, double i) { return cd{r, i}; }
__device__ __forceinline__ cd cadd(cd a, cd b) { return cd{a.re + b.re, a.im + b.im}; }
__device__ __forceinline__ cd csub(cd a, cd b) { return cd{a.re - b.re, a.im - b.im}; }
__device__ __forceinline__ cd cmul(cd a, cd b) {
  return cd{a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re};
}
__device__ __forceinline__ cd cscale(cd a, double s) { return cd{a.re * s, a.im * s}; }
__device__ __forceinline__ cd cconj(cd a) { return cd{a.re, -a.im}; }
__device__ __forceinline__ double cabs_(cd a) { return hypot(a.re, a.im); }
// z / (i hbar) = -i z / hbar
__device__ __forceinline__ cd cdiv_ih(cd z, double hbar) { return cd{z.im / hbar, -z.re / hbar}; }

__device__ __forceinline__ cd ld(const double *p, long i) {
  const double2 v = reinterpret_cast<const double2 *>(p)[i];
  return cd{v.x, v.y};
}
__device__ __forceinline__ void st(double *p, long i, cd v) {
  reinterpret_cast<double2 *>(p)[i] = make_double2(v.re, v.im);
}

constexpr double PI = 3.141592653589793;

// ---------------------------------------------------------------------------
// block reductions (256 threads, wave64)
// ---------------------------------------------------------------------------
__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}

// Deterministic block sum; all threads get the result.  `sh` >= 4 doubles.
__device__ __forceinline__ double block_sum(double v, double *sh) {
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  const int nw = (blockDim.x + 63) >> 6;
  v = wave_sum(v);
  __syncthreads();
  if (lane == 0) sh[wid] = v;
  __syncthreads();
  double t = 0.0;
  for (int i = 0; i < nw; ++i) t += sh[i];
  __syncthreads();
  return t;
}

__device__ __forceinline__ double block_max(double v, double *sh) {
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  const int nw = (blockDim.x + 63) >> 6;
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v = fmax(v, __shfl_xor(v, o, 64));
  __syncthreads();
  if (lane == 0) sh[wid] = v;
  __syncthreads();
  double t = sh[0];
  for (int i = 1; i < nw; ++i) t = fmax(t, sh[i]);
  __syncthreads();
  return t;
}

// (-1)^w and yw(w) = (-1)^floor((w-1)/2)  (qwoptics.py:1093-1106)
__device__ __forceinline__ double sgn_pow(int w) { return (w & 1) ? -1.0 : 1.0; }
__device__ __forceinline__ double yw_of(int w) {
  const int f = (w - 1) >= 0 ? (w - 1) / 2 : -((2 - w) / 2);
  return (f & 1) ? -1.0 : 1.0;
}

__device__ __forceinline__ long mat_off(const psa_group &g, int w) {
  return (long)w * g.N * g.N * 2;  // doubles
}

// The Hamiltonian buffers Heh, Heo, Hho hold complex64 in the split-BF16
// mode (packed from the start of the buffer: wire w at float offset
// mat_off(g, w)), complex128 otherwise.
__device__ __forceinline__ bool h32(const psa_group &g) { return g.prec == PSA_PREC_BF16X3; }
__device__ __forceinline__ cd ld_h32(const double *base, long o, long ij) {
  const float2 v = reinterpret_cast<const float2 *>(reinterpret_cast<const float *>(base) + o)[ij];
  return cd{(double)v.x, (double)v.y};
}
__device__ __forceinline__ void st_h32(double *base, long o, long ij, cd v) {
  reinterpret_cast<float2 *>(reinterpret_cast<float *>(base) + o)[ij] = make_float2((float)v.re, (float)v.im);
}
// element ij of wire offset o, in the group's storage precision
__device__ __forceinline__ cd ld_h(const psa_group &g, const double *base, long o, long ij) {
  return h32(g) ? ld_h32(base, o, ij) : ld(base + o, ij);
}
__device__ __forceinline__ void st_h(const psa_group &g, double *base, long o, long ij, cd v) {
  if (h32(g)) st_h32(base, o, ij, v);
  else st(base + o, ij, v);
}

}  // namespace psa
