// Split-BF16 right-hand sides (PSA_PREC_BF16X3) for gfx950.
//
// Every FP64 operand x of the RHS contractions (RhsProb, SBEs.py:1941-1965,
// 2308-2327, 2606-2625) is written once per step as two bf16 values,
//   f = float(x), hi = bf16(f), lo = bf16(f - float(hi))
// (v_cvt_pk_bf16_f32, round to nearest even; |x - hi - lo| <= ~2^-17 |x|),
// for the real and the imaginary part: four bf16 planes per operand slot
// (k_split for p, C, D; the Hamiltonians are split in registers while their
// slabs are staged, copy_h).  A real product a b is a_hi b_hi + a_hi b_lo + a_lo b_hi on
// v_mfma_f32_16x16x32_bf16 with FP32 accumulation (the lo*lo term, ~2^-16
// relative, is dropped); a complex product is four real ones, 12 MFMAs per
// 16 x 16 x 32 block.  The diagonal parts, the epilogue and the state stay
// in FP64 (RhsProb::store).
//
// Operand slots per wire: slot = 2 * matrix + orientation for the matrices
// 0 p, 1 C_o, 2 D_o (diagonal zeroed), 3 Heh, 4 Hee_o, 5 Hhh_o
// (orientation 0 the matrix, 1 its transpose); slot 12 the band
// i hbar Dh[m - x][x] and slot 13 i hbar De[m - x][x] of the off-diagonal
// dephasing (rows x, k = m; purely imaginary, real planes zero).  An A operand A(m, k)
// reads slot rows m, a B operand B(k, n) slot rows n, both k-contiguous
// as the MFMA lanes take them (A[l%16][8(l/16)+e], B[8(l/16)+e][l%16]).
// The Hamiltonian slots 6-11 are not kept in the plane buffer: the RHS
// kernel reads the complex64 Hamiltonians (8 bytes per value, as many as
// the four planes) and splits each slab in registers as it stages it
// (copy_h), into the same LDS layout.
//
// Layout: a slot is padded to NR = 128 ceil(N/128) rows and NK = 32
// ceil(N/32) k (zeros) and stored as [row block of 64][k block of 32]
// chunks of 16 KB = [4 planes][64 rows][32 k], with the 16-byte group of 8
// k in a row at position (k >> 3) ^ ((row >> 1) & 3): the layout the RHS
// kernel keeps in LDS, so a chunk is copied there byte for byte (direct
// global-to-LDS loads), and the MFMA operand reads (8 lanes = 8
// consecutive rows per 128-byte bank row) fall on distinct LDS banks.
#pragma once

#include "psa_gemm.h"

namespace psa {
namespace sp {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef __bf16 bf16x2 __attribute__((ext_vector_type(2)));
typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));

constexpr int TILE = 128, BK = 32, NSLOT = 14;

__host__ __device__ inline int nr(int N) { return (N + 127) / 128 * 128; }
__host__ __device__ inline int nk(int N) { return (N + 31) / 32 * 32; }
__host__ __device__ inline long slot_elems(int N) { return 4L * nr(N) * nk(N); }
__host__ __device__ inline long chunk(int N, int w, int slot, int rb, int kb) {
  return ((long)w * NSLOT + slot) * slot_elems(N) + ((long)rb * (nk(N) / 32) + kb) * (4 * 64 * 32);
}
// position of (plane, row, k) inside a chunk
__host__ __device__ inline int in_chunk(int pl, int row, int k) {
  return pl * 2048 + (row & 63) * 32 + ((((k & 31) >> 3) ^ (((row & 63) >> 1) & 3)) << 3) + (k & 7);
}

// hi/lo split of two values (the bf16 pair of f = float(x), then of f - hi)
__device__ __forceinline__ void split_pair(double x0, double x1, uint32_t &hi, uint32_t &lo) {
  const float f0 = (float)x0, f1 = (float)x1;
  const bf16x2 h = {(__bf16)f0, (__bf16)f1};
  const bf16x2 l = {(__bf16)(f0 - (float)h[0]), (__bf16)(f1 - (float)h[1])};
  hi = __builtin_bit_cast(uint32_t, h);
  lo = __builtin_bit_cast(uint32_t, l);
}

// ---------------------------------------------------------------------------
// k_split: block (64 x 64 tile of the matrix, matrix p / C / D, wire).
// The tile is read coalesced along the rows (two rows per thread and
// pass), split, and kept in LDS as row pairs per column (dword = rows
// 2r, 2r+1; column stride 33 dwords), from which both slots are written
// with 16-byte stores of 8 consecutive k.  Values past N (the padding) are
// written as zeros; groups of 8 k at or past NK are not written.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_split(psa_group g, int w0) {
  __shared__ uint32_t t[4][64][33];  // [plane][column j][row pair]
  const int N = g.N, NK = nk(N), mat = blockIdx.z % 3, w = w0 + blockIdx.z / 3;
  const int r0 = blockIdx.y * 64, c0 = blockIdx.x * 64, tid = threadIdx.x;
  const long o = mat_off(g, w);
  // the planes already hold this level (written by the previous step's
  // transport): nothing to do (block-uniform, before any barrier)
  if (g.wtag && g.wtag[2 * w] == (uint64_t)(g.pB + o)) return;
  const double *X = mat == 0 ? g.pB : mat == 1 ? g.CB : g.DB;
  const bool zd = (mat == 1 || mat == 2);
  uint16_t *spl = reinterpret_cast<uint16_t *>(g.spl);
  // slot of the matrix, and of its transpose
  const int s_same = 2 * mat, s_tr = 2 * mat + 1;
  {
    const int j = tid & 63, gj = c0 + j;
    cd v[8][2];
#pragma unroll
    for (int e = 0; e < 8; ++e) {
      const int rp = (tid >> 6) + 4 * e;
#pragma unroll
      for (int r = 0; r < 2; ++r) {
        const int gi = r0 + 2 * rp + r;
        const long ij = (long)min(gi, N - 1) * N + min(gj, N - 1);
        v[e][r] = ld(X + o, ij);
        if (gi >= N || gj >= N || (zd && gi == gj)) v[e][r] = cd{0.0, 0.0};
      }
    }
#pragma unroll
    for (int e = 0; e < 8; ++e) {
      const int rp = (tid >> 6) + 4 * e;
      uint32_t rh, rl, ih, il;  // planes of the row pair (rows 2 rp, 2 rp + 1) at column j
      split_pair(v[e][0].re, v[e][1].re, rh, rl);
      split_pair(v[e][0].im, v[e][1].im, ih, il);
      t[0][j][rp] = rh;
      t[1][j][rp] = rl;
      t[2][j][rp] = ih;
      t[3][j][rp] = il;
    }
  }
  __syncthreads();
  const int x = tid >> 2, kq = tid & 3;
  // the buffer as stored: row i = x, k = columns 16 kq .. 16 kq + 15
  {
    const int gi = r0 + x;
#pragma unroll
    for (int half = 0; half < 2; ++half) {
      const int k0 = c0 + 16 * kq + 8 * half;
      if (k0 >= NK) continue;
#pragma unroll
      for (int pl = 0; pl < 4; ++pl) {
        uint16_t hv[8];
#pragma unroll
        for (int q = 0; q < 8; ++q) hv[q] = (uint16_t)(t[pl][16 * kq + 8 * half + q][x >> 1] >> (16 * (x & 1)));
        *reinterpret_cast<u32x4 *>(&spl[chunk(N, w, s_same, gi >> 6, k0 >> 5) + in_chunk(pl, gi, k0)]) =
            __builtin_bit_cast(u32x4, hv);
      }
    }
  }
  // transposed: row j = x (a column of the buffer), k = rows 16 kq .. 16 kq + 15
  {
    const int gj = c0 + x;
#pragma unroll
    for (int half = 0; half < 2; ++half) {
      const int k0 = r0 + 16 * kq + 8 * half;
      if (k0 >= NK) continue;
#pragma unroll
      for (int pl = 0; pl < 4; ++pl) {
        u32x4 v;
#pragma unroll
        for (int q = 0; q < 4; ++q) v[q] = t[pl][x][8 * kq + 4 * half + q];
        *reinterpret_cast<u32x4 *>(&spl[chunk(N, w, s_tr, gj >> 6, k0 >> 5) + in_chunk(pl, gj, k0)]) = v;
      }
    }
  }
}

// ---------------------------------------------------------------------------
// k_split_band: slots 12 (kind 1, Dh) and 13 (kind 0, De): rows x, k = m,
// value i hbar D[m - x][x] from the tables dedh[w][kind][x][q + N - 1]
// (q = m - x, q = 0 excluded).  Block (64 rows, kind, wire); a thread takes
// one 8-k group of a row, so the lanes of a wave read consecutive table
// entries of a row and write consecutive 16-byte groups of it; the
// imaginary planes only (the real planes stay as allocated: zero).
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_split_band(psa_group g, int w0) {
  const int N = g.N, NK = nk(N), Q = 2 * N - 1, G = NK / 8;  // 8-k groups per row
  const int kind = blockIdx.y, w = w0 + blockIdx.z, slot = 13 - kind;
  uint16_t *spl = reinterpret_cast<uint16_t *>(g.spl);
  const int rows = 256 / G, gq = threadIdx.x % G, k0 = 8 * gq;
  for (int r = threadIdx.x / G; r < 64 && threadIdx.x < rows * G; r += rows) {
    const int x = blockIdx.x * 64 + r, gx = min(x, N - 1);
    const double *tab = g.dedh + ((long)w * 2 + kind) * (long)N * Q + (long)gx * Q + (N - 1 - gx);
    double v[8];
#pragma unroll
    for (int e = 0; e < 8; ++e) {
      const int k = k0 + e;
      const double raw = tab[min(k, N - 1)];  // q = k - x
      v[e] = (x < N && k < N && k != x) ? g.hbar * raw : 0.0;
    }
    u32x4 hi, lo;
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      uint32_t a, b;
      split_pair(v[2 * q], v[2 * q + 1], a, b);
      hi[q] = a;
      lo[q] = b;
    }
    const long c = chunk(N, w, slot, x >> 6, k0 >> 5);
    *reinterpret_cast<u32x4 *>(&spl[c + in_chunk(2, x, k0)]) = hi;
    *reinterpret_cast<u32x4 *>(&spl[c + in_chunk(3, x, k0)]) = lo;
  }
}

// ---------------------------------------------------------------------------
// The RHS products.  The output tiles are 128 x 128 of one (wire,
// equation); four waves (2 x 2) cover a tile, each a 64 x 64 sub-tile of
// 4 x 4 MFMA blocks (128 accumulator registers).  K runs in slabs of 32
// over the equation's terms.  The slab of each operand (two 16 KB chunks:
// rows 0-63 and 64-127) is copied from HBM into an LDS stage (64 KB) by
// direct global-to-LDS loads (1 KB per wave instruction, no registers).
// ---------------------------------------------------------------------------

// One term of an equation: A and B slots, conjugation and sign.
struct Term {
  int sa, sb;
  bool ca, cb;
  bool neg;
};

// Terms in the order each equation runs them.  The twelve workgroups of a
// wire run together on one XCD, so the orders are aligned for L2 reuse:
// at every position dC reads the same B slot as dp (p^T, Hee_o^T, C_o,
// Heh^T) and dD reads dp's A slot as its B (Hhh_o, p, Heh^T, D_o^T).
__device__ __forceinline__ Term term(int e, int t) {
  if (e == 1) t = (t + 3) & 3;               // dC: 3, 0, 1, 2
  else if (e == 2) t = (4 - t) & 3;          // dD: 0, 3, 2, 1
  switch (e * 8 + t) {
    case 0: return {10, 1, false, false, false};  // dp:  Hhh_o p
    case 1: return {0, 8, false, false, false};   //      p Hee_o^T
    case 2: return {7, 3, false, false, true};    //     -Heh^T C_o
    case 3: return {5, 6, false, false, true};    //     -D_o^T Heh^T
    case 4: return {12, 1, false, false, false};  //      (i hbar Dh band) p
    case 5: return {0, 13, false, false, false};  //      p (i hbar De band)
    case 8: return {2, 8, false, false, false};   // dC:  C_o Hee_o^T
    case 9: return {9, 3, false, false, true};    //     -Hee_o^T C_o
    case 10: return {1, 6, true, false, false};   //      p^H Heh^T
    case 11: return {6, 1, true, false, true};    //     -conj(Heh) p
    case 16: return {4, 10, false, false, false}; // dD:  D_o Hhh_o^T
    case 17: return {11, 5, false, false, true};  //     -Hhh_o^T D_o
    case 18: return {0, 7, true, false, false};   //      conj(p) Heh
    default: return {7, 0, true, false, true};    //     -Heh^H p^T
  }
}

struct Stage {
  uint16_t a[2][4 * 64 * 32];  // [half: rows 0-63, 64-127][chunk layout]
  uint16_t b[2][4 * 64 * 32];
};


typedef __attribute__((address_space(3))) void lds_void;

// Copy the slab (k block kb) of slot `slot`, rows of the 128-row tile at
// x0, into dst[2][chunk]: wave v copies the 1 KB pieces v, v + 4, ... of
// the 32.  IM_ONLY (the band slots, whose real planes are zero): only the
// 16 pieces of the imaginary planes.
template <bool IM_ONLY>
__device__ __forceinline__ void copy_slab(uint16_t (*dst)[4 * 64 * 32], const uint16_t *spl, int N, int w, int slot,
                                          int x0, int kb, int wave_ = -1) {
  const int lane = threadIdx.x & 63;
  const int wave = wave_ >= 0 ? wave_ : __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
#pragma unroll
  for (int j = 0; j < (IM_ONLY ? 4 : 8); ++j) {
    // IM_ONLY: pieces 8..15 of each half (planes 2 and 3)
    const int i = IM_ONLY ? 16 * (j >> 1) + 8 + wave + 4 * (j & 1) : wave + 4 * j;
    const int h = i >> 4, off = (i & 15) * 512;  // 64-row half of the tile, uint16 offset in the chunk
    const uint16_t *src = spl + chunk(N, w, slot, (x0 >> 6) + h, kb) + off + lane * 8;
    __builtin_amdgcn_global_load_lds(src, (lds_void *)&dst[h][off], 16, 0, 0);
  }
}

// Stage the slab (k block kb) of Hamiltonian slot `slot` (6-11: matrix
// slot / 2 = 3 Heh, 4 Hee_o, 5 Hhh_o, orientation slot & 1), rows of the
// 128-row tile at x0, from the complex64 buffer (stored transposed:
// buffer[k2][k1] = H[k1][k2], so orientation 1 reads the buffer as stored
// and orientation 0 its transpose), split as k_split splits, into dst[2]
// [chunk] in the chunk layout.  The four waves of the copying group take
// 512 items (tile row x, 8-k group kg), two per lane: for orientation 0
// the lanes of a wave read 64 neighbouring columns of each buffer row (8
// loads of 8 bytes per item); for orientation 1 with N even, 16 lanes read
// a row's 32 k as 16-byte pairs (N odd: four lanes per row, 8-byte loads).
// Values past N are zeros.
__device__ __forceinline__ void copy_h(uint16_t (*dst)[4 * 64 * 32], const psa_group &g, int N, int w, int slot,
                                       int x0, int kb, int wave) {
  const int lane = threadIdx.x & 63, mat = slot >> 1;
  const bool asis = (slot & 1) != 0;
  const float2 *H = reinterpret_cast<const float2 *>(
      reinterpret_cast<const float *>(mat == 3 ? g.Heh : (mat == 4 ? g.Heo : g.Hho)) + mat_off(g, w));
  if (asis && (N & 1) == 0) {
    // even N: 16 lanes per row, each a k pair (k, k + 1) as one 16-byte
    // load, four whole 256-byte row segments per instruction; a pair leaves
    // as one dword per plane
    const float4 *H4 = reinterpret_cast<const float4 *>(H);
    float4 u[8];
#pragma unroll
    for (int j = 0; j < 8; ++j) {
      const int row = x0 + wave * 32 + 4 * j + (lane >> 4), k = kb * 32 + 2 * (lane & 15);
      u[j] = H4[((long)min(row, N - 1) * N + min(k, N - 2)) >> 1];
      if (row >= N || k >= N) u[j] = make_float4(0.f, 0.f, 0.f, 0.f);
    }
#pragma unroll
    for (int j = 0; j < 8; ++j) {
      const int x = wave * 32 + 4 * j + (lane >> 4), kl = 2 * (lane & 15);
      uint32_t pl[4];
      split_pair(u[j].x, u[j].z, pl[0], pl[1]);
      split_pair(u[j].y, u[j].w, pl[2], pl[3]);
#pragma unroll
      for (int q = 0; q < 4; ++q) *reinterpret_cast<uint32_t *>(&dst[x >> 6][in_chunk(q, x, kl)]) = pl[q];
    }
    return;
  }
  float2 v[2][8];
#pragma unroll
  for (int it = 0; it < 2; ++it) {
    const int i = it * 256 + wave * 64 + lane;
    const int x = asis ? i >> 2 : i & 127, kg = asis ? i & 3 : i >> 7;
    const int row = x0 + x, k0 = kb * 32 + 8 * kg;
#pragma unroll
    for (int q = 0; q < 8; ++q) {
      const int k = k0 + q;
      const bool ok = row < N && k < N;
      const long ij = asis ? (long)min(row, N - 1) * N + min(k, N - 1) : (long)min(k, N - 1) * N + min(row, N - 1);
      v[it][q] = H[ij];
      if (!ok) v[it][q] = make_float2(0.f, 0.f);
    }
  }
#pragma unroll
  for (int it = 0; it < 2; ++it) {
    const int i = it * 256 + wave * 64 + lane;
    const int x = asis ? i >> 2 : i & 127, kg = asis ? i & 3 : i >> 7;
    u32x4 pl[4];
#pragma unroll
    for (int q = 0; q < 8; q += 2) {
      uint32_t rh, rl, ih, il;
      split_pair(v[it][q].x, v[it][q + 1].x, rh, rl);
      split_pair(v[it][q].y, v[it][q + 1].y, ih, il);
      pl[0][q >> 1] = rh;
      pl[1][q >> 1] = rl;
      pl[2][q >> 1] = ih;
      pl[3][q >> 1] = il;
    }
#pragma unroll
    for (int p = 0; p < 4; ++p) *reinterpret_cast<u32x4 *>(&dst[x >> 6][in_chunk(p, x, 8 * kg)]) = pl[p];
  }
}

__device__ __forceinline__ bf16x8 fread(const uint16_t *half, int pl, int row, int kc) {
  return __builtin_bit_cast(
      bf16x8, *reinterpret_cast<const u32x4 *>(&half[pl * 2048 + row * 32 + ((kc ^ ((row >> 1) & 3)) << 3)]));
}

__device__ __forceinline__ bf16x8 flip(bf16x8 x, uint32_t m) {
  u32x4 u = __builtin_bit_cast(u32x4, x);
  u ^= u32x4{m, m, m, m};
  return __builtin_bit_cast(bf16x8, u);
}

__device__ __forceinline__ fv4 mfma(bf16x8 a, bf16x8 b, fv4 c) {
  return __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, c, 0, 0, 0);
}

// acc[si][sj] += a[si] b[sj] over the three split products (hi hi, hi lo,
// lo hi); products on the same accumulator are 16 issues apart.
__device__ __forceinline__ void prod3(const bf16x8 (&ah)[4], const bf16x8 (&al)[4], const bf16x8 (&bh)[4],
                                      const bf16x8 (&bl)[4], fv4 (&acc)[4][4]) {
#pragma unroll
  for (int si = 0; si < 4; ++si)
#pragma unroll
    for (int sj = 0; sj < 4; ++sj) acc[si][sj] = mfma(ah[si], bh[sj], acc[si][sj]);
#pragma unroll
  for (int si = 0; si < 4; ++si)
#pragma unroll
    for (int sj = 0; sj < 4; ++sj) acc[si][sj] = mfma(ah[si], bl[sj], acc[si][sj]);
#pragma unroll
  for (int si = 0; si < 4; ++si)
#pragma unroll
    for (int sj = 0; sj < 4; ++sj) acc[si][sj] = mfma(al[si], bh[sj], acc[si][sj]);
}

// The products of one staged slab of a term into this wave's 4 x 4
// blocks.  The term's sign and conjugations are sign flips of the operand
// planes (exact): A' = sg conj?(A), B' = conj?(B), then
//   Re += Re A' Re B' - Im A' Im B',  Im += Re A' Im B' + Im A' Re B'.
// At most one A and one B plane pair (64 registers) are held at a time:
// Re A' Re B' (into Re), Re A' Im B' (Im), -Im A' Im B' (Re),
// Im A' Re B' (Im); each accumulator gets its products in that order.
// BAND 1 (A is a band slot) and 2 (B is): the products with the zero real
// planes of the band operand are skipped (they would add exact zeros).
template <int BAND>
__device__ __forceinline__ void slab(const Stage &st, uint32_t fre, uint32_t fim, uint32_t gim, fv4 (&aR)[4][4],
                                     fv4 (&aI)[4][4], int wm, int wn, int kl, int xl) {
  const uint16_t *A = st.a[wm], *B = st.b[wn];
  bf16x8 ah[4], al[4], bh[4], bl[4];
  auto readA = [&](int pl, uint32_t m) {
#pragma unroll
    for (int s = 0; s < 4; ++s) {
      ah[s] = flip(fread(A, pl, s * 16 + xl, kl), m);
      al[s] = flip(fread(A, pl + 1, s * 16 + xl, kl), m);
    }
  };
  auto readB = [&](int pl, uint32_t m) {
#pragma unroll
    for (int s = 0; s < 4; ++s) {
      bh[s] = flip(fread(B, pl, s * 16 + xl, kl), m);
      bl[s] = flip(fread(B, pl + 1, s * 16 + xl, kl), m);
    }
  };
  if constexpr (BAND == 0) {
    readA(0, fre);
    readB(0, 0u);
    prod3(ah, al, bh, bl, aR);
    __builtin_amdgcn_sched_barrier(0);
  }
  if constexpr (BAND != 1) {
    if constexpr (BAND == 2) readA(0, fre);
    readB(2, gim);
    prod3(ah, al, bh, bl, aI);
    __builtin_amdgcn_sched_barrier(0);
  } else {
    readB(2, gim);
  }
  readA(2, fim ^ 0x80008000u);  // -Im A'
  prod3(ah, al, bh, bl, aR);
  if constexpr (BAND != 2) {
    __builtin_amdgcn_sched_barrier(0);
#pragma unroll
    for (int s = 0; s < 4; ++s) {
      ah[s] = flip(ah[s], 0x80008000u);
      al[s] = flip(al[s], 0x80008000u);
    }
    readB(0, 0u);
    prod3(ah, al, bh, bl, aI);
  }
}

// ---------------------------------------------------------------------------
// The products, persistent and ping-pong.
// One workgroup of 8 waves per CU, two groups of 4 waves (the 2 x 2 layout
// of the tile each), two LDS stages (128 KB).  Both groups work on the same
// 128 x 128 tile: group g takes the K slabs s = g, g + 2, ...  In phase s
// the group of slab s multiplies from its stage while the other group
// copies slab s + 1 into its own stage and waits for it, so on every SIMD
// one wave issues MFMAs while the other one loads.  After the last slab the
// two partial sums are added in LDS (even slabs + odd slabs, fixed order)
// and all 512 threads run the FP64 epilogue.  Tiles are assigned
// statically: the tiles of wire w belong to XCD w % 8, listed wire by wire
// (the dp, dC, dD tiles of a wire together, so the tiles that read the same
// operand slabs run at the same time and meet in the XCD's L2); workgroup j
// of an XCD takes items j, j + P, j + 2P, ... of its list (P workgroups per
// XCD).  With 32 CUs per XCD and 4 tiles per equation, every CU gets one
// dp tile (six terms) and one dC and one dD tile (four terms each).
// ---------------------------------------------------------------------------
constexpr int PP_THREADS = 512;

constexpr size_t pp_lds_bytes() { return 2 * sizeof(Stage); }

// Item it of the list of XCD x: batch b (3 wire + equation) and tile index
// in the equation, or false past the list's end.
__device__ __forceinline__ bool pp_item(int x, int it, int nw, int tpb, int &b, int &tile) {
  const int nwx = nw > x ? (nw - x + 7) / 8 : 0;  // wires x, x + 8, ...
  if (it >= 3 * nwx * tpb) return false;
  const int u = it / (3 * tpb), r = it % (3 * tpb);
  tile = r % tpb;
  b = 3 * (x + 8 * u) + r / tpb;
  return true;
}

__device__ __forceinline__ void pp_tile(const RhsProb &p, int b, int m0, int n0, char *lds) {
  Stage *stg = reinterpret_cast<Stage *>(lds);
  const int N = p.g.N, tid = threadIdx.x, lane = tid & 63;
  const int wave = __builtin_amdgcn_readfirstlane(tid >> 6), grp = wave >> 2, wv4 = wave & 3;
  const int wm = wv4 >> 1, wn = wv4 & 1, kl = lane >> 4, xl = lane & 15;
  const int e = b % 3, w = p.w0 + b / 3;
  const int nt = p.nterms(b), KS = nk(N) / BK, S = nt * KS;
  const uint16_t *spl = reinterpret_cast<const uint16_t *>(p.g.spl);
  if (tid == 0 && e == 0 && m0 == 0 && n0 == 0 && p.g.wtag)
    p.g.wtag[2 * w] = p.mode <= 1 ? (uint64_t)(p.g.pB + mat_off(p.g, w)) : 0ull;
  fv4 aR[4][4], aI[4][4];
#pragma unroll
  for (int i = 0; i < 4; ++i)
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      aR[i][j] = fv4{0.f, 0.f, 0.f, 0.f};
      aI[i][j] = fv4{0.f, 0.f, 0.f, 0.f};
    }
  // slab s of this tile into stage st by the calling group (its four waves)
  // (the direct copies of the plane slots are issued before the register
  // loads of a Hamiltonian slot)
  auto copy = [&](Stage &st, int s) {
    const Term tm = term(e, s / KS);
    const int kb = s % KS;
    const bool ha = tm.sa >= 6 && tm.sa < 12, hb = tm.sb >= 6 && tm.sb < 12;
    if (tm.sa >= 12) copy_slab<true>(st.a, spl, N, w, tm.sa, m0, kb, wv4);
    else if (!ha) copy_slab<false>(st.a, spl, N, w, tm.sa, m0, kb, wv4);
    if (tm.sb >= 12) copy_slab<true>(st.b, spl, N, w, tm.sb, n0, kb, wv4);
    else if (!hb) copy_slab<false>(st.b, spl, N, w, tm.sb, n0, kb, wv4);
    if (ha) copy_h(st.a, p.g, N, w, tm.sa, m0, kb, wv4);
    if (hb) copy_h(st.b, p.g, N, w, tm.sb, n0, kb, wv4);
  };
  if (grp == 0) copy(stg[0], 0);
  __builtin_amdgcn_s_waitcnt(0);
  __syncthreads();
  // phases of term t; the terms with a band operand run in loops of their own
  auto run_term = [&](auto band_c, int t) {
    constexpr int BAND = decltype(band_c)::value;
    const Term tm = term(e, t);
    const uint32_t fre = tm.neg ? 0x80008000u : 0u, fim = (tm.neg != tm.ca) ? 0x80008000u : 0u;
    const uint32_t gim = tm.cb ? 0x80008000u : 0u;
    for (int kb = 0; kb < KS; ++kb) {
      const int s = t * KS + kb, c = s & 1;
      if (grp == c) {
        slab<BAND>(stg[c], fre, fim, gim, aR, aI, wm, wn, kl, xl);
      } else {
        if (s + 1 < S) copy(stg[1 - c], s + 1);
        __builtin_amdgcn_s_waitcnt(0);
      }
      __syncthreads();
    }
  };
  for (int t = 0; t < nt && t < 4; ++t) run_term(std::integral_constant<int, 0>{}, t);
  if (nt > 4) {
    run_term(std::integral_constant<int, 1>{}, 4);  // (i hbar Dh band) p
    run_term(std::integral_constant<int, 2>{}, 5);  // p (i hbar De band)
  }
  // even + odd slabs in LDS as [re/im][128 rows][128 columns]
  float *acc = reinterpret_cast<float *>(lds);
  auto put = [&](bool add) {
#pragma unroll
    for (int si = 0; si < 4; ++si)
#pragma unroll
      for (int sj = 0; sj < 4; ++sj)
#pragma unroll
        for (int v = 0; v < 4; ++v) {
          const int rr = wm * 64 + si * 16 + 4 * kl + v, cc = wn * 64 + sj * 16 + xl;
          float *re = &acc[rr * TILE + cc], *im = &acc[TILE * TILE + rr * TILE + cc];
          *re = add ? *re + aR[si][sj][v] : aR[si][sj][v];
          *im = add ? *im + aI[si][sj][v] : aI[si][sj][v];
        }
  };
  if (grp == 0) put(false);
  __syncthreads();
  if (grp == 1) put(true);
  __syncthreads();
  // the FP64 epilogue: a fixed column per thread, four rows at a time with
  // their FP64 operands loaded first
  const bool full = N % TILE == 0;
  const int col = n0 + (tid & (TILE - 1)), r0 = tid >> 7;  // 4 row phases
  const long o = mat_off(p.g, w);
  const double *X = (e == 0 ? p.g.pB : (e == 1 ? p.g.CB : p.g.DB)) + o;
  const double *H = e == 0 ? p.g.Heh : (e == 1 ? p.g.Heo : p.g.Hho);  // complex64 in this mode
  double *X1 = (e == 0 ? p.g.pA : (e == 1 ? p.g.CA : p.g.DA)) + o;
  double *out = (e == 0 ? p.g.pX : (e == 1 ? p.g.CX : p.g.DX)) + o;
  // deferred leapfrog average (psa_group.lazy): X1 = (A + B) / 2, written
  // back to A as this step's X1
  const bool lz = p.g.lazy && p.g.lazy[w];
  const RhsProb::Side cs = p.side(e, w, min(col, N - 1), true);
  auto rows = [&](auto leap) {
    constexpr bool LEAP = decltype(leap)::value;
    constexpr int EB = 4;
    for (int j0 = 0; j0 < TILE / 4; j0 += EB) {
      RhsProb::Pre q[EB];
#pragma unroll
      for (int u = 0; u < EB; ++u) {
        const int row = m0 + r0 + 4 * (j0 + u);
        const long ij = (long)row * N + col;
        q[u].x = ld(X, ij);
        const cd hv = ld_h32(H, o, ij);
        q[u].h = (e != 0 && row == col) ? cd{0.0, 0.0} : hv;
        if constexpr (LEAP) {
          q[u].x1 = ld(X1, ij);
          if (lz) q[u].x1 = cd{(q[u].x1.re + q[u].x.re) / 2.0, (q[u].x1.im + q[u].x.im) / 2.0};
        }
      }
#pragma unroll
      for (int u = 0; u < EB; ++u) {
        const int rr = r0 + 4 * (j0 + u), row = m0 + rr, cc = tid & (TILE - 1);
        const RhsProb::Side rs = p.side(e, w, row, false);
        const cd d = p.deriv(e, (double)acc[rr * TILE + cc], (double)acc[TILE * TILE + rr * TILE + cc], q[u], cs, rs);
        if constexpr (LEAP) {
          st(out, (long)row * N + col, cd{q[u].x1.re + d.re * p.dt * 2.0, q[u].x1.im + d.im * p.dt * 2.0});
          if (lz) st(X1, (long)row * N + col, q[u].x1);
        } else
          st(out, (long)row * N + col, d);
      }
    }
  };
  if (full && p.mode == 1) {
    rows(std::true_type{});
  } else if (full && p.mode == 0) {
    rows(std::false_type{});
  } else {
    for (int i = tid; i < TILE * TILE; i += PP_THREADS) {
      const int rr = i / TILE, cc = i % TILE;
      p.store(b, m0 + rr, n0 + cc, (double)acc[rr * TILE + cc], (double)acc[TILE * TILE + rr * TILE + cc]);
    }
  }
  __syncthreads();  // the next tile's first copy overwrites the accumulators
}

__global__ __launch_bounds__(PP_THREADS, 1) void k_rhs_pp(RhsProb p, int nw) {
  extern __shared__ char lds_raw[];
  const int nb = (p.g.N + TILE - 1) / TILE, tpb = nb * nb;
  const int x = blockIdx.x % 8, P = (gridDim.x - x + 7) / 8;  // workgroups of XCD x
  int b, tile;
  for (int it = blockIdx.x / 8; pp_item(x, it, nw, tpb, b, tile); it += P)
    pp_tile(p, b, (tile / nb) * TILE, (tile % nb) * TILE, lds_raw);
}

}  // namespace sp
}  // namespace psa
