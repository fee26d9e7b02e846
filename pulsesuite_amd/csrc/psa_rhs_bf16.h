// Split-BF16 right-hand sides (PSA_PREC_BF16X3) for gfx950.
//
// Every FP64 operand x of the RHS contractions is written once per step as
// hi = bf16(x), lo = bf16(x - hi) (relative residual about 2^-17), for the
// real and the imaginary part: four bf16 planes per matrix and orientation
// (k_split_bf16).  A real product a b is then a_hi b_hi + a_hi b_lo +
// a_lo b_hi (the lo*lo term, ~2^-16 relative, is dropped), each a
// v_mfma_f32_16x16x32_bf16 with FP32 accumulation; a complex product is
// four real ones, 12 MFMAs per 16x16x32 block.  The diagonal parts, the
// epilogue and the state stay in FP64 (RhsProb::store).
//
// Operand slots per wire (four bf16 planes, rows k-contiguous, blocked as
// described at spl_chunk):
// slot = 2 * matrix + orientation, matrix 0 p, 1 C_o, 2 D_o (diagonal
// zeroed), 3 Heh, 4 Hee_o, 5 Hhh_o; orientation 0 the matrix, 1 its
// transpose.  An A operand A(m, k) reads slot rows m; a B operand B(k, n)
// reads slot rows n (both k-contiguous, as the MFMA lanes take them:
// A[l%16][8(l/16)+e], B[8(l/16)+e][l%16]).
#pragma once

#include "psa_gemm.h"

namespace psa {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));

__device__ __forceinline__ uint16_t bf16_bits_rn(float f) {
  uint32_t u = __float_as_uint(f);
  u += 0x7FFFu + ((u >> 16) & 1u);
  return (uint16_t)(u >> 16);
}
__device__ __forceinline__ float bf16_to_f(uint16_t h) { return __uint_as_float((uint32_t)h << 16); }
__device__ __forceinline__ void split_bf16(double x, uint16_t &hi, uint16_t &lo) {
  hi = bf16_bits_rn((float)x);
  lo = bf16_bits_rn((float)(x - (double)bf16_to_f(hi)));
}

// Blocked operand layout: slot matrices padded to NR = 128 ceil(N/128) rows
// and NK = 32 ceil(N/32) columns (k), stored as [row block][k block]
// chunks of 4 planes x 64 rows x 32 k (16 KB, contiguous), so one K slab
// of one 64-row tile is a single contiguous chunk.  The padding stays zero.
__host__ __device__ inline int spl_nr(int N) { return (N + 127) / 128 * 128; }
__host__ __device__ inline int spl_nk(int N) { return (N + 31) / 32 * 32; }
__host__ __device__ inline long spl_slot_elems(int N) { return 4L * spl_nr(N) * spl_nk(N); }
__host__ __device__ inline long spl_chunk(int N, int w, int slot, int rb, int kb) {
  return ((long)w * 12 + slot) * spl_slot_elems(N) + ((long)rb * (spl_nk(N) / 32) + kb) * (4 * 64 * 32);
}
__device__ __forceinline__ long spl_index(int N, int w, int slot, int pl, int row, int k) {
  return spl_chunk(N, w, slot, row >> 6, k >> 5) + pl * (64 * 32) + (row & 63) * 32 + (k & 31);
}

// ---------------------------------------------------------------------------
// k_split_bf16: block (64x64 tile of the stored buffer, matrix, wire).
// Stored buffers: p, C, D as the matrix; Heh, Heo, Hho as its transpose.
// Loads are lane-contiguous along the buffer's rows (two rows per thread
// and pass); the bf16 planes go to an LDS tile of row pairs per column
// (dword = rows 2r, 2r+1; column stride 33 dwords), from which both slots
// are written with 16-byte stores of 8 consecutive k.  Groups of 8 k at or
// past the padded width NK are not written; values past N are written as
// zero (the padding).
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_split_bf16(psa_group g, int w0) {
  __shared__ uint32_t t[4][64][33];  // [plane][column j][row pair]
  const int N = g.N, NK = spl_nk(N), mat = blockIdx.z % 6, w = w0 + blockIdx.z / 6;
  const int r0 = blockIdx.y * 64, c0 = blockIdx.x * 64, tid = threadIdx.x;
  const long o = mat_off(g, w);
  const double *X = (mat == 0 ? g.pB : mat == 1 ? g.CB : mat == 2 ? g.DB
                     : mat == 3 ? g.Heh : mat == 4 ? g.Heo : g.Hho) + o;
  const bool zd = (mat == 1 || mat == 2);
  const bool stored_t = mat >= 3;  // buffer holds the transpose of the matrix
  uint16_t *spl = reinterpret_cast<uint16_t *>(g.spl);
  // slot receiving the buffer as stored, and the one receiving its transpose
  const int s_same = 2 * mat + (stored_t ? 1 : 0), s_tr = 2 * mat + (stored_t ? 0 : 1);
  {
    const int j = tid & 63, gj = c0 + j;
    cd v[8][2];
#pragma unroll
    for (int e = 0; e < 8; ++e) {
      const int rp = (tid >> 6) + 4 * e;
#pragma unroll
      for (int r = 0; r < 2; ++r) {
        const int gi = r0 + 2 * rp + r;
        v[e][r] = ld(X, (long)min(gi, N - 1) * N + min(gj, N - 1));
        if (gi >= N || gj >= N || (zd && gi == gj)) v[e][r] = cd{0.0, 0.0};
      }
    }
#pragma unroll
    for (int e = 0; e < 8; ++e) {
      const int rp = (tid >> 6) + 4 * e;
      uint16_t h0[4], h1[4];
      split_bf16(v[e][0].re, h0[0], h0[1]);
      split_bf16(v[e][0].im, h0[2], h0[3]);
      split_bf16(v[e][1].re, h1[0], h1[1]);
      split_bf16(v[e][1].im, h1[2], h1[3]);
#pragma unroll
      for (int pl = 0; pl < 4; ++pl) t[pl][j][rp] = (uint32_t)h0[pl] | ((uint32_t)h1[pl] << 16);
    }
  }
  __syncthreads();
  const int x = tid >> 2, kq = tid & 3;
  // the buffer as stored: row i = x, k = columns 16 kq .. 16 kq + 15
  {
    const int gi = r0 + x;
    if (gi < N) {
#pragma unroll
      for (int half = 0; half < 2; ++half) {
        const int k0 = c0 + 16 * kq + 8 * half;
        if (k0 >= NK) continue;
#pragma unroll
        for (int pl = 0; pl < 4; ++pl) {
          uint16_t hv[8];
#pragma unroll
          for (int q = 0; q < 8; ++q) hv[q] = (uint16_t)(t[pl][16 * kq + 8 * half + q][x >> 1] >> (16 * (x & 1)));
          *reinterpret_cast<u32x4 *>(&spl[spl_index(N, w, s_same, pl, gi, k0)]) = __builtin_bit_cast(u32x4, hv);
        }
      }
    }
  }
  // transposed: row j = x (a column of the buffer), k = rows 16 kq .. 16 kq + 15
  {
    const int gj = c0 + x;
    if (gj < N) {
#pragma unroll
      for (int half = 0; half < 2; ++half) {
        const int k0 = r0 + 16 * kq + 8 * half;
        if (k0 >= NK) continue;
#pragma unroll
        for (int pl = 0; pl < 4; ++pl) {
          u32x4 v;
#pragma unroll
          for (int q = 0; q < 4; ++q) v[q] = t[pl][x][8 * kq + 4 * half + q];
          *reinterpret_cast<u32x4 *>(&spl[spl_index(N, w, s_tr, pl, gj, k0)]) = v;
        }
      }
    }
  }
}

// One term of an equation: A and B slots (-1: the off-diagonal dephasing
// band, generated from dedh), conjugation and sign.
struct BTerm {
  int sa, sb;
  bool ca, cb;
  double sg;
};

__device__ __forceinline__ BTerm bterm(int e, int t) {
  switch (e * 8 + t) {
    case 0: return {10, 1, false, false, 1.0};   // dp:  Hhh_o p
    case 1: return {0, 8, false, false, 1.0};    //      p Hee_o^T
    case 2: return {7, 3, false, false, -1.0};   //     -Heh^T C_o
    case 3: return {5, 6, false, false, -1.0};   //     -D_o^T Heh^T
    case 4: return {-1, 1, false, false, 1.0};   //      (i hbar Dh band) p
    case 5: return {0, -1, false, false, 1.0};   //      p (i hbar De band)
    case 8: return {2, 8, false, false, 1.0};    // dC:  C_o Hee_o^T
    case 9: return {9, 3, false, false, -1.0};   //     -Hee_o^T C_o
    case 10: return {1, 6, true, false, 1.0};    //      p^H Heh^T
    case 11: return {6, 1, true, false, -1.0};   //     -conj(Heh) p
    case 16: return {4, 10, false, false, 1.0};  // dD:  D_o Hhh_o^T
    case 17: return {11, 5, false, false, -1.0}; //     -Hhh_o^T D_o
    case 18: return {0, 7, true, false, 1.0};    //      conj(p) Heh
    default: return {7, 0, true, false, -1.0};   //     -Heh^H p^T
  }
}

__device__ __forceinline__ bf16x8 bneg(bf16x8 x) {
  u32x4 u = __builtin_bit_cast(u32x4, x);
  u ^= u32x4{0x80008000u, 0x80008000u, 0x80008000u, 0x80008000u};
  return __builtin_bit_cast(bf16x8, u);
}

namespace t64 {

// ---------------------------------------------------------------------------
// k_rhs (64-tile variant): block = 64x64 output tile of one (wire, equation); 4 waves
// (2x2) of 32x32 = 2x2 MFMA blocks; K slabs of 32 staged in LDS (rows of
// 64 B, the 16-byte chunk index XOR-ed with (row >> 2) & 3 so the 16 rows
// a lane group reads, and the staging writes, fall on distinct banks).
// ---------------------------------------------------------------------------
constexpr int BK2 = 32, BLD2 = 32, TILE = 64;

struct BTile {
  uint16_t v[4][64][BLD2];  // [plane][row][k]
};


// Register fragment of one slab: 4 planes x one 16-byte chunk (8 k).
struct BFrag {
  u32x4 c[4];
  int row, kc;
};

// Load the slab [k0, k0+32) of rows [x0, x0+64) of a dense slot: one
// contiguous 16 KB chunk (x0, k0 are multiples of 64 and 32).
__device__ __forceinline__ void bload_dense(BFrag &f, const uint16_t *spl, int N, int w, int slot,
                                            int x0, int k0) {
  const int tid = threadIdx.x;
  f.row = tid >> 2;
  f.kc = (tid & 3) * 8;
  const uint16_t *src = spl + spl_chunk(N, w, slot, x0 >> 6, k0 >> 5) + f.row * 32 + f.kc;
#pragma unroll
  for (int pl = 0; pl < 4; ++pl) f.c[pl] = *reinterpret_cast<const u32x4 *>(src + pl * (64 * 32));
}

// The banded dephasing operand i hbar D[m - x][x] for rows x = x0.. and
// k = m (dedh table kind 1 = Dh as A, 0 = De as B; q = 0 excluded): purely
// imaginary, so the real planes are zero.
__device__ __forceinline__ void bload_band(BFrag &f, const psa_group &g, int w, int kind, int x0,
                                           int k0) {
  const int N = g.N, Q = 2 * N - 1, tid = threadIdx.x;
  f.row = tid >> 2;
  f.kc = (tid & 3) * 8;
  const int gx = x0 + f.row, gk = k0 + f.kc;
  const double *tab = g.dedh + ((long)w * 2 + kind) * (long)N * Q;
  const double *row = tab + (long)min(gx, N - 1) * Q;
  double raw[8];
#pragma unroll
  for (int e = 0; e < 8; ++e) raw[e] = row[min(max(gk + e - gx + N - 1, 0), Q - 1)];
  uint16_t hi[8], lo[8];
#pragma unroll
  for (int e = 0; e < 8; ++e) {
    const int m = gk + e;
    const double v = (gx < N && m < N && m != gx) ? g.hbar * raw[e] : 0.0;
    split_bf16(v, hi[e], lo[e]);
  }
  f.c[0] = u32x4{0u, 0u, 0u, 0u};
  f.c[1] = u32x4{0u, 0u, 0u, 0u};
  f.c[2] = __builtin_bit_cast(u32x4, hi);
  f.c[3] = __builtin_bit_cast(u32x4, lo);
}

__device__ __forceinline__ void bcommit(const BFrag &f, BTile &t) {
#pragma unroll
  for (int pl = 0; pl < 4; ++pl)
    *reinterpret_cast<u32x4 *>(&t.v[pl][f.row][((f.kc >> 3) ^ ((f.row >> 2) & 3)) << 3]) = f.c[pl];
}

__device__ __forceinline__ bf16x8 bfrag_read(const BTile &t, int pl, int row, int kq) {
  return __builtin_bit_cast(bf16x8,
                            *reinterpret_cast<const u32x4 *>(&t.v[pl][row][((kq >> 3) ^ ((row >> 2) & 3)) << 3]));
}


__device__ __forceinline__ fv4 mfma_bf16(bf16x8 a, bf16x8 b, fv4 c) {
  return __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, c, 0, 0, 0);
}

// The 12 (or 6, for a band operand) split products of one K slab of one
// term into this wave's 2x2 blocks of accumulators.  The term's sign and
// conjugations are put on the operands (exact bf16 sign flips):
// A' = sg conj?(A), B' = conj?(B), then
//   Re += Re A' Re B' - Im A' Im B',  Im += Re A' Im B' + Im A' Re B'.
// Each split product (hi hi, hi lo, lo hi) is issued over all eight
// accumulators before the next one, so dependent MFMAs are eight issues
// apart; every accumulator still receives its products in the order
// RR (or RI) hi hi, hi lo, lo hi, then II (or IR) hi hi, hi lo, lo hi.
__device__ __forceinline__ void bslab(const BTile &sA, const BTile &sB, const BTerm bt,
                                      fv4 (&aR)[2][2], fv4 (&aI)[2][2], int wm, int wn, int kl,
                                      int xl) {
  const bool nA = bt.sg < 0;
  const bool aRe0 = bt.sa < 0, bRe0 = bt.sb < 0;  // band operands are purely imaginary
  bf16x8 A[2][4], B[2][4], nAi[2][2];
#pragma unroll
  for (int s = 0; s < 2; ++s)
#pragma unroll
    for (int pl = 0; pl < 4; ++pl) {
      const bf16x8 x = bfrag_read(sA, pl, wm * 32 + s * 16 + xl, kl * 8);
      const bf16x8 y = bfrag_read(sB, pl, wn * 32 + s * 16 + xl, kl * 8);
      A[s][pl] = (nA != (pl >= 2 && bt.ca)) ? bneg(x) : x;
      B[s][pl] = (pl >= 2 && bt.cb) ? bneg(y) : y;
    }
#pragma unroll
  for (int s = 0; s < 2; ++s) {
    nAi[s][0] = bneg(A[s][2]);
    nAi[s][1] = bneg(A[s][3]);
  }
  // split product p: A plane offset (p == 2: lo), B plane offset (p == 1: lo)
  if (!aRe0 && !bRe0) {
#pragma unroll
    for (int p = 0; p < 3; ++p) {
      const int ha = p == 2, hb = p == 1;
#pragma unroll
      for (int si = 0; si < 2; ++si)
#pragma unroll
        for (int sj = 0; sj < 2; ++sj) {
          aR[si][sj] = mfma_bf16(A[si][ha], B[sj][hb], aR[si][sj]);      // Re Re
          aI[si][sj] = mfma_bf16(A[si][ha], B[sj][2 + hb], aI[si][sj]);  // Re Im
        }
    }
#pragma unroll
    for (int p = 0; p < 3; ++p) {
      const int ha = p == 2, hb = p == 1;
#pragma unroll
      for (int si = 0; si < 2; ++si)
#pragma unroll
        for (int sj = 0; sj < 2; ++sj) {
          aR[si][sj] = mfma_bf16(nAi[si][ha], B[sj][2 + hb], aR[si][sj]);  // -Im Im
          aI[si][sj] = mfma_bf16(A[si][2 + ha], B[sj][hb], aI[si][sj]);    // Im Re
        }
    }
  } else if (!aRe0) {  // B is a band: Re A Im B, -Im A Im B
#pragma unroll
    for (int p = 0; p < 3; ++p) {
      const int ha = p == 2, hb = p == 1;
#pragma unroll
      for (int si = 0; si < 2; ++si)
#pragma unroll
        for (int sj = 0; sj < 2; ++sj) {
          aI[si][sj] = mfma_bf16(A[si][ha], B[sj][2 + hb], aI[si][sj]);
          aR[si][sj] = mfma_bf16(nAi[si][ha], B[sj][2 + hb], aR[si][sj]);
        }
    }
  } else if (!bRe0) {  // A is a band: -Im A Im B, Im A Re B
#pragma unroll
    for (int p = 0; p < 3; ++p) {
      const int ha = p == 2, hb = p == 1;
#pragma unroll
      for (int si = 0; si < 2; ++si)
#pragma unroll
        for (int sj = 0; sj < 2; ++sj) {
          aR[si][sj] = mfma_bf16(nAi[si][ha], B[sj][2 + hb], aR[si][sj]);
          aI[si][sj] = mfma_bf16(A[si][2 + ha], B[sj][hb], aI[si][sj]);
        }
    }
  } else {  // both bands: -Im A Im B
#pragma unroll
    for (int p = 0; p < 3; ++p) {
      const int ha = p == 2, hb = p == 1;
#pragma unroll
      for (int si = 0; si < 2; ++si)
#pragma unroll
        for (int sj = 0; sj < 2; ++sj) aR[si][sj] = mfma_bf16(nAi[si][ha], B[sj][2 + hb], aR[si][sj]);
    }
  }
}

// One 64 x 64 output tile (batch b = 3 wire + equation, rows m0..,
// columns n0..); lds_raw holds lds_bytes() of operand tiles.
__device__ __forceinline__ void rhs_tile(const RhsProb &p, int b, int m0, int n0, char *lds_raw) {
  BTile *tiles = reinterpret_cast<BTile *>(lds_raw);  // [2 buffers][A, B]
  const psa_group &g = p.g;
  const int N = g.N, tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int wm = wave >> 1, wn = wave & 1;
  const int e = b % 3, w = p.w0 + b / 3;
  const int nt = p.nterms(b);
  const uint16_t *spl = reinterpret_cast<const uint16_t *>(g.spl);
  fv4 aR[2][2], aI[2][2];
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int j = 0; j < 2; ++j) {
      aR[i][j] = fv4{0, 0, 0, 0};
      aI[i][j] = fv4{0, 0, 0, 0};
    }
  auto fetch = [&](BFrag &fa, BFrag &fb, int t, int k0) {
    const BTerm bt = bterm(e, t);
    if (bt.sa >= 0) bload_dense(fa, spl, N, w, bt.sa, m0, k0);
    else bload_band(fa, g, w, 1, m0, k0);
    if (bt.sb >= 0) bload_dense(fb, spl, N, w, bt.sb, n0, k0);
    else bload_band(fb, g, w, 0, n0, k0);
  };
  // Flattened slabs s = t * KS + k0 / 32.  Two-slab register prefetch:
  // while slab s is multiplied from one LDS buffer, slab s+1 sits in the
  // other buffer or in registers and slab s+2 is in flight.
  const int KS = (N + BK2 - 1) / BK2, S = nt * KS;
  const int kl = lane >> 4, xl = lane & 15;
  BFrag f0a, f0b, f1a, f1b;
  fetch(f0a, f0b, 0, 0);
  if (S > 1) fetch(f1a, f1b, 1 / KS, (1 % KS) * BK2);
  bcommit(f0a, tiles[0]);
  bcommit(f0b, tiles[1]);
  if (S > 2) fetch(f0a, f0b, 2 / KS, (2 % KS) * BK2);
  __syncthreads();
  for (int s = 0; s < S; s += 2) {
    bslab(tiles[0], tiles[1], bterm(e, s / KS), aR, aI, wm, wn, kl, xl);
    if (s + 1 >= S) break;
    bcommit(f1a, tiles[2]);
    bcommit(f1b, tiles[3]);
    __syncthreads();
    if (s + 3 < S) fetch(f1a, f1b, (s + 3) / KS, ((s + 3) % KS) * BK2);
    bslab(tiles[2], tiles[3], bterm(e, (s + 1) / KS), aR, aI, wm, wn, kl, xl);
    if (s + 2 >= S) break;
    bcommit(f0a, tiles[0]);
    bcommit(f0b, tiles[1]);
    __syncthreads();
    if (s + 4 < S) fetch(f0a, f0b, (s + 4) / KS, ((s + 4) % KS) * BK2);
  }
#pragma unroll
  for (int si = 0; si < 2; ++si)
#pragma unroll
    for (int sj = 0; sj < 2; ++sj)
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        const int row = m0 + wm * 32 + si * 16 + 4 * kl + v;
        const int col = n0 + wn * 32 + sj * 16 + xl;
        p.store(b, row, col, (double)aR[si][sj][v], (double)aI[si][sj][v]);
      }
}

__global__ __launch_bounds__(256) void k_rhs(RhsProb p) {
  extern __shared__ char lds_raw[];
  // XCD-aware order: the three equations of wire j run on XCD j % 8 (they
  // share the wire's operand planes in that XCD's L2) when the wire count
  // is a multiple of 8; consecutive workgroups go round robin over XCDs.
  const unsigned gx = gridDim.x, gy = gridDim.y, tpb = gx * gy;
  unsigned lin = blockIdx.x + gx * (blockIdx.y + gy * blockIdx.z);
  int b;
  if (gridDim.z % 24 == 0 && p.g.tune[2] == 0) {
    const unsigned xcd = lin % 8, q = lin / 8, per = 3 * tpb;
    b = 3 * (xcd + 8 * (q / per)) + (q % per) / tpb;
    lin = q % tpb;
  } else {
    b = lin / tpb;
    lin %= tpb;
  }
  rhs_tile(p, b, (lin / gx) * GBM, (lin % gx) * GBN, lds_raw);
}

constexpr size_t lds_bytes() { return 4 * sizeof(BTile); }

}  // namespace t64

namespace t128 {

// ---------------------------------------------------------------------------
// k_rhs (128-tile variant): block = 128x128 output tile of one (wire, equation); 4 waves
// (2x2) of 64x64 = 4x4 MFMA blocks each; K slabs of 32 staged in LDS (two
// stages, 64 KB each).  LDS rows are 64 B with the 16-byte chunk index
// XOR-ed with (row >> 2) & 3, which keeps both the staging writes (4 rows x
// 4 chunks per 16 lanes) and the MFMA operand reads (16 rows x 1 chunk) on
// distinct banks.  The epilogue stages the accumulators through LDS and
// stores in row-major order.
// ---------------------------------------------------------------------------
constexpr int BT = 128, BK2 = 32, TILE = BT;

struct BStage {
  uint16_t a[4][BT][BK2];  // [plane][row][k]
  uint16_t b[4][BT][BK2];
};

__device__ __forceinline__ int bswz(int row, int chunk) { return chunk ^ ((row >> 2) & 3); }


// Register fragment of one slab of one operand: this thread's 8 16-byte
// chunks, piece j = plane (j >> 1), row (tid >> 2) + 64 (j & 1), chunk tid & 3.
struct BFrag {
  u32x4 c[8];
};

// Rows [x0, x0+128) of a dense slot, slab [k0, k0+32): two contiguous 16 KB
// chunks of the blocked layout (x0, k0 multiples of 128 and 32).
__device__ __forceinline__ void bload_dense(BFrag &f, const uint16_t *spl, int N, int w, int slot,
                                            int x0, int k0) {
  const int tid = threadIdx.x;
  const int off = (tid >> 2) * 32 + (tid & 3) * 8;
#pragma unroll
  for (int h = 0; h < 2; ++h) {
    const uint16_t *src = spl + spl_chunk(N, w, slot, (x0 >> 6) + h, k0 >> 5) + off;
#pragma unroll
    for (int pl = 0; pl < 4; ++pl) f.c[pl * 2 + h] = *reinterpret_cast<const u32x4 *>(src + pl * (64 * 32));
  }
}

// The banded dephasing operand i hbar D[m - x][x] for rows x and k = m
// (dedh table kind 1 = Dh as A, 0 = De as B; q = 0 excluded): purely
// imaginary, so the real planes are zero.
__device__ __forceinline__ void bload_band(BFrag &f, const psa_group &g, int w, int kind, int x0,
                                           int k0) {
  const int N = g.N, Q = 2 * N - 1, tid = threadIdx.x;
  const double *tab = g.dedh + ((long)w * 2 + kind) * (long)N * Q;
  const int gk = k0 + (tid & 3) * 8;
#pragma unroll
  for (int h = 0; h < 2; ++h) {
    const int gx = x0 + (tid >> 2) + 64 * h;
    // unconditional loads from clamped in-range addresses, masked after, so
    // the eight loads are in flight together (no wait per element)
    const double *row = tab + (long)min(gx, N - 1) * Q;
    double raw[8];
#pragma unroll
    for (int e = 0; e < 8; ++e) raw[e] = row[min(max(gk + e - gx + N - 1, 0), Q - 1)];
    uint16_t hi[8], lo[8];
#pragma unroll
    for (int e = 0; e < 8; ++e) {
      const int m = gk + e;
      const double v = (gx < N && m < N && m != gx) ? g.hbar * raw[e] : 0.0;
      split_bf16(v, hi[e], lo[e]);
    }
    f.c[0 * 2 + h] = u32x4{0u, 0u, 0u, 0u};
    f.c[1 * 2 + h] = u32x4{0u, 0u, 0u, 0u};
    f.c[2 * 2 + h] = __builtin_bit_cast(u32x4, hi);
    f.c[3 * 2 + h] = __builtin_bit_cast(u32x4, lo);
  }
}

__device__ __forceinline__ void bcommit(const BFrag &f, uint16_t (*t)[BT][BK2]) {
  const int tid = threadIdx.x, chunk = tid & 3;
#pragma unroll
  for (int j = 0; j < 8; ++j) {
    const int pl = j >> 1, row = (tid >> 2) + 64 * (j & 1);
    *reinterpret_cast<u32x4 *>(&t[pl][row][bswz(row, chunk) * 8]) = f.c[j];
  }
}

__device__ __forceinline__ bf16x8 bread(const uint16_t (*t)[BT][BK2], int pl, int row, int kc) {
  return __builtin_bit_cast(bf16x8, *reinterpret_cast<const u32x4 *>(&t[pl][row][bswz(row, kc) * 8]));
}


// acc[si][sj] += (sign) A_x B_y over this wave's 4x4 blocks with the three
// split products (hi hi, hi lo, lo hi), issued so that products on the same
// accumulator are 16 issues apart.
__device__ __forceinline__ void bpart(const BStage &st, int pa, int pb, bool neg, fv4 (&acc)[4][4],
                                      int wm, int wn, int kl, int xl) {
  bf16x8 ah[4], al[4], bh[4], bl[4];
#pragma unroll
  for (int s = 0; s < 4; ++s) {
    const int ra = wm * 64 + s * 16 + xl, rb = wn * 64 + s * 16 + xl;
    ah[s] = bread(st.a, pa, ra, kl);
    al[s] = bread(st.a, pa + 1, ra, kl);
    bh[s] = bread(st.b, pb, rb, kl);
    bl[s] = bread(st.b, pb + 1, rb, kl);
    if (neg) {
      ah[s] = bneg(ah[s]);
      al[s] = bneg(al[s]);
    }
  }
#pragma unroll
  for (int si = 0; si < 4; ++si)
#pragma unroll
    for (int sj = 0; sj < 4; ++sj)
      acc[si][sj] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah[si], bh[sj], acc[si][sj], 0, 0, 0);
#pragma unroll
  for (int si = 0; si < 4; ++si)
#pragma unroll
    for (int sj = 0; sj < 4; ++sj)
      acc[si][sj] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah[si], bl[sj], acc[si][sj], 0, 0, 0);
#pragma unroll
  for (int si = 0; si < 4; ++si)
#pragma unroll
    for (int sj = 0; sj < 4; ++sj)
      acc[si][sj] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(al[si], bh[sj], acc[si][sj], 0, 0, 0);
}

// The split products of one K slab of one term.
__device__ __forceinline__ void bslab(const BStage &st, const BTerm bt, fv4 (&aR)[4][4], fv4 (&aI)[4][4],
                                      int wm, int wn, int kl, int xl) {
  // signs of the four real products (conj flips the imaginary part)
  const double cA = bt.ca ? -1.0 : 1.0, cB = bt.cb ? -1.0 : 1.0;
  const bool nRR = bt.sg < 0, nII = -bt.sg * cA * cB < 0, nRI = bt.sg * cB < 0, nIR = bt.sg * cA < 0;
  const bool aRe0 = bt.sa < 0, bRe0 = bt.sb < 0;  // band operands are purely imaginary
  if (!aRe0 && !bRe0) bpart(st, 0, 0, nRR, aR, wm, wn, kl, xl);  // Re A Re B
  if (!aRe0) bpart(st, 0, 2, nRI, aI, wm, wn, kl, xl);           // Re A Im B
  bpart(st, 2, 2, nII, aR, wm, wn, kl, xl);                      // -Im A Im B
  if (!bRe0) bpart(st, 2, 0, nIR, aI, wm, wn, kl, xl);           // Im A Re B
}

__global__ __launch_bounds__(256) void k_rhs(RhsProb p) {
  extern __shared__ char lds_raw[];
  BStage *stage = reinterpret_cast<BStage *>(lds_raw);  // [2]
  const psa_group &g = p.g;
  const int N = g.N, tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int wm = wave >> 1, wn = wave & 1;
  // XCD-aware order: the three equations of wire j run on XCD j % 8 (they
  // share the wire's operand planes in that XCD's L2) when the wire count
  // is a multiple of 8; consecutive workgroups go round robin over XCDs.
  const unsigned gx = gridDim.x, gy = gridDim.y, tpb = gx * gy;
  unsigned lin = blockIdx.x + gx * (blockIdx.y + gy * blockIdx.z);
  int b;
  if (gridDim.z % 24 == 0 && g.tune[2] == 0) {
    const unsigned xcd = lin % 8, q = lin / 8, per = 3 * tpb;
    b = 3 * (xcd + 8 * (q / per)) + (q % per) / tpb;
    lin = q % tpb;
  } else {
    b = lin / tpb;
    lin %= tpb;
  }
  const int m0 = (lin / gx) * BT, n0 = (lin % gx) * BT;
  const int e = b % 3, w = p.w0 + b / 3;
  const int nt = p.nterms(b);
  const uint16_t *spl = reinterpret_cast<const uint16_t *>(g.spl);
  fv4 aR[4][4], aI[4][4];
#pragma unroll
  for (int i = 0; i < 4; ++i)
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      aR[i][j] = fv4{0, 0, 0, 0};
      aI[i][j] = fv4{0, 0, 0, 0};
    }
  auto fetch = [&](BFrag &fa, BFrag &fb, int s, int KS) {
    const BTerm bt = bterm(e, s / KS);
    const int k0 = (s % KS) * BK2;
    if (bt.sa >= 0) bload_dense(fa, spl, N, w, bt.sa, m0, k0);
    else bload_band(fa, g, w, 1, m0, k0);
    if (bt.sb >= 0) bload_dense(fb, spl, N, w, bt.sb, n0, k0);
    else bload_band(fb, g, w, 0, n0, k0);
  };
  // flattened slabs s = term * KS + k0 / 32, one-slab register prefetch
  const int KS = (N + BK2 - 1) / BK2, S = nt * KS;
  const int kl = lane >> 4, xl = lane & 15;
  BFrag fa, fb;
  fetch(fa, fb, 0, KS);
  bcommit(fa, stage[0].a);
  bcommit(fb, stage[0].b);
  __syncthreads();
  for (int s = 0; s < S; ++s) {
    if (s + 1 < S) fetch(fa, fb, s + 1, KS);
    bslab(stage[s & 1], bterm(e, s / KS), aR, aI, wm, wn, kl, xl);
    if (s + 1 >= S) break;
    bcommit(fa, stage[(s + 1) & 1].a);
    bcommit(fb, stage[(s + 1) & 1].b);
    __syncthreads();
  }
  // epilogue: accumulators -> LDS [re/im][128][128] f32 -> row-major stores
  __syncthreads();
  float *acc = reinterpret_cast<float *>(lds_raw);
#pragma unroll
  for (int si = 0; si < 4; ++si)
#pragma unroll
    for (int sj = 0; sj < 4; ++sj)
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        const int r = wm * 64 + si * 16 + 4 * kl + v, c = wn * 64 + sj * 16 + xl;
        acc[r * BT + c] = aR[si][sj][v];
        acc[BT * BT + r * BT + c] = aI[si][sj][v];
      }
  __syncthreads();
#pragma unroll 1
  for (int i = tid; i < BT * BT; i += 256) {
    const int r = i / BT, c = i % BT;
    p.store(b, m0 + r, n0 + c, (double)acc[i], (double)acc[BT * BT + i]);
  }
}

constexpr size_t lds_bytes() { return 2 * sizeof(BStage); }

}  // namespace t128

}  // namespace psa
