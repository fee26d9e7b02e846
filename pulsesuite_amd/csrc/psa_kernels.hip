// pulsesuite_amd: HIP kernels and C ABI for the quantum-wire SBE step.
// Target: gfx950 (MI355X).  See include/pulsesuite_amd.h for the ABI.
#include <hip/hip_runtime.h>

#include <cstdio>
#include <cstring>
#include <algorithm>
#include <map>
#include <mutex>
#include <type_traits>
#include <vector>
#include <string>

#include "psa_gemm.h"
#include "psa_rhs_split.h"
#include "psa_offdiag.h"
#include "psa_calch.h"
#include "psa_pstd.h"
#include "psa_modules.h"

using namespace psa;

namespace {

thread_local std::string g_err;

int fail(int code, const char *msg) {
  g_err = msg;
  return code;
}

int hip_check(hipError_t e, const char *what) {
  if (e == hipSuccess) return PSA_OK;
  g_err = std::string(what) + ": " + hipGetErrorString(e);
  return PSA_E_HIP;
}

#define PSA_LAUNCH_CHECK(what)                              \
  do {                                                      \
    int _rc = hip_check(hipGetLastError(), what);           \
    if (_rc) return _rc;                                    \
  } while (0)

constexpr double C0_LIGHT = 299792458.0;

int check_group(const psa_group *g, int w0, int nw) {
  if (!g) return fail(PSA_E_INVALID, "null group");
  if (g->N < 4 || g->N > 1023) return fail(PSA_E_INVALID, "N must be in [4, 1023]");
  if (g->Nr != 2 * g->N + 2) return fail(PSA_E_INVALID, "Nr must be 2N+2");
  if (g->Nw < 1) return fail(PSA_E_INVALID, "Nw must be >= 1");
  if (w0 < 0 || nw < 1 || w0 + nw > g->Nw) return fail(PSA_E_INVALID, "wire range outside the group");
  if (g->prec < 0 || g->prec > 2) return fail(PSA_E_INVALID, "unknown precision mode");
  return PSA_OK;
}

}  // namespace

// ===========================================================================
// Vint table (coulomb.py:38-177, 502-565): crude _K03_jit
// ===========================================================================
__device__ double k03_crude(double x) {
  if (x > 100.0) return 0.0;
  if (x < 1e-10) return 23.0;
  if (x < 2.0) {
    const double x2 = x * x;
    double res = -log(0.5 * x) - 0.5772156649015329;
    double term = 1.0;
    for (int k = 1; k < 20; ++k) {
      term = term * 0.25 * x2 / (double)(k * k);
      res = res + term * (1.0 / k - 0.5772156649015329);
      if (fabs(term) < 1e-15) break;
    }
    return res;
  }
  return sqrt(PI / (2.0 * x)) * exp(-x);
}

// one block per (component, Qy index); threads split the (i, j) double sum
__global__ __launch_bounds__(256) void k_vint(const double *Qy, int NQ, const double *y, int Ny,
                                              double ae, double ah, double Delta0, double pref,
                                              double *out) {
  __shared__ double sh[8];
  const int kq = blockIdx.x, c = blockIdx.y;
  const double a1 = (c == 2) ? ah : ae;  // (eh: ae,ah) (ee: ae,ae) (hh: ah,ah)
  const double a2 = (c == 1) ? ae : ah;
  const int N1 = Ny / 4, N2 = 3 * Ny / 4, M = N2 - N1 + 1;
  const double kmin = (a1 + a2) / 4.0;
  const double dk = fmax(fabs(Qy[kq]), kmin);
  const double dy = y[1] - y[0];
  const double mult = a1 * a2 / PI * dy * dy;
  double acc = 0.0;
  for (int t = threadIdx.x; t < M * M; t += blockDim.x) {
    const int i = N1 + t / M, j = N1 + t % M;
    const double yi = y[i], yj = y[j];
    const double r = sqrt((yi - yj) * (yi - yj) + Delta0 * Delta0);
    const double ex = exp(-(a1 * yi) * (a1 * yi) - (a2 * yj) * (a2 * yj));
    acc += ex * mult * k03_crude(dk * r);
  }
  const double s = block_sum(acc, sh);
  if (threadIdx.x == 0) out[(long)c * NQ + kq] = pref * s;
}

// ===========================================================================
// prep (SBEs.py:5173-5226 minus CalcH), in three kernels:
//  k_prep_a   one block per wire: occupations, screening (CalcScreenedArrays,
//             coulomb.py:1195-1230; Eps1D 1172-1193)
//  k_diagsums diagonal sums of C, D for the Hartree terms
//  k_prep_b   one wave per (wire, k): CalcGammaE/H (dephasing.py:28-103,
//             Vxx2 362-378) and SpontEmission (emission.py:94-198)
// ===========================================================================
__global__ __launch_bounds__(256) void k_prep_a(psa_group g, int w0) {
  __shared__ double sh[8];
  const int w = w0 + blockIdx.x, N = g.N, tid = threadIdx.x;
  const long o = mat_off(g, w);
  const double *C2 = g.CB + o, *D2 = g.DB + o;
  double *pr = g.prep + (long)w * PSA_PREP_STRIDE(N);
  // wire-fastest copies for k_mb and the off-diagonal dephasing products:
  // |ne|, |nh| of the checkout level and VC(d)^2 per channel
  const int W = g.Nw;
  double *vsq = g.wmb + 6L * N * W, *aocc = g.wmb + 9L * N * W;
  double part = 0.0;
  for (int k = tid; k < N; k += blockDim.x) {
    const cd a = ld(C2, (long)k * N + k), b = ld(D2, (long)k * N + k);
    st(pr + PSA_PREP_NE, k, a);
    st(pr + PSA_PREP_NH(N), k, b);
    const double ane = cabs_(a), anh = cabs_(b);
    pr[PSA_PREP_ANE(N) + k] = ane;
    pr[PSA_PREP_ANH(N) + k] = anh;
    aocc[(long)k * W + w] = ane;
    aocc[(long)(N + k) * W + w] = anh;
    part += a.re + b.re;
  }
  const double tot = block_sum(part, sh);
  const bool screened = (g.flags & PSA_F_SCREENED) != 0;
  double n1 = tot / 2.0 / g.L;
  n1 = fmin(n1, g.n1d_max);
  double *E1D = pr + PSA_PREP_E1D(N), *vc = pr + PSA_PREP_VC(N);
  for (int d = tid; d < N; d += blockDim.x) {
    double e = 1.0;
    if (screened) {
      const double qe = g.qeh[d], qh = g.qeh[N + d];
      e = 1.0 - g.chi[d] * 2.0 * log(fabs((qe - PI * n1) / (qe + n1))) -
          g.chi[N + d] * 2.0 * log(fabs((qh - PI * n1) / (qh + n1)));
    }
    E1D[d] = e;
    for (int c = 0; c < 3; ++c) {
      const double v = screened ? g.vtab[c * N + d] / e : g.vtab[c * N + d];
      vc[c * N + d] = v;
      vsq[((long)c * N + d) * W + w] = v * v;
    }
  }
}

// mode 0: dsC[s] = sum_k C2[k][k+s], dsD likewise (prep);
// mode 1: sums of the new level for the sources: dcv sum_kh (-1)^kh p[kh][kh+s],
//         dcv sum p[kh][kh+s], sum C3[k][k+s], sum conj(D3[k][k+s]).
// Block (wire, 64 diagonals): lane -> diagonal (coalesced along the rows),
// wave -> every fourth k; the four partial sums are added in wave order.
// Block = 64 diagonals (one per lane) x DS_WAVES waves over the rows k (k
// congruent to the wave index); each wave keeps four rows' loads in flight
// (partial sums per row slot, combined in a fixed order), then the waves'
// sums are added in a fixed order.
constexpr int DS_WAVES = 8;
__global__ __launch_bounds__(64 * DS_WAVES) void k_diagsums(psa_group g, int w0, int mode) {
  __shared__ cd red[4][DS_WAVES][64];
  const int w = w0 + blockIdx.x, N = g.N, Q = 2 * N - 1;
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  const int si = blockIdx.y * 64 + lane;
  const int s = si - (N - 1);
  const int k0 = s < 0 ? -s : 0, k1 = si < Q ? (s < 0 ? N : N - s) : 0;
  // rows run uniformly over the block's lanes (the union of their ranges),
  // so the lanes of one load read consecutive elements of one row
  const int s0 = blockIdx.y * 64 - (N - 1);
  const int kmin = max(0, -(s0 + 63)), kmax = N - max(0, s0);
  const long o = mat_off(g, w);
  constexpr int ST = DS_WAVES;
  if (mode == 0 && g.wtag && g.wtag[2 * w + 1] == (uint64_t)(g.CB + o)) {
    // the previous step's sources summed this level already (mode 1 sums
    // C3 and conj(D3) in the same order): copy them (block-uniform)
    if (wv == 0 && si < Q) {
      const double *ds = g.src + (long)w * PSA_SRC_STRIDE(N, g.Nr) + PSA_SRC_DS(g.Nr);
      double *pr = g.prep + (long)w * PSA_PREP_STRIDE(N);
      st(pr + PSA_PREP_DSC(N), si, ld(ds, 2 * Q + si));
      st(pr + PSA_PREP_DSD(N), si, cconj(ld(ds, 3 * Q + si)));
    }
    return;
  }
  cd a0 = {0, 0}, a1 = {0, 0}, a2 = {0, 0}, a3 = {0, 0};
  if (mode == 0) {
    const double *C2 = g.CB + o, *D2 = g.DB + o;
    cd pc[4] = {}, pd[4] = {};
    for (int kb = kmin + wv; kb < kmax; kb += 4 * ST) {
      cd vc_[4], vd_[4];
#pragma unroll
      for (int u = 0; u < 4; ++u) {
        const int k = min(max(kb + ST * u, k0), max(k1 - 1, k0));
        const long ix = k1 > k0 ? (long)k * N + k + s : 0;
        vc_[u] = ld(C2, ix);
        vd_[u] = ld(D2, ix);
      }
#pragma unroll
      for (int u = 0; u < 4; ++u) {
        const int k = kb + ST * u;
        if (k >= k0 && k < k1) {
          pc[u] = cadd(pc[u], vc_[u]);
          pd[u] = cadd(pd[u], vd_[u]);
        }
      }
    }
    a0 = cadd(cadd(pc[0], pc[1]), cadd(pc[2], pc[3]));
    a1 = cadd(cadd(pd[0], pd[1]), cadd(pd[2], pd[3]));
  } else {
    const double *p3 = g.pX + o, *C3 = g.CX + o, *D3 = g.DX + o;
    cd q0[4] = {}, q1[4] = {}, q2[4] = {}, q3[4] = {};
    for (int kb = kmin + wv; kb < kmax; kb += 4 * ST) {
      cd vp[4], vc_[4], vd_[4];
#pragma unroll
      for (int u = 0; u < 4; ++u) {
        const int k = min(max(kb + ST * u, k0), max(k1 - 1, k0));
        const long ix = k1 > k0 ? (long)k * N + k + s : 0;
        vp[u] = ld(p3, ix);
        vc_[u] = ld(C3, ix);
        vd_[u] = ld(D3, ix);
      }
#pragma unroll
      for (int u = 0; u < 4; ++u) {
        const int k = kb + ST * u;
        if (k >= k0 && k < k1) {
          q1[u] = cadd(q1[u], vp[u]);
          q0[u] = (k & 1) ? csub(q0[u], vp[u]) : cadd(q0[u], vp[u]);
          q2[u] = cadd(q2[u], vc_[u]);
          q3[u] = cadd(q3[u], cconj(vd_[u]));
        }
      }
    }
    a0 = cadd(cadd(q0[0], q0[1]), cadd(q0[2], q0[3]));
    a1 = cadd(cadd(q1[0], q1[1]), cadd(q1[2], q1[3]));
    a2 = cadd(cadd(q2[0], q2[1]), cadd(q2[2], q2[3]));
    a3 = cadd(cadd(q3[0], q3[1]), cadd(q3[2], q3[3]));
  }
  red[0][wv][lane] = a0;
  red[1][wv][lane] = a1;
  red[2][wv][lane] = a2;
  red[3][wv][lane] = a3;
  __syncthreads();
  if (wv != 0 || si >= Q) return;
  cd t[4];
#pragma unroll
  for (int j = 0; j < 4; ++j) {
    t[j] = red[j][0][lane];
#pragma unroll
    for (int u = 1; u < DS_WAVES; ++u) t[j] = cadd(t[j], red[j][u][lane]);
  }
  if (mode == 0) {
    double *pr = g.prep + (long)w * PSA_PREP_STRIDE(N);
    st(pr + PSA_PREP_DSC(N), si, t[0]);
    st(pr + PSA_PREP_DSD(N), si, t[1]);
  } else {
    double *ds = g.src + (long)w * PSA_SRC_STRIDE(N, g.Nr) + PSA_SRC_DS(g.Nr);
    st(ds, si, cscale(t[0], g.dcv));
    st(ds, Q + si, cscale(t[1], g.dcv));
    st(ds, 2 * Q + si, t[2]);
    st(ds, 3 * Q + si, t[3]);
    // src now holds the diagonal sums of C and D of this level (read by the
    // next step's prep once the level has become its level 3)
    if (blockIdx.y == 0 && lane == 0 && g.wtag) g.wtag[2 * w + 1] = (uint64_t)(g.CX + o);
  }
}
// Block = PB_WAVES rates k of one wire (a wave per k); the wire's
// occupations (real parts) and screened interactions are staged in LDS once
// per block, so the index-table gathers read LDS.
constexpr int PB_WAVES = 16;
__global__ __launch_bounds__(64 * PB_WAVES) void k_prep_b(psa_group g, int w0) {
  __shared__ double s_ne[1024], s_nh[1024], s_vc[3 * 1024];
  const int w = w0 + blockIdx.x, N = g.N, lane = threadIdx.x & 63;
  const int k = blockIdx.y * PB_WAVES + (threadIdx.x >> 6);
  double *pr = g.prep + (long)w * PSA_PREP_STRIDE(N);
  {
    const double *nev = pr + PSA_PREP_NE, *nhv = pr + PSA_PREP_NH(N), *vcg = pr + PSA_PREP_VC(N);
    for (int i = threadIdx.x; i < N; i += blockDim.x) {
      s_ne[i] = nev[2 * i];
      s_nh[i] = nhv[2 * i];
    }
    for (int i = threadIdx.x; i < 3 * N; i += blockDim.x) s_vc[i] = vcg[i];
  }
  __syncthreads();
  if (k >= N) return;  // whole wave; only wave-level reductions below
  const double *vc = s_vc;
  double *GE = pr + PSA_PREP_GAME(N), *GH = pr + PSA_PREP_GAMH(N), *Rsp = pr + PSA_PREP_RSP(N);
  if (!(g.flags & PSA_F_DIAGDPH)) {
    if (lane == 0) {
      GE[k] = g.gam_e;
      GH[k] = g.gam_h;
    }
  } else {
    const long NN = (long)N * N;
    const int *kpq = g.didx, *kmq = g.didx + NN, *k1pmq = g.didx + 2 * NN, *k1p = g.didx + 3 * NN,
              *k1mq = g.didx + 4 * NN, *k1 = g.didx + 5 * NN;  // stored [k][q]
    const double ph = PI / g.hbar;
    double ge = 0.0, gh = 0.0;
    // unrolled so that the index loads of several q are in flight together
#pragma unroll 4
    for (int q = lane; q < N; q += 64) {
      const long qk = (long)k * N + q;
      const int iq = g.iq[q];
      const double Vee = vc[N + iq], Veh = vc[iq], Vhh = vc[2 * N + iq];
      const int i = kpq[qk];
      if (i >= 0 && i < N) {
        const double ni = s_ne[i], hi = s_nh[i];
        ge += ph * (Vee * Vee) * ni * (1.0 - ni) * fabs(g.xe[q]);
        gh += ph * (Vhh * Vhh) * hi * (1.0 - hi) * fabs(g.xh[q]);
      }
      const int c = kmq[qk];
      int a = k1pmq[qk], b = k1p[qk];
      if (a >= 0 && a < N && b >= 0 && b < N && c >= 0 && c < N) {
        const double na = s_nh[a], nb = s_nh[b], ec = s_ne[c];
        ge += ph * (Veh * Veh) * (na * (1.0 - nb) * ec + (1.0 - na) * nb * (1.0 - ec)) * fabs(g.xh[q]);
      }
      a = k1mq[qk];
      b = k1[qk];
      if (a >= 0 && a < N && b >= 0 && b < N && c >= 0 && c < N) {
        const double ea = s_ne[a], eb = s_ne[b], hc = s_nh[c];
        gh += ph * (Veh * Veh) * (ea * (1.0 - eb) * hc + (1.0 - ea) * eb * (1.0 - hc)) * fabs(g.xe[q]);
      }
    }
    ge = wave_sum(ge);
    gh = wave_sum(gh);
    if (lane == 0) {
      GE[k] = ge;
      GH[k] = gh;
    }
  }
  if (!(g.flags & PSA_F_RECOMB)) {
    if (lane == 0) Rsp[k] = 0.0;
    return;
  }
  // Ec (emission.py:133-171) with the lower-triangular VC
  const double veh0 = vc[0], vee0 = vc[N], vhh0 = vc[2 * N];
  double s1 = 0.0;
  for (int i = lane; i < N; i += 64) {
    const double ne = s_ne[i], nh = s_nh[i];
    const double vee_ik = i >= k ? vc[N + i - k] : 0.0;
    const double vhh_ik = i >= k ? vc[2 * N + i - k] : 0.0;
    s1 += ne * (vee0 - vee_ik) + nh * (vhh0 - vhh_ik);
    if (i != k) s1 += ne * veh0 - nh * veh0;
  }
  s1 = wave_sum(s1);
  const double Ek = g.gap + g.Ee[k] + g.Eh[k] + (s1 - veh0);
  const double rho_c = 1.0 / (C0_LIGHT * C0_LIGHT * C0_LIGHT * PI * PI * g.hbar * g.hbar * g.hbar);
  double acc = 0.0;
  for (int n = lane; n < g.Nhw; n += 64) {
    const double hw = g.homega[n] + Ek;
    acc += hw * (hw * hw * rho_c) * g.hsq[n];
  }
  acc = wave_sum(acc);
  if (lane == 0) Rsp[k] = acc * g.dhw;
}

// Hamiltonians without excitons (SBEs.py:826-835): Heh = Meh, diagonal energies
__global__ __launch_bounds__(256) void k_h_simple(psa_group g, int w0) {
  const int w = w0 + blockIdx.y, N = g.N;
  const long o = mat_off(g, w);
  const int wid = g.wire_id[w];
  const double yv = yw_of(wid), zs = sgn_pow(wid);
  const double *eq = g.exq + (long)w * 3 * 2 * (2 * N - 1);
  for (long ij = blockIdx.x * (long)blockDim.x + threadIdx.x; ij < (long)N * N; ij += (long)gridDim.x * blockDim.x) {
    const int k2 = ij / N, k1 = ij % N;  // ij indexes the transposed storage [k2][k1]
    cd meh = {0, 0};
    if (g.flags & PSA_F_OPTICS) {
      const int dd = k1 - k2 + N - 1;
      const cd ex = ld(eq, dd), ey = ld(eq, (2 * N - 1) + dd), ez = ld(eq, 2 * (2 * N - 1) + dd);
      const double xcv = (k2 & 1) ? -g.dcv : g.dcv;
      const double a1 = -g.ehint * xcv, a2 = -g.ehint * g.dcv, a3 = g.ehint * g.dcv;
      meh.re = a1 * ex.re + (a2 * ey.re) * yv + (a3 * ez.re) * zs;
      meh.im = a1 * ex.im + (a2 * ey.im) * yv + (a3 * ez.im) * zs;
    }
    st_h(g, g.Heh, o, ij, meh);
    st_h(g, g.Heo, o, ij, cd{0, 0});
    st_h(g, g.Hho, o, ij, cd{0, 0});
    if (k1 == k2) {
      double *hd = g.hdiag + (long)w * 4 * N;
      st(hd, k1, cd{g.Ee[k1] + g.gap, 0.0});
      st(hd, N + k1, cd{g.Eh[k1], 0.0});
    }
  }
}

// ===========================================================================
// Many-body Coulomb scattering (MBCE/MBCH, coulomb.py:956-1071) over the
// sparse lists of non-negligible Gaussian-delta entries.  List l holds
// entries (i, j, l) of the stored table, grouped (CSR) by the rate index it
// feeds (see the slot mapping below).  Block = (MB_KG target k, slot,
// MB_WC wires): the block stages the occupations and the two V^2 channels
// of its slot for its wires in LDS (rows of MB_WC wires, from the
// wire-fastest tables built by k_diag_gather / k_prep_a); a lane is
// (entry group, wire) and the eight lanes of a group read one 64-byte LDS
// row.  The sums are reduced over the block in a fixed order.
//   wmb + 4 N Nw: occ[2][N][Nw]  |Re ne3|, |Re nh3|
//   wmb + 6 N Nw: vsq[3][N][Nw]  VC(d)^2 for eh, ee, hh
// ===========================================================================
constexpr int MB_WC = 8, MB_KG = 4, MB_THREADS = 512;

// The three occupation factors of list L for entry (i, j, l): rows of the
// electron (E) or hole (H) table, complemented (1 - n) or not.
template <int L>
struct MbList;
// case 0: (1 - nh[j]) nh[j+l-i] ne[l]
template <> struct MbList<0> { static constexpr int c1 = 1, c2 = 1, c3 = 0; static constexpr bool x1 = 1, x2 = 0, x3 = 0; };
// case 1: (1 - ne[i]) (1 - nh[j]) nh[j+l-i]
template <> struct MbList<1> { static constexpr int c1 = 0, c2 = 1, c3 = 1; static constexpr bool x1 = 1, x2 = 1, x3 = 0; };
// case 2: (1 - ne[j]) ne[i+j-l] ne[l]
template <> struct MbList<2> { static constexpr int c1 = 0, c2 = 0, c3 = 0; static constexpr bool x1 = 1, x2 = 0, x3 = 0; };
// case 3: (1 - ne[i]) (1 - ne[j]) ne[i+j-l]
template <> struct MbList<3> { static constexpr int c1 = 0, c2 = 0, c3 = 0; static constexpr bool x1 = 1, x2 = 1, x3 = 0; };
// case 4: (1 - ne[i]) nh[j+l-i] ne[l]
template <> struct MbList<4> { static constexpr int c1 = 0, c2 = 1, c3 = 0; static constexpr bool x1 = 1, x2 = 0, x3 = 0; };
// case 5: (1 - ne[i]) (1 - nh[j]) ne[l]
template <> struct MbList<5> { static constexpr int c1 = 0, c2 = 1, c3 = 0; static constexpr bool x1 = 1, x2 = 1, x3 = 0; };
// case 6: (1 - nh[j]) nh[i+j-l] nh[l]
template <> struct MbList<6> { static constexpr int c1 = 1, c2 = 1, c3 = 1; static constexpr bool x1 = 1, x2 = 0, x3 = 0; };
// case 7: (1 - nh[i]) (1 - nh[j]) nh[i+j-l]
template <> struct MbList<7> { static constexpr int c1 = 1, c2 = 1, c3 = 1; static constexpr bool x1 = 1, x2 = 1, x3 = 0; };

template <int L>
__device__ __forceinline__ void mb_indices(uint32_t pk, int &a, int &c, int &r1, int &r2, int &r3) {
  const int i = pk & 1023, j = (pk >> 10) & 1023, l = (pk >> 20) & 1023;
  a = (L == 5) ? l : i;  // V(a, c): Veh with the index order swapped for list 5
  c = (L == 5) ? i : l;
  switch (L) {
    case 0: r1 = j; r2 = j + l - i; r3 = l; break;
    case 1: r1 = i; r2 = j; r3 = j + l - i; break;
    case 2: r1 = j; r2 = i + j - l; r3 = l; break;
    case 3: r1 = i; r2 = j; r3 = i + j - l; break;
    case 4: r1 = i; r2 = j + l - i; r3 = l; break;
    case 5: r1 = i; r2 = j; r3 = l; break;
    case 6: r1 = j; r2 = i + j - l; r3 = l; break;
    default: r1 = i; r2 = j; r3 = i + j - l; break;
  }
}

// This lane's share of list L, for wire c = lane & 7: the wave loads 64
// consecutive entries (one per lane, coalesced) and hands them round by
// round to its eight 8-lane groups (lane >> 3 picks the entry), so the eight
// lanes of a group read one 64-byte row of a staged table (consecutive
// wires).  Padding entries (past the end) have value 0 and index 0, which
// is in range.
template <int L>
__device__ __forceinline__ double mb_list_sum(const uint32_t *ix, const double *vl, int beg, int end,
                                              const double *s_e, const double *s_h, const double *vv) {
  using T = MbList<L>;
  const double *t1 = T::c1 ? s_h : s_e, *t2 = T::c2 ? s_h : s_e, *t3 = T::c3 ? s_h : s_e;
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6, u = lane >> 3, c = lane & 7;
  double acc = 0.0;
  for (int e0 = beg + wv * 64; e0 < end; e0 += MB_THREADS) {
    const int e = e0 + lane;
    const uint32_t pk_l = e < end ? ix[e] : 0u;
    const double val_l = e < end ? vl[e] : 0.0;
#pragma unroll
    for (int r = 0; r < 8; ++r) {
      const int src = 8 * r + u;
      const uint32_t pk = (uint32_t)__shfl((int)pk_l, src, 64);
      const double val = __shfl(val_l, src, 64);
      int a, cc, r1, r2, r3;
      mb_indices<L>(pk, a, cc, r1, r2, r3);
      // VC is lower-triangular: V(a, c) = 0 for a < c
      const double v2 = vv[(a >= cc ? a - cc : 0) * MB_WC + c];
      const double w = (a >= cc) ? val : 0.0;
      const double p1 = t1[r1 * MB_WC + c], p2 = t2[r2 * MB_WC + c], p3 = t3[r3 * MB_WC + c];
      const double f1 = T::x1 ? 1.0 - p1 : p1;
      const double f2 = T::x2 ? 1.0 - p2 : p2;
      const double f3 = T::x3 ? 1.0 - p3 : p3;
      acc += v2 * (f1 * f2 * f3) * w;
    }
  }
  return acc;
}

__global__ __launch_bounds__(MB_THREADS) void k_mb(psa_group g, int w0, int nw) {
  extern __shared__ double smem[];
  __shared__ double red[MB_THREADS / 64][MB_WC];
  const int N = g.N, W = g.Nw, slot = blockIdx.y, kb = blockIdx.x * MB_KG;
  const int wc0 = blockIdx.z * MB_WC;  // first wire of the chunk (relative to w0)
  const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
  const int ch2 = slot < 2 ? 1 : 2;  // second V^2 channel: Vee for electrons, Vhh for holes
  double *s_ne = smem, *s_nh = smem + N * MB_WC;
  double *s_v0 = smem + 2 * N * MB_WC, *s_v1 = smem + 3 * N * MB_WC;
  const double *occ = g.wmb + 4L * N * W, *vsq = g.wmb + 6L * N * W;
  for (int idx = tid; idx < N * MB_WC; idx += MB_THREADS) {
    const int k = idx / MB_WC, cc = idx % MB_WC;
    const int w = w0 + min(wc0 + cc, nw - 1);  // idle wires of the last chunk repeat one wire
    s_ne[idx] = occ[(long)k * W + w];
    s_nh[idx] = occ[(long)(N + k) * W + w];
    s_v0[idx] = vsq[(long)k * W + w];
    s_v1[idx] = vsq[((long)ch2 * N + k) * W + w];
  }
  __syncthreads();
  for (int kk = 0; kk < MB_KG; ++kk) {
    const int k = kb + kk;
    if (k >= N) break;
    double acc = 0.0;
    for (int half = 0; half < 2; ++half) {
      // slot 0 WinE <- lists 0, 2; 1 WoutE <- 1, 3; 2 WinH <- 4, 6; 3 WoutH <- 5, 7
      const int L = (slot >> 1) * 4 + (slot & 1) + 2 * half;
      if (g.nnz_mb[L] == 0) continue;
      const int beg = g.mb_off[L][k], end = g.mb_off[L][k + 1];
      const uint32_t *ix = g.mb_idx[L];
      const double *vl = g.mb_val[L];
      // V^2 channel: lists 0, 1, 4, 5 use Veh, 2, 3 Vee, 6, 7 Vhh
      const double *vv = half == 0 ? s_v0 : s_v1;
      switch (L) {
        case 0: acc += mb_list_sum<0>(ix, vl, beg, end, s_ne, s_nh, vv); break;
        case 1: acc += mb_list_sum<1>(ix, vl, beg, end, s_ne, s_nh, vv); break;
        case 2: acc += mb_list_sum<2>(ix, vl, beg, end, s_ne, s_nh, vv); break;
        case 3: acc += mb_list_sum<3>(ix, vl, beg, end, s_ne, s_nh, vv); break;
        case 4: acc += mb_list_sum<4>(ix, vl, beg, end, s_ne, s_nh, vv); break;
        case 5: acc += mb_list_sum<5>(ix, vl, beg, end, s_ne, s_nh, vv); break;
        case 6: acc += mb_list_sum<6>(ix, vl, beg, end, s_ne, s_nh, vv); break;
        default: acc += mb_list_sum<7>(ix, vl, beg, end, s_ne, s_nh, vv); break;
      }
    }
    // reduce over the eight entry groups of a wave (lane bits 3..5), then
    // over the waves, in a fixed order
    acc += __shfl_xor(acc, 8, 64);
    acc += __shfl_xor(acc, 16, 64);
    acc += __shfl_xor(acc, 32, 64);
    if (lane < MB_WC) red[wv][lane] = acc;
    __syncthreads();
    if (tid < MB_WC && wc0 + tid < nw) {
      double sum = 0.0;
#pragma unroll
      for (int u = 0; u < MB_THREADS / 64; ++u) sum += red[u][tid];
      g.wmb[(long)(w0 + wc0 + tid) * 4 * N + slot * N + k] = sum;
    }
    __syncthreads();
  }
}

// The same sums for the reduced-precision modes: occupations and V^2 staged
// in FP32 (V^2 scaled by 2^100, exact, to stay in the FP32 normal range;
// unscaled with the FP64 result) for 16 wires per block.  A lane is (entry
// group lane >> 2, wires 4 (lane & 3) .. 4 (lane & 3) + 3): the four lanes
// of a group load the same list entry (one request), decode it once for
// their four wires and read each table row with one 16-byte LDS load; the
// sixteen groups of a wave take consecutive entries, and a lane loads the
// entries of four block passes before using them.  The occupation and
// V^2 products run in FP32; the table weight (~1e-270 .. 1e57) multiplies
// them in FP64 into FP64 sums.
constexpr int MB2_WC = 16;

template <int L>
__device__ __forceinline__ void mb_list_sum2(const uint32_t *ix, const double *vl, int beg, int end,
                                             const float *s_e, const float *s_h, const float *vv,
                                             double (&acc)[4]) {
  using T = MbList<L>;
  const float *t1 = T::c1 ? s_h : s_e, *t2 = T::c2 ? s_h : s_e, *t3 = T::c3 ? s_h : s_e;
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6, c4 = 4 * (lane & 3);
  constexpr int PASS = MB_THREADS / 4, B = 4;  // entries per block pass; passes loaded together
  for (int e0 = beg + 16 * wv + (lane >> 2); e0 < end; e0 += B * PASS) {
    uint32_t pks[B];
    double vals[B];
#pragma unroll
    for (int u = 0; u < B; ++u) {  // the entries of B passes in flight together (padding: weight 0, index 0)
      const int e = e0 + u * PASS;
      pks[u] = e < end ? ix[e] : 0u;
      vals[u] = e < end ? vl[e] : 0.0;
    }
#pragma unroll
    for (int u = 0; u < B; ++u) {
      const uint32_t pk = pks[u];
      const double val = vals[u];
      int a, cc, r1, r2, r3;
      mb_indices<L>(pk, a, cc, r1, r2, r3);
      // VC is lower-triangular: V(a, c) = 0 for a < c
      const float4 v2 = *reinterpret_cast<const float4 *>(&vv[(a >= cc ? a - cc : 0) * MB2_WC + c4]);
      const double w = (a >= cc) ? val : 0.0;
      const float4 p1 = *reinterpret_cast<const float4 *>(&t1[r1 * MB2_WC + c4]);
      const float4 p2 = *reinterpret_cast<const float4 *>(&t2[r2 * MB2_WC + c4]);
      const float4 p3 = *reinterpret_cast<const float4 *>(&t3[r3 * MB2_WC + c4]);
      const float pv[4][4] = {{v2.x, p1.x, p2.x, p3.x}, {v2.y, p1.y, p2.y, p3.y},
                              {v2.z, p1.z, p2.z, p3.z}, {v2.w, p1.w, p2.w, p3.w}};
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        const float f1 = T::x1 ? 1.0f - pv[q][1] : pv[q][1];
        const float f2 = T::x2 ? 1.0f - pv[q][2] : pv[q][2];
        const float f3 = T::x3 ? 1.0f - pv[q][3] : pv[q][3];
        acc[q] = fma((double)(pv[q][0] * (f1 * f2 * f3)), w, acc[q]);
      }
    }
  }
}

__global__ __launch_bounds__(MB_THREADS) void k_mb2(psa_group g, int w0, int nw) {
  extern __shared__ float smf[];
  __shared__ double red2[MB_THREADS / 64][MB2_WC];
  const int N = g.N, W = g.Nw, slot = blockIdx.y, kb = blockIdx.x * MB_KG;
  const int wc0 = blockIdx.z * MB2_WC;  // first wire of the chunk (relative to w0)
  const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
  const int ch2 = slot < 2 ? 1 : 2;  // second V^2 channel: Vee for electrons, Vhh for holes
  float *s_ne = smf, *s_nh = smf + N * MB2_WC;
  float *s_v0 = smf + 2 * N * MB2_WC, *s_v1 = smf + 3 * N * MB2_WC;
  const double *occ = g.wmb + 4L * N * W, *vsq = g.wmb + 6L * N * W;
  for (int idx = tid; idx < N * MB2_WC; idx += MB_THREADS) {
    const int k = idx / MB2_WC, cc = idx % MB2_WC;
    const int w = w0 + min(wc0 + cc, nw - 1);  // idle wires of the last chunk repeat one wire
    s_ne[idx] = (float)occ[(long)k * W + w];
    s_nh[idx] = (float)occ[(long)(N + k) * W + w];
    s_v0[idx] = (float)(0x1p100 * vsq[(long)k * W + w]);
    s_v1[idx] = (float)(0x1p100 * vsq[((long)ch2 * N + k) * W + w]);
  }
  __syncthreads();
  for (int kk = 0; kk < MB_KG; ++kk) {
    const int k = kb + kk;
    if (k >= N) break;
    double acc[4] = {0.0, 0.0, 0.0, 0.0};
    for (int half = 0; half < 2; ++half) {
      // slot 0 WinE <- lists 0, 2; 1 WoutE <- 1, 3; 2 WinH <- 4, 6; 3 WoutH <- 5, 7
      const int L = (slot >> 1) * 4 + (slot & 1) + 2 * half;
      if (g.nnz_mb[L] == 0) continue;
      const int beg = g.mb_off[L][k], end = g.mb_off[L][k + 1];
      const uint32_t *ix = g.mb_idx[L];
      const double *vl = g.mb_val[L];
      const float *vv = half == 0 ? s_v0 : s_v1;
      switch (L) {
        case 0: mb_list_sum2<0>(ix, vl, beg, end, s_ne, s_nh, vv, acc); break;
        case 1: mb_list_sum2<1>(ix, vl, beg, end, s_ne, s_nh, vv, acc); break;
        case 2: mb_list_sum2<2>(ix, vl, beg, end, s_ne, s_nh, vv, acc); break;
        case 3: mb_list_sum2<3>(ix, vl, beg, end, s_ne, s_nh, vv, acc); break;
        case 4: mb_list_sum2<4>(ix, vl, beg, end, s_ne, s_nh, vv, acc); break;
        case 5: mb_list_sum2<5>(ix, vl, beg, end, s_ne, s_nh, vv, acc); break;
        case 6: mb_list_sum2<6>(ix, vl, beg, end, s_ne, s_nh, vv, acc); break;
        default: mb_list_sum2<7>(ix, vl, beg, end, s_ne, s_nh, vv, acc); break;
      }
    }
    // reduce over the sixteen entry groups of a wave (lane bits 2..5), then
    // over the waves, in a fixed order
#pragma unroll
    for (int o = 4; o < 64; o *= 2)
#pragma unroll
      for (int q = 0; q < 4; ++q) acc[q] += __shfl_xor(acc[q], o, 64);
    if (lane < 4)
#pragma unroll
      for (int q = 0; q < 4; ++q) red2[wv][4 * lane + q] = acc[q];
    __syncthreads();
    if (tid < MB2_WC && wc0 + tid < nw) {
      double sum = 0.0;
#pragma unroll
      for (int u = 0; u < MB_THREADS / 64; ++u) sum += red2[u][tid];
      g.wmb[(long)(w0 + wc0 + tid) * 4 * N + slot * N + k] = sum * 0x1p-100;
    }
    __syncthreads();
  }
}

// Wire-fastest occupations of the new level's diagonals for k_mb (|Re|,
// coulomb.py:1307-1310).
__global__ void k_diag_gather(psa_group g, int w0) {
  const int w = w0 + blockIdx.x, N = g.N, W = g.Nw;
  const long o = mat_off(g, w);
  double *occ = g.wmb + 4L * N * W;
  for (int k = threadIdx.x; k < N; k += blockDim.x) {
    occ[(long)k * W + w] = fabs(ld(g.CX + o, (long)k * N + k).re);
    occ[(long)(N + k) * W + w] = fabs(ld(g.DX + o, (long)k * N + k).re);
  }
}

// Phonon rates (phonons.py:139-268), in/out update and Recomb
// (SBEs.py:5276-5325) on the diagonal of the new level.  Block (wire, 64 k):
// lane -> k, wave -> every fourth i of the phonon sums (EP[k][i] read as
// EPT[i][k], so both sums are coalesced along k); partials are added in
// wave order and wave 0 applies the update.
// Block = 64 rates k (one per lane) x RL_WAVES waves over the phonon sums
// (i congruent to the wave index), added in a fixed order.
constexpr int RL_WAVES = 16;
__global__ __launch_bounds__(64 * RL_WAVES) void k_relax(psa_group g, int w0, double dt) {
  __shared__ double s_ner[1024], s_nhr[1024];
  __shared__ double s_vep[1024], s_vhp[1024];  // VC / E1D * Vscale per |k - i|, electrons and holes
  __shared__ double part[4][RL_WAVES][64];
  const int w = w0 + blockIdx.x, N = g.N, tid = threadIdx.x;
  const int lane = tid & 63, wv = tid >> 6, k = blockIdx.y * 64 + lane;
  const long o = mat_off(g, w);
  double *C3 = g.CX + o, *D3 = g.DX + o;
  const double *pr = g.prep + (long)w * PSA_PREP_STRIDE(N);
  const double *E1D = pr + PSA_PREP_E1D(N), *vc = pr + PSA_PREP_VC(N), *Rsp = pr + PSA_PREP_RSP(N);
  double *mb = g.wmb + (long)w * 4 * N;
  const bool ehs = (g.flags & PSA_F_EHS) != 0, phon = (g.flags & PSA_F_PHONON) != 0;
  const bool recomb = (g.flags & PSA_F_RECOMB) != 0;
  for (int i = tid; i < N; i += blockDim.x) {
    s_ner[i] = ld(C3, (long)i * N + i).re;
    s_nhr[i] = ld(D3, (long)i * N + i).re;
    // the interaction factors depend on |k - i| only: one division per distance
    s_vep[i] = vc[N + i] / E1D[i] * g.Vscale;
    s_vhp[i] = vc[2 * N + i] / E1D[i] * g.Vscale;
  }
  __syncthreads();
  double a = 0, b = 0, c = 0, d = 0;
  if (phon && k < N) {
    // Vep[k][i] = VC[k,i,1] / E1D[k,i] * Vscale, VC lower-triangular
#pragma unroll 4
    for (int i = wv; i <= k; i += RL_WAVES) {
      const double vep = s_vep[k - i];
      a += vep * s_ner[i] * g.EPT[(long)i * N + k];
      b += vep * (1.0 - s_ner[i]) * g.EP[(long)i * N + k];
    }
#pragma unroll 4
    for (int i = k + wv; i < N; i += RL_WAVES) {
      const double vhp = s_vhp[i - k];
      c += vhp * s_nhr[i] * g.HPT[(long)i * N + k];
      d += vhp * (1.0 - s_nhr[i]) * g.HP[(long)i * N + k];
    }
  }
  part[0][wv][lane] = a;
  part[1][wv][lane] = b;
  part[2][wv][lane] = c;
  part[3][wv][lane] = d;
  __syncthreads();
  if (wv != 0 || k >= N) return;
  double wie = 0, woe = 0, wih = 0, woh = 0;
  if (ehs) {
    wie = mb[k];
    woe = mb[N + k];
    wih = mb[2 * N + k];
    woh = mb[3 * N + k];
  }
  if (phon) {
    double t[4];
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      t[j] = part[j][0][lane];
#pragma unroll
      for (int u = 1; u < RL_WAVES; ++u) t[j] += part[j][u][lane];
    }
    wie += t[0];
    woe += t[1];
    wih += t[2];
    woh += t[3];
  }
  cd ne = ld(C3, (long)k * N + k), nh = ld(D3, (long)k * N + k);
  if (ehs || phon) {
    // the rates as the reference's WriteFields snapshots print them
    // (SBEs.py:5282-5307): Win_e (1 - Re ne), Wout_e Re ne, Win_h (1 - Re nh),
    // Wout_h Re nh with the occupations before the update
    mb[k] = wie * (1.0 - ne.re);
    mb[N + k] = woe * ne.re;
    mb[2 * N + k] = wih * (1.0 - nh.re);
    mb[3 * N + k] = woh * nh.re;
    const double We = wie + woe, Wh = wih + woh;
    const double ee = exp(-We * dt), eep = exp(We * dt);
    const double eh = exp(-Wh * dt), ehp = exp(Wh * dt);
    cd t1 = {(-1.0 + eep + ne.re) * wie + ne.re * woe, ne.im * wie + ne.im * woe};
    ne = cd{ee * t1.re / (We + 1e-200), ee * t1.im / (We + 1e-200)};
    cd t2 = {(-1.0 + ehp + nh.re) * wih + nh.re * woh, nh.im * wih + nh.im * woh};
    nh = cd{eh * t2.re / (Wh + 1e-200), eh * t2.im / (Wh + 1e-200)};
    if (recomb) {
      // ne *= exp(-Rsp nh dt); nh *= exp(-Rsp ne dt) with the updated ne
      const double r = Rsp[k];
      cd z = {-r * nh.re * dt, -r * nh.im * dt};
      double sz, cz;
      sincos(z.im, &sz, &cz);
      const double m = exp(z.re);
      ne = cmul(ne, cd{m * cz, m * sz});
      z = cd{-r * ne.re * dt, -r * ne.im * dt};
      sincos(z.im, &sz, &cz);
      const double m2 = exp(z.re);
      nh = cmul(nh, cd{m2 * cz, m2 * sz});
    }
  }
  st(C3, (long)k * N + k, ne);
  st(D3, (long)k * N + k, nh);
}

// ===========================================================================
// Transport (dcfield.py:1176-1236): X <- ifft2(fft2(X) exp(-i(Y_a+Y_b) dk)).
// The phase is separable, so it is a row pass and a column pass of
// FFT -> phase -> inverse FFT, each in LDS (Stockham autosort, radix 4
// stages while 4 divides the remaining length, then 2 and 3, N = 2^a 3^b,
// twiddles from a per-N table).  One wave per line, 2, 4 or 8 lines per
// block (tune[1] for the row pass, tune[3] for the column pass of the
// compile-time-N kernel); the column pass reads and writes that many
// adjacent columns per row.
// ===========================================================================

__device__ __forceinline__ cd tw_get(const cd *tw, int idx, int sign) {
  cd w = tw[idx];  // exp(-2 pi i idx / N)
  if (sign > 0) w.im = -w.im;
  return w;
}

// multiply by i * sign
__device__ __forceinline__ cd mul_is(cd a, int sign) {
  return sign > 0 ? cd{-a.im, a.re} : cd{a.im, -a.re};
}

// LDS slot of element i of a line: one pad slot after every four elements,
// which spreads the strided Stockham accesses over the banks.
__device__ __forceinline__ int fpad(int i) { return i + (i >> 2); }

// Orders one wave's LDS writes before its later LDS reads by other lanes of
// the same wave (the line a wave transforms is private to that wave).
__device__ __forceinline__ void wave_lds_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Stockham autosort FFT of one line held in LDS (indexed through fpad), in
// place, by one wave: each stage reads and combines all its inputs into
// registers, then writes.  Stage with radix R and sub-length Ns: for
// j < N/R, k = j mod Ns, inputs x[j + r N/R] w^(r k N/(R Ns)), outputs at
// (j - k) R + k + s Ns.  sign -1 forward, +1 inverse (unscaled).
// N = 2^a 3^b <= 512, so a lane has at most 4 butterflies per stage.
__device__ void fft_line(cd *x, const cd *tw, int N, int sign, int lane) {
  int Ns = 1, rem = N;
  while (rem > 1) {
    const int R = (rem % 4 == 0) ? 4 : ((rem % 2 == 0) ? 2 : 3);
    const int M = N / R, step = N / (Ns * R);
    const uint64_t mg = ((1ull << 32) + Ns - 1) / Ns;  // j / Ns = (j mg) >> 32 for j < 2^20
    cd o[4][4];
#pragma unroll
    for (int it = 0; it < 4; ++it) {
      const int j = lane + 64 * it;
      if (j >= M) break;
      const int k = j - Ns * (int)(((uint64_t)j * mg) >> 32);
      const cd a0 = x[fpad(j)];
      if (R == 4) {
        const cd a1 = cmul(x[fpad(j + M)], tw_get(tw, k * step, sign));
        const cd a2 = cmul(x[fpad(j + 2 * M)], tw_get(tw, 2 * k * step, sign));
        const cd a3 = cmul(x[fpad(j + 3 * M)], tw_get(tw, 3 * k * step, sign));
        const cd s02 = cadd(a0, a2), d02 = csub(a0, a2);
        const cd s13 = cadd(a1, a3), d13 = mul_is(csub(a1, a3), sign);
        o[it][0] = cadd(s02, s13);
        o[it][1] = cadd(d02, d13);
        o[it][2] = csub(s02, s13);
        o[it][3] = csub(d02, d13);
      } else if (R == 2) {
        const cd b = cmul(x[fpad(j + M)], tw_get(tw, k * step, sign));
        o[it][0] = cadd(a0, b);
        o[it][1] = csub(a0, b);
      } else {
        const double c3 = -0.5, s3 = sign * 0.86602540378443864676;
        const cd w1 = {c3, s3}, w2 = {c3, -s3};
        const cd b = cmul(x[fpad(j + M)], tw_get(tw, k * step, sign));
        const cd c = cmul(x[fpad(j + 2 * M)], tw_get(tw, 2 * k * step, sign));
        o[it][0] = cadd(a0, cadd(b, c));
        o[it][1] = cadd(a0, cadd(cmul(b, w1), cmul(c, w2)));
        o[it][2] = cadd(a0, cadd(cmul(b, w2), cmul(c, w1)));
      }
    }
    wave_lds_sync();
#pragma unroll
    for (int it = 0; it < 4; ++it) {
      const int j = lane + 64 * it;
      if (j >= M) break;
      const int k = j - Ns * (int)(((uint64_t)j * mg) >> 32);
      const int base = (j - k) * R + k;
#pragma unroll
      for (int r = 0; r < 4; ++r)
        if (r < R) x[fpad(base + r * Ns)] = o[it][r];
    }
    wave_lds_sync();
    Ns *= R;
    rem /= R;
  }
}

// fft_line for nl lines of one wave at once (line l at x + l LS): the
// butterflies of all lines share the lanes, so short lines keep the wave
// busy.  nl (N / R) <= 256 for every radix R of the plan.
__device__ void fft_lines(cd *x, int LS, int nl, const cd *tw, int N, int sign, int lane) {
  int Ns = 1, rem = N;
  while (rem > 1) {
    const int R = (rem % 4 == 0) ? 4 : ((rem % 2 == 0) ? 2 : 3);
    const int M = N / R, step = N / (Ns * R), tot = nl * M;
    const uint64_t mg = ((1ull << 32) + Ns - 1) / Ns;
    cd o[4][4];
#pragma unroll
    for (int it = 0; it < 4; ++it) {
      const int jg = lane + 64 * it;
      if (jg >= tot) break;
      const int l = jg / M, j = jg - l * M;
      const cd *xl = x + l * LS;
      const int k = j - Ns * (int)(((uint64_t)j * mg) >> 32);
      const cd a0 = xl[fpad(j)];
      if (R == 4) {
        const cd a1 = cmul(xl[fpad(j + M)], tw_get(tw, k * step, sign));
        const cd a2 = cmul(xl[fpad(j + 2 * M)], tw_get(tw, 2 * k * step, sign));
        const cd a3 = cmul(xl[fpad(j + 3 * M)], tw_get(tw, 3 * k * step, sign));
        const cd s02 = cadd(a0, a2), d02 = csub(a0, a2);
        const cd s13 = cadd(a1, a3), d13 = mul_is(csub(a1, a3), sign);
        o[it][0] = cadd(s02, s13);
        o[it][1] = cadd(d02, d13);
        o[it][2] = csub(s02, s13);
        o[it][3] = csub(d02, d13);
      } else if (R == 2) {
        const cd b = cmul(xl[fpad(j + M)], tw_get(tw, k * step, sign));
        o[it][0] = cadd(a0, b);
        o[it][1] = csub(a0, b);
      } else {
        const double c3 = -0.5, s3 = sign * 0.86602540378443864676;
        const cd w1 = {c3, s3}, w2 = {c3, -s3};
        const cd b = cmul(xl[fpad(j + M)], tw_get(tw, k * step, sign));
        const cd c = cmul(xl[fpad(j + 2 * M)], tw_get(tw, 2 * k * step, sign));
        o[it][0] = cadd(a0, cadd(b, c));
        o[it][1] = cadd(a0, cadd(cmul(b, w1), cmul(c, w2)));
        o[it][2] = cadd(a0, cadd(cmul(b, w2), cmul(c, w1)));
      }
    }
    wave_lds_sync();
#pragma unroll
    for (int it = 0; it < 4; ++it) {
      const int jg = lane + 64 * it;
      if (jg >= tot) break;
      const int l = jg / M, j = jg - l * M;
      cd *xl = x + l * LS;
      const int k = j - Ns * (int)(((uint64_t)j * mg) >> 32);
      const int base = (j - k) * R + k;
#pragma unroll
      for (int r = 0; r < 4; ++r)
        if (r < R) xl[fpad(base + r * Ns)] = o[it][r];
    }
    wave_lds_sync();
    Ns *= R;
    rem /= R;
  }
}

// The same FFT with N known at compile time: the stage sequence, moduli
// and strides are constants, and each stage reads its own twiddle table
// tw[TOFF + (r-1) Ns + k] = w^(r k N / (R Ns)), contiguous in k, so the
// lanes of a wave read consecutive (or the same) slots.
template <int N, int Ns, int Rem, int TOFF>
__device__ __forceinline__ void build_stage_twiddles(cd *tw, const double *twid, int tid, int nthr) {
  if constexpr (Rem > 1) {
    constexpr int R = (Rem % 4 == 0) ? 4 : ((Rem % 2 == 0) ? 2 : 3);
    constexpr int step = N / (Ns * R);
    for (int i = tid; i < (R - 1) * Ns; i += nthr) {
      const int r = i / Ns + 1, k = i % Ns;
      tw[TOFF + i] = ld(twid, r * k * step);
    }
    build_stage_twiddles<N, Ns * R, Rem / R, TOFF + (R - 1) * Ns>(tw, twid, tid, nthr);
  }
}

template <int N, int Ns, int Rem, int SIGN, int TOFF>
__device__ __forceinline__ void fft_stages(cd *x, const cd *tw, int lane) {
  if constexpr (Rem > 1) {
    constexpr int R = (Rem % 4 == 0) ? 4 : ((Rem % 2 == 0) ? 2 : 3);
    constexpr int M = N / R, IT = (M + 63) / 64;
    const cd *tws = tw + TOFF;
    cd o[IT][R];
#pragma unroll
    for (int it = 0; it < IT; ++it) {
      const int j = lane + 64 * it;
      if (M % 64 == 0 || j < M) {
        const int k = j % Ns;
        const cd a0 = x[fpad(j)];
        if constexpr (R == 4) {
          // first stage (Ns = 1): k = 0, all twiddles are 1
          const cd a1 = Ns == 1 ? x[fpad(j + M)] : cmul(x[fpad(j + M)], tw_get(tws, k, SIGN));
          const cd a2 = Ns == 1 ? x[fpad(j + 2 * M)] : cmul(x[fpad(j + 2 * M)], tw_get(tws, Ns + k, SIGN));
          const cd a3 = Ns == 1 ? x[fpad(j + 3 * M)] : cmul(x[fpad(j + 3 * M)], tw_get(tws, 2 * Ns + k, SIGN));
          const cd s02 = cadd(a0, a2), d02 = csub(a0, a2);
          const cd s13 = cadd(a1, a3), d13 = mul_is(csub(a1, a3), SIGN);
          o[it][0] = cadd(s02, s13);
          o[it][1] = cadd(d02, d13);
          o[it][2] = csub(s02, s13);
          o[it][3] = csub(d02, d13);
        } else if constexpr (R == 2) {
          const cd b = cmul(x[fpad(j + M)], tw_get(tws, k, SIGN));
          o[it][0] = cadd(a0, b);
          o[it][1] = csub(a0, b);
        } else {
          constexpr double c3 = -0.5, s3 = SIGN * 0.86602540378443864676;
          const cd w1 = {c3, s3}, w2 = {c3, -s3};
          const cd b = cmul(x[fpad(j + M)], tw_get(tws, k, SIGN));
          const cd c = cmul(x[fpad(j + 2 * M)], tw_get(tws, Ns + k, SIGN));
          o[it][0] = cadd(a0, cadd(b, c));
          o[it][1] = cadd(a0, cadd(cmul(b, w1), cmul(c, w2)));
          o[it][2] = cadd(a0, cadd(cmul(b, w2), cmul(c, w1)));
        }
      }
    }
    wave_lds_sync();
#pragma unroll
    for (int it = 0; it < IT; ++it) {
      const int j = lane + 64 * it;
      if (M % 64 == 0 || j < M) {
        const int k = j % Ns;
        const int base = (j - k) * R + k;
#pragma unroll
        for (int r = 0; r < R; ++r) x[fpad(base + r * Ns)] = o[it][r];
      }
    }
    wave_lds_sync();
    fft_stages<N, Ns * R, Rem / R, SIGN, TOFF + (R - 1) * Ns>(x, tw, lane);
  }
}

template <int NT>
__device__ __forceinline__ void fft_line_n(cd *x, const cd *tw, int N, int sign, int lane) {
  if constexpr (NT == 0) {
    fft_line(x, tw, N, sign, lane);
  } else {
    if (sign < 0)
      fft_stages<NT, 1, NT, -1, 0>(x, tw, lane);
    else
      fft_stages<NT, 1, NT, +1, 0>(x, tw, lane);
  }
}

// per-wire phase exp(-i Y_j dk), dk = -e (Edc + Re Ex[q=0]) dt / hbar
__global__ void k_tphase(psa_group g, int w0, double dt) {
  const int w = w0 + blockIdx.x, N = g.N;
  const double *eq = g.exq + (long)w * 3 * 2 * (2 * N - 1);
  const double eac = eq[2 * (N - 1)];
  const double dk = -g.e0 * (g.Edc + eac) / g.hbar * dt;
  for (int j = threadIdx.x; j < N; j += blockDim.x) {
    double s, c;
    sincos(-g.ytr[j] * dk, &s, &c);
    st(g.tphase + (long)w * 2 * N, j, cd{c, s});
  }
}

// Transport with LF = 0 (dcfield.py:1229-1232): only the diagonal of each
// matrix moves, ShiftN1D (dcfield.py:1154-1174) = ifft(fft(n) exp(-i Y dk)).
// Block (wire, matrix): the diagonal and the twiddles in LDS, the forward
// and inverse transforms as direct sums (twiddle index m j mod N), FP64.
__global__ __launch_bounds__(256) void k_transport_diag(psa_group g, int w0) {
  __shared__ cd s_n[1024], s_f[1024], s_tw[1024];
  const int w = w0 + blockIdx.x, N = g.N, mat = blockIdx.y;  // 0 C, 1 D, 2 p
  const long o = mat_off(g, w);
  double *X = (mat == 0 ? g.CX : (mat == 1 ? g.DX : g.pX)) + o;
  const double *ph = g.tphase + (long)w * 2 * N;
  for (int j = threadIdx.x; j < N; j += blockDim.x) {
    s_n[j] = ld(X, (long)j * N + j);
    s_tw[j] = ld(g.twid, j);  // exp(-2 pi i j / N)
  }
  __syncthreads();
  for (int m = threadIdx.x; m < N; m += blockDim.x) {
    cd a = {0.0, 0.0};
    int t = 0;  // m j mod N
    for (int j = 0; j < N; ++j) {
      a = cadd(a, cmul(s_n[j], s_tw[t]));
      t += m;
      if (t >= N) t -= N;
    }
    s_f[m] = cmul(a, ld(ph, m));
  }
  __syncthreads();
  const double inv = 1.0 / N;
  for (int j = threadIdx.x; j < N; j += blockDim.x) {
    cd a = {0.0, 0.0};
    int t = 0;  // m j mod N
    for (int m = 0; m < N; ++m) {
      a = cadd(a, cmul(s_f[m], cconj(s_tw[t])));
      t += j;
      if (t >= N) t -= N;
    }
    st(X, (long)j * N + j, cscale(a, inv));
  }
}

// pass 0: rows, pass 1: columns.  With avg set (column pass only), also
// writes the leapfrog average A = (A + X3) / 2 (SBEs.py:5704-5706) for
// every element except the diagonals of C and D, which the e/h
// renormalisation still changes (k_finish_diag averages those).
// LDS per line: at least fpad(N) slots, rounded so that the line stride is
// 16 / LINES modulo 16 slots (of 16 bytes): the column pass's staging
// accesses (line index fastest, LINES lines x 16 / LINES rows per 16 lanes)
// then fall on 16 different bank groups.
__host__ __device__ inline int tr_line_stride(int N, int lines) {
  const int need = N + N / 4, want = 16 / lines;
  return need + ((want - need % 16) + 16) % 16;
}

// NT: N at compile time (0: taken from the group at run time); LINES lines
// (waves) per block.
template <int NT, int LINES>
__global__ __launch_bounds__(64 * LINES) void k_transport(psa_group g, int w0, int pass, int avg) {
  extern __shared__ double smem[];
  const int N = NT ? NT : g.N, tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
  const int LS = tr_line_stride(N, LINES);
  const int mat = blockIdx.y;  // 0 C, 1 D, 2 p
  const int w = w0 + blockIdx.z;
  const int line0 = blockIdx.x * LINES;
  cd *base = reinterpret_cast<cd *>(smem);
  cd *tw = base + LINES * LS;  // [N] twiddles
  if constexpr (NT == 0) {
    for (int j = tid; j < N; j += blockDim.x) tw[j] = ld(g.twid, j);
  } else {
    build_stage_twiddles<NT, 1, NT, 0>(tw, g.twid, tid, blockDim.x);
  }
  const long o = mat_off(g, w);
  double *X = (mat == 0 ? g.CX : (mat == 1 ? g.DX : g.pX)) + o;
  double *A = (mat == 0 ? g.CA : (mat == 1 ? g.DA : g.pA)) + o;
  const int line = line0 + wv;
  cd *x = base + wv * LS;
  if (pass == 0) {
    if (line < N)
      for (int j = lane; j < N; j += 64) x[fpad(j)] = ld(X, (long)line * N + j);
  } else {
    for (int idx = tid; idx < LINES * N; idx += blockDim.x) {
      const int c = idx % LINES, j = idx / LINES;
      if (line0 + c < N) base[c * LS + fpad(j)] = ld(X, (long)j * N + line0 + c);
    }
  }
  __syncthreads();
  fft_line_n<NT>(x, tw, N, -1, lane);
  const double *ph = g.tphase + (long)w * 2 * N;
  for (int j = lane; j < N; j += 64) x[fpad(j)] = cmul(x[fpad(j)], ld(ph, j));
  wave_lds_sync();
  fft_line_n<NT>(x, tw, N, +1, lane);
  __syncthreads();  // the column pass stores lines of other waves
  const double inv = 1.0 / N;
  if (pass == 0) {
    if (line < N)
      for (int j = lane; j < N; j += 64) st(X, (long)line * N + j, cscale(x[fpad(j)], inv));
  } else {
    for (int idx = tid; idx < LINES * N; idx += blockDim.x) {
      const int c = idx % LINES, j = idx / LINES, col = line0 + c;
      if (col < N) {
        const long ij = (long)j * N + col;
        const cd v = cscale(base[c * LS + fpad(j)], inv);
        st(X, ij, v);
        if (avg && !(mat < 2 && j == col)) {
          const cd a = ld(A, ij);
          st(A, ij, cd{(a.re + v.re) / 2.0, (a.im + v.im) / 2.0});
        }
      }
    }
  }
}

// ---------------------------------------------------------------------------
// Transport with N known at compile time, stages fused across the FFT pair.
// The forward FFT takes the radices 4 (while 4 divides the rest), then 2,
// then 3; the inverse takes the same radices in reverse order.  The forward
// FFT's last stage leaves lane j holding the elements j + r N/R, which are
// exactly the inputs of the inverse FFT's first stage (radix R, no
// twiddles) in the same lane, so the phase multiply and that stage run in
// registers between the two.  In the row pass the first forward stage reads
// its inputs straight from HBM and the last inverse stage writes straight
// back (its outputs j + r N/R are contiguous across lanes); the column pass
// stages its lines through LDS as before.  Each stage has its own twiddle
// table (w^(r k N/(R Ns)), contiguous in k), forward and inverse apart.
// ---------------------------------------------------------------------------
// LDS slot of element i of a line: i with bits 0-3 XOR-ed by 5 ((i >> 4) & 3),
// a permutation within each 16-element block that keeps the 16-lane groups
// of contiguous accesses and of the radix-4 write patterns (stride 4, and
// runs of four at stride 16) on 16 different 16-byte bank groups.
__device__ __forceinline__ int fswz(int i) { return i ^ (5 * ((i >> 4) & 3)); }
// line stride (slots) for fswz lines: at least N, = 16 / lines modulo 16
__host__ __device__ inline int tr_line_stride_s(int N, int lines) {
  const int want = 16 / lines;
  return N + ((want - N % 16) + 16) % 16;
}

struct RadixPlan {
  int S;
  int R[12];
};
__host__ __device__ constexpr RadixPlan radix_plan(int N) {
  RadixPlan p{0, {0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0}};
  int rem = N;
  while (rem > 1) {
    const int r = rem % 4 == 0 ? 4 : (rem % 2 == 0 ? 2 : 3);
    p.R[p.S++] = r;
    rem /= r;
  }
  return p;
}
// radix of stage s of direction dir (0 forward, 1 inverse)
__host__ __device__ constexpr int plan_r(int N, int dir, int s) {
  return dir == 0 ? radix_plan(N).R[s] : radix_plan(N).R[radix_plan(N).S - 1 - s];
}
__host__ __device__ constexpr int plan_ns(int N, int dir, int s) {
  int ns = 1;
  for (int t = 0; t < s; ++t) ns *= plan_r(N, dir, t);
  return ns;
}
// twiddle slots before stage s of direction dir (forward tables first)
__host__ __device__ constexpr int plan_toff(int N, int dir, int s) {
  int off = 0;
  if (dir == 1)
    for (int t = 0; t < radix_plan(N).S; ++t) off += (plan_r(N, 0, t) - 1) * plan_ns(N, 0, t);
  for (int t = 0; t < s; ++t) off += (plan_r(N, dir, t) - 1) * plan_ns(N, dir, t);
  return off;
}
__host__ __device__ constexpr int plan_twiddles(int N) { return plan_toff(N, 1, radix_plan(N).S); }

template <int N, int DIR, int s>
__device__ __forceinline__ void plan_build_twiddles(cd *tw, const double *twid, int tid, int nthr) {
  if constexpr (s < radix_plan(N).S) {
    constexpr int R = plan_r(N, DIR, s), Ns = plan_ns(N, DIR, s), TOFF = plan_toff(N, DIR, s);
    constexpr int step = N / (R * Ns);
    for (int i = tid; i < (R - 1) * Ns; i += nthr) {
      const int r = i / Ns + 1, k = i % Ns;
      tw[TOFF + i] = ld(twid, r * k * step);
    }
    plan_build_twiddles<N, DIR, s + 1>(tw, twid, tid, nthr);
  }
}

template <int R, int SIGN>
__device__ __forceinline__ void dft_small(const cd (&a)[R], cd (&o)[R]) {
  if constexpr (R == 4) {
    const cd s02 = cadd(a[0], a[2]), d02 = csub(a[0], a[2]);
    const cd s13 = cadd(a[1], a[3]), d13 = mul_is(csub(a[1], a[3]), SIGN);
    o[0] = cadd(s02, s13);
    o[1] = cadd(d02, d13);
    o[2] = csub(s02, s13);
    o[3] = csub(d02, d13);
  } else if constexpr (R == 2) {
    o[0] = cadd(a[0], a[1]);
    o[1] = csub(a[0], a[1]);
  } else {
    constexpr double c3 = -0.5, s3 = SIGN * 0.86602540378443864676;
    const cd w1 = {c3, s3}, w2 = {c3, -s3};
    o[0] = cadd(a[0], cadd(a[1], a[2]));
    o[1] = cadd(a[0], cadd(cmul(a[1], w1), cmul(a[2], w2)));
    o[2] = cadd(a[0], cadd(cmul(a[1], w2), cmul(a[2], w1)));
  }
}

// Inputs of stage s for this lane's butterflies, from the LDS line x.
template <int N, int DIR, int s, int IT, int R>
__device__ __forceinline__ void stage_read(const cd *x, const cd *tw, int lane, cd (&a)[IT][R]) {
  constexpr int Ns = plan_ns(N, DIR, s), M = N / R, TOFF = plan_toff(N, DIR, s);
  constexpr int SIGN = DIR == 0 ? -1 : +1;
#pragma unroll
  for (int it = 0; it < IT; ++it) {
    const int j = lane + 64 * it;
    if (M % 64 == 0 || j < M) {
      const int k = j % Ns;
      a[it][0] = x[fswz(j)];
#pragma unroll
      for (int r = 1; r < R; ++r) {
        const cd v = x[fswz(j + r * M)];
        a[it][r] = Ns == 1 ? v : cmul(v, tw_get(tw + TOFF, (r - 1) * Ns + k, SIGN));
      }
    }
  }
}

// Outputs of stage s to the LDS line x (positions (j - k) R + k + r Ns).
template <int N, int DIR, int s, int IT, int R>
__device__ __forceinline__ void stage_write(cd *x, int lane, const cd (&o)[IT][R]) {
  constexpr int Ns = plan_ns(N, DIR, s), M = N / R;
#pragma unroll
  for (int it = 0; it < IT; ++it) {
    const int j = lane + 64 * it;
    if (M % 64 == 0 || j < M) {
      const int k = j % Ns, base = (j - k) * R + k;
#pragma unroll
      for (int r = 0; r < R; ++r) x[fswz(base + r * Ns)] = o[it][r];
    }
  }
}

// Stages s .. e-1 of direction DIR, LDS to LDS.
template <int N, int DIR, int s, int e>
__device__ __forceinline__ void stages_lds(cd *x, const cd *tw, int lane) {
  if constexpr (s < e) {
    constexpr int R = plan_r(N, DIR, s), IT = (N / R + 63) / 64;
    constexpr int SIGN = DIR == 0 ? -1 : +1;
    cd a[IT][R], o[IT][R];
    stage_read<N, DIR, s, IT, R>(x, tw, lane, a);
#pragma unroll
    for (int it = 0; it < IT; ++it) dft_small<R, SIGN>(a[it], o[it]);
    wave_lds_sync();
    stage_write<N, DIR, s, IT, R>(x, lane, o);
    wave_lds_sync();
    stages_lds<N, DIR, s + 1, e>(x, tw, lane);
  }
}

// Forward last stage, phase multiply, inverse first stage, in registers;
// ph holds this lane's phases at the positions j + r N/R.
template <int N, int IT, int R>
__device__ __forceinline__ void fused_middle(cd *x, const cd *tw, int lane, const cd (&ph)[IT][R]) {
  constexpr int S = radix_plan(N).S;
  cd a[IT][R], o[IT][R];
  stage_read<N, 0, S - 1, IT, R>(x, tw, lane, a);
#pragma unroll
  for (int it = 0; it < IT; ++it) {
    dft_small<R, -1>(a[it], o[it]);
#pragma unroll
    for (int r = 0; r < R; ++r) a[it][r] = cmul(o[it][r], ph[it][r]);
    dft_small<R, +1>(a[it], o[it]);
  }
  wave_lds_sync();
  stage_write<N, 1, 0, IT, R>(x, lane, o);
  wave_lds_sync();
}

// Column pass with `planes` (split-BF16 step, LINES 8 or 16): the final
// p, C, D of the new level also leave as the bf16 hi/lo operand planes of
// the next step's RHS (the layout of sp::k_split, both orientations, C and
// D with their diagonals zeroed), and the p tag of the wire records it.
template <int NT, int LINES>
__global__ __launch_bounds__(64 * LINES) void k_transport_f(psa_group g, int w0, int pass, int avg, int planes) {
  extern __shared__ double smem[];
  constexpr int N = NT, S = radix_plan(N).S;
  constexpr int RL = plan_r(N, 0, S - 1), ITL = (N / RL + 63) / 64;  // the fused radix
  constexpr int R0 = plan_r(N, 0, 0), IT0 = (N / R0 + 63) / 64;      // first forward stage
  constexpr int RE = plan_r(N, 1, S - 1), ITE = (N / RE + 63) / 64;  // last inverse stage
  const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
  const int LS = tr_line_stride_s(N, LINES);
  const int mat = blockIdx.y;  // 0 C, 1 D, 2 p
  const int w = w0 + blockIdx.z;
  // column pass with planes: the blocks of neighbouring columns run on one
  // XCD (block i runs on XCD i % 8), so the partial 64-byte rows of the
  // matrix-orientation planes that they write meet in one L2
  int bx = blockIdx.x;
  if (pass == 1 && planes && gridDim.x % 8 == 0) bx = (bx % 8) * (gridDim.x / 8) + bx / 8;
  const int line0 = bx * LINES, line = line0 + wv;
  cd *base = reinterpret_cast<cd *>(smem);
  cd *tw = base + LINES * LS;
  cd *x = base + wv * LS;
  const long o = mat_off(g, w);
  double *X = (mat == 0 ? g.CX : (mat == 1 ? g.DX : g.pX)) + o;
  double *A = (mat == 0 ? g.CA : (mat == 1 ? g.DA : g.pA)) + o;
  const double *phg = g.tphase + (long)w * 2 * N;
  // this lane's phases for the fused middle: positions j + r N/RL
  cd ph[ITL][RL];
#pragma unroll
  for (int it = 0; it < ITL; ++it)
#pragma unroll
    for (int r = 0; r < RL; ++r) {
      const int j = min(lane + 64 * it, N / RL - 1);
      ph[it][r] = ld(phg, j + r * (N / RL));
    }
  const double inv = 1.0 / N;
  if (pass == 0) {
    // first forward stage straight from the row (radix R0, no twiddles)
    const bool on = line < N;
    const int row = on ? line : 0;
    cd a[IT0][R0];
#pragma unroll
    for (int it = 0; it < IT0; ++it)
#pragma unroll
      for (int r = 0; r < R0; ++r) {
        const int j = min(lane + 64 * it, N / R0 - 1);
        a[it][r] = ld(X, (long)row * N + j + r * (N / R0));
      }
    plan_build_twiddles<N, 0, 0>(tw, g.twid, tid, blockDim.x);
    plan_build_twiddles<N, 1, 0>(tw, g.twid, tid, blockDim.x);
    __syncthreads();
    if (!on) return;  // whole wave; no block barrier follows
    {
      cd ob[IT0][R0];
#pragma unroll
      for (int it = 0; it < IT0; ++it) dft_small<R0, -1>(a[it], ob[it]);
      stage_write<N, 0, 0, IT0, R0>(x, lane, ob);
      wave_lds_sync();
    }
    stages_lds<N, 0, 1, S - 1>(x, tw, lane);
    fused_middle<N, ITL, RL>(x, tw, lane, ph);
    stages_lds<N, 1, 1, S - 1>(x, tw, lane);
    // last inverse stage straight to the row: outputs j + r N/RE
    cd b[ITE][RE], ob[ITE][RE];
    stage_read<N, 1, S - 1, ITE, RE>(x, tw, lane, b);
#pragma unroll
    for (int it = 0; it < ITE; ++it) {
      dft_small<RE, +1>(b[it], ob[it]);
      const int j = lane + 64 * it;
      if ((N / RE) % 64 == 0 || j < N / RE)
#pragma unroll
        for (int r = 0; r < RE; ++r) st(X, (long)line * N + j + r * (N / RE), cscale(ob[it][r], inv));
    }
  } else {
    plan_build_twiddles<N, 0, 0>(tw, g.twid, tid, blockDim.x);
    plan_build_twiddles<N, 1, 0>(tw, g.twid, tid, blockDim.x);
    for (int idx = tid; idx < LINES * N; idx += blockDim.x) {
      const int c = idx % LINES, j = idx / LINES;
      if (line0 + c < N) base[c * LS + fswz(j)] = ld(X, (long)j * N + line0 + c);
    }
    __syncthreads();
    if (line < N) {
      stages_lds<N, 0, 0, S - 1>(x, tw, lane);
      fused_middle<N, ITL, RL>(x, tw, lane, ph);
      stages_lds<N, 1, 1, S>(x, tw, lane);
    }
    __syncthreads();  // the stores below take lines of other waves
    for (int idx = tid; idx < LINES * N; idx += blockDim.x) {
      const int c = idx % LINES, j = idx / LINES, col = line0 + c;
      if (col < N) {
        const long ij = (long)j * N + col;
        const cd v = cscale(base[c * LS + fswz(j)], inv);
        st(X, ij, v);
        if (avg && !(mat < 2 && j == col)) {
          const cd a = ld(A, ij);
          st(A, ij, cd{(a.re + v.re) / 2.0, (a.im + v.im) / 2.0});
        }
      }
    }
    if constexpr (LINES % 8 == 0 && NT % LINES == 0) {
      if (planes) {
        // items: (row j, 8-column group) of the matrix slot, then (column,
        // 8-row group) of the transposed slot; 16-byte stores of 8 k.  In
        // the transposed slot the four 8-row groups of a 32-k block are
        // neighbouring items, so four lanes write a slot row's whole 64
        // bytes of a plane in one instruction
        static_assert(NT % 32 == 0, "the plane-writing column pass needs N % 32 == 0");
        const int sm = mat == 2 ? 0 : mat + 1;  // slot matrix: p, C_o, D_o
        const bool zd = mat < 2;
        uint16_t *spl = reinterpret_cast<uint16_t *>(g.spl);
        constexpr int NA = N * (LINES / 8), NB = LINES * (N / 8);
        for (int it = tid; it < NA + NB; it += blockDim.x) {
          int row, k0, cb, jb, dc, dj;  // value q at line cb + q dc, row jb + q dj
          if (it < NA) {
            row = it / (LINES / 8);
            const int gq = it % (LINES / 8);
            k0 = line0 + 8 * gq;
            cb = 8 * gq, jb = row, dc = 1, dj = 0;
          } else {
            const int u = it - NA, c = (u >> 2) % LINES;
            row = line0 + c;
            k0 = 32 * ((u >> 2) / LINES) + 8 * (u & 3);
            cb = c, jb = k0, dc = 0, dj = 1;
          }
          sp::u32x4 ph[4];
#pragma unroll
          for (int q = 0; q < 8; q += 2) {
            cd v2[2];
#pragma unroll
            for (int r = 0; r < 2; ++r) {
              const int cc = cb + (q + r) * dc, jj = jb + (q + r) * dj;
              v2[r] = cscale(base[cc * LS + fswz(jj)], inv);
              if (zd && jj == line0 + cc) v2[r] = cd{0.0, 0.0};
            }
            uint32_t rh, rl, ih, il;
            sp::split_pair(v2[0].re, v2[1].re, rh, rl);
            sp::split_pair(v2[0].im, v2[1].im, ih, il);
            ph[0][q >> 1] = rh;
            ph[1][q >> 1] = rl;
            ph[2][q >> 1] = ih;
            ph[3][q >> 1] = il;
          }
          const int slot = 2 * sm + (it < NA ? 0 : 1);
          const long cbase = sp::chunk(N, w, slot, row >> 6, k0 >> 5);
#pragma unroll
          for (int pl = 0; pl < 4; ++pl)
            *reinterpret_cast<sp::u32x4 *>(&spl[cbase + sp::in_chunk(pl, row, k0)]) = ph[pl];
        }
        if (bx == 0 && mat == 2 && tid == 0) g.wtag[2 * w] = (uint64_t)(g.pX + o);
      }
    }
  }
}

// ===========================================================================
// finish: diagonal |.| and e/h renormalisation (SBEs.py:5692-5702)
// ===========================================================================
// avg: also the diagonal part of the leapfrog average (the rest was fused
// into the transport); lazy: the step defers the average (psa_group.lazy),
// so this wire's flag is set; otherwise an existing flag is cleared (the
// caller averages now or the level is not a leapfrog level).
__global__ __launch_bounds__(256) void k_finish_diag(psa_group g, int w0, int avg, int lazy) {
  __shared__ double sh[8];
  const int w = w0 + blockIdx.x, N = g.N, tid = threadIdx.x;
  const long o = mat_off(g, w);
  double *C3 = g.CX + o, *D3 = g.DX + o;
  double se = 0, shh = 0;
  for (int k = tid; k < N; k += blockDim.x) {
    se += cabs_(ld(C3, (long)k * N + k));
    shh += cabs_(ld(D3, (long)k * N + k));
  }
  const double Se = block_sum(se, sh), Sh = block_sum(shh, sh);
  const double total = (Se + Sh) / 2.0;
  for (int k = tid; k < N; k += blockDim.x) {
    const double a = cabs_(ld(C3, (long)k * N + k)), b = cabs_(ld(D3, (long)k * N + k));
    const cd c3 = {a * total / (Se + 1e-200), 0.0}, d3 = {b * total / (Sh + 1e-200), 0.0};
    st(C3, (long)k * N + k, c3);
    st(D3, (long)k * N + k, d3);
    if (avg) {  // diagonal part of the leapfrog average (the rest was fused into Transport)
      const cd ca = ld(g.CA + o, (long)k * N + k), da = ld(g.DA + o, (long)k * N + k);
      st(g.CA + o, (long)k * N + k, cd{(ca.re + c3.re) / 2.0, (ca.im + c3.im) / 2.0});
      st(g.DA + o, (long)k * N + k, cd{(da.re + d3.re) / 2.0, (da.im + d3.im) / 2.0});
    }
  }
  if (g.lazy && tid == 0) g.lazy[w] = lazy ? 1 : 0;
}

// X2 = (X1 + X3) / 2 for p, C, D of every wire (SBEs.py:5704-5706)
__global__ __launch_bounds__(256) void k_average(psa_group g, int w0, int nw) {
  const long per = (long)g.N * g.N;
  const long total = per * nw * 3;
  if (g.lazy && blockIdx.x == 0)
    for (int t = threadIdx.x; t < nw; t += blockDim.x) g.lazy[w0 + t] = 0;
  for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < total; i += (long)gridDim.x * blockDim.x) {
    const int m = i / (per * nw);
    const long r = i % (per * nw);
    const long off = (long)w0 * per + r;
    double *A = m == 0 ? g.pA : (m == 1 ? g.CA : g.DA);
    const double *X = m == 0 ? g.pX : (m == 1 ? g.CX : g.DX);
    const double2 a = reinterpret_cast<double2 *>(A)[off];
    const double2 x = reinterpret_cast<const double2 *>(X)[off];
    reinterpret_cast<double2 *>(A)[off] = make_double2((a.x + x.x) / 2.0, (a.y + x.y) / 2.0);
  }
}

// The deferred leapfrog average written out (psa_average_levels): for the
// wires with lazy[w] set, A <- (A + B) / 2 for p, C, D; k_lazy_clear then
// clears their flags (after every block has read them).
__global__ __launch_bounds__(256) void k_lazy_average(psa_group g, int w0, int nw) {
  const long per = (long)g.N * g.N;
  const long total = per * nw * 3;
  for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < total; i += (long)gridDim.x * blockDim.x) {
    const int m = i / (per * nw);
    const long r = i % (per * nw);
    if (!g.lazy[w0 + r / per]) continue;
    const long off = (long)w0 * per + r;
    double *A = m == 0 ? g.pA : (m == 1 ? g.CA : g.DA);
    const double *B = m == 0 ? g.pB : (m == 1 ? g.CB : g.DB);
    const double2 a = reinterpret_cast<double2 *>(A)[off];
    const double2 b = reinterpret_cast<const double2 *>(B)[off];
    reinterpret_cast<double2 *>(A)[off] = make_double2((a.x + b.x) / 2.0, (a.y + b.y) / 2.0);
  }
}

__global__ void k_lazy_clear(psa_group g, int w0, int nw) {
  for (int t = threadIdx.x; t < nw; t += blockDim.x) g.lazy[w0 + t] = 0;
}

// ===========================================================================
// sources: QWPolarization3 (qwoptics.py:809-940) and QWRho5 (942-1071) in
// real space.  Both are sums over (k, k') of a matrix element times
// exp(i (k'-k) r); with kr on a uniform grid that is sum_d S_d exp(i d dkr r)
// over the diagonal sums S_d (k_diagsums mode 1): one product with the
// static table phi[d][r] on the GEMM engine (SrcProb).  k_src_norm then applies the boundary-mean
// subtraction, the particle-number normalisation, the 2 e0 ehint / area
// scaling (SBEs.py:5749-5750) and QW2Prop's |.| / common-total step
// (qwoptics.py:785-790), and evaluates CalcI0 (dcfield.py:1110-1152).
// ===========================================================================
// One block of SN_THREADS per wire.  Every global load is issued up front
// (one round of latency): the source lines into registers, the diagonals of
// the new and the averaged level and the interaction rows into LDS.  The
// O(N^2) EkReNorm sums of CalcI0 split each k over SN_SPLIT neighbouring
// lanes (their partial sums added in lane order).
constexpr int SN_THREADS = 1024, SN_RC = 2, SN_SPLIT = 4;  // SN_RC * SN_THREADS >= Nr for N <= 1023
__global__ __launch_bounds__(SN_THREADS) void k_src_norm(psa_group g, int w0) {
  __shared__ double sh[SN_THREADS / 64];
  __shared__ double s_ne2[1024], s_nh2[1024], s_vee[1024], s_vhh[1024], ecE[1024], ecH[1024];
  const int w = w0 + blockIdx.x, N = g.N, Nr = g.Nr, tid = threadIdx.x;
  const long o = mat_off(g, w);
  double *src = g.src + (long)w * PSA_SRC_STRIDE(N, Nr);
  const cd *rcv = reinterpret_cast<const cd *>(src + PSA_SRC_RC(Nr));  // [2][Nr]
  double *re_ = src + 3 * Nr, *rh_ = src + 4 * Nr;
  const double *vc = g.prep + (long)w * PSA_PREP_STRIDE(N) + PSA_PREP_VC(N);
  const bool lzw = g.lazy && g.lazy[w];
  // --- loads ---
  cd ce[SN_RC], ch[SN_RC];
#pragma unroll
  for (int u = 0; u < SN_RC; ++u) {
    const int r = tid + u * SN_THREADS;
    ce[u] = r < Nr ? rcv[r] : cd{0.0, 0.0};
    ch[u] = r < Nr ? rcv[Nr + r] : cd{0.0, 0.0};
  }
  double ne_t = 0, nh_t = 0;
  for (int k = tid; k < N; k += SN_THREADS) {
    const long kk = (long)k * N + k;
    ne_t += cabs_(ld(g.CX + o, kk));
    nh_t += cabs_(ld(g.DX + o, kk));
    // CalcI0 on ne2, nh2 = |diag of the averaged level| (with the average
    // deferred: (A + X) / 2 of the new level X)
    cd a2 = ld(g.CA + o, kk), b2 = ld(g.DA + o, kk);
    if (lzw) {
      const cd ax = ld(g.CX + o, kk), bx = ld(g.DX + o, kk);
      a2 = cd{(a2.re + ax.re) / 2.0, (a2.im + ax.im) / 2.0};
      b2 = cd{(b2.re + bx.re) / 2.0, (b2.im + bx.im) / 2.0};
    }
    s_ne2[k] = cabs_(a2);
    s_nh2[k] = cabs_(b2);
    s_vee[k] = vc[N + k];
    s_vhh[k] = vc[2 * N + k];
  }
  // boundary mean of the real parts (qwoptics.py:1062-1064)
  double bpe = 0, bph = 0;
#pragma unroll
  for (int u = 0; u < SN_RC; ++u) {
    const int r = tid + u * SN_THREADS;
    if (r < 10 || (r >= Nr - 10 && r < Nr)) {
      bpe += ce[u].re;
      bph += ch[u].re;
    }
  }
  const double be = block_sum(bpe, sh) / 20.0, bh = block_sum(bph, sh) / 20.0;
  const double NeT = block_sum(ne_t, sh) + 1e-50, NhT = block_sum(nh_t, sh) + 1e-50;
  const double dr = g.rq[1] - g.rq[0];
  double ae = 0, ah = 0;
#pragma unroll
  for (int u = 0; u < SN_RC; ++u)
    if (tid + u * SN_THREADS < Nr) {
      ae += hypot(ce[u].re - be, ce[u].im);
      ah += hypot(ch[u].re - bh, ch[u].im);
    }
  const double Ae = block_sum(ae, sh), Ah = block_sum(ah, sh);
  const double fe = NeT / (Ae * dr + 1e-100), fh = NhT / (Ah * dr + 1e-100);
  const double sc = 2.0 * g.e0 / g.area * g.ehint;
  double av[SN_RC], bv[SN_RC], s1 = 0, s2 = 0;
#pragma unroll
  for (int u = 0; u < SN_RC; ++u) {
    av[u] = bv[u] = 0.0;
    if (tid + u * SN_THREADS < Nr) {
      av[u] = hypot((ce[u].re - be) * fe * sc, ce[u].im * fe * sc);
      bv[u] = hypot((ch[u].re - bh) * fh * sc, ch[u].im * fh * sc);
      s1 += av[u];
      s2 += bv[u];
    }
  }
  const double S1 = block_sum(s1, sh), S2 = block_sum(s2, sh);
  const double total = (S1 * dr + S2 * dr + 1e-100) / 2.0;
  // without LF the reference never calls QWRho5: re = rh = 0 (SBEs.py:5732)
  const bool lf = (g.flags & PSA_F_LF) != 0;
#pragma unroll
  for (int u = 0; u < SN_RC; ++u) {
    const int r = tid + u * SN_THREADS;
    if (r < Nr) {
      re_[r] = lf ? av[u] * total / (S1 * dr + 1e-100) : 0.0;
      rh_[r] = lf ? bv[u] * total / (S2 * dr + 1e-100) : 0.0;
    }
  }
  // EkReNorm (dcfield.py:338-345): V[k][i] nonzero for i <= k; the
  // SN_SPLIT neighbouring lanes of a k sum i = part, part + SN_SPLIT, ...
  // and their partial sums are added in lane order
  const double vee0 = s_vee[0], vhh0 = s_vhh[0];
  const int lane = tid & 63;
  for (int kb = 0; kb < N; kb += SN_THREADS / SN_SPLIT) {
    const int pt = tid % SN_SPLIT, k = kb + tid / SN_SPLIT;
    double se = 0, shh = 0;
    if (k < N) {
#pragma unroll 8
      for (int i = pt; i <= k; i += SN_SPLIT) {
        se += s_ne2[i] * (vee0 - s_vee[k - i]);
        shh += s_nh2[i] * (vhh0 - s_vhh[k - i]);
      }
#pragma unroll 8
      for (int i = k + 1 + (pt - (k + 1) % SN_SPLIT + SN_SPLIT) % SN_SPLIT; i < N; i += SN_SPLIT) {
        se += s_ne2[i] * vee0;
        shh += s_nh2[i] * vhh0;
      }
    }
    double a = 0, b = 0;
#pragma unroll
    for (int q = 0; q < SN_SPLIT; ++q) {
      a += __shfl(se, (lane & ~(SN_SPLIT - 1)) + q, 64);
      b += __shfl(shh, (lane & ~(SN_SPLIT - 1)) + q, 64);
    }
    if (pt == 0 && k < N) {
      ecE[k] = g.Ee[k] + a / 2.0;
      ecH[k] = g.Eh[k] + b / 2.0;
    }
  }
  __syncthreads();
  double pe = 0, phh = 0, sne = 0, snh = 0;
  const double dk = g.dkr;
  for (int i = tid; i < N; i += SN_THREADS) {  // DriftVt (dcfield.py:347-360)
    double de, dh;
    if (i >= 1 && i <= N - 2) {
      de = (ecE[i + 1] - ecE[i - 1]) / 2.0 / dk;
      dh = (ecH[i + 1] - ecH[i - 1]) / 2.0 / dk;
    } else if (i == 0) {
      de = 2.0 * ((ecE[2] - ecE[0]) / 2.0 / dk) - (ecE[3] - ecE[1]) / 2.0 / dk;
      dh = 2.0 * ((ecH[2] - ecH[0]) / 2.0 / dk) - (ecH[3] - ecH[1]) / 2.0 / dk;
    } else {
      de = 2.0 * ((ecE[N - 1] - ecE[N - 3]) / 2.0 / dk) - (ecE[N - 2] - ecE[N - 4]) / 2.0 / dk;
      dh = 2.0 * ((ecH[N - 1] - ecH[N - 3]) / 2.0 / dk) - (ecH[N - 2] - ecH[N - 4]) / 2.0 / dk;
    }
    pe += de * s_ne2[i];
    phh += dh * s_nh2[i];
    sne += s_ne2[i];
    snh += s_nh2[i];
  }
  const double Pe = block_sum(pe, sh), Ph = block_sum(phh, sh);
  const double Sne = block_sum(sne, sh), Snh = block_sum(snh, sh);
  if (tid == 0) {
    const double ve = Pe / (1e-100 + Sne) / g.hbar, vh = Ph / (1e-100 + Snh) / g.hbar;
    double *ob = g.obs + (long)w * PSA_OBS_N;
    ob[0] = -g.e0 * (ve + vh) * Sne * g.dkr * 2.0;
    ob[1] = Sne;
    ob[2] = Snh;
    ob[3] = total;
  }
}

// ===========================================================================
// QW2Prop (qwoptics.py:785-805) and Rho = RhoH - RhoE (SBEs.py:5917): the
// spline products run on the GEMM engine (Qw2pProb), which leaves RhoE and
// RhoH in the real and imaginary slots of Rho; k_qw2prop_b (one block per
// wire) renormalises both to the common total and forms Rho.
// ===========================================================================
__global__ __launch_bounds__(256) void k_qw2prop_b(psa_group g, int w0) {
  __shared__ double sh[8];
  const int w = w0 + blockIdx.x, Nx = g.Nx, tid = threadIdx.x;
  double *Rho = g.out + ((3 * (long)g.Nw + w) * (long)Nx) * 2;
  double se = 0.0, sh2 = 0.0;
  for (int x = g.bx0 + tid; x < g.bx0 + g.bnx; x += blockDim.x) {
    se += Rho[2 * x];
    sh2 += Rho[2 * x + 1];
  }
  const double SE = block_sum(se, sh), SH = block_sum(sh2, sh);
  const bool lf = (g.flags & PSA_F_LF) != 0;
  const double total = g.obs[(long)w * PSA_OBS_N + 3];
  const double fe = lf ? total / (SE * g.drr + 1e-100) : 1.0;
  const double fh = lf ? total / (SH * g.drr + 1e-100) : 1.0;
  for (int x = g.bx0 + tid; x < g.bx0 + g.bnx; x += blockDim.x) {
    const double a = Rho[2 * x] * fe, b = Rho[2 * x + 1] * fh;
    st(Rho, x, cd{b - a, 0.0});
  }
}

// Lorentzian table for the off-diagonal dephasing products: block (qi, kind)
// writes its N - |q| packed rows, 2N entries each.
__global__ __launch_bounds__(256) void k_lorentz_table(psa_group g) {
  const int N = g.N, qi = blockIdx.x, kind = blockIdx.y, q = qi - (N - 1);
  const int kb = max(0, -q), ke = min(N, N - q);
  double *row0 = g.lor + ((long)kind * N * N + lorentz_row0(N, qi)) * 2 * N;
  const long n = (long)(ke - kb) * 2 * N;
  for (long i = threadIdx.x; i < n; i += blockDim.x) {
    const int kr = kb + (int)(i / (2 * N)), pp = (int)(i % (2 * N));
    row0[i] = lorentz_value(g, kind, q, kr, pp);
  }
}

// ===========================================================================
// Driver side of a served simulation (sbetestprop.py:183-198, 210-224):
// the per-wire pulse, spatially uniform, evaluated on the device at the
// wire's own clock, and a per-step trace of the recorded observables.
// drive[w][8] = E0x, E0y, E0z, omega0, tw, tp, t, (unused); t += dt per call.
// ===========================================================================
__global__ __launch_bounds__(256) void k_drive(psa_group g, int w0, double *drive, double dt) {
  const int w = w0 + blockIdx.y, Nx = g.Nx;
  const double *d = drive + (long)w * 8;
  const double t = d[6], w0f = d[3], tw = d[4], tp = d[5];
  const double x = w0f * (t - tp), a = w0f * tw, b = 2.0 * tw * w0f;
  const double xa = x / b, x20 = pow(xa, 20.0);
  const double env = exp(-(x * x) / (a * a)) * cos(x) * exp(-x20);
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < Nx; i += gridDim.x * blockDim.x) {
#pragma unroll
    for (int f = 0; f < 3; ++f)
      st(g.efield + (((long)f * g.Nw + w) * Nx) * 2, i, cd{d[f] * env, 0.0});
  }
}

__global__ void k_drive_clock(int w0, int nw, double *drive, double dt) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < nw) drive[(long)(w0 + i) * 8 + 6] += dt;
}

// trace[w][n][4] = Re Pxx[xmid], Im Pxx[xmid], sum ne2, sum nh2 at step n = count[w]
__global__ void k_record(psa_group g, int w0, int nw, int xmid, double *trace, int *count, int maxn) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nw) return;
  const int w = w0 + i, n = count[w];
  if (n >= maxn) return;
  const cd px = ld(g.out + (((long)0 * g.Nw + w) * g.Nx) * 2, xmid);
  double *tr = trace + ((long)w * maxn + n) * 4;
  tr[0] = px.re;
  tr[1] = px.im;
  tr[2] = g.obs[(long)w * PSA_OBS_N + 1];
  tr[3] = g.obs[(long)w * PSA_OBS_N + 2];
  count[w] = n + 1;
}

// ===========================================================================
// Many-body scattering lists on the device (the tables of CalcMBArrays,
// coulomb.py:367-499, as _MBCE_core/_MBCH_core read them, coulomb.py:956-1071):
// entries (i, j, l) of the stored (N+1)^3 table, Gaussian or (LorentzDelta)
// Lorentzian (stored index = physical index + 1; index 0 planes are zero)
// whose weight is not zero,
// CSR-grouped by the rate index they feed and ordered (i, j, l) within a
// target.  Block = one target index; the threads walk the two free indices
// in that order, 256 at a time, and a block-wide ballot scan keeps the order.
//   list  kind  target     condition
//   0     eh    i          j+l-i in [0,N), i >= l
//   1     eh    l          j+l-i in [0,N), i >= l
//   2     ee    i          i+j-l in [0,N), i >= l
//   3     ee    l          i+j-l in [0,N), i >= l
//   4     eh    j          j+l-i in [0,N), i >= l
//   5     eh    j+l-i      j+l-i in [0,N), l >= i
//   6     hh    i          i+j-l in [0,N), i >= l
//   7     hh    l          i+j-l in [0,N), i >= l
// ===========================================================================
__device__ __forceinline__ double mb_weight(int kind, int i, int j, int l, int N, const double *Ee,
                                            const double *Eh, double ge, double gh, double hbar, int lorentz) {
  if (i < 1 || j < 1 || l < 1) return 0.0;
  const int a = i - 1, b = j - 1, c = l - 1;
  double Ed, gam;
  int k30;
  if (kind == 0) {  // eh
    k30 = c + b - a;
    if (k30 < 0 || k30 >= N) return 0.0;
    gam = (ge + gh) / 2.0;
    Ed = Ee[a] + Eh[b] - Eh[k30] - Ee[c];
  } else {
    const double *E = kind == 1 ? Ee : Eh;
    k30 = a + b - c;
    if (k30 < 0 || k30 >= N) return 0.0;
    gam = kind == 1 ? ge : gh;
    Ed = E[a] + E[b] - E[k30] - E[c];
  }
  if (a == c || b == k30) return 0.0;
  const double hb = hbar * gam;
  if (lorentz) return 2.0 * gam / (Ed * Ed + hb * hb);  // LorentzDelta branch, coulomb.py:405-446
  const double x = Ed / hb;
  return 2.0 * PI / hbar * (1.0 / (sqrt(PI) * hb) * exp(-(x * x)));
}

template <int L>
__global__ __launch_bounds__(256) void k_mb_list(int N, const double *Ee, const double *Eh, double ge,
                                                 double gh, double hbar, int lorentz, const int *off, int *counts,
                                                 uint32_t *idx, double *val) {
  __shared__ int wtot[4];
  const int t = blockIdx.x, tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
  constexpr int kind = (L == 2 || L == 3) ? 1 : ((L == 6 || L == 7) ? 2 : 0);
  constexpr bool jli = (kind == 0);  // third index j+l-i (eh) or i+j-l (ee, hh)
  const long NN = (long)N * N;
  int run = off ? off[t] : 0;
  for (long f0 = 0; f0 < NN; f0 += 256) {
    const long f = f0 + tid;
    const int x = (int)(f / N), y = (int)(f % N);
    int i, j, l;
    if (L == 0 || L == 2 || L == 6) { i = t; j = x; l = y; }
    else if (L == 1 || L == 3 || L == 7) { l = t; i = x; j = y; }
    else if (L == 4) { j = t; i = x; l = y; }
    else { i = x; j = y; l = t + i - j; }
    bool ok = f < NN && l >= 0 && l < N;
    const int i3 = jli ? j + l - i : i + j - l;
    ok = ok && i3 >= 0 && i3 < N && (L == 5 ? l >= i : i >= l);
    const double v = ok ? mb_weight(kind, i, j, l, N, Ee, Eh, ge, gh, hbar, lorentz) : 0.0;
    ok = ok && v > 0.0;
    const unsigned long long m = __ballot(ok);
    const int pre = __popcll(m & ((1ull << lane) - 1ull));
    if (lane == 0) wtot[wv] = __popcll(m);
    __syncthreads();
    int wbase = 0, tot = 0;
    for (int u = 0; u < 4; ++u) {
      if (u < wv) wbase += wtot[u];
      tot += wtot[u];
    }
    if (off && ok) {
      const int pos = run + wbase + pre;
      idx[pos] = (uint32_t)i | ((uint32_t)j << 10) | ((uint32_t)l << 20);
      val[pos] = v;
    }
    run += tot;
    __syncthreads();
  }
  if (!off && tid == 0) counts[t] = run;
}

// ===========================================================================
// host launchers
// ===========================================================================
namespace {

template <typename T, class P>
int launch_gemm_t(const P &p, dim3 grid, hipStream_t s, const char *what) {
  constexpr size_t shm = gemm_lds_bytes<T>();
  static bool attr_set = false;  // per instantiation
  if (shm > 65536 && !attr_set) {
    int rc = hip_check(hipFuncSetAttribute((const void *)gemm_kernel<T, P>,
                                           hipFuncAttributeMaxDynamicSharedMemorySize, (int)shm),
                       "hipFuncSetAttribute(gemm_kernel)");
    if (rc) return rc;
    attr_set = true;
  }
  hipLaunchKernelGGL((gemm_kernel<T, P>), grid, dim3(256), shm, s, p);
  PSA_LAUNCH_CHECK(what);
  return PSA_OK;
}

template <class P>
int launch_gemm(const P &p, int prec, int M, int Ncols, int batches, hipStream_t s, const char *what) {
  dim3 grid((Ncols + GBN - 1) / GBN, (M + GBM - 1) / GBM, batches);
  if (prec == PSA_PREC_FP64) return launch_gemm_t<double>(p, grid, s, what);
  return launch_gemm_t<float>(p, grid, s, what);
}

// A single-batch FP64 product of M rows x ncol columns over K, split over
// K when the tile grid alone would leave most of the GPU idle and the
// group provides the scratch (g.kscr): about two workgroups per CU, at most
// 8 splits, each at least one 16-deep slab.
template <class P>
int launch_gemm_splitk(const P &p, int M, int ncol, int K, hipStream_t s, const char *what) {
  const psa_group &g = p.g;
  static int cus = 0;
  if (!cus) {
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess || hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || cus <= 0)
      cus = 256;
  }
  const int tiles = ((M + GBM - 1) / GBM) * ((ncol + GBN - 1) / GBN);
  int S = min(SPLITK_MAX, max(1, (2 * cus + tiles - 1) / tiles));
  S = min(S, (K + GBK - 1) / GBK);
  while (S > 1 && (long)S * M * ncol * 2 > g.kscr_len) --S;
  if (S <= 1 || !g.kscr) return launch_gemm(p, PSA_PREC_FP64, M, ncol, 1, s, what);
  SplitK<P> q;
  q.p = p;
  q.g = g;
  q.ks = ((K + S - 1) / S + GBK - 1) / GBK * GBK;
  q.M = M;
  q.ncol = ncol;
  q.scr = g.kscr;
  S = (K + q.ks - 1) / q.ks;  // splits with work
  if (int rc = launch_gemm(q, PSA_PREC_FP64, M, ncol, S, s, what)) return rc;
  const long n = (long)M * ncol;
  hipLaunchKernelGGL(splitk_reduce<P>, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, p,
                     (const double *)g.kscr, S, M, ncol);
  PSA_LAUNCH_CHECK("splitk_reduce");
  return PSA_OK;
}

// The split-bf16 operand planes of wires [w0, w0 + nw) (sp::k_split, and
// sp::k_split_band for the off-diagonal dephasing bands).
// the operand planes of p, C, D (unless carried; the RHS kernel splits the
// Hamiltonians itself as it stages their slabs)
int split_main(const psa_group &g, int w0, int nw, hipStream_t s) {
  if (!g.spl) return fail(PSA_E_INVALID, "bf16x3 mode needs the spl operand buffer");
  const int nt = (g.N + 63) / 64;
  hipLaunchKernelGGL(sp::k_split, dim3(nt, nt, 3 * nw), dim3(256), 0, s, g, w0);
  PSA_LAUNCH_CHECK("sp::k_split");
  return PSA_OK;
}

// the band planes of the off-diagonal dephasing tables
int split_band(const psa_group &g, int w0, int nw, hipStream_t s) {
  if (!g.spl) return fail(PSA_E_INVALID, "bf16x3 mode needs the spl operand buffer");
  if (g.flags & PSA_F_OFFDIAGDPH) {
    const int nt = (g.N + 63) / 64;
    hipLaunchKernelGGL(sp::k_split_band, dim3(nt, 2, nw), dim3(256), 0, s, g, w0);
    PSA_LAUNCH_CHECK("sp::k_split_band");
  }
  return PSA_OK;
}

int split_planes(const psa_group &g, int w0, int nw, hipStream_t s) {
  if (int rc = split_main(g, w0, nw, s)) return rc;
  return split_band(g, w0, nw, s);
}

// Split-bf16 RHS: operand planes once per step, then the bf16 MFMA
// products (sp::k_rhs_pp, one persistent workgroup of 8 waves per CU).
int rhs_bf16(const RhsProb &p, int nw, hipStream_t s, bool split = true) {
  static int cus = 0;
  if (!cus) {
    int rc = hip_check(hipFuncSetAttribute((const void *)sp::k_rhs_pp, hipFuncAttributeMaxDynamicSharedMemorySize,
                                           (int)sp::pp_lds_bytes()),
                       "hipFuncSetAttribute(sp::k_rhs_pp)");
    if (rc) return rc;
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess ||
        hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || cus <= 0)
      cus = 256;
  }
  if (split)
    if (int rc = split_planes(p.g, p.w0, nw, s)) return rc;
  hipLaunchKernelGGL(sp::k_rhs_pp, dim3(cus), dim3(sp::PP_THREADS), sp::pp_lds_bytes(), s, p, nw);
  PSA_LAUNCH_CHECK("sp::k_rhs_pp");
  return PSA_OK;
}

// ---------------------------------------------------------------------------
// RHS of several wire groups (any mix of N, one precision mode) in one
// persistent launch over work queues.  Work items are the 64 x 64 output
// tiles of every (group, wire, equation).  The wires, in decreasing N, are
// dealt round robin to eight lists, one per XCD (workgroup i runs on XCD
// i % 8), each list wire-major so a wire's tiles share that XCD's L2.  A
// workgroup takes items from its own list's atomic counter and, when that
// list is exhausted, from the other lists in turn, until all are; the last
// workgroup to leave resets the counters, so the next launch (or graph
// replay) finds them at zero without a memset.  The tile bodies are the
// single-group ones, so every tile is computed exactly as psa_rhs does.
// ---------------------------------------------------------------------------
constexpr int NXCD = 8;

struct MultiHdr {
  int32_t n, prec, kind, nwires;
  int32_t xstart[NXCD + 1];  // wire entries of list x: [xstart[x], xstart[x+1])
  int32_t xcnt[NXCD];        // items in list x
  int32_t gx[PSA_RHS_MULTI_MAX];
  int32_t nw[PSA_RHS_MULTI_MAX];
};
struct WireEnt {
  int32_t gi, j, off, pad;  // group (packed order), wire offset in its range, first item in the list
};
constexpr size_t multi_hdr_bytes() { return (sizeof(MultiHdr) + 255) / 256 * 256; }
constexpr size_t multi_probs_bytes(int n) { return ((size_t)n * sizeof(RhsProb) + 255) / 256 * 256; }

struct RhsQueue {
  int32_t xstart[NXCD + 1];
  int32_t xcnt[NXCD];
  int32_t gx[PSA_RHS_MULTI_MAX];
};

template <int KIND>
constexpr size_t queue_tile_bytes() {
  return KIND == 2 ? sp::pp_lds_bytes() : (KIND == 0 ? gemm_lds_bytes<double>() : gemm_lds_bytes<float>());
}

// threads per workgroup and workgroups per CU of each kind
template <int KIND>
constexpr int queue_threads() {
  return KIND == 2 ? sp::PP_THREADS : 256;
}
template <int KIND>
constexpr int queue_occupancy() {
  return KIND == 2 ? 1 : 2;
}

// output tile edge of the work items of each kind
template <int KIND>
constexpr int queue_tile() {
  return KIND == 2 ? sp::TILE : GBM;
}

template <int KIND>  // 0: FP64 products, 1: FP32, 2: split-BF16
__global__ __launch_bounds__(queue_threads<KIND>(), queue_occupancy<KIND>()) void k_rhs_queue(const RhsProb *__restrict__ probs, const WireEnt *__restrict__ ents,
                                                   RhsQueue q, unsigned *queue) {
  // Dynamic LDS only: the operand tiles at offset 0 (aligned as the tile
  // bodies assume), the broadcast item index after them.
  extern __shared__ char lds_raw[];
  int *s_item = reinterpret_cast<int *>(lds_raw + queue_tile_bytes<KIND>());
  const int home = blockIdx.x % NXCD;
  for (int xo = 0; xo < NXCD; ++xo) {
    const int x = (home + xo) % NXCD;
    for (;;) {
      if (threadIdx.x == 0) *s_item = (int)atomicAdd(&queue[x], 1u);
      __syncthreads();
      const int it = __builtin_amdgcn_readfirstlane(*s_item);
      __syncthreads();
      if (it >= q.xcnt[x]) break;
      // the list's last wire entry starting at or before item it
      int lo = q.xstart[x], hi = q.xstart[x + 1] - 1;
      while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (ents[mid].off <= it) lo = mid;
        else hi = mid - 1;
      }
      const WireEnt e = ents[lo];
      const int gx = q.gx[e.gi], tpb = gx * gx, local = it - e.off;
      const int b = 3 * e.j + local / tpb, lin = local % tpb;
      const int m0 = (lin / gx) * queue_tile<KIND>(), n0 = (lin % gx) * queue_tile<KIND>();
      if constexpr (KIND == 2)
        sp::pp_tile(probs[e.gi], b, m0, n0, lds_raw);
      else
        gemm_tile<typename std::conditional<KIND == 0, double, float>::type, RhsProb>(probs[e.gi], b, m0, n0,
                                                                                      lds_raw);
    }
  }
  if (threadIdx.x == 0) {
    __threadfence();
    if (atomicAdd(&queue[NXCD], 1u) == gridDim.x - 1) {
      for (int x = 0; x <= NXCD; ++x) atomicExch(&queue[x], 0u);
    }
  }
}

template <int KIND>
int launch_rhs_queue(const RhsProb *probs_dev, const WireEnt *ents_dev, const RhsQueue &q, int total, unsigned *queue,
                     hipStream_t s) {
  constexpr size_t shm = queue_tile_bytes<KIND>() + 16;
  static int blocks = 0;  // resident workgroups on the whole device, per instantiation
  if (!blocks) {
    int rc = hip_check(hipFuncSetAttribute((const void *)k_rhs_queue<KIND>, hipFuncAttributeMaxDynamicSharedMemorySize,
                                           (int)shm),
                       "hipFuncSetAttribute(k_rhs_queue)");
    if (rc) return rc;
    int per = 0, dev = 0, cus = 0;
    rc = hip_check(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per, k_rhs_queue<KIND>, queue_threads<KIND>(), shm),
                   "hipOccupancyMaxActiveBlocksPerMultiprocessor");
    if (rc) return rc;
    if (hipGetDevice(&dev) != hipSuccess || hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess)
      return fail(PSA_E_HIP, "psa_rhs_multi: device query failed");
    blocks = std::max(1, per) * std::max(1, cus);
  }
  const int grid = std::min(blocks, std::max(total, NXCD));
  hipLaunchKernelGGL(k_rhs_queue<KIND>, dim3(grid), dim3(queue_threads<KIND>()), shm, s, probs_dev, ents_dev, q,
                     queue);
  PSA_LAUNCH_CHECK("k_rhs_queue");
  return PSA_OK;
}

}  // namespace

extern "C" {

int psa_abi_version(void) { return PSA_ABI_VERSION; }

const char *psa_last_error(void) { return g_err.c_str(); }

int psa_group_size(void) { return (int)sizeof(psa_group); }

int psa_vint_table(const double *Qy, int32_t NQ, const double *y, int32_t Ny, double alphae,
                   double alphah, double Delta0, double pref, double *out, void *stream) {
  if (!Qy || !y || !out || NQ < 1 || Ny < 4) return fail(PSA_E_INVALID, "psa_vint_table: bad arguments");
  hipLaunchKernelGGL(k_vint, dim3(NQ, 3), dim3(256), 0, (hipStream_t)stream, Qy, NQ, y, Ny, alphae,
                     alphah, Delta0, pref, out);
  PSA_LAUNCH_CHECK("k_vint");
  return PSA_OK;
}

int psa_prep(const psa_group *g, int32_t w0, int32_t nw, void *stream) {
  if (int rc = check_group(g, w0, nw)) return rc;
  const hipStream_t s = (hipStream_t)stream;
  hipLaunchKernelGGL(k_prep_a, dim3(nw), dim3(256), 0, s, *g, w0);
  PSA_LAUNCH_CHECK("k_prep_a");
  hipLaunchKernelGGL(k_diagsums, dim3(nw, (2 * g->N - 1 + 63) / 64), dim3(64 * DS_WAVES), 0, s, *g, w0, 0);
  PSA_LAUNCH_CHECK("k_diagsums");
  hipLaunchKernelGGL(k_prep_b, dim3(nw, (g->N + PB_WAVES - 1) / PB_WAVES), dim3(64 * PB_WAVES), 0, s, *g, w0);
  PSA_LAUNCH_CHECK("k_prep_b");
  return PSA_OK;
}

int psa_calc_h(const psa_group *g, int32_t w0, int32_t nw, void *stream) {
  if (int rc = check_group(g, w0, nw)) return rc;
  const hipStream_t s = (hipStream_t)stream;
  if (!(g->flags & PSA_F_EXCITONS)) {
    hipLaunchKernelGGL(k_h_simple, dim3((g->N * g->N + 255) / 256, nw), dim3(256), 0, s, *g, w0);
    PSA_LAUNCH_CHECK("k_h_simple");
    return PSA_OK;
  }
  // FP64, or FP32 MFMA in the reduced-precision modes (V ~ 1e-21 J and the
  // products stay inside the FP32 range); the Meh, Hartree and band-energy
  // parts are added in FP64 in the epilogue
  const int prec = g->prec == PSA_PREC_FP64 ? PSA_PREC_FP64 : PSA_PREC_FP32;
  // FP32: the shifted-Toeplitz kernel; FP64 runs faster on the GEMM engine
  // (the FP64 MFMA rate leaves the shifted kernel's larger row count behind)
  if (g->prec == PSA_PREC_BF16X3 && g->N <= chx::NMAX) {
    // split-BF16 products on the same shifted-Toeplitz blocks
    const int cols = (g->N + 31) / 32;
    constexpr int R = chx::bfx::ROWS;
    hipLaunchKernelGGL(chx::k_calch_bf, dim3(3 * nw, (g->N + R - 1) / R, cols), dim3(chx::bfx::THREADS), 0, s, *g,
                       w0);
    PSA_LAUNCH_CHECK("k_calch_bf");
    return PSA_OK;
  }
  if (prec == PSA_PREC_FP32 && g->N <= chx::NMAX) {
    const int cols = (g->N + 31) / 32;  // column blocks: 32 positive and 32 non-positive d each
    constexpr int R = chx::rows_per_block<float>();
    hipLaunchKernelGGL(chx::k_calch<float>, dim3(3 * nw, (g->N + R - 1) / R, cols),
                       dim3(chx::threads_per_block<float>()), 0, s, *g, w0);
    PSA_LAUNCH_CHECK("k_calch");
    return PSA_OK;
  }
  CalcHProb p;
  p.g = *g;
  p.w0 = w0;
  return launch_gemm(p, prec, g->N, 2 * g->N - 1, 3 * nw, s, "calc_h gemm");
}

int psa_mb_list(int32_t list, int32_t N, const double *Ee, const double *Eh, double ge, double gh,
                double hbar, int32_t lorentz, const int32_t *off, int32_t *counts, uint32_t *idx, double *val,
                void *stream) {
  if (list < 0 || list > 7 || N < 2 || N > 1023 || !Ee || !Eh)
    return fail(PSA_E_INVALID, "psa_mb_list: bad arguments");
  if (off ? (!idx || !val) : !counts) return fail(PSA_E_INVALID, "psa_mb_list: missing output");
  const hipStream_t s = (hipStream_t)stream;
  switch (list) {
#define PSA_MBL(L)                                                                                \
  case L:                                                                                         \
    hipLaunchKernelGGL(k_mb_list<L>, dim3(N), dim3(256), 0, s, N, Ee, Eh, ge, gh, hbar, lorentz, off,  \
                       counts, idx, val);                                                         \
    break;
    PSA_MBL(0) PSA_MBL(1) PSA_MBL(2) PSA_MBL(3) PSA_MBL(4) PSA_MBL(5) PSA_MBL(6) PSA_MBL(7)
#undef PSA_MBL
  }
  PSA_LAUNCH_CHECK("k_mb_list");
  return PSA_OK;
}

int psa_lorentz_table(const psa_group *g, void *stream) {
  if (!g || !g->lor || !g->Ee || !g->Eh || g->N < 2 || g->N > 1024)
    return fail(PSA_E_INVALID, "psa_lorentz_table: bad group");
  hipLaunchKernelGGL(k_lorentz_table, dim3(2 * g->N - 1, 2), dim3(256), 0, (hipStream_t)stream, *g);
  PSA_LAUNCH_CHECK("k_lorentz_table");
  return PSA_OK;
}

int psa_offdiag_tables(const psa_group *g, int32_t w0, int32_t nw, void *stream) {
  if (int rc = check_group(g, w0, nw)) return rc;
  // FP64, or (reduced-precision modes) FP32 MFMA or split-BF16 MFMA on
  // 2^100-scaled occupation columns (positive terms only: no cancellation);
  // the stored-table form runs FP32 in both reduced modes
  const int prec = g->prec == PSA_PREC_FP64 ? PSA_PREC_FP64 : PSA_PREC_FP32;
  const hipStream_t s = (hipStream_t)stream;
  if (!g->lor) {
    // Lorentzian generated in registers
    if (g->N > odg::NMAX) return fail(PSA_E_UNSUPPORTED, "psa_offdiag_tables: N > 512 needs the lor table");
    const dim3 grid((g->N + odg::ROWS - 1) / odg::ROWS, (2 * nw + odg::COLS - 1) / odg::COLS, 2 * (2 * g->N - 1));
    if (g->prec == PSA_PREC_BF16X3)
      hipLaunchKernelGGL(odg::k_offdiag_bf,
                         dim3((g->N + odg::bfx::ROWS - 1) / odg::bfx::ROWS, (nw + 63) / 64, 2 * (2 * g->N - 1)),
                         dim3(odg::bfx::THREADS), 0, s, *g, w0, nw);
    else if (prec == PSA_PREC_FP64)
      hipLaunchKernelGGL(odg::k_offdiag<double>, grid, dim3(odg::Cfg<double>::THREADS), 0, s, *g, w0, nw, 0);
    else
      hipLaunchKernelGGL(odg::k_offdiag<float>, grid, dim3(odg::Cfg<float>::THREADS), 0, s, *g, w0, nw, 1);
    PSA_LAUNCH_CHECK("k_offdiag");
    return PSA_OK;
  }
  OffdiagProb p;
  p.g = *g;
  p.w0 = w0;
  p.nw = nw;
  p.scaled = prec == PSA_PREC_FP32 ? 1 : 0;
  return launch_gemm(p, prec, g->N, 2 * nw, 2 * (2 * g->N - 1), s, "offdiag gemm");
}

}  // extern "C"

namespace {
// psa_rhs; with split false (split-BF16) the caller has written the
// operand planes already
int rhs_launch(const psa_group *g, int32_t w0, int32_t nw, int32_t mode, double dt, hipStream_t s, bool split) {
  if (int rc = check_group(g, w0, nw)) return rc;
  if (mode < 0 || mode > 5) return fail(PSA_E_INVALID, "psa_rhs: mode must be in [0, 5]");
  if (mode >= 2 && (!g->pR || !g->CR || !g->DR))
    return fail(PSA_E_INVALID, "psa_rhs: RK4 stages need pR, CR, DR");
  RhsProb p;
  p.g = *g;
  p.w0 = w0;
  p.mode = mode;
  p.dt = dt;
  if (g->prec == PSA_PREC_BF16X3) return rhs_bf16(p, nw, s, split);
  return launch_gemm(p, g->prec, g->N, g->N, 3 * nw, s, "rhs gemm");
}
}  // namespace

extern "C" {

int psa_rhs(const psa_group *g, int32_t w0, int32_t nw, int32_t mode, double dt, void *stream) {
  return rhs_launch(g, w0, nw, mode, dt, (hipStream_t)stream, true);
}

int64_t psa_rhs_multi_bytes(const int32_t *nw, int32_t n) {
  if (!nw || n < 1 || n > PSA_RHS_MULTI_MAX) return -1;
  long W = 0;
  for (int i = 0; i < n; ++i) {
    if (nw[i] < 1) return -1;
    W += nw[i];
  }
  return (int64_t)(multi_hdr_bytes() + multi_probs_bytes(n) + (size_t)W * sizeof(WireEnt));
}

int psa_rhs_multi_pack(const psa_group *const *groups, const int32_t *w0, const int32_t *nw, int32_t n, int32_t mode,
                       double dt, void *packed) {
  if (!groups || !w0 || !nw || !packed || n < 1 || n > PSA_RHS_MULTI_MAX)
    return fail(PSA_E_INVALID, "psa_rhs_multi_pack: bad arguments (1 <= n <= PSA_RHS_MULTI_MAX)");
  if (mode < 0 || mode > 5) return fail(PSA_E_INVALID, "psa_rhs_multi_pack: mode must be in [0, 5]");
  int order[PSA_RHS_MULTI_MAX];
  for (int i = 0; i < n; ++i) {
    if (int rc = check_group(groups[i], w0[i], nw[i])) return rc;
    if (groups[i]->prec != groups[0]->prec) return fail(PSA_E_INVALID, "psa_rhs_multi_pack: groups differ in precision");
    if (mode >= 2 && (!groups[i]->pR || !groups[i]->CR || !groups[i]->DR))
      return fail(PSA_E_INVALID, "psa_rhs_multi_pack: RK4 stages need pR, CR, DR");
    if (groups[i]->prec == PSA_PREC_BF16X3 && !groups[i]->spl)
      return fail(PSA_E_INVALID, "psa_rhs_multi_pack: bf16x3 mode needs the spl operand buffer");
    order[i] = i;
  }
  std::stable_sort(order, order + n, [&](int a, int b) { return groups[a]->N > groups[b]->N; });
  MultiHdr h;
  std::memset(&h, 0, sizeof(h));
  h.n = n;
  h.prec = groups[0]->prec;
  h.kind = h.prec == PSA_PREC_FP64 ? 0 : (h.prec == PSA_PREC_FP32 ? 1 : 2);
  char *base = static_cast<char *>(packed);
  RhsProb *probs = reinterpret_cast<RhsProb *>(base + multi_hdr_bytes());
  WireEnt *ents = reinterpret_cast<WireEnt *>(base + multi_hdr_bytes() + multi_probs_bytes(n));
  std::vector<WireEnt> lists[NXCD];
  long cnt[NXCD] = {0};
  int ord = 0;
  for (int j = 0; j < n; ++j) {
    const int i = order[j];
    RhsProb p;
    std::memset(&p, 0, sizeof(p));
    p.g = *groups[i];
    p.w0 = w0[i];
    p.mode = mode;
    p.dt = dt;
    std::memcpy(&probs[j], &p, sizeof(p));
    const int tile = h.kind == 2 ? sp::TILE : GBM;
    h.gx[j] = (groups[i]->N + tile - 1) / tile;
    h.nw[j] = nw[i];
    for (int k = 0; k < nw[i]; ++k, ++ord) {
      const int x = ord % NXCD;
      lists[x].push_back(WireEnt{j, k, (int32_t)cnt[x], 0});
      cnt[x] += 3L * h.gx[j] * h.gx[j];
      if (cnt[x] > 0x7fffffffL) return fail(PSA_E_INVALID, "psa_rhs_multi_pack: too many tiles");
    }
  }
  h.nwires = ord;
  int e = 0;
  for (int x = 0; x < NXCD; ++x) {
    h.xstart[x] = e;
    h.xcnt[x] = (int32_t)cnt[x];
    for (const WireEnt &w : lists[x]) ents[e++] = w;
  }
  h.xstart[NXCD] = e;
  std::memcpy(packed, &h, sizeof(h));
  return PSA_OK;
}

}  // extern "C"

namespace {
// psa_rhs_multi; with splits false the caller has written every group's
// bf16 operand planes already
int rhs_multi_impl(const void *packed_host, const void *packed_dev, uint32_t *queue, hipStream_t s, bool splits) {
  if (!packed_host || !packed_dev || !queue) return fail(PSA_E_INVALID, "psa_rhs_multi: null argument");
  MultiHdr h;
  std::memcpy(&h, packed_host, sizeof(h));
  if (h.n < 1 || h.n > PSA_RHS_MULTI_MAX || h.nwires < 1 || h.kind < 0 || h.kind > 2 || h.xstart[NXCD] != h.nwires)
    return fail(PSA_E_INVALID, "psa_rhs_multi: not a packed description");
  const char *bh = static_cast<const char *>(packed_host), *bd = static_cast<const char *>(packed_dev);
  if (h.kind == 2 && splits) {
    const RhsProb *probs_h = reinterpret_cast<const RhsProb *>(bh + multi_hdr_bytes());
    for (int j = 0; j < h.n; ++j) {
      RhsProb p;
      std::memcpy(&p, &probs_h[j], sizeof(p));
      if (int rc = split_planes(p.g, p.w0, h.nw[j], s)) return rc;
    }
  }
  RhsQueue q;
  long total = 0;
  for (int x = 0; x <= NXCD; ++x) q.xstart[x] = h.xstart[x];
  for (int x = 0; x < NXCD; ++x) {
    q.xcnt[x] = h.xcnt[x];
    total += h.xcnt[x];
  }
  for (int j = 0; j < PSA_RHS_MULTI_MAX; ++j) q.gx[j] = h.gx[j];
  const RhsProb *probs_d = reinterpret_cast<const RhsProb *>(bd + multi_hdr_bytes());
  const WireEnt *ents_d = reinterpret_cast<const WireEnt *>(bd + multi_hdr_bytes() + multi_probs_bytes(h.n));
  unsigned *qu = reinterpret_cast<unsigned *>(queue);
  const int tot = (int)std::min(total, 0x7fffffffL);
  if (h.kind == 0) return launch_rhs_queue<0>(probs_d, ents_d, q, tot, qu, s);
  if (h.kind == 1) return launch_rhs_queue<1>(probs_d, ents_d, q, tot, qu, s);
  return launch_rhs_queue<2>(probs_d, ents_d, q, tot, qu, s);
}
}  // namespace

extern "C" {

int psa_rhs_multi(const void *packed_host, const void *packed_dev, uint32_t *queue, void *stream) {
  return rhs_multi_impl(packed_host, packed_dev, queue, (hipStream_t)stream, true);
}

int psa_relaxation(const psa_group *g, int32_t w0, int32_t nw, double dt, void *stream) {
  if (int rc = check_group(g, w0, nw)) return rc;
  const hipStream_t s = (hipStream_t)stream;
  if (!(g->flags & (PSA_F_EHS | PSA_F_PHONON))) return PSA_OK;
  hipLaunchKernelGGL(k_diag_gather, dim3(nw), dim3(256), 0, s, *g, w0);
  PSA_LAUNCH_CHECK("k_diag_gather");
  if (g->flags & PSA_F_EHS) {
    const size_t shm = 4 * (size_t)g->N * MB_WC * sizeof(double);
    if (g->N > 512) return fail(PSA_E_UNSUPPORTED, "many-body scattering needs N <= 512 (LDS tables)");
    if (shm > 65536) {
      static bool attr_set = false;
      if (!attr_set) {
        // the largest table set (N = 512); k_mb also has a small static array
        int rc = hip_check(hipFuncSetAttribute((const void *)k_mb, hipFuncAttributeMaxDynamicSharedMemorySize,
                                               (int)(4 * 512 * MB_WC * sizeof(double))),
                           "hipFuncSetAttribute(k_mb)");
        if (rc) return rc;
        attr_set = true;
      }
    }
    if (g->prec != PSA_PREC_FP64) {
      // reduced modes: FP32 tables for 16 wires per block
      const size_t shm2 = 4 * (size_t)g->N * MB2_WC * sizeof(float);
      static bool attr2_set = false;
      if (shm2 > 65536 && !attr2_set) {
        int rc = hip_check(hipFuncSetAttribute((const void *)k_mb2, hipFuncAttributeMaxDynamicSharedMemorySize,
                                               (int)(4 * 512 * MB2_WC * sizeof(float))),
                           "hipFuncSetAttribute(k_mb2)");
        if (rc) return rc;
        attr2_set = true;
      }
      hipLaunchKernelGGL(k_mb2, dim3((g->N + MB_KG - 1) / MB_KG, 4, (nw + MB2_WC - 1) / MB2_WC),
                         dim3(MB_THREADS), shm2, s, *g, w0, nw);
      PSA_LAUNCH_CHECK("k_mb2");
    } else {
      hipLaunchKernelGGL(k_mb, dim3((g->N + MB_KG - 1) / MB_KG, 4, (nw + MB_WC - 1) / MB_WC), dim3(MB_THREADS),
                         shm, s, *g, w0, nw);
      PSA_LAUNCH_CHECK("k_mb");
    }
  }
  hipLaunchKernelGGL(k_relax, dim3(nw, (g->N + 63) / 64), dim3(64 * RL_WAVES), 0, s, *g, w0, dt);
  PSA_LAUNCH_CHECK("k_relax");
  return PSA_OK;
}

}  // extern "C"

namespace {

// one pass of the compile-time-N transport (0 rows, 1 columns) with LINES lines per block
template <int NT, int LINES>
int launch_transport_fp(const psa_group *g, int32_t w0, int nw, hipStream_t s, int pass, bool avg, bool planes) {
  const size_t shm = ((size_t)LINES * tr_line_stride_s(NT, LINES) + plan_twiddles(NT)) * sizeof(cd);
  static bool attr_set = false;  // per instantiation
  if (shm > 65536 && !attr_set) {
    int rc = hip_check(hipFuncSetAttribute((const void *)k_transport_f<NT, LINES>,
                                           hipFuncAttributeMaxDynamicSharedMemorySize, 160 * 1024),
                       "hipFuncSetAttribute(k_transport_f)");
    if (rc) return rc;
    attr_set = true;
  }
  const dim3 grid((NT + LINES - 1) / LINES, 3, nw);
  hipLaunchKernelGGL((k_transport_f<NT, LINES>), grid, dim3(64 * LINES), shm, s, *g, w0, pass,
                     (pass == 1 && avg) ? 1 : 0, (pass == 1 && planes && LINES % 8 == 0) ? 1 : 0);
  PSA_LAUNCH_CHECK("k_transport_f");
  return PSA_OK;
}

template <int NT>
int launch_transport_fpass(const psa_group *g, int32_t w0, int nw, hipStream_t s, int pass, bool avg, bool planes) {
  switch (pass == 0 ? g->tune[1] : g->tune[3]) {
    case 2: return launch_transport_fp<NT, 2>(g, w0, nw, s, pass, avg, planes);
    case 8: return launch_transport_fp<NT, 8>(g, w0, nw, s, pass, avg, planes);
    case 16: return launch_transport_fp<NT, 16>(g, w0, nw, s, pass, avg, planes);
    default: return launch_transport_fp<NT, 4>(g, w0, nw, s, pass, avg, planes);
  }
}

template <int NT, int LINES>
int launch_transport_l(const psa_group *g, int32_t w0, int nw, hipStream_t s, bool avg) {
  static_assert(NT == 0, "compile-time N uses k_transport_f");
  {
    const size_t shm = ((size_t)LINES * tr_line_stride(g->N, LINES) + g->N) * sizeof(cd);  // lines + twiddles
    if (shm > 65536) {
      static bool attr_set = false;  // per instantiation
      if (!attr_set) {
        int rc = hip_check(hipFuncSetAttribute((const void *)k_transport<NT, LINES>,
                                               hipFuncAttributeMaxDynamicSharedMemorySize, 160 * 1024),
                           "hipFuncSetAttribute(k_transport)");
        if (rc) return rc;
        attr_set = true;
      }
    }
    const dim3 grid((g->N + LINES - 1) / LINES, 3, nw);
    hipLaunchKernelGGL((k_transport<NT, LINES>), grid, dim3(64 * LINES), shm, s, *g, w0, 0, 0);
    PSA_LAUNCH_CHECK("k_transport");
    hipLaunchKernelGGL((k_transport<NT, LINES>), grid, dim3(64 * LINES), shm, s, *g, w0, 1, avg ? 1 : 0);
    PSA_LAUNCH_CHECK("k_transport");
    return PSA_OK;
  }
}

template <int NT>
int launch_transport(const psa_group *g, int32_t w0, int nw, hipStream_t s, bool avg, bool planes) {
  if constexpr (NT > 0) {
    // row pass with tune[1] lines per block, column pass with tune[3]
    if (int rc = launch_transport_fpass<NT>(g, w0, nw, s, 0, avg, false)) return rc;
    return launch_transport_fpass<NT>(g, w0, nw, s, 1, avg, planes);
  } else {
    switch (g->tune[1]) {
      case 2: return launch_transport_l<NT, 2>(g, w0, nw, s, avg);
      case 8: return launch_transport_l<NT, 8>(g, w0, nw, s, avg);
      default: return launch_transport_l<NT, 4>(g, w0, nw, s, avg);
    }
  }
}

// planes: the leapfrog step in split-BF16 mode may leave the next step's
// p, C, D operand planes behind (compile-time N, 8 or 16 column lines per
// block); otherwise the next RHS splits them itself.
int transport_impl(const psa_group *g, int32_t w0, int32_t nw, double dt, hipStream_t s, bool avg,
                   bool planes = false) {
  planes = planes && g->prec == PSA_PREC_BF16X3 && g->spl && g->wtag;
  if (!(g->flags & PSA_F_LF)) {  // diagonal-only shift; the caller averages
    hipLaunchKernelGGL(k_tphase, dim3(nw), dim3(256), 0, s, *g, w0, dt);
    PSA_LAUNCH_CHECK("k_tphase");
    hipLaunchKernelGGL(k_transport_diag, dim3(nw, 3), dim3(256), 0, s, *g, w0);
    PSA_LAUNCH_CHECK("k_transport_diag");
    return PSA_OK;
  }
  if (g->N > 512) return fail(PSA_E_UNSUPPORTED, "Transport needs N <= 512 (LDS lines)");
  int n = g->N;
  while (n % 2 == 0) n /= 2;
  while (n % 3 == 0) n /= 3;
  if (n != 1) return fail(PSA_E_UNSUPPORTED, "Transport needs N = 2^a 3^b");
  hipLaunchKernelGGL(k_tphase, dim3(nw), dim3(256), 0, s, *g, w0, dt);
  PSA_LAUNCH_CHECK("k_tphase");
  switch (g->N) {
    case 64: return launch_transport<64>(g, w0, nw, s, avg, planes);
    case 96: return launch_transport<96>(g, w0, nw, s, avg, planes);
    case 128: return launch_transport<128>(g, w0, nw, s, avg, planes);
    case 192: return launch_transport<192>(g, w0, nw, s, avg, planes);
    case 256: return launch_transport<256>(g, w0, nw, s, avg, planes);
    case 384: return launch_transport<384>(g, w0, nw, s, avg, planes);
    case 512: return launch_transport<512>(g, w0, nw, s, avg, planes);
    default: return launch_transport<0>(g, w0, nw, s, avg, false);
  }
}

// Library-owned streams that the steps fork into: one set per (device,
// caller stream), created on first use with the default priority and kept
// until the process ends, with the events of the forks and joins.  A set
// serves one caller stream, so steps on different caller streams (or host
// threads) never share streams or events.
struct AuxStreams {
  hipStream_t st[PSA_RHS_MULTI_MAX];  // created on first use each
  hipEvent_t fork, mid, join[2][PSA_RHS_MULTI_MAX];
};
std::map<std::pair<int, hipStream_t>, AuxStreams *> g_aux;
std::mutex g_aux_mu;

// The set of caller stream s with streams 0 .. n-1 created; with create
// false (under graph capture, where no stream may be created) only a set
// whose streams exist already, else nullptr.
int aux_streams(hipStream_t s, int n, bool create, AuxStreams **out) {
  *out = nullptr;
  int dev = -1;
  if (hipGetDevice(&dev) != hipSuccess) return fail(PSA_E_HIP, "aux streams: no current device");
  std::lock_guard<std::mutex> lock(g_aux_mu);
  auto it = g_aux.find({dev, s});
  AuxStreams *a = it == g_aux.end() ? nullptr : it->second;
  if (!a) {
    if (!create) return PSA_OK;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    if (int rc = hip_check(hipEventCreateWithFlags(&e0, hipEventDisableTiming), "aux streams: hipEventCreate"))
      return rc;
    if (int rc = hip_check(hipEventCreateWithFlags(&e1, hipEventDisableTiming), "aux streams: hipEventCreate")) {
      (void)hipEventDestroy(e0);
      return rc;
    }
    a = new AuxStreams();
    a->fork = e0;
    a->mid = e1;
    g_aux[{dev, s}] = a;
  }
  // stream j is usable once it and both its join events exist; a failed
  // creation leaves it null, to be created again by a later call
  for (int j = 0; j < n; ++j) {
    if (a->st[j]) continue;
    if (!create) return PSA_OK;
    hipEvent_t ev[2] = {nullptr, nullptr};
    hipStream_t st = nullptr;
    int rc = hip_check(hipEventCreateWithFlags(&ev[0], hipEventDisableTiming), "aux streams: hipEventCreate");
    if (!rc) rc = hip_check(hipEventCreateWithFlags(&ev[1], hipEventDisableTiming), "aux streams: hipEventCreate");
    if (!rc) rc = hip_check(hipStreamCreateWithFlags(&st, hipStreamNonBlocking), "aux streams: hipStreamCreate");
    if (rc) {
      for (hipEvent_t e : ev)
        if (e) (void)hipEventDestroy(e);
      return rc;
    }
    a->join[0][j] = ev[0];
    a->join[1][j] = ev[1];
    a->st[j] = st;
  }
  *out = a;
  return PSA_OK;
}

// aux_streams for a step on s: creates outside graph capture, and under
// capture uses the streams an eager call created (or none)
int step_aux(hipStream_t s, int n, AuxStreams **out) {
  *out = nullptr;
  hipStreamCaptureStatus cap = hipStreamCaptureStatusNone;
  if (int rc = hip_check(hipStreamIsCapturing(s, &cap), "hipStreamIsCapturing")) return rc;
  return aux_streams(s, n, cap == hipStreamCaptureStatusNone, out);
}

// fork: streams 0 .. n-1 of a wait for what s has queued so far
int aux_fork(AuxStreams *a, hipEvent_t ev, hipStream_t s, int n) {
  if (int rc = hip_check(hipEventRecord(ev, s), "hipEventRecord")) return rc;
  for (int j = 0; j < n; ++j)
    if (int rc = hip_check(hipStreamWaitEvent(a->st[j], ev, 0), "hipStreamWaitEvent")) return rc;
  return PSA_OK;
}

// join: s waits for what streams 0 .. n-1 of a have queued so far
int aux_join(AuxStreams *a, int k, hipStream_t s, int n) {
  for (int j = 0; j < n; ++j) {
    if (int rc = hip_check(hipEventRecord(a->join[k][j], a->st[j]), "hipEventRecord")) return rc;
    if (int rc = hip_check(hipStreamWaitEvent(s, a->join[k][j], 0), "hipStreamWaitEvent")) return rc;
  }
  return PSA_OK;
}

// lazy: the leapfrog average is deferred (psa_group.lazy), nothing is averaged
int finish_impl(const psa_group *g, int32_t w0, int32_t nw, hipStream_t s, bool averaged, bool lazy = false) {
  hipLaunchKernelGGL(k_finish_diag, dim3(nw), dim3(256), 0, s, *g, w0, (averaged && !lazy) ? 1 : 0, lazy ? 1 : 0);
  PSA_LAUNCH_CHECK("k_finish_diag");
  if (averaged || lazy) return PSA_OK;
  const long total = (long)g->N * g->N * nw * 3;
  const int blocks = (int)((total + 255) / 256 < 8192 ? (total + 255) / 256 : 8192);
  hipLaunchKernelGGL(k_average, dim3(blocks), dim3(256), 0, s, *g, w0, nw);
  PSA_LAUNCH_CHECK("k_average");
  return PSA_OK;
}

}  // namespace

extern "C" {

int psa_transport(const psa_group *g, int32_t w0, int32_t nw, double dt, void *stream) {
  if (int rc = check_group(g, w0, nw)) return rc;
  if (!(g->flags & PSA_F_DCTRANS)) return PSA_OK;
  return transport_impl(g, w0, nw, dt, (hipStream_t)stream, false);
}

int psa_finish(const psa_group *g, int32_t w0, int32_t nw, void *stream) {
  if (int rc = check_group(g, w0, nw)) return rc;
  return finish_impl(g, w0, nw, (hipStream_t)stream, false);
}

int psa_average_levels(const psa_group *g, int32_t w0, int32_t nw, void *stream) {
  if (int rc = check_group(g, w0, nw)) return rc;
  if (!g->lazy) return PSA_OK;
  const hipStream_t s = (hipStream_t)stream;
  const long total = (long)g->N * g->N * nw * 3;
  const int blocks = (int)((total + 255) / 256 < 8192 ? (total + 255) / 256 : 8192);
  hipLaunchKernelGGL(k_lazy_average, dim3(blocks), dim3(256), 0, s, *g, w0, nw);
  PSA_LAUNCH_CHECK("k_lazy_average");
  hipLaunchKernelGGL(k_lazy_clear, dim3(1), dim3(256), 0, s, *g, w0, nw);
  PSA_LAUNCH_CHECK("k_lazy_clear");
  return PSA_OK;
}

int psa_sources(const psa_group *g, int32_t w0, int32_t nw, void *stream) {
  if (int rc = check_group(g, w0, nw)) return rc;
  const hipStream_t s = (hipStream_t)stream;
  hipLaunchKernelGGL(k_diagsums, dim3(nw, (2 * g->N - 1 + 63) / 64), dim3(64 * DS_WAVES), 0, s, *g, w0, 1);
  PSA_LAUNCH_CHECK("k_diagsums");
  {
    SrcProb p;
    p.g = *g;
    p.w0 = w0;
    p.nw = nw;
    if (int rc = launch_gemm_splitk(p, g->Nr, 4 * nw, 2 * g->N - 1, s, "sources gemm")) return rc;
  }
  hipLaunchKernelGGL(k_src_norm, dim3(nw), dim3(SN_THREADS), 0, s, *g, w0);
  PSA_LAUNCH_CHECK("k_src_norm");
  return PSA_OK;
}

int psa_prop2qw(const psa_group *g, int32_t w0, int32_t nw, void *stream) {
  if (int rc = check_group(g, w0, nw)) return rc;
  Prop2qwProb p;
  p.g = *g;
  p.w0 = w0;
  p.nw = nw;
  // the field transform always runs in FP64 (it sets the drive amplitude)
  return launch_gemm_splitk(p, 2 * g->N - 1, 3 * nw, g->gnx, (hipStream_t)stream, "prop2qw gemm");
}

int psa_qw2prop(const psa_group *g, int32_t w0, int32_t nw, void *stream) {
  if (int rc = check_group(g, w0, nw)) return rc;
  const hipStream_t s = (hipStream_t)stream;
  if (g->bnx > 0) {
    Qw2pProb p;
    p.g = *g;
    p.w0 = w0;
    p.nw = nw;
    if (int rc = launch_gemm_splitk(p, g->bnx, 5 * nw, g->Nr, s, "qw2prop gemm")) return rc;
  }
  hipLaunchKernelGGL(k_qw2prop_b, dim3(nw), dim3(256), 0, s, *g, w0);
  PSA_LAUNCH_CHECK("k_qw2prop_b");
  return PSA_OK;
}

int psa_drive(const psa_group *g, int32_t w0, int32_t nw, double *drive, double dt, void *stream) {
  if (int rc = check_group(g, w0, nw)) return rc;
  if (!drive) return fail(PSA_E_INVALID, "psa_drive: null drive table");
  const hipStream_t s = (hipStream_t)stream;
  const int bx = (g->Nx + 255) / 256 < 16 ? (g->Nx + 255) / 256 : 16;
  hipLaunchKernelGGL(k_drive, dim3(bx, nw), dim3(256), 0, s, *g, w0, drive, dt);
  PSA_LAUNCH_CHECK("k_drive");
  hipLaunchKernelGGL(k_drive_clock, dim3((nw + 63) / 64), dim3(64), 0, s, w0, nw, drive, dt);
  PSA_LAUNCH_CHECK("k_drive_clock");
  return PSA_OK;
}

int psa_record(const psa_group *g, int32_t w0, int32_t nw, int32_t xmid, double *trace,
               int32_t *count, int32_t maxn, void *stream) {
  if (int rc = check_group(g, w0, nw)) return rc;
  if (!trace || !count || maxn < 1 || xmid < 0 || xmid >= g->Nx)
    return fail(PSA_E_INVALID, "psa_record: bad arguments");
  hipLaunchKernelGGL(k_record, dim3((nw + 63) / 64), dim3(64), 0, (hipStream_t)stream, *g, w0, nw, xmid,
                     trace, count, maxn);
  PSA_LAUNCH_CHECK("k_record");
  return PSA_OK;
}

int psa_step(const psa_group *g, int32_t w0, int32_t nw, double dt, void *stream) {
  if (int rc = check_group(g, w0, nw)) return rc;
  int rc;
  if ((rc = psa_prop2qw(g, w0, nw, stream))) return rc;
  if ((rc = psa_prep(g, w0, nw, stream))) return rc;
  if ((rc = psa_calc_h(g, w0, nw, stream))) return rc;
  if (g->flags & PSA_F_OFFDIAGDPH)
    if ((rc = psa_offdiag_tables(g, w0, nw, stream))) return rc;
  if ((rc = psa_rhs(g, w0, nw, 1, dt, stream))) return rc;
  if ((rc = psa_relaxation(g, w0, nw, dt, stream))) return rc;
  // Transport's column pass also writes the leapfrog average (all but the
  // C/D diagonals, which finish averages after the renormalisation); the
  // diagonal-only transport of LF = 0 leaves the average to finish
  const bool trans = (g->flags & PSA_F_DCTRANS) != 0, fused = trans && (g->flags & PSA_F_LF);
  const bool lazy = g->lazy != nullptr;  // the average deferred to the next step's RHS
  if (trans && (rc = transport_impl(g, w0, nw, dt, (hipStream_t)stream, !lazy, true))) return rc;
  if ((rc = finish_impl(g, w0, nw, (hipStream_t)stream, fused, lazy))) return rc;
  if ((rc = psa_sources(g, w0, nw, stream))) return rc;
  if ((rc = psa_qw2prop(g, w0, nw, stream))) return rc;
  return PSA_OK;
}

}  // extern "C"

// ===========================================================================
// Allocation-free step.  psa_step's split-K products (Prop2QW, the sources,
// QW2Prop) keep their partial sums in a workspace, normally the caller's
// psa_group.kscr.  psa_step_noalloc puts them in `scratch`, a file-scope
// device array of PSA_NOALLOC_SCRATCH_DOUBLES doubles that every GPU the
// library runs on holds until the process ends.  Its device address is
// taken with hipGetSymbolAddress on the first call on each device (that
// device current) and kept per device; host code uses only that address.
// Nothing in it needs to start at zero: every split writes its whole
// partial block before the fixed-order reduce reads it.  Calls that overlap
// on one GPU share the array and race.
// ===========================================================================
namespace {

static __device__ __align__(256) double scratch[PSA_NOALLOC_SCRATCH_DOUBLES];
constexpr int kScratchDevices = 64;
double *g_scratch_ptr[kScratchDevices];
std::mutex g_scratch_mu;

int scratch_address(double **out) {
  int dev = -1;
  if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= kScratchDevices)
    return fail(PSA_E_HIP, "psa_step_noalloc: no current device below index 64");
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch_ptr[dev]) {
    void *p = nullptr;
    if (int rc = hip_check(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)), "psa_step_noalloc: hipGetSymbolAddress"))
      return rc;
    if (!p) return fail(PSA_E_HIP, "psa_step_noalloc: hipGetSymbolAddress gave a null address");
    g_scratch_ptr[dev] = static_cast<double *>(p);
  }
  *out = g_scratch_ptr[dev];
  return PSA_OK;
}

}  // namespace

extern "C" {

int64_t psa_step_noalloc_scratch(const psa_group *g, int32_t nw) {
  if (!g || nw < 1) return -1;
  const int64_t Nr = 2L * g->N + 2, Q = 2L * g->N - 1, bnx = std::max(g->bnx, 1);
  return 16L * std::max({Nr * 4 * nw, Q * 3 * nw, bnx * 5 * nw});
}

int psa_step_noalloc(const psa_group *g, int32_t w0, int32_t nw, double dt, void *stream) {
  if (int rc = check_group(g, w0, nw)) return rc;
  const int64_t need = psa_step_noalloc_scratch(g, nw);
  if (need > PSA_NOALLOC_SCRATCH_DOUBLES) {
    static thread_local char msg[200];
    std::snprintf(msg, sizeof(msg),
                  "psa_step_noalloc: %d wires of N=%d with bnx=%d need %lld doubles of split-K scratch, "
                  "above the library's maximum of %lld",
                  nw, g->N, g->bnx, (long long)need, (long long)PSA_NOALLOC_SCRATCH_DOUBLES);
    return fail(PSA_E_INVALID, msg);
  }
  double *scr = nullptr;
  if (int rc = scratch_address(&scr)) return rc;
  psa_group gg = *g;
  gg.kscr = scr;
  gg.kscr_len = PSA_NOALLOC_SCRATCH_DOUBLES;
  return psa_step(&gg, w0, nw, dt, stream);
}

int psa_step_multi(const void *packed_host, const void *packed_dev, uint32_t *queue, double dt, void *stream) {
  if (!packed_host || !packed_dev || !queue) return fail(PSA_E_INVALID, "psa_step_multi: null argument");
  MultiHdr h;
  std::memcpy(&h, packed_host, sizeof(h));
  if (h.n < 1 || h.n > PSA_RHS_MULTI_MAX || h.nwires < 1 || h.xstart[NXCD] != h.nwires)
    return fail(PSA_E_INVALID, "psa_step_multi: not a packed description");
  const RhsProb *probs_h =
      reinterpret_cast<const RhsProb *>(static_cast<const char *>(packed_host) + multi_hdr_bytes());
  std::vector<psa_group> gs(h.n);
  std::vector<int> w0(h.n);
  for (int j = 0; j < h.n; ++j) {
    RhsProb p;
    std::memcpy(&p, &probs_h[j], sizeof(p));
    if (p.mode != 1 || p.dt != dt)
      return fail(PSA_E_INVALID, "psa_step_multi: the description must be packed with mode 1 and this dt");
    gs[j] = p.g;
    w0[j] = p.w0;
  }
  const hipStream_t s = (hipStream_t)stream;
  // The groups' stages before and after the RHS run on a stream per group
  // (forked from and joined back into the caller's stream by events), so
  // the small groups' kernels fill the GPU together.  A single group runs
  // on the caller's stream; so do all groups under graph capture when no
  // eager call has created the streams yet (streams are not created while
  // a capture is open).
  AuxStreams *aux = nullptr;
  if (h.n > 1)
    if (int rc = step_aux(s, h.n, &aux)) return rc;
  auto on = [&](int j) { return aux ? aux->st[j] : s; };
  auto fork = [&](bool mid) { return aux ? aux_fork(aux, mid ? aux->mid : aux->fork, s, h.n) : PSA_OK; };
  auto join = [&](int k) { return aux ? aux_join(aux, k, s, h.n) : PSA_OK; };
  int rc;
  // the single-group step (psa_step) up to its RHS, for every group
  // a failure after a fork still joins the forked streams back (their
  // queued work then precedes anything the caller queues next)
  auto fail_joined = [&](int k, int err) {
    join(k);
    return err;
  };
  if ((rc = fork(false))) return rc;
  for (int j = 0; j < h.n; ++j) {
    const psa_group *g = &gs[j];
    void *sj = on(j);
    if ((rc = psa_prop2qw(g, w0[j], h.nw[j], sj))) return fail_joined(0, rc);
    if ((rc = psa_prep(g, w0[j], h.nw[j], sj))) return fail_joined(0, rc);
    if ((rc = psa_calc_h(g, w0[j], h.nw[j], sj))) return fail_joined(0, rc);
    if (g->flags & PSA_F_OFFDIAGDPH)
      if ((rc = psa_offdiag_tables(g, w0[j], h.nw[j], sj))) return fail_joined(0, rc);
    if (g->prec == PSA_PREC_BF16X3 && (rc = split_planes(*g, w0[j], h.nw[j], on(j)))) return fail_joined(0, rc);
  }
  if ((rc = join(0))) return rc;
  if ((rc = rhs_multi_impl(packed_host, packed_dev, queue, s, false))) return rc;
  if ((rc = fork(true))) return rc;
  for (int j = 0; j < h.n; ++j) {
    const psa_group *g = &gs[j];
    void *sj = on(j);
    if ((rc = psa_relaxation(g, w0[j], h.nw[j], dt, sj))) return fail_joined(1, rc);
    const bool trans = (g->flags & PSA_F_DCTRANS) != 0, fused = trans && (g->flags & PSA_F_LF);
    const bool lazy = g->lazy != nullptr;
    if (trans && (rc = transport_impl(g, w0[j], h.nw[j], dt, (hipStream_t)sj, !lazy, true)))
      return fail_joined(1, rc);
    if ((rc = finish_impl(g, w0[j], h.nw[j], (hipStream_t)sj, fused, lazy))) return fail_joined(1, rc);
    if ((rc = psa_sources(g, w0[j], h.nw[j], sj))) return fail_joined(1, rc);
    if ((rc = psa_qw2prop(g, w0[j], h.nw[j], sj))) return fail_joined(1, rc);
  }
  return join(1);
}

int psa_step_rk4(const psa_group *g, int32_t w0, int32_t nw, double dt, void *stream) {
  if (int rc = check_group(g, w0, nw)) return rc;
  if (!g->pR || !g->CR || !g->DR) return fail(PSA_E_INVALID, "psa_step_rk4: needs pR, CR, DR");
  const hipStream_t s = (hipStream_t)stream;
  int rc;
  if ((rc = psa_prop2qw(g, w0, nw, stream))) return rc;
  // stage buffers: the state S = B stays the base; stage inputs alternate
  // between X and A; the last stage writes the new state over S in place
  double *const S[3] = {g->pB, g->CB, g->DB}, *const X[3] = {g->pX, g->CX, g->DX},
                *const A[3] = {g->pA, g->CA, g->DA};
  double *const *in[4] = {S, X, A, X}, *const *out[4] = {X, A, X, S};
  psa_group q = *g;
  q.lazy = nullptr;  // no leapfrog levels; k_finish_diag below clears the flags
  for (int st = 0; st < 4; ++st) {
    q.pA = S[0], q.CA = S[1], q.DA = S[2];
    q.pB = in[st][0], q.CB = in[st][1], q.DB = in[st][2];
    q.pX = out[st][0], q.CX = out[st][1], q.DX = out[st][2];
    if ((rc = psa_prep(&q, w0, nw, stream))) return rc;
    if ((rc = psa_calc_h(&q, w0, nw, stream))) return rc;
    if (q.flags & PSA_F_OFFDIAGDPH)
      if ((rc = psa_offdiag_tables(&q, w0, nw, stream))) return rc;
    if ((rc = psa_rhs(&q, w0, nw, 2 + st, dt, stream))) return rc;
  }
  // q: A = X = the new state.  No leapfrog average: the renormalised
  // diagonal is the state, and CalcI0 reads it through A.
  if ((rc = psa_relaxation(&q, w0, nw, dt, stream))) return rc;
  if ((q.flags & PSA_F_DCTRANS) && (rc = transport_impl(&q, w0, nw, dt, s, false))) return rc;
  {
    psa_group qf = q;
    qf.lazy = g->lazy;
    hipLaunchKernelGGL(k_finish_diag, dim3(nw), dim3(256), 0, s, qf, w0, 0, 0);
    PSA_LAUNCH_CHECK("k_finish_diag");
  }
  if ((rc = psa_sources(&q, w0, nw, stream))) return rc;
  if ((rc = psa_qw2prop(&q, w0, nw, stream))) return rc;
  return PSA_OK;
}

}  // extern "C"

// ===========================================================================
// Largest and smallest entry of the off-diagonal dephasing tables De, Dh of
// each wire (the MaxOffDiag.e/h.dat lines of OffDiagDephasing2,
// dephasing.py:733-743) into obs[w][4..7].  Unused (k, q) entries are 0, as
// in the reference's tables.
// ===========================================================================
namespace {

__global__ __launch_bounds__(256) void k_dedh_extrema(psa_group g, int w0) {
  __shared__ double sh[8];
  const int w = w0 + blockIdx.x, N = g.N;
  const long per = (long)N * (2 * N - 1);
  for (int kind = 0; kind < 2; ++kind) {
    const double *t = g.dedh + ((long)w * 2 + kind) * per;
    double mx = -INFINITY, mn = INFINITY;
    for (long i = threadIdx.x; i < per; i += 256) {
      mx = fmax(mx, t[i]);
      mn = fmin(mn, t[i]);
    }
    mx = block_max(mx, sh);
    mn = -block_max(-mn, sh);
    if (threadIdx.x == 0) {
      g.obs[(long)w * PSA_OBS_N + 4 + 2 * kind] = mx;
      g.obs[(long)w * PSA_OBS_N + 5 + 2 * kind] = mn;
    }
  }
}

}  // namespace

extern "C" int psa_offdiag_extrema(const psa_group *g, int32_t w0, int32_t nw, void *stream) {
  if (int rc = check_group(g, w0, nw)) return rc;
  if (!(g->flags & PSA_F_OFFDIAGDPH)) return fail(PSA_E_INVALID, "psa_offdiag_extrema: OffDiagDph is off");
  hipLaunchKernelGGL(k_dedh_extrema, dim3(nw), dim3(256), 0, (hipStream_t)stream, *g, w0);
  PSA_LAUNCH_CHECK("k_dedh_extrema");
  return PSA_OK;
}

// ===========================================================================
// chi(q, w) table recorded at start-up (_record_Xqw, SBEs.py:5395-5432) by
// _CalcXqw_jit (SBEs.py:3165-3243): one thread per (q, w), q = blockIdx.y in
// [0, N], w = iw dw.  The two sums run in the reference's order; each term
// is a real numerator over a complex denominator, divided as CPython
// divides complex numbers (Smith's scaling).
// ===========================================================================
namespace {

__device__ __forceinline__ cd real_over(double num, double dr, double di) {
  if (fabs(dr) >= fabs(di)) {
    const double r = di / dr, den = dr + di * r;
    return cd{num / den, (-num * r) / den};
  }
  const double r = dr / di, den = dr * r + di;
  return cd{(num * r) / den, -num / den};
}

__global__ __launch_bounds__(256) void k_xqw(int N, const double *Ee, const double *Eh, const double *fe,
                                             const double *fh, const double *game, const double *gamh,
                                             double gap, double x, double hbar, double hfloor, double dw,
                                             int nw, double *out) {
  __shared__ double sEc[1024], sEv[1024], sfc[1024], sfv[1024], shge[1024], shgh[1024];
  for (int k = threadIdx.x; k < N; k += 256) {
    sEc[k] = gap + Ee[k];
    sEv[k] = -Eh[k];
    sfc[k] = fe[k];
    sfv[k] = 1.0 - fh[k];
    shge[k] = fmax(hfloor, hbar * game[k]);
    shgh[k] = fmax(hfloor, hbar * gamh[k]);
  }
  __syncthreads();
  const int iq = blockIdx.y, iw = blockIdx.x * 256 + threadIdx.x;
  if (iw >= nw) return;
  const double hw = hbar * ((double)iw * dw);
  double re = 0.0, im = 0.0;
  for (int k = 0; k < N - iq; ++k) {
    const cd t = real_over(x * (sfc[k] - sfv[k + iq]), sEv[k + iq] - sEc[k] - hw, -(shgh[k + iq] + shge[k]));
    re += t.re;
    im += t.im;
  }
  for (int k = 0; k < N - iq; ++k) {
    const cd t = real_over(x * (sfv[k] - sfc[k + iq]), sEc[k + iq] - sEv[k] - hw, -(shge[k + iq] + shgh[k]));
    re += t.re;
    im += t.im;
  }
  st(out, (long)iq * nw + iw, cd{re, im});
}

}  // namespace

extern "C" int psa_xqw_table(int32_t N, const double *Ee, const double *Eh, const double *fe, const double *fh,
                             const double *game, const double *gamh, double gap, double x, double hbar,
                             double hfloor, double dw, int32_t nw, double *out, void *stream) {
  if (N < 2 || N > 1023 || nw < 1 || !Ee || !Eh || !fe || !fh || !game || !gamh || !out)
    return fail(PSA_E_INVALID, "psa_xqw_table: bad arguments");
  hipLaunchKernelGGL(k_xqw, dim3((nw + 255) / 256, N + 1), dim3(256), 0, (hipStream_t)stream, N, Ee, Eh, fe, fh,
                     game, gamh, gap, x, hbar, hfloor, dw, nw, out);
  PSA_LAUNCH_CHECK("k_xqw");
  return PSA_OK;
}

// ===========================================================================
// 3-D coupling (rhoPJ.QuantumWire): drive, current/charge/polarization
// lines, placement, and ElongfromRho's 3-D FFTs.  The 3-D FFT is three
// passes of line FFTs (fft_line, one wave per line, N = 2^a 3^b <= 512),
// eight lines per block; for the strided axes the eight lines of a block are
// neighbours in memory, so a row of eight elements is one 128-byte read.
// ===========================================================================
namespace {

constexpr int Q3_LINES = 8;

// lines of `narr` grids [n0][n1][n2] along `axis`; sign -1 forward, +1
// inverse (scaled by `scale` on the way out)
// LDS (dynamic, sized for the axis length): the block's lines of fpad(N)
// slots, then N twiddles.  Each wave transforms q3_lpw(N) lines together
// (short lines share a wave), so a block holds Q3_LINES q3_lpw(N) lines.
__host__ __device__ inline int q3_line_slots(int N) { return N + N / 4; }
__host__ __device__ inline int q3_lpw(int N) { return N >= 256 ? 1 : 256 / N; }
__global__ __launch_bounds__(64 * Q3_LINES) void k_q3_fft_axis(double *data, int n0, int n1, int n2, int axis,
                                                              int sign, const double *twid, double scale) {
  extern __shared__ cd q3_lds[];
  const int N = axis == 0 ? n0 : (axis == 1 ? n1 : n2);
  const int LS = q3_line_slots(N), LPW = q3_lpw(N), BL = Q3_LINES * LPW;
  cd *const tw = q3_lds + BL * LS;
  const long grid = (long)n0 * n1 * n2;
  const long es = axis == 0 ? (long)n1 * n2 : (axis == 1 ? n2 : 1);
  // lines: axis 2 -> (i0, i1); axis 1 -> (i0, i2); axis 0 -> (i1, i2)
  const long ninner = axis == 2 ? 1 : (axis == 1 ? n2 : (long)n1 * n2);
  const long hstride = axis == 2 ? n2 : (axis == 1 ? (long)n1 * n2 : 0);
  const long nlines = grid / N;
  const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
  for (int i = tid; i < N; i += 64 * Q3_LINES) tw[i] = ld(twid, i);
  const long L0 = (long)blockIdx.x * BL;  // first line of the block (within one grid)
  double *base_arr = data + 2 * grid * blockIdx.y;
  auto line_base = [&](long L) { return (L / ninner) * hstride + (L % ninner) * (axis == 2 ? n2 : 1); };
  // element e of block line l: thread t = e BL + l for the strided axes (the
  // block's lines are neighbours in memory), t = l N + e for the contiguous axis
  if (axis == 2) {
    for (int t = tid; t < N * BL; t += 64 * Q3_LINES) {
      const int l = t / N, e = t - l * N;
      const long L = L0 + l;
      if (L < nlines) q3_lds[l * LS + fpad(e)] = ld(base_arr, line_base(L) + e);
    }
  } else {
    for (int t = tid; t < N * BL; t += 64 * Q3_LINES) {
      const int l = t % BL, e = t / BL;
      const long L = L0 + l;
      if (L < nlines) q3_lds[l * LS + fpad(e)] = ld(base_arr, line_base(L) + e * es);
    }
  }
  __syncthreads();
  {
    const long Lw = L0 + (long)wv * LPW;
    const int nl = (int)(nlines - Lw < LPW ? nlines - Lw : LPW);
    if (nl > 0) fft_lines(q3_lds + wv * LPW * LS, LS, nl, tw, N, sign, lane);
  }
  __syncthreads();
  if (axis == 2) {
    for (int t = tid; t < N * BL; t += 64 * Q3_LINES) {
      const int l = t / N, e = t - l * N;
      const long L = L0 + l;
      if (L < nlines) st(base_arr, line_base(L) + e, cscale(q3_lds[l * LS + fpad(e)], scale));
    }
  } else {
    for (int t = tid; t < N * BL; t += 64 * Q3_LINES) {
      const int l = t % BL, e = t / BL;
      const long L = L0 + l;
      if (L < nlines) st(base_arr, line_base(L) + e * es, cscale(q3_lds[l * LS + fpad(e)], scale));
    }
  }
}

int q3_fft(const psa_qw3d &q, double *data, int narr, int sign, hipStream_t s) {
  const long grid = (long)q.nx * q.ny * q.nz;
  const int ns[3] = {q.nx, q.ny, q.nz};
  const double *tws[3] = {q.tw, q.tw + 2L * q.nx, q.tw + 2L * (q.nx + q.ny)};
  // inverse: z, y, x with the 1/(nx ny nz) of ifftn on the last pass
  for (int a = 0; a < 3; ++a) {
    const int axis = sign < 0 ? a : 2 - a;
    const double scale = (sign > 0 && a == 2) ? 1.0 / (double)grid : 1.0;
    const long nlines = grid / ns[axis];
    const int bl = Q3_LINES * q3_lpw(ns[axis]);
    const size_t shm = sizeof(cd) * (size_t)(bl * q3_line_slots(ns[axis]) + ns[axis]);
    static bool attr_set = false;
    if (!attr_set) {  // up to 88 KB for 512-point lines
      if (int rc = hip_check(hipFuncSetAttribute((const void *)k_q3_fft_axis,
                                                 hipFuncAttributeMaxDynamicSharedMemorySize,
                                                 (int)(sizeof(cd) * (Q3_LINES * q3_line_slots(512) + 512))),
                             "hipFuncSetAttribute(k_q3_fft_axis)"))
        return rc;
      attr_set = true;
    }
    hipLaunchKernelGGL(k_q3_fft_axis, dim3((unsigned)((nlines + bl - 1) / bl), narr),
                       dim3(64 * Q3_LINES), shm, s, data, q.nx, q.ny, q.nz, axis, sign, tws[axis], scale);
    PSA_LAUNCH_CHECK("k_q3_fft_axis");
  }
  return PSA_OK;
}

// EAtXYZ of Ex, Ey, Ez (usefulsubs.py:2057-2065 blend order) broadcast into
// efield[c][w][x] of every wire
__global__ __launch_bounds__(256) void k_q3_drive(psa_qw3d q, psa_group g) {
  const int x = blockIdx.x * 256 + threadIdx.x, w = blockIdx.y;
  if (x >= g.Nx) return;
  const double wx[2] = {1.0 - q.ax, q.ax}, wy[2] = {1.0 - q.ay, q.ay}, wz[2] = {1.0 - q.az, q.az};
  const double *F[3] = {q.Ex, q.Ey, q.Ez};
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    cd f0{0.0, 0.0};
    for (int r = 0; r < 2; ++r)
      for (int qq = 0; qq < 2; ++qq)
        for (int p = 0; p < 2; ++p) {
          const int i = min(q.i0 + p, q.nx - 1), j = min(q.j0 + qq, q.ny - 1), k = min(q.k0 + r, q.nz - 1);
          const double wgt = wx[p] * wy[qq] * wz[r];
          const cd v = ld(F[c], ((long)i * q.ny + j) * q.nz + k);
          f0 = cd{f0.re + wgt * v.re, f0.im + wgt * v.im};
        }
    st(g.efield, ((long)c * g.Nw + w) * g.Nx + x, f0);
  }
}

// per x: lines[0..2] = sum_w (P - P_old) / dt, [3] = sum_w Rho, [4..6] =
// sum_w P, [7] = (sum_w Rho - sum_w Rho_old) / dt.  The reference discards
// its previous P and Rho at the top of every call (rhoPJ.py:256-270), so
// P_old = Rho_old = 0 there; with PSA_QW3D_KEEP_OLD they are the values
// this kernel stored on the previous call (P_old <- P, Rho_old <- Rho).
__global__ __launch_bounds__(256) void k_q3_lines(psa_qw3d q, psa_group g) {
  const int i = blockIdx.x * 256 + threadIdx.x, nx = q.nx, W = g.Nw;
  if (i >= nx) return;
  const bool keep = (q.flags & PSA_QW3D_KEEP_OLD) != 0;
  cd d[3] = {}, ps[3] = {}, rs{0.0, 0.0}, ro{0.0, 0.0};
  for (int w = 0; w < W; ++w) {
#pragma unroll
    for (int c = 0; c < 4; ++c) {
      const long at = ((long)c * W + w) * nx + i;
      const cd v = ld(g.out, at), o = keep ? ld(q.old, at) : cd{0.0, 0.0};
      if (c < 3) {
        d[c] = cadd(d[c], csub(v, o));
        ps[c] = cadd(ps[c], v);
      } else {
        rs = cadd(rs, v);
        ro = cadd(ro, o);
      }
      if (keep) st(q.old, at, v);
    }
  }
  const double idt = 1.0 / q.dt;
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    st(q.lines, (long)c * nx + i, cscale(d[c], idt));
    st(q.lines, (long)(4 + c) * nx + i, ps[c]);
  }
  st(q.lines, 3L * nx + i, rs);
  st(q.lines, 7L * nx + i, cscale(csub(rs, ro), idt));
}

// CalcJfx (sequential, the reference's order) and the 1/Nw factors
__global__ __launch_bounds__(256) void k_q3_jfx(psa_qw3d q, int W) {
  __shared__ cd jf[512], dr[512];
  const int nx = q.nx;
  for (int i = threadIdx.x; i < nx; i += 256) dr[i] = ld(q.lines, 7L * nx + i);
  __syncthreads();
  if (threadIdx.x == 0) {
    cd a{0.0, 0.0};
    jf[0] = a;
    for (int i = 1; i < nx; ++i) {
      a = csub(a, cscale(dr[i], q.dx));
      jf[i] = a;
    }
  }
  __syncthreads();
  const double iw = 1.0 / (double)W;
  for (int i = threadIdx.x; i < nx; i += 256) {
    const cd jx = ld(q.lines, i);
    st(q.lines, i, cadd(cscale(jx, iw), cscale(jf[i], iw)));
    for (int c = 1; c < 7; ++c) st(q.lines, (long)c * nx + i, cscale(ld(q.lines, (long)c * nx + i), iw));
  }
}

// QWPlacement (rhoPJ.py:655-700): the grid is zeroed, then
// F[i][j][k] = 0 + (gx gy) gz f[i] with gx = 1, so a placed grid is
// overwritten: J into Jx, Jy, Jz (CURR), the charge into Rho (RHO), the
// polarization into Fk[1..3] (POL, with Fk[0] = Rho as ElongfromRho reads
// it).  Grids whose flag is off keep the caller's values.
__global__ __launch_bounds__(256) void k_q3_place(psa_qw3d q) {
  const long n = (long)q.nx * q.ny * q.nz, plane = (long)q.ny * q.nz;
  for (long t = blockIdx.x * 256L + threadIdx.x; t < n; t += (long)gridDim.x * 256) {
    const int i = (int)(t / plane), j = (int)((t / q.nz) % q.ny), k = (int)(t % q.nz);
    const double gg = (1.0 * q.gy[j]) * q.gz[k];
    if (q.flags & PSA_QW3D_CURR) {
      double *J[3] = {q.Jx, q.Jy, q.Jz};
#pragma unroll
      for (int c = 0; c < 3; ++c) st(J[c], t, cscale(ld(q.lines, (long)c * q.nx + i), gg));
    }
    cd rho;
    if (q.flags & PSA_QW3D_RHO) {
      rho = cscale(ld(q.lines, 3L * q.nx + i), gg);
      st(q.Rho, t, rho);
    } else {
      rho = ld(q.Rho, t);
    }
    if (q.flags & PSA_QW3D_POL) {
      st(q.Fk, t, rho);
#pragma unroll
      for (int c = 0; c < 3; ++c) st(q.Fk, (1 + c) * n + t, cscale(ld(q.lines, (long)(4 + c) * q.nx + i), gg));
    }
  }
}

// the k-space loop of ElongfromRho (rhoPJ.py:526-619)
__global__ __launch_bounds__(256) void k_q3_elong(psa_qw3d q) {
  const long n = (long)q.nx * q.ny * q.nz, plane = (long)q.ny * q.nz;
  const double small = (q.qx[1] - q.qx[0]) / 1e4;
  const double ie = 1.0 / q.eps;
  for (long t = blockIdx.x * 256L + threadIdx.x; t < n; t += (long)gridDim.x * 256) {
    const int i = (int)(t / plane), j = (int)((t / q.nz) % q.ny), k = (int)(t % q.nz);
    const double qx = q.qx[i], qy = q.qy[j], qz = q.qz[k];
    const double qsq = qx * qx + qy * qy + qz * qz + small;
    const cd rk = ld(q.Fk, t), px = ld(q.Fk, n + t), py = ld(q.Fk, 2 * n + t), pz = ld(q.Fk, 3 * n + t);
    const cd qdp = cadd(cadd(cscale(px, qx), cscale(py, qy)), cscale(pz, qz));
    const cd rb = {qdp.im, -qdp.re};                              // -i q.P
    const cd rt0 = {rk.im, -rk.re};                               // -i Rho
    const cd rt = {rt0.re * ie / qsq, rt0.im * ie / qsq};
    const cd pt = {-qdp.re * ie / qsq, -qdp.im * ie / qsq};
    const cd sp = cadd(rt, pt);
    const double qs[3] = {qx, qy, qz};
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      st(q.El, (long)c * n + t, cscale(sp, qs[c]));
      st(q.El, (long)(3 + c) * n + t, cscale(pt, qs[c]));
      st(q.El, (long)(6 + c) * n + t, cscale(rt, qs[c]));
    }
    st(q.El, 9L * n + t, rb);
  }
}

bool q3_len_ok(int n) {
  if (n < 2 || n > 512) return false;
  while (n % 2 == 0) n /= 2;
  while (n % 3 == 0) n /= 3;
  return n == 1;
}

int check_qw3d(const psa_qw3d *q, const psa_group *g, bool deposit) {
  if (!q || !g) return fail(PSA_E_INVALID, "psa_qw3d: null argument");
  if (int rc = check_group(g, 0, g->Nw)) return rc;
  if (!q3_len_ok(q->nx) || !q3_len_ok(q->ny) || !q3_len_ok(q->nz))
    return fail(PSA_E_INVALID, "psa_qw3d: nx, ny, nz must be 2^a 3^b in [2, 512]");
  if (g->Nx != q->nx) return fail(PSA_E_INVALID, "psa_qw3d: the group's Nx must equal nx");
  if (q->i0 < 0 || q->i0 >= q->nx || q->j0 < 0 || q->j0 >= q->ny || q->k0 < 0 || q->k0 >= q->nz)
    return fail(PSA_E_INVALID, "psa_qw3d: EAtXYZ cell outside the grid");
  if (!deposit && (!q->Ex || !q->Ey || !q->Ez)) return fail(PSA_E_INVALID, "psa_qw3d_drive: missing field");
  if (deposit && (!q->qx || !q->qy || !q->qz || !q->gy || !q->gz || !q->tw || !q->lines || !q->old ||
                  !q->Jx || !q->Jy || !q->Jz || !q->Rho || !q->Fk || !q->El))
    return fail(PSA_E_INVALID, "psa_qw3d_deposit: missing buffer");
  if (!(q->dt > 0.0) || !(q->eps > 0.0)) return fail(PSA_E_INVALID, "psa_qw3d: dt and eps must be positive");
  return PSA_OK;
}

}  // namespace

extern "C" {

int psa_qw3d_size(void) { return (int)sizeof(psa_qw3d); }

int psa_qw3d_drive(const psa_qw3d *q, const psa_group *g, void *stream) {
  if (int rc = check_qw3d(q, g, false)) return rc;
  hipLaunchKernelGGL(k_q3_drive, dim3((g->Nx + 255) / 256, g->Nw), dim3(256), 0, (hipStream_t)stream, *q, *g);
  PSA_LAUNCH_CHECK("k_q3_drive");
  return PSA_OK;
}

int psa_qw3d_deposit(const psa_qw3d *q, const psa_group *g, void *stream) {
  if (int rc = check_qw3d(q, g, true)) return rc;
  const hipStream_t s = (hipStream_t)stream;
  hipLaunchKernelGGL(k_q3_lines, dim3((q->nx + 255) / 256), dim3(256), 0, s, *q, *g);
  PSA_LAUNCH_CHECK("k_q3_lines");
  hipLaunchKernelGGL(k_q3_jfx, dim3(1), dim3(256), 0, s, *q, g->Nw);
  PSA_LAUNCH_CHECK("k_q3_jfx");
  const long n = (long)q->nx * q->ny * q->nz;
  const int blocks = (int)std::min<long>((n + 255) / 256, 4096);
  hipLaunchKernelGGL(k_q3_place, dim3(blocks), dim3(256), 0, s, *q);
  PSA_LAUNCH_CHECK("k_q3_place");
  if (!(q->flags & PSA_QW3D_POL)) return PSA_OK;
  if (int rc = q3_fft(*q, q->Fk, 4, -1, s)) return rc;
  hipLaunchKernelGGL(k_q3_elong, dim3(blocks), dim3(256), 0, s, *q);
  PSA_LAUNCH_CHECK("k_q3_elong");
  return q3_fft(*q, q->El, 10, +1, s);
}

}  // extern "C"

// ===========================================================================
// 1-D pseudo-spectral propagator (psa_pstd.h)
// ===========================================================================
namespace {

int check_pstd(const psa_pstd *p) {
  if (!p) return fail(PSA_E_INVALID, "null psa_pstd");
  if (p->N < 64 || p->N > 8192 || (p->N & (p->N - 1)))
    return fail(PSA_E_INVALID, "psa_pstd: N must be a power of two in [64, 8192]");
  if (p->nw < 0 || p->nw > PSA_PSTD_MAXW) return fail(PSA_E_INVALID, "psa_pstd: nw outside [0, 256]");
  if (p->nobs < 0 || p->nrec < 0) return fail(PSA_E_INVALID, "psa_pstd: negative nobs / nrec");
  if (!p->state || !p->mode || !p->twid || !p->count)
    return fail(PSA_E_INVALID, "psa_pstd: state, mode, twid and count are required");
  if (p->nw > 0 && (!p->jw || !p->pold || !p->jsrc || !p->ew))
    return fail(PSA_E_INVALID, "psa_pstd: wire slots need jw, pold, jsrc and ew");
  if (p->nobs > 0 && p->nrec > 0 && (!p->obs || !p->rec))
    return fail(PSA_E_INVALID, "psa_pstd: observation points need obs and rec");
  if ((p->flags & PSA_PSTD_ABC) && !p->abc) return fail(PSA_E_INVALID, "psa_pstd: PSA_PSTD_ABC needs abc");
  if (!(p->dy > 0.0) || !(p->eps > 0.0) || !(p->mu > 0.0) || !(p->dt > 0.0))
    return fail(PSA_E_INVALID, "psa_pstd: dy, eps, mu and dt must be positive");
  return PSA_OK;
}

template <int LOGN>
int launch_pstd_line_n(const psa_pstd &p, int mode, double *out, int blocks, hipStream_t s) {
  hipLaunchKernelGGL(pstd::k_pstd_line<LOGN>, dim3(blocks), dim3(pstd::Fft<LOGN>::T), 0, s, p, mode, out);
  PSA_LAUNCH_CHECK("k_pstd_line");
  return PSA_OK;
}

int launch_pstd_line(const psa_pstd &p, int mode, double *out, hipStream_t s) {
  const bool e_only = mode == pstd::SAMPLE || (mode == pstd::STEP && !(p.flags & PSA_PSTD_ABC));
  const int blocks = e_only ? 2 : 4;
  switch (p.N) {
    case 64: return launch_pstd_line_n<6>(p, mode, out, blocks, s);
    case 128: return launch_pstd_line_n<7>(p, mode, out, blocks, s);
    case 256: return launch_pstd_line_n<8>(p, mode, out, blocks, s);
    case 512: return launch_pstd_line_n<9>(p, mode, out, blocks, s);
    case 1024: return launch_pstd_line_n<10>(p, mode, out, blocks, s);
    case 2048: return launch_pstd_line_n<11>(p, mode, out, blocks, s);
    case 4096: return launch_pstd_line_n<12>(p, mode, out, blocks, s);
    case 8192: return launch_pstd_line_n<13>(p, mode, out, blocks, s);
  }
  return fail(PSA_E_INVALID, "psa_pstd: unsupported N");
}

int check_pstd_group(const psa_pstd *p, const psa_group *g, int w0, int nw, int slot0) {
  if (int rc = check_pstd(p)) return rc;
  if (int rc = check_group(g, w0, nw)) return rc;
  if (slot0 < 0 || slot0 + nw > p->nw) return fail(PSA_E_INVALID, "psa_pstd: wire slots outside [0, nw)");
  if (g->Nx < 1 || !g->out || !g->efield) return fail(PSA_E_INVALID, "psa_pstd: group has no rr grid");
  return PSA_OK;
}

}  // namespace

extern "C" {

int psa_pstd_size(void) { return (int)sizeof(psa_pstd); }

int psa_pstd_to_k(const psa_pstd *p, void *stream) {
  if (int rc = check_pstd(p)) return rc;
  return launch_pstd_line(*p, pstd::TO_K, nullptr, (hipStream_t)stream);
}

int psa_pstd_to_real(const psa_pstd *p, double *out, void *stream) {
  if (int rc = check_pstd(p)) return rc;
  if (!out) return fail(PSA_E_INVALID, "psa_pstd_to_real: null out");
  return launch_pstd_line(*p, pstd::TO_REAL, out, (hipStream_t)stream);
}

int psa_pstd_sample(const psa_pstd *p, void *stream) {
  if (int rc = check_pstd(p)) return rc;
  if (p->nw == 0) return PSA_OK;
  return launch_pstd_line(*p, pstd::SAMPLE, nullptr, (hipStream_t)stream);
}

int psa_pstd_advance(const psa_pstd *p, void *stream) {
  if (int rc = check_pstd(p)) return rc;
  const hipStream_t s = (hipStream_t)stream;
  hipLaunchKernelGGL(pstd::k_pstd_update, dim3((p->N + pstd::UPD_T - 1) / pstd::UPD_T), dim3(pstd::UPD_T), 0,
                     s, *p);
  PSA_LAUNCH_CHECK("k_pstd_update");
  if (!(p->flags & PSA_PSTD_ABC) && p->nw == 0 && (p->nobs == 0 || p->nrec == 0)) return PSA_OK;
  return launch_pstd_line(*p, pstd::STEP, nullptr, s);
}

int psa_pstd_gather(const psa_pstd *p, const psa_group *g, int32_t w0, int32_t nw, int32_t slot0,
                    void *stream) {
  if (int rc = check_pstd_group(p, g, w0, nw, slot0)) return rc;
  hipLaunchKernelGGL(pstd::k_pstd_gather, dim3(nw, 2), dim3(256), 0, (hipStream_t)stream, *p, *g, w0, slot0);
  PSA_LAUNCH_CHECK("k_pstd_gather");
  return PSA_OK;
}

int psa_pstd_scatter(const psa_pstd *p, const psa_group *g, int32_t w0, int32_t nw, int32_t slot0,
                     void *stream) {
  if (int rc = check_pstd_group(p, g, w0, nw, slot0)) return rc;
  hipLaunchKernelGGL(pstd::k_pstd_scatter, dim3((g->Nx + 255) / 256, nw, 2), dim3(256), 0, (hipStream_t)stream,
                     *p, *g, w0, slot0);
  PSA_LAUNCH_CHECK("k_pstd_scatter");
  return PSA_OK;
}

}  // extern "C"

// ---------------------------------------------------------------------------
// Module API on the caller's dense arrays (pulsesuite_amd.coulomb / .phonons)
// ---------------------------------------------------------------------------
extern "C" {

int psa_mod_mb_arrays(int32_t N, const double *Ee, const double *Eh, double ge, double gh, double hbar,
                      int32_t lorentz, int32_t *k3, double *Ceh, double *Cee, double *Chh, void *stream) {
  if (N < 1 || N > 1024 || !Ee || !Eh || !k3 || !Ceh || !Cee || !Chh)
    return fail(PSA_E_INVALID, "psa_mod_mb_arrays: bad arguments");
  hipLaunchKernelGGL(mod::k_mb_arrays, dim3(N, N), dim3(N >= 256 ? 256 : 64), 0, (hipStream_t)stream, N, Ee, Eh,
                     ge, gh, hbar, (int)lorentz, k3, Ceh, Cee, Chh);
  PSA_LAUNCH_CHECK("k_mb_arrays");
  return PSA_OK;
}

int psa_mod_screen(int32_t N, int32_t screened, double L, double nmax, const double *ne, const double *nh,
                   double n1d, const double *V0, const double *chi, const double *q, double *VC, double *E1D,
                   void *stream) {
  if (N < 1 || N > 1024 || !E1D || (!ne) != (!nh) || (VC && !V0) || (screened && (!chi || !q)) ||
      (screened && ne && !(L > 0.0)) || (!ne && !screened))
    return fail(PSA_E_INVALID, "psa_mod_screen: bad arguments");
  const long NN = (long)N * N;
  hipLaunchKernelGGL(mod::k_screen, dim3((unsigned)((NN + mod::RED_T - 1) / mod::RED_T)), dim3(mod::RED_T), 0,
                     (hipStream_t)stream, N, (int)(screened != 0), L, nmax, reinterpret_cast<const cd *>(ne),
                     reinterpret_cast<const cd *>(nh), n1d, V0, chi, q, VC, E1D);
  PSA_LAUNCH_CHECK("k_screen");
  return PSA_OK;
}

int psa_mod_mbc(int32_t N, int32_t hole, const double *ne0, const double *nh0, const double *VC, const int32_t *k3,
                const double *Ceh, const double *Cx, double *Win, double *Wout, void *stream) {
  if (N < 1 || N > 1024 || (hole != 0 && hole != 1) || !ne0 || !nh0 || !VC || !k3 || !Ceh || !Cx || !Win || !Wout)
    return fail(PSA_E_INVALID, "psa_mod_mbc: bad arguments");
  hipLaunchKernelGGL(mod::k_mbc, dim3(N), dim3(mod::RED_T), 0, (hipStream_t)stream, N, (int)hole,
                     reinterpret_cast<const cd *>(ne0), reinterpret_cast<const cd *>(nh0), VC, k3, Ceh, Cx, Win,
                     Wout);
  PSA_LAUNCH_CHECK("k_mbc");
  return PSA_OK;
}

int psa_mod_mbp(int32_t N, int32_t hole, const double *n, const double *VC, const double *E1D, const double *P,
                const double *PT, double Vscale, double *Win, double *Wout, void *stream) {
  if (N < 1 || N > 1024 || (hole != 0 && hole != 1) || !n || !VC || !E1D || !P || !PT || !Win || !Wout)
    return fail(PSA_E_INVALID, "psa_mod_mbp: bad arguments");
  hipLaunchKernelGGL(mod::k_mbp, dim3(N), dim3(mod::RED_T), 0, (hipStream_t)stream, N, (int)hole,
                     reinterpret_cast<const cd *>(n), VC, E1D, P, PT, Vscale, Win, Wout);
  PSA_LAUNCH_CHECK("k_mbp");
  return PSA_OK;
}

int psa_mod_dft(int32_t n, int32_t nvec, int32_t sign, double scale, const double *in, double *out, void *stream) {
  if (n < 1 || nvec < 1 || nvec > 65535 || (sign != 1 && sign != -1) || !in || !out || in == out)
    return fail(PSA_E_INVALID, "psa_mod_dft: bad arguments");
  hipLaunchKernelGGL(mod::k_dft, dim3(n, nvec), dim3(mod::RED_T), 0, (hipStream_t)stream, n, (int)sign, scale,
                     reinterpret_cast<const cd *>(in), reinterpret_cast<cd *>(out));
  PSA_LAUNCH_CHECK("k_dft");
  return PSA_OK;
}

int psa_mod_resample(int32_t nout, int32_t nin, int32_t nvec, const double *M, const double *w, int32_t real_in,
                     const double *in, double *out, void *stream) {
  if (nout < 1 || nin < 1 || nvec < 1 || nvec > 65535 || !M || !in || !out || in == out)
    return fail(PSA_E_INVALID, "psa_mod_resample: bad arguments");
  hipLaunchKernelGGL(mod::k_resample, dim3(nout, nvec), dim3(mod::RED_T), 0, (hipStream_t)stream, nout, nin, M, w,
                     (int)(real_in != 0), reinterpret_cast<const cd *>(in), reinterpret_cast<cd *>(out));
  PSA_LAUNCH_CHECK("k_resample");
  return PSA_OK;
}

int psa_mod_pol3(int32_t N, int32_t Nr, const double *p, const double *Xvc, const double *Yvc, const double *Zvc,
                 const double *E, const double *W, double syw, double sz, double scale, double *out, void *stream) {
  if (N < 1 || N > 1024 || Nr < 1 || !p || !Xvc || !Yvc || !Zvc || !E || !W || !out)
    return fail(PSA_E_INVALID, "psa_mod_pol3: bad arguments");
  hipLaunchKernelGGL(mod::k_pol3, dim3(Nr), dim3(mod::RED_T), 0, (hipStream_t)stream, N, Nr,
                     reinterpret_cast<const cd *>(p), reinterpret_cast<const cd *>(Xvc),
                     reinterpret_cast<const cd *>(Yvc), reinterpret_cast<const cd *>(Zvc),
                     reinterpret_cast<const cd *>(E), W, syw, sz, scale, out);
  PSA_LAUNCH_CHECK("k_pol3");
  return PSA_OK;
}

int psa_mod_rho5(int32_t N, int32_t Nr, const double *C, const double *D, const double *E, const double *W, double L,
                 double dr, const double *ne, const double *nh, double *out, void *stream) {
  if (N < 1 || N > 1024 || Nr < 20 || !C || !D || !E || !W || !ne || !nh || !out || !(L > 0.0))
    return fail(PSA_E_INVALID, "psa_mod_rho5: bad arguments (Nr >= 20)");
  const hipStream_t s = (hipStream_t)stream;
  hipLaunchKernelGGL(mod::k_rho5, dim3(Nr), dim3(mod::RED_T), 0, s, N, Nr, reinterpret_cast<const cd *>(C),
                     reinterpret_cast<const cd *>(D), reinterpret_cast<const cd *>(E), W, L, reinterpret_cast<cd *>(out));
  PSA_LAUNCH_CHECK("k_rho5");
  hipLaunchKernelGGL(mod::k_rho5_finish, dim3(1), dim3(mod::RED_T), 0, s, N, Nr, dr, reinterpret_cast<const cd *>(ne),
                     reinterpret_cast<const cd *>(nh), reinterpret_cast<cd *>(out));
  PSA_LAUNCH_CHECK("k_rho5_finish");
  return PSA_OK;
}

int psa_mod_plasmonics(int32_t n, double dx, int32_t mode, double *re, double *rh, double *total, void *stream) {
  if (n < 1 || (mode != 0 && mode != 1) || !re || !rh || !total)
    return fail(PSA_E_INVALID, "psa_mod_plasmonics: bad arguments");
  hipLaunchKernelGGL(mod::k_plasmonics, dim3(1), dim3(mod::RED_T), 0, (hipStream_t)stream, n, dx, (int)mode,
                     reinterpret_cast<cd *>(re), reinterpret_cast<cd *>(rh), total);
  PSA_LAUNCH_CHECK("k_plasmonics");
  return PSA_OK;
}

int psa_mod_gamma(int32_t N, int32_t hole, const double *ne0, const double *nh0, const double *VC, const int32_t *iq,
                  const int32_t *tabs, const double *xe, const double *xh, double hbar, double *out, void *stream) {
  if (N < 1 || N > 1024 || (hole != 0 && hole != 1) || !ne0 || !nh0 || !VC || !iq || !tabs || !xe || !xh || !out)
    return fail(PSA_E_INVALID, "psa_mod_gamma: bad arguments");
  const long NN = (long)N * N;
  hipLaunchKernelGGL(mod::k_gamma, dim3(N), dim3(mod::RED_T), 0, (hipStream_t)stream, N, (int)hole,
                     reinterpret_cast<const cd *>(ne0), reinterpret_cast<const cd *>(nh0), VC, iq, tabs, tabs + NN,
                     tabs + 2 * NN, tabs + 3 * NN, xe, xh, PI / hbar, out);
  PSA_LAUNCH_CHECK("k_gamma");
  return PSA_OK;
}

int psa_mod_offdiag2(int32_t N, const double *ne0, const double *nh0, const double *p, const double *VC,
                     const double *Ee, const double *Eh, double gee, double ghh, double geh, double hbar, double *D,
                     double *x, void *stream) {
  if (N < 1 || N > 1024 || !ne0 || !nh0 || !VC || !Ee || !Eh || !D || (x && !p))
    return fail(PSA_E_INVALID, "psa_mod_offdiag2: bad arguments");
  const hipStream_t s = (hipStream_t)stream;
  hipLaunchKernelGGL(mod::k_od2_tables, dim3(2 * N + 1, N, 2), dim3(mod::RED_T), 0, s, N,
                     reinterpret_cast<const cd *>(ne0), reinterpret_cast<const cd *>(nh0), VC, Ee, Eh, hbar * gee,
                     hbar * ghh, hbar * geh, hbar, D);
  PSA_LAUNCH_CHECK("k_od2_tables");
  if (x) {
    const long NN = (long)N * N;
    hipLaunchKernelGGL(mod::k_od2_apply, dim3((unsigned)((NN + mod::RED_T - 1) / mod::RED_T)), dim3(mod::RED_T), 0, s,
                       N, D, reinterpret_cast<const cd *>(p), hbar, reinterpret_cast<cd *>(x));
    PSA_LAUNCH_CHECK("k_od2_apply");
  }
  return PSA_OK;
}

int psa_mod_spont(int32_t N, int32_t Nhw, const double *ne0, const double *nh0, const double *VC, const double *Ee,
                  const double *Eh, double gap, const double *homega, const double *square, double dhw, double rden,
                  double *Rsp, double *ec, void *stream) {
  if (N < 1 || Nhw < 2 || !ne0 || !nh0 || !VC || !Ee || !Eh || !homega || !square || !Rsp)
    return fail(PSA_E_INVALID, "psa_mod_spont: bad arguments");
  hipLaunchKernelGGL(mod::k_spont, dim3(N), dim3(mod::RED_T), 0, (hipStream_t)stream, N, Nhw,
                     reinterpret_cast<const cd *>(ne0), reinterpret_cast<const cd *>(nh0), VC, Ee, Eh, gap, homega,
                     square, dhw, rden, Rsp, ec);
  PSA_LAUNCH_CHECK("k_spont");
  return PSA_OK;
}

static int mod_blocks(long n) { return (int)((n + mod::RED_T - 1) / mod::RED_T); }

int psa_mod_mveh(int32_t N, int32_t Nf, const double *p, const double *VC, const int32_t *k3, const double *UnDel,
                 double *MVeh, void *stream) {
  if (N < 1 || N > 4096 || Nf < 1 || !p || !VC || !k3 || !UnDel || !MVeh)
    return fail(PSA_E_INVALID, "psa_mod_mveh: bad arguments");
  hipLaunchKernelGGL(mod::k_mveh, dim3(mod_blocks((long)N * N * Nf)), dim3(mod::RED_T), 0, (hipStream_t)stream, N,
                     Nf, reinterpret_cast<const cd *>(p), VC, k3, UnDel, reinterpret_cast<cd *>(MVeh));
  PSA_LAUNCH_CHECK("k_mveh");
  return PSA_OK;
}

int psa_mod_renorm(int32_t N, int32_t which, const double *ne, const double *nh, const double *VC,
                   const double *UnDel, double *BGR, void *stream) {
  if (N < 1 || N > 4096 || which < 0 || which > 2 || !VC || !UnDel || !BGR || (which != 2 && !ne) ||
      (which != 1 && !nh))
    return fail(PSA_E_INVALID, "psa_mod_renorm: bad arguments");
  hipLaunchKernelGGL(mod::k_renorm, dim3(mod_blocks((long)N * N)), dim3(mod::RED_T), 0, (hipStream_t)stream, N,
                     which, reinterpret_cast<const cd *>(ne), reinterpret_cast<const cd *>(nh), VC, UnDel,
                     reinterpret_cast<cd *>(BGR));
  PSA_LAUNCH_CHECK("k_renorm");
  return PSA_OK;
}

int psa_mod_vn1n2(int32_t N, const double *rcv, const double *Hcc, const double *Hhh, const double *Hcv, double hbar,
                  double *Vcc, double *Vvv, double *Vcv, double *Vvc, void *stream) {
  if (N < 1 || N > 8192 || !rcv || !Hcc || !Hhh || !Hcv || !Vcc || !Vvv || !Vcv || !Vvc || !(hbar > 0.0))
    return fail(PSA_E_INVALID, "psa_mod_vn1n2: bad arguments");
  using C = cd;
  hipLaunchKernelGGL(mod::k_vn1n2, dim3(mod_blocks((long)N * N)), dim3(mod::RED_T), 0, (hipStream_t)stream, N,
                     reinterpret_cast<const C *>(rcv), reinterpret_cast<const C *>(Hcc),
                     reinterpret_cast<const C *>(Hhh), reinterpret_cast<const C *>(Hcv), hbar,
                     reinterpret_cast<C *>(Vcc), reinterpret_cast<C *>(Vvv), reinterpret_cast<C *>(Vcv),
                     reinterpret_cast<C *>(Vvc));
  PSA_LAUNCH_CHECK("k_vn1n2");
  return PSA_OK;
}

int psa_mod_plspectrum(int32_t N, int32_t Nhw, const double *ne, const double *nh, const double *Ee,
                       const double *Eh, double gap, double hgeh, const double *VC, const double *hw, double rden,
                       double kB, double rscale, double tfac, double *work, double *PLS, void *stream) {
  if (N < 2 || Nhw < 1 || !ne || !nh || !Ee || !Eh || !VC || !hw || !work || !PLS || !(hgeh > 0.0) ||
      !(rden > 0.0) || !(kB > 0.0))
    return fail(PSA_E_INVALID, "psa_mod_plspectrum: bad arguments");
  const cd *n1 = reinterpret_cast<const cd *>(ne), *n2 = reinterpret_cast<const cd *>(nh);
  hipLaunchKernelGGL(mod::k_pl_prep, dim3(N + 1), dim3(mod::RED_T), 0, (hipStream_t)stream, N, n1, n2, Ee, Eh, gap,
                     VC, kB, work);
  PSA_LAUNCH_CHECK("k_pl_prep");
  hipLaunchKernelGGL(mod::k_pl, dim3(Nhw), dim3(mod::RED_T), 0, (hipStream_t)stream, N, n1, n2, hw, work, hgeh,
                     rden, kB, rscale, tfac, PLS);
  PSA_LAUNCH_CHECK("k_pl");
  return PSA_OK;
}

int psa_mod_offdiag1(int32_t N, int32_t hole, const double *ne0, const double *nh0, const double *VC, const double *Ee,
                     const double *Eh, const int32_t *kpq, const int32_t *kmq, double gss, double gso, double hbar,
                     double *D, void *stream) {
  if (N < 1 || N > 4096 || hole < 0 || hole > 1 || !ne0 || !nh0 || !VC || !Ee || !Eh || !kpq || !kmq || !D ||
      !(hbar > 0.0))
    return fail(PSA_E_INVALID, "psa_mod_offdiag1: bad arguments");
  hipLaunchKernelGGL(mod::k_od1_tables, dim3(mod_blocks((long)N * N)), dim3(mod::RED_T), 0, (hipStream_t)stream, N,
                     hole, reinterpret_cast<const cd *>(ne0), reinterpret_cast<const cd *>(nh0), VC, Ee, Eh, kpq, kmq,
                     gss, gso, hbar, D);
  PSA_LAUNCH_CHECK("k_od1_tables");
  return PSA_OK;
}

int psa_mod_offdiag1_apply(int32_t N, const double *DE, const double *DH, const double *p, const int32_t *kpq,
                           const double *ky, double hbar, double *x, void *stream) {
  if (N < 1 || N > 4096 || !DE || !DH || !p || !kpq || !ky || !x || !(hbar > 0.0))
    return fail(PSA_E_INVALID, "psa_mod_offdiag1_apply: bad arguments");
  hipLaunchKernelGGL(mod::k_od1_apply, dim3(mod_blocks((long)N * N)), dim3(mod::RED_T), 0, (hipStream_t)stream, N,
                     DE, DH, reinterpret_cast<const cd *>(p), kpq, ky, hbar, reinterpret_cast<cd *>(x));
  PSA_LAUNCH_CHECK("k_od1_apply");
  return PSA_OK;
}

int psa_mod_vehint(double dky, const double *y, int32_t Ny, double alphae, double alphah, double Delta0, double *out,
                   void *stream) {
  if (Ny < 4 || !y || !out || !(alphae > 0.0) || !(alphah > 0.0))
    return fail(PSA_E_INVALID, "psa_mod_vehint: bad arguments");
  hipLaunchKernelGGL(mod::k_vehint, dim3(1), dim3(mod::RED_T), 0, (hipStream_t)stream, dky, y, Ny, alphae, alphah,
                     Delta0, out);
  PSA_LAUNCH_CHECK("k_vehint");
  return PSA_OK;
}

}  // extern "C"
