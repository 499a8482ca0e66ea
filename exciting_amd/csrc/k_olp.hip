// Muffin-tin overlap of the grouped setup by the addition theorem (EXC_BATCH_OLP_LEGENDRE).
//
// apwalm(G, io, lm, a) = (4 pi/sqrt(Omega)) i^l r_{a,io,l}(G) S_a(G) conj(Y_lm(G+k))
// (match.f90:83-187: r = D^-1 [j_l, |G+k| j_l'] is real, S_a = sfacgp), so the m-sum of the
// overlap's muffin-tin part collapses by sum_m Y_lm(u) conj(Y_lm(u')) = (2l+1)/(4 pi) P_l(u.u'):
//
//   O_MT(G, G') = sum_a conj(S_a(G)) S_a(G') sum_{l <= lmaxmat} P_l(u.u') sum_io rf(G) rf(G')
//
// with rf = (4 pi/sqrt(Omega)) sqrt((2l+1)/(4 pi)) r (written by the grouped match) and u the
// unit vector of G+k from tpgpc.  The l-sum is evaluated per element by Clenshaw's
// recurrence, l descending (P_{l+1} = (2l+1)/(l+1) x P_l - l/(l+1) P_{l-1}).  This is the
// same matrix as the reference's conj(A) A^T contraction (olpaa, overlapsetup), a different
// rounding: about 440 FP64 FMAs per element instead of 8 natm naa flops, on the VALU.
//
// Variants 0 and 1 — workgroup: one 64x64 tile of the upper triangle, 4 x CJ elements per
// thread (rows tr + 16 i, columns tc + (64 / CJ) j; variant 0: CJ = 2, 512 threads; variant
// 1: CJ = 4, 256 threads, half the LDS reads per FMA); per atom the rows' and columns' radial
// factors and structure factors are brought to LDS by DMA, the next atom's while the current
// one is evaluated (two stages).  Variant 2 (the default, below) keeps the atom loop out of
// LDS.  Epilogue as the MFMA tile: + Theta(G_i - G_j), store (i, j) and conj at (j, i).
#include "setup_dev.h"
#include "mt_dma.h"

namespace {

using namespace setup_dev;

constexpr int kOT = 64;        // tile edge
constexpr int kNLmax = 16;     // lmaxmat + 1 supported

struct d2 {
  double a, b;
};

// Clenshaw coefficients of the Legendre recurrence, P_{l+1} = al(l) x P_l + bl(l+1) P_{l-1}:
// al(l) = (2l+1)/(l+1), bl(l+1) = -(l+1)/(l+2) (correctly rounded at compile time, as the
// division at run time would be); read with scalar loads instead of two FP64 divisions per l
struct ClenshawTab {
  double al[kNLmax], bl[kNLmax];
};
constexpr ClenshawTab make_clenshaw() {
  ClenshawTab t{};
  for (int l = 0; l < kNLmax; ++l) {
    t.al[l] = (double)(2 * l + 1) / (double)(l + 1);
    t.bl[l] = -(double)(l + 1) / (double)(l + 2);
  }
  return t;
}
__constant__ ClenshawTab kClenshaw = make_clenshaw();
typedef double dv2 __attribute__((ext_vector_type(2)));

// One atom's operands in LDS, filled by DMA in lane order: the rows' and the columns'
// radial factors as (io0, io1) pairs, [row][l] with row stride nl pairs (for odd nl the
// 16-lane reads of 16 rows fall on distinct banks), then their structure factors.
struct __attribute__((aligned(16))) OlpStage {
  d2 rr[kOT * kNLmax], rc[kOT * kNLmax];
  cplx sr[kOT], sc[kOT];
};

// DMA of atom `at` into a stage: 2 nl + 2 wave-wide loads of 64 x 16 bytes, rows past N
// clamped to row N-1 (their products are never stored)
template <int NTHR>
__device__ __forceinline__ void olp_issue(OlpStage *st, const kdesc &k, int at, int nl, int row0, int col0) {
  const int lane = threadIdx.x & 63;
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int ngp = k.ngp;
  const d2 *rfa = (const d2 *)(k.rf + (size_t)at * ngp * 2 * nl);
  for (int j = wave; j < 2 * nl + 2; j += NTHR / 64) {
    const void *src;
    void *dst;
    if (j < 2 * nl) {
      const bool isr = j < nl;
      const int e = (isr ? j : j - nl) * 64 + lane;       // element of the [64][nl] block
      const int q = e / nl, l = e - q * nl;
      const int idx = min((isr ? row0 : col0) + q, ngp - 1);
      src = rfa + (size_t)idx * nl + l;
      dst = (isr ? st->rr : st->rc) + (isr ? j : j - nl) * 64;
    } else {
      const bool isr = j == 2 * nl;
      const int idx = min((isr ? row0 : col0) + lane, ngp - 1);
      src = k.sfacgp + idx + (size_t)k.ldg * at;
      dst = isr ? st->sr : st->sc;
    }
    __builtin_amdgcn_global_load_lds(src, (lds_void_t *)dst, 16, 0, 0);
  }
}

template <int CJ>
__global__ __launch_bounds__(64 * 64 / (4 * CJ)) __attribute__((amdgpu_waves_per_eu(CJ == 2 ? 4 : 2))) void olp_kernel(
    const kdesc *__restrict__ ks, const mt_item *__restrict__ items, int nitems, int natm, int nl, GTab gt) {
  constexpr int kThreads = 64 * 64 / (4 * CJ), CS = 64 / CJ;   // threads; column stride of a thread's elements
  __shared__ OlpStage st0;
  __shared__ OlpStage st1;
  __shared__ double ur[kOT][3], uc[kOT][3];
  __shared__ int gr[kOT][3], gc[kOT][3];
  const int item = blockIdx.x;
  if (item >= nitems) return;
  const mt_item w = items[item];
  const kdesc k = ks[w.k];
  const int ngp = k.ngp, row0 = w.ti * kOT, col0 = w.tj * kOT;
  const int tid = threadIdx.x, tr = tid & 15, tc = tid >> 4;
  olp_issue<kThreads>(&st0, k, 0, nl, row0, col0);
  for (int e = tid; e < 2 * kOT; e += kThreads) {
    const bool isr = e < kOT;
    const int q = isr ? e : e - kOT;
    const int idx = (isr ? row0 : col0) + q;
    double u0 = 0.0, u1 = 0.0, u2 = 0.0;
    int g0 = 0, g1 = 0, g2 = 0;
    if (idx < ngp) {
      double st, ct, sp, cp;
      sincos(k.tpgpc[2 * idx], &st, &ct);
      sincos(k.tpgpc[2 * idx + 1], &sp, &cp);
      u0 = st * cp; u1 = st * sp; u2 = ct;
      const int ig = k.igpig[idx] - 1;
      g0 = gt.ivg[3 * ig]; g1 = gt.ivg[3 * ig + 1]; g2 = gt.ivg[3 * ig + 2];
    }
    double(*u)[3] = isr ? ur : uc;
    int(*g)[3] = isr ? gr : gc;
    u[q][0] = u0; u[q][1] = u1; u[q][2] = u2;
    g[q][0] = g0; g[q][1] = g1; g[q][2] = g2;
  }
  __syncthreads();
  double x[4][CJ], are[4][CJ], aim[4][CJ];
#pragma unroll
  for (int i = 0; i < 4; ++i)
#pragma unroll
    for (int j = 0; j < CJ; ++j) {
      const int a = tr + 16 * i, b = tc + CS * j;
      x[i][j] = ur[a][0] * uc[b][0] + ur[a][1] * uc[b][1] + ur[a][2] * uc[b][2];
      are[i][j] = aim[i][j] = 0.0;
    }
  // one atom from stage cur while the next atom's DMA fills nxt
  auto atom = [&](const OlpStage *cur, OlpStage *nxt, int at) {
    wait_vm_lgkm0<0>();
    stage_barrier();                                  // atom `at` landed; nxt is no longer read
    if (at + 1 < natm) olp_issue<kThreads>(nxt, k, at + 1, nl, row0, col0);
    double b1[4][CJ], b2[4][CJ];
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
      for (int j = 0; j < CJ; ++j) b1[i][j] = b2[i][j] = 0.0;
#pragma unroll 1
    for (int l = nl - 1; l >= 0; --l) {
      const double al = kClenshaw.al[l], bl = kClenshaw.bl[l];
      dv2 p[4], q[CJ];
#pragma unroll
      for (int i = 0; i < 4; ++i) p[i] = *reinterpret_cast<const dv2 *>(&cur->rr[(tr + 16 * i) * nl + l]);
#pragma unroll
      for (int j = 0; j < CJ; ++j) q[j] = *reinterpret_cast<const dv2 *>(&cur->rc[(tc + CS * j) * nl + l]);
#pragma unroll
      for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int j = 0; j < CJ; ++j) {
          const double a = p[i].x * q[j].x + p[i].y * q[j].y;
          const double b0 = a + al * x[i][j] * b1[i][j] + bl * b2[i][j];
          b2[i][j] = b1[i][j];
          b1[i][j] = b0;
        }
    }
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
      for (int j = 0; j < CJ; ++j) {
        const cplx s = cur->sr[tr + 16 * i], t = cur->sc[tc + CS * j];
        are[i][j] += (s.re * t.re + s.im * t.im) * b1[i][j];     // conj(S(G)) S(G')
        aim[i][j] += (s.re * t.im - s.im * t.re) * b1[i][j];
      }
  };
  for (int at = 0; at < natm; at += 2) {
    atom(&st0, &st1, at);
    if (at + 1 < natm) atom(&st1, &st0, at + 1);
  }
#pragma unroll
  for (int j = 0; j < CJ; ++j) {
    const int cl = tc + CS * j, c = col0 + cl;
    if (c >= ngp) continue;
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      const int rl = tr + 16 * i, r = row0 + rl;
      if (r > c) continue;
      double re = are[i][j], im = aim[i][j];
      const int ig = lookup_g(gt, gr[rl][0] - gc[cl][0], gr[rl][1] - gc[cl][1], gr[rl][2] - gc[cl][2]);
      if (ig) {
        const cplx th = olp_theta(gt, ig, k.vgpc + 3 * r, k.vgpc + 3 * c);
        re += th.re;
        im += th.im;
      }
      k.o[r + (size_t)c * k.ldh] = cplx{re, im};
      if (r < c) k.o[c + (size_t)r * k.ldh] = cplx{re, -im};
    }
  }
}

// ---------------------------------------------------------------------------------------
// Variant 2: lanes along rows.  One 64 x 64 tile per workgroup of 8 waves; lane = row, wave w
// takes the 8 columns col0 + 8w .. +7, the same in all its lanes, so the columns' radial and
// structure factors are wave-uniform scalars: they are read through the scalar cache
// (constant-address-space loads: the factors are written by earlier kernels and only read
// here) and the atom loop touches no LDS (the LDS reads of variants 0 and 1 bound them).  The
// l-sum runs Clenshaw's recurrence on b~_l = gamma_l b_l with gamma_0 = 1,
// gamma_(l+1) = gamma_l (2l+1)/(l+1), which makes the coefficient of x b~_(l+1) one:
//   b~_l = gamma_l c_l + x b~_(l+1) + beta~_l b~_(l+2),  beta~_l = -(l+1)^2 / ((2l+1)(2l+3)),
// and b~_0 = b_0 = sum_l c_l P_l(x): four FP64 operations per element and l instead of five
// (rounded differently from variants 0 and 1).  The radial factors are first copied to
// rft[((atom (nl + 1) + nl - 1 - l) N + G) 2 + io] times sqrt(gamma_l) (olp_rft_kernel): one
// l of 64 rows is then 1 KiB contiguous, and the 8 columns of a wave 128 contiguous bytes.
// ---------------------------------------------------------------------------------------
struct ScaledClenshawTab {
  double gamma[kNLmax], beta[kNLmax];
};
constexpr ScaledClenshawTab make_scaled_clenshaw() {
  ScaledClenshawTab t{};
  double g = 1.0;
  for (int l = 0; l < kNLmax; ++l) {
    t.gamma[l] = g;
    g = g * (double)(2 * l + 1) / (double)(l + 1);
    t.beta[l] = -(double)((l + 1) * (l + 1)) / (double)((2 * l + 1) * (2 * l + 3));
  }
  return t;
}
__constant__ ScaledClenshawTab kScaled = make_scaled_clenshaw();

typedef const __attribute__((address_space(4))) double *cdbl_ptr;   // scalar-cache (constant) reads

// rows of rft per (atom, l): N rounded up to the 64 of a tile, the padding zero, so the
// columns of a tile never need clamping
__device__ __forceinline__ int rft_ld(int ngp) { return (ngp + kOT - 1) / kOT * kOT; }

// rf (per atom and row, l and io fastest) -> rft (per atom and slot, rows fastest): slot
// nl - 1 - l holds the radial factors of l times sqrt(gamma_l) (l descending, the order of
// the recurrence), slot nl the structure factor S_a(G) (re, im); rows past N are zero
__global__ __launch_bounds__(256) void olp_rft_kernel(const kdesc *__restrict__ ks, int nl) {
  const kdesc k = ks[blockIdx.z];
  const int row = blockIdx.x * 256 + threadIdx.x, at = blockIdx.y, ld = rft_ld(k.ngp);
  if (row >= ld) return;
  const d2 *src = (const d2 *)(k.rf + ((size_t)at * k.ngp + row) * 2 * nl);
  d2 *dst = (d2 *)k.rft + (size_t)at * (nl + 1) * ld + row;
  for (int l = 0; l < nl; ++l) {
    const double g = sqrt(kScaled.gamma[l]);
    const d2 v = row < k.ngp ? src[l] : d2{0.0, 0.0};
    dst[(size_t)(nl - 1 - l) * ld] = d2{v.a * g, v.b * g};
  }
  const cplx sf = row < k.ngp ? k.sfacgp[row + (size_t)k.ldg * at] : cplx{0.0, 0.0};
  dst[(size_t)nl * ld] = d2{sf.re, sf.im};
}

template <int NL>
__global__ __launch_bounds__(512) void olp_rows_kernel(const kdesc *__restrict__ ks, const mt_item *__restrict__ items,
                                                       int nitems, int natm, GTab gt) {
  constexpr int E = 8;                                  // columns per wave
  __shared__ double uc[kOT][3];
  __shared__ int gc[kOT][3];
  __shared__ cplx tr[8][E * 65];                        // mirror transpose, per wave [column][row] (+1 pad)
  const int item = blockIdx.x;
  if (item >= nitems) return;
  const mt_item w = items[item];
  const kdesc k = ks[w.k];
  const int ngp = k.ngp, row0 = w.ti * kOT, col0 = w.tj * kOT;
  const int lane = threadIdx.x & 63;
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int row = row0 + lane, rowc = min(row, ngp - 1);
  const int cw = col0 + E * wave;                       // first column of this wave
  if (threadIdx.x < kOT) {
    const int q = threadIdx.x, idx = col0 + q;
    double u0 = 0.0, u1 = 0.0, u2 = 0.0;
    int g0 = 0, g1 = 0, g2 = 0;
    if (idx < ngp) {
      double st, ct, sp, cp;
      sincos(k.tpgpc[2 * idx], &st, &ct);
      sincos(k.tpgpc[2 * idx + 1], &sp, &cp);
      u0 = st * cp; u1 = st * sp; u2 = ct;
      const int ig = k.igpig[idx] - 1;
      g0 = gt.ivg[3 * ig]; g1 = gt.ivg[3 * ig + 1]; g2 = gt.ivg[3 * ig + 2];
    }
    uc[q][0] = u0; uc[q][1] = u1; uc[q][2] = u2;
    gc[q][0] = g0; gc[q][1] = g1; gc[q][2] = g2;
  }
  double ur0, ur1, ur2;
  {
    double st, ct, sp, cp;
    sincos(k.tpgpc[2 * rowc], &st, &ct);
    sincos(k.tpgpc[2 * rowc + 1], &sp, &cp);
    ur0 = st * cp; ur1 = st * sp; ur2 = ct;
  }
  __syncthreads();
  double x[E], are[E], aim[E];
#pragma unroll
  for (int e = 0; e < E; ++e) {
    const int q = E * wave + e;
    x[e] = ur0 * uc[q][0] + ur1 * uc[q][1] + ur2 * uc[q][2];
    are[e] = aim[e] = 0.0;
  }
  // columns past N read the zero padding of rft (their products are never stored)
  const int ld = rft_ld(ngp);
  const d2 *rft = (const d2 *)k.rft;
#pragma unroll 1
  for (int at = 0; at < natm; ++at) {
    const cplx sr = k.sfacgp[rowc + (size_t)k.ldg * at];
    double b1[E], b2[E];
#pragma unroll
    for (int e = 0; e < E; ++e) b1[e] = b2[e] = 0.0;
    // the columns' factors of one l in SGPRs, those of the next l loaded while this one is
    // used; the scheduling barriers keep the scalar loads from being hoisted further (the
    // registers would spill) and let the rows' vector loads move ahead (mask: VMEM reads)
    auto col = [&](int l, d2 (&c)[E]) {
      cdbl_ptr pc = (cdbl_ptr)(rft + ((size_t)at * (NL + 1) + NL - 1 - l) * ld + cw);
#pragma unroll
      for (int e = 0; e < E; ++e) c[e] = d2{pc[2 * e], pc[2 * e + 1]};
    };
    d2 cur[E], nxt[E];
    col(NL - 1, cur);
#pragma unroll
    for (int l = NL - 1; l >= 0; --l) {
      const d2 r = rft[((size_t)at * (NL + 1) + NL - 1 - l) * ld + rowc];
      if (l > 0) col(l - 1, nxt);
      const double bt = kScaled.beta[l];
#pragma unroll
      for (int e = 0; e < E; ++e) {
        const double a = fma(r.b, cur[e].b, r.a * cur[e].a);
        const double b0 = fma(bt, b2[e], fma(x[e], b1[e], a));
        b2[e] = b1[e];
        b1[e] = b0;
      }
      __builtin_amdgcn_sched_barrier(0x0020);
#pragma unroll
      for (int e = 0; e < E; ++e) cur[e] = nxt[e];
    }
    // structure factors of the columns (rows ngp..ngkmax-1 of sfacgp are the caller's padding
    // and may be read: their products are never stored)
    cdbl_ptr ps = (cdbl_ptr)(k.sfacgp + (size_t)k.ldg * at);
#pragma unroll
    for (int e = 0; e < E; ++e) {
      const int ce = cw + e < k.ldg ? cw + e : 0;
      const cplx sc = {ps[2 * ce], ps[2 * ce + 1]};
      const double t1 = sc.re * b1[e], t2 = sc.im * b1[e];   // conj(S(G)) S(G') sum_l P_l c_l
      are[e] = fma(sr.re, t1, fma(sr.im, t2, are[e]));
      aim[e] = fma(sr.re, t2, fma(-sr.im, t1, aim[e]));
    }
  }
  // + Theta(G_i - G_j); store (i, j) (lanes along the column: coalesced) and conj at (j, i)
  // through this wave's LDS slice (8 lanes along each row of the mirror)
  int gr0 = 0, gr1 = 0, gr2 = 0;
  {
    const int ig = k.igpig[rowc] - 1;
    gr0 = gt.ivg[3 * ig]; gr1 = gt.ivg[3 * ig + 1]; gr2 = gt.ivg[3 * ig + 2];
  }
  cplx *tw = tr[wave];
#pragma unroll
  for (int e = 0; e < E; ++e) {
    const int q = E * wave + e, c = col0 + q;
    double re = are[e], im = aim[e];
    if (c < ngp && row <= c) {
      const int ig = lookup_g(gt, gr0 - gc[q][0], gr1 - gc[q][1], gr2 - gc[q][2]);
      if (ig) {
        const cplx th = olp_theta(gt, ig, k.vgpc + 3 * row, k.vgpc + 3 * c);
        re += th.re;
        im += th.im;
      }
      k.o[row + (size_t)c * k.ldh] = cplx{re, im};
    }
    tw[e * 65 + lane] = cplx{re, -im};
  }
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
#pragma unroll
  for (int i = 0; i < 8; ++i) {
    const int e = lane & 7, r = (lane >> 3) + 8 * i;     // row r of the tile, column e of the wave
    const int gr = row0 + r, c = cw + e;
    if (c < ngp && gr < c) k.o[c + (size_t)gr * k.ldh] = tw[e * 65 + r];
  }
}

template <int NL>
void launch_rows(const exc_dev &d, const kdesc *ks, const mt_item *items, int nitems, hipStream_t s) {
  hipLaunchKernelGGL(olp_rows_kernel<NL>, dim3(nitems), dim3(512), 0, s, ks, items, nitems, d.natm, make_gtab(d));
}

// ---------------------------------------------------------------------------------------
// Variant 3: variant 2's arithmetic (bit for bit) on a software pipeline.  Variant 2 waits for
// every l's operands as it needs them: the rows' factors (a vector load) and the columns'
// (scalar loads) come from L2 or further, and its waves stall on each.  Here one atom is
// NL + 1 steps (the l-sum, l descending, then the structure factors, from rft's last slot):
//  * the rows' operands of step s + 2 are loaded at step s (global loads, returned in order,
//    so the wait for step s leaves the later ones in flight);
//  * the columns' operands of step s + 1 are loaded at step s as soon as step s's have
//    arrived (scalar loads return out of order: any wait is for all of them, so the next
//    ones are issued only after it), with beta~ of step s + 1 from LDS;
//  * at most 128 VGPRs (a few spilled), so that two workgroups share a CU;
//  * workgroup b runs on XCD b % 8 and takes the next item of that XCD's contiguous eighth
//    of the list (which holds each k-point's tiles in 8 x 8 blocks): the tiles an XCD runs
//    together share their row and column factors in its L2.
// ---------------------------------------------------------------------------------------
typedef double gdv2 __attribute__((ext_vector_type(2)));
typedef const __attribute__((address_space(1))) gdv2 *gdv2_ptr;

// forces the wait for the scalar loads of `v` here (and of every scalar load before them);
// returns a zero the compiler cannot see through, so that loads indexed with it stay after
// this point (instead of being hoisted out of the loop into registers held throughout)
__device__ __forceinline__ int scalar_ready(double v) {
  int z;
  asm volatile("s_mov_b32 %0, 0" : "=s"(z) : "s"(v));
  return z;
}

template <int NL>
__global__ __launch_bounds__(512, 4) void olp_pipe_kernel(const kdesc *__restrict__ ks, const mt_item *__restrict__ items,
                                                       int nitems, int natm, GTab gt) {
  constexpr int E = 8, NS = NL + 1, D = 2;             // columns per wave; steps per atom; row prefetch distance
  static_assert(D <= NS, "prefetch within two atoms");
  __shared__ double uc[kOT][3];
  __shared__ int gc[kOT][3];
  __shared__ cplx tr[8][E * 65];
  __shared__ double bsh[NL];                            // beta~_l, read per step (held in no register)
  const int chunk = (nitems + 7) / 8, qb = blockIdx.x / 8;
  const int item = (blockIdx.x % 8) * chunk + qb;
  if (qb >= chunk || item >= nitems) return;
  const mt_item w = items[item];
  const kdesc k = ks[w.k];
  const int ngp = k.ngp, row0 = w.ti * kOT, col0 = w.tj * kOT;
  const int lane = threadIdx.x & 63;
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int row = row0 + lane, rowc = min(row, ngp - 1);
  const int cw = col0 + E * wave;
  if (threadIdx.x < kOT) {
    const int q = threadIdx.x, idx = col0 + q;
    double u0 = 0.0, u1 = 0.0, u2 = 0.0;
    int g0 = 0, g1 = 0, g2 = 0;
    if (idx < ngp) {
      double st, ct, sp, cp;
      sincos(k.tpgpc[2 * idx], &st, &ct);
      sincos(k.tpgpc[2 * idx + 1], &sp, &cp);
      u0 = st * cp; u1 = st * sp; u2 = ct;
      const int ig = k.igpig[idx] - 1;
      g0 = gt.ivg[3 * ig]; g1 = gt.ivg[3 * ig + 1]; g2 = gt.ivg[3 * ig + 2];
    }
    uc[q][0] = u0; uc[q][1] = u1; uc[q][2] = u2;
    gc[q][0] = g0; gc[q][1] = g1; gc[q][2] = g2;
  } else if (threadIdx.x < kOT + NL) {
    bsh[threadIdx.x - kOT] = kScaled.beta[threadIdx.x - kOT];
  }
  double ur0, ur1, ur2;
  {
    double st, ct, sp, cp;
    sincos(k.tpgpc[2 * rowc], &st, &ct);
    sincos(k.tpgpc[2 * rowc + 1], &sp, &cp);
    ur0 = st * cp; ur1 = st * sp; ur2 = ct;
  }
  __syncthreads();
  double x[E], are[E], aim[E];
#pragma unroll
  for (int e = 0; e < E; ++e) {
    const int q = E * wave + e;
    x[e] = ur0 * uc[q][0] + ur1 * uc[q][1] + ur2 * uc[q][2];
    are[e] = aim[e] = 0.0;
  }
  // rft is one stream of slots (atom-major, NL + 1 slots per atom, ld rows each), walked by two
  // running pointers: the rows' D slots ahead, the columns' one slot ahead.  Past the last
  // atom they read up to D slots further (the next k-point's factors, or the padding after
  // the last one): unused.  Rows and columns past N read rft's zero padding.
  const int ld = rft_ld(ngp);
  const size_t sb = (size_t)ld * 16;                    // bytes per slot
  const unsigned char *rsp = (const unsigned char *)k.rft;
  const unsigned char *csp = (const unsigned char *)k.rft + (size_t)cw * 16;
  const unsigned rofs = (unsigned)row * 16;
  auto rowld = [&]() -> gdv2 {
    const gdv2 v = *(gdv2_ptr)(rsp + rofs);
    rsp += sb;
    return v;
  };
  auto colld = [&](gdv2 (&c)[E]) {
    // the address is wave-uniform; readfirstlane keeps it (and the loads) scalar where the
    // compiler would otherwise fold it into the rows' per-lane pointer
    const unsigned long long a = (unsigned long long)csp;
    const unsigned lo = __builtin_amdgcn_readfirstlane((unsigned)a), hi = __builtin_amdgcn_readfirstlane((unsigned)(a >> 32));
    cdbl_ptr pc = (cdbl_ptr)(((unsigned long long)hi << 32) | lo);
#pragma unroll
    for (int e = 0; e < E; ++e) c[e] = gdv2{pc[2 * e], pc[2 * e + 1]};
    csp += sb;
  };
  gdv2 rr[NS + D];                                      // rr[s]: step s of this atom, rr[NS + t]: step t of the next
#pragma unroll
  for (int t = 0; t < D; ++t) rr[t] = rowld();
  gdv2 cur[E], nxt[E];
  double bcur = 0.0, bnxt = 0.0;
  colld(cur);
#pragma unroll 1
  for (int at = 0; at < natm; ++at) {
    double b1[E], b2[E];
#pragma unroll
    for (int s = 0; s < NS; ++s) {
      // the scheduling barriers (only vector-memory reads may cross them) keep the order
      // wait for this step's columns | issue the next step's | this step's arithmetic
      rr[s + D] = rowld();
      __builtin_amdgcn_sched_barrier(0x0020);
      const int z = scalar_ready(cur[0].x);
      __builtin_amdgcn_sched_barrier(0x0020);
      colld(nxt);
      // beta~ of the next step (steps 2 .. NL-1 use one) with the next columns
      if (s + 1 >= 2 && s + 1 < NL) bnxt = bsh[NL - 2 - s + z];
      __builtin_amdgcn_sched_barrier(0x0020);
      const gdv2 r = rr[s];
      if (s < NL) {
        const double bt = bcur;
#pragma unroll
        for (int e = 0; e < E; ++e) {
          const double a = fma(r.y, cur[e].y, r.x * cur[e].x);
          const double b0 = s == 0 ? a : (s == 1 ? fma(x[e], b1[e], a) : fma(bt, b2[e], fma(x[e], b1[e], a)));
          b2[e] = s == 0 ? 0.0 : b1[e];
          b1[e] = b0;
        }
      } else {
        // conj(S(G)) S(G') sum_l P_l c_l: r = S(G) of the row, cur = S(G') of the columns
#pragma unroll
        for (int e = 0; e < E; ++e) {
          const double t1 = cur[e].x * b1[e], t2 = cur[e].y * b1[e];
          are[e] = fma(r.x, t1, fma(r.y, t2, are[e]));
          aim[e] = fma(r.x, t2, fma(-r.y, t1, aim[e]));
        }
      }
#pragma unroll
      for (int e = 0; e < E; ++e) cur[e] = nxt[e];
      bcur = bnxt;
    }
#pragma unroll
    for (int t = 0; t < D; ++t) rr[t] = rr[NS + t];
  }
  // + Theta(G_i - G_j); store (i, j) and conj at (j, i) through this wave's LDS slice, as variant 2
  int gr0 = 0, gr1 = 0, gr2 = 0;
  {
    const int ig = k.igpig[rowc] - 1;
    gr0 = gt.ivg[3 * ig]; gr1 = gt.ivg[3 * ig + 1]; gr2 = gt.ivg[3 * ig + 2];
  }
  cplx *tw = tr[wave];
#pragma unroll
  for (int e = 0; e < E; ++e) {
    const int q = E * wave + e, c = col0 + q;
    double re = are[e], im = aim[e];
    if (c < ngp && row <= c) {
      const int ig = lookup_g(gt, gr0 - gc[q][0], gr1 - gc[q][1], gr2 - gc[q][2]);
      if (ig) {
        const cplx th = olp_theta(gt, ig, k.vgpc + 3 * row, k.vgpc + 3 * c);
        re += th.re;
        im += th.im;
      }
      k.o[row + (size_t)c * k.ldh] = cplx{re, im};
    }
    tw[e * 65 + lane] = cplx{re, -im};
  }
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
#pragma unroll
  for (int i = 0; i < 8; ++i) {
    const int e = lane & 7, r = (lane >> 3) + 8 * i;
    const int gr = row0 + r, c = cw + e;
    if (c < ngp && gr < c) k.o[c + (size_t)gr * k.ldh] = tw[e * 65 + r];
  }
}

template <int NL>
void launch_pipe(const exc_dev &d, const kdesc *ks, const mt_item *items, int nitems, hipStream_t s) {
  hipLaunchKernelGGL(olp_pipe_kernel<NL>, dim3((nitems + 7) / 8 * 8), dim3(512), 0, s, ks, items, nitems, d.natm,
                     make_gtab(d));
}

}  // namespace

int exc_olp_tile() { return kOT; }

int exc_launch_olp(const exc_dev &d, int nk, int maxngp, const kdesc *ks, const mt_item *items, int nitems, int variant,
                   hipStream_t s) {
  if (nitems <= 0) return EXC_OK;
  if (d.lmaxmat + 1 > kNLmax || d.apwordmax != 2) return EXC_ERR_UNSUPPORTED;
  switch (variant) {
    case 0:
      hipLaunchKernelGGL(olp_kernel<2>, dim3(nitems), dim3(512), 0, s, ks, items, nitems, d.natm, d.lmaxmat + 1,
                         make_gtab(d));
      break;
    case 1:
      hipLaunchKernelGGL(olp_kernel<4>, dim3(nitems), dim3(256), 0, s, ks, items, nitems, d.natm, d.lmaxmat + 1,
                         make_gtab(d));
      break;
    case 2:
      hipLaunchKernelGGL(olp_rft_kernel, dim3((maxngp + kOT - 1 + 255) / 256, d.natm, nk), dim3(256), 0, s, ks,
                         d.lmaxmat + 1);
      EXC_CHECK(hipGetLastError());
      switch (d.lmaxmat + 1) {
        case 2: launch_rows<2>(d, ks, items, nitems, s); break;
        case 3: launch_rows<3>(d, ks, items, nitems, s); break;
        case 4: launch_rows<4>(d, ks, items, nitems, s); break;
        case 5: launch_rows<5>(d, ks, items, nitems, s); break;
        case 6: launch_rows<6>(d, ks, items, nitems, s); break;
        case 7: launch_rows<7>(d, ks, items, nitems, s); break;
        case 8: launch_rows<8>(d, ks, items, nitems, s); break;
        case 9: launch_rows<9>(d, ks, items, nitems, s); break;
        case 10: launch_rows<10>(d, ks, items, nitems, s); break;
        case 11: launch_rows<11>(d, ks, items, nitems, s); break;
        case 12: launch_rows<12>(d, ks, items, nitems, s); break;
        case 13: launch_rows<13>(d, ks, items, nitems, s); break;
        default: return EXC_ERR_UNSUPPORTED;
      }
      break;
    case 3:
      hipLaunchKernelGGL(olp_rft_kernel, dim3((maxngp + kOT - 1 + 255) / 256, d.natm, nk), dim3(256), 0, s, ks,
                         d.lmaxmat + 1);
      EXC_CHECK(hipGetLastError());
      switch (d.lmaxmat + 1) {
        case 2: launch_pipe<2>(d, ks, items, nitems, s); break;
        case 3: launch_pipe<3>(d, ks, items, nitems, s); break;
        case 4: launch_pipe<4>(d, ks, items, nitems, s); break;
        case 5: launch_pipe<5>(d, ks, items, nitems, s); break;
        case 6: launch_pipe<6>(d, ks, items, nitems, s); break;
        case 7: launch_pipe<7>(d, ks, items, nitems, s); break;
        case 8: launch_pipe<8>(d, ks, items, nitems, s); break;
        case 9: launch_pipe<9>(d, ks, items, nitems, s); break;
        case 10: launch_pipe<10>(d, ks, items, nitems, s); break;
        case 11: launch_pipe<11>(d, ks, items, nitems, s); break;
        case 12: launch_pipe<12>(d, ks, items, nitems, s); break;
        case 13: launch_pipe<13>(d, ks, items, nitems, s); break;
        default: return EXC_ERR_UNSUPPORTED;
      }
      break;
    default:
      return EXC_ERR_ARG;
  }
  EXC_CHECK(hipGetLastError());
  return EXC_OK;
}
