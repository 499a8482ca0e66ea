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
// Workgroup: one 64x64 tile of the upper triangle, 4 x CJ elements per thread (rows tr + 16 i,
// columns tc + (64 / CJ) j; variant 0: CJ = 2, 512 threads; variant 1: CJ = 4, 256 threads,
// half the LDS reads per FMA); per atom the rows' and columns' radial factors and structure
// factors are brought to LDS by DMA, the next atom's while the current one is evaluated (two
// stages).  Epilogue as the MFMA tile: + Theta(G_i - G_j), store (i, j) and conj at (j, i).
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
        const cplx th = gt.cfunig[ig - 1];
        re += th.re;
        im += th.im;
      }
      k.o[r + (size_t)c * k.ldh] = cplx{re, im};
      if (r < c) k.o[c + (size_t)r * k.ldh] = cplx{re, -im};
    }
  }
}

}  // namespace

int exc_olp_tile() { return kOT; }

int exc_launch_olp(const exc_dev &d, const kdesc *ks, const mt_item *items, int nitems, int variant, hipStream_t s) {
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
    default:
      return EXC_ERR_ARG;
  }
  EXC_CHECK(hipGetLastError());
  return EXC_OK;
}
