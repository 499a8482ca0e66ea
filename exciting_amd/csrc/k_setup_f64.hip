// Plain FP64 per-k H+O setup on gfx950.
//
//  ystage : Yt_a = A_a . h_a^T for every atom a (hamiltonsetup.f90:75-88 computes its
//           transpose zm = h_a . apwi), batched over atoms in grid.z.
//  mt     : H(i,j) = sum_k conj(A(i,k)) Yt(j,k) + V(G_i-G_j) + 1/2 (G_i+k).(G_j+k) Theta(G_i-G_j)
//           O(i,j) = sum_k conj(A(i,k)) A(j,k)  + Theta(G_i-G_j)
//           over the upper triangle of tiles only, k running over all atoms' APW functions
//           (hamiltonsetup.f90:90-103 and overlapsetup.f90:95-108 summed over atoms, plus
//           hmlistln.f90:74-84 / olpistln.f90:69-82).  Each element with i <= j is written at
//           (i,j) and its conjugate at (j,i) (hamiltonoverlapocopy_UL.f90:15-27), so H and O
//           are written exactly once.
#include "zgemm_f64.h"

namespace {

using namespace zf64;

// ------------------------------------------------------------------------------------------
// Y stage
// ------------------------------------------------------------------------------------------
template <int BM, int BN, int BK, int WR, int WC>
__global__ __launch_bounds__(64 * WR * WC) void ystage_kernel(int ngp, int naa, const cplx *__restrict__ apwalm,
                                                           int ldg, const cplx *__restrict__ haa,
                                                           cplx *__restrict__ yt, int ldy,
                                                           const int *__restrict__ fperm) {
  __shared__ Smem<BM, BN, BK> sm;
  using C = Core<BM, BN, BK, WR, WC, false>;
  C core;
  const int ias = blockIdx.z;
  const int row0 = blockIdx.x * BM, col0 = blockIdx.y * BN;
  const cplx *A = apwalm + (size_t)ldg * naa * ias;
  const cplx *B = haa + (size_t)naa * naa * ias;
  core.run(sm, A, ldg, row0, ngp, B, naa, col0, naa, naa, fperm);
  cplx *Y = yt + (size_t)ldy * naa * ias;
  const int wave = threadIdx.x >> 6, wr = wave / WC, wc = wave % WC;
#pragma unroll
  for (int i = 0; i < C::MB; ++i)
#pragma unroll
    for (int j = 0; j < C::NB; ++j) {
      const int c = col0 + C::col_of(wc, j);
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        const int r = row0 + C::row_of(wr, i, v);
        if (r < ngp && c < naa) Y[r + (size_t)c * ldy] = cplx{core.cr[i][j][v], core.ci[i][j][v]};
      }
    }
}

// ------------------------------------------------------------------------------------------
// MT + interstitial + Hermitian mirror
// ------------------------------------------------------------------------------------------
struct GTab {
  const int *ivg;      // [3*ngrtot]
  const int *ivgig;    // box
  const cplx *cfunig;  // [ngrtot]
  const cplx *veffig;  // [ngvec]
  int ngvec;
  int lo0, lo1, lo2, n0, n1, n2;
};

__device__ __forceinline__ int lookup_g(const GTab &t, int d0, int d1, int d2) {
  const int a = d0 - t.lo0, b = d1 - t.lo1, c = d2 - t.lo2;
  if ((unsigned)a >= (unsigned)t.n0 || (unsigned)b >= (unsigned)t.n1 || (unsigned)c >= (unsigned)t.n2) return 0;
  const int ig = t.ivgig[a + t.n0 * (b + t.n1 * c)];
  return (ig > 0 && ig <= t.ngvec) ? ig : 0;
}

// upper-triangular tile index t -> (ti, tj), ti <= tj, enumerated column by column
__device__ __forceinline__ void tri_decode(int t, int &ti, int &tj) {
  int j = (int)((sqrt(8.0 * t + 1.0) - 1.0) * 0.5);
  while ((j + 1) * (j + 2) / 2 <= t) ++j;
  while (j * (j + 1) / 2 > t) --j;
  tj = j;
  ti = t - j * (j + 1) / 2;
}

template <int BM, int BK, int WR, int WC>
__global__ __launch_bounds__(64 * WR * WC) void mt_kernel(int ngp, int ktot, int ntile, const cplx *__restrict__ apwalm,
                                                       int ldg, const cplx *__restrict__ yt, int ldy,
                                                       const int *__restrict__ igpig, const double *__restrict__ vgpc,
                                                       GTab gt, cplx *__restrict__ h, cplx *__restrict__ o, int ldh,
                                                       const int *__restrict__ kperm) {
  constexpr int BN = BM;
  __shared__ Smem<BM, BN, BK> sm;
  __shared__ int s_gr[BM][3], s_gc[BN][3];
  __shared__ double s_vr[BM][3], s_vc[BN][3];
  using C = Core<BM, BN, BK, WR, WC, true>;
  constexpr int kThreads = C::kThreads;

  // work item: (tile, H or O); H and O of one tile are adjacent
  const int ntri = ntile * (ntile + 1) / 2;
  const int item = blockIdx.x;
  if (item >= 2 * ntri) return;
  const int which = item & 1;      // 0: H, 1: O
  int ti, tj;
  tri_decode(item >> 1, ti, tj);
  const int row0 = ti * BM, col0 = tj * BN;

  // G-vector integer coordinates and G+k of this tile's rows and columns
  for (int e = threadIdx.x; e < BM + BN; e += kThreads) {
    const bool isr = e < BM;
    const int idx = isr ? row0 + e : col0 + (e - BM);
    int g0 = 0, g1 = 0, g2 = 0;
    double v0 = 0.0, v1 = 0.0, v2 = 0.0;
    if (idx < ngp) {
      const int ig = igpig[idx] - 1;
      g0 = gt.ivg[3 * ig]; g1 = gt.ivg[3 * ig + 1]; g2 = gt.ivg[3 * ig + 2];
      v0 = vgpc[3 * idx]; v1 = vgpc[3 * idx + 1]; v2 = vgpc[3 * idx + 2];
    }
    if (isr) {
      s_gr[e][0] = g0; s_gr[e][1] = g1; s_gr[e][2] = g2;
      s_vr[e][0] = v0; s_vr[e][1] = v1; s_vr[e][2] = v2;
    } else {
      const int c = e - BM;
      s_gc[c][0] = g0; s_gc[c][1] = g1; s_gc[c][2] = g2;
      s_vc[c][0] = v0; s_vc[c][1] = v1; s_vc[c][2] = v2;
    }
  }

  C core;
  const cplx *B = which ? apwalm : yt;
  const int ldb = which ? ldg : ldy;
  core.run(sm, apwalm, ldg, row0, ngp, B, ldb, col0, ngp, ktot, kperm);

  cplx *out = which ? o : h;
  const int wave = threadIdx.x >> 6, wr = wave / WC, wc = wave % WC;
#pragma unroll
  for (int i = 0; i < C::MB; ++i)
#pragma unroll
    for (int j = 0; j < C::NB; ++j) {
      const int cl = C::col_of(wc, j);
      const int c = col0 + cl;
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        const int rl = C::row_of(wr, i, v);
        const int r = row0 + rl;
        if (r > c || c >= ngp) continue;
        double re = core.cr[i][j][v], im = core.ci[i][j][v];
        const int ig = lookup_g(gt, s_gr[rl][0] - s_gc[cl][0], s_gr[rl][1] - s_gc[cl][1], s_gr[rl][2] - s_gc[cl][2]);
        if (ig) {
          const cplx th = gt.cfunig[ig - 1];
          if (which) {
            re += th.re;
            im += th.im;
          } else {
            const cplx vv = gt.veffig[ig - 1];
            const double t1 = 0.5 * (s_vr[rl][0] * s_vc[cl][0] + s_vr[rl][1] * s_vc[cl][1] + s_vr[rl][2] * s_vc[cl][2]);
            re += vv.re + t1 * th.re;
            im += vv.im + t1 * th.im;
          }
        }
        out[r + (size_t)c * ldh] = cplx{re, im};
        if (r < c) out[c + (size_t)r * ldh] = cplx{re, -im};
      }
    }
}

GTab make_gtab(const exc_dev &d) {
  GTab t;
  t.ivg = d.ivg;
  t.ivgig = d.ivgig;
  t.cfunig = d.cfunig;
  t.veffig = d.veffig;
  t.ngvec = d.ngvec;
  t.lo0 = d.box_lo[0]; t.lo1 = d.box_lo[1]; t.lo2 = d.box_lo[2];
  t.n0 = d.box_n[0]; t.n1 = d.box_n[1]; t.n2 = d.box_n[2];
  return t;
}

template <int BM, int BN, int BK, int WR, int WC>
int launch_ystage_t(const exc_dev &d, int ngp, const cplx *apwalm, int ldg, cplx *yt, int ldy, hipStream_t s) {
  dim3 grid((ngp + BM - 1) / BM, (d.naa + BN - 1) / BN, d.natm);
  hipLaunchKernelGGL((ystage_kernel<BM, BN, BK, WR, WC>), grid, dim3(64 * WR * WC), 0, s, ngp, d.naa, apwalm, ldg,
                     d.haa, yt, ldy, d.fperm);
  EXC_CHECK(hipGetLastError());
  return EXC_OK;
}

template <int BM, int BK, int WR, int WC>
int launch_mt_t(const exc_dev &d, int ngp, const cplx *apwalm, int ldg, const cplx *yt, int ldy, const int *igpig,
                const double *vgpc, cplx *h, cplx *o, int ldh, hipStream_t s) {
  const int ntile = (ngp + BM - 1) / BM;
  const int nitem = ntile * (ntile + 1);
  hipLaunchKernelGGL((mt_kernel<BM, BK, WR, WC>), dim3(nitem), dim3(64 * WR * WC), 0, s, ngp, d.natm * d.naa, ntile, apwalm, ldg,
                     yt, ldy, igpig, vgpc, make_gtab(d), h, o, ldh, d.kperm);
  EXC_CHECK(hipGetLastError());
  return EXC_OK;
}

}  // namespace

int exc_launch_ystage(const exc_dev &d, int ngp, const cplx *apwalm, int ldg, cplx *yt, int ldy, int variant,
                      hipStream_t s) {
  if (ngp <= 0) return EXC_OK;
  switch (variant) {
    case 0: return launch_ystage_t<128, 128, 16, 2, 4>(d, ngp, apwalm, ldg, yt, ldy, s);
    case 1: return launch_ystage_t<64, 64, 16, 2, 2>(d, ngp, apwalm, ldg, yt, ldy, s);
    default: return EXC_ERR_ARG;
  }
}

int exc_launch_mt(const exc_dev &d, int ngp, const cplx *apwalm, int ldg, const cplx *yt, int ldy, const int *igpig,
                  const double *vgpc, cplx *h, cplx *o, int ldh, int variant, hipStream_t s) {
  if (ngp <= 0) return EXC_OK;
  switch (variant) {
    case 0: return launch_mt_t<128, 16, 2, 4>(d, ngp, apwalm, ldg, yt, ldy, igpig, vgpc, h, o, ldh, s);
    case 1: return launch_mt_t<64, 16, 2, 2>(d, ngp, apwalm, ldg, yt, ldy, igpig, vgpc, h, o, ldh, s);
    default: return EXC_ERR_ARG;
  }
}
