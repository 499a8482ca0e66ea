// Device code of the FP64 per-k H+O setup, shared by the per-k and the grouped kernels.
//
//  ystage : Yt_a = A_a . h_a^T for every atom a (hamiltonsetup.f90:75-88 computes its
//           transpose zm = h_a . apwi).
//  mt     : H(i,j) = sum_k conj(A(i,k)) Yt(j,k) + V(G_i-G_j) + 1/2 (G_i+k).(G_j+k) Theta(G_i-G_j)
//           O(i,j) = sum_k conj(A(i,k)) A(j,k)  + Theta(G_i-G_j)
//           over the upper triangle of tiles only, k running over all atoms' APW functions
//           (hamiltonsetup.f90:90-103 and overlapsetup.f90:95-108 summed over atoms, plus
//           hmlistln.f90:74-84 / olpistln.f90:69-82).  Each element with i <= j is written at
//           (i,j) and its conjugate at (j,i) (hamiltonoverlapocopy_UL.f90:15-27), so H and O
//           are written exactly once.
#pragma once
#include "zgemm_f64.h"

namespace setup_dev {

using namespace zf64;

struct GTab {
  const int *ivg;      // [3*ngrtot]
  const int *ivgig;    // box
  const cplx *cfunig;  // [ngrtot]
  const cplx *veffig;  // [ngvec]
  unsigned *miss;      // count of basis differences found outside the G set (exc_ctx_status)
  int ngvec;
  int lo0, lo1, lo2, n0, n1, n2;
};

inline GTab make_gtab(const exc_dev &d) {
  GTab t;
  t.ivg = d.ivg;
  t.ivgig = d.ivgig;
  t.cfunig = d.cfunig;
  t.veffig = d.veffig;
  t.miss = d.miss;
  t.ngvec = d.ngvec;
  t.lo0 = d.box_lo[0]; t.lo1 = d.box_lo[1]; t.lo2 = d.box_lo[2];
  t.n0 = d.box_n[0]; t.n1 = d.box_n[1]; t.n2 = d.box_n[2];
  return t;
}

// 1-based index of G = (d0,d1,d2) = G_i - G_j of two basis vectors in the G set, 0 if outside
// the box or beyond ngvec (the "ig > 0 .and. ig <= ngvec" test of hmlistln.f90:77).  The
// reference never reaches that case: it stops before any setup when 2 gkmax > gmaxvr
// (init1.f90:340-347), which is what keeps every basis difference inside the first ngvec
// G-vectors.  A miss is therefore counted, and exc_ctx_status reports it as an error.
__device__ __forceinline__ int lookup_g(const GTab &t, int d0, int d1, int d2) {
  const int a = d0 - t.lo0, b = d1 - t.lo1, c = d2 - t.lo2;
  int ig = 0;
  if ((unsigned)a < (unsigned)t.n0 && (unsigned)b < (unsigned)t.n1 && (unsigned)c < (unsigned)t.n2)
    ig = t.ivgig[a + t.n0 * (b + t.n1 * c)];
  if (ig > 0 && ig <= t.ngvec) return ig;
  if (t.miss) atomicAdd(t.miss, 1u);
  return 0;
}

// upper-triangular tile index t -> (ti, tj), ti <= tj, enumerated column by column
__device__ __forceinline__ void tri_decode(int t, int &ti, int &tj) {
  int j = (int)((sqrt(8.0 * t + 1.0) - 1.0) * 0.5);
  while ((j + 1) * (j + 2) / 2 <= t) ++j;
  while (j * (j + 1) / 2 > t) --j;
  tj = j;
  ti = t - j * (j + 1) / 2;
}

template <int BM, int BN, int BK, int WR, int WC>
struct YTile {
  using C = Core<BM, BN, BK, WR, WC, false>;
  // one (row tile, column tile) of Yt_a = A_a . h_a^T
  __device__ static void run(Smem<BM, BN, BK> &sm, int ngp, int naa, const cplx *__restrict__ A, int ldg,
                             const cplx *__restrict__ ha, cplx *__restrict__ Y, int ldy, int row0, int col0,
                             const int *__restrict__ fperm) {
    C core;
    core.run(sm, A, ldg, row0, ngp, ha, naa, col0, naa, naa, fperm);
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
};

template <int BM, int BK, int WR, int WC, bool DB = false>
struct MtTile {
  static constexpr int BN = BM;
  using C = Core<BM, BN, BK, WR, WC, true>;
  struct Shared {
    Smem<BM, BN, BK> sm[DB ? 2 : 1];
    int gr[BM][3], gc[BN][3];
    double vr[BM][3], vc[BN][3];
  };

  // one (tile, H-or-O) work item; the caller separates consecutive items with a barrier
  __device__ static void run(Shared &sh, int ngp, int ktot, const cplx *__restrict__ apwalm, int ldg,
                             const cplx *__restrict__ yt, int ldy, const int *__restrict__ igpig,
                             const double *__restrict__ vgpc, const GTab &gt, cplx *__restrict__ h,
                             cplx *__restrict__ o, int ldh, const int *__restrict__ kperm, int ti, int tj, int which) {
    constexpr int kThreads = C::kThreads;
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
        sh.gr[e][0] = g0; sh.gr[e][1] = g1; sh.gr[e][2] = g2;
        sh.vr[e][0] = v0; sh.vr[e][1] = v1; sh.vr[e][2] = v2;
      } else {
        const int c = e - BM;
        sh.gc[c][0] = g0; sh.gc[c][1] = g1; sh.gc[c][2] = g2;
        sh.vc[c][0] = v0; sh.vc[c][1] = v1; sh.vc[c][2] = v2;
      }
    }
    C core;
    const cplx *B = which ? apwalm : yt;
    const int ldb = which ? ldg : ldy;
    if constexpr (DB)
      core.run2(sh.sm, apwalm, ldg, row0, ngp, B, ldb, col0, ngp, ktot, kperm);
    else
      core.run(sh.sm[0], apwalm, ldg, row0, ngp, B, ldb, col0, ngp, ktot, kperm);

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
          const int ig = lookup_g(gt, sh.gr[rl][0] - sh.gc[cl][0], sh.gr[rl][1] - sh.gc[cl][1], sh.gr[rl][2] - sh.gc[cl][2]);
          if (ig) {
            const cplx th = gt.cfunig[ig - 1];
            if (which) {
              re += th.re;
              im += th.im;
            } else {
              const cplx vv = gt.veffig[ig - 1];
              const double t1 = 0.5 * (sh.vr[rl][0] * sh.vc[cl][0] + sh.vr[rl][1] * sh.vc[cl][1] + sh.vr[rl][2] * sh.vc[cl][2]);
              re += vv.re + t1 * th.re;
              im += vv.im + t1 * th.im;
            }
          }
          out[r + (size_t)c * ldh] = cplx{re, im};
          if (r < c) out[c + (size_t)r * ldh] = cplx{re, -im};
        }
      }
  }
};

// H and O of one (tile row, tile column) pair at once: a BM x 2BM tile over the columns
// [Yt_tj | A_tj], so the A row panel is fetched once for both (a quarter less operand
// traffic than two separate items).  Same summation order as MtTile: bit-identical output.
template <int BM, int BK, int WR, int WC>
struct MtTileHO {
  static constexpr int BN = 2 * BM;
  using C = Core<BM, BN, BK, WR, WC, true>;
  static_assert(C::WN <= BM && BM % C::WN == 0, "a wave's columns must not straddle H and O");
  struct Shared {
    Smem<BM, BN, BK> sm[2];
    int gr[BM][3], gc[BM][3];
    double vr[BM][3], vc[BM][3];
  };

  __device__ static void run(Shared &sh, int ngp, int ktot, const cplx *__restrict__ apwalm, int ldg,
                             const cplx *__restrict__ yt, int ldy, const int *__restrict__ igpig,
                             const double *__restrict__ vgpc, const GTab &gt, cplx *__restrict__ h,
                             cplx *__restrict__ o, int ldh, const int *__restrict__ kperm, int ti, int tj) {
    constexpr int kThreads = C::kThreads;
    const int row0 = ti * BM, col0 = tj * BM;
    for (int e = threadIdx.x; e < 2 * BM; e += kThreads) {
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
        sh.gr[e][0] = g0; sh.gr[e][1] = g1; sh.gr[e][2] = g2;
        sh.vr[e][0] = v0; sh.vr[e][1] = v1; sh.vr[e][2] = v2;
      } else {
        const int c = e - BM;
        sh.gc[c][0] = g0; sh.gc[c][1] = g1; sh.gc[c][2] = g2;
        sh.vc[c][0] = v0; sh.vc[c][1] = v1; sh.vc[c][2] = v2;
      }
    }
    C core;
    core.run2_ho(sh.sm, apwalm, ldg, row0, ngp, yt, ldy, col0, ngp, ktot, kperm);

    const int wave = threadIdx.x >> 6, wr = wave / WC, wc = wave % WC;
    const bool which = wc * C::WN >= BM;             // this wave's columns belong to O
    cplx *out = which ? o : h;
#pragma unroll
    for (int i = 0; i < C::MB; ++i)
#pragma unroll
      for (int j = 0; j < C::NB; ++j) {
        const int cl = C::col_of(wc, j) - (which ? BM : 0);
        const int c = col0 + cl;
#pragma unroll
        for (int v = 0; v < 4; ++v) {
          const int rl = C::row_of(wr, i, v);
          const int r = row0 + rl;
          if (r > c || c >= ngp) continue;
          double re = core.cr[i][j][v], im = core.ci[i][j][v];
          const int ig = lookup_g(gt, sh.gr[rl][0] - sh.gc[cl][0], sh.gr[rl][1] - sh.gc[cl][1], sh.gr[rl][2] - sh.gc[cl][2]);
          if (ig) {
            const cplx th = gt.cfunig[ig - 1];
            if (which) {
              re += th.re;
              im += th.im;
            } else {
              const cplx vv = gt.veffig[ig - 1];
              const double t1 = 0.5 * (sh.vr[rl][0] * sh.vc[cl][0] + sh.vr[rl][1] * sh.vc[cl][1] + sh.vr[rl][2] * sh.vc[cl][2]);
              re += vv.re + t1 * th.re;
              im += vv.im + t1 * th.im;
            }
          }
          out[r + (size_t)c * ldh] = cplx{re, im};
          if (r < c) out[c + (size_t)r * ldh] = cplx{re, -im};
        }
      }
  }
};

}  // namespace setup_dev
