// Complex FP64 MFMA tiles of the setup with their LDS filled by DMA (global_load_lds_dwordx4):
// the muffin-tin tile (MtTileDMA: conj(A) Yt^T or conj(A) A^T, then the interstitial term and
// the Hermitian mirror) and the Y-stage tile (YTileDMA: A_a h_a^T), without staging registers
// and with NST stages in flight.
//
// One wave-wide DMA moves one k-column of a 64-row panel: lane l reads the 16-byte complex
// element of row row0 + 64p + l and the LDS receives it at (stage base) + 16 l.  The
// column index of the k-th summand is kperm[k] (the l-descending order), identical for
// all lanes, so it is fetched with a scalar load.  Rows past the matrix are clamped to row
// 0 (their results are never stored); k-columns past K are zero-filled with LDS stores.
//
// The NST stages are separate __shared__ variables (EXC_DMA_STAGES) and the K loop rotates
// through them statically (unrolled by NST): the compiler can then tell the operand reads
// of the current stage from the DMA into the others and waits for that DMA only where this
// code asks for it (vmcnt <= the loads issued for the later stages).
//
// MFMA sequence per element identical to zf64::Core::compute (the adjoint's tiles) for the
// four-product form: all four-product variants are bit-identical to each other.
#pragma once
#include <type_traits>
#include "setup_dev.h"

namespace setup_dev {

typedef __attribute__((address_space(3))) void lds_void_t;

// the NST stage variables of a kernel: s0_, s1_ (and s2_, s3_ for NST 3, 4; a byte otherwise)
#define EXC_DMA_STAGES(St, NST)                                              \
  __shared__ St s0_;                                                         \
  __shared__ St s1_;                                                         \
  __shared__ typename std::conditional<((NST) > 2), St, char>::type s2_;     \
  __shared__ typename std::conditional<((NST) > 3), St, char>::type s3_;
#define EXC_DMA_STAGE_ARGS(St) &s0_, &s1_, (St *)&s2_, (St *)&s3_

// Workgroup barrier without a memory fence.  __syncthreads() carries a workgroup release
// fence, and since vmcnt counts the in-flight DMA as well as stores, the compiler then
// waits for all DMA before every barrier (in-flight stages would drain at each step).  The
// K loop instead waits explicitly (wait_vm_lgkm0: its own LDS stores and the DMA of the
// stage about to be read) and the empty asm statements keep the compiler from moving
// memory accesses across the barrier.
__device__ __forceinline__ void stage_barrier() {
  asm volatile("" ::: "memory");
  __builtin_amdgcn_s_barrier();
  asm volatile("" ::: "memory");
}

// s_waitcnt immediate (gfx9 encoding): vmcnt <= v, lgkmcnt = 0, expcnt not waited for
template <int V>
__device__ __forceinline__ void wait_vm_lgkm0() {
  static_assert(V >= 0 && V < 64, "vmcnt range");
  __builtin_amdgcn_s_waitcnt((V & 15) | (7 << 4) | (0 << 8) | ((V >> 4) << 14));
}

// C(i, j) = sum_k op(A(i, k)) B(j, k) over a BM x BN tile, A(i, k) at A[i + kperm[k] lda],
// B(j, k) at B[j + kperm[k] ldb], op = conj (CONJA) or identity; NST LDS stages filled by DMA.
// G3: three real products per complex product (Gauss's form, as zgemm3m) instead of four:
// P1 = sum ar br, P2 = sum ai bi, P3 = sum (ar -+ ai)(br + bi), combined after the K loop.
template <int BM, int BN, int BK, int NST, int WR, int WC, bool CONJA, bool G3 = false>
struct DmaCore {
  static constexpr int kWaves = WR * WC, kThreads = 64 * kWaves;
  static constexpr int WM = BM / WR, WN = BN / WC, MB = WM / 16, NB = WN / 16;
  static constexpr int RA = BM / 64, RB = BN / 64;   // 64-row blocks of the A and B panels
  static constexpr int NL = BK * (RA + RB);          // wave-wide DMAs per stage
  static constexpr int L = NL / kWaves;              // per wave
  static_assert(BM % 64 == 0 && BN % 64 == 0 && NL % kWaves == 0, "DMA split");
  static_assert(NST >= 2 && NST <= 4, "stages");
  static_assert(WM % 16 == 0 && WN % 16 == 0 && BK % 4 == 0, "MFMA blocks");
  struct __attribute__((aligned(16))) Stage {
    cplx a[BK][BM];
    cplx b[BK][BN];
  };
  // one complex operand as a 16-byte vector: a single ds_read_b128 (16 lanes per LDS pass,
  // contiguous), where two 8-byte reads of 32 lanes would collide on the banks of the k-rows
  // 1 KiB apart
  typedef double dv2 __attribute__((ext_vector_type(2)));

  // issue the DMAs of the K-step starting at k0 into stage st; rows past mrows / nrows are
  // clamped to row 0, k-columns past ktot are zero-filled
  __device__ __forceinline__ static void issue(Stage *st, const cplx *__restrict__ A, int lda, int row0, int mrows,
                                               const cplx *__restrict__ B, int ldb, int col0, int nrows, int ktot,
                                               const int *__restrict__ kperm, int k0) {
    const int lane = threadIdx.x & 63;
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    int kcs[L];
#pragma unroll
    for (int q = 0; q < L; ++q) {                     // wave-uniform: scalar loads, one wait
      const int j = wave + kWaves * q;
      const int kk = j < BK * RA ? j / RA : (j - BK * RA) / RB;
      const int gk = min(k0 + kk, ktot - 1);
      kcs[q] = kperm ? kperm[gk] : gk;
    }
#pragma unroll
    for (int q = 0; q < L; ++q) {
      const int j = wave + kWaves * q;
      const bool isb = j >= BK * RA;
      const int jj = isb ? j - BK * RA : j;
      const int kk = isb ? jj / RB : jj / RA, p = isb ? jj % RB : jj % RA;
      const int gk = k0 + kk;
      cplx *dst = (isb ? &st->b[kk][64 * p] : &st->a[kk][64 * p]);
      if (gk < ktot) {
        int gr = (isb ? col0 : row0) + 64 * p + lane;
        if (gr >= (isb ? nrows : mrows)) gr = 0;
        const cplx *src = (isb ? B + (size_t)kcs[q] * ldb : A + (size_t)kcs[q] * lda) + gr;
        __builtin_amdgcn_global_load_lds((const void *)src, (lds_void_t *)dst, 16, 0, 0);
      } else {
        dst[lane] = cplx{0.0, 0.0};
      }
    }
  }

  // the MFMA sequence of zf64::Core::compute, operands from the interleaved stage
  // kvalid: k-columns of this step inside [0, K); the 4-wide sub-steps past it hold only the
  // zero fill and are skipped (adding exact zeros would not change the sums)
  __device__ __forceinline__ static void compute(const Stage *st, d4_t (&cr)[MB][NB], d4_t (&ci)[MB][NB], int wr,
                                                 int wc, int kvalid) {
    const int lane = threadIdx.x & 63;
    const int lk = lane >> 4, li = lane & 15;
#pragma unroll
    for (int kk = 0; kk < BK; kk += 4) {
      if (kk >= kvalid) break;
      double ar[MB], ai[MB], br[NB], bi[NB], nai[MB];
#pragma unroll
      for (int i = 0; i < MB; ++i) {
        const dv2 v = *reinterpret_cast<const dv2 *>(&st->a[kk + lk][wr * WM + i * 16 + li]);
        ar[i] = v.x;
        ai[i] = v.y;
        nai[i] = -v.y;
      }
#pragma unroll
      for (int j = 0; j < NB; ++j) {
        const dv2 v = *reinterpret_cast<const dv2 *>(&st->b[kk + lk][wc * WN + j * 16 + li]);
        br[j] = v.x;
        bi[j] = v.y;
      }
#pragma unroll
      for (int i = 0; i < MB; ++i)
#pragma unroll
        for (int j = 0; j < NB; ++j) cr[i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(ar[i], br[j], cr[i][j], 0, 0, 0);
#pragma unroll
      for (int i = 0; i < MB; ++i)
#pragma unroll
        for (int j = 0; j < NB; ++j) ci[i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(ar[i], bi[j], ci[i][j], 0, 0, 0);
      if (CONJA) {
#pragma unroll
        for (int i = 0; i < MB; ++i)
#pragma unroll
          for (int j = 0; j < NB; ++j) cr[i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(ai[i], bi[j], cr[i][j], 0, 0, 0);
#pragma unroll
        for (int i = 0; i < MB; ++i)
#pragma unroll
          for (int j = 0; j < NB; ++j) ci[i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(nai[i], br[j], ci[i][j], 0, 0, 0);
      } else {
#pragma unroll
        for (int i = 0; i < MB; ++i)
#pragma unroll
          for (int j = 0; j < NB; ++j) cr[i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(nai[i], bi[j], cr[i][j], 0, 0, 0);
#pragma unroll
        for (int i = 0; i < MB; ++i)
#pragma unroll
          for (int j = 0; j < NB; ++j) ci[i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(ai[i], br[j], ci[i][j], 0, 0, 0);
      }
    }
  }

  // Gauss's form: p1 += ar br, p2 += ai bi, p3 += (ar + s ai)(br + bi), s = -1 for CONJA
  __device__ __forceinline__ static void compute3(const Stage *st, d4_t (&p1)[MB][NB], d4_t (&p2)[MB][NB],
                                                  d4_t (&p3)[MB][NB], int wr, int wc, int kvalid) {
    const int lane = threadIdx.x & 63;
    const int lk = lane >> 4, li = lane & 15;
#pragma unroll
    for (int kk = 0; kk < BK; kk += 4) {
      if (kk >= kvalid) break;
      double ar[MB], ai[MB], as[MB], br[NB], bi[NB], bs[NB];
#pragma unroll
      for (int i = 0; i < MB; ++i) {
        const dv2 v = *reinterpret_cast<const dv2 *>(&st->a[kk + lk][wr * WM + i * 16 + li]);
        ar[i] = v.x;
        ai[i] = v.y;
        as[i] = CONJA ? v.x - v.y : v.x + v.y;
      }
#pragma unroll
      for (int j = 0; j < NB; ++j) {
        const dv2 v = *reinterpret_cast<const dv2 *>(&st->b[kk + lk][wc * WN + j * 16 + li]);
        br[j] = v.x;
        bi[j] = v.y;
        bs[j] = v.x + v.y;
      }
#pragma unroll
      for (int i = 0; i < MB; ++i)
#pragma unroll
        for (int j = 0; j < NB; ++j) p1[i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(ar[i], br[j], p1[i][j], 0, 0, 0);
#pragma unroll
      for (int i = 0; i < MB; ++i)
#pragma unroll
        for (int j = 0; j < NB; ++j) p2[i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(ai[i], bi[j], p2[i][j], 0, 0, 0);
#pragma unroll
      for (int i = 0; i < MB; ++i)
#pragma unroll
        for (int j = 0; j < NB; ++j) p3[i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(as[i], bs[j], p3[i][j], 0, 0, 0);
    }
  }

  // the whole K loop.  On entry no DMA of this workgroup is in flight and every wave is done
  // with the stages; s2, s3 are unused below NST 3, 4.  A wave with active = false (its whole
  // sub-tile below the diagonal of a diagonal tile, never stored) issues its share of the DMA
  // and meets every barrier but skips the MFMAs.
  __device__ __forceinline__ static void run(Stage *s0, Stage *s1, Stage *s2, Stage *s3, const cplx *__restrict__ A,
                                             int lda, int row0, int mrows, const cplx *__restrict__ B, int ldb,
                                             int col0, int nrows, int ktot, const int *__restrict__ kperm,
                                             d4_t (&cr)[MB][NB], d4_t (&ci)[MB][NB], bool active = true) {
    const int nsteps = (ktot + BK - 1) / BK;
    Stage *st[4] = {s0, s1, s2, s3};
#pragma unroll
    for (int u = 0; u < NST - 1; ++u)
      if (u < nsteps) issue(st[u], A, lda, row0, mrows, B, ldb, col0, nrows, ktot, kperm, u * BK);
    const int wave = threadIdx.x >> 6, wr = wave / WC, wc = wave % WC;
    d4_t c3[G3 ? MB : 1][G3 ? NB : 1];   // P3 of Gauss's form (cr, ci hold P1, P2 meanwhile)
#pragma unroll
    for (int i = 0; i < MB; ++i)
#pragma unroll
      for (int j = 0; j < NB; ++j) cr[i][j] = ci[i][j] = d4_t{0.0, 0.0, 0.0, 0.0};
    if constexpr (G3) {
#pragma unroll
      for (int i = 0; i < MB; ++i)
#pragma unroll
        for (int j = 0; j < NB; ++j) c3[i][j] = d4_t{0.0, 0.0, 0.0, 0.0};
    }
    for (int base = 0; base < nsteps; base += NST) {
#pragma unroll
      for (int u = 0; u < NST; ++u) {
        const int step = base + u;
        if (step < nsteps) {
          // stage `step` has landed once at most the DMAs of the NST-2 later steps (all
          // full, else wait for everything) are outstanding
          if (step + NST - 2 < nsteps && (step + NST - 1) * BK <= ktot)
            wait_vm_lgkm0<(NST - 2) * L>();
          else
            wait_vm_lgkm0<0>();
          stage_barrier();
          if (step + NST - 1 < nsteps)
            issue(st[(u + NST - 1) % NST], A, lda, row0, mrows, B, ldb, col0, nrows, ktot, kperm,
                  (step + NST - 1) * BK);
          if (active) {
            if constexpr (G3)
              compute3(st[u], cr, ci, c3, wr, wc, ktot - step * BK);
            else
              compute(st[u], cr, ci, wr, wc, ktot - step * BK);
          }
        }
      }
    }
    if constexpr (G3) {   // conj(a) b: Re = P1 + P2, Im = P3 - P1 + P2;  a b: Re = P1 - P2, Im = P3 - P1 - P2
#pragma unroll
      for (int i = 0; i < MB; ++i)
#pragma unroll
        for (int j = 0; j < NB; ++j) {
          const d4_t p1 = cr[i][j], p2 = ci[i][j];
          cr[i][j] = CONJA ? p1 + p2 : p1 - p2;
          ci[i][j] = CONJA ? (c3[i][j] - p1) + p2 : (c3[i][j] - p1) - p2;
        }
    }
  }

  __device__ __forceinline__ static int row_of(int wr, int i, int v) {
    return wr * WM + i * 16 + 4 * v + ((threadIdx.x & 63) >> 4);
  }
  __device__ __forceinline__ static int col_of(int wc, int j) { return wc * WN + j * 16 + (threadIdx.x & 15); }
};

template <int BM, int BK, int NST, int WR, int WC, bool G3 = false>
struct MtTileDMA {
  static constexpr int BN = BM;
  using C = DmaCore<BM, BN, BK, NST, WR, WC, true, G3>;
  using Stage = typename C::Stage;
  static constexpr int kThreads = C::kThreads;
  struct Tables {
    int gr[BM][3], gc[BN][3];
    double vr[BM][3], vc[BN][3];
  };

  // one (tile, H-or-O) item; s0..s3 are the stage variables
  __device__ static void run(Tables &tb, Stage *s0, Stage *s1, Stage *s2, Stage *s3, int ngp, int ktot,
                             const cplx *__restrict__ apwalm, int ldg, const cplx *__restrict__ yt, int ldy,
                             const int *__restrict__ igpig, const double *__restrict__ vgpc, const GTab &gt,
                             cplx *__restrict__ h, cplx *__restrict__ o, int ldh, const int *__restrict__ kperm, int ti,
                             int tj, int which) {
    const int row0 = ti * BM, col0 = tj * BN;
    const cplx *B = which ? apwalm : yt;
    const int ldb = which ? ldg : ldy;
    // everything of the previous item (its stores and DMA) is complete and every wave is
    // past its epilogue before the stages and tables are refilled
    __builtin_amdgcn_s_waitcnt(0);
    __syncthreads();
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
        tb.gr[e][0] = g0; tb.gr[e][1] = g1; tb.gr[e][2] = g2;
        tb.vr[e][0] = v0; tb.vr[e][1] = v1; tb.vr[e][2] = v2;
      } else {
        const int c = e - BM;
        tb.gc[c][0] = g0; tb.gc[c][1] = g1; tb.gc[c][2] = g2;
        tb.vc[c][0] = v0; tb.vc[c][1] = v1; tb.vc[c][2] = v2;
      }
    }
    __builtin_amdgcn_s_waitcnt(0);                    // the table loads: only DMA from here on
    d4_t cr[C::MB][C::NB], ci[C::MB][C::NB];
    const int wave = threadIdx.x >> 6, wr = wave / WC, wc = wave % WC;
    const bool below = ti == tj && wr * C::WM > wc * C::WN + C::WN - 1;   // wave sub-tile entirely r > c
    C::run(s0, s1, s2, s3, apwalm, ldg, row0, ngp, B, ldb, col0, ngp, ktot, kperm, cr, ci, !below);

    cplx *out = which ? o : h;
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
          double re = cr[i][j][v], im = ci[i][j][v];
          const int ig = lookup_g(gt, tb.gr[rl][0] - tb.gc[cl][0], tb.gr[rl][1] - tb.gc[cl][1], tb.gr[rl][2] - tb.gc[cl][2]);
          if (ig) {
            if (which) {
              const cplx th = olp_theta(gt, ig, tb.vr[rl], tb.vc[cl]);
              re += th.re;
              im += th.im;
            } else {
              const cplx vv = gt.veffig[ig - 1];
              const double t1 = 0.5 * (tb.vr[rl][0] * tb.vc[cl][0] + tb.vr[rl][1] * tb.vc[cl][1] + tb.vr[rl][2] * tb.vc[cl][2]);
              re += vv.re + t1 * gt.kinig[ig - 1].re;
              im += vv.im + t1 * gt.kinig[ig - 1].im;
            }
          }
          out[r + (size_t)c * ldh] = cplx{re, im};
          if (r < c) out[c + (size_t)r * ldh] = cplx{re, -im};
        }
      }
  }
};

// One (row tile, column tile) of Yt_a = A_a h_a^T (no conjugation), summed over f in the
// order fperm; one tile per workgroup.  kfirst: the sum starts at position kfirst of that
// order (a table whose rows before it are zero for this tile's columns).
template <int BM, int BN, int BK, int NST, int WR, int WC, bool G3 = false>
struct YTileDMA {
  using C = DmaCore<BM, BN, BK, NST, WR, WC, false, G3>;
  using Stage = typename C::Stage;
  __device__ static void run(Stage *s0, Stage *s1, Stage *s2, Stage *s3, int ngp, int naa,
                             const cplx *__restrict__ A, int ldg, const cplx *__restrict__ ha, cplx *__restrict__ Y,
                             int ldy, int row0, int col0, const int *__restrict__ fperm, int kfirst = 0) {
    d4_t cr[C::MB][C::NB], ci[C::MB][C::NB];
    C::run(s0, s1, s2, s3, A, ldg, row0, ngp, ha, naa, col0, naa, naa - kfirst, fperm + kfirst, cr, ci);
    const int wave = threadIdx.x >> 6, wr = wave / WC, wc = wave % WC;
#pragma unroll
    for (int i = 0; i < C::MB; ++i)
#pragma unroll
      for (int j = 0; j < C::NB; ++j) {
        const int c = col0 + C::col_of(wc, j);
#pragma unroll
        for (int v = 0; v < 4; ++v) {
          const int r = row0 + C::row_of(wr, i, v);
          if (r < ngp && c < naa) Y[r + (size_t)c * ldy] = cplx{cr[i][j][v], ci[i][j][v]};
        }
      }
  }
};

}  // namespace setup_dev
