// FP64 MFMA complex "NT" GEMM main loop for gfx950:  C(i,j) = sum_k op(A(i,k)) * B(j,k)
// with A(i,k) at A[i + k*lda], B(j,k) at B[j + k*ldb] (both column-major, row index
// contiguous — the layout of apwalm viewed as N x (natm*naa)).  op = conj or identity.
//
// Workgroup = WR x WC waves, workgroup tile BM x BN, wave tile (BM/WR) x (BN/WC) built from
// 16x16 blocks of v_mfma_f64_16x16x4f64.  A complex product costs four real MFMAs:
//   conj(a) b : re += ar br + ai bi ; im += ar bi - ai br
//   a b       : re += ar br - ai bi ; im += ar bi + ai br
// Tiles of BK k-columns are staged through LDS with real and imaginary parts in separate
// arrays (so each MFMA operand is one ds_read_b64), and the next tile is prefetched into
// registers while the current one is multiplied.
#pragma once
#include "exc_common.h"

namespace zf64 {

template <int BM, int BN, int BK>
struct Smem {
  double a_re[BK][BM];
  double a_im[BK][BM];
  double b_re[BK][BN];
  double b_im[BK][BN];
};

template <int BM, int BN, int BK, int WR, int WC, bool CONJA>
struct Core {
  static constexpr int kThreads = 64 * WR * WC;
  static constexpr int WM = BM / WR, WN = BN / WC;
  static constexpr int MB = WM / 16, NB = WN / 16;
  static constexpr int LA = BM * BK / kThreads;   // complex loads per thread (A tile)
  static constexpr int LB = BN * BK / kThreads;
  static_assert(BM * BK % kThreads == 0 && BN * BK % kThreads == 0, "tile/thread mismatch");
  static_assert(BK % 4 == 0, "BK must be a multiple of 4");

  d4_t cr[MB][NB], ci[MB][NB];

  __device__ __forceinline__ void zero() {
#pragma unroll
    for (int i = 0; i < MB; ++i)
#pragma unroll
      for (int j = 0; j < NB; ++j) {
        cr[i][j] = d4_t{0.0, 0.0, 0.0, 0.0};
        ci[i][j] = d4_t{0.0, 0.0, 0.0, 0.0};
      }
  }

  // Load one BK slab of A (rows row0.., valid rows < mrows) and B into registers.  The
  // k-th step of the sum reads column kperm[k] of both operands: the callers order the
  // columns so that the small (high-l) terms are accumulated first, which keeps the long
  // sequential MFMA accumulation as accurate as a blocked BLAS sum (nullptr: natural order).
  __device__ __forceinline__ static void gload(const cplx *__restrict__ A, int lda, int row0, int mrows,
                                               const cplx *__restrict__ B, int ldb, int col0, int nrows,
                                               int k0, int ktot, const int *__restrict__ kperm, cplx (&ra)[LA],
                                               cplx (&rb)[LB]) {
    const int tid = threadIdx.x;
#pragma unroll
    for (int r = 0; r < LA; ++r) {
      const int e = tid + kThreads * r;
      const int row = e % BM, k = e / BM;
      const int gr = row0 + row, gk = k0 + k;
      ra[r] = (gr < mrows && gk < ktot) ? A[gr + (size_t)(kperm ? kperm[gk] : gk) * lda] : cplx{0.0, 0.0};
    }
#pragma unroll
    for (int r = 0; r < LB; ++r) {
      const int e = tid + kThreads * r;
      const int row = e % BN, k = e / BN;
      const int gr = col0 + row, gk = k0 + k;
      rb[r] = (gr < nrows && gk < ktot) ? B[gr + (size_t)(kperm ? kperm[gk] : gk) * ldb] : cplx{0.0, 0.0};
    }
  }

  __device__ __forceinline__ static void sstore(Smem<BM, BN, BK> &sm, const cplx (&ra)[LA], const cplx (&rb)[LB]) {
    const int tid = threadIdx.x;
#pragma unroll
    for (int r = 0; r < LA; ++r) {
      const int e = tid + kThreads * r;
      sm.a_re[e / BM][e % BM] = ra[r].re;
      sm.a_im[e / BM][e % BM] = ra[r].im;
    }
#pragma unroll
    for (int r = 0; r < LB; ++r) {
      const int e = tid + kThreads * r;
      sm.b_re[e / BN][e % BN] = rb[r].re;
      sm.b_im[e / BN][e % BN] = rb[r].im;
    }
  }

  __device__ __forceinline__ void compute(const Smem<BM, BN, BK> &sm, int wr, int wc) {
    const int lane = threadIdx.x & 63;
    const int lk = lane >> 4, li = lane & 15;
#pragma unroll
    for (int kk = 0; kk < BK; kk += 4) {
      double ar[MB], ai[MB], br[NB], bi[NB];
#pragma unroll
      for (int i = 0; i < MB; ++i) {
        ar[i] = sm.a_re[kk + lk][wr * WM + i * 16 + li];
        ai[i] = sm.a_im[kk + lk][wr * WM + i * 16 + li];
      }
#pragma unroll
      for (int j = 0; j < NB; ++j) {
        br[j] = sm.b_re[kk + lk][wc * WN + j * 16 + li];
        bi[j] = sm.b_im[kk + lk][wc * WN + j * 16 + li];
      }
      double nai[MB];
#pragma unroll
      for (int i = 0; i < MB; ++i) nai[i] = -ai[i];
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

  // Full K loop over [0, ktot).
  __device__ __forceinline__ void run(Smem<BM, BN, BK> &sm, const cplx *__restrict__ A, int lda, int row0,
                                      int mrows, const cplx *__restrict__ B, int ldb, int col0, int nrows,
                                      int ktot, const int *__restrict__ kperm) {
    const int wave = threadIdx.x >> 6;
    const int wr = wave / WC, wc = wave % WC;
    zero();
    cplx ra[LA], rb[LB];
    gload(A, lda, row0, mrows, B, ldb, col0, nrows, 0, ktot, kperm, ra, rb);
    for (int k0 = 0; k0 < ktot; k0 += BK) {
      __syncthreads();
      sstore(sm, ra, rb);
      __syncthreads();
      if (k0 + BK < ktot) gload(A, lda, row0, mrows, B, ldb, col0, nrows, k0 + BK, ktot, kperm, ra, rb);
      compute(sm, wr, wc);
    }
  }

  // Full K loop with two LDS stages: the next stage is stored while the current one is
  // multiplied, one barrier per stage.

  // Accumulator element (v) of block (i, j) sits at tile row / column (the output layout of
  // v_mfma_f64_16x16x4f64, measured on gfx950: register v of lane l holds row 4v + l/16):
  //   row = wr*WM + i*16 + 4*v + lane/16,  col = wc*WN + j*16 + lane%16.
  __device__ __forceinline__ static int row_of(int wr, int i, int v) {
    return wr * WM + i * 16 + 4 * v + ((threadIdx.x & 63) >> 4);
  }
  __device__ __forceinline__ static int col_of(int wc, int j) { return wc * WN + j * 16 + (threadIdx.x & 15); }
};

}  // namespace zf64
