// Device code of the APW-lo and lo-lo blocks of H and O, shared by the per-k and the
// grouped kernels.
//
//  H(N+lo, i)  = sum_f loa_a(lo, f) A_a(i, f)          (hamiltonsetup.f90:131-145)
//  H(i, N+lo)  = conj(H(N+lo, i))                         (hamiltonsetup.f90:147-149)
//  O(i, N+lo)  = sum_io conj(A_a(i, io, lm)) oalo(io, ilo) (overlapsetup.f90:119-131)
//  O(N+lo, i)  = conj(O(i, N+lo))
//  H(lo, lo')  = lolo_a block on the atom diagonal, 0 elsewhere (hamiltonsetup.f90:154-161)
//  O(lo, lo')  = ololo(ilo, ilo') for equal atom, l and m (overlapsetup.f90:136-148)
// with the strict lower triangle of the lo-lo block set to conj(upper) as
// hamiltonoverlapocopy_UL.f90:15-27 leaves it.
#pragma once
#include "exc_common.h"

namespace lo_dev {

constexpr int kLoThreads = 128;
constexpr int kMaxLo = 16;          // lo functions per atom supported
constexpr int kRows = 64;           // APW rows (G+k) per workgroup of the APW-lo kernel
constexpr int kSlots = 4;           // lo functions handled in parallel per row
constexpr int kPerSlot = kMaxLo / kSlots;

// One workgroup = 64 APW rows x 4 lo slots of one atom.  Thread (r, s) accumulates the lo
// functions s, s+4, s+8, s+12 of its row; results go through LDS so that both the
// (row, lo) and the (lo, row) blocks are written in contiguous column segments.
struct ApwLoShared {
  cplx h[kMaxLo][kRows];
  cplx o[kMaxLo][kRows];
  int first, cnt;
};

__device__ __forceinline__ void lo_apw_body(ApwLoShared &sh, int bx, int ias, int ngp, int naa, int maxnlo, int nlomax,
                                            int apwordmax, int nlotot, const lo_desc *__restrict__ lod,
                                            const cplx *__restrict__ apwalm, int ldg, const cplx *__restrict__ loa,
                                            const double *__restrict__ oalo, cplx *__restrict__ h,
                                            cplx *__restrict__ o, int ldh) {
  cplx(&s_h)[kMaxLo][kRows] = sh.h;
  cplx(&s_o)[kMaxLo][kRows] = sh.o;
  int &s_first = sh.first;
  int &s_cnt = sh.cnt;
  const int tid = threadIdx.x;
  if (tid == 0) {
    int first = -1, cnt = 0;
    for (int q = 0; q < nlotot; ++q)
      if (lod[q].ias == ias) {
        if (first < 0) first = q;
        ++cnt;
      }
    s_first = first;
    s_cnt = cnt;
  }
  __syncthreads();
  const int first = s_first, cnt = s_cnt;
  if (cnt == 0) return;
  const int r = tid % kRows, slot = tid / kRows;
  const int i0 = bx * kRows;
  const int i = i0 + r;
  const cplx *A = apwalm + (size_t)ldg * naa * ias;
  const cplx *la = loa + (size_t)maxnlo * naa * ias;
  cplx acc[kPerSlot];
#pragma unroll
  for (int u = 0; u < kPerSlot; ++u) acc[u] = cplx{0.0, 0.0};
  if (i < ngp) {
    for (int f = 0; f < naa; ++f) {
      const cplx a = A[i + (size_t)f * ldg];
#pragma unroll
      for (int u = 0; u < kPerSlot; ++u) {
        const int q = slot + kSlots * u;
        if (q < cnt) {
          const cplx w = la[q + (size_t)maxnlo * f];   // loa row q of the atom = its q-th lo function
          acc[u].re += w.re * a.re - w.im * a.im;
          acc[u].im += w.re * a.im + w.im * a.re;
        }
      }
    }
  }
#pragma unroll
  for (int u = 0; u < kPerSlot; ++u) {
    const int q = slot + kSlots * u;
    if (q < cnt) {
      s_h[q][r] = acc[u];
      cplx ov = {0.0, 0.0};
      if (i < ngp) {
        const lo_desc dq = lod[first + q];
        for (int io = 0; io < apwordmax; ++io) {
          const cplx a = A[i + (size_t)(io + apwordmax * dq.lm) * ldg];
          const double w = oalo[io + apwordmax * (dq.ilo + nlomax * ias)];
          ov.re += a.re * w;
          ov.im -= a.im * w;
        }
      }
      s_o[q][r] = ov;
    }
  }
  __syncthreads();
  const int nrow = min(kRows, ngp - i0);
  // (row, lo) blocks: O is the primary, H the conjugate — contiguous over rows
  for (int e = tid; e < cnt * kRows; e += kRows * kSlots) {
    const int q = e / kRows, rr = e % kRows;
    if (rr >= nrow) continue;
    const size_t idx = (size_t)(i0 + rr) + (size_t)(ngp + first + q) * ldh;
    const cplx hv = s_h[q][rr];
    h[idx] = cplx{hv.re, -hv.im};
    o[idx] = s_o[q][rr];
  }
  // (lo, row) blocks: H is the primary (loa . A^T), O the conjugate — contiguous over lo
  for (int e = tid; e < cnt * kRows; e += kRows * kSlots) {
    const int rr = e / cnt, q = e % cnt;
    if (rr >= nrow) continue;
    const size_t idx = (size_t)(ngp + first + q) + (size_t)(i0 + rr) * ldh;
    const cplx ov = s_o[q][rr];
    h[idx] = s_h[q][rr];
    o[idx] = cplx{ov.re, -ov.im};
  }
}

__device__ __forceinline__ void lo_lo_body(int r, int c, int ngp, int nlotot, int maxnlo, int nlomax,
                                           const lo_desc *__restrict__ lod, const cplx *__restrict__ lolo,
                                           const double *__restrict__ ololo, cplx *__restrict__ h,
                                           cplx *__restrict__ o, int ldh) {
  if (r >= nlotot) return;
  const lo_desc dr = lod[r], dc = lod[c];
  cplx hv = {0.0, 0.0}, ov = {0.0, 0.0};
  if (dr.ias == dc.ias) {
    const cplx *ll = lolo + (size_t)maxnlo * maxnlo * dr.ias;
    if (r <= c) {
      hv = ll[dr.row + maxnlo * dc.row];
    } else {
      const cplx t = ll[dc.row + maxnlo * dr.row];
      hv = cplx{t.re, -t.im};
    }
    if (dr.l == dc.l && dr.m == dc.m) {
      const double *oo = ololo + (size_t)nlomax * nlomax * dr.ias;
      ov.re = (r <= c) ? oo[dr.ilo + nlomax * dc.ilo] : oo[dc.ilo + nlomax * dr.ilo];
    }
  }
  const size_t idx = (size_t)(ngp + r) + (size_t)(ngp + c) * ldh;
  h[idx] = hv;
  o[idx] = ov;
}

}  // namespace lo_dev
