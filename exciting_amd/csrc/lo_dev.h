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
constexpr int kRows = 256;          // APW rows (G+k) per workgroup of the APW-lo kernel
constexpr int kSlots = 1;           // threads per row (the workgroup is kRows * kSlots threads)

// One workgroup = 256 APW rows of one atom, one thread per row accumulating all the atom's
// lo functions: each element of A is read once, the loa coefficients are wave-uniform
// (scalar loads).  Products are written as explicit FMAs so that the per-k and the grouped
// kernels, which inline this body, round identically.
struct ApwLoShared {
  int first, cnt;
};

__device__ __forceinline__ void lo_apw_body(ApwLoShared &sh, int bx, int ias, int ngp, int naa, int maxnlo, int nlomax,
                                            int apwordmax, int nlotot, const lo_desc *__restrict__ lod,
                                            const cplx *__restrict__ apwalm, int ldg, const cplx *__restrict__ loa,
                                            const double *__restrict__ oalo, const double *__restrict__ h1loa,
                                            cplx *__restrict__ h, cplx *__restrict__ o, int ldh) {
  const int tid = threadIdx.x;
  if (tid == 0) {
    int first = -1, cnt = 0;
    for (int q = 0; q < nlotot; ++q)
      if (lod[q].ias == ias) {
        if (first < 0) first = q;
        ++cnt;
      }
    sh.first = first;
    sh.cnt = cnt;
  }
  __syncthreads();
  const int first = sh.first, cnt = sh.cnt;
  const int i = bx * kRows + tid;
  if (cnt == 0 || i >= ngp) return;
  const cplx *A = apwalm + (size_t)ldg * naa * ias;
  const cplx *la = loa + (size_t)maxnlo * naa * ias;
  double hr[kMaxLo], hi[kMaxLo];
#pragma unroll
  for (int q = 0; q < kMaxLo; ++q) hr[q] = hi[q] = 0.0;
#pragma unroll 4
  for (int f = 0; f < naa; ++f) {
    const cplx a = A[i + (size_t)f * ldg];
#pragma unroll
    for (int q = 0; q < kMaxLo; ++q) {
      if (q < cnt) {
        const cplx w = la[q + (size_t)maxnlo * f];   // loa row q of the atom = its q-th lo function
        hr[q] = fma(w.re, a.re, hr[q]);
        hr[q] = fma(-w.im, a.im, hr[q]);
        hi[q] = fma(w.re, a.im, hi[q]);
        hi[q] = fma(w.im, a.re, hi[q]);
      }
    }
  }
  const size_t nlo0 = (size_t)ngp + first;
#pragma unroll
  for (int q = 0; q < kMaxLo; ++q) {
    if (q < cnt) {
      const lo_desc dq = lod[first + q];
      double orr = 0.0, oi = 0.0;
      for (int io = 0; io < apwordmax; ++io) {
        const cplx a = A[i + (size_t)(io + apwordmax * dq.lm) * ldg];
        // oalo + h1loa ('iora*', overlapsetup.f90:126; h1loa is zero otherwise)
        const size_t wi = io + apwordmax * (dq.ilo + nlomax * ias);
        const double w = h1loa ? oalo[wi] + h1loa[wi] : oalo[wi];
        orr = fma(a.re, w, orr);
        oi = fma(-a.im, w, oi);
      }
      // (row, lo), contiguous over the workgroup's rows: O primary, H = conj(H(lo, row));
      // (lo, row): H primary (loa . A^T), O = conj(O(row, lo)).  h or o is null when only
      // the other matrix is set up (a uniform branch).
      const size_t rl = (size_t)i + (nlo0 + q) * ldh;
      const size_t lr = (nlo0 + q) + (size_t)i * ldh;
      if (h) {
        h[rl] = cplx{hr[q], -hi[q]};
        h[lr] = cplx{hr[q], hi[q]};
      }
      if (o) {
        o[rl] = cplx{orr, oi};
        o[lr] = cplx{orr, -oi};
      }
    }
  }
}

__device__ __forceinline__ void lo_lo_body(int r, int c, int ngp, int nlotot, int maxnlo, int nlomax,
                                           const lo_desc *__restrict__ lod, const cplx *__restrict__ lolo,
                                           const double *__restrict__ ololo, const double *__restrict__ h1lolo,
                                           cplx *__restrict__ h, cplx *__restrict__ o, int ldh) {
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
      // ololo + h1lolo ('iora*', overlapsetup.f90:144)
      const size_t base = (size_t)nlomax * nlomax * dr.ias;
      const size_t oi = base + ((r <= c) ? dr.ilo + nlomax * dc.ilo : dc.ilo + nlomax * dr.ilo);
      ov.re = h1lolo ? ololo[oi] + h1lolo[oi] : ololo[oi];
    }
  }
  const size_t idx = (size_t)(ngp + r) + (size_t)(ngp + c) * ldh;
  if (h) h[idx] = hv;
  if (o) o[idx] = ov;
}

}  // namespace lo_dev
