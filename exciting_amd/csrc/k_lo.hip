// APW-lo and lo-lo blocks of H and O on gfx950.
//
//  H(N+lo, i)  = sum_f loa_a(lo, f) A_a(i, f)          (hamiltonsetup.f90:131-145)
//  H(i, N+lo)  = conj(H(N+lo, i))                         (hamiltonsetup.f90:147-149)
//  O(i, N+lo)  = sum_io conj(A_a(i, io, lm)) oalo(io, ilo) (overlapsetup.f90:119-131)
//  O(N+lo, i)  = conj(O(i, N+lo))
//  H(lo, lo')  = lolo_a block on the atom diagonal, 0 elsewhere (hamiltonsetup.f90:154-161)
//  O(lo, lo')  = ololo(ilo, ilo') for equal atom, l and m (overlapsetup.f90:136-148)
// with the strict lower triangle of the lo-lo block set to conj(upper) as
// hamiltonoverlapocopy_UL.f90:15-27 leaves it.
#include "exc_common.h"

namespace {

constexpr int kLoThreads = 128;
constexpr int kMaxLo = 16;

__global__ __launch_bounds__(kLoThreads) void lo_apw_kernel(int ngp, int naa, int maxnlo, int nlomax, int apwordmax,
                                                             int nlotot, const lo_desc *__restrict__ lod,
                                                             const cplx *__restrict__ apwalm, int ldg,
                                                             const cplx *__restrict__ loa, const double *__restrict__ oalo,
                                                             cplx *__restrict__ h, cplx *__restrict__ o, int ldh) {
  extern __shared__ cplx s_loa[];   // [naa][maxnlo] for this atom
  const int ias = blockIdx.y;
  // this atom's lo functions are a contiguous range [first, first+cnt)
  int first = -1, cnt = 0;
  for (int q = 0; q < nlotot; ++q)
    if (lod[q].ias == ias) {
      if (first < 0) first = q;
      ++cnt;
    }
  const cplx *la = loa + (size_t)maxnlo * naa * ias;
  for (int e = threadIdx.x; e < maxnlo * naa; e += kLoThreads) s_loa[e] = la[e];
  __syncthreads();
  const int i = blockIdx.x * kLoThreads + threadIdx.x;
  if (i >= ngp || cnt == 0) return;
  const cplx *A = apwalm + (size_t)ldg * naa * ias + i;
  cplx acc[kMaxLo];
#pragma unroll
  for (int q = 0; q < kMaxLo; ++q) acc[q] = cplx{0.0, 0.0};
  for (int f = 0; f < naa; ++f) {
    const cplx a = A[(size_t)f * ldg];
#pragma unroll
    for (int q = 0; q < kMaxLo; ++q) {
      if (q < cnt) {
        const cplx w = s_loa[q + maxnlo * f];
        acc[q].re += w.re * a.re - w.im * a.im;
        acc[q].im += w.re * a.im + w.im * a.re;
      }
    }
  }
  const int N = ngp;
#pragma unroll
  for (int q = 0; q < kMaxLo; ++q) {
    if (q >= cnt) break;
    const lo_desc dq = lod[first + q];
    const int col = N + first + q;
    // loa row q of the atom is its q-th lo function (idxlo is contiguous per atom)
    h[col + (size_t)i * ldh] = acc[q];
    h[i + (size_t)col * ldh] = cplx{acc[q].re, -acc[q].im};
    cplx ov = {0.0, 0.0};
    for (int io = 0; io < apwordmax; ++io) {
      const cplx a = A[(size_t)(io + apwordmax * dq.lm) * ldg];
      const double w = oalo[io + apwordmax * (dq.ilo + nlomax * ias)];
      ov.re += a.re * w;
      ov.im -= a.im * w;
    }
    o[i + (size_t)col * ldh] = ov;
    o[col + (size_t)i * ldh] = cplx{ov.re, -ov.im};
  }
}

__global__ __launch_bounds__(kLoThreads) void lo_lo_kernel(int ngp, int nlotot, int maxnlo, int nlomax,
                                                            const lo_desc *__restrict__ lod,
                                                            const cplx *__restrict__ lolo,
                                                            const double *__restrict__ ololo, cplx *__restrict__ h,
                                                            cplx *__restrict__ o, int ldh) {
  const int r = blockIdx.x * kLoThreads + threadIdx.x;
  const int c = blockIdx.y;
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

}  // namespace

int exc_launch_lo(const exc_dev &d, int ngp, const cplx *apwalm, int ldg, cplx *h, cplx *o, int ldh, hipStream_t s) {
  if (d.nlotot == 0) return EXC_OK;
  if (d.maxnlo > kMaxLo) return EXC_ERR_UNSUPPORTED;
  if (ngp > 0) {
    dim3 grid((ngp + kLoThreads - 1) / kLoThreads, d.natm);
    const size_t shm = sizeof(cplx) * d.maxnlo * d.naa;
    hipLaunchKernelGGL(lo_apw_kernel, grid, dim3(kLoThreads), shm, s, ngp, d.naa, d.maxnlo, d.nlomax, d.apwordmax,
                       d.nlotot, d.lod, apwalm, ldg, d.loa, d.oalo, h, o, ldh);
    EXC_CHECK(hipGetLastError());
  }
  dim3 grid2((d.nlotot + kLoThreads - 1) / kLoThreads, d.nlotot);
  hipLaunchKernelGGL(lo_lo_kernel, grid2, dim3(kLoThreads), 0, s, ngp, d.nlotot, d.maxnlo, d.nlomax, d.lod, d.lolo,
                     d.ololo, h, o, ldh);
  EXC_CHECK(hipGetLastError());
  return EXC_OK;
}
