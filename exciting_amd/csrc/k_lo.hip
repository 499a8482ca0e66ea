// APW-lo and lo-lo blocks of one k-point on gfx950 (device code in lo_dev.h).
#include "lo_dev.h"

namespace {

using namespace lo_dev;

__global__ __launch_bounds__(kRows * kSlots) void lo_apw_kernel(int ngp, int naa, int maxnlo, int nlomax, int apwordmax,
                                                                 int nlotot, const lo_desc *__restrict__ lod,
                                                                 const cplx *__restrict__ apwalm, int ldg,
                                                                 const cplx *__restrict__ loa,
                                                                 const double *__restrict__ oalo,
                                                                 const double *__restrict__ h1loa, cplx *__restrict__ h,
                                                                 cplx *__restrict__ o, int ldh) {
  __shared__ ApwLoShared sh;
  lo_apw_body(sh, blockIdx.x, blockIdx.y, ngp, naa, maxnlo, nlomax, apwordmax, nlotot, lod, apwalm, ldg, loa, oalo,
              h1loa, h, o, ldh);
}

__global__ __launch_bounds__(kLoThreads) void lo_lo_kernel(int ngp, int nlotot, int maxnlo, int nlomax,
                                                            const lo_desc *__restrict__ lod,
                                                            const cplx *__restrict__ lolo,
                                                            const double *__restrict__ ololo,
                                                            const double *__restrict__ h1lolo, cplx *__restrict__ h,
                                                            cplx *__restrict__ o, int ldh) {
  lo_lo_body(blockIdx.x * kLoThreads + threadIdx.x, blockIdx.y, ngp, nlotot, maxnlo, nlomax, lod, lolo, ololo, h1lolo,
             h, o, ldh);
}

}  // namespace

int exc_launch_lo(const exc_dev &d, int ngp, const cplx *apwalm, int ldg, cplx *h, cplx *o, int ldh, hipStream_t s) {
  if (d.nlotot == 0) return EXC_OK;
  if (d.maxnlo > kMaxLo) return EXC_ERR_UNSUPPORTED;
  if (ngp > 0) {
    dim3 grid((ngp + kRows - 1) / kRows, d.natm);
    hipLaunchKernelGGL(lo_apw_kernel, grid, dim3(kRows * kSlots), 0, s, ngp, d.naa, d.maxnlo, d.nlomax, d.apwordmax,
                       d.nlotot, d.lod, apwalm, ldg, d.loa, d.oalo, (d.iora & 2) ? d.h1loa : nullptr, h, o, ldh);
    EXC_CHECK(hipGetLastError());
  }
  dim3 grid2((d.nlotot + kLoThreads - 1) / kLoThreads, d.nlotot);
  hipLaunchKernelGGL(lo_lo_kernel, grid2, dim3(kLoThreads), 0, s, ngp, d.nlotot, d.maxnlo, d.nlomax, d.lod, d.lolo,
                     d.ololo, (d.iora & 2) ? d.h1lolo : nullptr, h, o, ldh);
  EXC_CHECK(hipGetLastError());
  return EXC_OK;
}
