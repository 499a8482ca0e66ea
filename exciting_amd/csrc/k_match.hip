// APW matching coefficients of one k-point on gfx950 (match.f90:13); the device code is in
// match_dev.h.
#include "match_dev.h"

namespace {

constexpr int kMatchThreads = 64;

template <int L>
__global__ __launch_bounds__(kMatchThreads) void match_kernel(
    int ngp, int apwordmax, const int *__restrict__ spidx, const double *__restrict__ rmt, double fpso,
    const double *__restrict__ dinv, const double *__restrict__ gpc, const double *__restrict__ tpgpc,
    const cplx *__restrict__ sfacgp, int ldsf, cplx *__restrict__ apwalm, int ldg) {
  const int ig = blockIdx.x * kMatchThreads + threadIdx.x;
  if (ig >= ngp) return;
  match_dev::match_body<L>(ig, blockIdx.y, apwordmax, spidx, rmt, fpso, dinv, gpc, tpgpc, sfacgp, ldsf, apwalm, ldg);
}

template <int L>
int launch_match_l(const exc_dev &d, int ngp, const double *gpc, const double *tpgpc,
                   const cplx *sfacgp, int ldsf, cplx *apwalm, int ldg, hipStream_t s) {
  dim3 grid((ngp + kMatchThreads - 1) / kMatchThreads, d.natm);
  const double fpso = 4.0 * M_PI / sqrt(d.omega);
  hipLaunchKernelGGL(match_kernel<L>, grid, dim3(kMatchThreads), 0, s, ngp, d.apwordmax, d.spidx, d.rmt, fpso,
                     d.dinv, gpc, tpgpc, sfacgp, ldsf, apwalm, ldg);
  EXC_CHECK(hipGetLastError());
  return EXC_OK;
}

}  // namespace
int exc_launch_match(const exc_dev &d, int ngp, const double *gpc, const double *tpgpc,
                     const cplx *sfacgp, int ldsf, cplx *apwalm, int ldg, hipStream_t s) {
  if (ngp <= 0) return EXC_OK;
  if (d.apwordmax != 2) return EXC_ERR_UNSUPPORTED;
  switch (d.lmaxapw) {
    case 1: return launch_match_l<1>(d, ngp, gpc, tpgpc, sfacgp, ldsf, apwalm, ldg, s);
    case 2: return launch_match_l<2>(d, ngp, gpc, tpgpc, sfacgp, ldsf, apwalm, ldg, s);
    case 3: return launch_match_l<3>(d, ngp, gpc, tpgpc, sfacgp, ldsf, apwalm, ldg, s);
    case 4: return launch_match_l<4>(d, ngp, gpc, tpgpc, sfacgp, ldsf, apwalm, ldg, s);
    case 5: return launch_match_l<5>(d, ngp, gpc, tpgpc, sfacgp, ldsf, apwalm, ldg, s);
    case 6: return launch_match_l<6>(d, ngp, gpc, tpgpc, sfacgp, ldsf, apwalm, ldg, s);
    case 7: return launch_match_l<7>(d, ngp, gpc, tpgpc, sfacgp, ldsf, apwalm, ldg, s);
    case 8: return launch_match_l<8>(d, ngp, gpc, tpgpc, sfacgp, ldsf, apwalm, ldg, s);
    case 9: return launch_match_l<9>(d, ngp, gpc, tpgpc, sfacgp, ldsf, apwalm, ldg, s);
    case 10: return launch_match_l<10>(d, ngp, gpc, tpgpc, sfacgp, ldsf, apwalm, ldg, s);
    case 11: return launch_match_l<11>(d, ngp, gpc, tpgpc, sfacgp, ldsf, apwalm, ldg, s);
    case 12: return launch_match_l<12>(d, ngp, gpc, tpgpc, sfacgp, ldsf, apwalm, ldg, s);
    default: return EXC_ERR_UNSUPPORTED;
  }
}
