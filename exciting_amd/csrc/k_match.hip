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

// The 'iora*' overlap operand of the per-k MFMA path: B = A L per atom and l, where
// I + h1aa(:,:,l,a) = L L^T (olpl), so that sum_f conj(B(i,f)) B(j,f) is the reference's
// sum_f conj(A(i,f)) ((I + h1aa)^T A^T)(f,j) (overlapsetup.f90:81-94 with h1aa symmetric).
// One thread per (row, lm, atom): B(:, io = 0) = L00 A0 + L10 A1, B(:, 1) = L11 A1.
__global__ __launch_bounds__(256) void iora_apw_kernel(int ngp, const cplx *__restrict__ A, int ldg, int lmmax, int nl,
                                                       const double *__restrict__ olpl, cplx *__restrict__ B,
                                                       int ldb) {
  const int g = blockIdx.x * 256 + threadIdx.x, lm = blockIdx.y, ias = blockIdx.z;
  if (g >= ngp) return;
  int l = (int)sqrt((double)lm);
  while (l * l > lm) --l;
  while ((l + 1) * (l + 1) <= lm) ++l;
  const size_t c0 = (size_t)2 * (lm + (size_t)lmmax * ias);
  const cplx a0 = A[g + c0 * ldg], a1 = A[g + (c0 + 1) * ldg];
  const double *lf = olpl + ((size_t)ias * nl + l) * 3;
  const double l00 = lf[0], l10 = lf[1], l11 = lf[2];
  B[g + c0 * ldb] = cplx{l00 * a0.re + l10 * a1.re, l00 * a0.im + l10 * a1.im};
  B[g + (c0 + 1) * ldb] = cplx{l11 * a1.re, l11 * a1.im};
}

}  // namespace

int exc_launch_iora_apw(const exc_dev &d, int ngp, const cplx *apwalm, int ldg, cplx *out, int ldo, hipStream_t s) {
  if (ngp <= 0) return EXC_OK;
  if (d.apwordmax != 2 || !(d.iora & 1) || !d.olpl) return EXC_ERR_STATE;
  const int nl = d.lmaxmat + 1;
  hipLaunchKernelGGL(iora_apw_kernel, dim3((ngp + 255) / 256, nl * nl, d.natm), dim3(256), 0, s, ngp, apwalm, ldg,
                     nl * nl, nl, d.olpl, out, ldo);
  EXC_CHECK(hipGetLastError());
  return EXC_OK;
}

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
