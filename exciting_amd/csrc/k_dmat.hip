// Muffin-tin density matrix of the first-variational states (gendmatmt.f90:59-172, spinless):
// per atom a and k-point, wf1 = [A_a^T C; C(lo rows of a)] and
//     dm_a += w_k wf1 diag(occ) wf1^H.
// The library GEMMs do the products (A^T C for all atoms in one product per k-point; the
// k-sum as one GEMM per atom whose contraction runs over the columns of a range of k-points,
// one range per stream); these kernels move data between them and add the partial sums:
//  * pack: rows of A_a^T C from apwalm's storage order (io, lm) into the reference's row order
//    f = (l, io, m), m fastest (gendmatmt.f90:92-101), then the atom's local-orbital
//    coefficients (:123-131), into column block col0.. of X;
//  * scale: X2 = X diag(w_k occ) for one k-point's columns (wf2prime, :151-153).
#include <algorithm>

#include "exc_common.h"

namespace {

constexpr int kDT = 256;

__global__ __launch_bounds__(kDT) void dmat_pack_kernel(int naa, int ws, const int *order, const int *lo,
                                                        const cplx *w, int nst, const cplx *evec, int ldv,
                                                        int ngp, cplx *x, int ldx, size_t xs, int col0) {
  const int f = blockIdx.x * kDT + threadIdx.x, j = blockIdx.y, a = blockIdx.z;
  if (f >= ws) return;
  cplx v = {0.0, 0.0};
  if (f < naa) {
    v = w[(size_t)a * naa + order[f] + (size_t)j * gridDim.z * naa];   // w: (natm naa) x nst
  } else {
    const int r = f - naa;
    if (r < lo[2 * a + 1]) v = evec[ngp + lo[2 * a] + r + (size_t)j * ldv];
  }
  x[a * xs + f + (size_t)(col0 + j) * ldx] = v;
}

__global__ __launch_bounds__(kDT) void dmat_scale_kernel(int ws, const cplx *x, cplx *x2, int ldx, size_t xs,
                                                         int col0, const double *occ, double wk) {
  const int f = blockIdx.x * kDT + threadIdx.x, j = blockIdx.y, a = blockIdx.z;
  if (f >= ws) return;
  const double s = wk * occ[j];
  const size_t i = a * xs + f + (size_t)(col0 + j) * ldx;
  x2[i] = {x[i].re * s, x[i].im * s};
}

// dm[i] += sum_p part_p[i], part_p at p0 + p * pstride (cplx elements)
__global__ __launch_bounds__(kDT) void dmat_reduce_kernel(cplx *dm, const cplx *p0, int np, size_t pstride,
                                                          size_t n) {
  for (size_t i = (size_t)blockIdx.x * kDT + threadIdx.x; i < n; i += (size_t)gridDim.x * kDT) {
    cplx v = dm[i];
    for (int p = 0; p < np; ++p) {
      const cplx q = p0[p * pstride + i];
      v.re += q.re;
      v.im += q.im;
    }
    dm[i] = v;
  }
}

}  // namespace

int exc_launch_dmat_reduce(cplx *dm, const cplx *per, int np, size_t pstride, size_t poff, size_t n,
                           hipStream_t s) {
  if (n == 0 || np <= 0) return EXC_OK;
  const size_t blocks = std::min<size_t>((n + kDT - 1) / kDT, 4096);
  hipLaunchKernelGGL(dmat_reduce_kernel, dim3((unsigned)blocks), dim3(kDT), 0, s, dm, per + poff, np, pstride, n);
  EXC_CHECK(hipGetLastError());
  return EXC_OK;
}

int exc_launch_dmat_pack(const exc_dev &d, int ws, int nst, const cplx *w, const cplx *evec, int ldv, int ngp,
                         cplx *x, int ldx, size_t xs, int col0, hipStream_t s) {
  if (nst <= 0) return EXC_OK;
  if (nst > 65535 || d.natm > 65535 || ws < d.naa || ldx < ws) return EXC_ERR_ARG;
  hipLaunchKernelGGL(dmat_pack_kernel, dim3((ws + kDT - 1) / kDT, nst, d.natm), dim3(kDT), 0, s, d.naa, ws,
                     d.dm_order, d.dm_lo, w, nst, evec, ldv, ngp, x, ldx, xs, col0);
  EXC_CHECK(hipGetLastError());
  return EXC_OK;
}

int exc_launch_dmat_scale(const exc_dev &d, int ws, int nst, const cplx *x, cplx *x2, int ldx, size_t xs, int col0,
                          const double *occ, double wk, hipStream_t s) {
  if (nst <= 0) return EXC_OK;
  if (nst > 65535 || d.natm > 65535 || ldx < ws) return EXC_ERR_ARG;
  hipLaunchKernelGGL(dmat_scale_kernel, dim3((ws + kDT - 1) / kDT, nst, d.natm), dim3(kDT), 0, s, ws, x, x2, ldx,
                     xs, col0, occ, wk);
  EXC_CHECK(hipGetLastError());
  return EXC_OK;
}
