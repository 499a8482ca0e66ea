// Per-SCF producers of the interstitial tables the setup reads: Theta(G) and Theta(r)
// (gencfun.f90:52-96, cfdamp = 0) and V(G) (genveffig.f90:23-33); and the interstitial
// density of the occupied states (genrhoir.f90:28-89): a chunk of states is scattered to the
// grid, transformed backward in one batched FFT and |psi|^2 accumulated per grid point.
// The FFT grid is the G box of the context (ngrid(i) = box extent, every grid point one G);
// position of G on the grid j1 + n1 (j2 + n2 j3) with negative components wrapped to n + i
// (gengvec.f90:100-117).  The transforms are hipFFT's double-complex 3-D ones with
// zfftifc.f90's convention: forward (sign -1) scaled by 1/N, backward (+1) unscaled.
#include <math.h>

#include "exc_common.h"

namespace {

constexpr int kST = 256;

struct Bvec {
  double b[9];   // reciprocal lattice vectors as columns: b[i + 3 j] = component i of b_j
};

__device__ __forceinline__ int grid_index(const exc_dev &d, const int *ivg, int g) {
  int j[3];
  for (int q = 0; q < 3; ++q) {
    const int v = ivg[3 * g + q];
    j[q] = v >= 0 ? v : d.box_n[q] + v;
  }
  return j[0] + d.box_n[0] * (j[1] + d.box_n[1] * j[2]);
}

// Theta(G) of every G, written at its grid position for the backward transform
__global__ __launch_bounds__(kST) void scf_cfunig_kernel(exc_dev d, Bvec B, int ngrtot, cplx *cf, cplx *z) {
  const int g = blockIdx.x * kST + threadIdx.x;
  if (g >= ngrtot) return;
  const int *iv = d.ivg + 3 * g;
  double G[3];
  for (int i = 0; i < 3; ++i) G[i] = B.b[i] * iv[0] + B.b[i + 3] * iv[1] + B.b[i + 6] * iv[2];
  const double gc = sqrt(G[0] * G[0] + G[1] * G[1] + G[2] * G[2]);
  const double t1 = 4.0 * M_PI / d.omega;
  cplx v = {g == 0 ? 1.0 : 0.0, 0.0};
  for (int a = 0; a < d.natm; ++a) {
    const double R = d.rmt[d.spidx[a]];
    double ff;
    if (gc > 1e-6) {
      const double t2 = gc * R;
      ff = t1 * (sin(t2) - t2 * cos(t2)) / (gc * gc * gc);
    } else {
      ff = (t1 / 3.0) * R * R * R;
    }
    const double *r = d.atposc + 3 * a;
    const double ph = -(G[0] * r[0] + G[1] * r[1] + G[2] * r[2]);
    v.re -= cos(ph) * ff;
    v.im -= sin(ph) * ff;
  }
  cf[g] = v;
  z[grid_index(d, d.ivg, g)] = v;
}

// gencfun.f90:108-114: Theta(G) *= exp(-(t1 |G|)^2), t1 = cfdamp / gmaxvr, after Theta(r) is
// taken from the undamped Theta(G); |G| as scf_cfunig_kernel forms it
__global__ __launch_bounds__(kST) void scf_cfdamp_kernel(exc_dev d, Bvec B, int ngrtot, double t1, cplx *cf) {
  const int g = blockIdx.x * kST + threadIdx.x;
  if (g >= ngrtot) return;
  const int *iv = d.ivg + 3 * g;
  double G[3];
  for (int i = 0; i < 3; ++i) G[i] = B.b[i] * iv[0] + B.b[i + 3] * iv[1] + B.b[i + 6] * iv[2];
  const double gc = sqrt(G[0] * G[0] + G[1] * G[1] + G[2] * G[2]);
  const double x = t1 * gc, t2 = exp(-(x * x));
  cf[g] = cplx{cf[g].re * t2, cf[g].im * t2};
}

__global__ __launch_bounds__(kST) void scf_real_kernel(int n, const cplx *z, double *out) {
  const int i = blockIdx.x * kST + threadIdx.x;
  if (i < n) out[i] = z[i].re;
}

__global__ __launch_bounds__(kST) void scf_product_kernel(int n, const double *a, const double *b, cplx *z) {
  const int i = blockIdx.x * kST + threadIdx.x;
  if (i < n) z[i] = {a[i] * b[i], 0.0};
}

__global__ __launch_bounds__(kST) void scf_gather_kernel(exc_dev d, int ngvec, double scale, const cplx *z,
                                                         cplx *out) {
  const int g = blockIdx.x * kST + threadIdx.x;
  if (g >= ngvec) return;
  const cplx v = z[grid_index(d, d.ivg, g)];
  out[g] = {v.re * scale, v.im * scale};
}

// z[b N + grid(G_igk)] = evec(igk, j0 + b): the plane-wave coefficients of state j0 + b
__global__ __launch_bounds__(kST) void rhoir_scatter_kernel(exc_dev d, int ngp, const int *igpig, const cplx *evec,
                                                            int ldv, int j0, int N, cplx *z) {
  const int igk = blockIdx.x * kST + threadIdx.x, b = blockIdx.y;
  if (igk >= ngp) return;
  const int g = igpig[igk] - 1;
  if (g < 0 || g >= N) return;
  z[(size_t)b * N + grid_index(d, d.ivg, g)] = evec[igk + (size_t)ldv * (j0 + b)];
}

// rho(r) += sum_b t_b |z_b(r)|^2, t_b = wk occ(j0 + b) / Omega, states with |wk occ| <= eps
// skipped; the weights of the chunk (at most kST) are formed once per workgroup
__global__ __launch_bounds__(kST) void rhoir_accum_kernel(int N, int nb, const cplx *z, const double *occ, int j0,
                                                          double wk, double eps, double omega, double *rho) {
  __shared__ double t2[kST];
  __shared__ int keep[kST];
  if (threadIdx.x < nb) {
    const double t1 = wk * occ[j0 + threadIdx.x];
    keep[threadIdx.x] = fabs(t1) > eps;
    t2[threadIdx.x] = t1 / omega;
  }
  __syncthreads();
  const int r = blockIdx.x * kST + threadIdx.x;
  if (r >= N) return;
  double s = 0.0;
  for (int b = 0; b < nb; ++b) {
    if (!keep[b]) continue;
    const cplx v = z[(size_t)b * N + r];
    s += t2[b] * (v.re * v.re + v.im * v.im);
  }
  rho[r] += s;
}

inline unsigned blocks(int n) { return (unsigned)((n + kST - 1) / kST); }

}  // namespace

int exc_launch_cfunig(const exc_dev &d, const double *bvec, int ngrtot, cplx *cf, cplx *z, hipStream_t s) {
  Bvec B;
  for (int i = 0; i < 9; ++i) B.b[i] = bvec[i];
  hipLaunchKernelGGL(scf_cfunig_kernel, dim3(blocks(ngrtot)), dim3(kST), 0, s, d, B, ngrtot, cf, z);
  EXC_CHECK(hipGetLastError());
  return EXC_OK;
}

int exc_launch_cfdamp(const exc_dev &d, const double *bvec, int ngrtot, double t1, cplx *cf, hipStream_t s) {
  Bvec B;
  for (int i = 0; i < 9; ++i) B.b[i] = bvec[i];
  hipLaunchKernelGGL(scf_cfdamp_kernel, dim3(blocks(ngrtot)), dim3(kST), 0, s, d, B, ngrtot, t1, cf);
  EXC_CHECK(hipGetLastError());
  return EXC_OK;
}

int exc_launch_scf_real(int n, const cplx *z, double *out, hipStream_t s) {
  hipLaunchKernelGGL(scf_real_kernel, dim3(blocks(n)), dim3(kST), 0, s, n, z, out);
  EXC_CHECK(hipGetLastError());
  return EXC_OK;
}

int exc_launch_scf_product(int n, const double *a, const double *b, cplx *z, hipStream_t s) {
  hipLaunchKernelGGL(scf_product_kernel, dim3(blocks(n)), dim3(kST), 0, s, n, a, b, z);
  EXC_CHECK(hipGetLastError());
  return EXC_OK;
}

int exc_launch_scf_gather(const exc_dev &d, int ngvec, double scale, const cplx *z, cplx *out, hipStream_t s) {
  hipLaunchKernelGGL(scf_gather_kernel, dim3(blocks(ngvec)), dim3(kST), 0, s, d, ngvec, scale, z, out);
  EXC_CHECK(hipGetLastError());
  return EXC_OK;
}

int exc_launch_rhoir_scatter(const exc_dev &d, int ngp, const int *igpig, const cplx *evec, int ldv, int j0, int nb,
                             int N, cplx *z, hipStream_t s) {
  if (ngp <= 0 || nb <= 0) return EXC_OK;
  if (nb > 65535) return EXC_ERR_ARG;
  hipLaunchKernelGGL(rhoir_scatter_kernel, dim3(blocks(ngp), nb), dim3(kST), 0, s, d, ngp, igpig, evec, ldv, j0, N, z);
  EXC_CHECK(hipGetLastError());
  return EXC_OK;
}

int exc_launch_rhoir_accum(int N, int nb, const cplx *z, const double *occ, int j0, double wk, double eps,
                           double omega, double *rho, hipStream_t s) {
  if (nb <= 0) return EXC_OK;
  if (nb > kST) return EXC_ERR_ARG;
  hipLaunchKernelGGL(rhoir_accum_kernel, dim3(blocks(N)), dim3(kST), 0, s, N, nb, z, occ, j0, wk, eps, omega, rho);
  EXC_CHECK(hipGetLastError());
  return EXC_OK;
}
