// Device helpers of the generalized eigensolve (exc_eigen_dev, exc_eigen_batched_dev).
//
// solvewithlapack (modfvsystem.f90:585-720) hands ZHEGVX the upper triangles of H and O.
// The level-3 reduction C = U^-H H U^-1 multiplies by the whole of H, so the lower triangle
// is first made the conjugate of the upper one (and the diagonal real, as ZHETRD assumes).
//
// The grouped solve of many k-points copies every k-point's pair into a strided batch of
// one common order n = max nmat_k, padded as
//     H_pad = [H_k 0; 0 sigma I],   O_pad = [O_k 0; 0 I],
// for the strided-batched Cholesky, triangular inverse, trmm and tridiagonal reduction (one
// launch per step for all k-points; the per-k-point reduction is thousands of small
// launches, one column at a time).  The pad stays out of the leading nmat_k block of every
// result: the Cholesky factor and its inverse are block diagonal with an identity block, the
// reduced matrix is [C_k 0; 0 sigma I], the Householder reflectors of the columns of C_k touch
// only rows above their column, and those of the pad columns annihilate vectors that are
// already zero (tau = 0).  The leading nmat_k part of the tridiagonal matrix, the reflectors
// and U^-1 are therefore those of the k-point alone (to rounding: the blocking of the
// batched routines follows n), and the rest of the solve runs per k-point on that part only.
#include "exc_common.h"

namespace {

constexpr int kET = 32;          // tile edge of the fill and pad kernels
constexpr int kEThreads = 256;   // 32 rows x 8 column groups
constexpr int kSelThreads = 1024;

__device__ __forceinline__ cplx czero() { return {0.0, 0.0}; }

// In place: h(r, c) = conj(h(c, r)) for r > c, Im h(c, c) = 0.  Block (ti, tj), ti <= tj,
// reads tile (ti, tj) of the upper triangle and writes its conjugate transpose at (tj, ti).
__global__ __launch_bounds__(kEThreads) void herm_fill_kernel(cplx *h, int n, int ldh) {
  const int ti = blockIdx.x, tj = blockIdx.y;
  if (ti > tj) return;
  __shared__ cplx s[kET][kET + 1];   // s[c][r] = h(row0 + r, col0 + c)
  const int r = threadIdx.x % kET, cg = threadIdx.x / kET;
  const int row0 = ti * kET, col0 = tj * kET;
  for (int c = cg; c < kET; c += kEThreads / kET) {
    const int R = row0 + r, C = col0 + c;
    s[c][r] = (R < n && C < n) ? h[R + (size_t)C * ldh] : czero();
  }
  __syncthreads();
  // destination (col0 + r, row0 + c) = conj(h(row0 + c, col0 + r)) = conj(s[r][c]),
  // for destination rows strictly below the diagonal
  for (int c = cg; c < kET; c += kEThreads / kET) {
    const int R = col0 + r, C = row0 + c;
    if (R >= n || C >= n) continue;
    if (R > C) {
      h[R + (size_t)C * ldh] = cconj(s[r][c]);
    } else if (R == C) {
      h[R + (size_t)C * ldh].im = 0.0;
    }
  }
}

// Padded copy of matrix `which` (0: H with pad sigma, 1: O with pad 1) of every k-point of
// the batch into p + k * n * n (leading dimension n), both triangles from the upper one.
__global__ __launch_bounds__(kEThreads) void eig_pad_kernel(const eig_kdesc *ks, int n, cplx *p, int which,
                                                            double sigma) {
  const int ti = blockIdx.x, tj = blockIdx.y, k = blockIdx.z;
  if (ti > tj) return;
  const eig_kdesc kd = ks[k];
  const cplx *src = which ? kd.o : kd.h;
  const double dpad = which ? 1.0 : sigma;
  cplx *dst = p + (size_t)k * n * n;
  __shared__ cplx s[kET][kET + 1];
  const int r = threadIdx.x % kET, cg = threadIdx.x / kET;
  const int row0 = ti * kET, col0 = tj * kET;
  for (int c = cg; c < kET; c += kEThreads / kET) {
    const int R = row0 + r, C = col0 + c;
    if (R >= n || C >= n) continue;
    cplx v;
    if (R > C) {
      v = czero();   // below the diagonal of a diagonal tile: written from s below
    } else if (C < kd.n) {
      v = src[R + (size_t)C * kd.ldh];
      if (R == C) v.im = 0.0;
    } else {
      v = R == C ? cplx{dpad, 0.0} : czero();
    }
    s[c][r] = v;
    if (R <= C) dst[R + (size_t)C * n] = v;
  }
  __syncthreads();
  for (int c = cg; c < kET; c += kEThreads / kET) {
    const int R = col0 + r, C = row0 + c;
    if (R >= n || C >= n || R <= C) continue;
    dst[R + (size_t)C * n] = cconj(s[r][c]);
  }
}

// zc(:, j) = zr(:, j) for j < ncols (m rows): real eigenvectors of the tridiagonal matrix
// into the complex workspace of the back-transform
__global__ __launch_bounds__(256) void eig_r2c_kernel(const double *zr, int m, int ldr, cplx *zc, int ldc) {
  const int r = blockIdx.x * 256 + threadIdx.x, j = blockIdx.y;
  if (r >= m) return;
  zc[r + (size_t)j * ldc] = {zr[r + (size_t)j * ldr], 0.0};
}

// Block i0 .. i0+ib-1 of the reflectors of the upper tridiagonal reduction (zhetrd 'U':
// reflector i has v(0:i-1) = A(0:i-1, i+1), v(i) = 1, zero below) as explicit columns,
// V[k sv + r + n c] for r < m = i0 + ib, of every k-point k of the batch
__global__ __launch_bounds__(256) void eig_vblock_kernel(const cplx *A, int n, int i0, int ib, size_t sv, cplx *V) {
  const int r = blockIdx.x * 256 + threadIdx.x, c = blockIdx.y, k = blockIdx.z;
  const int m = i0 + ib;
  if (r >= m) return;
  const int i = i0 + c;
  cplx v = {0.0, 0.0};
  if (r < i) v = A[(size_t)k * n * n + r + (size_t)n * (i + 1)];
  else if (r == i) v.re = 1.0;
  V[(size_t)k * sv + r + (size_t)n * c] = v;
}

// T (ib x ib, lower triangular) of the block reflector H(i0+ib-1) ... H(i0) = I - V T V^H from
// G = V^H V and tau (zlarft, direct 'B', storev 'C'): for i = ib-1 down to 0,
// T(i, i) = tau_i, T(i+1:, i) = -tau_i T(i+1:, i+1:) G(i+1:, i).  The reflectors of k-point k
// beyond its own order (i >= nmat_k - 1: the pad columns and the coupling column, zero
// vectors) are left out (tau taken as 0).  One workgroup of kTB threads per k-point; thread
// r owns row r of T.
constexpr int kTB = 64;
__global__ __launch_bounds__(kTB) void eig_tblock_kernel(const eig_kdesc *ks, const cplx *G, const cplx *tau, int n,
                                                         int i0, int ib, size_t sgt, cplx *T) {
  const int k = blockIdx.x, r = threadIdx.x;
  const int last = ks[k].n - 1;   // reflectors i0 + i < last are the k-point's
  const cplx *g = G + (size_t)k * sgt;   // G and T of k-point k: leading dimension ib, at k * sgt
  cplx *t = T + (size_t)k * sgt;
  __shared__ double tre[kTB][kTB], tim[kTB][kTB];   // T(r, q) at [q][r]: row r is thread r's
  __shared__ cplx tmp[kTB];
  for (int q = 0; q < kTB; ++q) tre[q][r] = tim[q][r] = 0.0;
  for (int i = ib - 1; i >= 0; --i) {
    const cplx ti = i0 + i < last ? tau[(size_t)k * n + i0 + i] : czero();
    if (r > i && r < ib) {   // tmp(r) = -tau_i G(r, i)
      const cplx p = cmul(ti, g[r + ib * i]);
      tmp[r] = {-p.re, -p.im};
    }
    __syncthreads();
    if (r > i && r < ib) {   // T(r, i) = sum_{q = i+1}^{r} T(r, q) tmp(q)
      double are = 0.0, aim = 0.0;
      for (int q = i + 1; q <= r; ++q) {
        const double xr = tre[q][r], xi = tim[q][r], yr = tmp[q].re, yi = tmp[q].im;
        are += xr * yr - xi * yi;
        aim += xr * yi + xi * yr;
      }
      tre[i][r] = are;
      tim[i][r] = aim;
    }
    if (r == i) {
      tre[i][i] = ti.re;
      tim[i][i] = ti.im;
    }
    __syncthreads();
  }
  if (r < ib)
    for (int q = 0; q < ib; ++q) t[r + (size_t)ib * q] = {tre[q][r], tim[q][r]};
}

// Eigenpairs of a tridiagonal matrix from bisection and inverse iteration (dstebz with eorder
// 'blocks', then dstein), ordered by split-off block; *nfound >= nev of them (the bisection
// keeps a whole cluster of equal eigenvalues at the upper end of the index range).  Column j
// moves to its rank among the nfound eigenvalues (ties by index) if that is below nev:
// x(r, rank_j) = z(r, j) as complex for r < m, eval(rank_j) = w(j).  Every workgroup of
// column j counts the rank itself (nfound reads); the grid covers m columns.
__global__ __launch_bounds__(256) void eig_sort_r2c_kernel(const double *zr, int m, int nev, int ldr,
                                                           const int *nfound, const double *w, cplx *x, int ldx,
                                                           double *eval) {
  const int r = blockIdx.x * 256 + threadIdx.x, j = blockIdx.y;
  const int nf = min(*nfound, m);
  if (j >= nf) return;   // uniform over the workgroup
  __shared__ int part[256];
  const double wj = w[j];
  int cnt = 0;
  for (int i = threadIdx.x; i < nf; i += 256) cnt += (w[i] < wj) || (w[i] == wj && i < j);
  part[threadIdx.x] = cnt;
  __syncthreads();
  for (int h = 128; h > 0; h >>= 1) {
    if (threadIdx.x < h) part[threadIdx.x] += part[threadIdx.x + h];
    __syncthreads();
  }
  const int rank = part[0];
  if (rank >= nev) return;
  if (blockIdx.x == 0 && threadIdx.x == 0) eval[rank] = wj;
  if (r < m) x[r + (size_t)rank * ldx] = {zr[r + (size_t)j * ldr], 0.0};
}

// evec_k(r, j) = x(r, j) for r < nmat_k, j < nev (x: n x nev per k-point, leading dimension n)
__global__ __launch_bounds__(256) void eig_copyout_kernel(const eig_kdesc *ks, int n, int nev, const cplx *x) {
  const int r = blockIdx.x * 256 + threadIdx.x, j = blockIdx.y, k = blockIdx.z;
  const eig_kdesc kd = ks[k];
  if (r >= kd.n) return;
  kd.evec[r + (size_t)j * kd.ldv] = x[(size_t)k * n * nev + r + (size_t)j * n];
}

inline int tiles(int n) { return (n + kET - 1) / kET; }

}  // namespace

int exc_launch_eig_hermfill(cplx *h, int n, int ldh, hipStream_t s) {
  if (n <= 0) return EXC_OK;
  if (ldh < n || tiles(n) > 65535) return EXC_ERR_ARG;
  hipLaunchKernelGGL(herm_fill_kernel, dim3(tiles(n), tiles(n)), dim3(kEThreads), 0, s, h, n, ldh);
  EXC_CHECK(hipGetLastError());
  return EXC_OK;
}

int exc_launch_eig_pad(const eig_kdesc *ks, int nb, int n, cplx *ph, cplx *po, double sigma, hipStream_t s) {
  if (nb <= 0 || n <= 0) return EXC_OK;
  if (nb > 65535 || tiles(n) > 65535) return EXC_ERR_ARG;
  const dim3 grid(tiles(n), tiles(n), nb);
  hipLaunchKernelGGL(eig_pad_kernel, grid, dim3(kEThreads), 0, s, ks, n, ph, 0, sigma);
  EXC_CHECK(hipGetLastError());
  hipLaunchKernelGGL(eig_pad_kernel, grid, dim3(kEThreads), 0, s, ks, n, po, 1, sigma);
  EXC_CHECK(hipGetLastError());
  return EXC_OK;
}

int exc_launch_eig_r2c(const double *zr, int m, int ncols, int ldr, cplx *zc, int ldc, hipStream_t s) {
  if (m <= 0 || ncols <= 0) return EXC_OK;
  if (ncols > 65535 || ldr < m || ldc < m) return EXC_ERR_ARG;
  hipLaunchKernelGGL(eig_r2c_kernel, dim3((m + 255) / 256, ncols), dim3(256), 0, s, zr, m, ldr, zc, ldc);
  EXC_CHECK(hipGetLastError());
  return EXC_OK;
}

int exc_launch_eig_vblock(const cplx *A, int nb, int n, int i0, int ib, size_t sv, cplx *V, hipStream_t s) {
  if (nb <= 0 || ib <= 0) return EXC_OK;
  if (sv < (size_t)n * ib) return EXC_ERR_ARG;
  if (nb > 65535 || ib > 64 || i0 < 0 || i0 + ib > n - 1) return EXC_ERR_ARG;   // reflector i in column i + 1
  hipLaunchKernelGGL(eig_vblock_kernel, dim3((i0 + ib + 255) / 256, ib, nb), dim3(256), 0, s, A, n, i0, ib, sv, V);
  EXC_CHECK(hipGetLastError());
  return EXC_OK;
}

int exc_launch_eig_tblock(const eig_kdesc *ks, const cplx *G, const cplx *tau, int nb, int n, int i0, int ib,
                          size_t sgt, cplx *T, hipStream_t s) {
  if (nb <= 0 || ib <= 0) return EXC_OK;
  if (ib > kTB || i0 + ib > n || sgt < (size_t)ib * ib) return EXC_ERR_ARG;
  hipLaunchKernelGGL(eig_tblock_kernel, dim3(nb), dim3(kTB), 0, s, ks, G, tau, n, i0, ib, sgt, T);
  EXC_CHECK(hipGetLastError());
  return EXC_OK;
}

int exc_launch_eig_sort_r2c(const double *zr, int m, int nev, int ldr, const int *nfound, const double *w, cplx *x,
                            int ldx, double *eval, hipStream_t s) {
  if (m <= 0 || nev <= 0) return EXC_OK;
  if (m > 65535 || nev > m) return EXC_ERR_ARG;
  hipLaunchKernelGGL(eig_sort_r2c_kernel, dim3((m + 255) / 256, m), dim3(256), 0, s, zr, m, nev, ldr, nfound, w, x,
                     ldx, eval);
  EXC_CHECK(hipGetLastError());
  return EXC_OK;
}

int exc_launch_eig_copyout(const eig_kdesc *ks, int nb, int n, int nev, const cplx *x, hipStream_t s) {
  if (nb <= 0 || nev <= 0) return EXC_OK;
  if (nb > 65535 || nev > 65535) return EXC_ERR_ARG;
  hipLaunchKernelGGL(eig_copyout_kernel, dim3((n + 255) / 256, nev, nb), dim3(256), 0, s, ks, n, nev, x);
  EXC_CHECK(hipGetLastError());
  return EXC_OK;
}
