// C-ABI entry points of the generalized eigensolve after the setup (exc_eigen, exc_eigen_dev,
// exc_eigen_batched_dev): the counterpart of solvewithlapack (modfvsystem.f90:585-720).
#include "exc_ctx.h"

using namespace exc_impl;

extern "C" {

namespace {

bool eig_args_ok(int nmat, int ldh, int nev, int ldv) {
  return nmat >= 0 && ldh >= (nmat > 1 ? nmat : 1) && nev >= 0 && nev <= nmat && ldv >= (nmat > 1 ? nmat : 1);
}

// Gives a worker's rocBLAS handle a workspace of its own, sized by a rocBLAS size query over
// the calls `calls` makes (recorded, not run): no workspace is then allocated or freed by the
// library while the handle's work is queued.  In its default mode rocBLAS grows a handle's
// workspace at enqueue time by freeing the old buffer and allocating a larger one, with
// earlier work of the handle still queued on its stream; with several worker streams another
// handle's allocation could then take over that memory.  The buffer here only grows through
// grow(), which drains the device before it frees.
extern "C++" template <class F>
int pin_workspace(exc_ctx::EigWorker &W, F &&calls) {
  if (rocblas_start_device_memory_size_query(W.blas) != rocblas_status_success) return EXC_ERR_HIP;
  calls();
  size_t need = 0;
  if (rocblas_stop_device_memory_size_query(W.blas, &need) != rocblas_status_success) return EXC_ERR_HIP;
  need = std::max(need, (size_t)256);
  const bool moved = need > W.wsb;
  int rc = grow(&W.ws, &W.wsb, need);
  if (rc) return rc;
  if ((moved || !W.ws_set) && rocblas_set_workspace(W.blas, W.ws, W.wsb) != rocblas_status_success) return EXC_ERR_HIP;
  W.ws_set = true;
  return EXC_OK;
}

// Enqueue one generalized eigensolve and the copies of its first nev eigenpairs on the
// worker's stream; d_ints receives (nev found, info).  Solver variants (all: Cholesky of O,
// reduction to standard form, then)
//  0 zhegvdx: tridiagonal divide and conquer, lowest nev     1 zhegvx: bisection + inverse iteration
//  2 zhegvd: divide and conquer, all                         3 zhegv: implicit QL/QR, all
//  4 (default) the reduction C = U^-H H U^-1 with level-3 operations (potrf, trtri, two trmm:
//    rocSOLVER's zhegst runs column by column), zheevd on C, and X = U^-1 Y for the nev
//    wanted columns only
// (rocSOLVER's Jacobi solver, zhegvj, misses the 1e-10 Ha eigenvalue bar on these spectra)
// d_ints: nev found, info of the solver / of potrf, info of trtri, info of heevd
// Variant 4 runs in two phases: phase 1 factors O (the caller checks its info on the host
// before phase 2, so a non-positive-definite O never feeds the later steps); the other
// variants do all their work in phase 2.
// a call that ran, or (in a size query) recorded its workspace
inline bool call_ok(rocblas_status st) {
  return st == rocblas_status_success || st == rocblas_status_size_unchanged || st == rocblas_status_size_increased;
}

// The rocBLAS / rocSOLVER calls of one solve (phase 1: the Cholesky factor of variant 4;
// phase 2: the rest), used both to run them and, in a size query, to size the handle's
// workspace; in_a: eigenvectors returned in A (all n columns) rather than in Z.
rocblas_status eig_calls(exc_ctx::EigWorker &W, int v, int phase, int nmat, int ldh, double *h, double *o, int nev,
                         int *d_ints, bool *in_a) {
  const auto itype = rocblas_eform_ax;
  const auto ev = rocblas_evect_original;
  const auto up = rocblas_fill_upper;
  rocblas_double_complex *A = (rocblas_double_complex *)h, *B = (rocblas_double_complex *)o;
  rocblas_double_complex *Z = (rocblas_double_complex *)W.z;
  *in_a = false;
  if (phase == 1) return v < 4 ? rocblas_status_success : rocsolver_zpotrf(W.blas, up, nmat, B, ldh, d_ints + 1);
  switch (v) {
    case 0:
      return rocsolver_zhegvdx(W.blas, itype, ev, rocblas_erange_index, up, nmat, A, ldh, B, ldh, 0.0, 0.0, 1, nev,
                               d_ints, W.w, Z, nmat, d_ints + 1);
    case 1:
      return rocsolver_zhegvx(W.blas, itype, ev, rocblas_erange_index, up, nmat, A, ldh, B, ldh, 0.0, 0.0, 1, nev, 0.0,
                              d_ints, W.w, Z, nmat, (rocblas_int *)W.e, d_ints + 1);
    case 2:
      *in_a = true;
      return rocsolver_zhegvd(W.blas, itype, ev, up, nmat, A, ldh, B, ldh, W.w, W.e, d_ints + 1);
    case 3:
      *in_a = true;
      return rocsolver_zhegv(W.blas, itype, ev, up, nmat, A, ldh, B, ldh, W.w, W.e, d_ints + 1);
    default: {
      const rocblas_double_complex one = {1.0, 0.0};
      rocblas_status st = rocsolver_ztrtri(W.blas, up, rocblas_diagonal_non_unit, nmat, B, ldh, d_ints + 2);   // B = U^-1
      if (call_ok(st))                                                                  // Z = H U^-1
        st = rocblas_ztrmm(W.blas, rocblas_side_right, up, rocblas_operation_none, rocblas_diagonal_non_unit, nmat,
                           nmat, &one, B, ldh, A, ldh, Z, nmat);
      if (call_ok(st))                                                                  // H = U^-H Z
        st = rocblas_ztrmm(W.blas, rocblas_side_left, up, rocblas_operation_conjugate_transpose,
                           rocblas_diagonal_non_unit, nmat, nmat, &one, B, ldh, Z, nmat, A, ldh);
      if (call_ok(st))                                                                  // C Y = Y diag(w)
        st = rocsolver_zheevd(W.blas, ev, up, nmat, A, ldh, W.w, W.e, d_ints + 3);
      if (call_ok(st))                                                                  // Z = U^-1 Y(:, 1:nev)
        st = rocblas_ztrmm(W.blas, rocblas_side_left, up, rocblas_operation_none, rocblas_diagonal_non_unit, nmat,
                           nev, &one, B, ldh, A, ldh, Z, nmat);
      return st;
    }
  }
}

int eig_enqueue(exc_ctx::EigWorker &W, int v, int phase, int nmat, int ldh, double *h, double *o, int nev,
                double *eval, double *evec, int ldv, int *d_ints) {
  if (!W.blas) {
    if (rocblas_create_handle(&W.blas) != rocblas_status_success) {
      W.blas = nullptr;
      return EXC_ERR_HIP;
    }
  }
  if (rocblas_set_stream(W.blas, W.s) != rocblas_status_success) return EXC_ERR_HIP;
  // W has n slots and Z n columns: rocSOLVER's divide-and-conquer path uses the columns past
  // nev as workspace, so the caller's nev columns are filled from an internal n x n buffer
  int rc;
  if ((rc = grow((void **)&W.w, &W.wb, sizeof(double) * nmat))) return rc;
  if ((rc = grow((void **)&W.z, &W.zb, sizeof(cplx) * (size_t)nmat * nmat))) return rc;
  if ((rc = grow((void **)&W.e, &W.eb, sizeof(double) * (nmat + 1)))) return rc;
  bool in_a = false;
  if (phase == 1) {
    // the workspace of both phases, before anything of this solve is queued
    if ((rc = pin_workspace(W, [&]() {
           eig_calls(W, v, 1, nmat, ldh, h, o, nev, d_ints, &in_a);
           eig_calls(W, v, 2, nmat, ldh, h, o, nev, d_ints, &in_a);
         })))
      return rc;
    // ZHEGVX('U') reads the upper triangles and the real parts of the diagonals only;
    // rocSOLVER's solvers and the products of variant 4 read more: lower := conj(upper)
    if ((rc = exc_launch_eig_hermfill((cplx *)h, nmat, ldh, W.s))) return rc;
    if ((rc = exc_launch_eig_hermfill((cplx *)o, nmat, ldh, W.s))) return rc;
    return call_ok(eig_calls(W, v, 1, nmat, ldh, h, o, nev, d_ints, &in_a)) ? EXC_OK : EXC_ERR_HIP;   // O = U^H U
  }
  if (!call_ok(eig_calls(W, v, 2, nmat, ldh, h, o, nev, d_ints, &in_a))) return EXC_ERR_HIP;
  EXC_CHECK(hipMemcpyAsync(eval, W.w, sizeof(double) * nev, hipMemcpyDeviceToDevice, W.s));
  EXC_CHECK(hipMemcpy2DAsync(evec, 16 * (size_t)ldv, in_a ? (const void *)h : (const void *)W.z,
                             16 * (size_t)(in_a ? ldh : nmat), 16 * (size_t)nmat, nev, hipMemcpyDeviceToDevice, W.s));
  return EXC_OK;
}

// status of a solve from its four ints (the full-spectrum solvers, v >= 2, report no count);
// a failed Cholesky of O is nmat + i as in LAPACK's zhegv
int eig_status(const int *hi, int nmat, int nev, int v) {
  if (v >= 4) {
    if (hi[1] > 0) return nmat + hi[1];
    if (hi[2] != 0) return EXC_ERR_HIP;
    return hi[3];
  }
  if (hi[1] != 0) return hi[1];
  return (v >= 2 || hi[0] == nev) ? EXC_OK : EXC_ERR_HIP;
}

// Variant 5 of the multi-k entry: every k-point's H, O padded to the largest nmat n and
// reduced as one strided batch -- Cholesky, U^-1, C = U^-H H U^-1 and the tridiagonal
// reduction of C, one launch per step for all k-points (the per-k-point reduction is
// thousands of small launches, one column at a time).  The pad stays out of the leading
// nmat_k block of every factor (k_eigen.hip).  Then per k-point, on the worker streams: the
// tridiagonal divide and conquer of the leading nmat_k part (real arithmetic), then its first
// nev eigenvectors through the Householder reflectors and U^-1 into the caller's evec:
//  variant 5: per k-point, zunmtr and trmm on the worker streams;
//  variant 6: for the whole batch, the reflectors applied kBT at a time as block reflectors
//    I - V T V^H (T from zlarft's recurrence, k_eigen.hip) with strided-batched GEMMs, then
//    one strided-batched trmm; the eigenvectors of every k-point sit in an n x nev slice
//    whose rows past nmat_k stay zero, so the pad's rows never mix in;
//  variant 7: as 6, with the tridiagonal eigenpairs of the lowest nev only, by bisection and
//    inverse iteration (dstebz + dstein, as ZHEGVX range 'I' with abstol = the reference's
//    default evaltol) instead of the full divide and conquer;
//  variant 8: as 6, the per-k divide and conquer of all k-points on one worker stream, one
//    after another (the serialized form of 6: the same calls on the same data).
// H and O are read, not modified.
constexpr int kBT = 64;            // reflectors per block of the batched back-transform
constexpr double kAbsTol = 1e-14;  // bisection tolerance (Ha): evaltol's default, groundstate.xsd:512-513
int eig_grouped(exc_ctx *c, int nk, const int *nmat, const int *ldh, double *const *h, double *const *o, int nev,
                double *const *eval, double *const *evec, const int *ldv, int v) {
  const bool blocked = v >= 6, bisect = v == 7;
  std::vector<eig_kdesc> kd;
  int n = 0;
  for (int k = 0; k < nk; ++k) {
    if (nmat[k] == 0) continue;
    kd.push_back({(const cplx *)h[k], (const cplx *)o[k], eval[k], (cplx *)evec[k], nmat[k], ldh[k], ldv[k], k});
    n = std::max(n, nmat[k]);
  }
  const int nb = (int)kd.size();
  if (nb == 0) return EXC_OK;
  if (nb > 65535) return EXC_ERR_ARG;
  // variant 8: variant 6 with every k-point's divide and conquer on one worker stream
  const int nw = v == 8 ? 1 : std::min(nb, (int)exc_ctx::kEigStreams);
  const size_t mat = (size_t)n * n, a256 = 255;
  auto al = [&](size_t b) { return (b + a256) & ~a256; };
  const size_t o_po = al(16 * mat * nb), o_pz = o_po + al(16 * mat * nb), o_d = o_pz + al(16 * mat * nb);
  const size_t o_e = o_d + al(8 * (size_t)n * nb);
  const size_t o_tau = o_e + al(8 * (size_t)n * nb), o_i = o_tau + al(16 * (size_t)n * nb);
  const size_t o_kd = o_i + al(12 * (size_t)nb);
  // variant 6: V (n x kBT), G and T (kBT x kBT), W1 and W2 (kBT x nev) per k-point
  const size_t o_v = o_kd + al(sizeof(eig_kdesc) * nb), o_g = o_v + (blocked ? al(16 * (size_t)n * kBT * nb) : 0);
  const size_t o_t = o_g + (blocked ? al(16 * (size_t)kBT * kBT * nb) : 0);
  const size_t o_w1 = o_t + (blocked ? al(16 * (size_t)kBT * kBT * nb) : 0);
  const size_t o_w2 = o_w1 + (blocked ? al(16 * (size_t)kBT * nev * nb) : 0);
  const size_t o_wv = o_w2 + (blocked ? al(16 * (size_t)kBT * nev * nb) : 0);
  // variant 7: eigenvalues, block and split indices, failure flags (n each) and 4 ints per k-point
  const size_t o_ib = o_wv + (bisect ? al(8 * (size_t)n * nb) : 0);
  const size_t o_is = o_ib + (bisect ? al(4 * (size_t)n * nb) : 0);
  const size_t o_if = o_is + (bisect ? al(4 * (size_t)n * nb) : 0);
  const size_t o_si = o_if + (bisect ? al(4 * (size_t)n * nb) : 0);
  const size_t total = o_si + (bisect ? al(16 * (size_t)nb) : 0);
  int rc;
  if ((rc = grow(&c->eg, &c->eg_bytes, total))) return rc;
  char *base = (char *)c->eg;
  cplx *ph = (cplx *)base, *po = (cplx *)(base + o_po), *pz = (cplx *)(base + o_pz), *tau = (cplx *)(base + o_tau);
  double *d = (double *)(base + o_d), *e = (double *)(base + o_e);
  int *info = (int *)(base + o_i);
  eig_kdesc *dkd = (eig_kdesc *)(base + o_kd);
  for (int i = 0; i <= nw; ++i) {
    exc_ctx::EigWorker &W = c->ew[i];
    if (i == 0) W.s = c->stream;
    if (!W.blas && rocblas_create_handle(&W.blas) != rocblas_status_success) {
      W.blas = nullptr;
      return EXC_ERR_HIP;
    }
    if (rocblas_set_stream(W.blas, W.s) != rocblas_status_success) return EXC_ERR_HIP;
    if (i > 0) {   // per worker: real eigenvectors of T (n x n), their first nev columns complex
      if ((rc = grow((void **)&W.z, &W.zb, sizeof(double) * mat))) return rc;
      if (!blocked && (rc = grow((void **)&W.zc, &W.zcb, sizeof(cplx) * (size_t)n * nev))) return rc;
    }
  }
  auto fail = [](const char *what, int a, int b) {
    fprintf(stderr, "exc_eigen_batched_dev (grouped): %s (%d, %d)\n", what, a, b);
    return EXC_ERR_HIP;
  };
  const auto up = rocblas_fill_upper;
  const auto nu = rocblas_diagonal_non_unit;
  const rocblas_double_complex one = {1.0, 0.0};
  rocblas_handle hb = c->ew[0].blas;
  auto *A = (rocblas_double_complex *)ph, *B = (rocblas_double_complex *)po, *Z = (rocblas_double_complex *)pz;
  const rocblas_stride sm = (rocblas_stride)mat;
  std::vector<int> hi(3 * (size_t)nb, 0);
  // every handle's workspace sized for its calls (the largest order n) before anything is queued
  {
    auto *V = (rocblas_double_complex *)(base + o_v), *G = (rocblas_double_complex *)(base + o_g);
    auto *T = (rocblas_double_complex *)(base + o_t), *W1 = (rocblas_double_complex *)(base + o_w1);
    auto *W2 = (rocblas_double_complex *)(base + o_w2);
    const rocblas_double_complex zero = {0.0, 0.0}, mone = {-1.0, 0.0};
    const auto N = rocblas_operation_none, CT = rocblas_operation_conjugate_transpose;
    const rocblas_stride sv = (rocblas_stride)n * kBT, st2 = (rocblas_stride)kBT * kBT;
    const rocblas_stride sw = (rocblas_stride)kBT * nev, sxx = (rocblas_stride)n * nev;
    const int ib = std::min(kBT, std::max(n - 1, 1));
    if ((rc = pin_workspace(c->ew[0], [&]() {
           rocsolver_zpotrf_strided_batched(hb, up, n, B, n, sm, info, nb);
           rocsolver_ztrtri_strided_batched(hb, up, nu, n, B, n, sm, info + nb, nb);
           rocblas_ztrmm_strided_batched(hb, rocblas_side_right, up, N, nu, n, n, &one, B, n, sm, A, n, sm, Z, n, sm, nb);
           rocblas_ztrmm_strided_batched(hb, rocblas_side_left, up, CT, nu, n, n, &one, B, n, sm, Z, n, sm, A, n, sm, nb);
           rocsolver_zhetrd_strided_batched(hb, up, n, A, n, sm, d, n, e, n, (rocblas_double_complex *)tau, n, nb);
           if (blocked) {
             rocblas_zgemm_strided_batched(hb, CT, N, ib, ib, n, &one, V, n, sv, V, n, sv, &zero, G, ib, st2, nb);
             rocblas_zgemm_strided_batched(hb, CT, N, ib, nev, n, &one, V, n, sv, Z, n, sxx, &zero, W1, ib, sw, nb);
             rocblas_zgemm_strided_batched(hb, N, N, ib, nev, ib, &one, T, ib, st2, W1, ib, sw, &zero, W2, ib, sw, nb);
             rocblas_zgemm_strided_batched(hb, N, N, n, nev, ib, &mone, V, n, sv, W2, ib, sw, &one, Z, n, sxx, nb);
             rocblas_ztrmm_strided_batched(hb, rocblas_side_left, up, N, nu, n, nev, &one, B, n, sm, Z, n, sxx, A, n, sxx,
                                           nb);
           }
         })))
      return rc;
    for (int i = 1; i <= nw; ++i) {
      exc_ctx::EigWorker &W = c->ew[i];
      if ((rc = pin_workspace(W, [&]() {
             if (bisect) {
               double *wv = (double *)(base + o_wv);
               int *ibk = (int *)(base + o_ib), *isp = (int *)(base + o_is), *ifl = (int *)(base + o_if);
               int *si = (int *)(base + o_si);
               rocsolver_dstebz(W.blas, rocblas_erange_index, rocblas_eorder_blocks, n, 0.0, 0.0, 1, nev, kAbsTol, d, e,
                                si, si + 1, wv, ibk, isp, si + 2);
               rocsolver_dstein(W.blas, n, d, e, si, wv, ibk, isp, (double *)W.z, n, ifl, si + 3);
               return;
             }
             rocsolver_dstedc(W.blas, rocblas_evect_tridiagonal, n, d, e, (double *)W.z, n, info + 2 * nb);
             if (!blocked) {
               rocsolver_zunmtr(W.blas, rocblas_side_left, up, N, n, nev, A, n, (rocblas_double_complex *)tau,
                                (rocblas_double_complex *)W.zc, n);
               rocblas_ztrmm(W.blas, rocblas_side_left, up, N, nu, n, nev, &one, B, n, (rocblas_double_complex *)W.zc, n,
                             (rocblas_double_complex *)W.zc, n);
             }
           })))
        return rc;
    }
  }
  StageTimer t(c, EXC_STAGE_EIGEN);
  EXC_CHECK(hipMemcpyAsync(dkd, kd.data(), sizeof(eig_kdesc) * nb, hipMemcpyHostToDevice, c->stream));
  if ((rc = exc_launch_eig_pad(dkd, nb, n, ph, po, 1.0, c->stream))) return rc;
  // O = U^H U for every k-point; checked before anything reads the factors
  rocblas_status st = rocsolver_zpotrf_strided_batched(hb, up, n, B, n, sm, info, nb);
  if (st != rocblas_status_success) return fail("zpotrf_strided_batched status", (int)st, n);
  EXC_CHECK(hipMemcpyAsync(hi.data(), info, sizeof(int) * nb, hipMemcpyDeviceToHost, c->stream));
  EXC_CHECK(hipStreamSynchronize(c->stream));
  for (int b = 0; b < nb; ++b)
    if (hi[b] > 0) return hi[b] <= kd[b].n ? kd[b].n + hi[b] : fail("pad not positive definite", b, hi[b]);
  st = rocsolver_ztrtri_strided_batched(hb, up, nu, n, B, n, sm, info + nb, nb);   // U^-1
  if (st == rocblas_status_success)   // Z = H U^-1
    st = rocblas_ztrmm_strided_batched(hb, rocblas_side_right, up, rocblas_operation_none, nu, n, n, &one, B, n, sm,
                                       A, n, sm, Z, n, sm, nb);
  if (st == rocblas_status_success)   // H = U^-H Z
    st = rocblas_ztrmm_strided_batched(hb, rocblas_side_left, up, rocblas_operation_conjugate_transpose, nu, n, n,
                                       &one, B, n, sm, Z, n, sm, A, n, sm, nb);
  if (st == rocblas_status_success)   // C = Q T Q^H
    st = rocsolver_zhetrd_strided_batched(hb, up, n, A, n, sm, d, n, e, n, (rocblas_double_complex *)tau, n, nb);
  if (st != rocblas_status_success) return fail("trtri / trmm / zhetrd strided_batched status", (int)st, n);
  // variant 6: the eigenvectors of T of k-point b in X = pz + b n nev (rows past nmat_k zero)
  const size_t sx = (size_t)n * nev;
  if (blocked) EXC_CHECK(hipMemsetAsync(pz, 0, sizeof(cplx) * sx * nb, c->stream));
  EXC_CHECK(hipEventRecord(c->eig_fork, c->stream));
  for (int i = 1; i <= nw; ++i) EXC_CHECK(hipStreamWaitEvent(c->ew[i].s, c->eig_fork, 0));
  // per k-point on the workers; every worker is joined below, also after an error
  int err = EXC_OK;
  for (int b = 0; b < nb && !err; ++b) {
    exc_ctx::EigWorker &W = c->ew[1 + b % nw];
    const int m = kd[b].n;
    double *db = d + (size_t)b * n, *eb = e + (size_t)b * n;
    if (bisect) {   // lowest nev of T_k: bisection, inverse iteration, then ordered by value
      double *wv = (double *)(base + o_wv) + (size_t)b * n;
      int *ibk = (int *)(base + o_ib) + (size_t)b * n, *isp = (int *)(base + o_is) + (size_t)b * n;
      int *ifl = (int *)(base + o_if) + (size_t)b * n, *si = (int *)(base + o_si) + 4 * b;
      st = rocsolver_dstebz(W.blas, rocblas_erange_index, rocblas_eorder_blocks, m, 0.0, 0.0, 1, nev, kAbsTol, db, eb,
                            si, si + 1, wv, ibk, isp, si + 2);
      if (st == rocblas_status_success)
        st = rocsolver_dstein(W.blas, m, db, eb, si, wv, ibk, isp, (double *)W.z, m, ifl, si + 3);
      if (st != rocblas_status_success) {
        err = fail("dstebz / dstein status", (int)st, b);
        break;
      }
      if ((err = exc_launch_eig_sort_r2c((const double *)W.z, m, nev, m, si, wv, pz + (size_t)b * sx, n,
                                         kd[b].eval, W.s)))
        break;
      continue;
    }
    st = rocsolver_dstedc(W.blas, rocblas_evect_tridiagonal, m, db, eb, (double *)W.z, m, info + 2 * nb + b);
    if (st != rocblas_status_success) {
      err = fail("dstedc status", (int)st, b);
      break;
    }
    if (blocked) {
      if ((err = exc_launch_eig_r2c((const double *)W.z, m, nev, m, pz + (size_t)b * sx, n, W.s))) break;
      if (hipMemcpyAsync(kd[b].eval, db, sizeof(double) * nev, hipMemcpyDeviceToDevice, W.s) != hipSuccess)
        err = EXC_ERR_HIP;
      continue;
    }
    if ((err = exc_launch_eig_r2c((const double *)W.z, m, nev, m, W.zc, m, W.s))) break;
    st = rocsolver_zunmtr(W.blas, rocblas_side_left, up, rocblas_operation_none, m, nev, A + (size_t)b * mat, n,
                          (rocblas_double_complex *)(tau + (size_t)b * n), (rocblas_double_complex *)W.zc, m);
    if (st == rocblas_status_success)
      st = rocblas_ztrmm(W.blas, rocblas_side_left, up, rocblas_operation_none, nu, m, nev, &one, B + (size_t)b * mat,
                         n, (rocblas_double_complex *)W.zc, m, (rocblas_double_complex *)kd[b].evec, kd[b].ldv);
    if (st != rocblas_status_success) {
      err = fail("zunmtr / trmm status", (int)st, b);
      break;
    }
    if (hipMemcpyAsync(kd[b].eval, db, sizeof(double) * nev, hipMemcpyDeviceToDevice, W.s) != hipSuccess)
      err = EXC_ERR_HIP;
  }
  for (int i = 1; i <= nw; ++i) {
    EXC_CHECK(hipEventRecord(c->ew[i].done, c->ew[i].s));
    EXC_CHECK(hipStreamWaitEvent(c->stream, c->ew[i].done, 0));
  }
  if (err) return err;
  if (blocked) {
    // X := Q X = H(n-2) ... H(0) X, block [i0, i0 + ib) of reflectors at a time, block 0 first
    auto *V = (rocblas_double_complex *)(base + o_v), *G = (rocblas_double_complex *)(base + o_g);
    auto *T = (rocblas_double_complex *)(base + o_t), *W1 = (rocblas_double_complex *)(base + o_w1);
    auto *W2 = (rocblas_double_complex *)(base + o_w2);
    const rocblas_double_complex zero = {0.0, 0.0}, mone = {-1.0, 0.0};
    const auto N = rocblas_operation_none, CT = rocblas_operation_conjugate_transpose;
    const rocblas_stride sv = (rocblas_stride)n * kBT, st2 = (rocblas_stride)kBT * kBT;
    const rocblas_stride sw = (rocblas_stride)kBT * nev, sxx = (rocblas_stride)sx;
    for (int i0 = 0; i0 < n - 1 && st == rocblas_status_success; i0 += kBT) {
      const int ib = std::min(kBT, n - 1 - i0), mb = i0 + ib;   // rows 0 .. mb-1 of the block's vectors
      if ((rc = exc_launch_eig_vblock(ph, nb, n, i0, ib, sv, (cplx *)V, c->stream))) return rc;
      st = rocblas_zgemm_strided_batched(hb, CT, N, ib, ib, mb, &one, V, n, sv, V, n, sv, &zero, G, ib, st2, nb);
      if (st != rocblas_status_success) break;
      if ((rc = exc_launch_eig_tblock(dkd, (const cplx *)G, tau, nb, n, i0, ib, st2, (cplx *)T, c->stream))) return rc;
      st = rocblas_zgemm_strided_batched(hb, CT, N, ib, nev, mb, &one, V, n, sv, Z, n, sxx, &zero, W1, ib, sw, nb);
      if (st == rocblas_status_success)
        st = rocblas_zgemm_strided_batched(hb, N, N, ib, nev, ib, &one, T, ib, st2, W1, ib, sw, &zero, W2, ib, sw, nb);
      if (st == rocblas_status_success)
        st = rocblas_zgemm_strided_batched(hb, N, N, mb, nev, ib, &mone, V, n, sv, W2, ib, sw, &one, Z, n, sxx, nb);
    }
    if (st == rocblas_status_success)   // U^-1 X into ph (its reflectors are no longer needed)
      st = rocblas_ztrmm_strided_batched(hb, rocblas_side_left, up, rocblas_operation_none, nu, n, nev, &one, B, n, sm,
                                         Z, n, sxx, A, n, sxx, nb);
    if (st != rocblas_status_success) return fail("blocked back-transform status", (int)st, n);
    if ((rc = exc_launch_eig_copyout(dkd, nb, n, nev, ph, c->stream))) return rc;
  }
  EXC_CHECK(hipMemcpyAsync(hi.data(), info, sizeof(int) * 3 * nb, hipMemcpyDeviceToHost, c->stream));
  EXC_CHECK(hipStreamSynchronize(c->stream));
  for (int b = 0; b < nb; ++b) {
    if (hi[nb + b] != 0) return fail("ztrtri info", b, hi[nb + b]);   // of a positive definite factor
    if (!bisect && hi[2 * nb + b] != 0) return hi[2 * nb + b];      // eigenvalues that did not converge
  }
  if (bisect) {   // per k-point: eigenvalues found, split blocks, dstebz info, dstein info
    std::vector<int> si(4 * (size_t)nb);
    EXC_CHECK(hipMemcpy(si.data(), base + o_si, sizeof(int) * 4 * nb, hipMemcpyDeviceToHost));
    for (int b = 0; b < nb; ++b) {
      // at least nev (a cluster of equal eigenvalues at the upper end comes whole)
      if (si[4 * b] < nev || si[4 * b + 2] != 0) return fail("dstebz count / info", si[4 * b], si[4 * b + 2]);
      if (si[4 * b + 3] != 0) return si[4 * b + 3];   // eigenvectors that did not converge (zhegvx info)
    }
  }
  return EXC_OK;
}

}  // namespace

int exc_eigen_dev(exc_ctx *c, int nmat, int ldh, double *h, double *o, int nev, double *eval, double *evec,
                  int ldv) {
  if (!ready(c)) return c ? EXC_ERR_STATE : EXC_ERR_ARG;
  if (!eig_args_ok(nmat, ldh, nev, ldv)) return EXC_ERR_ARG;
  if (nmat == 0 || nev == 0) return EXC_OK;
  if (!h || !o || !eval || !evec) return EXC_ERR_ARG;
  int rc = use_device(c);
  if (rc) return rc;
  if ((rc = grow((void **)&c->eig_int, &c->eig_int_bytes, 4 * sizeof(int)))) return rc;
  const int v = c->eig_variant >= 0 ? c->eig_variant : 4;
  exc_ctx::EigWorker &W = c->ew[0];
  W.s = c->stream;
  int hi[4] = {0, 0, 0, 0};
  {
    StageTimer t(c, EXC_STAGE_EIGEN);
    if ((rc = eig_enqueue(W, v, 1, nmat, ldh, h, o, nev, eval, evec, ldv, c->eig_int))) return rc;
    if (v >= 4) {
      EXC_CHECK(hipMemcpyAsync(hi, c->eig_int, sizeof(hi), hipMemcpyDeviceToHost, c->stream));
      EXC_CHECK(hipStreamSynchronize(c->stream));
      if (hi[1] > 0) return nmat + hi[1];
    }
    if ((rc = eig_enqueue(W, v, 2, nmat, ldh, h, o, nev, eval, evec, ldv, c->eig_int))) return rc;
  }
  EXC_CHECK(hipMemcpyAsync(hi, c->eig_int, sizeof(hi), hipMemcpyDeviceToHost, c->stream));
  EXC_CHECK(hipStreamSynchronize(c->stream));
  return eig_status(hi, nmat, nev, v);
}

int exc_eigen_batched_dev(exc_ctx *c, int nk, const int *nmat, const int *ldh, double *const *h, double *const *o,
                          int nev, double *const *eval, double *const *evec, const int *ldv) {
  if (!ready(c)) return c ? EXC_ERR_STATE : EXC_ERR_ARG;
  if (nk < 0 || (nk > 0 && (!nmat || !ldh || !h || !o || !eval || !evec || !ldv))) return EXC_ERR_ARG;
  for (int k = 0; k < nk; ++k) {
    if (!eig_args_ok(nmat[k], ldh[k], nev, ldv[k])) return EXC_ERR_ARG;
    if (nmat[k] > 0 && nev > 0 && (!h[k] || !o[k] || !eval[k] || !evec[k])) return EXC_ERR_ARG;
  }
  if (nk == 0 || nev == 0) return EXC_OK;
  int rc = use_device(c);
  if (rc) return rc;
  if ((rc = grow((void **)&c->eig_int, &c->eig_int_bytes, 4 * sizeof(int) * nk))) return rc;
  if (!c->eig_fork) EXC_CHECK(hipEventCreateWithFlags(&c->eig_fork, hipEventDisableTiming));
  const int nw = nk < exc_ctx::kEigStreams ? nk : exc_ctx::kEigStreams;
  for (int i = 1; i <= nw; ++i) {
    exc_ctx::EigWorker &W = c->ew[i];
    if (!W.s) EXC_CHECK(hipStreamCreateWithFlags(&W.s, hipStreamNonBlocking));   // default priority
    if (!W.done) EXC_CHECK(hipEventCreateWithFlags(&W.done, hipEventDisableTiming));
  }
  const int v = c->eig_variant >= 0 ? c->eig_variant : 6;
  if (v >= 5) return eig_grouped(c, nk, nmat, ldh, h, o, nev, eval, evec, ldv, v);
  // Variants 0-4 solve the k-points one after another on one worker stream: run
  // concurrently on the 4 worker streams, rocSOLVER's potrf / trtri / heevd chain returned one
  // k-point's eigenpairs wrong (ghost eigenvalues off by 0.1-0.3 Ha) in about one test run of
  // four, with no shared buffer or missing wait in this code.  The worker's buffers are sized
  // for the largest nmat before anything is queued, so eig_enqueue never reallocates a buffer
  // that the queued solve of an earlier k-point still uses.
  {
    int nmax = 1;
    for (int k = 0; k < nk; ++k) nmax = std::max(nmax, nmat[k]);
    for (int i = 1; i <= nw; ++i) {
      exc_ctx::EigWorker &W = c->ew[i];
      if ((rc = grow((void **)&W.w, &W.wb, sizeof(double) * nmax))) return rc;
      if ((rc = grow((void **)&W.z, &W.zb, sizeof(cplx) * (size_t)nmax * nmax))) return rc;
      if ((rc = grow((void **)&W.e, &W.eb, sizeof(double) * (nmax + 1)))) return rc;
    }
  }
  {
    StageTimer t(c, EXC_STAGE_EIGEN);
    EXC_CHECK(hipEventRecord(c->eig_fork, c->stream));
    for (int i = 1; i <= nw; ++i) EXC_CHECK(hipStreamWaitEvent(c->ew[i].s, c->eig_fork, 0));
    for (int phase = 1; phase <= 2; ++phase) {
      if (phase == 2 && v >= 4) {
        // every factorization done and positive definite before the rest runs
        for (int i = 1; i <= nw; ++i) EXC_CHECK(hipStreamSynchronize(c->ew[i].s));
        std::vector<int> hp(4 * (size_t)nk, 0);
        EXC_CHECK(hipMemcpy(hp.data(), c->eig_int, sizeof(int) * 4 * nk, hipMemcpyDeviceToHost));
        for (int k = 0; k < nk; ++k)
          if (nmat[k] > 0 && hp[4 * (size_t)k + 1] > 0) return nmat[k] + hp[4 * (size_t)k + 1];
      }
      for (int k = 0; k < nk && !rc; ++k) {
        if (nmat[k] == 0) continue;
        exc_ctx::EigWorker &W = c->ew[1];   // one after another (see below)
        rc = eig_enqueue(W, v, phase, nmat[k], ldh[k], h[k], o[k], nev, eval[k], evec[k], ldv[k], c->eig_int + 4 * k);
      }
      if (rc) break;
    }
    for (int i = 1; i <= nw; ++i) {   // joined also after an error
      EXC_CHECK(hipEventRecord(c->ew[i].done, c->ew[i].s));
      EXC_CHECK(hipStreamWaitEvent(c->stream, c->ew[i].done, 0));
    }
    if (rc) return rc;
  }
  std::vector<int> hi(4 * (size_t)nk, 0);
  EXC_CHECK(hipMemcpyAsync(hi.data(), c->eig_int, sizeof(int) * 4 * nk, hipMemcpyDeviceToHost, c->stream));
  EXC_CHECK(hipStreamSynchronize(c->stream));
  for (int k = 0; k < nk; ++k)
    if (nmat[k] > 0 && (rc = eig_status(&hi[4 * (size_t)k], nmat[k], nev, v))) return rc;
  return EXC_OK;
}


int exc_eigen(exc_ctx *c, int nmat, int ldh, double *h, double *o, int nev, double *eval, double *evec, int ldv) {
  if (!ready(c)) return c ? EXC_ERR_STATE : EXC_ERR_ARG;
  if (nmat < 0 || ldh < (nmat > 1 ? nmat : 1) || nev < 0 || nev > nmat || ldv < (nmat > 1 ? nmat : 1)) return EXC_ERR_ARG;
  if (nmat == 0 || nev == 0) return EXC_OK;
  if (!h || !o || !eval || !evec) return EXC_ERR_ARG;
  int rc = use_device(c);
  if (rc) return rc;
  const size_t b_m = 16 * (size_t)ldh * nmat, b_v = 16 * (size_t)ldv * nev, b_e = 8 * (size_t)nev;
  const size_t o_o = (b_m + 255) & ~(size_t)255, o_v = o_o + ((b_m + 255) & ~(size_t)255);
  const size_t o_e = o_v + ((b_v + 255) & ~(size_t)255);
  if ((rc = grow(&c->stage, &c->stage_bytes, o_e + b_e))) return rc;
  if ((rc = grow_host(&c->hstage, &c->hstage_bytes, std::max(2 * b_m, b_v)))) return rc;
  char *base = (char *)c->stage, *hs = (char *)c->hstage;
  const size_t pitch = 16 * (size_t)ldh, width = 16 * (size_t)nmat;
  copy_cols(hs, pitch, (const char *)h, pitch, width, nmat);
  copy_cols(hs + b_m, pitch, (const char *)o, pitch, width, nmat);
  EXC_CHECK(hipMemcpyAsync(base, hs, b_m, hipMemcpyHostToDevice, c->stream));
  EXC_CHECK(hipMemcpyAsync(base + o_o, hs + b_m, b_m, hipMemcpyHostToDevice, c->stream));
  rc = exc_eigen_dev(c, nmat, ldh, (double *)base, (double *)(base + o_o), nev, (double *)(base + o_e),
                     (double *)(base + o_v), ldv);
  if (rc) return rc;
  // H and O come back overwritten as ZHEGVX leaves them (O: its Cholesky factor)
  EXC_CHECK(hipMemcpyAsync(hs, base, b_m, hipMemcpyDeviceToHost, c->stream));
  EXC_CHECK(hipMemcpyAsync(hs + b_m, base + o_o, b_m, hipMemcpyDeviceToHost, c->stream));
  EXC_CHECK(hipStreamSynchronize(c->stream));
  copy_cols((char *)h, pitch, hs, pitch, width, nmat);
  copy_cols((char *)o, pitch, hs + b_m, pitch, width, nmat);
  EXC_CHECK(hipMemcpyAsync(hs, base + o_v, b_v, hipMemcpyDeviceToHost, c->stream));
  EXC_CHECK(hipMemcpyAsync(eval, base + o_e, b_e, hipMemcpyDeviceToHost, c->stream));
  EXC_CHECK(hipStreamSynchronize(c->stream));
  copy_cols((char *)evec, 16 * (size_t)ldv, hs, 16 * (size_t)ldv, width, nev);
  return EXC_OK;
}

}  // extern "C"
