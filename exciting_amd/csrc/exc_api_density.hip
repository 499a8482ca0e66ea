// C-ABI entry points of the muffin-tin density matrix (exc_gendmatmt*, exc_dmat_wfsize): the
// counterpart of gendmatmt (gendmatmt.f90:13).
#include "exc_ctx.h"

using namespace exc_impl;

extern "C" {

namespace {

// dm(ws, ws, natm) += sum over the nk k-points of w_k wf1 diag(occ) wf1^H (gendmatmt.f90:59-172).
// The k-points are split into contiguous ranges, one per worker stream (the products of one
// k-point fill only about half of the GPU: nst columns).  Per k: one product A^T C for all
// atoms (rows in apwalm's storage order), packed into the reference's row order with the lo
// rows appended (times evecsv when given) and the occupation-weighted copy.  Per worker: one
// product per atom whose contraction runs over the columns of its k-points, into a partial
// sum; the partials are added into dm on the context's stream.  All arrays are device memory.
int dmat_run(exc_ctx *c, int nk, const int *ngp, const int *ldg, const double *const *apwalm, int nst,
             const int *ldv, const double *const *evec, const double *const *evecsv, const double *const *occ,
             const double *wkpt, double *dm) {
  const exc_dev &d = c->d;
  const int ws = d.naa + d.dm_maxlo, na = d.natm, naa = d.naa;
  const size_t nt = (size_t)nk * nst;
  if (nt > (size_t)0x7fffffff) return EXC_ERR_ARG;
  const int nw = std::min(nk, (int)exc_ctx::kEigStreams);
  bool any_sv = false;
  for (int k = 0; k < nk; ++k) any_sv = any_sv || (evecsv && evecsv[k]);
  auto al = [](size_t b) { return (b + 255) & ~(size_t)255; };
  const size_t b_w = al(16 * (size_t)na * naa * nst), b_x = al(16 * (size_t)na * ws * nt);
  const size_t b_t = any_sv ? al(16 * (size_t)na * ws * nst) : 0, b_p = al(16 * (size_t)na * ws * ws);
  int rc;
  if ((rc = grow(&c->dmw, &c->dmw_bytes, nw * (b_w + b_t + b_p) + 2 * b_x))) return rc;
  char *base = (char *)c->dmw;
  cplx *x = (cplx *)base, *x2 = (cplx *)(base + b_x);
  char *per = base + 2 * b_x;   // per worker: W, Xt, partial
  if (!c->eig_fork) EXC_CHECK(hipEventCreateWithFlags(&c->eig_fork, hipEventDisableTiming));
  for (int i = 1; i <= nw; ++i) {
    exc_ctx::EigWorker &W = c->ew[i];
    if (!W.s) EXC_CHECK(hipStreamCreateWithFlags(&W.s, hipStreamNonBlocking));   // default priority
    if (!W.done) EXC_CHECK(hipEventCreateWithFlags(&W.done, hipEventDisableTiming));
    if (!W.blas && rocblas_create_handle(&W.blas) != rocblas_status_success) {
      W.blas = nullptr;
      return EXC_ERR_HIP;
    }
    if (rocblas_set_stream(W.blas, W.s) != rocblas_status_success) return EXC_ERR_HIP;
  }
  const rocblas_double_complex one = {1.0, 0.0}, zero = {0.0, 0.0};
  const size_t xs = (size_t)ws * nt;
  StageTimer t(c, EXC_STAGE_DMAT);
  EXC_CHECK(hipEventRecord(c->eig_fork, c->stream));
  // a worker's enqueue; every forked worker is joined below, also after an error
  auto work = [&](exc_ctx::EigWorker &W, int i) -> int {
    cplx *w = (cplx *)(per + i * (b_w + b_t + b_p)), *xt = (cplx *)((char *)w + b_w);
    cplx *part = (cplx *)((char *)w + b_w + b_t);
    const int k0 = (int)((long)nk * i / nw), k1 = (int)((long)nk * (i + 1) / nw);
    int rc;
    for (int k = k0; k < k1; ++k) {
      // W = A^T C_k for all atoms at once (the atoms' columns of apwalm are contiguous):
      // (natm naa) x nst, row a naa + io + apwordmax lm
      if (rocblas_zgemm(W.blas, rocblas_operation_transpose, rocblas_operation_none, na * naa, nst, ngp[k], &one,
                        (const rocblas_double_complex *)apwalm[k], ldg[k], (const rocblas_double_complex *)evec[k],
                        ldv[k], &zero, (rocblas_double_complex *)w, na * naa) != rocblas_status_success)
        return EXC_ERR_HIP;
      if (evecsv && evecsv[k]) {
        if ((rc = exc_launch_dmat_pack(d, ws, nst, w, (const cplx *)evec[k], ldv[k], ngp[k], xt, ws,
                                       (size_t)ws * nst, 0, W.s)))
          return rc;
        if (rocblas_zgemm_strided_batched(W.blas, rocblas_operation_none, rocblas_operation_none, ws, nst, nst, &one,
                                          (const rocblas_double_complex *)xt, ws, (rocblas_stride)ws * nst,
                                          (const rocblas_double_complex *)evecsv[k], nst, 0, &zero,
                                          (rocblas_double_complex *)(x + (size_t)k * nst * ws), ws,
                                          (rocblas_stride)xs, na) != rocblas_status_success)
          return EXC_ERR_HIP;
      } else if ((rc = exc_launch_dmat_pack(d, ws, nst, w, (const cplx *)evec[k], ldv[k], ngp[k], x, ws, xs,
                                            k * nst, W.s))) {
        return rc;
      }
      if ((rc = exc_launch_dmat_scale(d, ws, nst, x, x2, ws, xs, k * nst, occ[k], wkpt[k], W.s))) return rc;
    }
    // part_a = X_a X2_a^H over this worker's columns
    const size_t c0 = (size_t)k0 * nst * ws;
    if (rocblas_zgemm_strided_batched(W.blas, rocblas_operation_none, rocblas_operation_conjugate_transpose, ws, ws,
                                      (k1 - k0) * nst, &one, (const rocblas_double_complex *)(x + c0), ws,
                                      (rocblas_stride)xs, (const rocblas_double_complex *)(x2 + c0), ws,
                                      (rocblas_stride)xs, &zero, (rocblas_double_complex *)part, ws,
                                      (rocblas_stride)ws * ws, na) != rocblas_status_success)
      return EXC_ERR_HIP;
    return EXC_OK;
  };
  int err = EXC_OK;
  for (int i = 0; i < nw; ++i) {
    exc_ctx::EigWorker &W = c->ew[1 + i];
    EXC_CHECK(hipStreamWaitEvent(W.s, c->eig_fork, 0));
    if (!err) err = work(W, i);
    EXC_CHECK(hipEventRecord(W.done, W.s));
    EXC_CHECK(hipStreamWaitEvent(c->stream, W.done, 0));
  }
  if (err) return err;
  return exc_launch_dmat_reduce((cplx *)dm, (const cplx *)per, nw, (b_w + b_t + b_p) / sizeof(cplx),
                                (b_w + b_t) / sizeof(cplx), (size_t)na * ws * ws, c->stream);
}

bool dmat_args_ok(const exc_ctx *c, int ngp, int ngkmax, int nst, int ldv) {
  return ngp >= 0 && ngkmax >= (ngp > 1 ? ngp : 1) && nst >= 0 && nst <= 65535 && ldv >= ngp + c->d.nlotot &&
         ldv >= 1;
}

}  // namespace

int exc_dmat_wfsize(const exc_ctx *c) {
  if (!c) return EXC_ERR_ARG;
  if (!c->have_lo) return EXC_ERR_STATE;
  return c->d.naa + c->d.dm_maxlo;
}

int exc_gendmatmt_dev(exc_ctx *c, int ngp, int ngkmax, const double *apwalm, int nst, int ldv, const double *evecfv,
                      const double *evecsv, const double *occ, double wkpt, double *dm) {
  if (!ready(c)) return c ? EXC_ERR_STATE : EXC_ERR_ARG;
  if (!dmat_args_ok(c, ngp, ngkmax, nst, ldv)) return EXC_ERR_ARG;
  if (nst == 0) return EXC_OK;
  if (!apwalm || !evecfv || !occ || !dm) return EXC_ERR_ARG;
  int rc = use_device(c);
  if (rc) return rc;
  return dmat_run(c, 1, &ngp, &ngkmax, &apwalm, nst, &ldv, &evecfv, &evecsv, &occ, &wkpt, dm);
}

int exc_gendmatmt_batched_dev(exc_ctx *c, int nk, const exc_kpoint_dev *ks, int nst, const int *ldv,
                              const double *const *evecfv, const double *const *evecsv, const double *const *occ,
                              const double *wkpt, double *dm) {
  if (!ready(c)) return c ? EXC_ERR_STATE : EXC_ERR_ARG;
  if (nk < 0 || (nk > 0 && (!ks || !ldv || !evecfv || !occ || !wkpt || !dm))) return EXC_ERR_ARG;
  std::vector<int> ngp(nk), ldg(nk);
  std::vector<const double *> apw(nk);
  for (int k = 0; k < nk; ++k) {
    if (!dmat_args_ok(c, ks[k].ngp, ks[k].ngkmax, nst, ldv[k])) return EXC_ERR_ARG;
    if (nst > 0 && (!ks[k].apwalm || !evecfv[k] || !occ[k])) return EXC_ERR_ARG;
    ngp[k] = ks[k].ngp;
    ldg[k] = ks[k].ngkmax;
    apw[k] = ks[k].apwalm;
  }
  if (nk == 0 || nst == 0) return EXC_OK;
  int rc = use_device(c);
  if (rc) return rc;
  return dmat_run(c, nk, ngp.data(), ldg.data(), apw.data(), nst, ldv, evecfv, evecsv, occ, wkpt, dm);
}

int exc_gendmatmt(exc_ctx *c, int ngp, int ngkmax, const double *apwalm, int nst, int ldv, const double *evecfv,
                  const double *evecsv, const double *occ, double wkpt, double *dm) {
  if (!ready(c)) return c ? EXC_ERR_STATE : EXC_ERR_ARG;
  if (!dmat_args_ok(c, ngp, ngkmax, nst, ldv)) return EXC_ERR_ARG;
  if (nst == 0) return EXC_OK;
  if (!apwalm || !evecfv || !occ || !dm) return EXC_ERR_ARG;
  int rc = use_device(c);
  if (rc) return rc;
  const int ws = c->d.naa + c->d.dm_maxlo;
  const size_t b_a = 16 * (size_t)ngkmax * c->d.naa * c->d.natm, b_v = 16 * (size_t)ldv * nst;
  const size_t b_s = evecsv ? 16 * (size_t)nst * nst : 0, b_o = 8 * (size_t)nst;
  const size_t b_d = 16 * (size_t)ws * ws * c->d.natm;
  auto al = [](size_t b) { return (b + 255) & ~(size_t)255; };
  const size_t o_v = al(b_a), o_s = o_v + al(b_v), o_o = o_s + al(b_s), o_d = o_o + al(b_o);
  if ((rc = grow(&c->stage, &c->stage_bytes, o_d + b_d))) return rc;
  char *base = (char *)c->stage;
  EXC_CHECK(hipMemcpyAsync(base, apwalm, b_a, hipMemcpyHostToDevice, c->stream));
  EXC_CHECK(hipMemcpyAsync(base + o_v, evecfv, b_v, hipMemcpyHostToDevice, c->stream));
  if (evecsv) EXC_CHECK(hipMemcpyAsync(base + o_s, evecsv, b_s, hipMemcpyHostToDevice, c->stream));
  EXC_CHECK(hipMemcpyAsync(base + o_o, occ, b_o, hipMemcpyHostToDevice, c->stream));
  EXC_CHECK(hipMemcpyAsync(base + o_d, dm, b_d, hipMemcpyHostToDevice, c->stream));
  const double *pa = (const double *)base, *pv = (const double *)(base + o_v);
  const double *ps = evecsv ? (const double *)(base + o_s) : nullptr, *po = (const double *)(base + o_o);
  if ((rc = dmat_run(c, 1, &ngp, &ngkmax, &pa, nst, &ldv, &pv, &ps, &po, &wkpt, (double *)(base + o_d)))) return rc;
  EXC_CHECK(hipMemcpyAsync(dm, base + o_d, b_d, hipMemcpyDeviceToHost, c->stream));
  EXC_CHECK(hipStreamSynchronize(c->stream));
  return EXC_OK;
}

}  // extern "C"
