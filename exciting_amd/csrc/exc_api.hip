// C-ABI of the gfx950 LAPW H+O setup: context management, SCF-data uploads and the per-k
// entry points declared in include/exc_lapw.h.
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include <algorithm>
#include <thread>
#include <vector>

#include <hipfft/hipfft.h>
#include <rocsolver/rocsolver.h>

#include "exc_common.h"

struct exc_ctx {
  exc_dims dims;
  int device;
  hipStream_t stream;
  exc_dev d;
  bool have_atoms, have_lo, have_radial, have_mt, have_gvec;
  int radial_info;         // > 0: first singular matching matrix (1-based), as zgesv's info
  int mt_variant;          // tile variant of the MT stage (-1: chosen per call)
  int y_variant;
  // grow-only workspaces
  cplx *yt;
  size_t yt_bytes;
  // staging for the host-pointer entry points: device buffers and a pinned host buffer
  void *stage;
  size_t stage_bytes;
  void *hstage;
  size_t hstage_bytes;
  std::vector<void *> owned;
  // per-stage HIP-event timers (exc_ctx_set_timing)
  bool timing;
  struct mark {
    int stage;
    hipEvent_t a, b;
  };
  std::vector<mark> marks;
  std::vector<hipEvent_t> pool;
  // grouped launches
  int bmt_variant;          // -1: built-in choice
  int bsplit_variant;
  unsigned short *planes;   // split-BF16 slice planes of all k-points (A and Yt)
  size_t planes_bytes;
  int *expo;                // per-(row, atom) exponents
  size_t expo_bytes;
  cplx *vjp_work;           // adjoint workspace
  size_t vjp_bytes;
  cplx *byt;                // Y-stage workspace of all k-points
  size_t byt_bytes;
  kdesc *d_ks;
  size_t ks_cap;
  mt_item *d_items;
  size_t items_cap;
  double *rfbuf;            // radial factors of all k-points (EXC_BATCH_OLP_LEGENDRE)
  size_t rf_bytes;
  mt_item *d_oitems;        // overlap tiles of the addition-theorem kernel
  size_t oitems_cap;
  int batch_noitems;
  std::vector<char> batch_sig;   // the batch the descriptors / graph were built for
  int batch_nitems, batch_nblocks, batch_maxngp, batch_nk;
  hipGraph_t graph;
  hipGraphExec_t graph_exec;
  // generalized eigensolve: worker 0 runs on the context's stream (exc_eigen_dev), workers
  // 1..kEigStreams on streams of their own (exc_eigen_batched_dev)
  struct EigWorker {
    hipStream_t s;
    rocblas_handle blas;
    double *w;              // all n eigenvalue slots rocSOLVER writes
    size_t wb;
    cplx *z;                // n x n eigenvector workspace (rocSOLVER may use all n columns)
    size_t zb;
    cplx *zc;               // n x nev eigenvectors of the grouped path's back-transform
    size_t zcb;
    double *e;              // n-length workspace (tridiagonal off-diagonal, ifail)
    size_t eb;
    hipEvent_t done;
  };
  static constexpr int kEigStreams = 4;
  EigWorker ew[kEigStreams + 1];
  int eig_variant;          // -1: built-in choice
  int *eig_int;             // per solve: 4 ints (device)
  size_t eig_int_bytes;
  void *eg;                 // grouped solve: padded batch, tridiagonal factors, infos, descriptors
  size_t eg_bytes;
  void *dmw;                // density matrix: A^T C, packed and weighted wavefunction blocks
  size_t dmw_bytes;
  // muffin-tin integrals (exc_ctx_hmlint): host copies of nlorb / lorbl, the Gaunt table of
  // gnt_lmaxvr (device), radial inputs + integrals (device)
  std::vector<int> h_nlorb, h_lorbl;
  cplx *gnt;
  size_t gnt_bytes;
  int gnt_lmaxvr;
  void *hml;
  size_t hml_bytes;
  int hml_lmaxvr, hml_nrmtmax;   // of the last exc_ctx_hmlint (-1: none)
  // Theta(r) / V(G) (exc_ctx_gencfun, exc_ctx_genveffig): 3-D FFT plan of the G box, the grid
  // work array, Theta(r) on the grid, a staging array for V(r)
  hipfftHandle fft;
  bool have_fft, have_cfunir;
  hipfftHandle fftb;          // batched backward FFT of genrhoir, fftb_n transforms per call
  int fftb_n;
  cplx *zb;
  size_t zb_bytes;
  cplx *fz;
  double *cfunir, *fin;
  hipEvent_t eig_fork;
  // second branch of a captured batch (overlap + lo kernels beside the Y + MT chain)
  hipStream_t side;
  hipEvent_t ev_fork, ev_join;
  int ncu;
};

namespace {

template <typename T>
int dev_alloc(exc_ctx *c, T **p, size_t n) {
  *p = nullptr;
  if (n == 0) return EXC_OK;
  void *q = nullptr;
  EXC_CHECK(hipMalloc(&q, n * sizeof(T)));
  c->owned.push_back(q);
  *p = (T *)q;
  return EXC_OK;
}

template <typename T>
int dev_upload(exc_ctx *c, T **p, const void *src, size_t n) {
  if (*p == nullptr) {
    int rc = dev_alloc(c, p, n);
    if (rc) return rc;
  }
  if (n) EXC_CHECK(hipMemcpyAsync(*p, src, n * sizeof(T), hipMemcpyHostToDevice, c->stream));
  return EXC_OK;
}

// pinned host staging (grow-only); every host entry synchronises before it returns, so no
// copy into the old buffer is in flight when it is replaced
int grow_host(void **p, size_t *have, size_t need) {
  if (*have >= need) return EXC_OK;
  if (*p) hipHostFree(*p);
  *p = nullptr;
  *have = 0;
  EXC_CHECK(hipHostMalloc(p, need, hipHostMallocDefault));
  *have = need;
  return EXC_OK;
}

// ncols columns of rowbytes bytes between column-major host arrays (pitches dld, sld);
// large transfers are split over host threads (the pageable side is the caller's array)
void copy_cols(char *dst, size_t dld, const char *src, size_t sld, size_t rowbytes, size_t ncols) {
  auto work = [=](size_t j0, size_t j1) {
    for (size_t j = j0; j < j1; ++j) memcpy(dst + j * dld, src + j * sld, rowbytes);
  };
  const unsigned hc = std::thread::hardware_concurrency();
  const unsigned nt = (rowbytes * ncols >= ((size_t)8 << 20)) ? std::min(8u, std::max(1u, hc)) : 1u;
  if (nt <= 1 || ncols < nt) {
    work(0, ncols);
    return;
  }
  const size_t per = (ncols + nt - 1) / nt;
  std::vector<std::thread> th;
  size_t done = per;                      // columns [0, per) run on this thread
  try {
    for (unsigned t = 1; t < nt; ++t) {
      const size_t a = t * per, b = std::min(ncols, a + per);
      if (a >= b) break;
      th.emplace_back(work, a, b);
      done = b;
    }
  } catch (...) {                         // no threads available: copy the rest here
  }
  work(0, std::min(ncols, per));
  if (done < ncols) work(done, ncols);
  for (auto &x : th) x.join();
}

int grow(void **p, size_t *have, size_t need) {
  if (need <= *have) return EXC_OK;
  if (*p) EXC_CHECK(hipFree(*p));
  *p = nullptr;
  *have = 0;
  EXC_CHECK(hipMalloc(p, need));
  *have = need;
  return EXC_OK;
}

int use_device(exc_ctx *c) {
  int cur = -1;
  EXC_CHECK(hipGetDevice(&cur));
  if (cur != c->device) EXC_CHECK(hipSetDevice(c->device));
  return EXC_OK;
}

bool ready(const exc_ctx *c) { return c && c->have_atoms && c->have_lo && c->have_radial && c->have_mt && c->have_gvec; }

hipEvent_t take_event(exc_ctx *c) {
  if (!c->pool.empty()) {
    hipEvent_t e = c->pool.back();
    c->pool.pop_back();
    return e;
  }
  hipEvent_t e = nullptr;
  if (hipEventCreate(&e) != hipSuccess) return nullptr;
  return e;
}

// Brackets one stage's launches with events when timing is on.
struct StageTimer {
  exc_ctx *c;
  int stage;
  hipEvent_t a = nullptr;
  StageTimer(exc_ctx *ctx, int st) : c(ctx), stage(st) {
    if (st >= 0 && c->timing && (a = take_event(c))) hipEventRecord(a, c->stream);
  }
  ~StageTimer() {
    if (!a) return;
    hipEvent_t b = take_event(c);
    if (!b) {
      c->pool.push_back(a);
      return;
    }
    hipEventRecord(b, c->stream);
    c->marks.push_back({stage, a, b});
  }
};

}  // namespace

extern "C" {

int exc_ctx_create(const exc_dims *dims, exc_ctx **out) {
  if (!dims || !out) return EXC_ERR_ARG;
  *out = nullptr;
  const exc_dims &m = *dims;
  if (m.natmtot <= 0 || m.nspecies <= 0 || m.lmaxapw < 0 || m.ngrtot <= 0 || m.ngvec <= 0 ||
      m.ngvec > m.ngrtot || m.omega <= 0.0 || m.nlomax < 0 || m.maxnlo < 0)
    return EXC_ERR_ARG;
  if (m.apwordmax != 2 || m.lmaxmat != m.lmaxapw || m.lmaxapw > 12 || m.lmaxapw < 1) return EXC_ERR_UNSUPPORTED;
  for (int i = 0; i < 3; ++i)
    if (m.intgv[3 + i] < m.intgv[i]) return EXC_ERR_ARG;
  exc_ctx *c = new exc_ctx();
  c->dims = m;
  if (hipGetDevice(&c->device) != hipSuccess) {
    delete c;
    return EXC_ERR_HIP;
  }
  c->stream = nullptr;
  memset(&c->d, 0, sizeof(c->d));
  exc_dev &d = c->d;
  d.natm = m.natmtot;
  d.nspecies = m.nspecies;
  d.lmaxapw = m.lmaxapw;
  d.lmaxmat = m.lmaxmat;
  d.apwordmax = m.apwordmax;
  d.lmmaxapw = (m.lmaxapw + 1) * (m.lmaxapw + 1);
  d.naa = m.apwordmax * (m.lmaxmat + 1) * (m.lmaxmat + 1);
  d.maxnlo = m.maxnlo;
  d.nlomax = m.nlomax;
  d.ngrtot = m.ngrtot;
  d.ngvec = m.ngvec;
  for (int i = 0; i < 3; ++i) {
    d.box_lo[i] = m.intgv[i];
    d.box_n[i] = m.intgv[3 + i] - m.intgv[i] + 1;
  }
  d.omega = m.omega;
  // Summation order of the muffin-tin sums: l descending (the small high-l terms first),
  // atoms interleaved inside each l, then m and the APW order io.
  {
    const int L = m.lmaxmat, na = m.natmtot, ao = m.apwordmax, naa = d.naa;
    std::vector<int> kp, fp;
    for (int l = L; l >= 0; --l) {
      for (int ias = 0; ias < na; ++ias)
        for (int mm = -l; mm <= l; ++mm)
          for (int io = 0; io < ao; ++io) kp.push_back(ias * naa + ao * (l * (l + 1) + mm) + io);
      for (int mm = -l; mm <= l; ++mm)
        for (int io = 0; io < ao; ++io) fp.push_back(ao * (l * (l + 1) + mm) + io);
    }
    int rc = dev_upload(c, &d.kperm, kp.data(), kp.size());
    if (!rc) rc = dev_upload(c, &d.fperm, fp.data(), fp.size());
    if (!rc && hipStreamSynchronize(nullptr) != hipSuccess) rc = EXC_ERR_HIP;
    if (rc) {
      for (void *p : c->owned) hipFree(p);
      delete c;
      return rc;
    }
  }
  c->mt_variant = -1;
  c->y_variant = -1;
  c->bmt_variant = -1;
  c->bsplit_variant = -1;
  c->eig_variant = -1;
  c->gnt_lmaxvr = -1;
  c->hml_lmaxvr = -1;
  {
    hipDeviceProp_t prop;
    c->ncu = (hipGetDeviceProperties(&prop, c->device) == hipSuccess) ? prop.multiProcessorCount : 256;
  }
  *out = c;
  return EXC_OK;
}

int exc_ctx_destroy(exc_ctx *c) {
  if (!c) return EXC_OK;
  use_device(c);
  if (c->stream) hipStreamSynchronize(c->stream);
  else hipDeviceSynchronize();
  for (void *p : c->owned) hipFree(p);
  if (c->yt) hipFree(c->yt);
  for (int i = 0; i <= exc_ctx::kEigStreams; ++i) {
    exc_ctx::EigWorker &w = c->ew[i];
    if (w.blas) rocblas_destroy_handle(w.blas);
    if (w.w) hipFree(w.w);
    if (w.z) hipFree(w.z);
    if (w.zc) hipFree(w.zc);
    if (w.e) hipFree(w.e);
    if (w.done) hipEventDestroy(w.done);
    if (i > 0 && w.s) hipStreamDestroy(w.s);
  }
  if (c->eig_int) hipFree(c->eig_int);
  if (c->eg) hipFree(c->eg);
  if (c->dmw) hipFree(c->dmw);
  if (c->gnt) hipFree(c->gnt);
  if (c->have_fft) hipfftDestroy(c->fft);
  if (c->fftb_n > 0) hipfftDestroy(c->fftb);
  if (c->zb) hipFree(c->zb);
  if (c->hml) hipFree(c->hml);
  if (c->eig_fork) hipEventDestroy(c->eig_fork);
  if (c->side) hipStreamDestroy(c->side);
  if (c->ev_fork) hipEventDestroy(c->ev_fork);
  if (c->ev_join) hipEventDestroy(c->ev_join);
  if (c->stage) hipFree(c->stage);
  if (c->hstage) hipHostFree(c->hstage);
  for (auto &m : c->marks) {
    hipEventDestroy(m.a);
    hipEventDestroy(m.b);
  }
  for (hipEvent_t e : c->pool) hipEventDestroy(e);
  if (c->graph_exec) hipGraphExecDestroy(c->graph_exec);
  if (c->graph) hipGraphDestroy(c->graph);
  if (c->byt) hipFree(c->byt);
  if (c->d_ks) hipFree(c->d_ks);
  if (c->d_items) hipFree(c->d_items);
  if (c->rfbuf) hipFree(c->rfbuf);
  if (c->d_oitems) hipFree(c->d_oitems);
  if (c->planes) hipFree(c->planes);
  if (c->expo) hipFree(c->expo);
  if (c->vjp_work) hipFree(c->vjp_work);
  delete c;
  return EXC_OK;
}

int exc_ctx_set_stream(exc_ctx *c, void *stream) {
  if (!c) return EXC_ERR_ARG;
  c->stream = (hipStream_t)stream;
  return EXC_OK;
}

namespace {

bool eig_args_ok(int nmat, int ldh, int nev, int ldv) {
  return nmat >= 0 && ldh >= (nmat > 1 ? nmat : 1) && nev >= 0 && nev <= nmat && ldv >= (nmat > 1 ? nmat : 1);
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
  const auto itype = rocblas_eform_ax;
  const auto ev = rocblas_evect_original;
  const auto up = rocblas_fill_upper;
  rocblas_double_complex *A = (rocblas_double_complex *)h, *B = (rocblas_double_complex *)o;
  rocblas_double_complex *Z = (rocblas_double_complex *)W.z;
  rocblas_status st;
  bool in_a = false;      // eigenvectors returned in A (all n columns) rather than in Z (v 0, 1, 4)
  if (phase == 1) {
    // ZHEGVX('U') reads the upper triangles and the real parts of the diagonals only;
    // rocSOLVER's solvers and the products of variant 4 read more: lower := conj(upper)
    if ((rc = exc_launch_eig_hermfill((cplx *)h, nmat, ldh, W.s))) return rc;
    if ((rc = exc_launch_eig_hermfill((cplx *)o, nmat, ldh, W.s))) return rc;
    if (v < 4) return EXC_OK;
    return rocsolver_zpotrf(W.blas, up, nmat, B, ldh, d_ints + 1) == rocblas_status_success ? EXC_OK
                                                                                           : EXC_ERR_HIP;   // O = U^H U
  }
  switch (v) {
    case 0:
      st = rocsolver_zhegvdx(W.blas, itype, ev, rocblas_erange_index, up, nmat, A, ldh, B, ldh, 0.0, 0.0, 1, nev,
                             d_ints, W.w, Z, nmat, d_ints + 1);
      break;
    case 1:
      st = rocsolver_zhegvx(W.blas, itype, ev, rocblas_erange_index, up, nmat, A, ldh, B, ldh, 0.0, 0.0, 1, nev, 0.0,
                            d_ints, W.w, Z, nmat, (rocblas_int *)W.e, d_ints + 1);
      break;
    case 2:
      st = rocsolver_zhegvd(W.blas, itype, ev, up, nmat, A, ldh, B, ldh, W.w, W.e, d_ints + 1);
      in_a = true;
      break;
    case 3:
      st = rocsolver_zhegv(W.blas, itype, ev, up, nmat, A, ldh, B, ldh, W.w, W.e, d_ints + 1);
      in_a = true;
      break;
    default: {
      const rocblas_double_complex one = {1.0, 0.0};
      st = rocsolver_ztrtri(W.blas, up, rocblas_diagonal_non_unit, nmat, B, ldh, d_ints + 2);   // B = U^-1
      if (st == rocblas_status_success)                                                 // Z = H U^-1
        st = rocblas_ztrmm(W.blas, rocblas_side_right, up, rocblas_operation_none, rocblas_diagonal_non_unit, nmat,
                           nmat, &one, B, ldh, A, ldh, Z, nmat);
      if (st == rocblas_status_success)                                                 // H = U^-H Z
        st = rocblas_ztrmm(W.blas, rocblas_side_left, up, rocblas_operation_conjugate_transpose,
                           rocblas_diagonal_non_unit, nmat, nmat, &one, B, ldh, Z, nmat, A, ldh);
      if (st == rocblas_status_success)                                                 // C Y = Y diag(w)
        st = rocsolver_zheevd(W.blas, ev, up, nmat, A, ldh, W.w, W.e, d_ints + 3);
      if (st == rocblas_status_success)                                                 // Z = U^-1 Y(:, 1:nev)
        st = rocblas_ztrmm(W.blas, rocblas_side_left, up, rocblas_operation_none, rocblas_diagonal_non_unit, nmat,
                           nev, &one, B, ldh, A, ldh, Z, nmat);
      break;
    }
  }
  if (st != rocblas_status_success) return EXC_ERR_HIP;
  EXC_CHECK(hipMemcpyAsync(eval, W.w, sizeof(double) * nev, hipMemcpyDeviceToDevice, W.s));
  EXC_CHECK(hipMemcpy2DAsync(evec, 16 * (size_t)ldv, in_a ? (const void *)A : (const void *)Z,
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
// tridiagonal divide and conquer of the leading nmat_k part (real arithmetic), its first nev
// eigenvectors through the Householder reflectors (zunmtr) and U^-1 (trmm) into the caller's
// evec.  H and O are read, not modified.
int eig_grouped(exc_ctx *c, int nk, const int *nmat, const int *ldh, double *const *h, double *const *o, int nev,
                double *const *eval, double *const *evec, const int *ldv) {
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
  const int nw = std::min(nb, (int)exc_ctx::kEigStreams);
  const size_t mat = (size_t)n * n, a256 = 255;
  auto al = [&](size_t b) { return (b + a256) & ~a256; };
  const size_t o_po = al(16 * mat * nb), o_pz = o_po + al(16 * mat * nb), o_d = o_pz + al(16 * mat * nb);
  const size_t o_e = o_d + al(8 * (size_t)n * nb);
  const size_t o_tau = o_e + al(8 * (size_t)n * nb), o_i = o_tau + al(16 * (size_t)n * nb);
  const size_t o_kd = o_i + al(12 * (size_t)nb);
  const size_t total = o_kd + sizeof(eig_kdesc) * nb;
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
      if ((rc = grow((void **)&W.zc, &W.zcb, sizeof(cplx) * (size_t)n * nev))) return rc;
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
  EXC_CHECK(hipEventRecord(c->eig_fork, c->stream));
  for (int i = 1; i <= nw; ++i) EXC_CHECK(hipStreamWaitEvent(c->ew[i].s, c->eig_fork, 0));
  // per k-point on the workers; every worker is joined below, also after an error
  int err = EXC_OK;
  for (int b = 0; b < nb && !err; ++b) {
    exc_ctx::EigWorker &W = c->ew[1 + b % nw];
    const int m = kd[b].n;
    double *db = d + (size_t)b * n, *eb = e + (size_t)b * n;
    st = rocsolver_dstedc(W.blas, rocblas_evect_tridiagonal, m, db, eb, (double *)W.z, m, info + 2 * nb + b);
    if (st != rocblas_status_success) {
      err = fail("dstedc status", (int)st, b);
      break;
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
  EXC_CHECK(hipMemcpyAsync(hi.data(), info, sizeof(int) * 3 * nb, hipMemcpyDeviceToHost, c->stream));
  EXC_CHECK(hipStreamSynchronize(c->stream));
  for (int b = 0; b < nb; ++b) {
    if (hi[nb + b] != 0) return fail("ztrtri info", b, hi[nb + b]);   // of a positive definite factor
    if (hi[2 * nb + b] != 0) return hi[2 * nb + b];                 // eigenvalues that did not converge
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
  const int v = c->eig_variant >= 0 ? c->eig_variant : 5;
  if (v == 5) return eig_grouped(c, nk, nmat, ldh, h, o, nev, eval, evec, ldv);
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
        exc_ctx::EigWorker &W = c->ew[1 + k % nw];
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

int exc_ctx_nlotot(const exc_ctx *c) { return c ? c->d.nlotot : EXC_ERR_ARG; }

int exc_ctx_set_variant(exc_ctx *c, int stage, int variant) {
  if (!c) return EXC_ERR_ARG;
  switch (stage) {
    case EXC_STAGE_YSTAGE:
      if (variant < -1 || variant >= kYstageVariants) return EXC_ERR_ARG;
      c->y_variant = variant;
      return EXC_OK;
    case EXC_STAGE_MT:
      if (variant < -1 || variant >= kMtVariants) return EXC_ERR_ARG;
      c->mt_variant = variant;
      return EXC_OK;
    case EXC_STAGE_BMT:
      if (variant < -1 || variant >= kBmtVariants) return EXC_ERR_ARG;
      c->bmt_variant = variant;
      return EXC_OK;
    case EXC_STAGE_BSPLIT:
      if (variant < -1 || variant >= kBsplitVariants) return EXC_ERR_ARG;
      c->bsplit_variant = variant;
      return EXC_OK;
    case EXC_STAGE_EIGEN:
      if (variant < -1 || variant >= kEigenVariants) return EXC_ERR_ARG;
      c->eig_variant = variant;
      return EXC_OK;
    default:
      return EXC_ERR_ARG;
  }
}

int exc_ctx_num_variants(int stage) {
  switch (stage) {
    case EXC_STAGE_YSTAGE: return kYstageVariants;
    case EXC_STAGE_MT: return kMtVariants;
    case EXC_STAGE_BMT: return kBmtVariants;
    case EXC_STAGE_BSPLIT: return kBsplitVariants;
    case EXC_STAGE_EIGEN: return kEigenVariants;
    default: return 0;
  }
}

int exc_ctx_set_timing(exc_ctx *c, int enable) {
  if (!c) return EXC_ERR_ARG;
  c->timing = enable != 0;
  return EXC_OK;
}

int exc_ctx_read_timing(exc_ctx *c, double *ms, int *launches) {
  if (!c || !ms || !launches) return EXC_ERR_ARG;
  int rc = use_device(c);
  if (rc) return rc;
  for (int i = 0; i < EXC_NSTAGE; ++i) {
    ms[i] = 0.0;
    launches[i] = 0;
  }
  if (!c->marks.empty()) EXC_CHECK(hipStreamSynchronize(c->stream));
  for (auto &m : c->marks) {
    float t = 0.f;
    EXC_CHECK(hipEventElapsedTime(&t, m.a, m.b));
    if (m.stage >= 0 && m.stage < EXC_NSTAGE) {
      ms[m.stage] += t;
      launches[m.stage] += 1;
    }
    c->pool.push_back(m.a);
    c->pool.push_back(m.b);
  }
  c->marks.clear();
  return EXC_OK;
}

int exc_ctx_set_atoms(exc_ctx *c, const int *natoms, const double *rmt, const double *atposc) {
  if (!c || !natoms || !rmt || !atposc) return EXC_ERR_ARG;
  int rc = use_device(c);
  if (rc) return rc;
  const int ns = c->dims.nspecies, na = c->dims.natmtot;
  std::vector<int> sp;
  for (int s = 0; s < ns; ++s) {
    if (natoms[s] < 0 || rmt[s] <= 0.0) return EXC_ERR_ARG;
    for (int a = 0; a < natoms[s]; ++a) sp.push_back(s);
  }
  if ((int)sp.size() != na) return EXC_ERR_ARG;
  if ((rc = dev_upload(c, &c->d.spidx, sp.data(), na))) return rc;
  if ((rc = dev_upload(c, &c->d.rmt, rmt, ns))) return rc;
  if ((rc = dev_upload(c, &c->d.atposc, atposc, 3 * (size_t)na))) return rc;
  EXC_CHECK(hipStreamSynchronize(c->stream));
  c->have_atoms = true;
  return EXC_OK;
}

int exc_ctx_set_lo(exc_ctx *c, const int *nlorb, const int *lorbl, const int *idxlo) {
  if (!c || !nlorb || (c->dims.nlomax > 0 && (!lorbl || !idxlo))) return EXC_ERR_ARG;
  if (!c->have_atoms) return EXC_ERR_STATE;
  int rc = use_device(c);
  if (rc) return rc;
  const int ns = c->dims.nspecies, na = c->dims.natmtot, nlomax = c->dims.nlomax, lolm = c->dims.lolmmax;
  std::vector<int> sp(na);
  EXC_CHECK(hipMemcpy(sp.data(), c->d.spidx, sizeof(int) * na, hipMemcpyDeviceToHost));
  std::vector<lo_desc> lod;
  int maxlos = 0;
  for (int ias = 0; ias < na; ++ias) {
    const int is = sp[ias];
    if (nlorb[is] < 0 || nlorb[is] > nlomax) return EXC_ERR_ARG;
    const int first = (int)lod.size();
    for (int ilo = 0; ilo < nlorb[is]; ++ilo) {
      const int l = lorbl[ilo + nlomax * is];
      if (l < 0 || (l + 1) * (l + 1) > lolm) return EXC_ERR_ARG;
      for (int m = -l; m <= l; ++m) {
        lo_desc q;
        q.ias = ias;
        q.row = (int)lod.size() - first;
        q.ilo = ilo;
        q.l = l;
        q.m = m;
        q.lm = l * (l + 1) + m;
        q.first = first;
        q.pad = 0;
        // the setup routines assume the running numbering of genidxlo.f90:24-43
        if (idxlo[q.lm + lolm * (ilo + nlomax * ias)] != (int)lod.size() + 1) return EXC_ERR_ARG;
        lod.push_back(q);
      }
    }
    maxlos = (int)lod.size() - first > maxlos ? (int)lod.size() - first : maxlos;
  }
  if (maxlos > c->dims.maxnlo) return EXC_ERR_ARG;
  c->d.nlotot = (int)lod.size();
  // density-matrix tables: per atom (first lo, number of lo), and the row order f = (l, io, m)
  std::vector<int> dlo(2 * (size_t)na, 0);
  for (const lo_desc &q : lod) {
    dlo[2 * q.ias] = q.first;
    dlo[2 * q.ias + 1] += 1;
  }
  std::vector<int> dord;
  for (int l = 0; l <= c->d.lmaxmat; ++l)
    for (int io = 0; io < c->d.apwordmax; ++io)
      for (int m = -l; m <= l; ++m) dord.push_back(io + c->d.apwordmax * (l * l + l + m));
  if ((int)dord.size() != c->d.naa) return EXC_ERR_ARG;
  if (!c->d.dm_lo && (rc = dev_alloc(c, &c->d.dm_lo, dlo.size()))) return rc;
  if (!c->d.dm_order && (rc = dev_alloc(c, &c->d.dm_order, dord.size()))) return rc;
  EXC_CHECK(hipMemcpy(c->d.dm_lo, dlo.data(), sizeof(int) * dlo.size(), hipMemcpyHostToDevice));
  EXC_CHECK(hipMemcpy(c->d.dm_order, dord.data(), sizeof(int) * dord.size(), hipMemcpyHostToDevice));
  c->d.dm_maxlo = maxlos;
  c->h_nlorb.assign(nlorb, nlorb + ns);
  c->h_lorbl.assign((size_t)nlomax * ns, 0);
  for (int is = 0; is < ns; ++is)
    for (int ilo = 0; ilo < nlorb[is]; ++ilo) c->h_lorbl[ilo + (size_t)nlomax * is] = lorbl[ilo + nlomax * is];
  if (c->d.lod == nullptr && !lod.empty()) {
    if ((rc = dev_alloc(c, &c->d.lod, lod.size()))) return rc;
  }
  if (!lod.empty()) EXC_CHECK(hipMemcpy(c->d.lod, lod.data(), sizeof(lo_desc) * lod.size(), hipMemcpyHostToDevice));
  c->have_lo = true;
  return EXC_OK;
}

int exc_ctx_set_radial(exc_ctx *c, const double *dmat) {
  if (!c || !dmat) return EXC_ERR_ARG;
  int rc = use_device(c);
  if (rc) return rc;
  const int na = c->dims.natmtot, L = c->dims.lmaxapw, ao = c->dims.apwordmax;
  std::vector<double> inv((size_t)na * (L + 1) * 4);
  int info = 0;
  for (int ias = 0; ias < na; ++ias)
    for (int l = 0; l <= L; ++l) {
      const double *m = dmat + (size_t)ao * ao * (l + (L + 1) * ias);
      // m column-major 2x2: m[0]=D11 m[1]=D21 m[2]=D12 m[3]=D22
      const double det = m[0] * m[3] - m[2] * m[1];
      double *o = inv.data() + 4 * (l + (L + 1) * (size_t)ias);
      if (det == 0.0 || !isfinite(det)) {
        if (!info) info = 1 + l + (L + 1) * ias;
        o[0] = o[1] = o[2] = o[3] = 0.0;
      } else {
        o[0] = m[3] / det;
        o[1] = -m[1] / det;
        o[2] = -m[2] / det;
        o[3] = m[0] / det;
      }
    }
  if ((rc = dev_upload(c, &c->d.dinv, inv.data(), inv.size()))) return rc;
  EXC_CHECK(hipStreamSynchronize(c->stream));
  c->radial_info = info;
  c->have_radial = true;
  return info;
}

int exc_ctx_set_mt(exc_ctx *c, const double *haa, const double *loa, const double *lolo, const double *oalo,
                   const double *ololo) {
  if (!c || !haa) return EXC_ERR_ARG;
  if (c->dims.maxnlo > 0 && (!loa || !lolo || !oalo || !ololo)) return EXC_ERR_ARG;
  int rc = use_device(c);
  if (rc) return rc;
  const size_t na = c->dims.natmtot, naa = c->d.naa, mn = c->dims.maxnlo, nl = c->dims.nlomax, ao = c->dims.apwordmax;
  if ((rc = dev_upload(c, &c->d.haa, haa, naa * naa * na))) return rc;
  if (mn) {
    if ((rc = dev_upload(c, &c->d.loa, loa, mn * naa * na))) return rc;
    if ((rc = dev_upload(c, &c->d.lolo, lolo, mn * mn * na))) return rc;
  }
  if (nl) {
    if ((rc = dev_upload(c, &c->d.oalo, oalo, ao * nl * na))) return rc;
    if ((rc = dev_upload(c, &c->d.ololo, ololo, nl * nl * na))) return rc;
  }
  EXC_CHECK(hipStreamSynchronize(c->stream));
  c->have_mt = true;
  return EXC_OK;
}

namespace {
// m-th derivative (m = 0, 1) at x of the cubic through four points (polynom.f90 with np = 4):
// the cubic in t = x - xa[0] from the offsets of the other three points
double cubic_deriv(int m, const double *xa, const double *ya, double x) {
  const double x1 = xa[1] - xa[0], x2 = xa[2] - xa[0], x3 = xa[3] - xa[0];
  const double y1 = ya[1] - ya[0], y2 = ya[2] - ya[0], y3 = ya[3] - ya[0];
  const double t1 = 1.0 / (x1 * x2 * x3 * (x1 - x2) * (x1 - x3) * (x2 - x3));
  const double t2 = x1 * x2 * y3, t3 = x2 * x3 * y1, t4 = x3 * x1 * y2;
  const double t5 = x1 * x1, t6 = x2 * x2, t7 = x3 * x3;
  const double c2 = t1 * (t3 * (x3 * t6 - x2 * t7) + t4 * (x1 * t7 - x3 * t5) + t2 * (x2 * t5 - x1 * t6));
  const double c3 = t1 * (t3 * (t7 - t6) + t4 * (t5 - t7) + t2 * (t6 - t5));
  const double c4 = t1 * (t3 * (x2 - x3) + t4 * (x3 - x1) + t2 * (x1 - x2));
  const double t = x - xa[0];
  return m == 0 ? ya[0] + t * (c2 + t * (c3 + c4 * t)) : c2 + t * (2.0 * c3 + 3.0 * c4 * t);
}
}  // namespace

int exc_ctx_set_radial_fr(exc_ctx *c, int nrmtmax, const int *nrmt, const double *spr, const double *apwfr) {
  if (!c || !nrmt || !spr || !apwfr) return EXC_ERR_ARG;
  if (!c->have_atoms) return EXC_ERR_STATE;
  const int ns = c->dims.nspecies, na = c->dims.natmtot, L = c->dims.lmaxapw, ao = c->dims.apwordmax;
  if (std::max(ao + 1, 4) != 4) return EXC_ERR_UNSUPPORTED;   // match.f90:85 uses np = 4 here
  for (int is = 0; is < ns; ++is)
    if (nrmt[is] < 4 || nrmt[is] > nrmtmax) return EXC_ERR_ARG;
  int rc = use_device(c);
  if (rc) return rc;
  std::vector<int> sp(na);
  EXC_CHECK(hipMemcpy(sp.data(), c->d.spidx, sizeof(int) * na, hipMemcpyDeviceToHost));
  std::vector<double> rmt(ns);
  EXC_CHECK(hipMemcpy(rmt.data(), c->d.rmt, sizeof(double) * ns, hipMemcpyDeviceToHost));
  // dmat(io1, io2, l, ias) = d^(io1-1) u_(io2,l) / dr^(io1-1) at R (match.f90:139-148)
  std::vector<double> dmat((size_t)ao * ao * (L + 1) * na);
  for (int ias = 0; ias < na; ++ias) {
    const int is = sp[ias], n = nrmt[is];
    const double *xa = spr + (size_t)nrmtmax * is + (n - 4);
    for (int l = 0; l <= L; ++l)
      for (int io2 = 0; io2 < ao; ++io2) {
        const double *ya = apwfr + (size_t)nrmtmax * 2 * (io2 + (size_t)ao * (l + (size_t)(L + 1) * ias)) + (n - 4);
        for (int io1 = 0; io1 < ao; ++io1)
          dmat[io1 + ao * (io2 + (size_t)ao * (l + (size_t)(L + 1) * ias))] = cubic_deriv(io1, xa, ya, rmt[is]);
      }
  }
  return exc_ctx_set_radial(c, dmat.data());
}

namespace {
// device buffer of exc_ctx_hmlint: weights, meshes, apwfr, lofr, veffmt, radial integrals, ints
struct HmlLayout {
  int nf;
  size_t n_w, n_apw, n_lo, n_v, n_rad, o_spr, o_apw, o_lo, o_v, o_rad, o_i, total;
};
HmlLayout hml_layout(const exc_ctx *c, int lmaxvr, int nrmtmax) {
  const exc_dev &d = c->d;
  const int ns = c->dims.nspecies, na = d.natm, nlomax = c->dims.nlomax, ao = d.apwordmax;
  const int lmv = (lmaxvr + 1) * (lmaxvr + 1);
  HmlLayout L;
  L.nf = ao * (d.lmaxmat + 1) + (nlomax > 0 ? nlomax : 0);
  L.n_w = (size_t)nrmtmax * ns;
  L.n_apw = (size_t)nrmtmax * 2 * ao * (d.lmaxapw + 1) * na;
  L.n_lo = (size_t)nrmtmax * 2 * (nlomax > 0 ? nlomax : 0) * na;
  L.n_v = (size_t)lmv * nrmtmax * na;
  L.n_rad = (size_t)na * (lmv + 2) * L.nf * L.nf;
  auto al = [](size_t b) { return (b + 255) & ~(size_t)255; };
  L.o_spr = al(8 * L.n_w);
  L.o_apw = L.o_spr + al(8 * L.n_w);
  L.o_lo = L.o_apw + al(8 * L.n_apw);
  L.o_v = L.o_lo + al(8 * L.n_lo);
  L.o_rad = L.o_v + al(8 * L.n_v);
  L.o_i = L.o_rad + al(8 * L.n_rad);
  L.total = L.o_i + sizeof(int) * (ns + ns + (size_t)(nlomax > 0 ? nlomax : 1) * ns);
  return L;
}
}  // namespace

int exc_ctx_hmlint(exc_ctx *c, int lmaxvr, int nrmtmax, const int *nrmt, const double *spr, const double *apwfr,
                   const double *lofr, const double *veffmt, double *haa, double *loa, double *lolo, double *oalo,
                   double *ololo) {
  if (!c || !nrmt || !spr || !apwfr || !veffmt) return EXC_ERR_ARG;
  if (!c->have_atoms || !c->have_lo) return EXC_ERR_STATE;
  const exc_dev &d = c->d;
  const int ns = c->dims.nspecies, na = d.natm, nlomax = c->dims.nlomax, ao = d.apwordmax;
  if (nlomax > 0 && !lofr) return EXC_ERR_ARG;
  if (lmaxvr < 0 || lmaxvr > 2 * d.lmaxmat || nrmtmax < 4 || nrmtmax > 8192) return EXC_ERR_ARG;
  for (int is = 0; is < ns; ++is) {
    if (nrmt[is] < 4 || nrmt[is] > nrmtmax) return EXC_ERR_ARG;
    const double *x = spr + (size_t)nrmtmax * is;
    if (!(x[0] > 0.0)) return EXC_ERR_ARG;
    for (int i = 1; i < nrmt[is]; ++i)
      if (!(x[i] > x[i - 1])) return EXC_ERR_ARG;
    for (int ilo = 0; ilo < c->h_nlorb[is]; ++ilo)
      if (c->h_lorbl[ilo + (size_t)nlomax * is] > d.lmaxmat) return EXC_ERR_UNSUPPORTED;
  }
  int rc = use_device(c);
  if (rc) return rc;
  // integration weights of each species' mesh, the Gaunt table (per lmaxvr)
  std::vector<double> wt((size_t)nrmtmax * ns, 0.0);
  for (int is = 0; is < ns; ++is) exc_hml_spline_weights(nrmt[is], spr + (size_t)nrmtmax * is, &wt[(size_t)nrmtmax * is]);
  const int lmm = (d.lmaxmat + 1) * (d.lmaxmat + 1), lmv = (lmaxvr + 1) * (lmaxvr + 1);
  if (c->gnt_lmaxvr != lmaxvr) {
    std::vector<cplx> g((size_t)lmm * lmm * (lmaxvr + 1) * 2);
    exc_hml_gaunt_table(d.lmaxmat, lmaxvr, g.data());
    if ((rc = grow((void **)&c->gnt, &c->gnt_bytes, sizeof(cplx) * g.size()))) return rc;
    EXC_CHECK(hipMemcpy(c->gnt, g.data(), sizeof(cplx) * g.size(), hipMemcpyHostToDevice));
    c->gnt_lmaxvr = lmaxvr;
  }
  // inputs and the radial integrals in one device buffer
  const HmlLayout H = hml_layout(c, lmaxvr, nrmtmax);
  const size_t n_w = H.n_w, n_apw = H.n_apw, n_lo = H.n_lo, n_v = H.n_v;
  const size_t o_spr = H.o_spr, o_apw = H.o_apw, o_lo = H.o_lo, o_v = H.o_v, o_rad = H.o_rad, o_i = H.o_i;
  c->hml_lmaxvr = -1;
  if ((rc = grow(&c->hml, &c->hml_bytes, H.total))) return rc;
  char *b = (char *)c->hml;
  std::vector<int> ints(nrmt, nrmt + ns);
  ints.insert(ints.end(), c->h_nlorb.begin(), c->h_nlorb.end());
  if (nlomax > 0) ints.insert(ints.end(), c->h_lorbl.begin(), c->h_lorbl.end());
  else ints.insert(ints.end(), (size_t)ns, 0);
  EXC_CHECK(hipMemcpy(b, wt.data(), 8 * n_w, hipMemcpyHostToDevice));
  EXC_CHECK(hipMemcpy(b + o_spr, spr, 8 * n_w, hipMemcpyHostToDevice));
  EXC_CHECK(hipMemcpy(b + o_apw, apwfr, 8 * n_apw, hipMemcpyHostToDevice));
  if (n_lo) EXC_CHECK(hipMemcpy(b + o_lo, lofr, 8 * n_lo, hipMemcpyHostToDevice));
  EXC_CHECK(hipMemcpy(b + o_v, veffmt, 8 * n_v, hipMemcpyHostToDevice));
  EXC_CHECK(hipMemcpy(b + o_i, ints.data(), sizeof(int) * ints.size(), hipMemcpyHostToDevice));
  // output arrays of the context (what exc_ctx_set_mt uploads)
  const size_t mn = d.maxnlo, nl = nlomax > 0 ? nlomax : 0;
  if (!c->d.haa && (rc = dev_alloc(c, &c->d.haa, (size_t)d.naa * d.naa * na))) return rc;
  if (mn && !c->d.loa && (rc = dev_alloc(c, &c->d.loa, mn * d.naa * na))) return rc;
  if (mn && !c->d.lolo && (rc = dev_alloc(c, &c->d.lolo, mn * mn * na))) return rc;
  if (nl && !c->d.oalo && (rc = dev_alloc(c, &c->d.oalo, (size_t)ao * nl * na))) return rc;
  if (nl && !c->d.ololo && (rc = dev_alloc(c, &c->d.ololo, nl * nl * na))) return rc;
  const int *di = (const int *)(b + o_i);
  {
    StageTimer t(c, EXC_STAGE_HMLINT);
    if ((rc = exc_launch_hmlint(c->d, lmv, nrmtmax, di, di + ns, di + 2 * ns, (const double *)b,
                                (const double *)(b + o_spr), (const double *)(b + o_apw), (const double *)(b + o_lo),
                                (const double *)(b + o_v), (double *)(b + o_rad), c->gnt, c->stream)))
      return rc;
  }
  if (haa) EXC_CHECK(hipMemcpyAsync(haa, c->d.haa, 16 * (size_t)d.naa * d.naa * na, hipMemcpyDeviceToHost, c->stream));
  if (mn && loa) EXC_CHECK(hipMemcpyAsync(loa, c->d.loa, 16 * mn * d.naa * na, hipMemcpyDeviceToHost, c->stream));
  if (mn && lolo) EXC_CHECK(hipMemcpyAsync(lolo, c->d.lolo, 16 * mn * mn * na, hipMemcpyDeviceToHost, c->stream));
  if (nl && oalo) EXC_CHECK(hipMemcpyAsync(oalo, c->d.oalo, 8 * (size_t)ao * nl * na, hipMemcpyDeviceToHost, c->stream));
  if (nl && ololo) EXC_CHECK(hipMemcpyAsync(ololo, c->d.ololo, 8 * nl * nl * na, hipMemcpyDeviceToHost, c->stream));
  EXC_CHECK(hipStreamSynchronize(c->stream));
  c->have_mt = true;
  c->hml_lmaxvr = lmaxvr;
  c->hml_nrmtmax = nrmtmax;
  return EXC_OK;
}

int exc_genrhomt_dev(exc_ctx *c, const double *dm, double *rhomt) {
  if (!c || !dm || !rhomt) return EXC_ERR_ARG;
  if (c->hml_lmaxvr < 0 || c->gnt_lmaxvr != c->hml_lmaxvr) return EXC_ERR_STATE;
  int rc = use_device(c);
  if (rc) return rc;
  const int lmaxvr = c->hml_lmaxvr, nrmtmax = c->hml_nrmtmax, ns = c->dims.nspecies;
  const HmlLayout H = hml_layout(c, lmaxvr, nrmtmax);
  const char *b = (const char *)c->hml;
  const int *di = (const int *)(b + H.o_i);
  StageTimer t(c, EXC_STAGE_HMLINT);
  return exc_launch_genrhomt(c->d, (lmaxvr + 1) * (lmaxvr + 1), nrmtmax, di, di + ns, di + 2 * ns,
                             (const double *)(b + H.o_apw), (const double *)(b + H.o_lo), c->gnt, (const cplx *)dm,
                             c->d.naa + c->d.dm_maxlo, (double *)(b + H.o_rad), rhomt, c->stream);
}

int exc_genrhomt(exc_ctx *c, const double *dm, double *rhomt) {
  if (!c || !dm || !rhomt) return EXC_ERR_ARG;
  if (c->hml_lmaxvr < 0) return EXC_ERR_STATE;
  int rc = use_device(c);
  if (rc) return rc;
  const size_t ws = c->d.naa + c->d.dm_maxlo, na = c->d.natm;
  const size_t b_d = 16 * ws * ws * na, b_r = 8 * (size_t)(c->hml_lmaxvr + 1) * (c->hml_lmaxvr + 1) * c->hml_nrmtmax * na;
  const size_t o_r = (b_d + 255) & ~(size_t)255;
  if ((rc = grow(&c->stage, &c->stage_bytes, o_r + b_r))) return rc;
  char *base = (char *)c->stage;
  EXC_CHECK(hipMemcpyAsync(base, dm, b_d, hipMemcpyHostToDevice, c->stream));
  if ((rc = exc_genrhomt_dev(c, (const double *)base, (double *)(base + o_r)))) return rc;
  EXC_CHECK(hipMemcpyAsync(rhomt, base + o_r, b_r, hipMemcpyDeviceToHost, c->stream));
  EXC_CHECK(hipStreamSynchronize(c->stream));
  return EXC_OK;
}

namespace {
// the FFT of the G box (every grid point one G) and its work arrays, made once per context
int scf_prepare(exc_ctx *c) {
  const exc_dev &d = c->d;
  const size_t box = (size_t)d.box_n[0] * d.box_n[1] * d.box_n[2];
  if (box != (size_t)c->dims.ngrtot) return EXC_ERR_UNSUPPORTED;   // the G list must fill the box
  int rc;
  if (!c->fz && (rc = dev_alloc(c, &c->fz, box))) return rc;
  if (!c->cfunir && (rc = dev_alloc(c, &c->cfunir, box))) return rc;
  if (!c->fin && (rc = dev_alloc(c, &c->fin, box))) return rc;
  if (!c->have_fft) {
    // Fortran order: the first index runs fastest, so it is hipFFT's last dimension
    if (hipfftPlan3d(&c->fft, d.box_n[2], d.box_n[1], d.box_n[0], HIPFFT_Z2Z) != HIPFFT_SUCCESS) return EXC_ERR_HIP;
    c->have_fft = true;
  }
  if (hipfftSetStream(c->fft, c->stream) != HIPFFT_SUCCESS) return EXC_ERR_HIP;
  return EXC_OK;
}
}  // namespace

int exc_ctx_gencfun(exc_ctx *c, const double *bvec, double *cfunig, double *cfunir) {
  if (!c || !bvec) return EXC_ERR_ARG;
  if (!c->have_atoms || !c->have_gvec) return EXC_ERR_STATE;
  int rc = use_device(c);
  if (rc) return rc;
  if ((rc = scf_prepare(c))) return rc;
  const int n = c->dims.ngrtot;
  {
    StageTimer t(c, EXC_STAGE_SCF);
    if ((rc = exc_launch_cfunig(c->d, bvec, n, c->d.cfunig, c->fz, c->stream))) return rc;
    if (hipfftExecZ2Z(c->fft, (hipfftDoubleComplex *)c->fz, (hipfftDoubleComplex *)c->fz, HIPFFT_BACKWARD) !=
        HIPFFT_SUCCESS)
      return EXC_ERR_HIP;
    if ((rc = exc_launch_scf_real(n, c->fz, c->cfunir, c->stream))) return rc;
  }
  if (cfunig) EXC_CHECK(hipMemcpyAsync(cfunig, c->d.cfunig, 16 * (size_t)n, hipMemcpyDeviceToHost, c->stream));
  if (cfunir) EXC_CHECK(hipMemcpyAsync(cfunir, c->cfunir, 8 * (size_t)n, hipMemcpyDeviceToHost, c->stream));
  EXC_CHECK(hipStreamSynchronize(c->stream));
  c->have_cfunir = true;
  return EXC_OK;
}

int exc_ctx_genveffig(exc_ctx *c, const double *veffir, double *veffig) {
  if (!c || !veffir) return EXC_ERR_ARG;
  if (!c->have_cfunir) return EXC_ERR_STATE;
  int rc = use_device(c);
  if (rc) return rc;
  if ((rc = scf_prepare(c))) return rc;
  const int n = c->dims.ngrtot, nv = c->dims.ngvec;
  EXC_CHECK(hipMemcpyAsync(c->fin, veffir, 8 * (size_t)n, hipMemcpyHostToDevice, c->stream));
  {
    StageTimer t(c, EXC_STAGE_SCF);
    if ((rc = exc_launch_scf_product(n, c->fin, c->cfunir, c->fz, c->stream))) return rc;
    if (hipfftExecZ2Z(c->fft, (hipfftDoubleComplex *)c->fz, (hipfftDoubleComplex *)c->fz, HIPFFT_FORWARD) !=
        HIPFFT_SUCCESS)
      return EXC_ERR_HIP;
    if ((rc = exc_launch_scf_gather(c->d, nv, 1.0 / (double)n, c->fz, c->d.veffig, c->stream))) return rc;
  }
  if (veffig) EXC_CHECK(hipMemcpyAsync(veffig, c->d.veffig, 16 * (size_t)nv, hipMemcpyDeviceToHost, c->stream));
  EXC_CHECK(hipStreamSynchronize(c->stream));
  return EXC_OK;
}

namespace {
constexpr int kRhoirChunk = 64;   // states per batched FFT of genrhoir

// rhoir += the interstitial density of the nst states of each of nk k-points (device arrays)
int rhoir_run(exc_ctx *c, int nk, const int *ngp, const int *const *igpig, int nst, const int *ldv,
              const double *const *evec, const double *const *occ, const double *wkpt, double epsocc, double *rhoir) {
  int rc;
  if ((rc = scf_prepare(c))) return rc;
  const exc_dev &d = c->d;
  const int N = c->dims.ngrtot, B = std::min(nst, kRhoirChunk);
  if ((rc = grow((void **)&c->zb, &c->zb_bytes, sizeof(cplx) * (size_t)B * N))) return rc;
  if (c->fftb_n != B) {
    if (c->fftb_n > 0) hipfftDestroy(c->fftb);
    c->fftb_n = 0;
    int dims[3] = {d.box_n[2], d.box_n[1], d.box_n[0]};   // Fortran order: first index fastest
    if (hipfftPlanMany(&c->fftb, 3, dims, nullptr, 1, N, nullptr, 1, N, HIPFFT_Z2Z, B) != HIPFFT_SUCCESS)
      return EXC_ERR_HIP;
    c->fftb_n = B;
  }
  if (hipfftSetStream(c->fftb, c->stream) != HIPFFT_SUCCESS) return EXC_ERR_HIP;
  StageTimer t(c, EXC_STAGE_SCF);
  for (int k = 0; k < nk; ++k)
    for (int j0 = 0; j0 < nst; j0 += B) {
      const int nb = std::min(B, nst - j0);
      EXC_CHECK(hipMemsetAsync(c->zb, 0, sizeof(cplx) * (size_t)B * N, c->stream));
      if ((rc = exc_launch_rhoir_scatter(d, ngp[k], igpig[k], (const cplx *)evec[k], ldv[k], j0, nb, N, c->zb,
                                         c->stream)))
        return rc;
      if (hipfftExecZ2Z(c->fftb, (hipfftDoubleComplex *)c->zb, (hipfftDoubleComplex *)c->zb, HIPFFT_BACKWARD) !=
          HIPFFT_SUCCESS)
        return EXC_ERR_HIP;
      if ((rc = exc_launch_rhoir_accum(N, nb, c->zb, occ[k], j0, wkpt[k], epsocc, d.omega, rhoir, c->stream)))
        return rc;
    }
  return EXC_OK;
}
}  // namespace

int exc_genrhoir_dev(exc_ctx *c, int ngp, const int *igpig, int nst, int ldv, const double *evecfv, const double *occ,
                     double wkpt, double epsocc, double *rhoir) {
  if (!ready(c)) return c ? EXC_ERR_STATE : EXC_ERR_ARG;
  if (ngp < 0 || nst < 0 || ldv < std::max(ngp, 1)) return EXC_ERR_ARG;
  if (nst == 0 || ngp == 0) return EXC_OK;
  if (!igpig || !evecfv || !occ || !rhoir) return EXC_ERR_ARG;
  int rc = use_device(c);
  if (rc) return rc;
  return rhoir_run(c, 1, &ngp, &igpig, nst, &ldv, &evecfv, &occ, &wkpt, epsocc, rhoir);
}

int exc_genrhoir_batched_dev(exc_ctx *c, int nk, const exc_kpoint_dev *ks, int nst, const int *ldv,
                             const double *const *evecfv, const double *const *occ, const double *wkpt, double epsocc,
                             double *rhoir) {
  if (!ready(c)) return c ? EXC_ERR_STATE : EXC_ERR_ARG;
  if (nk < 0 || nst < 0 || (nk > 0 && (!ks || !ldv || !evecfv || !occ || !wkpt || !rhoir))) return EXC_ERR_ARG;
  std::vector<int> ngp(nk);
  std::vector<const int *> ig(nk);
  for (int k = 0; k < nk; ++k) {
    if (ks[k].ngp < 0 || ldv[k] < std::max(ks[k].ngp, 1)) return EXC_ERR_ARG;
    if (nst > 0 && (!ks[k].igpig || !evecfv[k] || !occ[k])) return EXC_ERR_ARG;
    ngp[k] = ks[k].ngp;
    ig[k] = ks[k].igpig;
  }
  if (nk == 0 || nst == 0) return EXC_OK;
  int rc = use_device(c);
  if (rc) return rc;
  return rhoir_run(c, nk, ngp.data(), ig.data(), nst, ldv, evecfv, occ, wkpt, epsocc, rhoir);
}

int exc_genrhoir(exc_ctx *c, int ngp, const int *igpig, int nst, int ldv, const double *evecfv, const double *occ,
                 double wkpt, double epsocc, double *rhoir) {
  if (!ready(c)) return c ? EXC_ERR_STATE : EXC_ERR_ARG;
  if (ngp < 0 || nst < 0 || ldv < std::max(ngp, 1)) return EXC_ERR_ARG;
  if (nst == 0 || ngp == 0) return EXC_OK;
  if (!igpig || !evecfv || !occ || !rhoir) return EXC_ERR_ARG;
  int rc = use_device(c);
  if (rc) return rc;
  const size_t N = c->dims.ngrtot, b_i = 4 * (size_t)ngp, b_v = 16 * (size_t)ldv * nst, b_o = 8 * (size_t)nst;
  auto al = [](size_t b) { return (b + 255) & ~(size_t)255; };
  const size_t o_v = al(b_i), o_o = o_v + al(b_v), o_r = o_o + al(b_o);
  if ((rc = grow(&c->stage, &c->stage_bytes, o_r + 8 * N))) return rc;
  char *base = (char *)c->stage;
  EXC_CHECK(hipMemcpyAsync(base, igpig, b_i, hipMemcpyHostToDevice, c->stream));
  EXC_CHECK(hipMemcpyAsync(base + o_v, evecfv, b_v, hipMemcpyHostToDevice, c->stream));
  EXC_CHECK(hipMemcpyAsync(base + o_o, occ, b_o, hipMemcpyHostToDevice, c->stream));
  EXC_CHECK(hipMemcpyAsync(base + o_r, rhoir, 8 * N, hipMemcpyHostToDevice, c->stream));
  const int *pi = (const int *)base;
  const double *pv = (const double *)(base + o_v), *po = (const double *)(base + o_o);
  if ((rc = rhoir_run(c, 1, &ngp, &pi, nst, &ldv, &pv, &po, &wkpt, epsocc, (double *)(base + o_r)))) return rc;
  EXC_CHECK(hipMemcpyAsync(rhoir, base + o_r, 8 * N, hipMemcpyDeviceToHost, c->stream));
  EXC_CHECK(hipStreamSynchronize(c->stream));
  return EXC_OK;
}

int exc_ctx_set_gvec(exc_ctx *c, const int *ivg, const int *ivgig, const double *cfunig, const double *veffig,
                     const double *ffacg) {
  if (!c || !ivg || !ivgig || !cfunig || !veffig) return EXC_ERR_ARG;
  int rc = use_device(c);
  if (rc) return rc;
  const size_t ng = c->dims.ngrtot, nv = c->dims.ngvec;
  const size_t box = (size_t)c->d.box_n[0] * c->d.box_n[1] * c->d.box_n[2];
  for (size_t i = 0; i < box; ++i)
    if (ivgig[i] < 0 || (size_t)ivgig[i] > ng) return EXC_ERR_ARG;
  for (size_t g = 0; g < ng; ++g)
    for (int q = 0; q < 3; ++q) {
      const int v = ivg[3 * g + q];
      if (v < c->d.box_lo[q] || v >= c->d.box_lo[q] + c->d.box_n[q]) return EXC_ERR_ARG;
    }
  if ((rc = dev_upload(c, &c->d.ivg, ivg, 3 * ng))) return rc;
  if ((rc = dev_upload(c, &c->d.ivgig, ivgig, box))) return rc;
  if ((rc = dev_upload(c, &c->d.cfunig, cfunig, ng))) return rc;
  if ((rc = dev_upload(c, &c->d.veffig, veffig, nv))) return rc;
  if (ffacg) {
    // reorder ffacg(ngvec, nspecies) -> contiguous per species (same layout)
    if ((rc = dev_upload(c, &c->d.ffacg, ffacg, nv * c->dims.nspecies))) return rc;
  }
  EXC_CHECK(hipStreamSynchronize(c->stream));
  c->have_gvec = true;
  return EXC_OK;
}

int exc_match_dev(exc_ctx *c, int ngp, int ngkmax, const double *gpc, const double *tpgpc, const double *sfacgp,
                  double *apwalm) {
  if (!ready(c)) return c ? EXC_ERR_STATE : EXC_ERR_ARG;
  if (ngp < 0 || ngkmax < ngp || (ngp > 0 && (!gpc || !tpgpc || !sfacgp || !apwalm))) return EXC_ERR_ARG;
  int rc = use_device(c);
  if (rc) return rc;
  {
    StageTimer t(c, EXC_STAGE_MATCH);
    rc = exc_launch_match(c->d, ngp, gpc, tpgpc, (const cplx *)sfacgp, ngkmax, (cplx *)apwalm, ngkmax, c->stream);
  }
  if (rc) return rc;
  return c->radial_info;
}

int exc_match(exc_ctx *c, int ngp, int ngkmax, const double *gpc, const double *tpgpc, const double *sfacgp,
              double *apwalm) {
  if (!ready(c)) return c ? EXC_ERR_STATE : EXC_ERR_ARG;
  if (ngp < 0 || ngkmax < ngp || (ngp > 0 && (!gpc || !tpgpc || !sfacgp || !apwalm))) return EXC_ERR_ARG;
  int rc = use_device(c);
  if (rc) return rc;
  const size_t na = c->d.natm, lmm = c->d.lmmaxapw, ao = c->d.apwordmax, nk = ngkmax;
  const size_t b_gpc = 8 * nk, b_tp = 16 * nk, b_sf = 16 * nk * na, b_apw = 16 * nk * ao * lmm * na;
  if ((rc = grow(&c->stage, &c->stage_bytes, b_gpc + b_tp + b_sf + b_apw))) return rc;
  char *base = (char *)c->stage;
  double *d_gpc = (double *)base;
  double *d_tp = (double *)(base + b_gpc);
  double *d_sf = (double *)(base + b_gpc + b_tp);
  double *d_apw = (double *)(base + b_gpc + b_tp + b_sf);
  EXC_CHECK(hipMemcpyAsync(d_gpc, gpc, b_gpc, hipMemcpyHostToDevice, c->stream));
  EXC_CHECK(hipMemcpyAsync(d_tp, tpgpc, b_tp, hipMemcpyHostToDevice, c->stream));
  EXC_CHECK(hipMemcpyAsync(d_sf, sfacgp, b_sf, hipMemcpyHostToDevice, c->stream));
  rc = exc_match_dev(c, ngp, ngkmax, d_gpc, d_tp, d_sf, d_apw);
  if (rc < 0) return rc;
  // only rows 1..ngp of every column are results; the caller's padding rows stay untouched
  if (ngp > 0) {
    int r2 = grow_host(&c->hstage, &c->hstage_bytes, b_apw);
    if (r2) return r2;
    EXC_CHECK(hipMemcpyAsync(c->hstage, d_apw, b_apw, hipMemcpyDeviceToHost, c->stream));
    EXC_CHECK(hipStreamSynchronize(c->stream));
    copy_cols((char *)apwalm, 16 * nk, (const char *)c->hstage, 16 * nk, 16 * (size_t)ngp, ao * lmm * na);
  }
  EXC_CHECK(hipStreamSynchronize(c->stream));
  return rc;
}

int exc_hamiltonandoverlapsetup_dev(exc_ctx *c, int ngp, int ngkmax, const double *apwalm, const int *igpig,
                                    const double *vgpc, int ldh, double *h, double *o) {
  if (!ready(c)) return c ? EXC_ERR_STATE : EXC_ERR_ARG;
  const int nmat = ngp + c->d.nlotot;
  if (ngp < 0 || ngkmax < ngp || ldh < nmat || !h || !o || (ngp > 0 && (!apwalm || !igpig || !vgpc)))
    return EXC_ERR_ARG;
  int rc = use_device(c);
  if (rc) return rc;
  if (ngp > 0) {
    const size_t need = sizeof(cplx) * (size_t)ngp * c->d.naa * c->d.natm;
    if ((rc = grow((void **)&c->yt, &c->yt_bytes, need))) return rc;
    // defaults from the autotuner's sweep on MI355X (64x64 MT and Y tiles, LDS filled by DMA)
    const int yv = c->y_variant >= 0 ? c->y_variant : 5;
    const int mv = c->mt_variant >= 0 ? c->mt_variant : 6;
    {
      StageTimer t(c, EXC_STAGE_YSTAGE);
      rc = exc_launch_ystage(c->d, ngp, (const cplx *)apwalm, ngkmax, c->yt, ngp, yv, c->stream);
    }
    if (rc) return rc;
    {
      StageTimer t(c, EXC_STAGE_MT);
      rc = exc_launch_mt(c->d, ngp, (const cplx *)apwalm, ngkmax, c->yt, ngp, igpig, vgpc, (cplx *)h, (cplx *)o, ldh,
                         mv, c->stream);
    }
    if (rc) return rc;
  }
  {
    StageTimer t(c, EXC_STAGE_LO);
    rc = exc_launch_lo(c->d, ngp, (const cplx *)apwalm, ngkmax, (cplx *)h, (cplx *)o, ldh, c->stream);
  }
  return rc;
}

int exc_hamiltonandoverlapsetup(exc_ctx *c, int ngp, int ngkmax, const double *apwalm, const int *igpig,
                                const double *vgpc, int ldh, double *h, double *o) {
  if (!ready(c)) return c ? EXC_ERR_STATE : EXC_ERR_ARG;
  const int nmat = ngp + c->d.nlotot;
  if (ngp < 0 || ngkmax < ngp || ldh < nmat || !h || !o || (ngp > 0 && (!apwalm || !igpig || !vgpc)))
    return EXC_ERR_ARG;
  int rc = use_device(c);
  if (rc) return rc;
  const size_t na = c->d.natm, lmm = c->d.lmmaxapw, ao = c->d.apwordmax, nk = ngkmax;
  const size_t b_apw = 16 * nk * ao * lmm * na, b_ig = 4 * nk, b_vg = 24 * nk, b_h = 16 * (size_t)ldh * nmat;
  const size_t o_ig = (b_apw + 255) & ~(size_t)255, o_vg = o_ig + ((b_ig + 255) & ~(size_t)255);
  const size_t o_h = o_vg + ((b_vg + 255) & ~(size_t)255), o_o = o_h + ((b_h + 255) & ~(size_t)255);
  if ((rc = grow(&c->stage, &c->stage_bytes, o_o + b_h))) return rc;
  if ((rc = grow_host(&c->hstage, &c->hstage_bytes, std::max(b_apw, 2 * b_h)))) return rc;
  char *base = (char *)c->stage;
  char *hs = (char *)c->hstage;
  if (ngp > 0) {   // rows 1..ngp of apwalm are the input; the caller's padding rows are never read
    copy_cols(hs, 16 * nk, (const char *)apwalm, 16 * nk, 16 * (size_t)ngp, ao * lmm * na);
    EXC_CHECK(hipMemcpyAsync(base, hs, b_apw, hipMemcpyHostToDevice, c->stream));
  }
  EXC_CHECK(hipMemcpyAsync(base + o_ig, igpig, b_ig, hipMemcpyHostToDevice, c->stream));
  EXC_CHECK(hipMemcpyAsync(base + o_vg, vgpc, b_vg, hipMemcpyHostToDevice, c->stream));
  rc = exc_hamiltonandoverlapsetup_dev(c, ngp, ngkmax, (const double *)base, (const int *)(base + o_ig),
                                       (const double *)(base + o_vg), ldh, (double *)(base + o_h),
                                       (double *)(base + o_o));
  if (rc) return rc;
  // rows nmat+1..ldh of the caller's matrices are not part of the result: copy rows 1..nmat
  const size_t pitch = 16 * (size_t)ldh, width = 16 * (size_t)nmat;
  EXC_CHECK(hipMemcpyAsync(hs, base + o_h, b_h, hipMemcpyDeviceToHost, c->stream));
  EXC_CHECK(hipMemcpyAsync(hs + b_h, base + o_o, b_h, hipMemcpyDeviceToHost, c->stream));
  EXC_CHECK(hipStreamSynchronize(c->stream));
  copy_cols((char *)h, pitch, hs, pitch, width, nmat);
  copy_cols((char *)o, pitch, hs + b_h, pitch, width, nmat);
  return EXC_OK;
}

int exc_setup_batched_dev(exc_ctx *c, int nk, const exc_kpoint_dev *ks, int flags) {
  if (!ready(c)) return c ? EXC_ERR_STATE : EXC_ERR_ARG;
  if (nk < 0 || (nk > 0 && !ks) ||
      (flags & ~(EXC_BATCH_MATCH | EXC_BATCH_GRAPH | EXC_BATCH_SPLIT_BF16 | EXC_BATCH_OLP_LEGENDRE)))
    return EXC_ERR_ARG;
  const bool split = flags & EXC_BATCH_SPLIT_BF16;
  // the addition-theorem overlap needs the radial factors that only the grouped match writes
  const bool olp = flags & EXC_BATCH_OLP_LEGENDRE;
  if (olp && !(flags & EXC_BATCH_MATCH)) return EXC_ERR_ARG;
  if (split && c->d.naa > kSplitKB) return EXC_ERR_UNSUPPORTED;
  if (nk == 0) return EXC_OK;
  int rc = use_device(c);
  if (rc) return rc;
  const bool do_match = flags & EXC_BATCH_MATCH;
  for (int k = 0; k < nk; ++k) {
    const exc_kpoint_dev &q = ks[k];
    if (q.ngp < 0 || q.ngkmax < q.ngp || q.ldh < q.ngp + c->d.nlotot || !q.h || !q.o) return EXC_ERR_ARG;
    if (q.ngp > 0 && (!q.apwalm || !q.igpig || !q.vgpc)) return EXC_ERR_ARG;
    if (do_match && q.ngp > 0 && (!q.gpc || !q.tpgpc || !q.sfacgp)) return EXC_ERR_ARG;
  }
  // defaults: the fastest variants measured on MI355X for N = 1150..5000 (grouped MT: 64x64
  // tile, 8 waves, two LDS stages filled by DMA, one item per workgroup; split: 64x64 tile,
  // LDS filled by DMA)
  const int bv = split ? (c->bsplit_variant >= 0 ? c->bsplit_variant : 3) : (c->bmt_variant >= 0 ? c->bmt_variant : 18);
  const int yv = c->y_variant >= 0 ? c->y_variant : 8;     // 64x128 Y tiles, DMA-staged
  // batch signature: descriptors, flags and variants
  std::vector<char> sig(sizeof(exc_kpoint_dev) * nk + 4 * sizeof(int));
  memcpy(sig.data(), ks, sizeof(exc_kpoint_dev) * nk);
  int tail[4] = {nk, flags, bv, yv};
  memcpy(sig.data() + sizeof(exc_kpoint_dev) * nk, tail, sizeof(tail));
  const bool same = (sig == c->batch_sig);
  if (same && (flags & EXC_BATCH_GRAPH) && c->graph_exec) {
    StageTimer t(c, EXC_STAGE_GRAPH);
    EXC_CHECK(hipGraphLaunch(c->graph_exec, c->stream));
    return EXC_OK;
  }
  if (!same) {
    // new batch: lay out the Y workspace, descriptors and the muffin-tin work list
    EXC_CHECK(hipStreamSynchronize(c->stream));   // earlier launches may still read the old tables
    if (c->graph_exec) {
      hipGraphExecDestroy(c->graph_exec);
      c->graph_exec = nullptr;
    }
    if (c->graph) {
      hipGraphDestroy(c->graph);
      c->graph = nullptr;
    }
    size_t ytot = 0, ptot = 0, etot = 0, rtot = 0;
    const int nl2 = 2 * (c->d.lmaxmat + 1);
    for (int k = 0; k < nk; ++k) {
      ytot += (size_t)ks[k].ngp * c->d.naa * c->d.natm;
      rtot += (size_t)ks[k].ngp * c->d.natm * nl2;
      ptot += 2 * exc_split_plane_elems(ks[k].ngp, c->d.natm);
      etot += 2 * (size_t)ks[k].ngp * c->d.natm;
    }
    if ((rc = grow((void **)&c->byt, &c->byt_bytes, sizeof(cplx) * (ytot ? ytot : 1)))) return rc;
    if (split) {
      if ((rc = grow((void **)&c->planes, &c->planes_bytes, 2 * (ptot ? ptot : 1)))) return rc;
      if ((rc = grow((void **)&c->expo, &c->expo_bytes, sizeof(int) * (etot ? etot : 1)))) return rc;
    }
    if (olp && (rc = grow((void **)&c->rfbuf, &c->rf_bytes, sizeof(double) * (rtot ? rtot : 1)))) return rc;
    std::vector<kdesc> kd(nk);
    std::vector<mt_item> items, oitems;
    const int T = split ? exc_bsplit_tile(bv) : exc_bmt_tile(bv);
    const int TO = exc_olp_tile();
    size_t yoff = 0, poff = 0, eoff = 0, roff = 0;
    int maxngp = 0;
    for (int k = 0; k < nk; ++k) {
      const exc_kpoint_dev &q = ks[k];
      kdesc &e = kd[k];
      e.ngp = q.ngp;
      e.ldg = q.ngkmax;
      e.ldh = q.ldh;
      e.ldy = q.ngp;
      e.gpc = q.gpc;
      e.tpgpc = q.tpgpc;
      e.sfacgp = (const cplx *)q.sfacgp;
      e.apwalm = (cplx *)q.apwalm;
      e.igpig = q.igpig;
      e.vgpc = q.vgpc;
      e.yt = c->byt + yoff;
      e.h = (cplx *)q.h;
      e.o = (cplx *)q.o;
      yoff += (size_t)q.ngp * c->d.naa * c->d.natm;
      e.sa = e.sy = nullptr;
      e.ea = e.ey = nullptr;
      if (split) {
        const size_t pe = exc_split_plane_elems(q.ngp, c->d.natm);
        e.sa = c->planes + poff;
        e.sy = c->planes + poff + pe;
        poff += 2 * pe;
        e.ea = c->expo + eoff;
        e.ey = c->expo + eoff + (size_t)q.ngp * c->d.natm;
        eoff += 2 * (size_t)q.ngp * c->d.natm;
      }
      e.rf = nullptr;
      if (olp) {
        e.rf = c->rfbuf + roff;
        roff += (size_t)q.ngp * c->d.natm * nl2;
      }
      maxngp = q.ngp > maxngp ? q.ngp : maxngp;
      const int nt = (q.ngp + T - 1) / T;
      const bool fused = !split && exc_bmt_fused(bv);
      // with the addition-theorem overlap the MT work list holds the H tiles only (a fused
      // variant still computes both; its O is then overwritten)
      for (int tj = 0; tj < nt; ++tj)
        for (int ti = 0; ti <= tj; ++ti) {
          items.push_back({k, ti, tj, 0});
          if (!fused && !olp) items.push_back({k, ti, tj, 1});
        }
      if (olp) {
        const int no = (q.ngp + TO - 1) / TO;
        for (int tj = 0; tj < no; ++tj)
          for (int ti = 0; ti <= tj; ++ti) oitems.push_back({k, ti, tj, 1});
      }
    }
    if (olp) {
      if ((rc = grow((void **)&c->d_oitems, &c->oitems_cap, sizeof(mt_item) * (oitems.empty() ? 1 : oitems.size()))))
        return rc;
      if (!oitems.empty())
        EXC_CHECK(hipMemcpy(c->d_oitems, oitems.data(), sizeof(mt_item) * oitems.size(), hipMemcpyHostToDevice));
    }
    c->batch_noitems = (int)oitems.size();
    if ((rc = grow((void **)&c->d_ks, &c->ks_cap, sizeof(kdesc) * nk))) return rc;
    if ((rc = grow((void **)&c->d_items, &c->items_cap, sizeof(mt_item) * (items.empty() ? 1 : items.size()))))
      return rc;
    EXC_CHECK(hipMemcpy(c->d_ks, kd.data(), sizeof(kdesc) * nk, hipMemcpyHostToDevice));
    if (!items.empty())
      EXC_CHECK(hipMemcpy(c->d_items, items.data(), sizeof(mt_item) * items.size(), hipMemcpyHostToDevice));
    c->batch_nitems = (int)items.size();
    int nb = c->ncu * (split ? exc_bsplit_max_blocks(bv) : exc_bmt_max_blocks(bv));
    nb = nb < c->batch_nitems ? nb : c->batch_nitems;
    if (nb >= 8) nb -= nb % 8;
    c->batch_nblocks = nb > 0 ? nb : 1;
    c->batch_maxngp = maxngp;
    c->batch_nk = nk;
    c->batch_sig = sig;
  }
  const bool capture = flags & EXC_BATCH_GRAPH;
  if (capture && c->stream == nullptr) return EXC_ERR_ARG;   // the null stream cannot be captured
  // inside a graph the overlap and lo kernels (they read only A, the radial factors and the
  // structure factors, and write parts of H and O the MT stage does not) form a branch of
  // their own beside the Y and MT stages; with a persistent MT kernel (static work split,
  // all workgroups assumed resident) the branches join before it.  Eager calls keep one
  // stream so the stage timers stay exact
  const bool fork = capture;
  const bool join_early = split || exc_bmt_persistent(bv);
  if (fork && !c->side) {
    EXC_CHECK(hipStreamCreateWithFlags(&c->side, hipStreamNonBlocking));
    EXC_CHECK(hipEventCreateWithFlags(&c->ev_fork, hipEventDisableTiming));
    EXC_CHECK(hipEventCreateWithFlags(&c->ev_join, hipEventDisableTiming));
  }
  hipStream_t s2 = fork ? c->side : c->stream;
  if (capture) EXC_CHECK(hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
  auto launches = [&]() -> int {
    int r = EXC_OK;
    if (do_match) {
      StageTimer t(c, capture ? -1 : EXC_STAGE_MATCH);
      if ((r = exc_launch_bmatch(c->d, nk, c->batch_maxngp, c->d_ks, c->stream))) return r;
    }
    if (fork) {
      if (hipEventRecord(c->ev_fork, c->stream) != hipSuccess) return EXC_ERR_HIP;
      if (hipStreamWaitEvent(c->side, c->ev_fork, 0) != hipSuccess) return EXC_ERR_HIP;
      if (olp && (r = exc_launch_olp(c->d, c->d_ks, c->d_oitems, c->batch_noitems, s2))) return r;
      if ((r = exc_launch_blo(c->d, nk, c->batch_maxngp, c->d_ks, s2))) return r;
      if (hipEventRecord(c->ev_join, c->side) != hipSuccess) return EXC_ERR_HIP;
    }
    {
      StageTimer t(c, capture ? -1 : EXC_STAGE_YSTAGE);
      if ((r = exc_launch_bystage(c->d, nk, c->batch_maxngp, c->d_ks, yv, c->stream))) return r;
    }
    if (split) {
      {
        StageTimer t(c, capture ? -1 : EXC_STAGE_SPLIT);
        if ((r = exc_launch_split(c->d, nk, c->batch_maxngp, c->d_ks, c->stream))) return r;
      }
      if (fork && hipStreamWaitEvent(c->stream, c->ev_join, 0) != hipSuccess) return EXC_ERR_HIP;
      StageTimer t(c, capture ? -1 : EXC_STAGE_BSPLIT);
      if ((r = exc_launch_bsplit(c->d, c->d_ks, c->d_items, c->batch_nitems, bv, c->batch_nblocks, c->stream)))
        return r;
    } else {
      if (fork && join_early && hipStreamWaitEvent(c->stream, c->ev_join, 0) != hipSuccess) return EXC_ERR_HIP;
      StageTimer t(c, capture ? -1 : EXC_STAGE_BMT);
      if ((r = exc_launch_bmt(c->d, c->d_ks, c->d_items, c->batch_nitems, bv, c->batch_nblocks, c->stream))) return r;
    }
    if (fork && !join_early && hipStreamWaitEvent(c->stream, c->ev_join, 0) != hipSuccess) return EXC_ERR_HIP;
    if (fork) return r;
    if (olp) {
      StageTimer t(c, EXC_STAGE_OLP);
      if ((r = exc_launch_olp(c->d, c->d_ks, c->d_oitems, c->batch_noitems, c->stream))) return r;
    }
    {
      StageTimer t(c, EXC_STAGE_LO);
      if ((r = exc_launch_blo(c->d, nk, c->batch_maxngp, c->d_ks, c->stream))) return r;
    }
    return r;
  };
  rc = launches();
  if (capture) {
    hipGraph_t g = nullptr;
    hipError_t e = hipStreamEndCapture(c->stream, &g);
    if (rc) {
      if (g) hipGraphDestroy(g);
      return rc;
    }
    if (e != hipSuccess) return EXC_ERR_HIP;
    c->graph = g;
    EXC_CHECK(hipGraphInstantiate(&c->graph_exec, c->graph, nullptr, nullptr, 0));
    StageTimer t(c, EXC_STAGE_GRAPH);
    EXC_CHECK(hipGraphLaunch(c->graph_exec, c->stream));
  }
  return rc;
}

int exc_ho_vjp_dev(exc_ctx *c, int ngp, int ngkmax, const double *apwalm, const int *igpig, const double *vgpc,
                   const double *sfacgp, int nst, int ldc, const double *evec, const double *wh, const double *wo,
                   double *dedtau, double *dedv) {
  if (!ready(c)) return c ? EXC_ERR_STATE : EXC_ERR_ARG;
  if (!c->d.ffacg) return EXC_ERR_STATE;
  const int nmat = ngp + c->d.nlotot;
  if (ngp < 0 || ngkmax < ngp || nst < 0 || ldc < nmat || !dedtau || !dedv) return EXC_ERR_ARG;
  if (ngp > 0 && nst > 0 && (!apwalm || !igpig || !vgpc || !sfacgp || !evec || !wh || !wo)) return EXC_ERR_ARG;
  int rc = use_device(c);
  if (rc) return rc;
  const size_t need = sizeof(cplx) * exc_vjp_work_elems(c->d, ngp, nst);
  if ((rc = grow((void **)&c->vjp_work, &c->vjp_bytes, need ? need : 1))) return rc;
  StageTimer t(c, EXC_STAGE_VJP);
  return exc_launch_vjp(c->d, ngp, nst, (const cplx *)apwalm, ngkmax, igpig, vgpc, (const cplx *)sfacgp,
                        (const cplx *)evec, ldc, wh, wo, c->vjp_work, dedtau, dedv, c->stream);
}

int exc_ho_vjp(exc_ctx *c, int ngp, int ngkmax, const double *apwalm, const int *igpig, const double *vgpc,
               const double *sfacgp, int nst, int ldc, const double *evec, const double *wh, const double *wo,
               double *dedtau, double *dedv) {
  if (!ready(c)) return c ? EXC_ERR_STATE : EXC_ERR_ARG;
  const int nmat = ngp + c->d.nlotot;
  if (ngp < 0 || ngkmax < ngp || nst < 0 || ldc < nmat || !dedtau) return EXC_ERR_ARG;
  if (ngp > 0 && nst > 0 && (!apwalm || !igpig || !vgpc || !sfacgp || !evec || !wh || !wo)) return EXC_ERR_ARG;
  int rc = use_device(c);
  if (rc) return rc;
  const size_t na = c->d.natm, nk = ngkmax;
  auto al = [](size_t b) { return (b + 255) & ~(size_t)255; };
  const size_t b_apw = 16 * nk * c->d.apwordmax * c->d.lmmaxapw * na, b_ig = 4 * nk, b_vg = 24 * nk;
  const size_t b_sf = 16 * nk * na, b_ev = 16 * (size_t)ldc * nst, b_w = 8 * (size_t)nst;
  const size_t b_dt = 8 * 3 * na, b_dv = 16 * (size_t)c->d.ngvec;
  size_t off[9];
  off[0] = 0;
  const size_t sz[9] = {b_apw, b_ig, b_vg, b_sf, b_ev, b_w, b_w, b_dt, b_dv};
  for (int i = 1; i < 9; ++i) off[i] = off[i - 1] + al(sz[i - 1]);
  if ((rc = grow(&c->stage, &c->stage_bytes, off[8] + b_dv))) return rc;
  char *base = (char *)c->stage;
  const void *src[7] = {apwalm, igpig, vgpc, sfacgp, evec, wh, wo};
  for (int i = 0; i < 7; ++i)
    if (sz[i]) EXC_CHECK(hipMemcpyAsync(base + off[i], src[i], sz[i], hipMemcpyHostToDevice, c->stream));
  EXC_CHECK(hipMemsetAsync(base + off[7], 0, b_dt + (off[8] - off[7] - b_dt) + b_dv, c->stream));
  rc = exc_ho_vjp_dev(c, ngp, ngkmax, (const double *)(base + off[0]), (const int *)(base + off[1]),
                      (const double *)(base + off[2]), (const double *)(base + off[3]), nst, ldc,
                      (const double *)(base + off[4]), (const double *)(base + off[5]),
                      (const double *)(base + off[6]), (double *)(base + off[7]), (double *)(base + off[8]));
  if (rc) return rc;
  std::vector<double> dt(3 * na), dv(dedv ? 2 * (size_t)c->d.ngvec : 0);
  EXC_CHECK(hipMemcpyAsync(dt.data(), base + off[7], b_dt, hipMemcpyDeviceToHost, c->stream));
  if (dedv) EXC_CHECK(hipMemcpyAsync(dv.data(), base + off[8], b_dv, hipMemcpyDeviceToHost, c->stream));
  EXC_CHECK(hipStreamSynchronize(c->stream));
  for (size_t i = 0; i < dt.size(); ++i) dedtau[i] += dt[i];
  for (size_t i = 0; i < dv.size(); ++i) dedv[i] += dv[i];
  return EXC_OK;
}

}  // extern "C"
