// Private to the library: the context behind the C-ABI handle and the helpers the entry-point
// files share (device allocation and uploads owned by the context, grow-only buffers, the
// stage timers).  Not part of the C-ABI.
#ifndef EXC_CTX_H
#define EXC_CTX_H

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
  cplx *apwo;                 // per-k 'iora*' overlap operand A L (exc_launch_iora_apw)
  size_t apwo_bytes;
  // staging for the host-pointer entry points: device buffers and a pinned host buffer
  void *stage;
  size_t stage_bytes;
  // apwalm of the last host-array exc_match, kept on the device for the setup calls that
  // pass apwalm = NULL (match and the setup of one k-point without a second PCIe copy)
  cplx *apw_res;
  size_t apw_res_bytes;
  int apw_res_ngp, apw_res_ngkmax;
  bool apw_res_valid;
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
  std::vector<hipEvent_t> copy_ev;   // chunked device-to-host copies (d2h_cols)
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
  // modular split (EXC_BATCH_SPLIT_CRT): moduli count, GEMM variant, residue planes, row scale
  // exponents, residues of every (tile, modulus), the tile and (tile, modulus) work lists
  int crt_nmod, crt_variant;
  int olp_variant;          // thread blocking of the addition-theorem overlap (-1: built-in)
  int fork_variant;         // grouped graph: 0 overlap + lo on a side branch, 1 one chain (-1: 1)
  void *crt_q;
  size_t crt_q_bytes;
  int *crt_x;
  size_t crt_x_bytes;
  double *crt_parts;          // squared row norms of A per atom (modular split)
  size_t crt_parts_bytes;
  void *gk_buf;               // G+k sets: k-point coordinates, counts, and the host entries' staging
  size_t gk_buf_bytes;
  double cfdamp_t1;           // gencfun's damping exponent cfdamp / gmaxvr (0: none)
  // Gram form of the modular split: the table conj(U) (Lambda + sigma)^(1/2) per atom, the
  // eigenvectors and values of h_a, sigma (host and device); valid until h_a changes
  int crt_form;               // -1 built-in (Gram where it applies), 0 A and Yt, 1 Gram
  int crt_last_form, crt_last_nmod;   // what the last modular-split setup used
  void *gram_buf;
  size_t gram_bytes;
  rocblas_handle gram_blas;   // its own handle: the eigensolve workers pin their workspaces
  cplx *gram_table;           // in gram_buf: the Y stage's table conj(U) (Lambda + sigma)^(1/2)
  double *gram_sig_dev;       // in gram_buf: sigma for the merge
  bool gram_valid;
  double gram_sigma, gram_lmax;
  bool gram_tri;              // the table is the triangular (Cholesky) one
  signed char *crt_res;
  size_t crt_res_bytes;
  mt_item *d_citems;
  size_t citems_cap;
  int2 *d_cwork;
  size_t cwork_cap;
  int batch_ncitems, batch_ncwork, batch_maxcnp;
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
    void *ws;               // the rocBLAS handle's workspace, owned here (rocblas_set_workspace)
    size_t wsb;
    bool ws_set;
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
  void *rvw;                // rhovalk: A^T C of one k-point
  size_t rvw_bytes;
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

namespace exc_impl {

template <typename T>
inline int dev_alloc(exc_ctx *c, T **p, size_t n) {
  *p = nullptr;
  if (n == 0) return EXC_OK;
  void *q = nullptr;
  EXC_CHECK(hipMalloc(&q, n * sizeof(T)));
  c->owned.push_back(q);
  *p = (T *)q;
  return EXC_OK;
}

template <typename T>
inline int dev_upload(exc_ctx *c, T **p, const void *src, size_t n) {
  if (*p == nullptr) {
    int rc = dev_alloc(c, p, n);
    if (rc) return rc;
  }
  if (n) EXC_CHECK(hipMemcpyAsync(*p, src, n * sizeof(T), hipMemcpyHostToDevice, c->stream));
  return EXC_OK;
}

// pinned host staging (grow-only); every host entry synchronises before it returns, so no
// copy into the old buffer is in flight when it is replaced
inline int grow_host(void **p, size_t *have, size_t need) {
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
inline void copy_cols(char *dst, size_t dld, const char *src, size_t sld, size_t rowbytes, size_t ncols) {
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

// Device columns (leading dimension sld bytes, rowbytes of each wanted) to the caller's host
// columns (leading dimension dld) through the pinned staging buffer `stage` (ncols * sld
// bytes), in chunks of about 16 MiB: every chunk's DMA is queued with an event after it, and
// the host copy of a chunk runs as soon as its event has passed, while the later chunks'
// DMAs are still in flight.  Returns after the last host copy.
inline int d2h_cols(exc_ctx *c, char *dst, size_t dld, const char *dsrc, size_t sld, size_t rowbytes, size_t ncols,
                    char *stage) {
  constexpr size_t kMaxChunks = 64;
  if (ncols == 0) return EXC_OK;
  size_t per = std::max<size_t>(1, ((size_t)16 << 20) / sld);
  per = std::max(per, (ncols + kMaxChunks - 1) / kMaxChunks);
  const size_t nch = (ncols + per - 1) / per;
  while (c->copy_ev.size() < nch) {
    hipEvent_t e;
    EXC_CHECK(hipEventCreateWithFlags(&e, hipEventDisableTiming));
    c->copy_ev.push_back(e);
  }
  for (size_t j = 0; j < nch; ++j) {
    const size_t a = j * per, b = std::min(ncols, a + per);
    EXC_CHECK(hipMemcpyAsync(stage + a * sld, dsrc + a * sld, (b - a) * sld, hipMemcpyDeviceToHost, c->stream));
    EXC_CHECK(hipEventRecord(c->copy_ev[j], c->stream));
  }
  for (size_t j = 0; j < nch; ++j) {
    const size_t a = j * per, b = std::min(ncols, a + per);
    EXC_CHECK(hipEventSynchronize(c->copy_ev[j]));
    copy_cols(dst + a * dld, dld, stage + a * sld, sld, rowbytes, b - a);
  }
  return EXC_OK;
}

// A larger buffer replaces the old one; work already queued on any stream may still read or
// write the old one, so the device drains before it is freed.
inline int grow(void **p, size_t *have, size_t need) {
  if (need <= *have) return EXC_OK;
  if (*p) {
    EXC_CHECK(hipDeviceSynchronize());
    EXC_CHECK(hipFree(*p));
  }
  *p = nullptr;
  *have = 0;
  EXC_CHECK(hipMalloc(p, need));
  *have = need;
  return EXC_OK;
}

inline int use_device(exc_ctx *c) {
  int cur = -1;
  EXC_CHECK(hipGetDevice(&cur));
  if (cur != c->device) EXC_CHECK(hipSetDevice(c->device));
  return EXC_OK;
}

inline bool ready(const exc_ctx *c) { return c && c->have_atoms && c->have_lo && c->have_radial && c->have_mt && c->have_gvec; }

inline hipEvent_t take_event(exc_ctx *c) {
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

}  // namespace exc_impl

#endif
