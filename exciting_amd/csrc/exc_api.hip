// C-ABI of the gfx950 LAPW H+O setup (include/exc_lapw.h): context management, SCF-data
// uploads, match, the per-k and grouped setup, the adjoint and the stage timers.  The entry
// points of the eigensolve, the density matrix and the per-SCF producers are in
// exc_api_eigen.hip, exc_api_density.hip and exc_api_scf.hip.
#include <mutex>

#include "exc_ctx.h"

using namespace exc_impl;

// Fixed device workspace of exc_hamiltonandoverlapsetup_noalloc_dev (the Y stage's
// Yt = A h^T, ngp x natm naa complex), for callers that cannot allocate.  In host code
// `scratch` names a host-side stand-in, not device memory: exc_ctx_create takes its device
// address once per device, with that device current (hipGetSymbolAddress), and only that
// address is handed to kernels.  Every GPU the library runs on holds its own copy until the
// process ends.
static constexpr size_t kScratchBytes = (size_t)1 << 30;   // 1 GiB: ngp * natm * naa <= 64 Mi complex
static __device__ __align__(256) unsigned char scratch[kScratchBytes];

namespace {

constexpr int kMaxScratchDevices = 64;
std::mutex g_scratch_lock;
void *g_scratch[kMaxScratchDevices];   // device address of `scratch` per device (null: not taken yet)

int register_scratch(int device) {
  if (device < 0 || device >= kMaxScratchDevices) return EXC_ERR_UNSUPPORTED;
  std::lock_guard<std::mutex> g(g_scratch_lock);
  if (g_scratch[device]) return EXC_OK;
  void *p = nullptr;
  const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
  if (e != hipSuccess || !p) {
    fprintf(stderr, "exc_lapw: hipGetSymbolAddress(scratch) failed on device %d: %s\n", device,
            hipGetErrorString(e));
    return EXC_ERR_HIP;
  }
  g_scratch[device] = p;
  return EXC_OK;
}

void *scratch_of(int device) {
  std::lock_guard<std::mutex> g(g_scratch_lock);
  return device >= 0 && device < kMaxScratchDevices ? g_scratch[device] : nullptr;
}

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
  if (int rc = register_scratch(c->device)) {
    delete c;
    return rc;
  }
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
    const unsigned zero = 0u;
    if (!rc) rc = dev_upload(c, &d.miss, &zero, 1);
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
  c->crt_variant = -1;
  c->crt_nmod = 0;                 // built-in: 10 moduli with the Gram form, 11 without
  c->olp_variant = -1;
  c->fork_variant = -1;
  c->crt_form = -1;
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
  if (c->apwo) hipFree(c->apwo);
  for (int i = 0; i <= exc_ctx::kEigStreams; ++i) {
    exc_ctx::EigWorker &w = c->ew[i];
    if (w.blas) rocblas_destroy_handle(w.blas);
    if (w.w) hipFree(w.w);
    if (w.z) hipFree(w.z);
    if (w.zc) hipFree(w.zc);
    if (w.e) hipFree(w.e);
    if (w.ws) hipFree(w.ws);
    if (w.done) hipEventDestroy(w.done);
    if (i > 0 && w.s) hipStreamDestroy(w.s);
  }
  if (c->eig_int) hipFree(c->eig_int);
  if (c->eg) hipFree(c->eg);
  if (c->dmw) hipFree(c->dmw);
  if (c->rvw) hipFree(c->rvw);
  if (c->gnt) hipFree(c->gnt);
  if (c->have_fft) hipfftDestroy(c->fft);
  if (c->fftb_n > 0) hipfftDestroy(c->fftb);
  if (c->zb) hipFree(c->zb);
  if (c->hml) hipFree(c->hml);
  if (c->eig_fork) hipEventDestroy(c->eig_fork);
  if (c->side) hipStreamDestroy(c->side);
  if (c->ev_fork) hipEventDestroy(c->ev_fork);
  if (c->ev_join) hipEventDestroy(c->ev_join);
  for (hipEvent_t e : c->copy_ev) hipEventDestroy(e);
  if (c->stage) hipFree(c->stage);
  if (c->apw_res) hipFree(c->apw_res);
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
  if (c->crt_q) hipFree(c->crt_q);
  if (c->crt_x) hipFree(c->crt_x);
  if (c->crt_parts) hipFree(c->crt_parts);
  if (c->gram_buf) hipFree(c->gram_buf);
  if (c->gram_blas) rocblas_destroy_handle(c->gram_blas);
  if (c->gk_buf) hipFree(c->gk_buf);
  if (c->crt_res) hipFree(c->crt_res);
  if (c->d_citems) hipFree(c->d_citems);
  if (c->d_cwork) hipFree(c->d_cwork);
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

int exc_ctx_nlotot(const exc_ctx *c) { return c ? c->d.nlotot : EXC_ERR_ARG; }

int exc_ctx_status(exc_ctx *c) {
  if (!c) return EXC_ERR_ARG;
  int rc = use_device(c);
  if (rc) return rc;
  EXC_CHECK(hipStreamSynchronize(c->stream));
  unsigned miss = 0;
  EXC_CHECK(hipMemcpy(&miss, c->d.miss, sizeof(unsigned), hipMemcpyDeviceToHost));
  if (!miss) return EXC_OK;
  const unsigned zero = 0u;
  EXC_CHECK(hipMemcpy(c->d.miss, &zero, sizeof(unsigned), hipMemcpyHostToDevice));
  fprintf(stderr,
          "exc_lapw: %u pair(s) of basis vectors with G_i - G_j outside the first ngvec G-vectors "
          "(2 gkmax > gmaxvr, init1.f90:340-347)\n", miss);
  return EXC_ERR_ARG;
}

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
    case EXC_STAGE_CRTGEMM:
      if (variant < -1 || variant >= kCrtVariants) return EXC_ERR_ARG;
      c->crt_variant = variant;
      return EXC_OK;
    case EXC_STAGE_OLP:
      if (variant < -1 || variant >= kOlpVariants) return EXC_ERR_ARG;
      c->olp_variant = variant;
      return EXC_OK;
    case EXC_STAGE_GRAPH:
      if (variant < -1 || variant >= 2) return EXC_ERR_ARG;
      c->fork_variant = variant;
      return EXC_OK;
    case EXC_STAGE_CRTFORM:
      if (variant < -1 || variant >= 2) return EXC_ERR_ARG;
      c->crt_form = variant;
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
    case EXC_STAGE_CRTGEMM: return kCrtVariants;
    case EXC_STAGE_OLP: return kOlpVariants;
    case EXC_STAGE_GRAPH: return 2;   // the overlap and lo kernels on a side branch, or in line
    case EXC_STAGE_CRTFORM: return 2; // the modular split's operands: A and Yt, or the Gram form
    default: return 0;
  }
}

int exc_ctx_set_moduli(exc_ctx *c, int n) {
  if (!c || (n != 0 && (n < 8 || n > kCrtMaxUsed))) return EXC_ERR_ARG;
  c->crt_nmod = n;
  return EXC_OK;
}

int exc_ctx_crt_form(const exc_ctx *c, int *form, int *nmod, double *sigma) {
  if (!c || !form || !nmod || !sigma) return EXC_ERR_ARG;
  *form = c->crt_last_form;
  *nmod = c->crt_last_nmod;
  *sigma = c->crt_last_form == 1 ? c->gram_sigma : 0.0;
  return EXC_OK;
}

int exc_crt_moduli(int n, int *moduli, int *beta) {
  if (n < 8 || n > kCrtMaxUsed || !moduli || !beta) return EXC_ERR_ARG;
  CrtTab t;
  exc_crt_tables(n, &t);
  for (int l = 0; l < n; ++l) moduli[l] = t.m[l];
  *beta = t.beta;
  return EXC_OK;
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
  // the tables below are replaced by synchronous copies on the null stream, which does not wait
  // for a non-blocking context stream: queued work that reads them completes first
  EXC_CHECK(hipStreamSynchronize(c->stream));
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
        // zgesv stops after the factorization with info > 0 and leaves the right-hand sides
        // unsolved; match.f90:166-185 prints the error and stores them as the coefficients,
        // so the inverse is taken as the identity here
        if (!info) info = 1 + l + (L + 1) * ias;
        o[0] = 1.0;
        o[1] = o[2] = 0.0;
        o[3] = 1.0;
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
  c->gram_valid = false;
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

int exc_ctx_set_meffig(exc_ctx *c, const double *meffig) {
  if (!c) return EXC_ERR_ARG;
  int rc = use_device(c);
  if (rc) return rc;
  if (meffig) {
    if ((rc = dev_upload(c, &c->d.meffig, meffig, (size_t)c->dims.ngvec))) return rc;
    EXC_CHECK(hipStreamSynchronize(c->stream));
  }
  c->d.zora = meffig ? 1 : 0;
  c->batch_sig.clear();        // a captured graph holds the tables' addresses: capture again
  return EXC_OK;
}

int exc_ctx_set_iora(exc_ctx *c, const double *h1aa, const double *h1loa, const double *h1lolo,
                     const double *m2effig) {
  if (!c) return EXC_ERR_ARG;
  if (!c->have_atoms) return EXC_ERR_STATE;
  const int na = c->dims.natmtot, ao = c->dims.apwordmax, la = c->dims.lmaxapw, nl = c->d.lmaxmat + 1;
  const int nlo = c->dims.nlomax;
  if (nlo > 0 && (!h1loa != !h1lolo)) return EXC_ERR_ARG;
  int rc = use_device(c);
  if (rc) return rc;
  int iora = 0;
  if (h1aa) {
    // I + h1aa(:,:,l,a) = L L^T per atom and l <= lmaxmat (the l-range of overlapsetup.f90:60-68)
    std::vector<double> lf((size_t)na * nl * 3);
    for (int ias = 0; ias < na; ++ias)
      for (int l = 0; l < nl; ++l) {
        const double *h = h1aa + (size_t)ao * ao * (l + (size_t)(la + 1) * ias);   // h(io1 + ao io2)
        const double m00 = 1.0 + h[0], m11 = 1.0 + h[3], m10 = h[1], m01 = h[2];
        if (!std::isfinite(m00) || !std::isfinite(m11) || !std::isfinite(m10) || !std::isfinite(m01))
          return EXC_ERR_ARG;
        // olprad makes h1aa symmetric (the same radial integrand for (io1, io2) and (io2, io1))
        if (!(fabs(m10 - m01) <= 1e-12 * std::max(1.0, std::max(fabs(m10), fabs(m01))))) return EXC_ERR_ARG;
        if (!(m00 > 0.0)) return EXC_ERR_UNSUPPORTED;
        const double l00 = sqrt(m00), l10 = 0.5 * (m10 + m01) / l00, dd = m11 - l10 * l10;
        if (!(dd > 0.0)) return EXC_ERR_UNSUPPORTED;   // I + h1aa not positive definite
        double *q = lf.data() + ((size_t)ias * nl + l) * 3;
        q[0] = l00;
        q[1] = l10;
        q[2] = sqrt(dd);
      }
    if ((rc = dev_upload(c, &c->d.olpl, lf.data(), lf.size()))) return rc;
    iora |= 1;
  }
  if (nlo > 0 && h1loa) {
    if ((rc = dev_upload(c, &c->d.h1loa, h1loa, (size_t)ao * nlo * na))) return rc;
    if ((rc = dev_upload(c, &c->d.h1lolo, h1lolo, (size_t)nlo * nlo * na))) return rc;
    iora |= 2;
  }
  if (m2effig) {
    if ((rc = dev_upload(c, &c->d.m2effig, m2effig, (size_t)c->dims.ngvec))) return rc;
    iora |= 4;
  }
  EXC_CHECK(hipStreamSynchronize(c->stream));
  c->d.iora = iora;
  c->batch_sig.clear();
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
  c->apw_res_valid = false;
  if ((rc = grow(&c->stage, &c->stage_bytes, b_gpc + b_tp + b_sf))) return rc;
  if ((rc = grow((void **)&c->apw_res, &c->apw_res_bytes, b_apw ? b_apw : 1))) return rc;
  char *base = (char *)c->stage;
  double *d_gpc = (double *)base;
  double *d_tp = (double *)(base + b_gpc);
  double *d_sf = (double *)(base + b_gpc + b_tp);
  double *d_apw = (double *)c->apw_res;
  EXC_CHECK(hipMemcpyAsync(d_gpc, gpc, b_gpc, hipMemcpyHostToDevice, c->stream));
  EXC_CHECK(hipMemcpyAsync(d_tp, tpgpc, b_tp, hipMemcpyHostToDevice, c->stream));
  EXC_CHECK(hipMemcpyAsync(d_sf, sfacgp, b_sf, hipMemcpyHostToDevice, c->stream));
  rc = exc_match_dev(c, ngp, ngkmax, d_gpc, d_tp, d_sf, d_apw);
  if (rc < 0) return rc;
  // only rows 1..ngp of every column are results; the caller's padding rows stay untouched
  if (ngp > 0) {
    int r2 = grow_host(&c->hstage, &c->hstage_bytes, b_apw);
    if (r2) return r2;
    if ((r2 = d2h_cols(c, (char *)apwalm, 16 * nk, (const char *)d_apw, 16 * nk, 16 * (size_t)ngp, ao * lmm * na,
                       (char *)c->hstage)))
      return r2;
  }
  EXC_CHECK(hipStreamSynchronize(c->stream));
  c->apw_res_valid = true;
  c->apw_res_ngp = ngp;
  c->apw_res_ngkmax = ngkmax;
  return rc;
}

}  // extern "C"

namespace {

// H (h given), O (o given) or both of one k-point from device arrays
// ws: the Y stage's workspace; NULL takes the context's grow-only buffer.  With a caller's
// workspace nothing is allocated and no stage timer runs (a timer may create its events).
int setup_dev_sel(exc_ctx *c, int ngp, int ngkmax, const double *apwalm, const int *igpig, const double *vgpc, int ldh,
                  double *h, double *o, cplx *ws = nullptr) {
  if (!ready(c)) return c ? EXC_ERR_STATE : EXC_ERR_ARG;
  const int nmat = ngp + c->d.nlotot;
  if (ngp < 0 || ngkmax < ngp || ldh < nmat || (!h && !o) || (ngp > 0 && (!apwalm || !igpig || !vgpc)))
    return EXC_ERR_ARG;
  int rc = use_device(c);
  if (rc) return rc;
  const bool untimed = ws != nullptr;   // a StageTimer of stage -1 records nothing
  if (ngp > 0) {
    // defaults from the autotuner's sweep on MI355X (64x64 MT and Y tiles, LDS filled by DMA;
    // the MT tile with Gauss's three real products per complex product)
    const int yv = c->y_variant >= 0 ? c->y_variant : 3;
    const int mv = c->mt_variant >= 0 ? c->mt_variant : 0;
    cplx *yt = ws;
    if (h) {   // the Y stage feeds only H
      if (!ws) {
        const size_t need = sizeof(cplx) * (size_t)ngp * c->d.naa * c->d.natm;
        if ((rc = grow((void **)&c->yt, &c->yt_bytes, need))) return rc;
        yt = c->yt;
      }
      StageTimer t(c, untimed ? -1 : EXC_STAGE_YSTAGE);
      rc = exc_launch_ystage(c->d, ngp, (const cplx *)apwalm, ngkmax, yt, ngp, yv, c->stream);
    }
    if (rc) return rc;
    {
      StageTimer t(c, untimed ? -1 : EXC_STAGE_MT);
      if ((c->d.iora & 1) && o) {
        // 'iora*': O's tiles take the operand A L, H's keep A (two launches)
        if (ws) return EXC_ERR_UNSUPPORTED;
        if ((rc = grow((void **)&c->apwo, &c->apwo_bytes, sizeof(cplx) * (size_t)ngp * c->d.naa * c->d.natm)))
          return rc;
        rc = exc_launch_iora_apw(c->d, ngp, (const cplx *)apwalm, ngkmax, c->apwo, ngp, c->stream);
        if (!rc && h)
          rc = exc_launch_mt(c->d, ngp, (const cplx *)apwalm, ngkmax, yt, ngp, igpig, vgpc, (cplx *)h, nullptr, ldh,
                             mv, c->stream);
        if (!rc)
          rc = exc_launch_mt(c->d, ngp, c->apwo, ngp, yt, ngp, igpig, vgpc, nullptr, (cplx *)o, ldh, mv, c->stream);
      } else {
        rc = exc_launch_mt(c->d, ngp, (const cplx *)apwalm, ngkmax, yt, ngp, igpig, vgpc, (cplx *)h, (cplx *)o, ldh,
                           mv, c->stream);
      }
    }
    if (rc) return rc;
  }
  StageTimer t(c, untimed ? -1 : EXC_STAGE_LO);
  return exc_launch_lo(c->d, ngp, (const cplx *)apwalm, ngkmax, (cplx *)h, (cplx *)o, ldh, c->stream);
}

// the same from host arrays; apwalm = NULL takes the device copy of the last exc_match
int setup_host_sel(exc_ctx *c, int ngp, int ngkmax, const double *apwalm, const int *igpig, const double *vgpc,
                   int ldh, double *h, double *o) {
  if (!ready(c)) return c ? EXC_ERR_STATE : EXC_ERR_ARG;
  const int nmat = ngp + c->d.nlotot;
  if (ngp < 0 || ngkmax < ngp || ldh < nmat || (!h && !o) || (ngp > 0 && (!igpig || !vgpc))) return EXC_ERR_ARG;
  const bool resident = apwalm == nullptr;
  if (resident && ngp > 0 && !(c->apw_res_valid && c->apw_res_ngp == ngp && c->apw_res_ngkmax == ngkmax))
    return EXC_ERR_STATE;
  int rc = use_device(c);
  if (rc) return rc;
  const size_t na = c->d.natm, lmm = c->d.lmmaxapw, ao = c->d.apwordmax, nk = ngkmax;
  const size_t b_apw = resident ? 0 : 16 * nk * ao * lmm * na, b_ig = 4 * nk, b_vg = 24 * nk;
  const size_t b_h = 16 * (size_t)ldh * nmat;
  auto al = [](size_t b) { return (b + 255) & ~(size_t)255; };
  const size_t o_ig = al(b_apw), o_vg = o_ig + al(b_ig), o_h = o_vg + al(b_vg);
  const size_t o_o = o_h + (h ? al(b_h) : 0);
  if ((rc = grow(&c->stage, &c->stage_bytes, o_o + (o ? b_h : 0)))) return rc;
  const int nout = (h ? 1 : 0) + (o ? 1 : 0);
  if ((rc = grow_host(&c->hstage, &c->hstage_bytes, std::max(b_apw, nout * b_h)))) return rc;
  char *base = (char *)c->stage;
  char *hs = (char *)c->hstage;
  const double *d_apw = resident ? (const double *)c->apw_res : (const double *)base;
  if (!resident && ngp > 0) {   // rows 1..ngp of apwalm are the input; the caller's padding rows are never read
    copy_cols(hs, 16 * nk, (const char *)apwalm, 16 * nk, 16 * (size_t)ngp, ao * lmm * na);
    EXC_CHECK(hipMemcpyAsync(base, hs, b_apw, hipMemcpyHostToDevice, c->stream));
  }
  EXC_CHECK(hipMemcpyAsync(base + o_ig, igpig, b_ig, hipMemcpyHostToDevice, c->stream));
  EXC_CHECK(hipMemcpyAsync(base + o_vg, vgpc, b_vg, hipMemcpyHostToDevice, c->stream));
  double *dh = h ? (double *)(base + o_h) : nullptr, *dout = o ? (double *)(base + o_o) : nullptr;
  rc = setup_dev_sel(c, ngp, ngkmax, d_apw, (const int *)(base + o_ig), (const double *)(base + o_vg), ldh, dh, dout);
  if (rc) return rc;
  // rows nmat+1..ldh of the caller's matrices are not part of the result: copy rows 1..nmat
  const size_t pitch = 16 * (size_t)ldh, width = 16 * (size_t)nmat;
  if (h && (rc = d2h_cols(c, (char *)h, pitch, (const char *)dh, pitch, width, nmat, hs))) return rc;
  if (o && (rc = d2h_cols(c, (char *)o, pitch, (const char *)dout, pitch, width, nmat, hs + (h ? b_h : 0)))) return rc;
  return exc_ctx_status(c);
}

// The Gram form's table from every h_a (rocSOLVER, strided batches): the eigenvalues give
// sigma = -min lambda (times 1 + 1e-6) when some lambda < 0, plus 1e-10 of the largest
// |lambda| (h + sigma I is then positive definite beyond the eigensolver's rounding); the
// table is the lower Cholesky factor of h + sigma I in the Y stage's summation order
// (exc_launch_crt_gram_tri: a triangular table, whose Y stage skips the zero part), or, if
// that factorization fails, conj(U) (Lambda + sigma)^(1/2) from the eigenvectors.  Once per
// change of h_a; a failed eigensolve leaves the form unused.
int gram_build(exc_ctx *c) {
  if (c->gram_valid) return EXC_OK;
  const int na = c->d.natm, n = c->d.naa;
  const size_t nn = (size_t)n * n * na;
  auto al = [](size_t b) { return (b + 255) & ~(size_t)255; };
  const size_t o_u = al(sizeof(cplx) * nn), o_l = o_u + al(sizeof(cplx) * nn), o_e = o_l + al(8 * (size_t)n * na);
  const size_t o_i = o_e + al(8 * (size_t)n * na), o_s = o_i + al(4 * (size_t)na);
  const size_t o_m = o_s + al(8);
  int rc = grow(&c->gram_buf, &c->gram_bytes, o_m + sizeof(cplx) * nn);
  if (rc) return rc;
  char *b = (char *)c->gram_buf;
  cplx *T = (cplx *)b, *U = (cplx *)(b + o_u);
  double *lam = (double *)(b + o_l), *E = (double *)(b + o_e), *dsig = (double *)(b + o_s);
  int *info = (int *)(b + o_i);
  cplx *Mp = (cplx *)(b + o_m);
  c->gram_table = T;
  c->gram_sig_dev = dsig;
  EXC_CHECK(hipMemcpyAsync(U, c->d.haa, sizeof(cplx) * nn, hipMemcpyDeviceToDevice, c->stream));
  // a handle of its own with the library-managed workspace (the eigensolve workers' handles
  // hold fixed workspaces sized for their own calls); the stream is drained below
  if (!c->gram_blas && rocblas_create_handle(&c->gram_blas) != rocblas_status_success) {
    c->gram_blas = nullptr;
    return EXC_ERR_HIP;
  }
  if (rocblas_set_stream(c->gram_blas, c->stream) != rocblas_status_success) return EXC_ERR_HIP;
  const bool solved = rocsolver_zheevd_strided_batched(c->gram_blas, rocblas_evect_original, rocblas_fill_upper, n,
                                                       (rocblas_double_complex *)U, n, (rocblas_stride)n * n, lam, n,
                                                       E, n, info, na) == rocblas_status_success;
  std::vector<double> hl((size_t)n * na);
  std::vector<int> hi(na, 1);
  if (solved) {
    EXC_CHECK(hipMemcpyAsync(hl.data(), lam, 8 * hl.size(), hipMemcpyDeviceToHost, c->stream));
    EXC_CHECK(hipMemcpyAsync(hi.data(), info, 4 * (size_t)na, hipMemcpyDeviceToHost, c->stream));
  }
  EXC_CHECK(hipStreamSynchronize(c->stream));
  bool ok = solved;
  double lmin = 0.0, lmax = 0.0;
  for (int a = 0; a < na; ++a) ok = ok && hi[a] == 0;
  for (double v : hl) {
    ok = ok && std::isfinite(v);
    lmin = std::min(lmin, v);
    lmax = std::max(lmax, fabs(v));
  }
  c->gram_sigma = ok ? (lmin < 0.0 ? -lmin * (1.0 + 1e-6) : 0.0) + 1e-10 * lmax : HUGE_VAL;
  c->gram_lmax = lmax;
  c->gram_tri = false;
  if (ok) {
    if ((rc = exc_launch_crt_gram_perm(c->d, c->gram_sigma, Mp, c->stream))) return rc;
    bool tri = rocsolver_zpotrf_strided_batched(c->gram_blas, rocblas_fill_lower, n, (rocblas_double_complex *)Mp, n,
                                                (rocblas_stride)n * n, info, na) == rocblas_status_success;
    if (tri) {
      EXC_CHECK(hipMemcpyAsync(hi.data(), info, 4 * (size_t)na, hipMemcpyDeviceToHost, c->stream));
      EXC_CHECK(hipStreamSynchronize(c->stream));
      for (int a = 0; a < na; ++a) tri = tri && hi[a] == 0;
    }
    rc = tri ? exc_launch_crt_gram_tri(c->d, Mp, T, c->stream)
             : exc_launch_crt_gram_table(c->d, U, lam, c->gram_sigma, T, c->stream);
    if (rc) return rc;
    c->gram_tri = tri;
    EXC_CHECK(hipMemcpyAsync(dsig, &c->gram_sigma, 8, hipMemcpyHostToDevice, c->stream));
    EXC_CHECK(hipStreamSynchronize(c->stream));
  }
  c->gram_valid = true;
  return EXC_OK;
}

}  // namespace

extern "C" {

int exc_hamiltonandoverlapsetup_dev(exc_ctx *c, int ngp, int ngkmax, const double *apwalm, const int *igpig,
                                    const double *vgpc, int ldh, double *h, double *o) {
  if (!h || !o) return EXC_ERR_ARG;
  return setup_dev_sel(c, ngp, ngkmax, apwalm, igpig, vgpc, ldh, h, o);
}

int exc_hamiltonandoverlapsetup_noalloc_dev(exc_ctx *c, int ngp, int ngkmax, const double *apwalm, const int *igpig,
                                            const double *vgpc, int ldh, double *h, double *o) {
  if (!h || !o) return EXC_ERR_ARG;
  if (!ready(c)) return c ? EXC_ERR_STATE : EXC_ERR_ARG;
  if (c->d.iora & 1) return EXC_ERR_UNSUPPORTED;   // the 'iora*' APW-APW overlap needs a second operand buffer
  const size_t need = sizeof(cplx) * (size_t)std::max(ngp, 0) * c->d.naa * c->d.natm;
  if (need > kScratchBytes) {
    fprintf(stderr,
            "exc_lapw: exc_hamiltonandoverlapsetup_noalloc_dev: ngp %d needs %zu bytes of workspace, the fixed limit "
            "is %zu (ngp * natm * naa <= %zu)\n",
            ngp, need, kScratchBytes, kScratchBytes / sizeof(cplx));
    return EXC_ERR_ARG;
  }
  cplx *ws = (cplx *)scratch_of(c->device);
  if (!ws) return EXC_ERR_STATE;
  return setup_dev_sel(c, ngp, ngkmax, apwalm, igpig, vgpc, ldh, h, o, ws);
}

int exc_hamiltonsetup_dev(exc_ctx *c, int ngp, int ngkmax, const double *apwalm, const int *igpig,
                          const double *vgpc, int ldh, double *h) {
  if (!h) return EXC_ERR_ARG;
  return setup_dev_sel(c, ngp, ngkmax, apwalm, igpig, vgpc, ldh, h, nullptr);
}

int exc_overlapsetup_dev(exc_ctx *c, int ngp, int ngkmax, const double *apwalm, const int *igpig, const double *vgpc,
                         int ldh, double *o) {
  if (!o) return EXC_ERR_ARG;
  return setup_dev_sel(c, ngp, ngkmax, apwalm, igpig, vgpc, ldh, nullptr, o);
}

int exc_hamiltonandoverlapsetup(exc_ctx *c, int ngp, int ngkmax, const double *apwalm, const int *igpig,
                                const double *vgpc, int ldh, double *h, double *o) {
  if (!h || !o) return EXC_ERR_ARG;
  return setup_host_sel(c, ngp, ngkmax, apwalm, igpig, vgpc, ldh, h, o);
}

int exc_hamiltonsetup(exc_ctx *c, int ngp, int ngkmax, const double *apwalm, const int *igpig, const double *vgpc,
                      int ldh, double *h) {
  if (!h) return EXC_ERR_ARG;
  return setup_host_sel(c, ngp, ngkmax, apwalm, igpig, vgpc, ldh, h, nullptr);
}

int exc_overlapsetup(exc_ctx *c, int ngp, int ngkmax, const double *apwalm, const int *igpig, const double *vgpc,
                     int ldh, double *o) {
  if (!o) return EXC_ERR_ARG;
  return setup_host_sel(c, ngp, ngkmax, apwalm, igpig, vgpc, ldh, nullptr, o);
}

int exc_setup_batched_dev(exc_ctx *c, int nk, const exc_kpoint_dev *ks, int flags) {
  if (!ready(c)) return c ? EXC_ERR_STATE : EXC_ERR_ARG;
  if (nk < 0 || (nk > 0 && !ks) ||
      (flags & ~(EXC_BATCH_MATCH | EXC_BATCH_GRAPH | EXC_BATCH_SPLIT_BF16 | EXC_BATCH_OLP_LEGENDRE |
                 EXC_BATCH_SPLIT_CRT)))
    return EXC_ERR_ARG;
  const bool split = flags & EXC_BATCH_SPLIT_BF16;
  const bool crt = flags & EXC_BATCH_SPLIT_CRT;
  if (split && crt) return EXC_ERR_ARG;
  // exact int32 sums of K residue products: K 128^2 < 2^31
  if (crt && c->d.natm * c->d.naa > 131000) return EXC_ERR_UNSUPPORTED;
  // the addition-theorem overlap needs the radial factors that only the grouped match writes
  const bool olp = flags & EXC_BATCH_OLP_LEGENDRE;
  if (olp && !(flags & EXC_BATCH_MATCH)) return EXC_ERR_ARG;
  if (split && c->d.naa > kSplitKB) return EXC_ERR_UNSUPPORTED;
  // 'iora*' overlap terms: the APW-APW factor I + h1aa in the addition-theorem form only; the
  // split products' O tiles take none of the terms
  if (((c->d.iora & 1) || ((split || crt) && c->d.iora)) && !olp) return EXC_ERR_UNSUPPORTED;
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
  const int bv = split ? (c->bsplit_variant >= 0 ? c->bsplit_variant : 3) : (c->bmt_variant >= 0 ? c->bmt_variant : 0);
  const int yv = c->y_variant >= 0 ? c->y_variant : 0;     // 64x128 Y tiles, DMA-staged, 3 real products
  // residue GEMM: the register-pipelined 256 x 256 form from N = 2048, 128 x 128 tiles below
  int maxn = 0;
  for (int k = 0; k < nk; ++k) maxn = std::max(maxn, ks[k].ngp);
  const int cv = c->crt_variant >= 0 ? c->crt_variant : (maxn >= 2048 ? 8 : 4);
  const int ov = c->olp_variant >= 0 ? c->olp_variant : 3;   // rows along lanes, column factors as scalars, pipelined
  int fv = c->fork_variant;
  // the modular split's Gram form: needs O_APW from the addition theorem, equal to A's Gram
  // matrix (no 'iora*' h1aa); built in by default unless the shift sigma exceeds half of
  // h_a's largest |eigenvalue| (the form's rounding grows with sigma)
  bool gram = false;
  if (crt && olp && !(c->d.iora & 1) && c->crt_form != 0) {
    if ((rc = gram_build(c))) return rc;
    gram = std::isfinite(c->gram_sigma) && (c->crt_form == 1 || c->gram_sigma <= 0.5 * c->gram_lmax);
  }
  // graph shape: one chain (faster for the FP64 path on workloads 1-4: the overlap kernel only
  // competes with the MT kernel), the overlap and lo on a side branch for the Gram form, whose
  // residue GEMM leaves the vector units idle (+1% at workload 3; the merge waits for the branch)
  if (fv < 0) fv = gram ? 0 : 1;
  const cplx *gram_t = gram ? c->gram_table : nullptr;
  const double *gram_s = gram ? c->gram_sig_dev : nullptr;
  // moduli: the caller's count, else 10 with the Gram form (its rounding error is far below
  // the (A, Yt) form's: 5.7e-10 Ha at 10 moduli against 2.6e-9 at 11 on workload 3) and 11
  const int nmod = c->crt_nmod > 0 ? c->crt_nmod : (gram ? 10 : 11);
  CrtTab ct;
  if (crt) {
    exc_crt_tables(nmod, &ct);
    c->crt_last_form = gram ? 1 : 0;
    c->crt_last_nmod = nmod;
  }
  constexpr int kCrtPlanes = 2;                             // residue planes per operand and modulus
  const int kp = exc_crt_kpad(c->d.natm * c->d.naa);
  // batch signature: descriptors, flags and variants
  // the Gram form's E stage on its triangular table: narrow column tiles skip more of it
  // (64 x 64 three-product tiles, 5.7 against 6.7 ms at workload 3; the dense Y stage keeps
  // 64 x 128, 8.8 against 9.1 ms)
  const bool etri = gram && c->gram_tri;
  const int ye = c->y_variant >= 0 ? c->y_variant : (etri ? 4 : 0);
  int tail[10] = {nk, flags, bv, yv, crt ? cv : -1, (crt ? nmod : -1) + 100 * ov, fv, gram, etri, ye};
  std::vector<char> sig(sizeof(exc_kpoint_dev) * nk + sizeof(tail));
  memcpy(sig.data(), ks, sizeof(exc_kpoint_dev) * nk);
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
    // radial factors and their transposed copy (overlap variants 2 and 3: N rounded up to 64
    // rows per atom and slot, NL + 1 slots per atom: the radial factors of each l and the
    // structure factors; variant 3 reads up to 3 slots past the last k-point's: padding)
    size_t rttot = 0;
    size_t rtpad = 0;
    for (int k = 0; k < nk; ++k) {
      rttot += (size_t)(ks[k].ngp + 63) / 64 * 64 * c->d.natm * (nl2 + 2);
      rtpad = std::max(rtpad, (size_t)(ks[k].ngp + 63) / 64 * 64 * 2 * 4);
    }
    rttot += rtpad;
    if (olp && (rc = grow((void **)&c->rfbuf, &c->rf_bytes, sizeof(double) * (rtot + rttot ? rtot + rttot : 1))))
      return rc;
    size_t qtot = 0, xtot = 0;
    int cbt = 0, cwr = 0, cwc = 0;
    if (crt && exc_crt_shape(cv, &cbt, &cwr, &cwc)) return EXC_ERR_ARG;
    // squared row norms of A per atom (cnp each), when the grouped match runs here
    const size_t nparts = do_match ? (size_t)c->d.natm : 0;
    size_t ptot2 = 0;
    if (crt) {
      for (int k = 0; k < nk; ++k) {
        const size_t cnp = (size_t)(ks[k].ngp + 255) / 256 * 256;
        qtot += (size_t)(gram ? 1 : 2) * kCrtPlanes * ct.n * kp * cnp;   // the Gram form splits E only
        xtot += 2 * cnp;
        ptot2 += nparts * cnp;
      }
      if ((rc = grow(&c->crt_q, &c->crt_q_bytes, qtot ? qtot : 1))) return rc;
      if ((rc = grow((void **)&c->crt_x, &c->crt_x_bytes, sizeof(int) * (xtot ? xtot : 1)))) return rc;
      if ((rc = grow((void **)&c->crt_parts, &c->crt_parts_bytes, sizeof(double) * (ptot2 ? ptot2 : 1)))) return rc;
    }
    std::vector<mt_item> citems;
    std::vector<int2> cwork;
    size_t qoff = 0, xoff = 0, poff2 = 0;
    int maxcnp = 0;
    std::vector<kdesc> kd(nk);
    std::vector<mt_item> items, oitems;
    const int T = split ? exc_bsplit_tile(bv) : exc_bmt_tile(bv);
    const int TO = exc_olp_tile();
    size_t yoff = 0, poff = 0, eoff = 0, roff = 0, rtoff = 0;
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
      e.rf = e.rft = nullptr;
      if (olp) {
        e.rf = c->rfbuf + roff;
        e.rft = c->rfbuf + rtot + rtoff;
        roff += (size_t)q.ngp * c->d.natm * nl2;
        rtoff += (size_t)(q.ngp + 63) / 64 * 64 * c->d.natm * (nl2 + 2);
      }
      e.cnp = 0;
      e.qa = e.qy = nullptr;
      e.pa2 = nullptr;
      e.xa = e.xy = nullptr;
      if (crt) {
        // planes [modulus][plane][K/16][cnp][16] of A then of Yt; tiles (k, ti <= tj, H | O)
        // in 4 x 4 blocks of tiles, and the (tile, modulus, product) work modulus- and
        // product-major within the k-point, the tiles in that block order
        const int cnp = (q.ngp + 255) / 256 * 256;
        e.cnp = cnp;
        e.qa = (char *)c->crt_q + qoff;
        e.qy = gram ? nullptr : (char *)c->crt_q + qoff + (size_t)kCrtPlanes * ct.n * kp * cnp;
        qoff += (size_t)(gram ? 1 : 2) * kCrtPlanes * ct.n * kp * cnp;
        e.xa = c->crt_x + xoff;
        e.xy = c->crt_x + xoff + cnp;
        e.pa2 = (do_match && !gram) ? c->crt_parts + poff2 : nullptr;
        poff2 += nparts * cnp;
        xoff += 2 * (size_t)cnp;
        maxcnp = cnp > maxcnp ? cnp : maxcnp;
        const int nt = (q.ngp + cbt - 1) / cbt, first = (int)citems.size(), S = 4;
        for (int which = 0; which < (olp ? 1 : 2); ++which)
          for (int bj = 0; bj < nt; bj += S)
            for (int bi = 0; bi <= bj; bi += S)
              for (int tj = bj; tj < std::min(bj + S, nt); ++tj)
                for (int ti = bi; ti < std::min(bi + S, tj + 1); ++ti)
                  citems.push_back({k, ti, tj, gram ? 2 : which});
        for (int l = 0; l < ct.n; ++l)
          for (int p = 0; p < 2; ++p)
            for (int i = first; i < (int)citems.size(); ++i) cwork.push_back(make_int2(i, 4 * l + p));
      }
      maxngp = q.ngp > maxngp ? q.ngp : maxngp;
      const int nt = (q.ngp + T - 1) / T;
      // with the addition-theorem overlap the MT work list holds the H tiles only
      for (int tj = 0; tj < nt; ++tj)
        for (int ti = 0; ti <= tj; ++ti) {
          items.push_back({k, ti, tj, 0});
          if (!olp) items.push_back({k, ti, tj, 1});
        }
      if (olp) {
        // in 8 x 8 blocks of tiles: the 64 workgroups an XCD runs at a time (the overlap
        // kernel hands each XCD a contiguous part of the list) share 8 row and 8 column
        // panels of the radial factors in its L2
        const int no = (q.ngp + TO - 1) / TO, S = 8;
        for (int bj = 0; bj < no; bj += S)
          for (int bi = 0; bi <= bj; bi += S)
            for (int tj = bj; tj < std::min(bj + S, no); ++tj)
              for (int ti = bi; ti < std::min(bi + S, tj + 1); ++ti) oitems.push_back({k, ti, tj, 1});
      }
    }
    if (olp) {
      if ((rc = grow((void **)&c->d_oitems, &c->oitems_cap, sizeof(mt_item) * (oitems.empty() ? 1 : oitems.size()))))
        return rc;
      if (!oitems.empty())
        EXC_CHECK(hipMemcpy(c->d_oitems, oitems.data(), sizeof(mt_item) * oitems.size(), hipMemcpyHostToDevice));
    }
    c->batch_noitems = (int)oitems.size();
    c->batch_ncitems = (int)citems.size();
    c->batch_ncwork = (int)cwork.size();
    c->batch_maxcnp = maxcnp;
    if (crt) {
      const size_t rbytes = (size_t)citems.size() * ct.n * 2 * cbt * cbt;
      if ((rc = grow((void **)&c->crt_res, &c->crt_res_bytes, rbytes ? rbytes : 1))) return rc;
      if ((rc = grow((void **)&c->d_citems, &c->citems_cap, sizeof(mt_item) * (citems.empty() ? 1 : citems.size()))))
        return rc;
      if ((rc = grow((void **)&c->d_cwork, &c->cwork_cap, sizeof(int2) * (cwork.empty() ? 1 : cwork.size()))))
        return rc;
      if (!citems.empty())
        EXC_CHECK(hipMemcpy(c->d_citems, citems.data(), sizeof(mt_item) * citems.size(), hipMemcpyHostToDevice));
      if (!cwork.empty())
        EXC_CHECK(hipMemcpy(c->d_cwork, cwork.data(), sizeof(int2) * cwork.size(), hipMemcpyHostToDevice));
    }
    if ((rc = grow((void **)&c->d_ks, &c->ks_cap, sizeof(kdesc) * nk))) return rc;
    if ((rc = grow((void **)&c->d_items, &c->items_cap, sizeof(mt_item) * (items.empty() ? 1 : items.size()))))
      return rc;
    EXC_CHECK(hipMemcpy(c->d_ks, kd.data(), sizeof(kdesc) * nk, hipMemcpyHostToDevice));
    if (!items.empty())
      EXC_CHECK(hipMemcpy(c->d_items, items.data(), sizeof(mt_item) * items.size(), hipMemcpyHostToDevice));
    c->batch_nitems = (int)items.size();
    int nb = c->ncu * (split ? exc_bsplit_max_blocks(bv) : 1);
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
  // structure factors, and write parts of H and O the MT stage does not) can form a branch of
  // their own beside the Y and MT stages; with a persistent split-BF16 kernel (static work
  // split, all workgroups assumed resident) the branches join before it.  Eager calls keep one
  // stream so the stage timers stay exact.  With every stage filling the GPU the branches
  // only compete (workload 3: 111 ms per replay with the branch, 105 ms as one chain; workload
  // 1: 4.79 against 4.60 ms), so the default graph is one chain (variant 1 of
  // EXC_STAGE_GRAPH) and the branch is variant 0
  const bool fork = capture && fv == 0;
  const bool join_early = split;
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
      if (olp && (r = exc_launch_olp(c->d, nk, c->batch_maxngp, c->d_ks, c->d_oitems, c->batch_noitems, ov, s2))) return r;
      if ((r = exc_launch_blo(c->d, nk, c->batch_maxngp, c->d_ks, s2))) return r;
      if (hipEventRecord(c->ev_join, c->side) != hipSuccess) return EXC_ERR_HIP;
    }
    {
      // the Gram form's E = A conj(U) (Lambda + sigma)^(1/2) is the Y stage on its table
      StageTimer t(c, capture ? -1 : EXC_STAGE_YSTAGE);
      exc_dev dy = c->d;
      if (gram) {
        dy.haa = (cplx *)gram_t;
        dy.ytri = etri ? 1 : 0;
      }
      if ((r = exc_launch_bystage(dy, nk, c->batch_maxngp, c->d_ks, gram ? ye : yv, c->stream))) return r;
    }
    // the Gram form's merge reads O's APW block: the overlap runs before it
    const bool olp_first = gram && !fork;
    if (olp_first) {
      StageTimer t(c, capture ? -1 : EXC_STAGE_OLP);
      if ((r = exc_launch_olp(c->d, nk, c->batch_maxngp, c->d_ks, c->d_oitems, c->batch_noitems, ov, c->stream))) return r;
    }
    if (crt) {
      // the residue GEMM and the merge write only H's (and without the addition theorem O's)
      // APW-APW tiles: no conflict with the side branch
      {
        StageTimer t(c, capture ? -1 : EXC_STAGE_CRTPREP);
        if ((r = exc_launch_crt_prep(c->d, nk, c->batch_maxcnp, c->d_ks, ct, do_match, gram, c->stream))) return r;
      }
      {
        StageTimer t(c, capture ? -1 : EXC_STAGE_CRTGEMM);
        if ((r = exc_launch_crt_gemm(c->d, c->d_ks, c->d_citems, c->d_cwork, c->batch_ncwork, ct, c->crt_res, cv,
                                     c->stream)))
          return r;
      }
      if (gram && fork && hipStreamWaitEvent(c->stream, c->ev_join, 0) != hipSuccess) return EXC_ERR_HIP;
      StageTimer t(c, capture ? -1 : EXC_STAGE_CRTMERGE);
      if ((r = exc_launch_crt_merge(c->d, c->d_ks, c->d_citems, c->batch_ncitems, ct, c->crt_res, cv, gram_s,
                                    c->stream)))
        return r;
    } else if (split) {
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
      if ((r = exc_launch_bmt(c->d, c->d_ks, c->d_items, c->batch_nitems, bv, c->stream))) return r;
    }
    if (fork && !join_early && hipStreamWaitEvent(c->stream, c->ev_join, 0) != hipSuccess) return EXC_ERR_HIP;
    if (fork) return r;
    if (olp && !olp_first) {
      StageTimer t(c, capture ? -1 : EXC_STAGE_OLP);
      if ((r = exc_launch_olp(c->d, nk, c->batch_maxngp, c->d_ks, c->d_oitems, c->batch_noitems, ov, c->stream))) return r;
    }
    {
      StageTimer t(c, capture ? -1 : EXC_STAGE_LO);
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
  // the adjoint differentiates the Theta-form kinetic term and the 'none' overlap only
  if (c->d.zora || c->d.iora) return EXC_ERR_UNSUPPORTED;
  const int nmat = ngp + c->d.nlotot;
  if (ngp < 0 || ngkmax < ngp || nst < 0 || ldc < nmat || !dedtau || !dedv) return EXC_ERR_ARG;
  if (ngp > 0 && nst > 0 && (!apwalm || !igpig || !vgpc || !sfacgp || !evec || !wh || !wo)) return EXC_ERR_ARG;
  int rc = use_device(c);
  if (rc) return rc;
  const size_t need = sizeof(cplx) * exc_vjp_work_elems(c->d, ngp, nst);
  if ((rc = grow((void **)&c->vjp_work, &c->vjp_bytes, need ? need : 1))) return rc;
  exc_ctx::EigWorker &W = c->ew[0];
  W.s = c->stream;
  if (!W.blas && rocblas_create_handle(&W.blas) != rocblas_status_success) {
    W.blas = nullptr;
    return EXC_ERR_HIP;
  }
  if (rocblas_set_stream(W.blas, W.s) != rocblas_status_success) return EXC_ERR_HIP;
  StageTimer t(c, EXC_STAGE_VJP);
  if (ngp > 0 && nst > 0) {
    // uv(:, l K + c) = (q_l C)^T A(:, c), l = 0..3 (q_0 = 1): the q-scaled copies of C, then
    // one strided batch of four library products into uv (nst x 4K, leading dimension nst)
    const int ktot = c->d.natm * c->d.naa;
    cplx *uv = c->vjp_work;
    cplx *cs = uv + (size_t)nst * 4 * ktot + 2 * (size_t)ngp * nst;
    if ((rc = exc_launch_vjp_qscale(ngp, nst, (const cplx *)evec, ldc, vgpc, cs, c->stream))) return rc;
    const rocblas_double_complex one = {1.0, 0.0}, zero = {0.0, 0.0};
    if (rocblas_zgemm_strided_batched(W.blas, rocblas_operation_transpose, rocblas_operation_none, nst, ktot, ngp,
                                      &one, (const rocblas_double_complex *)cs, ngp, (rocblas_stride)ngp * nst,
                                      (const rocblas_double_complex *)apwalm, ngkmax, 0, &zero,
                                      (rocblas_double_complex *)uv, nst, (rocblas_stride)ktot * nst,
                                      4) != rocblas_status_success)
      return EXC_ERR_HIP;
  }
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
  return exc_ctx_status(c);
}

}  // extern "C"
