// C-ABI entry points of the per-SCF producers around the setup: match's boundary values from
// the radial functions (exc_ctx_set_radial_fr), the muffin-tin integrals and the density from
// a density matrix (exc_ctx_hmlint, exc_genrhomt*), Theta / V(G) (exc_ctx_gencfun,
// exc_ctx_genveffig) and the interstitial density (exc_genrhoir*).
#include "exc_ctx.h"

using namespace exc_impl;

extern "C" {

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
  // the uploads below are synchronous copies on the null stream, which does not wait for a
  // non-blocking context stream: work still queued there (genrhomt, rhovalk) may read the
  // Gaunt table and the radial buffer they overwrite, so it completes first
  EXC_CHECK(hipStreamSynchronize(c->stream));
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
  c->gram_valid = false;
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

int exc_rhovalk_dev(exc_ctx *c, int lradstep, const double *zbshtvr, const double *rfshtvr, int ngp, int ngkmax,
                    const double *apwalm, int nst, int ldv, const double *evecfv, const double *evecsv,
                    const double *occ, double wkpt, double epsocc, double *rhomt) {
  if (!ready(c)) return c ? EXC_ERR_STATE : EXC_ERR_ARG;
  if (c->hml_lmaxvr < 0) return EXC_ERR_STATE;   // the radial functions of exc_ctx_hmlint
  const exc_dev &d = c->d;
  const int lmaxvr = c->hml_lmaxvr, nrmtmax = c->hml_nrmtmax, ns = c->dims.nspecies;
  if (lmaxvr > d.lmaxapw) return EXC_ERR_ARG;   // wavefmt.f90:76-81
  if (lmaxvr > 8 || d.apwordmax > 4) return EXC_ERR_UNSUPPORTED;
  if (lradstep < 1 || nst < 0 || nst > 65535 || ngp < 0 || ngkmax < (ngp > 1 ? ngp : 1) || ldv < ngp + d.nlotot ||
      ldv < 1)
    return EXC_ERR_ARG;
  if (nst == 0) return EXC_OK;
  if (!zbshtvr || !rfshtvr || !apwalm || !evecfv || !occ || !rhomt) return EXC_ERR_ARG;
  int rc = use_device(c);
  if (rc) return rc;
  const int ldw = d.natm * d.naa, nmat = ngp + d.nlotot;
  auto al = [](size_t x) { return (x + 255) & ~(size_t)255; };
  const size_t o_c = al(sizeof(cplx) * (size_t)ldw * nst);   // W, then evecfv evecsv, then the kernels' work
  const size_t o_k = o_c + (evecsv ? al(sizeof(cplx) * (size_t)nmat * nst) : 0);
  if ((rc = grow(&c->rvw, &c->rvw_bytes, o_k + exc_rhovalk_work_bytes(d, lmaxvr, c->dims.nlomax, nst)))) return rc;
  exc_ctx::EigWorker &W = c->ew[0];
  W.s = c->stream;
  if (!W.blas && rocblas_create_handle(&W.blas) != rocblas_status_success) {
    W.blas = nullptr;
    return EXC_ERR_HIP;
  }
  if (rocblas_set_stream(W.blas, W.s) != rocblas_status_success) return EXC_ERR_HIP;
  const HmlLayout H = hml_layout(c, lmaxvr, nrmtmax);
  const char *b = (const char *)c->hml;
  const int *di = (const int *)(b + H.o_i);
  const rocblas_double_complex one = {1.0, 0.0}, zero = {0.0, 0.0};
  StageTimer t(c, EXC_STAGE_SCF);
  if (evecsv) {   // the states evecfv evecsv (wavefmt is linear in the coefficients)
    cplx *c2 = (cplx *)((char *)c->rvw + o_c);
    if (rocblas_zgemm(W.blas, rocblas_operation_none, rocblas_operation_none, nmat, nst, nst, &one,
                      (const rocblas_double_complex *)evecfv, ldv, (const rocblas_double_complex *)evecsv, nst,
                      &zero, (rocblas_double_complex *)c2, nmat) != rocblas_status_success)
      return EXC_ERR_HIP;
    evecfv = (const double *)c2;
    ldv = nmat;
  }
  // W = A^T C (the zdotu of wavefmt.f90:95 for every (atom, io, lm <= lmaxvr) and state at
  // once): per atom the first apwordmax lmmaxvr columns of its apwalm block, one strided batch
  // over the atoms; row a naa + io + apwordmax lm of W (leading dimension natm naa)
  if (ngp > 0) {
    const int mrows = d.apwordmax * (lmaxvr + 1) * (lmaxvr + 1);
    if (rocblas_zgemm_strided_batched(W.blas, rocblas_operation_transpose, rocblas_operation_none, mrows, nst, ngp,
                                      &one, (const rocblas_double_complex *)apwalm, ngkmax,
                                      (rocblas_stride)ngkmax * d.naa, (const rocblas_double_complex *)evecfv, ldv, 0,
                                      &zero, (rocblas_double_complex *)c->rvw, ldw, (rocblas_stride)d.naa,
                                      d.natm) != rocblas_status_success)
      return EXC_ERR_HIP;
  } else {
    EXC_CHECK(hipMemsetAsync(c->rvw, 0, sizeof(cplx) * (size_t)ldw * nst, c->stream));
  }
  return exc_launch_rhovalk(d, lmaxvr, lradstep, nrmtmax, c->dims.nlomax, di, di + ns, di + 2 * ns,
                            (const double *)(b + H.o_apw), (const double *)(b + H.o_lo), (const cplx *)c->rvw, ldw,
                            ngp, (const cplx *)evecfv, ldv, nst, occ, wkpt, epsocc, (const cplx *)zbshtvr, rfshtvr,
                            (char *)c->rvw + o_k, rhomt, c->stream);
}

int exc_rhovalk(exc_ctx *c, int lradstep, const double *zbshtvr, const double *rfshtvr, int ngp, int ngkmax,
                const double *apwalm, int nst, int ldv, const double *evecfv, const double *evecsv, const double *occ,
                double wkpt, double epsocc, double *rhomt) {
  if (!ready(c)) return c ? EXC_ERR_STATE : EXC_ERR_ARG;
  if (c->hml_lmaxvr < 0) return EXC_ERR_STATE;
  if (nst < 0 || ldv < 1 || ngkmax < 1) return EXC_ERR_ARG;
  if (nst == 0) return EXC_OK;
  if (!zbshtvr || !rfshtvr || !apwalm || !evecfv || !occ || !rhomt) return EXC_ERR_ARG;
  int rc = use_device(c);
  if (rc) return rc;
  const exc_dev &d = c->d;
  const size_t lmv = (size_t)(c->hml_lmaxvr + 1) * (c->hml_lmaxvr + 1);
  const size_t b_a = 16 * (size_t)ngkmax * d.naa * d.natm, b_v = 16 * (size_t)ldv * nst, b_o = 8 * (size_t)nst;
  const size_t b_z = 16 * lmv * lmv, b_f = 8 * lmv * lmv, b_r = 8 * lmv * c->hml_nrmtmax * d.natm;
  const size_t b_s = evecsv ? 16 * (size_t)nst * nst : 0;
  auto al = [](size_t x) { return (x + 255) & ~(size_t)255; };
  const size_t o_v = al(b_a), o_o = o_v + al(b_v), o_z = o_o + al(b_o), o_f = o_z + al(b_z), o_r = o_f + al(b_f);
  const size_t o_s = o_r + al(b_r);
  if ((rc = grow(&c->stage, &c->stage_bytes, o_s + b_s))) return rc;
  char *base = (char *)c->stage;
  EXC_CHECK(hipMemcpyAsync(base, apwalm, b_a, hipMemcpyHostToDevice, c->stream));
  EXC_CHECK(hipMemcpyAsync(base + o_v, evecfv, b_v, hipMemcpyHostToDevice, c->stream));
  EXC_CHECK(hipMemcpyAsync(base + o_o, occ, b_o, hipMemcpyHostToDevice, c->stream));
  EXC_CHECK(hipMemcpyAsync(base + o_z, zbshtvr, b_z, hipMemcpyHostToDevice, c->stream));
  EXC_CHECK(hipMemcpyAsync(base + o_f, rfshtvr, b_f, hipMemcpyHostToDevice, c->stream));
  EXC_CHECK(hipMemcpyAsync(base + o_r, rhomt, b_r, hipMemcpyHostToDevice, c->stream));
  if (evecsv) EXC_CHECK(hipMemcpyAsync(base + o_s, evecsv, b_s, hipMemcpyHostToDevice, c->stream));
  if ((rc = exc_rhovalk_dev(c, lradstep, (const double *)(base + o_z), (const double *)(base + o_f), ngp, ngkmax,
                            (const double *)base, nst, ldv, (const double *)(base + o_v),
                            evecsv ? (const double *)(base + o_s) : nullptr, (const double *)(base + o_o), wkpt,
                            epsocc, (double *)(base + o_r))))
    return rc;
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
    if (c->cfdamp_t1 > 0.0 && (rc = exc_launch_cfdamp(c->d, bvec, n, c->cfdamp_t1, c->d.cfunig, c->stream)))
      return rc;
  }
  if (cfunig) EXC_CHECK(hipMemcpyAsync(cfunig, c->d.cfunig, 16 * (size_t)n, hipMemcpyDeviceToHost, c->stream));
  if (cfunir) EXC_CHECK(hipMemcpyAsync(cfunir, c->cfunir, 8 * (size_t)n, hipMemcpyDeviceToHost, c->stream));
  EXC_CHECK(hipStreamSynchronize(c->stream));
  c->have_cfunir = true;
  return EXC_OK;
}

int exc_ctx_set_cfdamp(exc_ctx *c, double cfdamp, double gmaxvr) {
  if (!c || !(cfdamp >= 0.0) || !(gmaxvr > 0.0)) return EXC_ERR_ARG;
  c->cfdamp_t1 = cfdamp / gmaxvr;
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

// ---------------------------------------------------------------------------------------
// G+k sets (gengpvec.f90, gensfacgp.f90)
// ---------------------------------------------------------------------------------------
namespace {
// the gk_buf layout: [kv: 6 nk doubles][ngp: nk ints, 8-byte aligned][staging]
size_t gk_head(int nk) { return 48 * (size_t)nk + ((4 * (size_t)nk + 255) / 256) * 256 + 256; }
}  // namespace

int exc_gengpvec_dev(exc_ctx *c, int nk, const double *bvec, const double *vpl, const double *vpc, double gkmax,
                     int ngkmax, int *ngp, int *igpig, double *vgpl, double *vgpc, double *gpc, double *tpgpc,
                     double *sfacgp) {
  if (!c || nk < 0 || !bvec || ngkmax <= 0 || !(gkmax > 0.0)) return EXC_ERR_ARG;
  if (nk == 0) return EXC_OK;
  if (!vpl || !vpc || !ngp || !igpig || !vgpl || !vgpc || !gpc || !tpgpc) return EXC_ERR_ARG;
  if (!c->have_atoms || !c->have_gvec) return EXC_ERR_STATE;
  int rc = use_device(c);
  if (rc) return rc;
  const size_t head = gk_head(nk);
  if ((rc = grow(&c->gk_buf, &c->gk_buf_bytes, head))) return rc;
  std::vector<double> kv(6 * (size_t)nk);
  for (int k = 0; k < nk; ++k)
    for (int d = 0; d < 3; ++d) {
      kv[6 * k + d] = vpl[3 * k + d];
      kv[6 * k + 3 + d] = vpc[3 * k + d];
    }
  double *dkv = (double *)c->gk_buf;
  int *dngp = (int *)((char *)c->gk_buf + 48 * (size_t)nk);
  EXC_CHECK(hipMemcpyAsync(dkv, kv.data(), sizeof(double) * kv.size(), hipMemcpyHostToDevice, c->stream));
  if ((rc = exc_launch_gengpvec(c->d, nk, bvec, dkv, gkmax, ngkmax, dngp, igpig, vgpl, vgpc, gpc, tpgpc,
                                (cplx *)sfacgp, c->stream)))
    return rc;
  EXC_CHECK(hipMemcpyAsync(ngp, dngp, sizeof(int) * (size_t)nk, hipMemcpyDeviceToHost, c->stream));
  EXC_CHECK(hipStreamSynchronize(c->stream));
  for (int k = 0; k < nk; ++k)
    if (ngp[k] > ngkmax) return EXC_ERR_ARG;
  return EXC_OK;
}

int exc_gengpvec(exc_ctx *c, const double *bvec, const double *vpl, const double *vpc, double gkmax, int ngkmax,
                 int *ngp, int *igpig, double *vgpl, double *vgpc, double *gpc, double *tpgpc) {
  if (!c || !bvec || !vpl || !vpc || ngkmax <= 0 || !(gkmax > 0.0) || !ngp || !igpig || !vgpl || !vgpc || !gpc ||
      !tpgpc)
    return EXC_ERR_ARG;
  if (!c->have_atoms || !c->have_gvec) return EXC_ERR_STATE;
  int rc = use_device(c);
  if (rc) return rc;
  // device staging after the head: igpig, vgpl, vgpc, gpc, tpgpc (9 doubles + 1 int per row)
  const size_t head = gk_head(1), n = (size_t)ngkmax;
  if ((rc = grow(&c->gk_buf, &c->gk_buf_bytes, head + 72 * n + 4 * n + 256))) return rc;
  char *st = (char *)c->gk_buf + head;
  double *d_vgpl = (double *)st, *d_vgpc = d_vgpl + 3 * n, *d_gpc = d_vgpc + 3 * n, *d_tp = d_gpc + n;
  int *d_ig = (int *)(d_tp + 2 * n);
  int cnt = 0;
  const int rc2 = exc_gengpvec_dev(c, 1, bvec, vpl, vpc, gkmax, ngkmax, &cnt, d_ig, d_vgpl, d_vgpc, d_gpc, d_tp,
                                   nullptr);
  if (rc2 && rc2 != EXC_ERR_ARG) return rc2;
  *ngp = cnt;
  EXC_CHECK(hipMemcpyAsync(igpig, d_ig, 4 * n, hipMemcpyDeviceToHost, c->stream));
  EXC_CHECK(hipMemcpyAsync(vgpl, d_vgpl, 24 * n, hipMemcpyDeviceToHost, c->stream));
  EXC_CHECK(hipMemcpyAsync(vgpc, d_vgpc, 24 * n, hipMemcpyDeviceToHost, c->stream));
  EXC_CHECK(hipMemcpyAsync(gpc, d_gpc, 8 * n, hipMemcpyDeviceToHost, c->stream));
  EXC_CHECK(hipMemcpyAsync(tpgpc, d_tp, 16 * n, hipMemcpyDeviceToHost, c->stream));
  EXC_CHECK(hipStreamSynchronize(c->stream));
  return rc2;
}

int exc_gensfacgp(exc_ctx *c, int ngp, const double *vgpc, int ld, double *sfacgp) {
  if (!c || ngp < 0 || ld < ngp || (ngp > 0 && (!vgpc || !sfacgp))) return EXC_ERR_ARG;
  if (!c->have_atoms) return EXC_ERR_STATE;
  if (ngp == 0) return EXC_OK;
  int rc = use_device(c);
  if (rc) return rc;
  const size_t head = gk_head(1), nv = 3 * (size_t)ngp, ns = (size_t)ld * c->d.natm;
  if ((rc = grow(&c->gk_buf, &c->gk_buf_bytes, head + 8 * nv + 16 * ns + 256))) return rc;
  double *d_v = (double *)((char *)c->gk_buf + head);
  cplx *d_s = (cplx *)(d_v + nv);
  EXC_CHECK(hipMemcpyAsync(d_v, vgpc, 8 * nv, hipMemcpyHostToDevice, c->stream));
  EXC_CHECK(hipMemcpyAsync(d_s, sfacgp, 16 * ns, hipMemcpyHostToDevice, c->stream));   // rows past ngp kept
  if ((rc = exc_launch_gensfacgp(c->d, ngp, d_v, ld, d_s, c->stream))) return rc;
  EXC_CHECK(hipMemcpyAsync(sfacgp, d_s, 16 * ns, hipMemcpyDeviceToHost, c->stream));
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

}  // extern "C"
