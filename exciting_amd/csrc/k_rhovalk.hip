// Muffin-tin density of the direct route, rhovalk (rhovalk.f90:99-207, spinless, first-
// variational states): per atom a and coarse radial point irc (mesh point ir = irc lradstep),
//   psi_j(lm)  = sum_io W(a, io, lm; j) u_{io,l}(ir) + sum_ilo C(lo(a, ilo, m); j) v_ilo(ir)
//                (wavefmt.f90:91-124; W = A^T C, the zdotu of apwalm with the coefficients),
//   rfmt(itp)  = sum_j w occ_j |sum_lm zbshtvr(itp, lm) psi_j(lm)|^2   (|w occ_j| > epsocc),
//   rhomt(:, ir, a) += rfshtvr rfmt                                 (rhovalk.f90:183-207).
// One workgroup of 256 threads per (irc, atom) loops over the states kRvS at a time: the
// backward transform's matrix sits in LDS, the pass's psi_j in LDS beside it; thread
// (itp, g), itp = t mod 128 and g = t / 128, sums |z|^2 of the pass's states g, g + 2, ...
// in registers, so each matrix element read from LDS serves kRvS / 2 complex products.  The
// two halves' sums meet in LDS at the end, the forward transform follows, and each
// workgroup writes its own column of rhomt.  States with |w occ| <= epsocc are left out.
#include "exc_common.h"

namespace {

constexpr int kRvThreads = 256;       // 2 x 128 >= 2 x lmmaxvr
constexpr int kRvS = 16;              // states per pass
constexpr int kRvMaxAo = 4;
constexpr size_t kRvMaxLds = 163840;  // gfx950's LDS per workgroup: lmaxvr <= 8 (81 x 81 complex)

struct RvArgs {
  int lmv, lmaxvr, lradstep, nrmtmax, nlomax, lmaxapw, ao, naa, ldw, ngp, ldv, nst;
  const int *nrmt, *nlorb, *lorbl;   // per species; lorbl[ilo + nlomax is]
  const double *apwfr, *lofr;        // (nrmtmax, 2, ao, 0:lmaxapw, natm), (nrmtmax, 2, nlomax, natm)
  const cplx *w;                     // W(a naa + io + ao lm, j), leading dimension ldw
  const cplx *evec;                  // C(ldv, nst)
  const double *occ;
  double wk, eps;
  const cplx *zb;                    // zbshtvr(itp, lm), lmv x lmv column-major
  const double *rf;                  // rfshtvr(lm, itp), lmv x lmv column-major
  double *rhomt;                     // (lmv, nrmtmax, natm)
};

__global__ __launch_bounds__(kRvThreads) void rhovalk_kernel(exc_dev d, RvArgs p) {
  extern __shared__ cplx zbs[];   // lmv * lmv, then psi (kRvS x lmv), then 2 x 128 partial sums
  const int irc = blockIdx.x, a = blockIdx.y, t = threadIdx.x;
  const int is = d.spidx[a], lmv = p.lmv;
  const int ir = irc * p.lradstep;
  if (ir >= p.nrmt[is]) return;   // uniform over the workgroup
  cplx *psi = zbs + lmv * lmv;
  double *part = (double *)(psi + kRvS * lmv);
  for (int i = t; i < lmv * lmv; i += kRvThreads) zbs[i] = p.zb[i];
  const int itp = t & 127, g = t >> 7;
  const int lo0 = p.ngp + d.dm_lo[2 * a];   // the atom's first lo row of C
  const int nlo = p.nlorb[is];
  const double *ua = p.apwfr + ir + (size_t)p.nrmtmax * 2 * p.ao * (size_t)(p.lmaxapw + 1) * a;
  double acc = 0.0;
  for (int j0 = 0; j0 < p.nst; j0 += kRvS) {
    // psi(s, lm) of the pass's states (wavefmt.f90:91-124)
    for (int e = t; e < kRvS * lmv; e += kRvThreads) {
      const int s = e / lmv, lm = e - s * lmv, j = j0 + s;
      cplx v = {0.0, 0.0};
      if (j < p.nst) {
        int l = 0;
        while ((l + 1) * (l + 1) <= lm) ++l;
        const int m = lm - l * l - l;
        const cplx *wj = p.w + (size_t)p.ldw * j + (size_t)a * p.naa + p.ao * lm;
        for (int io = 0; io < p.ao; ++io) {
          const double u = ua[(size_t)p.nrmtmax * 2 * (io + (size_t)p.ao * l)];
          v.re += wj[io].re * u;
          v.im += wj[io].im * u;
        }
        int off = 0;   // lo rows of the atom: (ilo, m), m fastest
        for (int ilo = 0; ilo < nlo; ++ilo) {
          const int ll = p.lorbl[ilo + p.nlomax * is];
          if (ll == l) {
            const double ul = p.lofr[ir + (size_t)p.nrmtmax * 2 * (ilo + (size_t)p.nlomax * a)];
            const cplx c = p.evec[lo0 + off + m + l + (size_t)p.ldv * j];
            v.re += c.re * ul;
            v.im += c.im * ul;
          }
          off += 2 * ll + 1;
        }
      }
      psi[e] = v;
    }
    __syncthreads();
    if (itp < lmv) {
      double zr[kRvS / 2], zi[kRvS / 2];
      for (int s = 0; s < kRvS / 2; ++s) zr[s] = zi[s] = 0.0;
      for (int q = 0; q < lmv; ++q) {
        const cplx b = zbs[itp + lmv * q];
        for (int s = 0; s < kRvS / 2; ++s) {
          const cplx c = psi[(g + 2 * s) * lmv + q];
          zr[s] += b.re * c.re - b.im * c.im;
          zi[s] += b.re * c.im + b.im * c.re;
        }
      }
      for (int s = 0; s < kRvS / 2; ++s) {
        const int j = j0 + g + 2 * s;
        if (j < p.nst) {
          const double t1 = p.wk * p.occ[j];
          if (fabs(t1) > p.eps) acc += t1 * (zr[s] * zr[s] + zi[s] * zi[s]);   // rhovalk.f90:101-102
        }
      }
    }
    __syncthreads();
  }
  part[t] = acc;
  __syncthreads();
  if (t < lmv) {   // rflm = rfshtvr (rfmt of both halves)
    double s = 0.0;
    for (int q = 0; q < lmv; ++q) s += p.rf[t + (size_t)lmv * q] * (part[q] + part[128 + q]);
    p.rhomt[t + (size_t)lmv * (ir + (size_t)p.nrmtmax * a)] += s;
  }
}

}  // namespace

int exc_launch_rhovalk(const exc_dev &d, int lmaxvr, int lradstep, int nrmtmax, int nlomax, const int *nrmt,
                       const int *nlorb, const int *lorbl, const double *apwfr, const double *lofr, const cplx *w,
                       int ldw, int ngp, const cplx *evec, int ldv, int nst, const double *occ, double wk,
                       double eps, const cplx *zb, const double *rf, double *rhomt, hipStream_t s) {
  const int lmv = (lmaxvr + 1) * (lmaxvr + 1);
  if (nst <= 0) return EXC_OK;
  if (lmv > kRvThreads / 2 || d.apwordmax > kRvMaxAo || lradstep < 1 || d.natm > 65535) return EXC_ERR_ARG;
  RvArgs p;
  p.lmv = lmv;
  p.lmaxvr = lmaxvr;
  p.lradstep = lradstep;
  p.nrmtmax = nrmtmax;
  p.nlomax = nlomax > 0 ? nlomax : 1;
  p.lmaxapw = d.lmaxapw;
  p.ao = d.apwordmax;
  p.naa = d.naa;
  p.ldw = ldw;
  p.ngp = ngp;
  p.ldv = ldv;
  p.nst = nst;
  p.nrmt = nrmt;
  p.nlorb = nlorb;
  p.lorbl = lorbl;
  p.apwfr = apwfr;
  p.lofr = lofr;
  p.w = w;
  p.evec = evec;
  p.occ = occ;
  p.wk = wk;
  p.eps = eps;
  p.zb = zb;
  p.rf = rf;
  p.rhomt = rhomt;
  const int nrc = (nrmtmax + lradstep - 1) / lradstep;
  const size_t lds = sizeof(cplx) * ((size_t)lmv * lmv + (size_t)kRvS * lmv) + sizeof(double) * kRvThreads;
  if (lds > kRvMaxLds) return EXC_ERR_ARG;
  hipLaunchKernelGGL(rhovalk_kernel, dim3(nrc, d.natm), dim3(kRvThreads), lds, s, d, p);
  EXC_CHECK(hipGetLastError());
  return EXC_OK;
}
