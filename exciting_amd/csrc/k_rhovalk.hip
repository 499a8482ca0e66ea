// Muffin-tin density of the direct route, rhovalk (rhovalk.f90:99-207, spinless; C holds the
// states' coefficients): per atom a and coarse radial point irc (mesh point ir = irc lradstep),
//   psi_j(lm)  = sum_io W(a, io, lm; j) u_{io,l}(ir) + sum_ilo C(lo(a, ilo, m); j) v_ilo(ir)
//                (wavefmt.f90:91-124; W = A^T C, the zdotu of apwalm with the coefficients),
//   rfmt(itp)  = sum_j w occ_j |sum_lm zbshtvr(itp, lm) psi_j(lm)|^2   (|w occ_j| > epsocc),
//   rhomt(:, ir, a) += rfshtvr rfmt                                 (rhovalk.f90:183-207).
// The transform is linear, so the sum over states factors through the radial functions
// f = (io, l <= lmaxvr) and the atom's local orbitals (the same sampled-point density, summed
// in another order):
//   P(itp, f, j) = sum_m zbshtvr(itp, lm(l_f, m)) coef_f(m, j)     coef: W rows or lo rows of C
//   z(itp, ir, j) = sum_f u_f(ir) P(itp, f, j)
//   rfmt(itp, ir) = sum_j t_j |z|^2 = sum_{f <= f'} u_f(ir) u_f'(ir) R(itp, f, f'),
//   R = (2 - delta_ff') Re sum_j t_j P(itp, f, j) conj(P(itp, f', j)),   t_j = w occ_j,
// with t_j = 0 for |w occ_j| <= epsocc.  Three kernels: rv_p (per state and atom: the
// coefficients in LDS, P for every (itp, f)), rv_q (per point and atom: R over the states in
// chunks staged through LDS), rv_r (per 8 mesh points and atom: rfmt from R and u, then
// rfshtvr).  Flops per k-point 8 lmmaxvr (2 lmaxvr + 1) nf nst + 4 lmmaxvr nf^2 nst per atom
// and 2 lmmaxvr nf^2 + 2 lmmaxvr^2 per atom and mesh point, nf = apwordmax (lmaxvr + 1) +
// nlomax, instead of 8 lmmaxvr^2 per atom, mesh point and state.
#include "exc_common.h"

namespace {

constexpr int kRvThreads = 256;
constexpr int kRvChunk = 32;     // states per LDS stage of rv_q
constexpr int kRvPts = 8;        // mesh points per workgroup of rv_r
constexpr int kRvMaxM = 17;      // 2 lmaxvr + 1, lmaxvr <= 8

struct RvArgs {
  int lmv, lmaxvr, lradstep, nrmtmax, nlomax, lmaxapw, ao, naa, ldw, ngp, ldv, nst, nf, npair;
  const int *nrmt, *nlorb, *lorbl;   // per species; lorbl[ilo + nlomax is]
  const double *apwfr, *lofr;        // (nrmtmax, 2, ao, 0:lmaxapw, natm), (nrmtmax, 2, nlomax, natm)
  const cplx *w;                     // W(a naa + io + ao lm, j), leading dimension ldw
  const cplx *evec;                  // C(ldv, nst)
  const double *occ;
  double wk, eps;
  const cplx *zb;                    // zbshtvr(itp, lm), lmv x lmv column-major
  const double *rf;                  // rfshtvr(lm, itp), lmv x lmv column-major
  cplx *p;                           // P[((a lmv + itp) nst + j) nf + f]
  double *r;                         // R[(a npair + pair) lmv + itp], pairs f <= f' row by row
  double *rhomt;                     // (lmv, nrmtmax, natm)
};

// l of radial function f of species is (-1: none, an lo of another l range or beyond nlorb)
__device__ __forceinline__ int rv_fun_l(const RvArgs &p, int is, int f) {
  const int nfa = p.ao * (p.lmaxvr + 1);
  if (f < nfa) return f / p.ao;
  const int ilo = f - nfa;
  if (ilo >= p.nlorb[is]) return -1;
  const int l = p.lorbl[ilo + p.nlomax * is];
  return l <= p.lmaxvr ? l : -1;
}

__global__ __launch_bounds__(kRvThreads) void rv_p_kernel(exc_dev d, RvArgs p) {
  __shared__ cplx ca[4 * 81];            // W(a, io, lm; j) at io + ao lm (ao <= 4, lmv <= 81)
  __shared__ cplx cl[8 * kRvMaxM];       // lo coefficients (ilo, m + l), nlomax <= 8
  const int j = blockIdx.x, a = blockIdx.y, t = threadIdx.x;
  const int is = d.spidx[a], lmv = p.lmv, nfa = p.ao * (p.lmaxvr + 1);
  for (int e = t; e < p.ao * lmv; e += kRvThreads) ca[e] = p.w[(size_t)p.ldw * j + (size_t)a * p.naa + e];
  if (t == 0) {   // the atom's lo rows: (ilo, m), m fastest (genidxlo order)
    const int lo0 = p.ngp + d.dm_lo[2 * a];
    int off = 0;
    for (int ilo = 0; ilo < p.nlorb[is]; ++ilo) {
      const int l = p.lorbl[ilo + p.nlomax * is];
      if (l <= p.lmaxvr)
        for (int mm = 0; mm < 2 * l + 1; ++mm) cl[ilo * kRvMaxM + mm] = p.evec[lo0 + off + mm + (size_t)p.ldv * j];
      off += 2 * l + 1;
    }
  }
  __syncthreads();
  for (int e = t; e < lmv * p.nf; e += kRvThreads) {
    const int itp = e / p.nf, f = e - itp * p.nf;
    const int l = rv_fun_l(p, is, f);
    cplx v = {0.0, 0.0};
    if (l >= 0) {
      for (int mm = 0; mm < 2 * l + 1; ++mm) {
        const cplx b = p.zb[itp + (size_t)lmv * (l * l + mm)];
        const cplx c = f < nfa ? ca[(f % p.ao) + p.ao * (l * l + mm)] : cl[(f - nfa) * kRvMaxM + mm];
        v.re += b.re * c.re - b.im * c.im;
        v.im += b.re * c.im + b.im * c.re;
      }
    }
    p.p[(((size_t)a * lmv + itp) * p.nst + j) * p.nf + f] = v;
  }
}

__global__ __launch_bounds__(kRvThreads) void rv_q_kernel(exc_dev d, RvArgs p) {
  extern __shared__ cplx ps[];   // kRvChunk x nf, then kRvChunk weights
  const int itp = blockIdx.x, a = blockIdx.y, t = threadIdx.x, nf = p.nf;
  double *tw = (double *)(ps + kRvChunk * nf);
  // this thread's pairs f <= f' (at most 2 per thread: npair <= 2 kRvThreads)
  int f1[2] = {-1, -1}, g1[2] = {0, 0};
  for (int k = 0; k < 2; ++k) {
    int q = t + k * kRvThreads;
    if (q >= p.npair) break;
    int f = 0;
    while (q >= nf - f) {
      q -= nf - f;
      ++f;
    }
    f1[k] = f;
    g1[k] = f + q;
  }
  double acc[2] = {0.0, 0.0};
  const cplx *pa = p.p + ((size_t)a * p.lmv + itp) * p.nst * nf;
  for (int j0 = 0; j0 < p.nst; j0 += kRvChunk) {
    const int nj = min(kRvChunk, p.nst - j0);
    for (int e = t; e < nj * nf; e += kRvThreads) ps[e] = pa[(size_t)j0 * nf + e];
    if (t < nj) {
      const double t1 = p.wk * p.occ[j0 + t];
      tw[t] = fabs(t1) > p.eps ? t1 : 0.0;   // rhovalk.f90:101-102
    }
    __syncthreads();
    for (int k = 0; k < 2; ++k) {
      if (f1[k] < 0) continue;
      double s = 0.0;
      for (int jj = 0; jj < nj; ++jj) {
        const cplx x = ps[jj * nf + f1[k]], y = ps[jj * nf + g1[k]];
        s += tw[jj] * (x.re * y.re + x.im * y.im);
      }
      acc[k] += s;
    }
    __syncthreads();
  }
  for (int k = 0; k < 2; ++k) {
    if (f1[k] < 0) continue;
    const int q = t + k * kRvThreads;
    p.r[((size_t)a * p.npair + q) * p.lmv + itp] = (f1[k] == g1[k] ? 1.0 : 2.0) * acc[k];
  }
}

__global__ __launch_bounds__(128) void rv_r_kernel(exc_dev d, RvArgs p) {
  extern __shared__ double us[];   // u[kRvPts][nf], then rfmt[kRvPts][lmv]
  const int a = blockIdx.y, t = threadIdx.x, nf = p.nf, lmv = p.lmv;
  const int is = d.spidx[a], nr = p.nrmt[is];
  const int irc0 = blockIdx.x * kRvPts;
  if (irc0 * p.lradstep >= nr) return;   // uniform over the workgroup
  double *rfm = us + kRvPts * nf;
  const int nfa = p.ao * (p.lmaxvr + 1);
  for (int e = t; e < kRvPts * nf; e += 128) {
    const int r = e / nf, f = e - r * nf, ir = (irc0 + r) * p.lradstep;
    double u = 0.0;
    if (ir < nr && rv_fun_l(p, is, f) >= 0)
      u = f < nfa ? p.apwfr[ir + (size_t)p.nrmtmax * 2 * ((f % p.ao) + (size_t)p.ao * (f / p.ao + (size_t)(p.lmaxapw + 1) * a))]
                  : p.lofr[ir + (size_t)p.nrmtmax * 2 * ((f - nfa) + (size_t)p.nlomax * a)];
    us[e] = u;
  }
  __syncthreads();
  if (t < lmv) {
    double s[kRvPts];
    for (int r = 0; r < kRvPts; ++r) s[r] = 0.0;
    const double *ra = p.r + (size_t)a * p.npair * lmv + t;
    int q = 0;
    for (int f = 0; f < nf; ++f)
      for (int g = f; g < nf; ++g, ++q) {
        const double rv = ra[(size_t)q * lmv];
        for (int r = 0; r < kRvPts; ++r) s[r] += rv * us[r * nf + f] * us[r * nf + g];
      }
    for (int r = 0; r < kRvPts; ++r) rfm[r * lmv + t] = s[r];
  }
  __syncthreads();
  if (t < lmv) {   // rhomt(:, ir, a) += rfshtvr rfmt(:, ir)   (rhovalk.f90:183-207)
    for (int r = 0; r < kRvPts; ++r) {
      const int ir = (irc0 + r) * p.lradstep;
      if (ir >= nr) break;
      double v = 0.0;
      for (int q = 0; q < lmv; ++q) v += p.rf[t + (size_t)lmv * q] * rfm[r * lmv + q];
      p.rhomt[t + (size_t)lmv * (ir + (size_t)p.nrmtmax * a)] += v;
    }
  }
}

}  // namespace

size_t exc_rhovalk_work_bytes(const exc_dev &d, int lmaxvr, int nlomax, int nst) {
  const size_t lmv = (size_t)(lmaxvr + 1) * (lmaxvr + 1);
  const size_t nf = (size_t)d.apwordmax * (lmaxvr + 1) + (nlomax > 0 ? nlomax : 0);
  const size_t np = nf * (nf + 1) / 2;
  return ((sizeof(cplx) * d.natm * lmv * nst * nf + 255) & ~(size_t)255) + sizeof(double) * d.natm * np * lmv;
}

int exc_launch_rhovalk(const exc_dev &d, int lmaxvr, int lradstep, int nrmtmax, int nlomax, const int *nrmt,
                       const int *nlorb, const int *lorbl, const double *apwfr, const double *lofr, const cplx *w,
                       int ldw, int ngp, const cplx *evec, int ldv, int nst, const double *occ, double wk,
                       double eps, const cplx *zb, const double *rf, void *work, double *rhomt, hipStream_t s) {
  const int lmv = (lmaxvr + 1) * (lmaxvr + 1);
  if (nst <= 0) return EXC_OK;
  const int nl = nlomax > 0 ? nlomax : 0;
  RvArgs p;
  p.nf = d.apwordmax * (lmaxvr + 1) + nl;
  p.npair = p.nf * (p.nf + 1) / 2;
  if (lmaxvr > 8 || d.apwordmax > 4 || nl > 8 || p.npair > 2 * kRvThreads || lradstep < 1 || d.natm > 65535 ||
      nst > 65535)
    return EXC_ERR_ARG;
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
  p.p = (cplx *)work;
  p.r = (double *)((char *)work + ((sizeof(cplx) * d.natm * (size_t)lmv * nst * p.nf + 255) & ~(size_t)255));
  p.rhomt = rhomt;
  hipLaunchKernelGGL(rv_p_kernel, dim3(nst, d.natm), dim3(kRvThreads), 0, s, d, p);
  EXC_CHECK(hipGetLastError());
  const size_t lq = sizeof(cplx) * kRvChunk * p.nf + sizeof(double) * kRvChunk;
  hipLaunchKernelGGL(rv_q_kernel, dim3(lmv, d.natm), dim3(kRvThreads), lq, s, d, p);
  EXC_CHECK(hipGetLastError());
  const int nrc = (nrmtmax + lradstep - 1) / lradstep;
  const size_t lr = sizeof(double) * kRvPts * ((size_t)p.nf + lmv);
  hipLaunchKernelGGL(rv_r_kernel, dim3((nrc + kRvPts - 1) / kRvPts, d.natm), dim3(128), lr, s, d, p);
  EXC_CHECK(hipGetLastError());
  return EXC_OK;
}
