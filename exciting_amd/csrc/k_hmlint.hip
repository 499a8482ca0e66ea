// Muffin-tin Hamiltonian and overlap integrals of one SCF iteration, the producers of the
// setup's h_a (hmlint.f90:33-34 = mt_kin.f90 + mt_pot.f90; ValenceRelativity 'none',
// spinless) and of oalo / ololo (olprad.f90:90-125).
//
// Every radial integral of the reference is fderiv(-1) of a product on the species' mesh:
// the cubic spline of the product integrated interval by interval (fderiv.f90:64-72).  That
// functional is linear in the sampled values, so it is a fixed weight vector per mesh,
//     int f = sum_r wt_r f_r,
// computed once on the host by the same spline recurrences (spline.f90:46-96) applied to the
// unit vectors.  On the device:
//  * hml_radial_kernel, one workgroup per (atom, q): for q < lmmaxvr the potential integrals
//    P_q(a, b) = sum_r wt_r r^2 V_q(r) u_a u_b (mt_pot.f90:142-212), q = lmmaxvr the kinetic
//    ones K(a, b) = sum_r wt_r (1/2 u_a' u_b' r^2 + 1/2 l(l+1) u_a u_b) for l_a = l_b
//    (mt_kin.f90:120-167), q = lmmaxvr + 1 the overlaps S(a, b) = sum_r wt_r r^2 u_a u_b
//    (olprad.f90:94-125); functions a: APW (l, io) io fastest for l <= lmaxmat, then lo;
//  * hml_assemble_kernel, one thread per matrix element:
//      h(f1, f3) = sum_lm2 G(lm1, lm2, lm3) P_lm2(a1, a3) + [lm1 = lm3] K(a1, a3)
//    with G the Gaunt coefficient int conj(Y_lm1) R_lm2 Y_lm3 (gauntyry.f90, nonzero only for
//    m2 = +-(m1 - m3)), stored per (lm1, lm3) as [l2][sign of m2]; rows and columns in the
//    setup's order (APW (l, m, io) io fastest, lo (ilo, m) m fastest); and oalo, ololo from S.
// And the reverse, the muffin-tin density from a density matrix (genrhomt.f90:34-150):
//  * rho_factor_kernel: F_lm2(a, b) = w sum_{m1, m3} Re(dm(f1, f3) conj(G(lm1, lm2, lm3)))
//    (w = 2 for APW-lo pairs, the lo-APW block then 0), one thread per (atom, a, b, l2);
//  * rho_radial_kernel: rho_lm2(r) = sum_{a, b} F_lm2(a, b) u_a(r) u_b(r), one thread per
//    (atom, lm2, r), on the radial functions of the last hmlint call.
#include <math.h>

#include <vector>

#include "exc_common.h"

namespace {

constexpr int kHT = 256;

__device__ __forceinline__ int isqrt_lm(int lm) {
  int l = (int)sqrt((double)lm);
  while (l * l > lm) --l;
  while ((l + 1) * (l + 1) <= lm) ++l;
  return l;
}

struct HmlArgs {
  int natm, ao, lmaxapw, lmaxmat, naf, nf, nlomax, lmv, nrmtmax, naa, maxnlo;
  const int *spidx, *nrmt, *nlorb, *lorbl;   // lorbl[is * nlomax + ilo]
  const double *wt, *spr, *apwfr, *lofr, *veffmt;
  double *rad;                                // [natm][lmv + 2][nf][nf]
  const cplx *gnt;                            // [lmmaxmat][lmmaxmat][lmaxvr + 1][2]
  const lo_desc *lod;
  const int *dm_lo;                           // [natm][2]: first lo, count
  cplx *haa, *loa, *lolo;
  double *oalo, *ololo;
  const cplx *dm;                             // genrhomt: dm(ws, ws, natm)
  int ws;
  double *rho;                                // genrhomt: rhomt(lmv, nrmtmax, natm)
};

__device__ __forceinline__ const double *radial_u(const HmlArgs &A, int atom, int a, int *l) {
  if (a < A.naf) {
    const int io = a % A.ao, la = a / A.ao;
    *l = la;
    return A.apwfr + (size_t)A.nrmtmax * 2 * (io + (size_t)A.ao * (la + (size_t)(A.lmaxapw + 1) * atom));
  }
  const int ilo = a - A.naf, is = A.spidx[atom];
  if (ilo >= A.nlorb[is]) return nullptr;
  *l = A.lorbl[is * A.nlomax + ilo];
  return A.lofr + (size_t)A.nrmtmax * 2 * (ilo + (size_t)A.nlomax * atom);
}

__global__ __launch_bounds__(kHT) void hml_radial_kernel(HmlArgs A) {
  const int q = blockIdx.x, atom = blockIdx.y;
  const int is = A.spidx[atom], nr = A.nrmt[is];
  extern __shared__ double sw[];   // [nr] weights of this block's integrand
  const double *wt = A.wt + (size_t)A.nrmtmax * is, *x = A.spr + (size_t)A.nrmtmax * is;
  for (int r = threadIdx.x; r < nr; r += kHT) {
    const double r2 = x[r] * x[r];
    double v;
    if (q < A.lmv) v = wt[r] * r2 * A.veffmt[q + (size_t)A.lmv * (r + (size_t)A.nrmtmax * atom)];
    else if (q == A.lmv) v = wt[r];          // kinetic: the r^2 and l(l+1) factors per pair
    else v = wt[r] * r2;
    sw[r] = v;
  }
  __syncthreads();
  double *out = A.rad + ((size_t)atom * (A.lmv + 2) + q) * A.nf * A.nf;
  for (int p = threadIdx.x; p < A.nf * A.nf; p += kHT) {
    const int a = p / A.nf, b = p % A.nf;
    int la = 0, lb = 0;
    const double *ua = radial_u(A, atom, a, &la), *ub = radial_u(A, atom, b, &lb);
    double s = 0.0;
    if (ua && ub && (q != A.lmv || la == lb)) {
      if (q == A.lmv) {
        const double ang = 0.5 * la * (la + 1);
        const double *dua = ua + A.nrmtmax, *dub = ub + A.nrmtmax;
        for (int r = 0; r < nr; ++r) {
          const double r2 = x[r] * x[r];
          s = fma(sw[r], fma(0.5 * r2, dua[r] * dub[r], ang * (ua[r] * ub[r])), s);
        }
      } else {
        for (int r = 0; r < nr; ++r) s = fma(sw[r], ua[r] * ub[r], s);
      }
    }
    out[a * A.nf + b] = s;
  }
}

// the Gaunt sum + kinetic term of one element; a1, a3 radial function indices
__device__ cplx hml_element(const HmlArgs &A, int atom, int l1, int m1, int a1, int l3, int m3, int a3) {
  const int lmm = (A.lmaxmat + 1) * (A.lmaxmat + 1), lmaxvr = isqrt_lm(A.lmv - 1);
  const int lm1 = l1 * l1 + l1 + m1, lm3 = l3 * l3 + l3 + m3;
  const cplx *g = A.gnt + ((size_t)lm1 * lmm + lm3) * (lmaxvr + 1) * 2;
  const double *P = A.rad + (size_t)atom * (A.lmv + 2) * A.nf * A.nf;
  const size_t ab = (size_t)a1 * A.nf + a3, qs = (size_t)A.nf * A.nf;
  const int d = m1 > m3 ? m1 - m3 : m3 - m1;
  cplx z = {0.0, 0.0};
  for (int l2 = d; l2 <= lmaxvr; ++l2) {
    const double pp = P[(size_t)(l2 * l2 + l2 + d) * qs + ab], pm = P[(size_t)(l2 * l2 + l2 - d) * qs + ab];
    const cplx gp = g[2 * l2], gm = g[2 * l2 + 1];
    z.re = fma(gp.re, pp, fma(gm.re, pm, z.re));
    z.im = fma(gp.im, pp, fma(gm.im, pm, z.im));
  }
  if (lm1 == lm3) z.re += P[(size_t)A.lmv * qs + ab];
  return z;
}

__global__ __launch_bounds__(kHT) void hml_assemble_kernel(HmlArgs A) {
  const int atom = blockIdx.y;
  const size_t e = (size_t)blockIdx.x * kHT + threadIdx.x;
  const size_t n_aa = (size_t)A.naa * A.naa, n_la = (size_t)A.maxnlo * A.naa, n_ll = (size_t)A.maxnlo * A.maxnlo;
  const size_t n_oa = (size_t)A.ao * A.nlomax, n_oo = (size_t)A.nlomax * A.nlomax;
  const int first = A.dm_lo[2 * atom], cnt = A.dm_lo[2 * atom + 1];
  auto apw = [&](int f, int *l, int *m, int *a) {
    const int io = f % A.ao, lm = f / A.ao;
    *l = isqrt_lm(lm);
    *m = lm - *l * *l - *l;
    *a = io + A.ao * *l;
  };
  if (e < n_aa) {
    const int f1 = (int)(e % A.naa), f3 = (int)(e / A.naa);
    int l1, m1, a1, l3, m3, a3;
    apw(f1, &l1, &m1, &a1);
    apw(f3, &l3, &m3, &a3);
    A.haa[(size_t)atom * n_aa + e] = hml_element(A, atom, l1, m1, a1, l3, m3, a3);
    return;
  }
  size_t i = e - n_aa;
  if (i < n_la) {
    const int r = (int)(i % A.maxnlo), f3 = (int)(i / A.maxnlo);
    cplx z = {0.0, 0.0};
    if (r < cnt) {
      const lo_desc q = A.lod[first + r];
      int l3, m3, a3;
      apw(f3, &l3, &m3, &a3);
      z = hml_element(A, atom, q.l, q.m, A.naf + q.ilo, l3, m3, a3);
    }
    A.loa[(size_t)atom * n_la + i] = z;
    return;
  }
  i -= n_la;
  if (i < n_ll) {
    const int r1 = (int)(i % A.maxnlo), r3 = (int)(i / A.maxnlo);
    cplx z = {0.0, 0.0};
    if (r1 < cnt && r3 < cnt) {
      const lo_desc q1 = A.lod[first + r1], q3 = A.lod[first + r3];
      z = hml_element(A, atom, q1.l, q1.m, A.naf + q1.ilo, q3.l, q3.m, A.naf + q3.ilo);
    }
    A.lolo[(size_t)atom * n_ll + i] = z;
    return;
  }
  i -= n_ll;
  const double *S = A.rad + ((size_t)atom * (A.lmv + 2) + A.lmv + 1) * A.nf * A.nf;
  const int is = A.spidx[atom];
  if (i < n_oa) {
    const int io = (int)(i % A.ao), ilo = (int)(i / A.ao);
    double v = 0.0;
    if (ilo < A.nlorb[is]) v = S[(size_t)(io + A.ao * A.lorbl[is * A.nlomax + ilo]) * A.nf + A.naf + ilo];
    A.oalo[(size_t)atom * n_oa + i] = v;
    return;
  }
  i -= n_oa;
  if (i < n_oo) {
    const int i1 = (int)(i % A.nlomax), i2 = (int)(i / A.nlomax);
    double v = 0.0;
    if (i1 < A.nlorb[is] && i2 < A.nlorb[is] && A.lorbl[is * A.nlomax + i1] == A.lorbl[is * A.nlomax + i2])
      v = S[(size_t)(A.naf + i1) * A.nf + A.naf + i2];
    A.ololo[(size_t)atom * n_oo + i] = v;
  }
}

// row of function a's m = -l in the density matrix: APW (l, io) at apwordmax l^2 + io (2l+1)
// (rows (l, io, m) m fastest, gendmatmt.f90:92-101), lo ilo after maxaa (ilo, m)
__device__ __forceinline__ int dm_row(const HmlArgs &A, int atom, int a) {
  if (a < A.naf) {
    const int io = a % A.ao, l = a / A.ao;
    return A.ao * l * l + io * (2 * l + 1);
  }
  const int is = A.spidx[atom];
  int r = A.naa;
  for (int i = 0; i < a - A.naf; ++i) r += 2 * A.lorbl[is * A.nlomax + i] + 1;
  return r;
}

constexpr int kMaxL2 = 16;   // lmaxvr supported by genrhomt

// F[atom][lm2][a][b] = w sum_{m1, m3} Re(dm(f1, f3) conj(G(lm1, lm2, lm3))) for the l2 of this
// thread (genrhomt.f90:60-71, 96-107, 131-142; w = 2 for APW-lo, the lo-APW block is 0)
__global__ __launch_bounds__(kHT) void rho_factor_kernel(HmlArgs A) {
  const int atom = blockIdx.y, L2 = isqrt_lm(A.lmv - 1) + 1;
  const int t = blockIdx.x * kHT + threadIdx.x;
  if (t >= A.nf * A.nf * L2) return;
  const int l2 = t % L2, ab = t / L2, a = ab / A.nf, b = ab % A.nf;
  const int lmm = (A.lmaxmat + 1) * (A.lmaxmat + 1);
  double *F = A.rad + (size_t)atom * A.lmv * A.nf * A.nf + (size_t)a * A.nf + b;
  const size_t qs = (size_t)A.nf * A.nf;
  double acc[2 * kMaxL2 + 1];
  for (int i = 0; i <= 2 * l2; ++i) acc[i] = 0.0;
  int la = 0, lb = 0;
  const double *ua = radial_u(A, atom, a, &la), *ub = radial_u(A, atom, b, &lb);
  if (ua && ub && !(a >= A.naf && b < A.naf)) {
    const double w = (a < A.naf && b >= A.naf) ? 2.0 : 1.0;
    const int o1 = dm_row(A, atom, a), o3 = dm_row(A, atom, b);
    const cplx *dm = A.dm + (size_t)atom * A.ws * A.ws;
    for (int m1 = -la; m1 <= la; ++m1)
      for (int m3 = -lb; m3 <= lb; ++m3) {
        const int d = m1 > m3 ? m1 - m3 : m3 - m1;
        if (d > l2) continue;
        const cplx z = dm[(o1 + m1 + la) + (size_t)A.ws * (o3 + m3 + lb)];
        const cplx *g = A.gnt + (((size_t)(la * la + la + m1) * lmm + (lb * lb + lb + m3)) * L2 + l2) * 2;
        acc[l2 + d] = fma(z.re, g[0].re, fma(z.im, g[0].im, acc[l2 + d]));
        if (d > 0) acc[l2 - d] = fma(z.re, g[1].re, fma(z.im, g[1].im, acc[l2 - d]));
      }
    for (int i = 0; i <= 2 * l2; ++i) acc[i] *= w;
  }
  for (int m2 = -l2; m2 <= l2; ++m2) F[(size_t)(l2 * l2 + l2 + m2) * qs] = acc[l2 + m2];
}

// rhomt(lm2, r, atom) = sum_{a, b} F[atom][lm2][a][b] u_a(r) u_b(r) (genrhomt.f90:72-77)
__global__ __launch_bounds__(kHT) void rho_radial_kernel(HmlArgs A) {
  const int r = blockIdx.x * kHT + threadIdx.x, lm2 = blockIdx.y, atom = blockIdx.z;
  if (r >= A.nrmtmax) return;
  const int nr = A.nrmt[A.spidx[atom]];
  double s = 0.0;
  if (r < nr) {
    const double *F = A.rad + ((size_t)atom * A.lmv + lm2) * A.nf * A.nf;
    for (int a = 0; a < A.nf; ++a) {
      int la;
      const double *ua = radial_u(A, atom, a, &la);
      if (!ua) continue;
      const double va = ua[r];
      double t = 0.0;
      for (int b = 0; b < A.nf; ++b) {
        int lb;
        const double *ub = radial_u(A, atom, b, &lb);
        if (ub) t = fma(F[a * A.nf + b], ub[r], t);
      }
      s = fma(va, t, s);
    }
  }
  A.rho[lm2 + (size_t)A.lmv * (r + (size_t)A.nrmtmax * atom)] = s;
}

// ------------------------------------------------------------------ host helpers

long double lfact(int n) {
  long double f = 1.0L;
  for (int i = 2; i <= n; ++i) f *= i;
  return f;
}

// (j1 j2 j3; m1 m2 m3) by the Racah formula
double wigner3j(int j1, int j2, int j3, int m1, int m2, int m3) {
  if (m1 + m2 + m3 != 0 || abs(m1) > j1 || abs(m2) > j2 || abs(m3) > j3) return 0.0;
  if (j3 < abs(j1 - j2) || j3 > j1 + j2) return 0.0;
  const long double tri = lfact(j1 + j2 - j3) * lfact(j1 - j2 + j3) * lfact(-j1 + j2 + j3) / lfact(j1 + j2 + j3 + 1);
  const long double pre = tri * lfact(j1 + m1) * lfact(j1 - m1) * lfact(j2 + m2) * lfact(j2 - m2) * lfact(j3 + m3) *
                          lfact(j3 - m3);
  long double s = 0.0L;
  for (int k = 0; k <= j1 + j2 + j3; ++k) {
    const int d[6] = {k, j3 - j2 + k + m1, j3 - j1 + k - m2, j1 + j2 - j3 - k, j1 - k - m1, j2 - k + m2};
    bool ok = true;
    long double den = 1.0L;
    for (int v : d) {
      if (v < 0) {
        ok = false;
        break;
      }
      den *= lfact(v);
    }
    if (ok) s += ((k & 1) ? -1.0L : 1.0L) / den;
  }
  const long double sign = ((j1 - j2 - m3) & 1) ? -1.0L : 1.0L;
  return (double)(sign * sqrtl(pre) * s);
}

// int conj(Y_l1m1) Y_l2m2 Y_l3m3 dOmega
double gaunt_c(int l1, int l2, int l3, int m1, int m2, int m3) {
  if (m1 - m2 - m3 != 0 || ((l1 + l2 + l3) & 1)) return 0.0;
  const double t = sqrt((2.0 * l1 + 1.0) * (2.0 * l2 + 1.0) * (2.0 * l3 + 1.0) / (4.0 * M_PI));
  return ((m1 & 1) ? -t : t) * wigner3j(l1, l2, l3, 0, 0, 0) * wigner3j(l1, l2, l3, -m1, m2, m3);
}

// int conj(Y_l1m1) R_l2m2 Y_l3m3 dOmega, R_lm = sqrt2 Re Y_lm (m > 0), Y_l0, sqrt2 Im Y_lm (m < 0)
cplx gaunt_yry(int l1, int l2, int l3, int m1, int m2, int m3) {
  const double c1 = 0.70710678118654752440;
  if (m2 > 0) {
    const double t1 = (m2 % 2 == 0) ? 1.0 : -1.0;
    return {c1 * (gaunt_c(l1, l2, l3, m1, m2, m3) + t1 * gaunt_c(l1, l2, l3, m1, -m2, m3)), 0.0};
  }
  if (m2 < 0) {
    const double t1 = ((-m2) % 2 == 0) ? 1.0 : -1.0;
    return {0.0, -c1 * (gaunt_c(l1, l2, l3, m1, m2, m3) - t1 * gaunt_c(l1, l2, l3, m1, -m2, m3))};
  }
  return {gaunt_c(l1, l2, l3, m1, m2, m3), 0.0};
}

}  // namespace

// Integration weights of fderiv(-1) on the mesh x (n >= 4 points): the spline recurrences run
// on all n unit vectors at once (column j of the right-hand side is e_j).
void exc_hml_spline_weights(int n, const double *x, double *wt) {
  std::vector<double> w(n), c1((size_t)n * n, 0.0), c3((size_t)n * n, 0.0), c2s(n);
  auto F = [&](int i, int j) { return i == j ? 1.0 : 0.0; };
  for (int i = 0; i < n - 1; ++i) {
    w[i] = 1.0 / (x[i + 1] - x[i]);
    for (int j = 0; j < n; ++j) c1[(size_t)i * n + j] = w[i] * (F(i + 1, j) - F(i, j));
  }
  c2s[0] = 1.0;
  for (int j = 0; j < n; ++j) c3[j] = (c1[n + j] - c1[j]) / (x[2] - x[0]);
  double t1 = (x[1] - x[0]) * w[1], t2 = t1 * c2s[0], t3 = 1.0 / (2.0 * (t1 + 1.0));
  c2s[1] = t3;
  std::vector<double> t4(n);
  for (int j = 0; j < n; ++j) {
    t4[j] = 3.0 * (c1[n + j] - c1[j]) * w[1] - t2 * c3[j];
    c3[n + j] = t4[j];
  }
  for (int i = 2; i < n - 1; ++i) {
    t1 = (x[i] - x[i - 1]) * w[i];
    t2 = t1 * t3;
    t3 = 1.0 / (2.0 * t1 + 2.0 - t2);
    c2s[i] = t3;
    for (int j = 0; j < n; ++j) {
      t4[j] = 3.0 * (c1[(size_t)i * n + j] - c1[(size_t)(i - 1) * n + j]) * w[i] - t2 * t4[j];
      c3[(size_t)i * n + j] = t4[j];
    }
  }
  c2s[n - 1] = 1.0;
  std::vector<double> tb(n);
  for (int j = 0; j < n; ++j) {
    tb[j] = (c1[(size_t)(n - 2) * n + j] - c1[(size_t)(n - 3) * n + j]) / (x[n - 1] - x[n - 3]);
    c3[(size_t)(n - 1) * n + j] = tb[j];
  }
  for (int i = n - 2; i >= 1; --i)
    for (int j = 0; j < n; ++j) {
      tb[j] = c3[(size_t)i * n + j] - tb[j] * c2s[i + 1];
      c3[(size_t)i * n + j] = tb[j];
    }
  // c2 = c2s * c3; then c3, c1 of each interval, and the integral
  for (int j = 0; j < n; ++j) wt[j] = 0.0;
  for (int i = 0; i < n - 1; ++i) {
    const double dx = x[i + 1] - x[i];
    for (int j = 0; j < n; ++j) {
      const double c2i = c2s[i] * c3[(size_t)i * n + j], c2n = c2s[i + 1] * c3[(size_t)(i + 1) * n + j];
      const double c3i = 0.3333333333333333333 * (c2n - c2i) * w[i];
      const double c1i = c1[(size_t)i * n + j] - (c2i + c3i * dx) * dx;
      wt[j] += (((0.25 * c3i * dx + 0.3333333333333333333 * c2i) * dx + 0.5 * c1i) * dx + F(i, j)) * dx;
    }
  }
}

// G[((lm1 * lmmaxmat + lm3) * (lmaxvr + 1) + l2) * 2 + s] = gauntyry(l1, l2, l3, m1, m2, m3)
// with m2 = +d (s = 0) and m2 = -d (s = 1, zero when d = 0), d = |m1 - m3|
void exc_hml_gaunt_table(int lmaxmat, int lmaxvr, cplx *g) {
  const int lmm = (lmaxmat + 1) * (lmaxmat + 1);
  for (int l1 = 0; l1 <= lmaxmat; ++l1)
    for (int m1 = -l1; m1 <= l1; ++m1)
      for (int l3 = 0; l3 <= lmaxmat; ++l3)
        for (int m3 = -l3; m3 <= l3; ++m3) {
          const int lm1 = l1 * l1 + l1 + m1, lm3 = l3 * l3 + l3 + m3, d = abs(m1 - m3);
          cplx *o = g + ((size_t)lm1 * lmm + lm3) * (lmaxvr + 1) * 2;
          for (int l2 = 0; l2 <= lmaxvr; ++l2) {
            o[2 * l2] = o[2 * l2 + 1] = {0.0, 0.0};
            if (d > l2) continue;
            o[2 * l2] = gaunt_yry(l1, l2, l3, m1, d, m3);
            if (d > 0) o[2 * l2 + 1] = gaunt_yry(l1, l2, l3, m1, -d, m3);
          }
        }
}

namespace {
HmlArgs hml_args(const exc_dev &d, int lmv, int nrmtmax, const int *nrmt, const int *nlorb, const int *lorbl,
                 const double *wt, const double *spr, const double *apwfr, const double *lofr, const double *veffmt,
                 double *rad, const cplx *gnt) {
  HmlArgs A;
  A.natm = d.natm;
  A.ao = d.apwordmax;
  A.lmaxapw = d.lmaxapw;
  A.lmaxmat = d.lmaxmat;
  A.naf = d.apwordmax * (d.lmaxmat + 1);
  A.nlomax = d.nlomax > 0 ? d.nlomax : 0;
  A.nf = A.naf + A.nlomax;
  A.lmv = lmv;
  A.nrmtmax = nrmtmax;
  A.naa = d.naa;
  A.maxnlo = d.maxnlo;
  A.spidx = d.spidx;
  A.nrmt = nrmt;
  A.nlorb = nlorb;
  A.lorbl = lorbl;
  A.wt = wt;
  A.spr = spr;
  A.apwfr = apwfr;
  A.lofr = lofr;
  A.veffmt = veffmt;
  A.rad = rad;
  A.gnt = gnt;
  A.lod = d.lod;
  A.dm_lo = d.dm_lo;
  A.haa = d.haa;
  A.loa = d.loa;
  A.lolo = d.lolo;
  A.oalo = d.oalo;
  A.ololo = d.ololo;
  A.dm = nullptr;
  A.ws = 0;
  A.rho = nullptr;
  return A;
}
}  // namespace

int exc_launch_hmlint(const exc_dev &d, int lmv, int nrmtmax, const int *nrmt, const int *nlorb, const int *lorbl,
                      const double *wt, const double *spr, const double *apwfr, const double *lofr,
                      const double *veffmt, double *rad, const cplx *gnt, hipStream_t s) {
  HmlArgs A = hml_args(d, lmv, nrmtmax, nrmt, nlorb, lorbl, wt, spr, apwfr, lofr, veffmt, rad, gnt);
  if (d.natm > 65535 || lmv + 2 > 65535) return EXC_ERR_ARG;
  if ((size_t)nrmtmax * sizeof(double) > 64 * 1024) return EXC_ERR_ARG;
  hipLaunchKernelGGL(hml_radial_kernel, dim3(lmv + 2, d.natm), dim3(kHT), nrmtmax * sizeof(double), s, A);
  EXC_CHECK(hipGetLastError());
  const size_t n = (size_t)d.naa * d.naa + (size_t)d.maxnlo * d.naa + (size_t)d.maxnlo * d.maxnlo +
                   (size_t)A.ao * A.nlomax + (size_t)A.nlomax * A.nlomax;
  hipLaunchKernelGGL(hml_assemble_kernel, dim3((unsigned)((n + kHT - 1) / kHT), d.natm), dim3(kHT), 0, s, A);
  EXC_CHECK(hipGetLastError());
  return EXC_OK;
}

int exc_launch_genrhomt(const exc_dev &d, int lmv, int nrmtmax, const int *nrmt, const int *nlorb, const int *lorbl,
                        const double *apwfr, const double *lofr, const cplx *gnt, const cplx *dm, int ws, double *fbuf,
                        double *rhomt, hipStream_t s) {
  HmlArgs A = hml_args(d, lmv, nrmtmax, nrmt, nlorb, lorbl, nullptr, nullptr, apwfr, lofr, nullptr, fbuf, gnt);
  A.dm = dm;
  A.ws = ws;
  A.rho = rhomt;
  const int L2 = (int)lround(sqrt((double)lmv));
  if (L2 - 1 > kMaxL2 || d.natm > 65535 || lmv > 65535 || ws < d.naa + A.maxnlo) return EXC_ERR_ARG;
  const size_t nt = (size_t)A.nf * A.nf * L2;
  hipLaunchKernelGGL(rho_factor_kernel, dim3((unsigned)((nt + kHT - 1) / kHT), d.natm), dim3(kHT), 0, s, A);
  EXC_CHECK(hipGetLastError());
  hipLaunchKernelGGL(rho_radial_kernel, dim3((nrmtmax + kHT - 1) / kHT, lmv, d.natm), dim3(kHT), 0, s, A);
  EXC_CHECK(hipGetLastError());
  return EXC_OK;
}
