// Device code of the APW matching coefficients (replaces match.f90:83-187), shared by the
// per-k and the grouped (all k-points) kernels.
//
// One thread per (G+k vector, atom).  The thread evaluates j_l(|G+k|R) and
// |G+k| j_l'(|G+k|R) for l <= L, the complex spherical harmonics Y_lm(G+k)
// (Condon–Shortley phase) by the three-term Legendre recurrence in l at fixed m,
// and writes
//   apwalm(ig, io, lm, ias) = (4 pi i^l / sqrt(Omega)) S(ig, ias) conj(Y_lm)
//                             * (Dinv(io,0) j_l + Dinv(io,1) |G+k| j_l')
// where Dinv is the inverse of the 2x2 matching matrix of (ias, l) (the zgesv solve of
// match.f90:163 done once per SCF iteration on the host).  Consecutive threads take
// consecutive ig, so every store of a fixed (io, lm, ias) column is coalesced.
#pragma once
#include "exc_common.h"

namespace match_dev {

// j_l(x), l = 0..L+1, into jl[]: series for x < 1e-8, upward recurrence for x > L,
// Miller downward recurrence (normalised by j_0 or j_1 in closed form) otherwise.
template <int L>
__device__ __forceinline__ void sbessel_seq(double x, double (&jl)[L + 2]) {
  // no FMA contraction: the per-k and the grouped match kernels inline this code into
  // different surroundings and must still round identically
#pragma clang fp contract(off)
  constexpr int LL = L + 1;
  if (x < 1e-8) {
    double t = 1.0, dfac = 1.0;
    jl[0] = 1.0;
#pragma unroll
    for (int l = 1; l <= LL; ++l) {
      dfac *= (double)(2 * l + 1);
      t *= x;
      jl[l] = t / dfac;
    }
    return;
  }
  const double xi = 1.0 / x;
  double sn, cs;
  sincos(x, &sn, &cs);
  if (x > (double)LL) {
    double j0 = sn * xi;
    double j1 = (j0 - cs) * xi;
    jl[0] = j0;
    jl[1] = j1;
#pragma unroll
    for (int l = 1; l < LL; ++l) {
      double j2 = (double)(2 * l + 1) * xi * j1 - j0;
      jl[l + 1] = j2;
      j0 = j1;
      j1 = j2;
    }
    return;
  }
  // downward: start well above L; values grow towards l = 0
  const int lstart = LL + 20 + (int)ceil(x);
  double jp1 = 0.0, jc = 1e-300;
  for (int l = lstart; l > LL; --l) {
    double jm1 = (double)(2 * l + 1) * xi * jc - jp1;
    jp1 = jc;
    jc = jm1;
    if (fabs(jc) > 1e250) {
      jc *= 1e-250;
      jp1 *= 1e-250;
    }
  }
#pragma unroll
  for (int l = LL; l >= 0; --l) {
    jl[l] = jc;
    double jm1 = (double)(2 * l + 1) * xi * jc - jp1;
    jp1 = jc;
    jc = jm1;
    if (fabs(jc) > 1e250) {
      jc *= 1e-250;
      jp1 *= 1e-250;
#pragma unroll
      for (int q = l; q <= LL; ++q) jl[q] *= 1e-250;
    }
  }
  const double j0t = sn * xi;
  const double j1t = sn * xi * xi - cs * xi;
  const double scale = (fabs(j0t) >= fabs(j1t)) ? j0t / jl[0] : j1t / jl[1];
#pragma unroll
  for (int l = 0; l <= LL; ++l) jl[l] *= scale;
}

template <int L>
__device__ __forceinline__ void match_body(int ig, int ias, int apwordmax, const int *__restrict__ spidx,
                                           const double *__restrict__ rmt, double fpso,
                                           const double *__restrict__ dinv, const double *__restrict__ gpc,
                                           const double *__restrict__ tpgpc, const cplx *__restrict__ sfacgp,
                                           int ldsf, cplx *__restrict__ apwalm, int ldg, double *__restrict__ rf = nullptr,
                                           int nl = 0, int ngp = 0, double *__restrict__ s2out = nullptr,
                                           const double *__restrict__ olpl = nullptr) {
#pragma clang fp contract(off)   // see sbessel_seq
  constexpr int lmmax = (L + 1) * (L + 1);
  const double g = gpc[ig];
  const double x = g * rmt[spidx[ias]];
  double jl[L + 2];
  sbessel_seq<L>(x, jl);
  double djl[L + 1];
#pragma unroll
  for (int l = 0; l <= L; ++l) djl[l] = (x > 0.0) ? g * ((double)l / x * jl[l] - jl[l + 1]) : 0.0;

  double st, ct;
  sincos(tpgpc[2 * ig], &st, &ct);
  const double phi = tpgpc[2 * ig + 1];
  const cplx sf = sfacgp[ig + (size_t)ldsf * ias];
  const double *di = dinv + (size_t)ias * (L + 1) * 4;
  const size_t colstride = (size_t)ldg;
  cplx *out = apwalm + ig + (size_t)ldg * apwordmax * lmmax * ias;

  if (s2out) {
    // sum over this atom's (lm, io) of |apwalm|^2 = (4 pi/sqrt(Omega))^2 |S|^2 |Y_lm|^2 r_io^2,
    // by sum_m |Y_lm|^2 = (2l+1)/(4 pi): O(lmax) instead of a sum over every value
    double t = 0.0;
#pragma unroll
    for (int l = 0; l <= L; ++l) {
      const double r0 = di[l * 4 + 0] * jl[l] + di[l * 4 + 2] * djl[l];
      const double r1 = di[l * 4 + 1] * jl[l] + di[l * 4 + 3] * djl[l];
      t = fma((double)(2 * l + 1), fma(r1, r1, r0 * r0), t);
    }
    *s2out = fpso * fpso * (sf.re * sf.re + sf.im * sf.im) * t * 0.079577471545947667884;   // 1/(4 pi)
  }
  double pmm = 0.28209479177387814347;  // 1/sqrt(4 pi)
#pragma unroll
  for (int m = 0; m <= L; ++m) {
    if (m > 0) pmm = -sqrt((double)(2 * m + 1) / (2.0 * m)) * st * pmm;
    double sm, cm;
    sincos((double)m * phi, &sm, &cm);
    double p2 = 0.0, p1 = pmm;
#pragma unroll
    for (int l = m; l <= L; ++l) {
      double p;
      if (l == m) {
        p = pmm;
      } else if (l == m + 1) {
        p = sqrt((double)(2 * m + 3)) * ct * pmm;
      } else {
        const double a = sqrt((4.0 * l * l - 1.0) / ((double)(l * l - m * m)));
        const double b = sqrt(((l - 1.0) * (l - 1.0) - (double)(m * m)) / (4.0 * (l - 1.0) * (l - 1.0) - 1.0));
        p = a * (ct * p1 - b * p2);
      }
      if (l > m) {
        p2 = p1;
        p1 = p;
      }
      // coefficient (4 pi / sqrt(Omega)) i^l S conj(Y_lm) for +m; -m uses Y_{l,-m} = (-1)^m conj(Y_lm)
      const cplx ylm = {p * cm, p * sm};
      cplx il;
      switch (l & 3) {
        case 0: il = {fpso, 0.0}; break;
        case 1: il = {0.0, fpso}; break;
        case 2: il = {-fpso, 0.0}; break;
        default: il = {0.0, -fpso}; break;
      }
      const cplx c0 = {il.re * sf.re - il.im * sf.im, il.re * sf.im + il.im * sf.re};
      const cplx cp = {c0.re * ylm.re + c0.im * ylm.im, c0.im * ylm.re - c0.re * ylm.im};   // c0 conj(Y), m >= 0
      const double sgn = (m & 1) ? -1.0 : 1.0;
      const cplx cn = {sgn * (c0.re * ylm.re - c0.im * ylm.im),   // conj(Y_{l,-m}) = (-1)^m Y_lm
                       sgn * (c0.re * ylm.im + c0.im * ylm.re)};
      const double d00 = di[l * 4 + 0], d01 = di[l * 4 + 2];    // Dinv(1,1), Dinv(1,2) column-major
      const double d10 = di[l * 4 + 1], d11 = di[l * 4 + 3];
      const double r0 = d00 * jl[l] + d01 * djl[l];
      const double r1 = d10 * jl[l] + d11 * djl[l];
      if (rf && m == 0 && l < nl) {
        // radial factors for the overlap by the addition theorem, weighted by
        // (4 pi / sqrt(Omega)) sqrt((2l+1) / 4 pi): rf[(ias*ngp + ig)*2nl + 2l + io]
        // With the 'iora*' overlap factor I + h1aa = L L^T (olpl, lower L per atom and l) the
        // pair (r0, r1) is replaced by (r0, r1) L, so that the sum over io of the two rows'
        // products is r M r'^T (overlapsetup.f90:81-94)
        const double wl = fpso * sqrt((double)(2 * l + 1) * 0.079577471545947667884);   // 1/(4 pi)
        double *q = rf + ((size_t)ias * ngp + ig) * 2 * nl + 2 * l;
        double s0 = r0, s1 = r1;
        if (olpl) {
          const double *lf = olpl + ((size_t)ias * nl + l) * 3;
          s0 = lf[0] * r0 + lf[1] * r1;
          s1 = lf[2] * r1;
        }
        q[0] = wl * s0;
        q[1] = wl * s1;
      }
      const int lmp = l * (l + 1) + m;
      out[colstride * (0 + 2 * lmp)] = {cp.re * r0, cp.im * r0};
      out[colstride * (1 + 2 * lmp)] = {cp.re * r1, cp.im * r1};
      if (m > 0) {
        const int lmn = l * (l + 1) - m;
        out[colstride * (0 + 2 * lmn)] = {cn.re * r0, cn.im * r0};
        out[colstride * (1 + 2 * lmn)] = {cn.re * r1, cn.im * r1};
      }
    }
  }
}


}  // namespace match_dev
