// Device code of the FP64 per-k H+O setup, shared by the per-k and the grouped kernels.
//
//  ystage : Yt_a = A_a . h_a^T for every atom a (hamiltonsetup.f90:75-88 computes its
//           transpose zm = h_a . apwi).
//  mt     : H(i,j) = sum_k conj(A(i,k)) Yt(j,k) + V(G_i-G_j) + 1/2 (G_i+k).(G_j+k) Theta(G_i-G_j)
//           O(i,j) = sum_k conj(A(i,k)) A(j,k)  + Theta(G_i-G_j)
//           over the upper triangle of tiles only, k running over all atoms' APW functions
//           (hamiltonsetup.f90:90-103 and overlapsetup.f90:95-108 summed over atoms, plus
//           hmlistln.f90:74-84 / olpistln.f90:69-82).  Each element with i <= j is written at
//           (i,j) and its conjugate at (j,i) (hamiltonoverlapocopy_UL.f90:15-27), so H and O
//           are written exactly once.
#pragma once
#include "zgemm_f64.h"

namespace setup_dev {

using namespace zf64;

struct GTab {
  const int *ivg;      // [3*ngrtot]
  const int *ivgig;    // box
  const cplx *cfunig;  // [ngrtot]
  const cplx *veffig;  // [ngvec]
  const cplx *kinig;   // the factor of H's interstitial kinetic term: cfunig, or meffig [ngvec]
                       // with ValenceRelativity /= 'none' (hmlistln.f90:61-72 against :74-84)
  const cplx *m2eff;   // 'iora*': O's interstitial term a2h (G_i+k).(G_j+k) m2effig [ngvec], or null
  double a2h;          // 1/2 a2 = 1/4 alpha^2 (olpistln.f90:46-47, 60)
  unsigned *miss;      // count of basis differences found outside the G set (exc_ctx_status)
  int ngvec;
  int lo0, lo1, lo2, n0, n1, n2;
};

inline GTab make_gtab(const exc_dev &d) {
  GTab t;
  t.ivg = d.ivg;
  t.ivgig = d.ivgig;
  t.cfunig = d.cfunig;
  t.veffig = d.veffig;
  t.kinig = d.zora ? d.meffig : d.cfunig;
  t.m2eff = (d.iora & 4) ? d.m2effig : nullptr;
  const double alpha = 1.0 / 137.03599911, a2 = 0.5 * (alpha * alpha);
  t.a2h = 0.5 * a2;
  t.miss = d.miss;
  t.ngvec = d.ngvec;
  t.lo0 = d.box_lo[0]; t.lo1 = d.box_lo[1]; t.lo2 = d.box_lo[2];
  t.n0 = d.box_n[0]; t.n1 = d.box_n[1]; t.n2 = d.box_n[2];
  return t;
}

// 1-based index of G = (d0,d1,d2) = G_i - G_j of two basis vectors in the G set, 0 if outside
// the box or beyond ngvec (the "ig > 0 .and. ig <= ngvec" test of hmlistln.f90:77).  The
// reference never reaches that case: it stops before any setup when 2 gkmax > gmaxvr
// (init1.f90:340-347), which is what keeps every basis difference inside the first ngvec
// G-vectors.  A miss is therefore counted, and exc_ctx_status reports it as an error.
__device__ __forceinline__ int lookup_g(const GTab &t, int d0, int d1, int d2) {
  const int a = d0 - t.lo0, b = d1 - t.lo1, c = d2 - t.lo2;
  int ig = 0;
  if ((unsigned)a < (unsigned)t.n0 && (unsigned)b < (unsigned)t.n1 && (unsigned)c < (unsigned)t.n2)
    ig = t.ivgig[a + t.n0 * (b + t.n1 * c)];
  if (ig > 0 && ig <= t.ngvec) return ig;
  if (t.miss) atomicAdd(t.miss, 1u);
  return 0;
}

// O's interstitial value at G = G_i - G_j (index ig > 0): Theta(G) (olpistln.f90:69-82), and
// with 'iora*' 1/2 a2 (G_i+k).(G_j+k) m2effig(G) + Theta(G) (olpistln.f90:53-62); vi, vj are
// the two G+k vectors, read only in that case
__device__ __forceinline__ cplx olp_theta(const GTab &t, int ig, const double *vi, const double *vj) {
  const cplx th = t.cfunig[ig - 1];
  if (!t.m2eff) return th;
  const double t1 = t.a2h * (vi[0] * vj[0] + vi[1] * vj[1] + vi[2] * vj[2]);
  const cplx m = t.m2eff[ig - 1];
  return cplx{t1 * m.re + th.re, t1 * m.im + th.im};
}

// upper-triangular tile index t -> (ti, tj), ti <= tj, enumerated column by column
__device__ __forceinline__ void tri_decode(int t, int &ti, int &tj) {
  int j = (int)((sqrt(8.0 * t + 1.0) - 1.0) * 0.5);
  while ((j + 1) * (j + 2) / 2 <= t) ++j;
  while (j * (j + 1) / 2 > t) --j;
  tj = j;
  ti = t - j * (j + 1) / 2;
}

}  // namespace setup_dev
