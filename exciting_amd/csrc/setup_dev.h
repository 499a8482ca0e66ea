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

// upper-triangular tile index t -> (ti, tj), ti <= tj, enumerated column by column
__device__ __forceinline__ void tri_decode(int t, int &ti, int &tj) {
  int j = (int)((sqrt(8.0 * t + 1.0) - 1.0) * 0.5);
  while ((j + 1) * (j + 2) / 2 <= t) ++j;
  while (j * (j + 1) / 2 > t) --j;
  tj = j;
  ti = t - j * (j + 1) / 2;
}

}  // namespace setup_dev
