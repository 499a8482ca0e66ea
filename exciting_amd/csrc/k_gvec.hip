// G+k vectors of a set of k-points and their structure factors (gengpvec.f90:13-94,
// gensfacgp.f90:13-58), one workgroup per k-point.
//
// The G vectors are the context's ivg, in their order (sorted by length, gengvec.f90:74); the
// k-point's set is the G with |G + k| < gkmax (compared squared, as the reference), kept in
// that order: an ordered compaction, 1024 G at a time — per wave a ballot ranks the selected
// lanes, the waves' counts are added in wave order — so position p of the set is the p-th
// selected G, whatever the scheduling.  For each one: igpig = ig (1-based), vgpl = ivg + vpl,
// vgpc = G + k (G = i1 b1 + i2 b2 + i3 b3 as gengvec.f90:106-107, no contraction), gpc and
// tpgpc as sphcrd.f90:27-49, and S_a(G + k) = exp(i (G + k).r_a) for every atom
// (gensfacgp.f90:38-49).  Rows ngp .. ngkmax-1 are zeroed; a set larger than ngkmax keeps its
// first ngkmax and reports its full size (the reference stops there, gengpvec.f90:53-58).
#include "exc_common.h"

namespace {

constexpr int kGT = 1024;

struct Bvec {
  double b[9];   // bvec(3,3), column-major: b[3 j + i] = bvec(i+1, j+1)
};

__device__ __forceinline__ void sphcrd_dev(const double (&v)[3], double &r, double &th, double &ph) {
#pragma clang fp contract(off)
  constexpr double twopi = 6.2831853071795864769, eps = 1e-14;
  r = sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
  th = 0.0;
  ph = 0.0;
  if (r > eps) {
    double t1 = v[2] / r;
    t1 = t1 > 0.0 ? fmin(t1, 1.0) : fmax(t1, -1.0);
    th = acos(t1);
    if (fabs(v[0]) > eps || fabs(v[1]) > eps) {
      ph = atan2(v[1], v[0]);
      if (ph < 0.0) ph += twopi;
    }
  }
}

__global__ __launch_bounds__(kGT) void gengpvec_kernel(const int *__restrict__ ivg, int ngvec, Bvec bv,
                                                       const double *__restrict__ kv, double gk2, int ngkmax,
                                                       int *__restrict__ ngp_out, int *__restrict__ igpig,
                                                       double *__restrict__ vgpl, double *__restrict__ vgpc,
                                                       double *__restrict__ gpc, double *__restrict__ tpgpc,
                                                       const double *__restrict__ atposc, int natm,
                                                       cplx *__restrict__ sfacgp) {
#pragma clang fp contract(off)
  const int k = blockIdx.x;
  const double *vpl = kv + 6 * k, *vpc = kv + 6 * k + 3;
  const size_t o1 = (size_t)k * ngkmax;
  __shared__ int wsum[kGT / 64];
  __shared__ int base;
  if (threadIdx.x == 0) base = 0;
  __syncthreads();
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  for (int c0 = 0; c0 < ngvec; c0 += kGT) {
    const int ig = c0 + threadIdx.x;
    bool sel = false;
    double v[3] = {0.0, 0.0, 0.0};
    int iv[3] = {0, 0, 0};
    if (ig < ngvec) {
      iv[0] = ivg[3 * ig];
      iv[1] = ivg[3 * ig + 1];
      iv[2] = ivg[3 * ig + 2];
#pragma unroll
      for (int d = 0; d < 3; ++d)
        v[d] = ((double)iv[0] * bv.b[d] + (double)iv[1] * bv.b[3 + d] + (double)iv[2] * bv.b[6 + d]) + vpc[d];
      sel = v[0] * v[0] + v[1] * v[1] + v[2] * v[2] < gk2;
    }
    const unsigned long long m = __ballot(sel);
    const int rank = __popcll(m & ((1ull << lane) - 1ull));
    if (lane == 0) wsum[wave] = __popcll(m);
    __syncthreads();
    int off = base;
    for (int w = 0; w < wave; ++w) off += wsum[w];
    const int p = off + rank;
    if (sel && p < ngkmax) {
      igpig[o1 + p] = ig + 1;
#pragma unroll
      for (int d = 0; d < 3; ++d) {
        vgpl[3 * (o1 + p) + d] = (double)iv[d] + vpl[d];
        vgpc[3 * (o1 + p) + d] = v[d];
      }
      double r, th, ph;
      sphcrd_dev(v, r, th, ph);
      gpc[o1 + p] = r;
      tpgpc[2 * (o1 + p)] = th;
      tpgpc[2 * (o1 + p) + 1] = ph;
      if (sfacgp) {
        for (int a = 0; a < natm; ++a) {
          const double t1 = v[0] * atposc[3 * a] + v[1] * atposc[3 * a + 1] + v[2] * atposc[3 * a + 2];
          double s, c;
          sincos(t1, &s, &c);
          sfacgp[((size_t)k * natm + a) * ngkmax + p] = cplx{c, s};
        }
      }
    }
    __syncthreads();
    if (threadIdx.x == 0) {
      int t = 0;
      for (int w = 0; w < kGT / 64; ++w) t += wsum[w];
      base += t;
    }
    __syncthreads();
  }
  const int ngp = base;
  if (threadIdx.x == 0) ngp_out[k] = ngp;
  // zero padding rows
  for (int p = ngp + threadIdx.x; p < ngkmax; p += kGT) {
    igpig[o1 + p] = 0;
#pragma unroll
    for (int d = 0; d < 3; ++d) {
      vgpl[3 * (o1 + p) + d] = 0.0;
      vgpc[3 * (o1 + p) + d] = 0.0;
    }
    gpc[o1 + p] = 0.0;
    tpgpc[2 * (o1 + p)] = 0.0;
    tpgpc[2 * (o1 + p) + 1] = 0.0;
    if (sfacgp)
      for (int a = 0; a < natm; ++a) sfacgp[((size_t)k * natm + a) * ngkmax + p] = cplx{0.0, 0.0};
  }
}

// S_a(G + k) of given G+k vectors (gensfacgp.f90:38-49): vgpc(3, ngp) -> sfacgp(ld, natm)
__global__ __launch_bounds__(256) void gensfacgp_kernel(int ngp, const double *__restrict__ vgpc, int ld,
                                                        const double *__restrict__ atposc, int natm,
                                                        cplx *__restrict__ sfacgp) {
#pragma clang fp contract(off)
  const int p = blockIdx.x * 256 + threadIdx.x, a = blockIdx.y;
  if (p >= ngp || a >= natm) return;
  const double t1 = vgpc[3 * p] * atposc[3 * a] + vgpc[3 * p + 1] * atposc[3 * a + 1] + vgpc[3 * p + 2] * atposc[3 * a + 2];
  double s, c;
  sincos(t1, &s, &c);
  sfacgp[(size_t)a * ld + p] = cplx{c, s};
}

}  // namespace

int exc_launch_gengpvec(const exc_dev &d, int nk, const double *bvec, const double *kv, double gkmax, int ngkmax,
                        int *ngp, int *igpig, double *vgpl, double *vgpc, double *gpc, double *tpgpc, cplx *sfacgp,
                        hipStream_t s) {
  if (nk <= 0) return EXC_OK;
  Bvec bv;
  for (int i = 0; i < 9; ++i) bv.b[i] = bvec[i];
  hipLaunchKernelGGL(gengpvec_kernel, dim3(nk), dim3(kGT), 0, s, d.ivg, d.ngvec, bv, kv, gkmax * gkmax, ngkmax, ngp,
                     igpig, vgpl, vgpc, gpc, tpgpc, d.atposc, d.natm, sfacgp);
  EXC_CHECK(hipGetLastError());
  return EXC_OK;
}

int exc_launch_gensfacgp(const exc_dev &d, int ngp, const double *vgpc, int ld, cplx *sfacgp, hipStream_t s) {
  if (ngp <= 0) return EXC_OK;
  hipLaunchKernelGGL(gensfacgp_kernel, dim3((ngp + 255) / 256, d.natm), dim3(256), 0, s, ngp, vgpc, ld, d.atposc,
                     d.natm, sfacgp);
  EXC_CHECK(hipGetLastError());
  return EXC_OK;
}
