// Adjoint (vector-Jacobian product) of the setup on gfx950: dL/dtau and dL/dV(G) of
// L = Re sum_ij (D_ji H_ij + W_ji O_ij) for cotangents of rank nst,
//   D = C diag(wh) C^H,   W = C diag(wo) C^H,
// which with wh = w_k f and wo = -w_k f e is the k-point's contribution to dE/dtau, i.e.
// minus forcek's IBS force (forcek.f90:267-383, F = -dE/dtau).
//
// With q_i = G_i + k and U = C_apw^T A, V_l = C_apw^T diag(q_l) A (per atom a, columns of A_a):
//  MT APW-APW :  2 sum_n wh_n Im(V_l,n^H h_a U_n) + 2 sum_n wo_n Im(V_l,n^H U_n)
//  APW-lo     :  2 sum_n wh_n Im(sum_f L_nf conj(V_l,nf)),  L_nf = sum_{lo j of a} C_jn conj(loa_jf)
//                2 sum_n wo_n Im(sum_{lo j of a} C_jn sum_io oalo(io, ilo_j) conj(V_l,n(io, lm_j)))
//  Theta(G)   :  sum_ij (q_il - q_jl) Im(-ffac_a(G_i-G_j) conj(S_ia) S_ja (dp_ij D_ji + W_ji))
//                (the tau dependence of Theta(G) = delta_G0 - sum_a exp(-iG.r_a) ffac_a(G))
//  dL/dV(g)   =  sum_{ij: G_i - G_j = g} D_ji  (holomorphic partial over the full matrix)
// U and V_l come from library products the entry point runs before exc_launch_vjp (the
// q-scaled copies of C from q_scale_kernel times A); kernels here: atom (per (atom, state)
// traces), dw_theta (FP64 MFMA tiles of D^T, W^T over the APW block, reduced in the epilogue).
#include "setup_dev.h"

namespace {

using namespace zf64;
using setup_dev::GTab;
using setup_dev::lookup_g;

// ---------------------------------------------------------------------------------------
// atom: one workgroup per (atom, state n); MT and APW-lo terms
// ---------------------------------------------------------------------------------------
constexpr int ATH = 256;

__device__ __forceinline__ double block_sum(double v, double *red) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
  const int w = threadIdx.x >> 6;
  __syncthreads();
  if ((threadIdx.x & 63) == 0) red[w] = v;
  __syncthreads();
  double s = 0.0;
  for (int i = 0; i < ATH / 64; ++i) s += red[i];
  return s;
}

__global__ __launch_bounds__(ATH) void atom_kernel(int ngp, int nst, int natm, int naa, int maxnlo, int nlomax,
                                                   int apwordmax, int nlotot, const lo_desc *__restrict__ lod,
                                                   const cplx *__restrict__ haa, const cplx *__restrict__ loa,
                                                   const double *__restrict__ oalo, const cplx *__restrict__ evec,
                                                   int ldc, const double *__restrict__ wh, const double *__restrict__ wo,
                                                   const cplx *__restrict__ uv, double *__restrict__ dedtau) {
  __shared__ double red[ATH / 64];
  const int ias = blockIdx.x, n = blockIdx.y;
  const int ktot = natm * naa;
  const int f = threadIdx.x;
  const cplx *U = uv + n;                                   // U(n, c) = uv[n + c*nst]
  const size_t cs = (size_t)nst;
  const size_t col0 = (size_t)ias * naa;
  double tH[3] = {0.0, 0.0, 0.0}, tO[3] = {0.0, 0.0, 0.0};
  if (f < naa) {
    // (h_a U_n)_f
    const cplx *h = haa + (size_t)naa * naa * ias;
    cplx hu = {0.0, 0.0};
    for (int fp = 0; fp < naa; ++fp) {
      const cplx hv = h[f + (size_t)fp * naa];
      const cplx u = U[(col0 + fp) * cs];
      hu.re += hv.re * u.re - hv.im * u.im;
      hu.im += hv.re * u.im + hv.im * u.re;
    }
    const cplx uf = U[(col0 + f) * cs];
    // L_nf = sum over this atom's lo functions j of C(N+j, n) conj(loa(j, f))
    cplx L = {0.0, 0.0};
    for (int q = 0; q < nlotot; ++q) {
      const lo_desc d = lod[q];
      if (d.ias != ias) continue;
      const cplx c = evec[ngp + q + (size_t)n * ldc];
      const cplx w = loa[d.row + (size_t)maxnlo * (f + (size_t)naa * ias)];
      L.re += c.re * w.re + c.im * w.im;
      L.im += c.im * w.re - c.re * w.im;
    }
#pragma unroll
    for (int l = 0; l < 3; ++l) {
      const cplx v = U[((size_t)(l + 1) * ktot + col0 + f) * cs];   // V_l(n, f)
      // Im(conj(v) * x) = v.re x.im - v.im x.re
      tH[l] = v.re * (hu.im + L.im) - v.im * (hu.re + L.re);
      tO[l] = v.re * uf.im - v.im * uf.re;
    }
  }
  // APW-lo overlap term: threads over this atom's lo functions
  if (f < nlotot && lod[f].ias == ias) {
    const lo_desc d = lod[f];
    const cplx c = evec[ngp + f + (size_t)n * ldc];
#pragma unroll
    for (int l = 0; l < 3; ++l) {
      cplx acc = {0.0, 0.0};
      for (int io = 0; io < apwordmax; ++io) {
        const cplx v = U[((size_t)(l + 1) * ktot + col0 + io + apwordmax * d.lm) * cs];
        const double w = oalo[io + apwordmax * (d.ilo + nlomax * ias)];
        acc.re += w * v.re;
        acc.im -= w * v.im;                                   // conj(V)
      }
      tO[l] += c.re * acc.im + c.im * acc.re;                 // Im(c * acc)
    }
  }
  for (int l = 0; l < 3; ++l) {
    const double sH = block_sum(tH[l], red);
    const double sO = block_sum(tO[l], red);
    if (threadIdx.x == 0) atomicAdd(&dedtau[l + 3 * ias], 2.0 * (wh[n] * sH + wo[n] * sO));
  }
}

// ---------------------------------------------------------------------------------------
// dw_theta: tiles of D^T (i,j) = D_ji and W^T over the APW block, Theta(G) and V(G) terms
// ---------------------------------------------------------------------------------------
constexpr int DBM = 64, DBK = 16;
using DCore = Core<DBM, DBM, DBK, 2, 2, true>;

__global__ __launch_bounds__(256) void scale_cols_kernel(int ngp, int nst, const cplx *__restrict__ evec, int ldc,
                                                         const double *__restrict__ wh, const double *__restrict__ wo,
                                                         cplx *__restrict__ cf, cplx *__restrict__ cg) {
  const int i = blockIdx.x * 256 + threadIdx.x;
  const int n = blockIdx.y;
  if (i >= ngp) return;
  const cplx c = evec[i + (size_t)n * ldc];
  cf[i + (size_t)n * ngp] = cplx{c.re * wh[n], c.im * wh[n]};
  cg[i + (size_t)n * ngp] = cplx{c.re * wo[n], c.im * wo[n]};
}

__global__ __launch_bounds__(256) void dw_theta_kernel(int ngp, int nst, int natm, const cplx *__restrict__ cf,
                                                       const cplx *__restrict__ cg, const cplx *__restrict__ evec,
                                                       int ldc, const int *__restrict__ igpig,
                                                       const double *__restrict__ vgpc, const cplx *__restrict__ sfacgp,
                                                       int ldg, const int *__restrict__ spidx,
                                                       const double *__restrict__ ffacg, GTab gt,
                                                       double *__restrict__ dedtau, double *__restrict__ dedv) {
  __shared__ Smem<DBM, DBM, DBK> sm;
  __shared__ int s_gr[DBM][3], s_gc[DBM][3];
  __shared__ double s_qr[DBM][3], s_qc[DBM][3];
  __shared__ double s_red[4][3];
  const int row0 = blockIdx.x * DBM, col0 = blockIdx.y * DBM;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6, wr = wave >> 1, wc = wave & 1;
  for (int e = tid; e < 2 * DBM; e += 256) {
    const bool isr = e < DBM;
    const int idx = isr ? row0 + e : col0 + e - DBM;
    int g0 = 0, g1 = 0, g2 = 0;
    double v0 = 0.0, v1 = 0.0, v2 = 0.0;
    if (idx < ngp) {
      const int ig = igpig[idx] - 1;
      g0 = gt.ivg[3 * ig]; g1 = gt.ivg[3 * ig + 1]; g2 = gt.ivg[3 * ig + 2];
      v0 = vgpc[3 * idx]; v1 = vgpc[3 * idx + 1]; v2 = vgpc[3 * idx + 2];
    }
    int(&g)[DBM][3] = isr ? s_gr : s_gc;
    double(&q)[DBM][3] = isr ? s_qr : s_qc;
    const int r = isr ? e : e - DBM;
    g[r][0] = g0; g[r][1] = g1; g[r][2] = g2;
    q[r][0] = v0; q[r][1] = v1; q[r][2] = v2;
  }
  // D^T tile: out(i, j) = sum_n conj(C_in wh_n) C_jn;  W^T with wo
  DCore cd, cw;
  cd.run(sm, cf, ngp, row0, ngp, evec, ldc, col0, ngp, nst, nullptr);
  cw.run(sm, cg, ngp, row0, ngp, evec, ldc, col0, ngp, nst, nullptr);

  // per element: lookup, dL/dV scatter, and the Theta terms of all atoms
  constexpr int EPT = DCore::MB * DCore::NB * 4;
  int eig[EPT];
  double xr[EPT], xi[EPT];
  int er[EPT], ec[EPT];
  int p = 0;
#pragma unroll
  for (int i = 0; i < DCore::MB; ++i)
#pragma unroll
    for (int j = 0; j < DCore::NB; ++j)
#pragma unroll
      for (int v = 0; v < 4; ++v, ++p) {
        const int rl = DCore::row_of(wr, i, v), cl = DCore::col_of(wc, j);
        const int r = row0 + rl, c = col0 + cl;
        er[p] = rl;
        ec[p] = cl;
        eig[p] = 0;
        xr[p] = xi[p] = 0.0;
        if (r < ngp && c < ngp) {
          const int ig = lookup_g(gt, s_gr[rl][0] - s_gc[cl][0], s_gr[rl][1] - s_gc[cl][1], s_gr[rl][2] - s_gc[cl][2]);
          eig[p] = ig;
          if (ig) {
            atomicAdd(&dedv[2 * (ig - 1)], cd.cr[i][j][v]);
            atomicAdd(&dedv[2 * (ig - 1) + 1], cd.ci[i][j][v]);
            const double dp = 0.5 * (s_qr[rl][0] * s_qc[cl][0] + s_qr[rl][1] * s_qc[cl][1] + s_qr[rl][2] * s_qc[cl][2]);
            xr[p] = dp * cd.cr[i][j][v] + cw.cr[i][j][v];
            xi[p] = dp * cd.ci[i][j][v] + cw.ci[i][j][v];
          }
        }
      }
  for (int a = 0; a < natm; ++a) {
    const int sp = spidx[a];
    const cplx *S = sfacgp + (size_t)ldg * a;
    double acc[3] = {0.0, 0.0, 0.0};
#pragma unroll
    for (int q = 0; q < EPT; ++q) {
      if (!eig[q]) continue;
      const int r = row0 + er[q], c = col0 + ec[q];
      const cplx si = S[r], sj = S[c];
      // p = conj(S_i) S_j = exp(-i (G_i - G_j).r_a);  val = -ffac * p * X
      const double pr = si.re * sj.re + si.im * sj.im;
      const double pi = si.re * sj.im - si.im * sj.re;
      const double th = -ffacg[(eig[q] - 1) + (size_t)gt.ngvec * sp];
      const double vi = th * (pr * xi[q] + pi * xr[q]);        // Im(th * p * X)
#pragma unroll
      for (int l = 0; l < 3; ++l) acc[l] += (s_qr[er[q]][l] - s_qc[ec[q]][l]) * vi;
    }
#pragma unroll
    for (int l = 0; l < 3; ++l) {
      double v = acc[l];
#pragma unroll
      for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
      if (lane == 0) s_red[wave][l] = v;
    }
    __syncthreads();
    if (tid < 3) atomicAdd(&dedtau[tid + 3 * a], s_red[0][tid] + s_red[1][tid] + s_red[2][tid] + s_red[3][tid]);
    __syncthreads();
  }
}

// cs_l(G, m) = q_l(G) C(G, m) for l = 1..3 and cs_0 = C: the left operands of
// uv(:, l K + c) = cs_l^T A(:, c), one strided batch of library products
__global__ __launch_bounds__(256) void q_scale_kernel(int ngp, int nst, const cplx *__restrict__ evec, int ldc,
                                                      const double *__restrict__ vgpc, cplx *__restrict__ cs) {
  const int g = blockIdx.x * 256 + threadIdx.x, m = blockIdx.y;
  if (g >= ngp) return;
  const cplx v = evec[g + (size_t)ldc * m];
  const size_t o = g + (size_t)ngp * m, st = (size_t)ngp * nst;
  cs[o] = v;
  for (int l = 1; l <= 3; ++l) {
    const double q = vgpc[3 * g + l - 1];
    cs[o + l * st] = cplx{v.re * q, v.im * q};
  }
}

}  // namespace

// the four q-scaled copies of the APW rows of C (4 ngp nst elements at cs)
int exc_launch_vjp_qscale(int ngp, int nst, const cplx *evec, int ldc, const double *vgpc, cplx *cs, hipStream_t s) {
  if (ngp <= 0 || nst <= 0) return EXC_OK;
  if (nst > 65535) return EXC_ERR_ARG;
  hipLaunchKernelGGL(q_scale_kernel, dim3((ngp + 255) / 256, nst), dim3(256), 0, s, ngp, nst, evec, ldc, vgpc, cs);
  EXC_CHECK(hipGetLastError());
  return EXC_OK;
}

int exc_launch_vjp(const exc_dev &d, int ngp, int nst, const cplx *apwalm, int ldg, const int *igpig,
                   const double *vgpc, const cplx *sfacgp, const cplx *evec, int ldc, const double *wh,
                   const double *wo, cplx *work, double *dedtau, double *dedv, hipStream_t s) {
  if (ngp <= 0 || nst <= 0) return EXC_OK;
  const int ktot = d.natm * d.naa;
  cplx *uv = work;                                   // nst x 4K, filled by the caller
  cplx *cf = uv + (size_t)nst * 4 * ktot;            // ngp x nst
  cplx *cg = cf + (size_t)ngp * nst;
  if (d.naa > ATH || d.nlotot > ATH) return EXC_ERR_UNSUPPORTED;
  hipLaunchKernelGGL(atom_kernel, dim3(d.natm, nst), dim3(ATH), 0, s, ngp, nst, d.natm, d.naa, d.maxnlo, d.nlomax,
                     d.apwordmax, d.nlotot, d.lod, d.haa, d.loa, d.oalo, evec, ldc, wh, wo, uv, dedtau);
  EXC_CHECK(hipGetLastError());
  hipLaunchKernelGGL(scale_cols_kernel, dim3((ngp + 255) / 256, nst), dim3(256), 0, s, ngp, nst, evec, ldc, wh, wo,
                     cf, cg);
  EXC_CHECK(hipGetLastError());
  const int nt = (ngp + DBM - 1) / DBM;
  hipLaunchKernelGGL(dw_theta_kernel, dim3(nt, nt), dim3(256), 0, s, ngp, nst, d.natm, cf, cg, evec, ldc, igpig, vgpc,
                     sfacgp, ldg, d.spidx, d.ffacg, setup_dev::make_gtab(d), dedtau, dedv);
  EXC_CHECK(hipGetLastError());
  return EXC_OK;
}

// uv (nst x 4K), cf and cg (ngp x nst each), then the q-scaled copies of C (4 ngp nst)
size_t exc_vjp_work_elems(const exc_dev &d, int ngp, int nst) {
  return (size_t)nst * 4 * d.natm * d.naa + 6 * (size_t)ngp * nst;
}
