// Modular split of the muffin-tin contraction on gfx950 (EXC_BATCH_SPLIT_CRT).
//
// H_MT = conj(A) Yt^T (and O_MT = conj(A) A^T) from exact integer products instead of FP64
// MFMA, the other terms of the setup staying FP64:
//
//  scale : per row i of an operand X (N x K complex, K = natm naa) the exponent e_i with
//          |X_i|_2 < 2^e_i and s_i = beta - e_i, beta the largest with 2^(2 beta + 2) <= P,
//          P = m_0 ... m_(n-1) the product of the pairwise coprime moduli (241, 233, 229, ...).
//  split : X'_ik = rint(Re X_ik 2^s_i) + i rint(Im X_ik 2^s_i), exact integers below 2^beta.
//          Every modulus m is odd with prime factors 1 mod 4, so it has a square root iota of
//          -1 and Z_m[i] splits into two copies of Z_m by the ring maps x + iy -> x +- iota y.
//          Per modulus the two planes of an operand are those maps, as symmetric int8
//          residues: Re X' - iota Im X' and Re X' + iota Im X' for the conjugated left operand,
//          Re X' + iota Im X' and Re X' - iota Im X' for the right one; layout
//          [m][plane][K/16][row][16 bytes]: one 16-byte k-chunk of 64 consecutive rows is one
//          1 KiB wave-wide DMA.
//  gemm  : per modulus, tile and map, U = u_a u_b^T or V = v_a v_b^T by
//          v_mfma_i32_32x32x32_i8 with exact int32 accumulation (|P| <= K 120^2 < 2^31),
//          reduced modulo m to int8 residues: two real products per modulus, where Gauss's
//          form needs three.
//  merge : per tile element Re = (U + V)/2, Im = (U - V)/(2 iota) modulo each m, then
//          Garner's mixed-radix digits of Re', Im' in [-(P-1)/2, (P-1)/2] (exact:
//          |Re'|, |Im'| <= |A'_i| |B'_j| < P/2), the FP64 value correctly rounded as
//          fma(x_hi, M_6, x_lo), times 2^-(s_i + t_j); then the interstitial term and the
//          Hermitian mirror exactly as the FP64 epilogue (hmlistln.f90:74-84,
//          olpistln.f90:69-82, hamiltonoverlapocopy_UL.f90:15-27).
// Integer sums do not depend on their order, so every variant of every kernel gives the
// same bits; the only rounding between the inputs and H_MT is that of the operands to
// beta-bit integers and the final conversion.
//
// Gram form (the default where it applies): with h_a = U_a Lambda_a U_a^H and a shift
// sigma >= -min lambda, E_a = A_a conj(U_a) (Lambda_a + sigma)^(1/2) (the Y stage on that
// table instead of h_a) gives conj(E) E^T = conj(A) (h + sigma) A^T = H_MT + sigma O_APW, so
// H_MT = conj(E) E^T - sigma O_APW with O_APW the FP64 addition-theorem overlap (O minus
// its interstitial term).  One operand is split instead of two, and the right operand's
// planes are the left's swapped, as for O = conj(A) A^T; merge items of this form carry
// which = 2.
#include "setup_dev.h"
#include "mt_dma.h"

namespace {

using setup_dev::GTab;
using setup_dev::lookup_g;
using setup_dev::stage_barrier;
using setup_dev::wait_vm_lgkm0;

typedef int v4i __attribute__((ext_vector_type(4)));
typedef int v16i __attribute__((ext_vector_type(16)));
typedef __attribute__((address_space(3))) void lds_void;

constexpr int kMaxTile = 256;    // largest tile edge of the gemm and merge kernels; rows padded to it

// symmetric residue of an exact integer held in a double, |x| < 2^52, modulo an odd m
// (|x/m| is never within 1/(2m) of a half-integer, far above the rounding of x/m)
__device__ __forceinline__ int dres(double x, int m, double minv) { return (int)fma(-(double)m, rint(x * minv), x); }

// ---------------------------------------------------------------------------------------
// scale: the per-row exponents s = beta - e, |X_i| < 2^e, of Yt, and of A when the caller
// supplies apwalm (with the grouped match, crt_expo_kernel sums A's from the norms the match
// wrote): one workgroup per (k, 64 rows); wave w sums |x|^2 over its quarter of K, the
// quarters added in wave order (a fixed order: the exponents do not depend on timing)
// ---------------------------------------------------------------------------------------
__global__ __launch_bounds__(256) void crt_scale_kernel(const kdesc *__restrict__ ks, int ktot, int beta, int op) {
  const kdesc k = ks[blockIdx.z];             // op 0: A (apwalm), 1: Yt, 2: E (in Yt's place) as the left operand
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int row = blockIdx.x * 64 + lane;
  __shared__ double part[4][64];
  if (blockIdx.x * 64 >= k.cnp) return;       // uniform over the workgroup
  const cplx *src = op ? k.yt : k.apwalm;
  const int ld = op ? k.ldy : k.ldg;
  const int kq = (ktot + 3) / 4, k0 = wave * kq, k1 = min(ktot, k0 + kq);   // op 2 writes xa
  double acc = 0.0;
  if (row < k.ngp) {
    for (int c = k0; c < k1; ++c) {
      const cplx v = src[row + (size_t)c * ld];
      acc = fma(v.re, v.re, acc);
      acc = fma(v.im, v.im, acc);
    }
  }
  part[wave][lane] = acc;
  __syncthreads();
  if (wave == 0) {
    const double s2 = ((part[0][lane] + part[1][lane]) + part[2][lane]) + part[3][lane];
    int e = 0;
    if (s2 > 0.0) frexp(sqrt(s2), &e);       // |X_i| < 2^e
    (op == 1 ? k.xy : k.xa)[row] = row < k.ngp ? beta - e : 0;
  }
}

// A's exponents from the squared norms per (atom, row) that the grouped match wrote (in
// closed form), summed in atom order: no pass over A.  One thread per row.
__global__ __launch_bounds__(256) void crt_expo_kernel(const kdesc *__restrict__ ks, int natm, int beta) {
  const kdesc k = ks[blockIdx.z];
  const int row = blockIdx.x * 256 + threadIdx.x;
  if (row >= k.cnp) return;
  double s2 = 0.0;
  if (row < k.ngp)
    for (int a = 0; a < natm; ++a) s2 += k.pa2[(size_t)a * k.cnp + row];
  int e = 0;
  if (s2 > 0.0) frexp(sqrt(s2), &e);
  k.xa[row] = row < k.ngp ? beta - e : 0;
}

// ---------------------------------------------------------------------------------------
// split: 4 waves; wave w of workgroup (bx, by) converts rows 64 bx .. +63 of k-chunk
// 4 by + w (16 values of k) into the residue planes of every modulus
// ---------------------------------------------------------------------------------------
// the low bytes of four ints as one word (byte 0 from a)
__device__ __forceinline__ unsigned pack4(int a, int b, int c, int d) {
  return __builtin_amdgcn_perm((unsigned)b, (unsigned)a, 0x0c0c0400u) |
         __builtin_amdgcn_perm((unsigned)d, (unsigned)c, 0x04000c0cu);
}

// The two residue planes of one modulus for 16 values of k of one row: the maps
// Re X' + sg iota Im X' and Re X' - sg iota Im X' (sg = -1 for the conjugated left operand A,
// +1 for the right operand Yt), formed exactly in FP64 from Re X' and the residue of Im X'
// (|Re X'| + iota (m-1)/2 < 2^45: the quotient by m is then far more accurate than the
// 1/(2m) margin of its rounding) and reduced to their symmetric residues.
template <int SG>
__device__ __forceinline__ void split_planes(const double (&xr)[16], const double (&xi)[16], const CrtTab &t,
                                             signed char *base, size_t pstride, size_t off) {
#pragma unroll 1
  for (int l = 0; l < t.n; ++l) {
    const int m = t.m[l];
    const double mi = t.dminv[l], io = (double)(SG * t.iota[l]);
    unsigned w[2][4];
#pragma unroll
    for (int g = 0; g < 4; ++g) {
      int ru[4], rv[4];
#pragma unroll
      for (int b = 0; b < 4; ++b) {
        const double a = xr[4 * g + b], c = xi[4 * g + b];
        const double ci = fma(-(double)m, rint(c * mi), c);     // Im X' mod m, exact
        ru[b] = dres(fma(io, ci, a), m, mi);
        rv[b] = dres(fma(-io, ci, a), m, mi);
      }
      w[0][g] = pack4(ru[0], ru[1], ru[2], ru[3]);
      w[1][g] = pack4(rv[0], rv[1], rv[2], rv[3]);
    }
#pragma unroll
    for (int p = 0; p < 2; ++p)
      *(uint4 *)(base + (size_t)(2 * l + p) * pstride + off) = make_uint4(w[p][0], w[p][1], w[p][2], w[p][3]);
  }
}

// gram: one operand per k-point (grid z = k), E from Yt's buffer into the left planes
__global__ __launch_bounds__(256) void crt_split_kernel(const kdesc *__restrict__ ks, int ktot, int kp, CrtTab t,
                                                        int gram) {
  const kdesc k = ks[gram ? blockIdx.z : blockIdx.z >> 1];
  const int op = gram ? 0 : blockIdx.z & 1;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int row = blockIdx.x * 64 + lane;
  const int kc = blockIdx.y * 4 + wave;
  if (blockIdx.x * 64 >= k.cnp || kc * 16 >= kp) return;
  const cplx *src = (op || gram) ? k.yt : k.apwalm;
  const int ld = (op || gram) ? k.ldy : k.ldg;
  const int s = row < k.ngp ? (op ? k.xy : k.xa)[row] : 0;
  double xr[16], xi[16];
#pragma unroll
  for (int q = 0; q < 16; ++q) {
    const int c = kc * 16 + q;
    cplx v = {0.0, 0.0};
    if (row < k.ngp && c < ktot) v = src[row + (size_t)c * ld];
    xr[q] = rint(ldexp(v.re, s));
    xi[q] = rint(ldexp(v.im, s));
  }
  signed char *base = (signed char *)(op ? k.qy : k.qa);
  const size_t pstride = (size_t)kp * k.cnp;            // bytes of one plane
  const size_t off = ((size_t)kc * k.cnp + row) * 16;
  if (op)
    split_planes<1>(xr, xi, t, base, pstride, off);
  else
    split_planes<-1>(xr, xi, t, base, pstride, off);
}

// ---------------------------------------------------------------------------------------
// gemm: one real residue product (tile, modulus, product p) per workgroup on an XCD-mapped
// grid (workgroup b runs on XCD b % 8 and takes the next work of that XCD's contiguous eighth
// of the list; the list keeps the tiles of one (k-point, modulus, product) in 4 x 4 blocks, so
// the workgroups an XCD runs together share their row and column panels in its L2).  BT x BT
// tiles, WR x WC waves with (BT / WR) x (BT / WC) wave tiles of 32 x 32 MFMA blocks, LDS
// stages of BK bytes of k of one plane of both panels, filled by DMA, NST stages in flight.
//   p = 0: U = u_a u_b^T, p = 1: V = v_a v_b^T (the two maps of Z_m[i]), each reduced modulo
//   m to int8 residues; the merge combines them.
// ---------------------------------------------------------------------------------------
template <int BT, int BK, int NST, int WR, int WC>
struct GemmCfg {
  static constexpr int kWaves = WR * WC, kThreads = 64 * kWaves;
  static constexpr int KC = BK / 16;                          // 16-byte k-chunks per stage
  static constexpr int PB = KC * BT * 16;                     // bytes of one panel per stage
  static constexpr int NL = 2 * KC * (BT / 64);               // wave-wide 1 KiB DMAs per stage
  static constexpr int L = NL / kWaves;
  static constexpr int MI = BT / WR / 32, NJ = BT / WC / 32;
  static_assert(NL % kWaves == 0 && BK % 32 == 0 && BT % 64 == 0, "DMA split / MFMA k-step");
  struct __attribute__((aligned(16))) Stage {
    unsigned char a[PB];
    unsigned char b[PB];
  };
};

template <int BT, int BK, int NST, int WR, int WC>
__device__ __forceinline__ void gemm_issue(typename GemmCfg<BT, BK, NST, WR, WC>::Stage *st, const unsigned char *pa,
                                           const unsigned char *pb, size_t chunk_stride, int row0, int col0,
                                           int kc0) {
  using C = GemmCfg<BT, BK, NST, WR, WC>;
  const int lane = threadIdx.x & 63;
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
#pragma unroll
  for (int q = 0; q < C::L; ++q) {
    const int j = wave + C::kWaves * q;               // 0 .. NL-1: (panel, chunk, 64-row quarter)
    const bool isb = j >= C::NL / 2;
    const int jj = isb ? j - C::NL / 2 : j;
    const int h = jj % (BT / 64), kc = jj / (BT / 64);
    const unsigned char *src = (isb ? pb : pa) + (size_t)(kc0 + kc) * chunk_stride +
                               (size_t)((isb ? col0 : row0) + 64 * h + lane) * 16;
    unsigned char *dst = (isb ? st->b : st->a) + kc * (BT * 16) + h * 1024;
    __builtin_amdgcn_global_load_lds((const void *)src, (lds_void *)dst, 16, 0, 0);
  }
}

template <int BT, int BK, int NST, int WR, int WC>
__device__ __forceinline__ void gemm_compute(const typename GemmCfg<BT, BK, NST, WR, WC>::Stage *st,
                                             v16i (&acc)[GemmCfg<BT, BK, NST, WR, WC>::MI][GemmCfg<BT, BK, NST, WR, WC>::NJ],
                                             int wr, int wc) {
  using C = GemmCfg<BT, BK, NST, WR, WC>;
  const int lane = threadIdx.x & 63;
  const int r = lane & 31, hk = lane >> 5;
#pragma unroll
  for (int s = 0; s < BK / 32; ++s) {
    v4i af[C::MI], bf[C::NJ];
#pragma unroll
    for (int i = 0; i < C::MI; ++i)
      af[i] = *(const v4i *)(st->a + ((2 * s + hk) * BT + wr * (32 * C::MI) + 32 * i + r) * 16);
#pragma unroll
    for (int j = 0; j < C::NJ; ++j)
      bf[j] = *(const v4i *)(st->b + ((2 * s + hk) * BT + wc * (32 * C::NJ) + 32 * j + r) * 16);
#pragma unroll
    for (int i = 0; i < C::MI; ++i)
#pragma unroll
      for (int j = 0; j < C::NJ; ++j) acc[i][j] = __builtin_amdgcn_mfma_i32_32x32x32_i8(af[i], bf[j], acc[i][j], 0, 0, 0);
  }
}

template <int BT, int BK, int NST, int WR, int WC, int MINW>
__global__ __launch_bounds__(64 * WR * WC, MINW) void crt_gemm_kernel(const kdesc *__restrict__ ks,
                                                                      const mt_item *__restrict__ items,
                                                                      const int2 *__restrict__ work, int nwork, int kp,
                                                                      CrtTab t, signed char *__restrict__ res) {
  using C = GemmCfg<BT, BK, NST, WR, WC>;
  using St = typename C::Stage;
  EXC_DMA_STAGES(St, NST)
  const int chunk = (nwork + 7) / 8;
  const int x = blockIdx.x % 8, qb = blockIdx.x / 8;
  const int wi = x * chunk + qb;
  if (qb >= chunk || wi >= nwork) return;
  const int2 wk = work[wi];                      // (item, 4 * modulus + product)
  const mt_item it = items[wk.x];
  const kdesc k = ks[it.k];
  const int l = wk.y >> 2, p = wk.y & 3;
  const size_t pstride = (size_t)kp * k.cnp, chunk_stride = (size_t)k.cnp * 16;
  // left operand conj(A): its planes u, v; right operand Yt (H): its planes u, v; or A (O),
  // whose maps as a right operand are the left planes swapped
  const unsigned char *pa = (const unsigned char *)k.qa + (size_t)(2 * l + p) * pstride;
  const unsigned char *pb = it.which ? (const unsigned char *)k.qa + (size_t)(2 * l + 1 - p) * pstride
                                     : (const unsigned char *)k.qy + (size_t)(2 * l + p) * pstride;
  const int row0 = it.ti * BT, col0 = it.tj * BT;
  const int wave = threadIdx.x >> 6, wr = wave / WC, wc = wave % WC;
  v16i acc[C::MI][C::NJ];
#pragma unroll
  for (int i = 0; i < C::MI; ++i)
#pragma unroll
    for (int j = 0; j < C::NJ; ++j) acc[i][j] = v16i{};
  St *st[4] = {&s0_, &s1_, (St *)&s2_, (St *)&s3_};
  const int nsteps = kp / BK;
#pragma unroll
  for (int u = 0; u < NST - 1; ++u)
    if (u < nsteps) gemm_issue<BT, BK, NST, WR, WC>(st[u], pa, pb, chunk_stride, row0, col0, u * C::KC);
  for (int base = 0; base < nsteps; base += NST) {
#pragma unroll
    for (int u = 0; u < NST; ++u) {
      const int step = base + u;
      if (step < nsteps) {
        // stage `step` has landed once at most the DMAs of the NST-2 later stages are out
        if (step + NST - 2 < nsteps)
          wait_vm_lgkm0<(NST - 2) * C::L>();
        else
          wait_vm_lgkm0<0>();
        stage_barrier();
        if (step + NST - 1 < nsteps)
          gemm_issue<BT, BK, NST, WR, WC>(st[(u + NST - 1) % NST], pa, pb, chunk_stride, row0, col0,
                                          (step + NST - 1) * C::KC);
        gemm_compute<BT, BK, NST, WR, WC>(st[u], acc, wr, wc);
      }
    }
  }
  // residues modulo m of one 32 x 32 block and lane are 16 contiguous bytes:
  // res[item][modulus][product][wave][block][lane][16]
  const int m = t.m[l];
  const double mi = t.dminv[l];
  const int lane = threadIdx.x & 63;
  signed char *out = res + (((size_t)wk.x * t.n + l) * 2 + p) * (size_t)(BT * BT);
#pragma unroll
  for (int i = 0; i < C::MI; ++i)
#pragma unroll
    for (int j = 0; j < C::NJ; ++j) {
      unsigned u4[4] = {0u, 0u, 0u, 0u};
#pragma unroll
      for (int v = 0; v < 16; ++v) u4[v >> 2] |= (unsigned)(dres((double)acc[i][j][v], m, mi) & 255) << (8 * (v & 3));
      const int blk = wave * (C::MI * C::NJ) + i * C::NJ + j;
      *(uint4 *)(out + ((size_t)blk * 64 + lane) * 16) = make_uint4(u4[0], u4[1], u4[2], u4[3]);
    }
}

// ---------------------------------------------------------------------------------------
// gemm, register-pipelined form for one wave per SIMD (4 waves, 2 x 2, 128 x 128 wave tiles:
// 8 LDS fragment reads per 16 MFMAs instead of 6 per 8 — the LDS, read by the MFMA operands
// and written by the DMA, bounds the 2 x 4 form).  With no second wave on the SIMD to cover
// the LDS latency, the fragments of MFMA k-step s + 1 are read while the MFMAs of step s run
// (two register sets), and the workgroup barrier that hands an LDS stage over to the DMA is
// taken once per stage, at its last k-step, after its last fragments are in registers:
//   wait for the next stage's DMA | barrier | read its first fragments | DMA into the stage
//   just finished | MFMAs of the last k-step
// so that the next stage's first fragments are in flight during those MFMAs.  Same
// products, same integer sums, same output layout as crt_gemm_kernel with a 2 x 2 wave grid.
// ---------------------------------------------------------------------------------------
template <int BT, int BK, int NST, int WR, int WC>
__device__ __forceinline__ void gemm_frags(const typename GemmCfg<BT, BK, NST, WR, WC>::Stage *st, int s,
                                           v4i (&af)[GemmCfg<BT, BK, NST, WR, WC>::MI],
                                           v4i (&bf)[GemmCfg<BT, BK, NST, WR, WC>::NJ], int wr, int wc) {
  using C = GemmCfg<BT, BK, NST, WR, WC>;
  const int lane = threadIdx.x & 63;
  const int r = lane & 31, hk = lane >> 5;
#pragma unroll
  for (int i = 0; i < C::MI; ++i)
    af[i] = *(const v4i *)(st->a + ((2 * s + hk) * BT + wr * (32 * C::MI) + 32 * i + r) * 16);
#pragma unroll
  for (int j = 0; j < C::NJ; ++j)
    bf[j] = *(const v4i *)(st->b + ((2 * s + hk) * BT + wc * (32 * C::NJ) + 32 * j + r) * 16);
}

template <int BT, int BK, int NST, int WR, int WC>
__device__ __forceinline__ void gemm_mfma(const v4i (&af)[GemmCfg<BT, BK, NST, WR, WC>::MI],
                                          const v4i (&bf)[GemmCfg<BT, BK, NST, WR, WC>::NJ],
                                          v16i (&acc)[GemmCfg<BT, BK, NST, WR, WC>::MI][GemmCfg<BT, BK, NST, WR, WC>::NJ]) {
  using C = GemmCfg<BT, BK, NST, WR, WC>;
#pragma unroll
  for (int i = 0; i < C::MI; ++i)
#pragma unroll
    for (int j = 0; j < C::NJ; ++j) acc[i][j] = __builtin_amdgcn_mfma_i32_32x32x32_i8(af[i], bf[j], acc[i][j], 0, 0, 0);
}

// the L wave-wide DMAs of one stage from per-slot sources fixed for the tile (slot q: panel,
// 16-byte chunk, 64-row quarter), `off` bytes further along k: no address arithmetic at the
// hand-over beyond one add per slot
template <int BT, int BK, int NST, int WR, int WC>
__device__ __forceinline__ void gemm_issue_pre(typename GemmCfg<BT, BK, NST, WR, WC>::Stage *st, const unsigned char *pak,
                                               const unsigned char *pbk,
                                               const unsigned (&voff)[GemmCfg<BT, BK, NST, WR, WC>::L],
                                               const int (&doff)[GemmCfg<BT, BK, NST, WR, WC>::L]) {
  using C = GemmCfg<BT, BK, NST, WR, WC>;
  static_assert((C::NL / 2) % C::kWaves == 0, "slot q of every wave in one panel");
#pragma unroll
  for (int q = 0; q < C::L; ++q)
    __builtin_amdgcn_global_load_lds((const void *)((q * C::kWaves >= C::NL / 2 ? pbk : pak) + voff[q]),
                                     (lds_void *)((unsigned char *)st + doff[q]), 16, 0, 0);
}

template <int BT, int BK, int NST, int WR, int WC>
__global__ __launch_bounds__(64 * WR * WC, 1) void crt_gemm_pipe_kernel(const kdesc *__restrict__ ks,
                                                                        const mt_item *__restrict__ items,
                                                                        const int2 *__restrict__ work, int nwork,
                                                                        int kp, CrtTab t, signed char *__restrict__ res) {
  using C = GemmCfg<BT, BK, NST, WR, WC>;
  using St = typename C::Stage;
  constexpr int S = BK / 32;                    // MFMA k-steps per stage
  static_assert(S % 2 == 0 && NST >= 2 && NST <= 4, "register sets alternate within a stage; LDS size");
  EXC_DMA_STAGES(St, NST)
  const int chunk = (nwork + 7) / 8;
  const int x = blockIdx.x % 8, qb = blockIdx.x / 8;
  const int wi = x * chunk + qb;
  if (qb >= chunk || wi >= nwork) return;
  const int2 wk = work[wi];
  const mt_item it = items[wk.x];
  const kdesc k = ks[it.k];
  const int l = wk.y >> 2, p = wk.y & 3;
  const size_t pstride = (size_t)kp * k.cnp, chunk_stride = (size_t)k.cnp * 16;
  const unsigned char *pa = (const unsigned char *)k.qa + (size_t)(2 * l + p) * pstride;
  const unsigned char *pb = it.which ? (const unsigned char *)k.qa + (size_t)(2 * l + 1 - p) * pstride
                                     : (const unsigned char *)k.qy + (size_t)(2 * l + p) * pstride;
  const int row0 = it.ti * BT, col0 = it.tj * BT;
  const int wave = threadIdx.x >> 6, wr = wave / WC, wc = wave % WC;
  v16i acc[C::MI][C::NJ];
#pragma unroll
  for (int i = 0; i < C::MI; ++i)
#pragma unroll
    for (int j = 0; j < C::NJ; ++j) acc[i][j] = v16i{};
  St *st[4] = {&s0_, &s1_, (St *)&s2_, (St *)&s3_};
  const int nsteps = kp / BK;
  // DMA slot q of this wave: panel (A for the first half of the slots), 16-byte k-chunk kc
  // and 64-row quarter h; its source is the panel's base for the stage plus voff[q] (below
  // 2^32: kc < KC, rows < cnp), its destination doff[q] bytes into the stage
  unsigned voff[C::L];
  int doff[C::L];
  {
    const int lane = threadIdx.x & 63;
    const int wv = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
#pragma unroll
    for (int q = 0; q < C::L; ++q) {
      const int j = wv + C::kWaves * q;
      const bool isb = j >= C::NL / 2;
      const int jj = isb ? j - C::NL / 2 : j;
      const int h = jj % (BT / 64), kc = jj / (BT / 64);
      voff[q] = (unsigned)((size_t)kc * chunk_stride + (size_t)((isb ? col0 : row0) + 64 * h + lane) * 16);
      doff[q] = (isb ? C::PB : 0) + kc * (BT * 16) + h * 1024;
    }
  }
  const size_t sstride = (size_t)C::KC * chunk_stride;       // bytes of k per stage
  // stage j's DMA is issued at the hand-over of stage j - NST (stages 0 .. NST-2 here, NST-1
  // at the first hand-over below); stages past the end load the last one again (never read),
  // so that exactly NST - 2 stages' DMA after the next one is in flight at every hand-over
  const size_t klast = (size_t)(nsteps - 1) * sstride;
#pragma unroll
  for (int u = 0; u < NST - 1; ++u) {
    const size_t o = std::min((size_t)u * sstride, klast);
    gemm_issue_pre<BT, BK, NST, WR, WC>(st[u], pa + o, pb + o, voff, doff);
  }
  v4i fa[2][C::MI], fb[2][C::NJ];
  // hand-over before stage 0
  wait_vm_lgkm0<(NST - 2) * C::L>();
  stage_barrier();
  gemm_frags<BT, BK, NST, WR, WC>(st[0], 0, fa[0], fb[0], wr, wc);
  {
    const size_t o = std::min((size_t)(NST - 1) * sstride, klast);
    gemm_issue_pre<BT, BK, NST, WR, WC>(st[NST - 1], pa + o, pb + o, voff, doff);
  }
  // one stage per iteration, its LDS buffer chosen at run time, and no branch inside: the
  // accumulators never flow through one (they would be copied between register files), and
  // the scheduler can put one fragment read and one DMA between two MFMAs.  The last stage
  // hands over too (its reads unused).
  size_t koff = (size_t)(NST - 1) * sstride;    // k offset of the DMA at the next hand-over, less one stage
#pragma unroll 1
  for (int step = 0; step < nsteps; ++step) {
    const int u = step % NST;
    const int un = u + 1 < NST ? u + 1 : 0;
    koff = koff + sstride < klast ? koff + sstride : klast;    // stage min(step + NST, nsteps - 1)
    // (selects, not an indexed array: a run-time index into an array of LDS addresses would
    // make it a static table)
    St *const cs = u == 0 ? &s0_ : (u == 1 ? &s1_ : (u == 2 ? (St *)&s2_ : (St *)&s3_));
    St *const ns = un == 0 ? &s0_ : (un == 1 ? &s1_ : (un == 2 ? (St *)&s2_ : (St *)&s3_));
#pragma unroll
    for (int s = 0; s < S; ++s) {
      __builtin_amdgcn_sched_barrier(0);
      if (s + 1 < S) {
        gemm_frags<BT, BK, NST, WR, WC>(cs, s + 1, fa[(s + 1) & 1], fb[(s + 1) & 1], wr, wc);
        gemm_mfma<BT, BK, NST, WR, WC>(fa[s & 1], fb[s & 1], acc);
#pragma unroll
        for (int q = 0; q < C::MI + C::NJ; ++q) {
          __builtin_amdgcn_sched_group_barrier(0x008, 1, 0);
          __builtin_amdgcn_sched_group_barrier(0x100, 1, 0);
        }
      } else {
        // hand-over: stage step + 1 has landed once at most the DMA of the NST - 2 stages
        // after it is out
        wait_vm_lgkm0<(NST - 2) * C::L>();
        stage_barrier();
        gemm_frags<BT, BK, NST, WR, WC>(ns, 0, fa[0], fb[0], wr, wc);
        gemm_issue_pre<BT, BK, NST, WR, WC>(cs, pa + koff, pb + koff, voff, doff);
        gemm_mfma<BT, BK, NST, WR, WC>(fa[s & 1], fb[s & 1], acc);
#pragma unroll
        for (int q = 0; q < C::MI + C::NJ; ++q) {
          __builtin_amdgcn_sched_group_barrier(0x008, 1, 0);
          __builtin_amdgcn_sched_group_barrier(0x100, 1, 0);
          if (q < C::L) __builtin_amdgcn_sched_group_barrier(0x010, 1, 0);
        }
      }
    }
  }
  wait_vm_lgkm0<0>();                            // no DMA into this workgroup's LDS outlives it
  const int m = t.m[l];
  const double mi = t.dminv[l];
  const int lane = threadIdx.x & 63;
  signed char *out = res + (((size_t)wk.x * t.n + l) * 2 + p) * (size_t)(BT * BT);
#pragma unroll
  for (int i = 0; i < C::MI; ++i)
#pragma unroll
    for (int j = 0; j < C::NJ; ++j) {
      unsigned u4[4] = {0u, 0u, 0u, 0u};
#pragma unroll
      for (int v = 0; v < 16; ++v) u4[v >> 2] |= (unsigned)(dres((double)acc[i][j][v], m, mi) & 255) << (8 * (v & 3));
      const int blk = wave * (C::MI * C::NJ) + i * C::NJ + j;
      *(uint4 *)(out + ((size_t)blk * 64 + lane) * 16) = make_uint4(u4[0], u4[1], u4[2], u4[3]);
    }
}

// ---------------------------------------------------------------------------------------
// merge: one tile per workgroup (4 waves).  Wave w takes the 32 x 32 blocks w, w + 4, ...
// of the gemm layout (wave-major as the gemm wrote them, decoded with that variant's wave
// grid), one lane per gemm lane with its 16 values.
// ---------------------------------------------------------------------------------------
// The moduli and Garner's tables are fixed at compile time for each modulus count (the
// moduli list is constant; only how many of them is chosen), so every reduction below
// multiplies by immediates.  Moduli: odd, pairwise coprime, every prime factor 1 mod 4
// (221 = 13 17, 205 = 5 41), the largest first.
constexpr int kModList[kMaxMod] = {241, 233, 229, 221, 205, 197, 193, 181, 173, 157, 149, 137, 113, 109};

constexpr int inv_mod(int a, int m) {   // a^-1 mod m, gcd(a, m) = 1 (extended Euclid)
  int r0 = m, r1 = ((a % m) + m) % m, t0 = 0, t1 = 1;
  while (r1) {
    const int q = r0 / r1, r2 = r0 - q * r1, t2 = t0 - q * t1;
    r0 = r1; r1 = r2; t0 = t1; t1 = t2;
  }
  return t0 < 0 ? t0 + m : t0;
}

constexpr int iota_of(int m) {          // the smallest square root of -1 modulo m
  for (int x = 1; x < m; ++x)
    if ((x * x + 1) % m == 0) return x;
  return 0;
}

// Garner's mixed-radix digits d_l = (r_l - sum_j d_j M_j) M_l^-1 mod m_l (M_j = m_0 ... m_(j-1))
// of Re' and Im', from U and V: r_l = (U + V) 2^-1 for Re, (U - V) (2 iota_l)^-1 for Im, so
// the digit is s_l c_l - sum_j d_j w[l][j] reduced once, with s_l = U +- V unreduced,
// c_l = (2^-1 or (2 iota_l)^-1) M_l^-1 mod m_l and w[l][j] = (M_j mod m_l) M_l^-1 mod m_l.
template <int NM>
struct GarnerTab {
  int w[NM][NM];
  int cre[NM], cim[NM];
  double mk;        // M_6 (or P for fewer moduli): x = x_hi M_6 + x_lo
  constexpr GarnerTab() : w(), cre(), cim(), mk(1.0) {
    for (int l = 0; l < NM; ++l) {
      const int m = kModList[l];
      long long ml = 1;
      int mj[NM] = {};
      for (int j = 0; j < l; ++j) {
        mj[j] = (int)(ml % m);
        ml = ml * kModList[j] % m;
      }
      const int v = inv_mod((int)ml, m);
      for (int j = 0; j < l; ++j) w[l][j] = mj[j] * v % m;
      cre[l] = inv_mod(2, m) * v % m;
      cim[l] = inv_mod(2 * iota_of(m) % m, m) * v % m;
    }
    for (int j = 0; j < (NM < kCrtLoDigits ? NM : kCrtLoDigits); ++j) mk *= kModList[j];
  }
};

// symmetric residue modulo m = kModList[L] of |x| < 2^22; the float quotient is within
// 2^-21 |x| / m of x / m, far inside the 1/(2m) margin of the rounding
template <int L>
__device__ __forceinline__ int cres(int x) {
  constexpr int m = kModList[L];
  constexpr float finv = 1.0f / (float)m;
  return x - __mul24((int)rintf((float)x * finv), m);
}

// |s_l c_l| <= 240 240 and |d_j w| <= 120 240: the sum stays below 2^22 for up to 12 moduli
template <int NM, int L, bool IM>
__device__ __forceinline__ void garner_digit(int (&d)[NM], const int (&s)[NM]) {
  constexpr GarnerTab<NM> G{};
  int acc = __mul24(s[L], IM ? G.cim[L] : G.cre[L]);
#pragma unroll
  for (int j = 0; j < L; ++j) acc -= __mul24(d[j], G.w[L][j]);
  d[L] = cres<L>(acc);
  if constexpr (L + 1 < NM) garner_digit<NM, L + 1, IM>(d, s);
}

// Re' (IM false) or Im' in [-(P-1)/2, (P-1)/2] from s_l = U + V (or U - V) modulo each m_l:
// Garner's digits, then the correctly rounded fma(x_hi, M_6, x_lo), the two parts exact in
// FP64 (|x_lo| < M_6 / 2 < 2^52)
template <int NM, bool IM>
__device__ __forceinline__ double garner(const int (&s)[NM]) {
  constexpr GarnerTab<NM> G{};
  int d[NM];
  garner_digit<NM, 0, IM>(d, s);
  constexpr int NLO = NM < kCrtLoDigits ? NM : kCrtLoDigits;
  double lo = (double)d[NLO - 1];
#pragma unroll
  for (int j = NLO - 2; j >= 0; --j) lo = fma(lo, (double)kModList[j], (double)d[j]);
  double hi = 0.0;
#pragma unroll
  for (int j = NM - 1; j >= NLO; --j) hi = fma(hi, (double)kModList[j], (double)d[j]);
  return fma(hi, G.mk, lo);
}

template <int NM>
__global__ __launch_bounds__(256) void crt_merge_kernel(const kdesc *__restrict__ ks, const mt_item *__restrict__ items,
                                                        const signed char *__restrict__ res, int bt, int wr_n, int wc_n,
                                                        GTab gt, const double *__restrict__ gsig) {
  const mt_item it = items[blockIdx.x];
  const kdesc k = ks[it.k];
  const int ngp = k.ngp;
  const int row0 = it.ti * bt, col0 = it.tj * bt;
  __shared__ int gr[kMaxTile][3], gc[kMaxTile][3];
  __shared__ double vr[kMaxTile][3], vc[kMaxTile][3];
  __shared__ int er[kMaxTile], ec[kMaxTile];
  for (int e = threadIdx.x; e < 2 * bt; e += 256) {
    const bool isr = e < bt;
    const int q = isr ? e : e - bt;
    const int idx = (isr ? row0 : col0) + q;
    int g0 = 0, g1 = 0, g2 = 0, ex = 0;
    double v0 = 0.0, v1 = 0.0, v2 = 0.0;
    if (idx < ngp) {
      const int ig = k.igpig[idx] - 1;
      g0 = gt.ivg[3 * ig]; g1 = gt.ivg[3 * ig + 1]; g2 = gt.ivg[3 * ig + 2];
      v0 = k.vgpc[3 * idx]; v1 = k.vgpc[3 * idx + 1]; v2 = k.vgpc[3 * idx + 2];
      ex = isr ? k.xa[idx] : (it.which ? k.xa[idx] : k.xy[idx]);
    }
    int(&g)[kMaxTile][3] = isr ? gr : gc;
    double(&vv)[kMaxTile][3] = isr ? vr : vc;
    g[q][0] = g0; g[q][1] = g1; g[q][2] = g2;
    vv[q][0] = v0; vv[q][1] = v1; vv[q][2] = v2;
    (isr ? er : ec)[q] = ex;
  }
  __syncthreads();
  const int mi_n = bt / wr_n / 32, nj_n = bt / wc_n / 32, per_wave = mi_n * nj_n;
  const int nblk = (bt / 32) * (bt / 32);
  const size_t ptile = (size_t)bt * bt;                     // residue bytes of one (modulus, product)
  const signed char *base = res + (size_t)blockIdx.x * NM * 2 * ptile;
  cplx *out = it.which == 1 ? k.o : k.h;
  const double sig = it.which == 2 ? *gsig : 0.0;           // Gram form: H_MT = conj(E) E^T - sigma O_APW
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  // Element (row, col) of a 32 x 32 block sits in lane (col mod 32) of the gemm layout, so the
  // conjugate store at (col, row) is coalesced along the lanes; the store at (row, col) goes
  // through this wave's LDS slice, 16 rows of the block at a time: stored [col][row] with one
  // pad element per column, read back with 16 lanes along a column (256 contiguous bytes)
  __shared__ cplx tb[4][32 * 17];
  cplx *tw = tb[wave];
  for (int b4 = 0; b4 < nblk; b4 += 4) {                    // the same trip count in every wave
    const int blk = b4 + wave;
    const int wv = blk / per_wave, bi = (blk % per_wave) / nj_n, bj = blk % nj_n;
    const int wr = wv / wc_n, wc = wv % wc_n;
    const int rb = wr * (32 * mi_n) + 32 * bi, cb = wc * (32 * nj_n) + 32 * bj;   // block origin in the tile
    const int cl = cb + (lane & 31);
    const int col = col0 + cl;
    uint4 q1[NM], q2[NM];
#pragma unroll
    for (int l = 0; l < NM; ++l) {
      const signed char *b = base + (size_t)l * 2 * ptile + ((size_t)blk * 64 + lane) * 16;
      q1[l] = *(const uint4 *)b;
      q2[l] = *(const uint4 *)(b + ptile);
    }
#pragma unroll
    for (int h = 0; h < 2; ++h) {
      // Gram form: this half's eight O elements, loaded together ahead of the reconstruction;
      // read at (col, row), O's conjugate mirror (bit for bit), where the lanes are contiguous
      cplx om[8];
      if (it.which == 2) {
#pragma unroll
        for (int vv = 0; vv < 8; ++vv) {
          const int v = 8 * h + vv;
          const int row = row0 + rb + 16 * h + 8 * (vv >> 2) + 4 * (lane >> 5) + (v & 3);
          om[vv] = (col < ngp && row <= col) ? k.o[col + (size_t)row * k.ldh] : cplx{0.0, 0.0};
        }
      }
#pragma unroll
      for (int vv = 0; vv < 8; ++vv) {
        const int v = 8 * h + vv;
        const int rh = 8 * (vv >> 2) + 4 * (lane >> 5) + (v & 3);   // row within the half
        const int rl = rb + 16 * h + rh;
        const int row = row0 + rl;
        cplx val = {0.0, 0.0};
        if (col < ngp && row <= col) {
          // conj(a) b: Re = (U + V)/2, Im = (U - V)/(2 iota) modulo each m
          const int g = v >> 2, sh = 8 * (v & 3);
          int sr[NM], si[NM];
#pragma unroll
          for (int l = 0; l < NM; ++l) {
            const unsigned w1 = g == 0 ? q1[l].x : g == 1 ? q1[l].y : g == 2 ? q1[l].z : q1[l].w;
            const unsigned w2 = g == 0 ? q2[l].x : g == 1 ? q2[l].y : g == 2 ? q2[l].z : q2[l].w;
            const int u = (signed char)(w1 >> sh), vv = (signed char)(w2 >> sh);
            sr[l] = u + vv;
            si[l] = u - vv;
          }
          const int sc = -(er[rl] + ec[cl]);
          double re = ldexp(garner<NM, false>(sr), sc);
          double im = ldexp(garner<NM, true>(si), sc);
          const int ig = lookup_g(gt, gr[rl][0] - gc[cl][0], gr[rl][1] - gc[cl][1], gr[rl][2] - gc[cl][2]);
          if (it.which == 2) {
            // O_APW = O - its interstitial term, O as the overlap kernel wrote it
            cplx oa = {om[vv].re, -om[vv].im};
            if (ig) {
              const cplx tho = setup_dev::olp_theta(gt, ig, vr[rl], vc[cl]);
              oa = cplx{oa.re - tho.re, oa.im - tho.im};
            }
            re = fma(-sig, oa.re, re);
            im = fma(-sig, oa.im, im);
          }
          if (ig) {
            const cplx th = gt.cfunig[ig - 1];
            if (it.which == 1) {
              re += th.re;
              im += th.im;
            } else {
              const cplx vv2 = gt.veffig[ig - 1];
              const double t1 = 0.5 * (vr[rl][0] * vc[cl][0] + vr[rl][1] * vc[cl][1] + vr[rl][2] * vc[cl][2]);
              re += vv2.re + t1 * gt.kinig[ig - 1].re;
              im += vv2.im + t1 * gt.kinig[ig - 1].im;
            }
          }
          val = cplx{re, im};
          if (row < col) out[col + (size_t)row * k.ldh] = cplx{re, -im};
        }
        tw[(lane & 31) * 17 + rh] = val;
      }
      __syncthreads();
      // 16 rows x 32 columns of this half: lane -> row (lane & 15), columns (lane >> 4) + 4 i
#pragma unroll
      for (int i = 0; i < 8; ++i) {
        const int cc = (lane >> 4) + 4 * i, r = lane & 15;
        const int grow = row0 + rb + 16 * h + r, gcol = col0 + cb + cc;
        if (gcol < ngp && grow <= gcol) out[grow + (size_t)gcol * k.ldh] = tw[cc * 17 + r];
      }
      __syncthreads();                                       // the slice is rewritten next
    }
  }
}

template <int NM>
int launch_merge_t(const exc_dev &d, const kdesc *ks, const mt_item *items, int nitems, const CrtTab &t,
                   const signed char *res, int bt, int wr, int wc, const double *gsig, hipStream_t s) {
  hipLaunchKernelGGL((crt_merge_kernel<NM>), dim3(nitems), dim3(256), 0, s, ks, items, res, bt, wr, wc,
                     setup_dev::make_gtab(d), gsig);
  EXC_CHECK(hipGetLastError());
  return EXC_OK;
}

template <int BT, int BK, int NST, int WR, int WC, int MINW>
int launch_gemm_t(const kdesc *ks, const mt_item *items, const int2 *work, int nwork, int kp, const CrtTab &t,
                  signed char *res, hipStream_t s) {
  if (kp % BK) return EXC_ERR_UNSUPPORTED;      // the k loop runs whole stages of BK bytes
  const int chunk = (nwork + 7) / 8;
  hipLaunchKernelGGL((crt_gemm_kernel<BT, BK, NST, WR, WC, MINW>), dim3(8 * chunk), dim3(64 * WR * WC), 0, s, ks,
                     items, work, nwork, kp, t, res);
  EXC_CHECK(hipGetLastError());
  return EXC_OK;
}

template <int BT, int BK, int NST>
int launch_gemm_pipe_t(const kdesc *ks, const mt_item *items, const int2 *work, int nwork, int kp, const CrtTab &t,
                       signed char *res, hipStream_t s) {
  if (kp % BK) return EXC_ERR_UNSUPPORTED;
  const int chunk = (nwork + 7) / 8;
  hipLaunchKernelGGL((crt_gemm_pipe_kernel<BT, BK, NST, 2, 2>), dim3(8 * chunk), dim3(256), 0, s, ks, items, work,
                     nwork, kp, t, res);
  EXC_CHECK(hipGetLastError());
  return EXC_OK;
}

}  // namespace

void exc_crt_tables(int n, CrtTab *t) {
  *t = CrtTab{};
  t->n = n;
  double lp = 0.0;
  for (int l = 0; l < n; ++l) {
    t->m[l] = kModList[l];
    t->iota[l] = iota_of(kModList[l]);
    t->dminv[l] = 1.0 / kModList[l];
    lp += log2((double)kModList[l]);
  }
  // beta: the largest b with 2^(2b + 2) <= P, from log2 P with a margin against its rounding
  t->beta = (int)floor((lp - 2.0 - 1e-9) / 2.0);
}

// a_parts: the grouped match wrote A's squared norms per atom (else A takes a pass of its
// own, as Yt always does); gram: E in Yt's buffer is the only operand
int exc_launch_crt_prep(const exc_dev &d, int nk, int maxnp, const kdesc *ks, const CrtTab &t, bool a_parts,
                        bool gram, hipStream_t s) {
  if (nk <= 0 || maxnp <= 0) return EXC_OK;
  const int ktot = d.natm * d.naa, kp = exc_crt_kpad(ktot);
  if (gram) {
    hipLaunchKernelGGL(crt_scale_kernel, dim3(maxnp / 64, 1, nk), dim3(256), 0, s, ks, ktot, t.beta, 2);
    EXC_CHECK(hipGetLastError());
    hipLaunchKernelGGL(crt_split_kernel, dim3(maxnp / 64, kp / 64, nk), dim3(256), 0, s, ks, ktot, kp, t, 1);
    EXC_CHECK(hipGetLastError());
    return EXC_OK;
  }
  if (a_parts) {
    hipLaunchKernelGGL(crt_expo_kernel, dim3((maxnp + 255) / 256, 1, nk), dim3(256), 0, s, ks, d.natm, t.beta);
    EXC_CHECK(hipGetLastError());
  } else {
    hipLaunchKernelGGL(crt_scale_kernel, dim3(maxnp / 64, 1, nk), dim3(256), 0, s, ks, ktot, t.beta, 0);
    EXC_CHECK(hipGetLastError());
  }
  hipLaunchKernelGGL(crt_scale_kernel, dim3(maxnp / 64, 1, nk), dim3(256), 0, s, ks, ktot, t.beta, 1);
  EXC_CHECK(hipGetLastError());
  hipLaunchKernelGGL(crt_split_kernel, dim3(maxnp / 64, kp / 64, 2 * nk), dim3(256), 0, s, ks, ktot, kp, t, 0);
  EXC_CHECK(hipGetLastError());
  return EXC_OK;
}

// variants: (tile edge, k bytes per LDS stage, stages, wave grid, minimum waves per SIMD)
//   0: 256, 64, 3, 2 x 4 (128 x 64 wave tiles), 1    1: 256, 32, 4, 2 x 4, 1
//   2: 128, 64, 3, 2 x 2 (64 x 64 wave tiles), 2     3: 256, 64, 2, 2 x 4, 1
//   4: 128, 32, 4, 2 x 2, 2
//   5: 256, 64, 3, 2 x 2 (128 x 128 wave tiles: two thirds of the LDS reads per MFMA), 1
//   6: 256, 32, 4, 2 x 2, 1                           7: 256, 64, 4, 2 x 2, 1
//   8: 256, 64, 3, 2 x 2, register-pipelined          9: 256, 64, 2, 2 x 2, register-pipelined
//  10: 256, 64, 4, 2 x 2, register-pipelined
int exc_crt_shape(int variant, int *bt, int *wr, int *wc) {
  switch (variant) {
    case 0: case 1: case 3: *bt = 256; *wr = 2; *wc = 4; return EXC_OK;
    case 5: case 6: case 7: case 8: case 9: case 10: *bt = 256; *wr = 2; *wc = 2; return EXC_OK;
    case 2: case 4: *bt = 128; *wr = 2; *wc = 2; return EXC_OK;
    default: return EXC_ERR_ARG;
  }
}

int exc_launch_crt_gemm(const exc_dev &d, const kdesc *ks, const mt_item *items, const int2 *work, int nwork,
                        const CrtTab &t, signed char *res, int variant, hipStream_t s) {
  if (nwork <= 0) return EXC_OK;
  const int kp = exc_crt_kpad(d.natm * d.naa);
  switch (variant) {
    case 0: return launch_gemm_t<256, 64, 3, 2, 4, 1>(ks, items, work, nwork, kp, t, res, s);
    case 1: return launch_gemm_t<256, 32, 4, 2, 4, 1>(ks, items, work, nwork, kp, t, res, s);
    case 2: return launch_gemm_t<128, 64, 3, 2, 2, 2>(ks, items, work, nwork, kp, t, res, s);
    case 3: return launch_gemm_t<256, 64, 2, 2, 4, 1>(ks, items, work, nwork, kp, t, res, s);
    case 4: return launch_gemm_t<128, 32, 4, 2, 2, 2>(ks, items, work, nwork, kp, t, res, s);
    case 5: return launch_gemm_t<256, 64, 3, 2, 2, 1>(ks, items, work, nwork, kp, t, res, s);
    case 6: return launch_gemm_t<256, 32, 4, 2, 2, 1>(ks, items, work, nwork, kp, t, res, s);
    case 7: return launch_gemm_t<256, 64, 4, 2, 2, 1>(ks, items, work, nwork, kp, t, res, s);
    case 8: return launch_gemm_pipe_t<256, 64, 3>(ks, items, work, nwork, kp, t, res, s);
    case 9: return launch_gemm_pipe_t<256, 64, 2>(ks, items, work, nwork, kp, t, res, s);
    case 10: return launch_gemm_pipe_t<256, 64, 4>(ks, items, work, nwork, kp, t, res, s);
    default: return EXC_ERR_ARG;
  }
}

int exc_launch_crt_merge(const exc_dev &d, const kdesc *ks, const mt_item *items, int nitems, const CrtTab &t,
                         const signed char *res, int variant, const double *gsig, hipStream_t s) {
  if (nitems <= 0) return EXC_OK;
  int bt = 0, wr = 0, wc = 0;
  if (exc_crt_shape(variant, &bt, &wr, &wc)) return EXC_ERR_ARG;
  switch (t.n) {
    case 8: return launch_merge_t<8>(d, ks, items, nitems, t, res, bt, wr, wc, gsig, s);
    case 9: return launch_merge_t<9>(d, ks, items, nitems, t, res, bt, wr, wc, gsig, s);
    case 10: return launch_merge_t<10>(d, ks, items, nitems, t, res, bt, wr, wc, gsig, s);
    case 11: return launch_merge_t<11>(d, ks, items, nitems, t, res, bt, wr, wc, gsig, s);
    case 12: return launch_merge_t<12>(d, ks, items, nitems, t, res, bt, wr, wc, gsig, s);
    default: return EXC_ERR_ARG;
  }
}

namespace {
// T_a(r, f) = conj(U_a(f, r)) sqrt(lambda_r + sigma), column-major per atom: the table the Y
// stage reads in h_a's place (it forms A_a T_a^T = A_a conj(U_a) (Lambda_a + sigma)^(1/2))
__global__ __launch_bounds__(256) void crt_gram_table_kernel(int naa, const cplx *__restrict__ U,
                                                             const double *__restrict__ lam, double sigma,
                                                             cplx *__restrict__ T) {
  const int r = blockIdx.x * 256 + threadIdx.x, f = blockIdx.y, a = blockIdx.z;
  if (r >= naa) return;
  const size_t o = (size_t)a * naa * naa;
  const cplx u = U[o + f + (size_t)naa * r];
  const double w = sqrt(fmax(lam[(size_t)a * naa + r] + sigma, 0.0));
  T[o + r + (size_t)naa * f] = cplx{u.re * w, -u.im * w};
}
}  // namespace

int exc_launch_crt_gram_table(const exc_dev &d, const cplx *U, const double *lam, double sigma, cplx *T,
                              hipStream_t s) {
  hipLaunchKernelGGL(crt_gram_table_kernel, dim3((d.naa + 255) / 256, d.naa, d.natm), dim3(256), 0, s, d.naa, U, lam,
                     sigma, T);
  EXC_CHECK(hipGetLastError());
  return EXC_OK;
}

namespace {
// M_p = P (h + sigma I) P^T per atom: h's rows and columns in the Y stage's summation order
__global__ __launch_bounds__(256) void crt_gram_perm_kernel(int naa, const cplx *__restrict__ h,
                                                            const int *__restrict__ fperm, double sigma,
                                                            cplx *__restrict__ Mp) {
  const int a = blockIdx.x * 256 + threadIdx.x, b = blockIdx.y, at = blockIdx.z;
  if (a >= naa) return;
  const size_t o = (size_t)at * naa * naa;
  cplx v = h[o + fperm[a] + (size_t)naa * fperm[b]];
  if (a == b) v.re += sigma;
  Mp[o + a + (size_t)naa * b] = v;
}

// T(r, fperm[kk]) = conj(L(kk, r)) for kk >= r and 0 above, L the lower Cholesky factor of M_p:
// the Y stage on T forms E = A_p conj(L), conj(E) E^T = conj(A) (h + sigma) A^T, and the sum of
// a column tile starting at r0 may start at position r0 of the order
__global__ __launch_bounds__(256) void crt_gram_tri_kernel(int naa, const cplx *__restrict__ L,
                                                           const int *__restrict__ fperm, cplx *__restrict__ T) {
  const int r = blockIdx.x * 256 + threadIdx.x, kk = blockIdx.y, at = blockIdx.z;
  if (r >= naa) return;
  const size_t o = (size_t)at * naa * naa;
  cplx v = {0.0, 0.0};
  if (kk >= r) {
    const cplx l = L[o + kk + (size_t)naa * r];
    v = cplx{l.re, -l.im};
  }
  T[o + r + (size_t)naa * fperm[kk]] = v;
}
}  // namespace

int exc_launch_crt_gram_perm(const exc_dev &d, double sigma, cplx *Mp, hipStream_t s) {
  hipLaunchKernelGGL(crt_gram_perm_kernel, dim3((d.naa + 255) / 256, d.naa, d.natm), dim3(256), 0, s, d.naa, d.haa,
                     d.fperm, sigma, Mp);
  EXC_CHECK(hipGetLastError());
  return EXC_OK;
}

int exc_launch_crt_gram_tri(const exc_dev &d, const cplx *L, cplx *T, hipStream_t s) {
  hipLaunchKernelGGL(crt_gram_tri_kernel, dim3((d.naa + 255) / 256, d.naa, d.natm), dim3(256), 0, s, d.naa, L, d.fperm,
                     T);
  EXC_CHECK(hipGetLastError());
  return EXC_OK;
}
