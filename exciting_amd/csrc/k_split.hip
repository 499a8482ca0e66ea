// Split-BF16 muffin-tin contraction on gfx950 (the precision variant of the grouped setup).
//
// H_MT = conj(A) Yt^T and O_MT = conj(A) A^T are computed from BF16 slices with FP32
// MFMA accumulation (Ozaki-style), the rest of the setup staying FP64:
//
//  split : every real value x of an operand row r and atom block a is written as
//          x = 2^(e-8) (s0 + s1 + s2),  e = exponent of max |x| over (r, a),
//          s0 = q1, s1 = q2 / 2^8, s2 = q3 / 2^16 with integers |q1| <= 256, |q2|, |q3| <= 128,
//          so every slice is exact in BF16 (8 significant bits) and every slice product
//          is exact in FP32.
//  bsplit: per atom block (256 k, naa padded), the six slice pairs with s + t <= 2 are
//          accumulated by v_mfma_f32_32x32x16_bf16 into two FP32 levels — the leading pair
//          (s0 x s0, integer sums, exact up to 2^24) and the five smaller pairs — then folded
//          into FP64 totals with the power-of-two factor 2^(ea(r,a) + eb(c,a) - 16).  The
//          interstitial epilogue and the Hermitian mirror are the FP64 path's.
//  The dropped pairs (s + t >= 3) and the slice truncation leave a relative error of
//  about 2^-24 of each (row, atom) block's largest element.
#include "setup_dev.h"

namespace {

using setup_dev::GTab;
using setup_dev::lookup_g;

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x16 __attribute__((ext_vector_type(16)));

constexpr int KB = kSplitKB;          // k per atom block
constexpr int NP = kSplitPlanes;      // planes per k: (s0,re) (s0,im) (s1,re) (s1,im) (s2,re) (s2,im)
constexpr int CH = 16;                // k per chunk (one MFMA k-step)
constexpr int NCH = KB / CH;          // chunks per atom block
constexpr int ROWB = NP * CH * 2;     // bytes of one (row, chunk): 192

__device__ __forceinline__ unsigned short bf16_bits_exact(double v) {
  // v is q * 2^k with |q| <= 256: exact in float and in bf16 (low 16 float bits are zero)
  return (unsigned short)(__float_as_uint((float)v) >> 16);
}

// ---------------------------------------------------------------------------------------
// split: one thread per (row, atom) of one operand (blockIdx.z = 2*k + operand)
// ---------------------------------------------------------------------------------------
constexpr int kSplitThreads = 64;

__global__ __launch_bounds__(kSplitThreads) void split_kernel(const kdesc *__restrict__ ks, int natm, int naa) {
  const kdesc k = ks[blockIdx.z >> 1];
  const int which = blockIdx.z & 1;              // 0: A (apwalm), 1: Yt
  const int row = blockIdx.x * kSplitThreads + threadIdx.x;
  const int atom = blockIdx.y;
  if (row >= k.ngp) return;
  const int ld = which ? k.ldy : k.ldg;
  const cplx *src = (which ? k.yt : k.apwalm) + row + (size_t)ld * naa * atom;
  unsigned short *dst = (which ? k.sy : k.sa) + ((size_t)row * natm + atom) * NCH * NP * CH;
  int *ex = (which ? k.ey : k.ea) + (size_t)row * natm + atom;
  double m = 0.0;
  for (int f = 0; f < naa; ++f) {
    const cplx v = src[(size_t)f * ld];
    m = fmax(m, fmax(fabs(v.re), fabs(v.im)));
  }
  int e = 0;
  if (m > 0.0) frexp(m, &e);                     // m < 2^e
  *ex = e;
  for (int c = 0; c < NCH; ++c) {
    unsigned short sl[NP][CH];
#pragma unroll
    for (int kk = 0; kk < CH; ++kk) {
      const int f = c * CH + kk;
      cplx v = {0.0, 0.0};
      if (f < naa) v = src[(size_t)f * ld];
#pragma unroll
      for (int part = 0; part < 2; ++part) {
        const double x = ldexp(part ? v.im : v.re, 8 - e);
        const double q1 = rint(x);
        const double r1 = (x - q1) * 256.0;
        const double q2 = rint(r1);
        const double q3 = rint((r1 - q2) * 256.0);
        sl[0 + part][kk] = bf16_bits_exact(q1);
        sl[2 + part][kk] = bf16_bits_exact(q2 * (1.0 / 256.0));
        sl[4 + part][kk] = bf16_bits_exact(q3 * (1.0 / 65536.0));
      }
    }
    uint4 *o4 = (uint4 *)(dst + (size_t)c * NP * CH);
#pragma unroll
    for (int p = 0; p < NP; ++p) {
#pragma unroll
      for (int h = 0; h < 2; ++h) {
        uint4 w;
        w.x = sl[p][8 * h + 0] | ((unsigned)sl[p][8 * h + 1] << 16);
        w.y = sl[p][8 * h + 2] | ((unsigned)sl[p][8 * h + 3] << 16);
        w.z = sl[p][8 * h + 4] | ((unsigned)sl[p][8 * h + 5] << 16);
        w.w = sl[p][8 * h + 6] | ((unsigned)sl[p][8 * h + 7] << 16);
        o4[p * 2 + h] = w;
      }
    }
  }
}

// ---------------------------------------------------------------------------------------
// bsplit: persistent over (k, tile, H/O) items; workgroup tile BM x BM, waves WR x WC,
// wave tile (32*MI) x (32*NJ) of 32x32 MFMA blocks
// ---------------------------------------------------------------------------------------
// v_mfma_f32_32x32x16_bf16 output layout: register v of lane l holds
//   row 8*(v/4) + 4*(l/32) + v%4, column l%32.
__device__ __forceinline__ int d32_row(int v, int lane) { return 8 * (v >> 2) + 4 * (lane >> 5) + (v & 3); }

template <int BM, int WR, int WC>
struct SplitCfg {
  static constexpr int kThreads = 64 * WR * WC;
  static constexpr int WM = BM / WR, WN = BM / WC;
  static constexpr int MI = WM / 32, NJ = WN / 32;
  static constexpr int BKC = 2;                          // chunks per LDS stage
  static constexpr int RS = BKC * ROWB + 16;             // LDS row stride (bytes), padded against bank conflicts
  static constexpr int SEGS = BKC * ROWB / 16;           // 16-byte segments per row and stage
  static constexpr int LD = (2 * BM * SEGS + kThreads - 1) / kThreads;
  struct Shared {
    unsigned char a[BM * RS];
    unsigned char b[BM * RS];
    int ea[BM], eb[BM];
    int gr[BM][3], gc[BM][3];
    double vr[BM][3], vc[BM][3];
  };
};

template <int BM, int WR, int WC>
__global__ __launch_bounds__(64 * WR * WC) void bsplit_kernel(const kdesc *__restrict__ ks,
                                                              const mt_item *__restrict__ items, int nitems, int natm,
                                                              GTab gt) {
  using S = SplitCfg<BM, WR, WC>;
  constexpr int MI = S::MI, NJ = S::NJ, RS = S::RS, SEGS = S::SEGS, LD = S::LD;
  __shared__ typename S::Shared sh;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int wr = wave / WC, wc = wave % WC;
  const int nx = gridDim.x >= 8 ? 8 : 1;
  const int x = blockIdx.x % nx, lb = blockIdx.x / nx, per = gridDim.x / nx;
  const int chunk = (nitems + nx - 1) / nx;
  const int end = min(nitems, (x + 1) * chunk);
  for (int it = x * chunk + lb; it < end; it += per) {
    const mt_item w = items[it];
    const kdesc k = ks[w.k];
    const int ngp = k.ngp;
    const int row0 = w.ti * BM, col0 = w.tj * BM;
    const unsigned short *PA = k.sa;
    const unsigned short *PB = w.which ? k.sa : k.sy;
    const int *EA = k.ea;
    const int *EB = w.which ? k.ea : k.ey;
    __syncthreads();
    for (int e = tid; e < 2 * BM; e += S::kThreads) {
      const bool isr = e < BM;
      const int idx = isr ? row0 + e : col0 + (e - BM);
      int g0 = 0, g1 = 0, g2 = 0;
      double v0 = 0.0, v1 = 0.0, v2 = 0.0;
      if (idx < ngp) {
        const int ig = k.igpig[idx] - 1;
        g0 = gt.ivg[3 * ig]; g1 = gt.ivg[3 * ig + 1]; g2 = gt.ivg[3 * ig + 2];
        v0 = k.vgpc[3 * idx]; v1 = k.vgpc[3 * idx + 1]; v2 = k.vgpc[3 * idx + 2];
      }
      const int q = isr ? e : e - BM;
      int(&g)[BM][3] = isr ? sh.gr : sh.gc;
      double(&vv)[BM][3] = isr ? sh.vr : sh.vc;
      g[q][0] = g0; g[q][1] = g1; g[q][2] = g2;
      vv[q][0] = v0; vv[q][1] = v1; vv[q][2] = v2;
    }
    double tre[MI][NJ][16], tim[MI][NJ][16];
#pragma unroll
    for (int i = 0; i < MI; ++i)
#pragma unroll
      for (int j = 0; j < NJ; ++j)
#pragma unroll
        for (int v = 0; v < 16; ++v) tre[i][j][v] = tim[i][j][v] = 0.0;

    for (int atom = 0; atom < natm; ++atom) {
      f32x16 a0r[MI][NJ], a0i[MI][NJ], a1r[MI][NJ], a1i[MI][NJ];
#pragma unroll
      for (int i = 0; i < MI; ++i)
#pragma unroll
        for (int j = 0; j < NJ; ++j) a0r[i][j] = a0i[i][j] = a1r[i][j] = a1i[i][j] = f32x16{};
      for (int c0 = 0; c0 < NCH; c0 += S::BKC) {
        // stage BKC chunks of this atom block for the tile's rows (A) and columns (B)
        uint4 reg[LD];
#pragma unroll
        for (int r = 0; r < LD; ++r) {
          const int e = tid + S::kThreads * r;
          reg[r] = make_uint4(0u, 0u, 0u, 0u);
          if (e < 2 * BM * SEGS) {
            const bool isa = e < BM * SEGS;
            const int ee = isa ? e : e - BM * SEGS;
            const int rr = ee / SEGS, seg = ee % SEGS;
            const int gr = (isa ? row0 : col0) + rr;
            if (gr < ngp) {
              const unsigned short *P = isa ? PA : PB;
              const uint4 *src = (const uint4 *)(P + (((size_t)gr * natm + atom) * NCH + c0) * NP * CH);
              reg[r] = src[seg];
            }
          }
        }
        __syncthreads();
#pragma unroll
        for (int r = 0; r < LD; ++r) {
          const int e = tid + S::kThreads * r;
          if (e < 2 * BM * SEGS) {
            const bool isa = e < BM * SEGS;
            const int ee = isa ? e : e - BM * SEGS;
            const int rr = ee / SEGS, seg = ee % SEGS;
            *(uint4 *)((isa ? sh.a : sh.b) + rr * RS + seg * 16) = reg[r];
          }
        }
        if (c0 == 0) {
          for (int e = tid; e < 2 * BM; e += S::kThreads) {
            const bool isr = e < BM;
            const int idx = isr ? row0 + e : col0 + (e - BM);
            const int ex = idx < ngp ? (isr ? EA : EB)[(size_t)idx * natm + atom] : 0;
            (isr ? sh.ea : sh.eb)[isr ? e : e - BM] = ex;
          }
        }
        __syncthreads();
#pragma unroll
        for (int cc = 0; cc < S::BKC; ++cc) {
          bf16x8 av[MI][NP], bv[NJ][NP], nai[MI][3];
#pragma unroll
          for (int i = 0; i < MI; ++i) {
            const int rr = wr * S::WM + i * 32 + (lane & 31);
#pragma unroll
            for (int p = 0; p < NP; ++p)
              av[i][p] = *(const bf16x8 *)(sh.a + rr * RS + (cc * NP + p) * 32 + (lane >> 5) * 16);
#pragma unroll
            for (int s = 0; s < 3; ++s) {
              uint4 t = *(const uint4 *)&av[i][2 * s + 1];
              t.x ^= 0x80008000u; t.y ^= 0x80008000u; t.z ^= 0x80008000u; t.w ^= 0x80008000u;
              nai[i][s] = *(const bf16x8 *)&t;
            }
          }
#pragma unroll
          for (int j = 0; j < NJ; ++j) {
            const int rr = wc * S::WN + j * 32 + (lane & 31);
#pragma unroll
            for (int p = 0; p < NP; ++p)
              bv[j][p] = *(const bf16x8 *)(sh.b + rr * RS + (cc * NP + p) * 32 + (lane >> 5) * 16);
          }
#pragma unroll
          for (int i = 0; i < MI; ++i)
#pragma unroll
            for (int j = 0; j < NJ; ++j) {
              // level 0: (s0, s0); conj(a) b: re = ar br + ai bi, im = ar bi - ai br
              a0r[i][j] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(av[i][0], bv[j][0], a0r[i][j], 0, 0, 0);
              a0r[i][j] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(av[i][1], bv[j][1], a0r[i][j], 0, 0, 0);
              a0i[i][j] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(av[i][0], bv[j][1], a0i[i][j], 0, 0, 0);
              a0i[i][j] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(nai[i][0], bv[j][0], a0i[i][j], 0, 0, 0);
              // level 1: (s0,s1) (s1,s0) (s0,s2) (s1,s1) (s2,s0)
#pragma unroll
              for (int pr = 0; pr < 5; ++pr) {
                const int s = (pr == 1 || pr == 3) ? 1 : (pr == 4 ? 2 : 0);
                const int t = (pr == 0 || pr == 3) ? 1 : (pr == 2 ? 2 : 0);
                a1r[i][j] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(av[i][2 * s], bv[j][2 * t], a1r[i][j], 0, 0, 0);
                a1r[i][j] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(av[i][2 * s + 1], bv[j][2 * t + 1], a1r[i][j], 0, 0, 0);
                a1i[i][j] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(av[i][2 * s], bv[j][2 * t + 1], a1i[i][j], 0, 0, 0);
                a1i[i][j] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(nai[i][s], bv[j][2 * t], a1i[i][j], 0, 0, 0);
              }
            }
        }
      }
      // fold this atom block into the FP64 totals: x 2^(ea(row) + eb(col) - 16)
#pragma unroll
      for (int i = 0; i < MI; ++i)
#pragma unroll
        for (int j = 0; j < NJ; ++j) {
          const int cl = wc * S::WN + j * 32 + (lane & 31);
          const int ebc = sh.eb[cl] - 16;
#pragma unroll
          for (int v = 0; v < 16; ++v) {
            const int rl = wr * S::WM + i * 32 + d32_row(v, lane);
            const int sc = sh.ea[rl] + ebc;
            tre[i][j][v] += ldexp((double)a0r[i][j][v] + (double)a1r[i][j][v], sc);
            tim[i][j][v] += ldexp((double)a0i[i][j][v] + (double)a1i[i][j][v], sc);
          }
        }
    }
    // interstitial + mirror epilogue (as the FP64 path)
    cplx *out = w.which ? k.o : k.h;
#pragma unroll
    for (int i = 0; i < MI; ++i)
#pragma unroll
      for (int j = 0; j < NJ; ++j) {
        const int cl = wc * S::WN + j * 32 + (lane & 31);
        const int c = col0 + cl;
#pragma unroll
        for (int v = 0; v < 16; ++v) {
          const int rl = wr * S::WM + i * 32 + d32_row(v, lane);
          const int r = row0 + rl;
          if (r > c || c >= ngp) continue;
          double re = tre[i][j][v], im = tim[i][j][v];
          const int ig = lookup_g(gt, sh.gr[rl][0] - sh.gc[cl][0], sh.gr[rl][1] - sh.gc[cl][1], sh.gr[rl][2] - sh.gc[cl][2]);
          if (ig) {
            const cplx th = gt.cfunig[ig - 1];
            if (w.which) {
              re += th.re;
              im += th.im;
            } else {
              const cplx vv = gt.veffig[ig - 1];
              const double t1 = 0.5 * (sh.vr[rl][0] * sh.vc[cl][0] + sh.vr[rl][1] * sh.vc[cl][1] + sh.vr[rl][2] * sh.vc[cl][2]);
              re += vv.re + t1 * gt.kinig[ig - 1].re;
              im += vv.im + t1 * gt.kinig[ig - 1].im;
            }
          }
          out[r + (size_t)c * k.ldh] = cplx{re, im};
          if (r < c) out[c + (size_t)r * k.ldh] = cplx{re, -im};
        }
      }
  }
}

// ---------------------------------------------------------------------------------------
// bsplit2: the same arithmetic with a larger tile.  Four waves (2 x 2), wave tile
// (32 MI) x (32 NJ), workgroup tile BM x BN = (64 MI) x (64 NJ).  One 16-k chunk per LDS
// stage, two stages in flight, filled by global_load_lds_dwordx4 (no staging registers);
// operand slices are read from LDS per slice pair just before their MFMAs; FP64 totals,
// folded per atom block exactly as bsplit_kernel.  An LDS row is 13 segments of 16 bytes (12 data + 1 pad) so that the
// 32 rows one MFMA operand read touches fall on distinct banks; the pad lanes are masked
// off the DMA.
// ---------------------------------------------------------------------------------------
constexpr int SEG = 16;                     // bytes per load
constexpr int RSEG = ROWB / SEG;            // data segments per row and chunk: 12
constexpr int PSEG = RSEG + 1;              // with the pad: 13

template <int MI, int NJ>
struct Split2Cfg {
  static constexpr int BM = 64 * MI, BN = 64 * NJ;
  static constexpr int ABYTES = BM * PSEG * SEG, BBYTES = BN * PSEG * SEG;
  static constexpr int ALOADS = BM * PSEG / 64, BLOADS = BN * PSEG / 64;   // wave-wide loads per stage
  static constexpr int STAGE16 = (ABYTES + BBYTES) / 16;   // one LDS stage in 16-byte units
  struct Shared {
    int ea[2][BM], eb[2][BN];
  };
};

typedef __attribute__((address_space(3))) void lds_void;

// DMA of one chunk of the A and B row panels into an LDS stage.  rowb = bytes between
// consecutive rows of a slice plane set (natm * NCH records of ROWB bytes).
template <int MI, int NJ>
__device__ __forceinline__ void split2_issue(unsigned char *stage, const unsigned char *pa, const unsigned char *pb,
                                             int row0, int col0, int ngp, unsigned rowb) {
  using S = Split2Cfg<MI, NJ>;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
#pragma nounroll
  for (int b = wave; b < S::ALOADS + S::BLOADS; b += 4) {
    const bool isa = b < S::ALOADS;
    const int bb = isa ? b : b - S::ALOADS;
    const int L = bb * 64 + lane;
    const int rr = L / PSEG, seg = L - rr * PSEG;
    int gr = (isa ? row0 : col0) + rr;
    if (gr >= ngp) gr = 0;                  // rows past N: any valid row, never stored
    const unsigned char *src = (isa ? pa : pb) + (gr * rowb + (unsigned)seg * SEG);
    unsigned char *dst = stage + (isa ? 0 : S::ABYTES) + bb * 64 * SEG;
    if (seg < RSEG) __builtin_amdgcn_global_load_lds((const void *)src, (lds_void *)dst, 16, 0, 0);
  }
}

template <int MI, int NJ>
__global__ __launch_bounds__(256) void bsplit2_kernel(const kdesc *__restrict__ ks, const mt_item *__restrict__ items,
                                                      int nitems, int natm, GTab gt) {
  using S = Split2Cfg<MI, NJ>;
  constexpr int BM = S::BM, BN = S::BN;
  __shared__ typename S::Shared sh;
  // the two LDS stages are separate variables and the chunk loop alternates them statically:
  // the compiler can then tell the operand reads of one stage from the DMA into the other
  // and does not wait for that DMA before the reads (it would with one indexed array)
  __shared__ uint4 st0[S::STAGE16];
  __shared__ uint4 st1[S::STAGE16];
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int wr = wave >> 1, wc = wave & 1;
  const int nx = gridDim.x >= 8 ? 8 : 1;
  const int x = blockIdx.x % nx, lb = blockIdx.x / nx, per = gridDim.x / nx;
  const int chunk = (nitems + nx - 1) / nx;
  const int end = min(nitems, (x + 1) * chunk);
  const unsigned rowb = (unsigned)natm * NCH * ROWB;
  for (int it = x * chunk + lb; it < end; it += per) {
    const mt_item w = items[it];
    const kdesc k = ks[w.k];
    const int ngp = k.ngp;
    const int row0 = w.ti * BM, col0 = w.tj * BN;
    const unsigned char *PA = (const unsigned char *)k.sa;
    const unsigned char *PB = (const unsigned char *)(w.which ? k.sa : k.sy);
    const int *EA = k.ea;
    const int *EB = w.which ? k.ea : k.ey;
    double tre[MI][NJ][16], tim[MI][NJ][16];
#pragma unroll
    for (int i = 0; i < MI; ++i)
#pragma unroll
      for (int j = 0; j < NJ; ++j)
#pragma unroll
        for (int v = 0; v < 16; ++v) tre[i][j][v] = tim[i][j][v] = 0.0;
    __syncthreads();                         // previous item done with both stages
    split2_issue<MI, NJ>((unsigned char *)st0, PA, PB, row0, col0, ngp, rowb);
    for (int atom = 0; atom < natm; ++atom) {
      // exponents of this atom block; buffer atom&1 was last read by the fold of atom-2,
      // which precedes at least NCH barriers
      int *ea = sh.ea[atom & 1], *eb = sh.eb[atom & 1];
      for (int e = tid; e < BM + BN; e += 256) {
        const bool isr = e < BM;
        const int idx = isr ? row0 + e : col0 + (e - BM);
        const int ex = idx < ngp ? (isr ? EA : EB)[(size_t)idx * natm + atom] : 0;
        if (isr) ea[e] = ex; else eb[e - BM] = ex;
      }
      f32x16 a0r[MI][NJ], a0i[MI][NJ], a1r[MI][NJ], a1i[MI][NJ];
#pragma unroll
      for (int i = 0; i < MI; ++i)
#pragma unroll
        for (int j = 0; j < NJ; ++j) a0r[i][j] = a0i[i][j] = a1r[i][j] = a1i[i][j] = f32x16{};
      auto chunk = [&](const unsigned char *cur, unsigned char *nxt, int cg) {
        __builtin_amdgcn_s_waitcnt(0);
        __syncthreads();                     // chunk cg landed; all waves done with the other stage
        if (cg + 1 < natm * NCH) {
          const unsigned off = (unsigned)(cg + 1) * ROWB;      // record (atom, c) of row 0
          split2_issue<MI, NJ>(nxt, PA + off, PB + off, row0, col0, ngp, rowb);
        }
        const unsigned char *sa = cur + (wr * 32 * MI + (lane & 31)) * PSEG * SEG + (lane >> 5) * 16;
        const unsigned char *sb = cur + S::ABYTES + (wc * 32 * NJ + (lane & 31)) * PSEG * SEG + (lane >> 5) * 16;
        // pairs in the order of bsplit_kernel: (0,0) | (0,1) (1,0) | (0,2) (1,1) (2,0)
#pragma unroll
        for (int pr = 0; pr < 6; ++pr) {
          const int sl = (pr == 2 || pr == 4) ? 1 : (pr == 5 ? 2 : 0);
          const int tt = (pr == 1 || pr == 4) ? 1 : (pr == 3 ? 2 : 0);
          bf16x8 ar[MI], ai[MI], na[MI], br[NJ], bi[NJ];
#pragma unroll
          for (int i = 0; i < MI; ++i) {
            ar[i] = *(const bf16x8 *)(sa + i * 32 * PSEG * SEG + (2 * sl) * 32);
            ai[i] = *(const bf16x8 *)(sa + i * 32 * PSEG * SEG + (2 * sl + 1) * 32);
            uint4 u = *(const uint4 *)&ai[i];
            u.x ^= 0x80008000u; u.y ^= 0x80008000u; u.z ^= 0x80008000u; u.w ^= 0x80008000u;
            na[i] = *(const bf16x8 *)&u;
          }
#pragma unroll
          for (int j = 0; j < NJ; ++j) {
            br[j] = *(const bf16x8 *)(sb + j * 32 * PSEG * SEG + (2 * tt) * 32);
            bi[j] = *(const bf16x8 *)(sb + j * 32 * PSEG * SEG + (2 * tt + 1) * 32);
          }
#pragma unroll
          for (int i = 0; i < MI; ++i)
#pragma unroll
            for (int j = 0; j < NJ; ++j) {
              if (pr == 0) {
                a0r[i][j] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(ar[i], br[j], a0r[i][j], 0, 0, 0);
                a0r[i][j] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(ai[i], bi[j], a0r[i][j], 0, 0, 0);
                a0i[i][j] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(ar[i], bi[j], a0i[i][j], 0, 0, 0);
                a0i[i][j] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(na[i], br[j], a0i[i][j], 0, 0, 0);
              } else {
                a1r[i][j] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(ar[i], br[j], a1r[i][j], 0, 0, 0);
                a1r[i][j] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(ai[i], bi[j], a1r[i][j], 0, 0, 0);
                a1i[i][j] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(ar[i], bi[j], a1i[i][j], 0, 0, 0);
                a1i[i][j] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(na[i], br[j], a1i[i][j], 0, 0, 0);
              }
            }
        }
      };
      static_assert(NCH % 2 == 0, "the stage alternation needs an even chunk count per atom");
      for (int c = 0; c < NCH; c += 2) {
        chunk((const unsigned char *)st0, (unsigned char *)st1, atom * NCH + c);
        chunk((const unsigned char *)st1, (unsigned char *)st0, atom * NCH + c + 1);
      }
      // fold the atom block into the FP64 totals: x 2^(ea(row) + eb(col) - 16)
#pragma unroll
      for (int i = 0; i < MI; ++i)
#pragma unroll
        for (int j = 0; j < NJ; ++j) {
          const int cl = wc * (32 * NJ) + j * 32 + (lane & 31);
          const int ebc = eb[cl] - 16;
#pragma unroll
          for (int v = 0; v < 16; ++v) {
            const int rl = wr * (32 * MI) + i * 32 + d32_row(v, lane);
            const int sc = ea[rl] + ebc;
            tre[i][j][v] += ldexp((double)a0r[i][j][v] + (double)a1r[i][j][v], sc);
            tim[i][j][v] += ldexp((double)a0i[i][j][v] + (double)a1i[i][j][v], sc);
          }
        }
    }
    // interstitial + mirror epilogue (as the FP64 path); G vectors straight from global
    cplx *out = w.which ? k.o : k.h;
#pragma unroll
    for (int j = 0; j < NJ; ++j) {
      const int c = col0 + wc * (32 * NJ) + j * 32 + (lane & 31);
      if (c >= ngp) continue;
      const int igc = k.igpig[c] - 1;
      const int gc0 = gt.ivg[3 * igc], gc1 = gt.ivg[3 * igc + 1], gc2 = gt.ivg[3 * igc + 2];
      const double vc0 = k.vgpc[3 * c], vc1 = k.vgpc[3 * c + 1], vc2 = k.vgpc[3 * c + 2];
#pragma unroll
      for (int i = 0; i < MI; ++i)
#pragma unroll
        for (int v = 0; v < 16; ++v) {
          const int r = row0 + wr * (32 * MI) + i * 32 + d32_row(v, lane);
          if (r > c) continue;
          double re = tre[i][j][v], im = tim[i][j][v];
          const int igr = k.igpig[r] - 1;
          const int ig = lookup_g(gt, gt.ivg[3 * igr] - gc0, gt.ivg[3 * igr + 1] - gc1, gt.ivg[3 * igr + 2] - gc2);
          if (ig) {
            const cplx th = gt.cfunig[ig - 1];
            if (w.which) {
              re += th.re;
              im += th.im;
            } else {
              const cplx vv = gt.veffig[ig - 1];
              const double t1 = 0.5 * (k.vgpc[3 * r] * vc0 + k.vgpc[3 * r + 1] * vc1 + k.vgpc[3 * r + 2] * vc2);
              re += vv.re + t1 * gt.kinig[ig - 1].re;
              im += vv.im + t1 * gt.kinig[ig - 1].im;
            }
          }
          out[r + (size_t)c * k.ldh] = cplx{re, im};
          if (r < c) out[c + (size_t)r * k.ldh] = cplx{re, -im};
        }
    }
  }
}

template <int MI, int NJ>
int launch_bsplit2_t(const exc_dev &d, const kdesc *ks, const mt_item *items, int nitems, int nblocks, hipStream_t s) {
  hipLaunchKernelGGL((bsplit2_kernel<MI, NJ>), dim3(nblocks), dim3(256), 0, s, ks, items, nitems, d.natm,
                     setup_dev::make_gtab(d));
  EXC_CHECK(hipGetLastError());
  return EXC_OK;
}

template <int MI, int NJ>
int max_blocks2_t() {
  int n = 0;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, bsplit2_kernel<MI, NJ>, 256, 0) != hipSuccess) return 1;
  return n > 0 ? n : 1;
}

template <int BM, int WR, int WC>
int launch_bsplit_t(const exc_dev &d, const kdesc *ks, const mt_item *items, int nitems, int nblocks, hipStream_t s) {
  hipLaunchKernelGGL((bsplit_kernel<BM, WR, WC>), dim3(nblocks), dim3(64 * WR * WC), 0, s, ks, items, nitems, d.natm,
                     setup_dev::make_gtab(d));
  EXC_CHECK(hipGetLastError());
  return EXC_OK;
}

template <int BM, int WR, int WC>
int max_blocks_t() {
  int n = 0;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, bsplit_kernel<BM, WR, WC>, 64 * WR * WC, 0) != hipSuccess)
    return 1;
  return n > 0 ? n : 1;
}

}  // namespace

size_t exc_split_plane_elems(int ngp, int natm) { return (size_t)ngp * natm * KB * NP; }

int exc_launch_split(const exc_dev &d, int nk, int maxngp, const kdesc *ks, hipStream_t s) {
  if (nk <= 0 || maxngp <= 0) return EXC_OK;
  if (d.naa > KB) return EXC_ERR_UNSUPPORTED;
  dim3 grid((maxngp + kSplitThreads - 1) / kSplitThreads, d.natm, 2 * nk);
  hipLaunchKernelGGL(split_kernel, grid, dim3(kSplitThreads), 0, s, ks, d.natm, d.naa);
  EXC_CHECK(hipGetLastError());
  return EXC_OK;
}

int exc_bsplit_tile(int variant) {
  switch (variant) {
    case 0: return 64;
    case 1: return 32;
    case 2: return 128;
    case 3: return 64;
    default: return 0;
  }
}

int exc_bsplit_max_blocks(int variant) {
  switch (variant) {
    case 0: return max_blocks_t<64, 2, 2>();
    case 1: return max_blocks_t<32, 1, 1>();
    case 2: return max_blocks2_t<2, 2>();
    case 3: return max_blocks2_t<1, 1>();
    default: return 1;
  }
}

int exc_launch_bsplit(const exc_dev &d, const kdesc *ks, const mt_item *items, int nitems, int variant, int nblocks,
                      hipStream_t s) {
  if (nitems <= 0) return EXC_OK;
  switch (variant) {
    case 0: return launch_bsplit_t<64, 2, 2>(d, ks, items, nitems, nblocks, s);
    case 1: return launch_bsplit_t<32, 1, 1>(d, ks, items, nitems, nblocks, s);
    case 2: return launch_bsplit2_t<2, 2>(d, ks, items, nitems, nblocks, s);
    case 3: return launch_bsplit2_t<1, 1>(d, ks, items, nitems, nblocks, s);
    default: return EXC_ERR_ARG;
  }
}
