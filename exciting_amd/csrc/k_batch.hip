// Grouped launches over all k-points of a batch (each k with its own basis size N_k).
//
// match, Y stage and the lo blocks are grids with a k dimension; the muffin-tin stage runs a
// work list of (k, tile row, tile column, H or O) items over all k-points, one item per
// workgroup, so the uneven tile counts of the k-points are balanced over the whole chip
// instead of leaving CUs idle at the end of each k-point.  The device code is the same as
// the per-k path (match_dev.h, setup_dev.h, mt_dma.h, lo_dev.h).
#include "lo_dev.h"
#include "match_dev.h"
#include "setup_dev.h"
#include "mt_dma.h"

namespace {

using namespace setup_dev;

constexpr int kMatchThreads = 64;

template <int L>
__global__ __launch_bounds__(kMatchThreads) void bmatch_kernel(const kdesc *__restrict__ ks, int apwordmax,
                                                                const int *__restrict__ spidx,
                                                                const double *__restrict__ rmt, double fpso,
                                                                const double *__restrict__ dinv, int nl,
                                                                const double *__restrict__ olpl) {
  const kdesc k = ks[blockIdx.z];
  const int ig = blockIdx.x * kMatchThreads + threadIdx.x;
  if (ig >= k.ngp) return;
  match_dev::match_body<L>(ig, blockIdx.y, apwordmax, spidx, rmt, fpso, dinv, k.gpc, k.tpgpc, k.sfacgp, k.ldg,
                           k.apwalm, k.ldg, k.rf, nl, k.ngp,
                           k.pa2 ? k.pa2 + (size_t)blockIdx.y * k.cnp + ig : nullptr, olpl);
}

template <int L>
int launch_bmatch_l(const exc_dev &d, int nk, int maxngp, const kdesc *ks, hipStream_t s) {
  dim3 grid((maxngp + kMatchThreads - 1) / kMatchThreads, d.natm, nk);
  hipLaunchKernelGGL(bmatch_kernel<L>, grid, dim3(kMatchThreads), 0, s, ks, d.apwordmax, d.spidx, d.rmt,
                     4.0 * M_PI / sqrt(d.omega), d.dinv, d.lmaxmat + 1, (d.iora & 1) ? d.olpl : nullptr);
  EXC_CHECK(hipGetLastError());
  return EXC_OK;
}

template <int BM, int BN, int BK, int NST, int WR, int WC, bool G3 = false>
__global__ __launch_bounds__(64 * WR * WC) void bystage_dma_kernel(const kdesc *__restrict__ ks, int natm, int naa,
                                                                   const cplx *__restrict__ haa,
                                                                   const int *__restrict__ fperm, int tri) {
  using T = YTileDMA<BM, BN, BK, NST, WR, WC, G3>;
  using St = typename T::Stage;
  EXC_DMA_STAGES(St, NST)
  const int kk = blockIdx.z / natm, ias = blockIdx.z % natm;
  const kdesc k = ks[kk];
  const int row0 = blockIdx.x * BM;
  if (row0 >= k.ngp) return;
  // tri: the table is zero at positions of fperm before its column (the Gram form's
  // triangular factor), so this column tile's sum starts at its first column
  const int col0 = blockIdx.y * BN;
  T::run(EXC_DMA_STAGE_ARGS(St), k.ngp, naa, k.apwalm + (size_t)k.ldg * naa * ias, k.ldg,
         haa + (size_t)naa * naa * ias, k.yt + (size_t)k.ldy * naa * ias, k.ldy, row0, col0, fperm,
         tri ? min(col0, naa) : 0);
}

template <int BM, int BN, int BK, int NST, int WR, int WC, bool G3 = false>
int launch_bystage_dma_t(const exc_dev &d, int nk, int maxngp, const kdesc *ks, hipStream_t s) {
  dim3 grid((maxngp + BM - 1) / BM, (d.naa + BN - 1) / BN, d.natm * nk);
  hipLaunchKernelGGL((bystage_dma_kernel<BM, BN, BK, NST, WR, WC, G3>), grid, dim3(64 * WR * WC), 0, s, ks, d.natm, d.naa,
                     d.haa, d.fperm, d.ytri);
  EXC_CHECK(hipGetLastError());
  return EXC_OK;
}

// MT items on a plain grid, one item per workgroup, LDS filled by DMA with NST stages in
// flight: workgroup b runs on XCD b % 8 and takes the next item of that XCD's contiguous
// eighth of the list (the k-points an XCD works on stay in its L2), while the dispatcher
// balances the load dynamically
template <int BM, int BK, int NST, int WR, int WC, bool G3 = false>
__global__ __launch_bounds__(64 * WR * WC) void bmt_dma_grid_kernel(const kdesc *__restrict__ ks,
                                                                    const mt_item *__restrict__ items, int nitems,
                                                                    int ktot, GTab gt,
                                                                    const int *__restrict__ kperm) {
  using T = MtTileDMA<BM, BK, NST, WR, WC, G3>;
  using St = typename T::Stage;
  __shared__ typename T::Tables tb;
  EXC_DMA_STAGES(St, NST)
  const int chunk = (nitems + 7) / 8;
  const int x = blockIdx.x % 8, q = blockIdx.x / 8;
  const int it = x * chunk + q;
  if (q >= chunk || it >= nitems) return;
  const mt_item w = items[it];
  const kdesc k = ks[w.k];
  T::run(tb, EXC_DMA_STAGE_ARGS(St), k.ngp, ktot, k.apwalm, k.ldg, k.yt, k.ldy, k.igpig, k.vgpc, gt, k.h, k.o,
         k.ldh, kperm, w.ti, w.tj, w.which);
}

template <int BM, int BK, int NST, int WR, int WC, bool G3 = false>
int launch_bmt_dma_grid_t(const exc_dev &d, const kdesc *ks, const mt_item *items, int nitems, hipStream_t s) {
  const int chunk = (nitems + 7) / 8;
  hipLaunchKernelGGL((bmt_dma_grid_kernel<BM, BK, NST, WR, WC, G3>), dim3(8 * chunk), dim3(64 * WR * WC), 0, s, ks,
                     items, nitems, d.natm * d.naa, make_gtab(d), d.kperm);
  EXC_CHECK(hipGetLastError());
  return EXC_OK;
}

__global__ __launch_bounds__(lo_dev::kRows * lo_dev::kSlots) void blo_apw_kernel(const kdesc *__restrict__ ks, int naa,
                                                                                  int maxnlo, int nlomax, int apwordmax,
                                                                                  int nlotot,
                                                                                  const lo_desc *__restrict__ lod,
                                                                                  const cplx *__restrict__ loa,
                                                                                  const double *__restrict__ oalo,
                                                                                  const double *__restrict__ h1loa) {
  __shared__ lo_dev::ApwLoShared sh;
  const kdesc k = ks[blockIdx.z];
  if ((int)blockIdx.x * lo_dev::kRows >= k.ngp) return;
  lo_dev::lo_apw_body(sh, blockIdx.x, blockIdx.y, k.ngp, naa, maxnlo, nlomax, apwordmax, nlotot, lod, k.apwalm, k.ldg,
                      loa, oalo, h1loa, k.h, k.o, k.ldh);
}

__global__ __launch_bounds__(lo_dev::kLoThreads) void blo_lo_kernel(const kdesc *__restrict__ ks, int nlotot, int maxnlo,
                                                                     int nlomax, const lo_desc *__restrict__ lod,
                                                                     const cplx *__restrict__ lolo,
                                                                     const double *__restrict__ ololo,
                                                                     const double *__restrict__ h1lolo) {
  const kdesc k = ks[blockIdx.z];
  lo_dev::lo_lo_body(blockIdx.x * lo_dev::kLoThreads + threadIdx.x, blockIdx.y, k.ngp, nlotot, maxnlo, nlomax, lod,
                     lolo, ololo, h1lolo, k.h, k.o, k.ldh);
}

}  // namespace

int exc_launch_bmatch(const exc_dev &d, int nk, int maxngp, const kdesc *ks, hipStream_t s) {
  if (nk <= 0 || maxngp <= 0) return EXC_OK;
  switch (d.lmaxapw) {
    case 1: return launch_bmatch_l<1>(d, nk, maxngp, ks, s);
    case 2: return launch_bmatch_l<2>(d, nk, maxngp, ks, s);
    case 3: return launch_bmatch_l<3>(d, nk, maxngp, ks, s);
    case 4: return launch_bmatch_l<4>(d, nk, maxngp, ks, s);
    case 5: return launch_bmatch_l<5>(d, nk, maxngp, ks, s);
    case 6: return launch_bmatch_l<6>(d, nk, maxngp, ks, s);
    case 7: return launch_bmatch_l<7>(d, nk, maxngp, ks, s);
    case 8: return launch_bmatch_l<8>(d, nk, maxngp, ks, s);
    case 9: return launch_bmatch_l<9>(d, nk, maxngp, ks, s);
    case 10: return launch_bmatch_l<10>(d, nk, maxngp, ks, s);
    case 11: return launch_bmatch_l<11>(d, nk, maxngp, ks, s);
    case 12: return launch_bmatch_l<12>(d, nk, maxngp, ks, s);
    default: return EXC_ERR_UNSUPPORTED;
  }
}

// Y-stage variants: as the per-k Y stage's (exc_launch_ystage); 0 is the grouped default
int exc_launch_bystage(const exc_dev &d, int nk, int maxngp, const kdesc *ks, int variant, hipStream_t s) {
  if (nk <= 0 || maxngp <= 0) return EXC_OK;
  switch (variant) {
    case 0: return launch_bystage_dma_t<64, 128, 8, 2, 2, 4, true>(d, nk, maxngp, ks, s);   // Gauss's 3 products
    case 1: return launch_bystage_dma_t<64, 128, 16, 2, 2, 4, true>(d, nk, maxngp, ks, s);
    case 2: return launch_bystage_dma_t<64, 128, 8, 2, 2, 4>(d, nk, maxngp, ks, s);
    case 3: return launch_bystage_dma_t<64, 64, 16, 2, 2, 4>(d, nk, maxngp, ks, s);
    case 4: return launch_bystage_dma_t<64, 64, 8, 2, 2, 4, true>(d, nk, maxngp, ks, s);   // narrow, Gauss's form
    default: return EXC_ERR_ARG;
  }
}

int exc_bmt_tile(int variant) { return variant >= 0 && variant < kBmtVariants ? 64 : 0; }

// grouped MT variants: 64 x 64 tile, 8 waves; 0: BK 16, two DMA stages, Gauss's three real
// products per complex product (the default); 1: the same with four products; 2: BK 8, four
// stages, three products; 3: BK 16, three stages, three products
int exc_launch_bmt(const exc_dev &d, const kdesc *ks, const mt_item *items, int nitems, int variant, hipStream_t s) {
  if (nitems <= 0) return EXC_OK;
  switch (variant) {
    case 0: return launch_bmt_dma_grid_t<64, 16, 2, 2, 4, true>(d, ks, items, nitems, s);
    case 1: return launch_bmt_dma_grid_t<64, 16, 2, 2, 4>(d, ks, items, nitems, s);
    case 2: return launch_bmt_dma_grid_t<64, 8, 4, 2, 4, true>(d, ks, items, nitems, s);
    case 3: return launch_bmt_dma_grid_t<64, 16, 3, 2, 4, true>(d, ks, items, nitems, s);
    default: return EXC_ERR_ARG;
  }
}

int exc_launch_blo(const exc_dev &d, int nk, int maxngp, const kdesc *ks, hipStream_t s) {
  if (nk <= 0 || d.nlotot == 0) return EXC_OK;
  if (d.maxnlo > lo_dev::kMaxLo) return EXC_ERR_UNSUPPORTED;
  if (maxngp > 0) {
    dim3 grid((maxngp + lo_dev::kRows - 1) / lo_dev::kRows, d.natm, nk);
    hipLaunchKernelGGL(blo_apw_kernel, grid, dim3(lo_dev::kRows * lo_dev::kSlots), 0, s, ks, d.naa, d.maxnlo, d.nlomax,
                       d.apwordmax, d.nlotot, d.lod, d.loa, d.oalo, (d.iora & 2) ? d.h1loa : nullptr);
    EXC_CHECK(hipGetLastError());
  }
  dim3 grid2((d.nlotot + lo_dev::kLoThreads - 1) / lo_dev::kLoThreads, d.nlotot, nk);
  hipLaunchKernelGGL(blo_lo_kernel, grid2, dim3(lo_dev::kLoThreads), 0, s, ks, d.nlotot, d.maxnlo, d.nlomax, d.lod,
                     d.lolo, d.ololo, (d.iora & 2) ? d.h1lolo : nullptr);
  EXC_CHECK(hipGetLastError());
  return EXC_OK;
}
