// Plain FP64 per-k H+O setup on gfx950: the per-k launches of the Y stage and the fused
// muffin-tin + interstitial + mirror stage (device code in setup_dev.h).
#include "setup_dev.h"
#include "mt_dma.h"

namespace {

using namespace setup_dev;

// the matrices to compute: H when h is given, O when o is given
inline int sel_count(const cplx *h, const cplx *o) { return (h ? 1 : 0) + (o ? 1 : 0); }

template <int BM, int BN, int BK, int NST, int WR, int WC, bool G3 = false>
__global__ __launch_bounds__(64 * WR * WC) void ystage_dma_kernel(int ngp, int naa, const cplx *__restrict__ apwalm,
                                                                  int ldg, const cplx *__restrict__ haa,
                                                                  cplx *__restrict__ yt, int ldy,
                                                                  const int *__restrict__ fperm) {
  using T = YTileDMA<BM, BN, BK, NST, WR, WC, G3>;
  using St = typename T::Stage;
  EXC_DMA_STAGES(St, NST)
  const int ias = blockIdx.z;
  T::run(EXC_DMA_STAGE_ARGS(St), ngp, naa, apwalm + (size_t)ldg * naa * ias, ldg, haa + (size_t)naa * naa * ias,
         yt + (size_t)ldy * naa * ias, ldy, blockIdx.x * BM, blockIdx.y * BN, fperm);
}

template <int BM, int BN, int BK, int NST, int WR, int WC, bool G3 = false>
int launch_ystage_dma_t(const exc_dev &d, int ngp, const cplx *apwalm, int ldg, cplx *yt, int ldy, hipStream_t s) {
  dim3 grid((ngp + BM - 1) / BM, (d.naa + BN - 1) / BN, d.natm);
  hipLaunchKernelGGL((ystage_dma_kernel<BM, BN, BK, NST, WR, WC, G3>), grid, dim3(64 * WR * WC), 0, s, ngp, d.naa, apwalm,
                     ldg, d.haa, yt, ldy, d.fperm);
  EXC_CHECK(hipGetLastError());
  return EXC_OK;
}

// one (tile, H-or-O) item per workgroup, LDS filled by DMA
template <int BM, int BK, int NST, int WR, int WC, bool G3 = false>
__global__ __launch_bounds__(64 * WR * WC) void mt_dma_kernel(int ngp, int ktot, int ntile,
                                                              const cplx *__restrict__ apwalm, int ldg,
                                                              const cplx *__restrict__ yt, int ldy,
                                                              const int *__restrict__ igpig,
                                                              const double *__restrict__ vgpc, GTab gt,
                                                              cplx *__restrict__ h, cplx *__restrict__ o, int ldh,
                                                              const int *__restrict__ kperm, int nsel, int which0) {
  using T = MtTileDMA<BM, BK, NST, WR, WC, G3>;
  using St = typename T::Stage;
  __shared__ typename T::Tables tb;
  EXC_DMA_STAGES(St, NST)
  const int ntri = ntile * (ntile + 1) / 2;
  const int item = blockIdx.x;
  if (item >= nsel * ntri) return;
  int ti, tj;
  tri_decode(item / nsel, ti, tj);
  T::run(tb, EXC_DMA_STAGE_ARGS(St), ngp, ktot, apwalm, ldg, yt, ldy, igpig, vgpc, gt, h, o, ldh, kperm, ti, tj,
         nsel == 2 ? (item & 1) : which0);
}

template <int BM, int BK, int NST, int WR, int WC, bool G3 = false>
int launch_mt_dma_t(const exc_dev &d, int ngp, const cplx *apwalm, int ldg, const cplx *yt, int ldy, const int *igpig,
                    const double *vgpc, cplx *h, cplx *o, int ldh, hipStream_t s) {
  const int ntile = (ngp + BM - 1) / BM;
  const int nsel = sel_count(h, o);
  const int nitem = nsel * ntile * (ntile + 1) / 2;
  hipLaunchKernelGGL((mt_dma_kernel<BM, BK, NST, WR, WC, G3>), dim3(nitem), dim3(64 * WR * WC), 0, s, ngp,
                     d.natm * d.naa, ntile, apwalm, ldg, yt, ldy, igpig, vgpc, make_gtab(d), h, o, ldh, d.kperm, nsel,
                     h ? 0 : 1);
  EXC_CHECK(hipGetLastError());
  return EXC_OK;
}

}  // namespace

// Y-stage variants (the same shapes as the grouped Y stage's): 0: 64 x 128 tile, BK 8, Gauss's
// three real products per complex product; 1: the same with BK 16; 2: 64 x 128, BK 8, four
// products; 3: 64 x 64, BK 16, four DMA stages, four products (the per-k default); 4: 64 x 64,
// BK 8, three products (narrow column tiles: more of the Gram form's triangular table skipped)
int exc_launch_ystage(const exc_dev &d, int ngp, const cplx *apwalm, int ldg, cplx *yt, int ldy, int variant,
                      hipStream_t s) {
  if (ngp <= 0) return EXC_OK;
  switch (variant) {
    case 0: return launch_ystage_dma_t<64, 128, 8, 2, 2, 4, true>(d, ngp, apwalm, ldg, yt, ldy, s);
    case 1: return launch_ystage_dma_t<64, 128, 16, 2, 2, 4, true>(d, ngp, apwalm, ldg, yt, ldy, s);
    case 2: return launch_ystage_dma_t<64, 128, 8, 2, 2, 4>(d, ngp, apwalm, ldg, yt, ldy, s);
    case 3: return launch_ystage_dma_t<64, 64, 16, 2, 2, 4>(d, ngp, apwalm, ldg, yt, ldy, s);
    case 4: return launch_ystage_dma_t<64, 64, 8, 2, 2, 4, true>(d, ngp, apwalm, ldg, yt, ldy, s);
    default: return EXC_ERR_ARG;
  }
}

// MT variants: 64 x 64 tile, 8 waves; 0: BK 16, two DMA stages, Gauss's three products (the
// default); 1: the same with four products; 2: BK 8, four stages, four products
int exc_launch_mt(const exc_dev &d, int ngp, const cplx *apwalm, int ldg, const cplx *yt, int ldy, const int *igpig,
                  const double *vgpc, cplx *h, cplx *o, int ldh, int variant, hipStream_t s) {
  if (ngp <= 0 || (!h && !o)) return EXC_OK;
  switch (variant) {
    case 0: return launch_mt_dma_t<64, 16, 2, 2, 4, true>(d, ngp, apwalm, ldg, yt, ldy, igpig, vgpc, h, o, ldh, s);
    case 1: return launch_mt_dma_t<64, 16, 2, 2, 4>(d, ngp, apwalm, ldg, yt, ldy, igpig, vgpc, h, o, ldh, s);
    case 2: return launch_mt_dma_t<64, 8, 4, 2, 4>(d, ngp, apwalm, ldg, yt, ldy, igpig, vgpc, h, o, ldh, s);
    default: return EXC_ERR_ARG;
  }
}
