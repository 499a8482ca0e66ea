// Shared definitions of the gfx950 LAPW H+O setup library.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>
#include <stdio.h>

#include "../../include/exc_lapw.h"

#define EXC_CHECK(call)                                                              \
  do {                                                                               \
    hipError_t e_ = (call);                                                          \
    if (e_ != hipSuccess) {                                                          \
      fprintf(stderr, "exc_lapw: HIP error %s at %s:%d\n", hipGetErrorString(e_),   \
              __FILE__, __LINE__);                                                   \
      return EXC_ERR_HIP;                                                            \
    }                                                                                \
  } while (0)

typedef double __attribute__((ext_vector_type(4))) d4_t;

struct cplx {
  double re, im;
};

__device__ __forceinline__ cplx cmul(cplx a, cplx b) {
  return {a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re};
}
__device__ __forceinline__ cplx cconj(cplx a) { return {a.re, -a.im}; }

// Per-local-orbital description (one entry per lo column, 0..nlotot-1), built on the
// host from idxlo/lorbl (genidxlo.f90:24-43).
struct lo_desc {
  int ias;     // atom
  int row;     // row inside the atom's loa/lolo block (0..losize-1)
  int ilo;     // local-orbital index inside the species (0-based)
  int l, m;    // angular momentum of the function
  int lm;      // l(l+1)+m, 0-based
  int first;   // global lo index of the atom's first lo function
  int pad;
};

// Device-resident SCF data of one context.
struct exc_dev {
  int natm, nspecies, lmaxapw, lmaxmat, apwordmax, lmmaxapw, naa;
  int nlotot, maxnlo, nlomax;
  int ngrtot, ngvec;
  int box_lo[3], box_n[3];
  double omega;
  int *spidx;          // [natm]
  double *rmt;         // [nspecies]
  double *atposc;      // [3*natm]
  double *dinv;        // [natm][lmax+1][2][2]: inverse of the 2x2 matching matrix
  int *dsing;          // [natm*(lmax+1)]: 1 where the matching matrix is singular
  cplx *haa;           // [natm][naa][naa] column-major per atom
  cplx *loa;           // [natm][naa][maxnlo]
  cplx *lolo;          // [natm][maxnlo][maxnlo]
  double *oalo;        // [natm][nlomax][apwordmax]
  double *ololo;       // [natm][nlomax][nlomax]
  lo_desc *lod;        // [nlotot]
  int *ivg;            // [3*ngrtot]
  int *ivgig;          // box
  cplx *cfunig;        // [ngrtot]
  cplx *veffig;        // [ngvec]
  cplx *meffig;        // [ngvec]: the kinetic term's factor with ValenceRelativity /= 'none'
  int zora;            // 1: H's interstitial kinetic term takes meffig instead of cfunig
  // ValenceRelativity 'iora*' overlap terms (exc_ctx_set_iora); iora bits: 1 the APW-APW
  // factor I + h1aa, 2 h1loa / h1lolo, 4 the interstitial term with m2effig
  int iora;
  int ytri;            // grouped Y stage: haa is the Gram form's triangular table (tile sums start at the column)
  double *olpl;        // [natm][lmaxmat+1][3]: lower Cholesky factor (L00, L10, L11) of I + h1aa(:,:,l,a)
  double *h1loa;       // [natm][nlomax][apwordmax], added to oalo
  double *h1lolo;      // [natm][nlomax][nlomax], added to ololo
  cplx *m2effig;       // [ngvec]
  double *ffacg;       // [nspecies][ngvec]
  int *kperm;          // [natm*naa]: summation order of the MT product (column of apwalm)
  int *fperm;          // [naa]: summation order inside one atom (Y stage)
  // muffin-tin density matrix (gendmatmt.f90)
  int dm_maxlo;        // largest number of lo functions of an atom (mt_dm%maxnlo)
  int *dm_lo;          // [natm][2]: global index of the atom's first lo function, their number
  int *dm_order;       // [naa]: apwalm storage column (io + apwordmax lm) of density-matrix row f
  unsigned *miss;      // [1]: basis differences G_i - G_j met outside the first ngvec G-vectors
};

// Device-side description of one k-point of a grouped launch.
struct kdesc {
  int ngp, ldg, ldh, ldy;
  const double *gpc, *tpgpc;
  const cplx *sfacgp;
  cplx *apwalm;
  const int *igpig;
  const double *vgpc;
  cplx *yt;
  cplx *h, *o;
  // split-BF16 operands (exc_setup_batched_dev with EXC_BATCH_SPLIT_BF16)
  unsigned short *sa, *sy;   // slice planes of A and of Yt, layout [row][atom][k/16][plane][16]
  int *ea, *ey;              // per-(row, atom) exponents, [row*natm + atom]
  // radial factors of apwalm for the overlap by the addition theorem (EXC_BATCH_OLP_LEGENDRE),
  // written by the grouped match: [(atom*ngp + row)*2(lmaxmat+1) + 2l + io]
  double *rf;
  double *rft;               // the same per (atom, l) with rows fastest, times sqrt(gamma_l) (overlap variant 2)
  // modular split (EXC_BATCH_SPLIT_CRT): rows padded to cnp (a multiple of 256), int8 residue
  // planes of A and Yt, layout [modulus][plane][K/16][cnp][16], and per-row scale exponents
  int cnp, pad_;
  void *qa, *qy;
  int *xa, *xy;
  // squared row norms of A per atom, written by the grouped match, [atom][cnp]: A's
  // exponents xa are summed from them
  double *pa2;
};

// Work item of the grouped muffin-tin launch: (k, tile row, tile column, H=0 / O=1).
struct mt_item {
  int k, ti, tj, which;
};

// Launchers (one translation unit each).
int exc_launch_match(const exc_dev &d, int ngp, const double *gpc, const double *tpgpc,
                     const cplx *sfacgp, int ldsf, cplx *apwalm, int ldg, hipStream_t s);
int exc_launch_ystage(const exc_dev &d, int ngp, const cplx *apwalm, int ldg, cplx *yt, int ldy,
                      int variant, hipStream_t s);
int exc_launch_mt(const exc_dev &d, int ngp, const cplx *apwalm, int ldg, const cplx *yt, int ldy,
                  const int *igpig, const double *vgpc, cplx *h, cplx *o, int ldh, int variant,
                  hipStream_t s);
constexpr int kYstageVariants = 5;   // per-k and grouped Y stage (the same shapes)
constexpr int kMtVariants = 3;       // per-k MT
constexpr int kBmtVariants = 4;      // grouped MT
// Grouped launches over nk k-points described by ks (device memory); maxngp bounds the grids.
int exc_launch_bmatch(const exc_dev &d, int nk, int maxngp, const kdesc *ks, hipStream_t s);
int exc_launch_bystage(const exc_dev &d, int nk, int maxngp, const kdesc *ks, int variant, hipStream_t s);
int exc_bmt_tile(int variant);   // square tile edge of a grouped-MT variant
int exc_launch_bmt(const exc_dev &d, const kdesc *ks, const mt_item *items, int nitems, int variant, hipStream_t s);
int exc_launch_blo(const exc_dev &d, int nk, int maxngp, const kdesc *ks, hipStream_t s);
// adjoint (vector-Jacobian product) of the setup for rank-nst cotangents
size_t exc_vjp_work_elems(const exc_dev &d, int ngp, int nst);
int exc_launch_vjp(const exc_dev &d, int ngp, int nst, const cplx *apwalm, int ldg, const int *igpig,
                   const double *vgpc, const cplx *sfacgp, const cplx *evec, int ldc, const double *wh,
                   const double *wo, cplx *work, double *dedtau, double *dedv, hipStream_t s);   // uv in work first
int exc_launch_vjp_qscale(int ngp, int nst, const cplx *evec, int ldc, const double *vgpc, cplx *cs, hipStream_t s);
// split-BF16 muffin-tin stage
constexpr int kSplitKB = 256;       // APW functions per atom block, padded (naa <= 256)
constexpr int kSplitPlanes = 6;     // 3 slices x (re, im)
constexpr int kBsplitVariants = 4;
constexpr int kEigenVariants = 9;   // generalized eigensolver paths (exc_eigen_dev)
size_t exc_split_plane_elems(int ngp, int natm);   // bf16 elements of one operand's planes
int exc_launch_split(const exc_dev &d, int nk, int maxngp, const kdesc *ks, hipStream_t s);
int exc_bsplit_tile(int variant);
int exc_bsplit_max_blocks(int variant);
int exc_olp_tile();   // tile edge of the addition-theorem overlap kernel
constexpr int kOlpVariants = 4;
int exc_launch_olp(const exc_dev &d, int nk, int maxngp, const kdesc *ks, const mt_item *items, int nitems, int variant,
                   hipStream_t s);
int exc_launch_bsplit(const exc_dev &d, const kdesc *ks, const mt_item *items, int nitems, int variant, int nblocks,
                      hipStream_t s);
// modular split of the muffin-tin contraction (k_crt.hip)
constexpr int kMaxMod = 14;          // moduli available: 241, 233, 229, 221, 205, 197, 193, ...
constexpr int kCrtLoDigits = 6;      // mixed-radix digits of the low part (their product < 2^53)
constexpr int kCrtMaxUsed = 12;      // beyond 12 moduli beta > 44: re -+ iota im of the row
                                     // integers no longer exact in FP64
constexpr int kCrtVariants = 11;     // tile / staging variants of the residue GEMM
struct CrtTab {                      // passed by value to the kernels (kernel-argument segment)
  int n, beta;
  int m[kMaxMod];                    // odd, pairwise coprime, every prime factor 1 mod 4
  int iota[kMaxMod];                 // a square root of -1 modulo m
  double dminv[kMaxMod];
};
void exc_crt_tables(int n, CrtTab *t);
inline int exc_crt_kpad(int ktot) { return (ktot + 63) / 64 * 64; }
int exc_crt_shape(int variant, int *bt, int *wr, int *wc);   // tile edge and wave grid of a GEMM variant
int exc_launch_crt_prep(const exc_dev &d, int nk, int maxnp, const kdesc *ks, const CrtTab &t, bool a_parts,
                        bool gram, hipStream_t s);
int exc_launch_crt_gemm(const exc_dev &d, const kdesc *ks, const mt_item *items, const int2 *work, int nwork,
                        const CrtTab &t, signed char *res, int variant, hipStream_t s);
int exc_launch_crt_merge(const exc_dev &d, const kdesc *ks, const mt_item *items, int nitems, const CrtTab &t,
                         const signed char *res, int variant, const double *gsig, hipStream_t s);
int exc_launch_crt_gram_table(const exc_dev &d, const cplx *U, const double *lam, double sigma, cplx *T,
                              hipStream_t s);
int exc_launch_crt_gram_perm(const exc_dev &d, double sigma, cplx *Mp, hipStream_t s);
int exc_launch_crt_gram_tri(const exc_dev &d, const cplx *L, cplx *T, hipStream_t s);
// generalized eigensolve helpers (k_eigen.hip)
struct eig_kdesc {   // one k-point of a grouped solve: nmat n, device arrays
  const cplx *h, *o;
  double *eval;
  cplx *evec;
  int n, ldh, ldv, pad;
};
int exc_launch_eig_hermfill(cplx *h, int n, int ldh, hipStream_t s);
int exc_launch_eig_pad(const eig_kdesc *ks, int nb, int n, cplx *ph, cplx *po, double sigma, hipStream_t s);
int exc_launch_eig_r2c(const double *zr, int m, int ncols, int ldr, cplx *zc, int ldc, hipStream_t s);
int exc_launch_eig_vblock(const cplx *A, int nb, int n, int i0, int ib, size_t sv, cplx *V, hipStream_t s);
int exc_launch_eig_tblock(const eig_kdesc *ks, const cplx *G, const cplx *tau, int nb, int n, int i0, int ib,
                          size_t sgt, cplx *T, hipStream_t s);
int exc_launch_eig_sort_r2c(const double *zr, int m, int nev, int ldr, const int *nfound, const double *w, cplx *x,
                            int ldx, double *eval, hipStream_t s);
int exc_launch_eig_copyout(const eig_kdesc *ks, int nb, int n, int nev, const cplx *x, hipStream_t s);
// direct-route muffin-tin density (k_rhovalk.hip)
size_t exc_rhovalk_work_bytes(const exc_dev &d, int lmaxvr, int nlomax, int nst);
int exc_launch_rhovalk(const exc_dev &d, int lmaxvr, int lradstep, int nrmtmax, int nlomax, const int *nrmt,
                       const int *nlorb, const int *lorbl, const double *apwfr, const double *lofr, const cplx *w,
                       int ldw, int ngp, const cplx *evec, int ldv, int nst, const double *occ, double wk,
                       double eps, const cplx *zb, const double *rf, void *work, double *rhomt, hipStream_t s);
// muffin-tin density matrix helpers (k_dmat.hip)
int exc_launch_dmat_pack(const exc_dev &d, int ws, int nst, const cplx *w, const cplx *evec, int ldv, int ngp,
                         cplx *x, int ldx, size_t xs, int col0, hipStream_t s);
int exc_launch_dmat_scale(const exc_dev &d, int ws, int nst, const cplx *x, cplx *x2, int ldx, size_t xs, int col0,
                          const double *occ, double wk, hipStream_t s);
int exc_launch_dmat_reduce(cplx *dm, const cplx *per, int np, size_t pstride, size_t poff, size_t n,
                           hipStream_t s);
// muffin-tin integrals (k_hmlint.hip)
void exc_hml_spline_weights(int n, const double *x, double *wt);
void exc_hml_gaunt_table(int lmaxmat, int lmaxvr, cplx *g);
int exc_launch_genrhomt(const exc_dev &d, int lmv, int nrmtmax, const int *nrmt, const int *nlorb, const int *lorbl,
                        const double *apwfr, const double *lofr, const cplx *gnt, const cplx *dm, int ws, double *fbuf,
                        double *rhomt, hipStream_t s);
int exc_launch_hmlint(const exc_dev &d, int lmv, int nrmtmax, const int *nrmt, const int *nlorb, const int *lorbl,
                      const double *wt, const double *spr, const double *apwfr, const double *lofr,
                      const double *veffmt, double *rad, const cplx *gnt, hipStream_t s);
// Theta(G), Theta(r), V(G) (k_scf.hip)
int exc_launch_cfunig(const exc_dev &d, const double *bvec, int ngrtot, cplx *cf, cplx *z, hipStream_t s);
int exc_launch_cfdamp(const exc_dev &d, const double *bvec, int ngrtot, double t1, cplx *cf, hipStream_t s);
// G+k sets (k_gvec.hip): kv holds (vpl, vpc) per k-point, 6 doubles each, in device memory
int exc_launch_gengpvec(const exc_dev &d, int nk, const double *bvec, const double *kv, double gkmax, int ngkmax,
                        int *ngp, int *igpig, double *vgpl, double *vgpc, double *gpc, double *tpgpc, cplx *sfacgp,
                        hipStream_t s);
int exc_launch_gensfacgp(const exc_dev &d, int ngp, const double *vgpc, int ld, cplx *sfacgp, hipStream_t s);
int exc_launch_scf_real(int n, const cplx *z, double *out, hipStream_t s);
int exc_launch_scf_product(int n, const double *a, const double *b, cplx *z, hipStream_t s);
int exc_launch_scf_gather(const exc_dev &d, int ngvec, double scale, const cplx *z, cplx *out, hipStream_t s);
int exc_launch_rhoir_scatter(const exc_dev &d, int ngp, const int *igpig, const cplx *evec, int ldv, int j0, int nb,
                             int N, cplx *z, hipStream_t s);
int exc_launch_rhoir_accum(int N, int nb, const cplx *z, const double *occ, int j0, double wk, double eps,
                           double omega, double *rho, hipStream_t s);
int exc_launch_lo(const exc_dev &d, int ngp, const cplx *apwalm, int ldg, cplx *h, cplx *o, int ldh,
                  hipStream_t s);
int exc_launch_iora_apw(const exc_dev &d, int ngp, const cplx *apwalm, int ldg, cplx *out, int ldo, hipStream_t s);
