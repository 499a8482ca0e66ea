/*
 * exc_lapw.h — C-ABI of the gfx950 (MI355X) LAPW Hamiltonian/overlap setup.
 *
 * The functions below replace, one for one, the reference routines of exciting's
 * per-k first-variational setup (paths relative to the reference's src/):
 *
 *   exc_match                      match.f90:13            Subroutine match(ngp,gpc,tpgpc,sfacgp,apwalm)
 *   exc_hamiltonandoverlapsetup    src_eigensystem/hamiltonandoverlapsetup.f90:8
 *                                  (== hamiltonsetup.f90:8 followed by overlapsetup.f90:8,
 *                                   including hmlistln.f90:13, olpistln.f90:13 and
 *                                   hamiltonoverlapocopy_UL.f90:6)
 *   exc_hamiltonsetup              src_eigensystem/hamiltonsetup.f90:8
 *   exc_overlapsetup               src_eigensystem/overlapsetup.f90:8
 *
 * Module state that the reference routines read implicitly (mt_hscf, oalo, ololo, idxlo,
 * ivg, ivgig, cfunig, veffig, apwfr-derived boundary values ...) is uploaded once per SCF
 * iteration into a context with the exc_ctx_set_* calls.
 *
 * Conventions
 *  - Arrays have the reference's Fortran shapes and column-major layout; integer index
 *    arrays hold the reference's 1-based values.  Complex arrays are interleaved
 *    (re, im) doubles, i.e. Fortran complex(8) / C double _Complex.
 *  - Functions whose name ends in _dev take device pointers (memory of the GPU the
 *    context was created on) and run on the context's stream asynchronously; all other
 *    pointers are host memory and the call returns when the results are in host memory.
 *  - H and O are fully overwritten ("=" semantics): the reference adds to a matrix that
 *    newsystem (modfvsystem.f90:92-101) has just zeroed, so the results are the same.
 *    Both triangles are written; the strict lower triangle is conj(upper) exactly, as
 *    after hamiltonoverlapocopy_UL.
 *  - Return value: 0 = success; > 0 mirrors a LAPACK info (exc_match: 1 + the 0-based
 *    index (l + (lmaxapw+1)*ias) of the first singular matching matrix, like zgesv's
 *    info in match.f90:163-175 — the coefficients are still written, as the reference
 *    continues: for a singular block they are the unsolved right-hand sides, which is what
 *    zgesv leaves behind with info > 0); < 0 = EXC_ERR_*.
 *  - One context serves one host thread at a time (the reference calls the setup from one
 *    thread per MPI rank, scf_cycle.f90:241-276).  Calls on one context are serialised on
 *    its stream.
 */
#ifndef EXC_LAPW_H
#define EXC_LAPW_H

#ifdef __cplusplus
extern "C" {
#endif

#define EXC_OK 0
#define EXC_ERR_ARG (-1)
#define EXC_ERR_HIP (-2)
#define EXC_ERR_UNSUPPORTED (-3)
#define EXC_ERR_STATE (-4)

typedef struct exc_ctx exc_ctx;

/* Sizes of the module state (mod_atoms, mod_APW_LO, mod_muffin_tin, mod_Gvector). */
typedef struct exc_dims {
  int natmtot;    /* number of atoms (mod_atoms natmtot)                              */
  int nspecies;   /* number of species                                               */
  int lmaxapw;    /* input%groundstate%lmaxapw (<= 14)                               */
  int lmaxmat;    /* input%groundstate%lmaxmat (must equal lmaxapw here)             */
  int apwordmax;  /* APW order; must be 2 with apword(l,is) = 2 for every l          */
  int nlomax;     /* max nlorb over species (first dimension of lorbl, oalo, ololo)   */
  int lolmmax;    /* first dimension of idxlo, (lolmax+1)^2                           */
  int maxnlo;     /* mt_hscf%maxnlo (leading dimension of loa and lolo)              */
  int ngrtot;     /* number of G vectors in ivg / cfunig                             */
  int ngvec;      /* G vectors with |G| < gmaxvr (length of veffig)                  */
  int intgv[6];   /* intgv(3,2) in Fortran order: lower bounds (1..3), upper (1..3)   */
  double omega;   /* unit-cell volume                                                */
} exc_dims;

/* Create a context on the current HIP device.  Replaces nothing in the reference: the
 * reference keeps this state in Fortran modules (mod_atoms, mod_muffin_tin, ...). */
int exc_ctx_create(const exc_dims *dims, exc_ctx **out);
int exc_ctx_destroy(exc_ctx *ctx);

/* Run the context's work on this HIP stream (hipStream_t passed as void*; NULL = the
 * null stream).  The library never creates streams of non-default priority. */
int exc_ctx_set_stream(exc_ctx *ctx, void *stream);

/* Atoms: natoms(nspecies), rmt(nspecies), atposc(3, natmtot) Cartesian, atoms in the
 * reference's idxas order (species-major).  mod_atoms / mod_muffin_tin state. */
int exc_ctx_set_atoms(exc_ctx *ctx, const int *natoms, const double *rmt, const double *atposc);

/* Local orbitals: nlorb(nspecies), lorbl(nlomax, nspecies), idxlo(lolmmax, nlomax, natmtot)
 * as built by genidxlo.f90:24-43.  Returns EXC_ERR_ARG if idxlo is not the running
 * species->atom->ilo->m numbering the setup routines assume. */
int exc_ctx_set_lo(exc_ctx *ctx, const int *nlorb, const int *lorbl, const int *idxlo);

/* Radial boundary values for match: dmat(apwordmax, apwordmax, 0:lmaxapw, natmtot),
 * dmat(i, j, l, ias) = d^(i-1) u_(j,l) / dr^(i-1) at r = R, i.e. the zd matrix that
 * match.f90:141-148 builds with polynom from apwfr.  Returns > 0 (as exc_match would) if
 * a matching matrix is singular. */
int exc_ctx_set_radial(exc_ctx *ctx, const double *dmat);
/* The same from the radial functions themselves, as match.f90:139-148 builds zd: the value
 * and first derivative at R of the cubic through the last four mesh points (polynom.f90 with
 * np = max(apwordmax+1, 4) = 4, match.f90:85) of apwfr(nrmtmax, 2, apwordmax, 0:lmaxapw,
 * natmtot) (only the u part is read) on the meshes spr(nrmtmax, nspecies), nrmt(nspecies). */
int exc_ctx_set_radial_fr(exc_ctx *ctx, int nrmtmax, const int *nrmt, const double *spr,
                          const double *apwfr);

/* Muffin-tin matrices of this SCF iteration (mod_eigensystem.F90:60-80, olprad.f90):
 *   haa(maxaa, maxaa, natmtot)   complex  mt_hscf%main%aa
 *   loa(maxnlo, maxaa, natmtot)  complex  mt_hscf%main%loa
 *   lolo(maxnlo, maxnlo, natmtot) complex mt_hscf%main%lolo
 *   oalo(apwordmax, nlomax, natmtot) real, ololo(nlomax, nlomax, natmtot) real
 * with maxaa = apwordmax*(lmaxmat+1)^2.  The h1* (IORA) terms are zero in this setup. */
/* The same matrices computed on the GPU from this SCF iteration's radial data, the
 * counterpart of hmlint (hmlint.f90:14, mt_kin.f90 + mt_pot.f90 with ValenceRelativity
 * 'none', spinless) and olprad (olprad.f90:13): every radial integral is the reference's
 * cubic-spline integral (fderiv.f90 m = -1, spline.f90) on the species' mesh, the angular
 * part the Gaunt coefficients int conj(Y_lm1) R_lm2 Y_lm3 (gauntyry.f90).  Inputs (host):
 *   nrmt(nspecies), spr(nrmtmax, nspecies)                      the radial meshes
 *   apwfr(nrmtmax, 2, apwordmax, 0:lmaxapw, natmtot)              u and du/dr of the APW functions
 *   lofr(nrmtmax, 2, nlomax, natmtot)                              the same of the local orbitals
 *   veffmt((lmaxvr+1)^2, nrmtmax, natmtot)                         V in real spherical harmonics
 * The results replace the context's muffin-tin matrices (as exc_ctx_set_mt would) and are
 * also copied to the host arrays given (each may be NULL), in exc_ctx_set_mt's shapes.
 * Needs exc_ctx_set_atoms and exc_ctx_set_lo first; lo angular momenta <= lmaxmat. */
int exc_ctx_hmlint(exc_ctx *ctx, int lmaxvr, int nrmtmax, const int *nrmt, const double *spr,
                   const double *apwfr, const double *lofr, const double *veffmt, double *haa,
                   double *loa, double *lolo, double *oalo, double *ololo);
int exc_ctx_set_mt(exc_ctx *ctx, const double *haa, const double *loa, const double *lolo,
                   const double *oalo, const double *ololo);

/* G-vector tables (mod_Gvector / mod_potential_and_density):
 *   ivg(3, ngrtot) integer, ivgig(intgv(1,1):intgv(1,2), ..2.., ..3..) integer (1-based),
 *   cfunig(ngrtot) complex (gencfun.f90), veffig(ngvec) complex (genveffig.f90),
 *   ffacg(ngvec, nspecies) real (genffacg.f90; may be NULL, needed only by the VJP). */
int exc_ctx_set_gvec(exc_ctx *ctx, const int *ivg, const int *ivgig, const double *cfunig,
                     const double *veffig, const double *ffacg);

/* ValenceRelativity /= 'none' in the interstitial (hmlistln.f90:61-72): meffig(ngvec) complex,
 * the Fourier components of the effective-mass factor (mod_potential_and_density), takes
 * cfunig's place in H's kinetic term, H(i,j) += V(G_i - G_j) + (1/2)(G_i+k).(G_j+k)
 * meffig(G_i - G_j); O keeps cfunig (olpistln.f90:69-82; the 'iora*' overlap terms are
 * exc_ctx_set_iora's).  NULL restores the 'none' form.  The muffin-tin matrices are the
 * caller's (exc_ctx_set_mt: exc_ctx_hmlint forms the 'none' ones only), and exc_ho_vjp
 * returns EXC_ERR_UNSUPPORTED while meffig is set. */
int exc_ctx_set_meffig(exc_ctx *ctx, const double *meffig);

/* ValenceRelativity 'iora*' overlap terms (seceqnfv.f90:74 h1on; overlapsetup.f90:8):
 *   h1aa(apwordmax, apwordmax, 0:lmaxapw, natmtot) real (olprad.f90:75): the APW-APW block
 *     sum_a conj(A_a) (I + h1aa)^T A_a^T over l <= lmaxmat (overlapsetup.f90:81-94);
 *   h1loa(apwordmax, nlomax, natmtot), h1lolo(nlomax, nlomax, natmtot) real: added to oalo
 *     and ololo in the APW-lo and lo-lo blocks (overlapsetup.f90:126, 144);
 *   m2effig(ngvec) complex (genmeffig.f90): O's interstitial element becomes
 *     1/2 a2 (G_i+k).(G_j+k) m2effig(G_i - G_j) + cfunig(G_i - G_j), a2 = alpha^2 / 2
 *     (olpistln.f90:53-62).
 * Each may be NULL (that term off); all NULL restores the 'none' overlap.  h1aa must be
 * symmetric in (io1, io2) (to 1e-12 relative; EXC_ERR_ARG) and I + h1aa positive definite
 * (EXC_ERR_UNSUPPORTED otherwise): the library factors it as L L^T.  H is unchanged (its
 * 'iora*' muffin-tin and kinetic parts are the caller's haa and exc_ctx_set_meffig).  With
 * h1aa set, exc_setup_batched_dev needs EXC_BATCH_OLP_LEGENDRE and the no-allocation entry
 * returns EXC_ERR_UNSUPPORTED; with any term set, the split flags need
 * EXC_BATCH_OLP_LEGENDRE and exc_ho_vjp returns EXC_ERR_UNSUPPORTED. */
int exc_ctx_set_iora(exc_ctx *ctx, const double *h1aa, const double *h1loa, const double *h1lolo,
                     const double *m2effig);

/* match.f90:13 — APW matching coefficients of one k-point.
 *   gpc(ngkmax), tpgpc(2, ngkmax), sfacgp(ngkmax, natmtot) complex,
 *   apwalm(ngkmax, apwordmax, lmmaxapw, natmtot) complex (output; rows > ngp untouched). */
int exc_match(exc_ctx *ctx, int ngp, int ngkmax, const double *gpc, const double *tpgpc,
              const double *sfacgp, double *apwalm);
int exc_match_dev(exc_ctx *ctx, int ngp, int ngkmax, const double *gpc, const double *tpgpc,
                  const double *sfacgp, double *apwalm);

/* hamiltonandoverlapsetup.f90:8 (= hamiltonsetup + overlapsetup) — H and O of one k-point.
 *   apwalm(ngkmax, apwordmax, lmmaxapw, natmtot), igpig(ngkmax), vgpc(3, ngkmax),
 *   H(ldh, nmat), O(ldh, nmat) complex with nmat = ngp + nlotot <= ldh.
 * Host-array entries: apwalm = NULL takes the coefficients of the context's last exc_match
 * call, still on the device (same ngp and ngkmax; EXC_ERR_STATE otherwise), so match and the
 * setup of one k-point move apwalm over PCIe once; the caller must not have changed its copy. */
int exc_hamiltonandoverlapsetup(exc_ctx *ctx, int ngp, int ngkmax, const double *apwalm,
                                const int *igpig, const double *vgpc, int ldh, double *h,
                                double *o);
int exc_hamiltonandoverlapsetup_dev(exc_ctx *ctx, int ngp, int ngkmax, const double *apwalm,
                                    const int *igpig, const double *vgpc, int ldh, double *h,
                                    double *o);
/* No counterpart in the reference: exc_hamiltonandoverlapsetup_dev for callers that cannot
 * allocate (code under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc).  The same kernels, variants and results; the Y stage's workspace is a fixed
 * device array of the library (1 GiB on each GPU, held until the process ends) instead of
 * the context's grow-only buffer, and no stage timer runs.  ngp * natm * naa above 64 Mi
 * (complex elements of that workspace) gives EXC_ERR_ARG before anything is queued.  Every
 * call on one GPU shares the array: calls that overlap on one GPU, from any context or
 * stream, race. */
int exc_hamiltonandoverlapsetup_noalloc_dev(exc_ctx *ctx, int ngp, int ngkmax,
                                            const double *apwalm, const int *igpig,
                                            const double *vgpc, int ldh, double *h, double *o);
/* hamiltonsetup.f90:8, Subroutine hamiltonsetup(system, ngp, apwalm, igpig, vgpc) — H alone
 * (muffin-tin APW-APW, APW-lo and lo-lo blocks, hmlistln.f90:13) — and overlapsetup.f90:8,
 * Subroutine overlapsetup(system, ngp, apwalm, igpig, vgpc) — O alone (olpistln.f90:13), the
 * two calls seceqnfv.f90:76-77 makes one after the other.  Same arguments as above with the
 * one output matrix; each writes its matrix completely, both triangles (what the reference
 * holds after overlapsetup's hamiltonoverlapocopy_UL, overlapsetup.f90:164-166). */
int exc_hamiltonsetup(exc_ctx *ctx, int ngp, int ngkmax, const double *apwalm, const int *igpig,
                      const double *vgpc, int ldh, double *h);
int exc_hamiltonsetup_dev(exc_ctx *ctx, int ngp, int ngkmax, const double *apwalm,
                          const int *igpig, const double *vgpc, int ldh, double *h);
int exc_overlapsetup(exc_ctx *ctx, int ngp, int ngkmax, const double *apwalm, const int *igpig,
                     const double *vgpc, int ldh, double *o);
int exc_overlapsetup_dev(exc_ctx *ctx, int ngp, int ngkmax, const double *apwalm,
                         const int *igpig, const double *vgpc, int ldh, double *o);

/* ---- grouped setup of many k-points (no reference counterpart: the reference loops over
 * k in scf_cycle.f90:241-276 and calls match + hamiltonsetup + overlapsetup per k). ----
 * One entry per k-point, all pointers in device memory, layouts as for the per-k calls:
 *   gpc(ngkmax), tpgpc(2,ngkmax), sfacgp(ngkmax,natmtot)  match inputs (used with EXC_BATCH_MATCH)
 *   igpig(ngkmax), vgpc(3,ngkmax)                          basis of the k-point
 *   apwalm(ngkmax,apwordmax,lmmaxapw,natmtot)              output of match / input of the setup
 *   h(ldh,nmat), o(ldh,nmat)                               outputs, nmat = ngp + nlotot <= ldh */
typedef struct exc_kpoint_dev {
  int ngp, ngkmax, ldh, reserved;
  const double *gpc, *tpgpc, *sfacgp;
  const int *igpig;
  const double *vgpc;
  double *apwalm;
  double *h, *o;
} exc_kpoint_dev;

#define EXC_BATCH_MATCH 1  /* compute apwalm of every k first (else apwalm is an input)        */
#define EXC_BATCH_GRAPH 2  /* record the launches in a hipGraph and replay it on later calls   */
                           /* with an identical batch (same pointers, sizes, flags, variants); */
                           /* needs a context stream other than the null stream               */
#define EXC_BATCH_SPLIT_BF16 4 /* muffin-tin contraction from split-BF16 slices with FP32 MFMA   */
                               /* accumulation and per-(row, atom) power-of-two scaling (3 slices,*/
                               /* 6 slice pairs); Y stage, interstitial, lo blocks stay FP64      */
#define EXC_BATCH_SPLIT_CRT 16  /* muffin-tin contraction from exact integer products: each row of  */
                                /* A and Yt scaled by a power of two to integers below 2^beta (row */
                                /* 2-norm), their residues modulo n pairwise coprime odd moduli    */
                                /* < 256 in which -1 is a square, so each complex product splits   */
                                /* into two real int8 MFMA products with int32 accumulation, the   */
                                /* results reconstructed by the Chinese remainder theorem and      */
                                /* rounded once to FP64; n from exc_ctx_set_moduli.  Y stage,      */
                                /* interstitial, lo blocks stay FP64.  With                        */
                                /* EXC_BATCH_OLP_LEGENDRE (and no 'iora*' h1aa) the Gram form:     */
                                /* E = A conj(U) (Lambda + sigma)^(1/2) from h_a = U Lambda U^H,   */
                                /* H_MT = conj(E) E^T - sigma O_APW, one operand to split          */
#define EXC_BATCH_OLP_LEGENDRE 8 /* needs EXC_BATCH_MATCH: the muffin-tin part of O from the      */
                                 /* radial factors of match by the spherical-harmonic addition   */
                                 /* theorem, sum_a conj(S_a) S_a' sum_l P_l(u.u') sum_io r r',   */
                                 /* FP64 on the vector units (the same matrix as conj(A) A^T,    */
                                 /* rounded differently); H is unchanged                         */

/* match (optional) + H and O of nk k-points, each with its own basis size, in one grouped
 * launch per stage on the context's stream; the muffin-tin stage is a persistent kernel over
 * a work list of (k, tile, H/O) items.  Results equal nk calls of the per-k entry points. */
int exc_setup_batched_dev(exc_ctx *ctx, int nk, const exc_kpoint_dev *ks, int flags);

/* Adjoint of the setup (forcek.f90:13, the IBS force term of one k-point).
 * For the cotangents D = C diag(wh) C^H of H and W = C diag(wo) C^H of O, with
 * C = evec(ldc, nst) (first-variational eigenvectors, ldc >= nmat), accumulates
 *   dedtau(3, natmtot) += d/dtau Re sum_ij (D_ji H_ij + W_ji O_ij)
 *   dedv(ngvec) complex += sum_{ij: G_i - G_j = g} D_ji  (holomorphic partial in V(g))
 * where H and O depend on the atomic positions through the structure factors in apwalm and
 * through Theta(G) (gencfun.f90), with V(G), h_a, loa, oalo held fixed as forcek does.
 * With wh = w_k*occ and wo = -w_k*occ*eval, dedtau is minus forcek's contribution to
 * forceibs.  sfacgp(ngkmax, natmtot) is the k-point's structure factors (gensfacgp.f90);
 * ffacg must have been given to exc_ctx_set_gvec.  All pointers are device memory. */
int exc_ho_vjp_dev(exc_ctx *ctx, int ngp, int ngkmax, const double *apwalm, const int *igpig,
                   const double *vgpc, const double *sfacgp, int nst, int ldc, const double *evec,
                   const double *wh, const double *wo, double *dedtau, double *dedv);
/* The same with host arrays (for the Fortran driver's forcek); dedtau and dedv are
 * accumulated in host memory.  dedv may be NULL. */
int exc_ho_vjp(exc_ctx *ctx, int ngp, int ngkmax, const double *apwalm, const int *igpig,
               const double *vgpc, const double *sfacgp, int nst, int ldc, const double *evec,
               const double *wh, const double *wo, double *dedtau, double *dedv);

/* Generalized Hermitian eigenproblem of one k-point after the setup, the counterpart of
 * solvewithlapack (modfvsystem.f90:585-720), which calls ZHEGVX(itype 1, 'V', range 'I',
 * 'U', il = 1, iu = nev) (:691-693): H c = e O c for the lowest nev eigenpairs.
 * h, o: ldh x nmat (upper triangles and real parts of the diagonals used, as ZHEGVX('U');
 * both overwritten, O's upper triangle by its Cholesky factor, the lower triangles by the
 * conjugates of the upper ones);
 * eval(nev) ascending, evec(ldv, nev) O-orthonormal (ldv >= nmat).  Returns 0, or LAPACK's
 * info > 0 (nmat + i: the leading minor of order i of O is not positive definite; i <= nmat:
 * i eigenvectors did not converge), or < 0 on an argument/HIP error.
 * exc_eigen_dev: device arrays, on the context's stream.  exc_eigen: host arrays. */
int exc_eigen_dev(exc_ctx *ctx, int nmat, int ldh, double *h, double *o, int nev, double *eval,
                  double *evec, int ldv);
int exc_eigen(exc_ctx *ctx, int nmat, int ldh, double *h, double *o, int nev, double *eval,
              double *evec, int ldv);
/* The lowest nev eigenpairs of nk k-points at once (device arrays; per k: nmat[k], ldh[k],
 * h[k], o[k], eval[k], evec[k], ldv[k] as for exc_eigen_dev, the pointer arrays themselves
 * in host memory).  The reference solves the k-points one after another in the loop of
 * scf_cycle.f90:241-276.  Default (variant 6): all k-points padded to the largest nmat and
 * solved as one strided batch, one launch per step for all of them, the tridiagonal divide
 * and conquer per k-point on worker streams (h, o left unchanged); variant 5 back-transforms
 * the eigenvectors per k-point instead of as one batch; variant 7 finds the tridiagonal
 * eigenpairs by bisection and inverse iteration as ZHEGVX does (slow); variant 8 is variant 6
 * with the per-k divide and conquer on one worker stream.  Every rocBLAS handle of the solve
 * gets a workspace sized before anything is queued, so no workspace is allocated or freed
 * while a solve is in flight.  Solver variants 0-4
 * (exc_ctx_set_variant, EXC_STAGE_EIGEN) instead run the single-k solves one after another
 * on a worker stream forked from and joined back into the context's stream (h, o
 * overwritten).  Returns 0, or the first failing k-point's status. */
int exc_eigen_batched_dev(exc_ctx *ctx, int nk, const int *nmat, const int *ldh, double *const *h,
                          double *const *o, int nev, double *const *eval, double *const *evec,
                          const int *ldv);

/* Muffin-tin density matrix of the first-variational states, the counterpart of
 * gendmatmt(ik, evecfv, evecsv) (gendmatmt.f90:13, called per k at scf_cycle.f90:351; the
 * spinless part, ValenceRelativity-independent).  Accumulates, per atom a,
 *   dm(:, :, a) += wkpt * wf diag(occ) wf^H,   wf = wf1 evecsv,
 *   wf1(1:naa, j) = A_a^T evecfv(1:ngp, j), rows f = (l, io, m), m fastest (:92-121),
 *   wf1(naa+1 : naa+losize, j) = evecfv(lo rows of a, j) (:123-131),
 * dm(wfsize, wfsize, natmtot) as mt_dm%alpha%ff, wfsize = exc_dmat_wfsize(ctx) = maxaa +
 * the largest number of lo functions of an atom (mod_potential_and_density.F90:186-230).
 * apwalm(ngkmax, apwordmax, lmmaxapw, natmtot) as returned by exc_match; evecfv(ldv, nst)
 * with ldv >= ngp + nlotot; evecsv(nst, nst) or NULL for the identity (first-variational
 * states, tevecsv false); occ(nst) = occsv(:, ik).  exc_gendmatmt_dev: device arrays, on the
 * context's stream.  exc_gendmatmt: host arrays (dm accumulated in host memory). */
int exc_dmat_wfsize(const exc_ctx *ctx);
int exc_gendmatmt_dev(exc_ctx *ctx, int ngp, int ngkmax, const double *apwalm, int nst, int ldv,
                      const double *evecfv, const double *evecsv, const double *occ, double wkpt,
                      double *dm);
int exc_gendmatmt(exc_ctx *ctx, int ngp, int ngkmax, const double *apwalm, int nst, int ldv,
                  const double *evecfv, const double *evecsv, const double *occ, double wkpt, double *dm);
/* The same for the nk k-points of a batch (apwalm, ngp, ngkmax from ks[k]; per k evecfv[k]
 * with ldv[k], evecsv[k] (the array or its entries may be NULL), occ[k], wkpt[k]; pointer
 * arrays in host memory, the arrays they point to on the device).  The sum over k-points
 * runs as one product per atom over the columns of all k-points. */
int exc_gendmatmt_batched_dev(exc_ctx *ctx, int nk, const exc_kpoint_dev *ks, int nst, const int *ldv,
                              const double *const *evecfv, const double *const *evecsv,
                              const double *const *occ, const double *wkpt, double *dm);
/* The muffin-tin density from the density matrix, the counterpart of genrhomt(basis, basis,
 * rhomt) (genrhomt.f90:13, called after the k loop when useDensityMatrix,
 * scf_cycle.f90:365-366; the spinless part):
 *   rhomt(lm2, r, a) = sum u_f(r) u_f'(r) Re(dm(f, f', a) conj(G(lm_f, lm2, lm_f')))
 * over APW-APW, twice APW-lo and lo-lo pairs (genrhomt.f90:34-150), G = int conj(Y) R Y
 * (gauntyry.f90), on the radial functions, meshes and lmaxvr of the last exc_ctx_hmlint call.
 * dm(wfsize, wfsize, natmtot) as exc_gendmatmt accumulates it; rhomt((lmaxvr+1)^2, nrmtmax,
 * natmtot) real (zero beyond each species' nrmt).  _dev: device arrays on the context's
 * stream; without: host arrays. */
int exc_genrhomt_dev(exc_ctx *ctx, const double *dm, double *rhomt);
int exc_genrhomt(exc_ctx *ctx, const double *dm, double *rhomt);
/* The muffin-tin density of one k-point by the direct route, the counterpart of
 * rhovalk(ik, evecfv, evecsv) (rhovalk.f90:1, called per k when useDensityMatrix is false,
 * scf_cycle.f90:410; spinless): the states are evecfv evecsv (evecsv(nst, nst), or NULL for the
 * identity: first-variational states, tevecsv false; the reference also leaves out entries of
 * evecsv with |Re| + |Im| <= epsocc, rhovalk.f90:115-117, which are kept here).  Per atom, for
 * every state j with |wkpt occ(j)| > epsocc, the wavefunction in complex harmonics on every
 * lradstep-th mesh point (wavefmt.f90:86-124, l <= lmaxvr), transformed to the lmmaxvr points
 * of the spherical covering by zbshtvr, squared and weighted, and transformed back by rfshtvr:
 *   rhomt(:, ir, a) += rfshtvr sum_j wkpt occ(j) |zbshtvr wfmt_j(:, ir)|^2,
 * ir = 1, 1 + lradstep, ... <= nrmt (rhovalk.f90:99-207; the other mesh points are not
 * touched).  Radial functions, meshes and lmaxvr (<= lmaxapw, <= 8; nlomax <= 8) of the last
 * exc_ctx_hmlint call; zbshtvr(lmmaxvr, lmmaxvr) complex and rfshtvr(lmmaxvr, lmmaxvr) real
 * as genshtmat makes them (genshtmat.f90:78-108); apwalm as exc_match returns it, evecfv(ldv,
 * nst), occ(nst); rhomt((lmaxvr+1)^2, nrmtmax, natmtot) accumulated.  _dev: device arrays on
 * the context's stream; without: host arrays. */
int exc_rhovalk_dev(exc_ctx *ctx, int lradstep, const double *zbshtvr, const double *rfshtvr, int ngp,
                    int ngkmax, const double *apwalm, int nst, int ldv, const double *evecfv,
                    const double *evecsv, const double *occ, double wkpt, double epsocc,
                    double *rhomt);
int exc_rhovalk(exc_ctx *ctx, int lradstep, const double *zbshtvr, const double *rfshtvr, int ngp,
                int ngkmax, const double *apwalm, int nst, int ldv, const double *evecfv,
                const double *evecsv, const double *occ, double wkpt, double epsocc, double *rhomt);
/* The interstitial density of one k-point, the counterpart of genrhoir(ik, evecfv, evecsv)
 * (genrhoir.f90:13, called beside gendmatmt at scf_cycle.f90:352; spinless, first-variational
 * states): for every state j with |wkpt occ(j)| > epsocc, psi_j(r) = the backward FFT (unscaled)
 * of evecfv(1:ngp, j) placed at the grid positions of G(igpig(1:ngp)), and
 *   rhoir(r) += wkpt occ(j) / Omega |psi_j(r)|^2
 * on the FFT grid of exc_ctx_gencfun (the G box), rhoir(ngrtot) accumulated.  _dev: device
 * arrays (evecfv(ldv, nst), occ(nst), igpig(ngp) 1-based), on the context's stream; the
 * batched form takes ngp and igpig from ks[k] and one evecfv, occ, wkpt per k-point, with the
 * states of every k-point transformed in batches of 64; without _dev: host arrays. */
int exc_genrhoir_dev(exc_ctx *ctx, int ngp, const int *igpig, int nst, int ldv, const double *evecfv,
                     const double *occ, double wkpt, double epsocc, double *rhoir);
int exc_genrhoir_batched_dev(exc_ctx *ctx, int nk, const exc_kpoint_dev *ks, int nst, const int *ldv,
                             const double *const *evecfv, const double *const *occ, const double *wkpt,
                             double epsocc, double *rhoir);
int exc_genrhoir(exc_ctx *ctx, int ngp, const int *igpig, int nst, int ldv, const double *evecfv,
                 const double *occ, double wkpt, double epsocc, double *rhoir);

/* Per-SCF producers of the interstitial tables, on the GPU.  The FFT grid is the G box of
 * exc_ctx_set_gvec (ngrid(i) = intgv(2,i) - intgv(1,i) + 1, every grid point one of the ngrtot
 * G-vectors; EXC_ERR_UNSUPPORTED otherwise), arrays on it in Fortran order, positions as
 * igfft (gengvec.f90:100-117); transforms as zfftifc.f90: forward scaled by 1/N.
 * exc_ctx_gencfun — gencfun.f90:13 with cfdamp = 0: Theta(G) = delta_G0 - sum_a
 *   e^(-i G.r_a) (4 pi/Omega)(sin GR - GR cos GR)/G^3 for every G (bvec(3,3): reciprocal
 *   lattice vectors as columns, mod_lattice), which replaces the context's cfunig, and
 *   Theta(r) = its backward transform (real part), kept for exc_ctx_genveffig.  Optional host
 *   outputs cfunig(ngrtot) complex, cfunir(ngrtot) real.
 * exc_ctx_genveffig — genveffig.f90:13: V(G) = forward transform of veffir(ngrtot) * Theta(r)
 *   at the first ngvec G, which replaces the context's veffig; optional host output
 *   veffig(ngvec) complex.  Needs exc_ctx_gencfun first. */
int exc_ctx_gencfun(exc_ctx *ctx, const double *bvec, double *cfunig, double *cfunir);
/* input%groundstate%cfdamp and gmaxvr for exc_ctx_gencfun (gencfun.f90:108-114): with
 * cfdamp > 0, Theta(G) is damped by exp(-(cfdamp |G| / gmaxvr)^2) after Theta(r) is taken
 * from the undamped values, as the reference does.  0 (the default): no damping. */
int exc_ctx_set_cfdamp(exc_ctx *ctx, double cfdamp, double gmaxvr);
int exc_ctx_genveffig(exc_ctx *ctx, const double *veffir, double *veffig);

/* G+k sets of k-points on the GPU (gengpvec.f90:13-94 and gensfacgp.f90:13-58; once per
 * k-set in init1, where the reference calls them): the G vectors of exc_ctx_set_gvec's ivg
 * (the reference's sorted order, gengvec.f90:74), G = i1 b1 + i2 b2 + i3 b3 (bvec(3,3) as in
 * exc_ctx_gencfun), that satisfy |G + k| < gkmax, in that order.
 * exc_gengpvec — one k-point, host arrays, the reference's arguments: vpl(3), vpc(3) the
 *   k-point in lattice and Cartesian coordinates; out ngp, igpig(ngkmax) (1-based G index),
 *   vgpl(3,ngkmax), vgpc(3,ngkmax), gpc(ngkmax), tpgpc(2,ngkmax) (sphcrd.f90: theta, phi),
 *   rows past ngp zero.  More than ngkmax vectors: EXC_ERR_ARG with *ngp set to the full
 *   count and the first ngkmax rows written (the reference stops, gengpvec.f90:53-58).
 * exc_gensfacgp — S_a(G+k) = exp(i (G+k).r_a) of ngp given vectors vgpc(3,ngp) for every atom
 *   (exc_ctx_set_atoms' positions) into sfacgp(ld, natmtot), host arrays.
 * exc_gengpvec_dev — nk k-points at once, one workgroup each: vpl, vpc (3,nk) host arrays,
 *   ngp(nk) host output; device outputs per k-point with leading dimension ngkmax:
 *   igpig(ngkmax,nk), vgpl(3,ngkmax,nk), vgpc(3,ngkmax,nk), gpc(ngkmax,nk), tpgpc(2,ngkmax,nk),
 *   and sfacgp(ngkmax,natmtot,nk) complex (null: not formed) — the layouts the setup entries
 *   read.  Synchronizes the context's stream (ngp comes back to the host). */
int exc_gengpvec(exc_ctx *ctx, const double *bvec, const double *vpl, const double *vpc, double gkmax,
                 int ngkmax, int *ngp, int *igpig, double *vgpl, double *vgpc, double *gpc, double *tpgpc);
int exc_gensfacgp(exc_ctx *ctx, int ngp, const double *vgpc, int ld, double *sfacgp);
int exc_gengpvec_dev(exc_ctx *ctx, int nk, const double *bvec, const double *vpl, const double *vpc, double gkmax,
                     int ngkmax, int *ngp, int *igpig, double *vgpl, double *vgpc, double *gpc, double *tpgpc,
                     double *sfacgp);

/* Number of local-orbital columns (nlotot) of the context, for sizing H and O. */
int exc_ctx_nlotot(const exc_ctx *ctx);

/* Deferred argument check of the setup calls, the counterpart of init1.f90:340-347, where the
 * reference stops when 2 gkmax > gmaxvr: that guard keeps every difference G_i - G_j of two
 * basis vectors of a k-point inside the first ngvec G-vectors (hmlistln.f90:77 and
 * olpistln.f90 look them up there).  The setup kernels count every pair they find outside
 * (that term would be dropped); exc_ctx_status waits for the context's stream and returns
 * EXC_ERR_ARG (and clears the count) if any setup, grouped setup or adjoint since the last
 * call met one, EXC_OK otherwise.  The host-array entry points call it before they return. */
int exc_ctx_status(exc_ctx *ctx);

/* Per-stage kernel timers, the counterpart of the reference's timing accumulators
 * (mod_timing.F90: timematch, time_hmlaan, time_olpaan, time_hmlalon, time_hmlistln, ...).
 * When enabled, every launch of a stage is bracketed by HIP events on the context's stream.
 * exc_ctx_read_timing waits for the stream, returns per stage the summed kernel time (ms)
 * and launch count since the last read, and clears them.  Stages, in array order:
 *   EXC_STAGE_MATCH, EXC_STAGE_YSTAGE (h.A^T), EXC_STAGE_MT (A^H.Y, A^H.A + interstitial),
 *   EXC_STAGE_LO (APW-lo and lo-lo blocks); arrays must hold EXC_NSTAGE entries. */
#define EXC_STAGE_MATCH 0
#define EXC_STAGE_YSTAGE 1
#define EXC_STAGE_MT 2
#define EXC_STAGE_LO 3
#define EXC_STAGE_BMT 4      /* muffin-tin stage of exc_setup_batched_dev                     */
#define EXC_STAGE_GRAPH 5    /* one replay of a batch's hipGraph                              */
#define EXC_STAGE_SPLIT 6    /* slicing of A and Yt into BF16 planes (EXC_BATCH_SPLIT_BF16)   */
#define EXC_STAGE_BSPLIT 7   /* split-BF16 muffin-tin stage of exc_setup_batched_dev          */
#define EXC_STAGE_VJP 8      /* exc_ho_vjp_dev                                                */
#define EXC_STAGE_OLP 9      /* addition-theorem overlap (EXC_BATCH_OLP_LEGENDRE)             */
#define EXC_STAGE_EIGEN 10   /* exc_eigen_dev                                                 */
#define EXC_STAGE_DMAT 11    /* exc_gendmatmt_dev (muffin-tin density matrix)                 */
#define EXC_STAGE_HMLINT 12  /* exc_ctx_hmlint (muffin-tin radial integrals and Gaunt sums)   */
#define EXC_STAGE_SCF 13     /* exc_ctx_gencfun, exc_ctx_genveffig (Theta, V(G), FFTs)        */
#define EXC_STAGE_CRTPREP 14 /* row scales and residue planes (EXC_BATCH_SPLIT_CRT)           */
#define EXC_STAGE_CRTGEMM 15 /* residue products, every tile and modulus                      */
#define EXC_STAGE_CRTMERGE 16 /* reconstruction, interstitial and mirror                      */
#define EXC_STAGE_CRTFORM 17 /* the modular split's operands: 0 A and Yt, 1 the Gram form      */
#define EXC_NSTAGE 17
int exc_ctx_set_timing(exc_ctx *ctx, int enable);

/* Number of moduli of EXC_BATCH_SPLIT_CRT (8 .. 12; 0, the default: 10 with the Gram form,
 * 11 without).  With n moduli the row integers have beta bits (the largest with
 * 2^(2 beta + 2) <= their product): 29 bits at n = 8, 33 at 9, 37 at 10, 40 at 11, 44 at 12;
 * the residue products cost 2 n int8 GEMMs.  No reference counterpart (the reference
 * computes in FP64). */
int exc_ctx_set_moduli(exc_ctx *ctx, int n);
/* What the last EXC_BATCH_SPLIT_CRT setup used: form 0 (A and Yt) or 1 (Gram), the number of
 * moduli, and the Gram form's shift sigma (0 for form 0).  The form is chosen per call:
 * EXC_STAGE_CRTFORM pins it (variant 0 or 1; the Gram form still needs
 * EXC_BATCH_OLP_LEGENDRE and no h1aa), the built-in choice takes the Gram form unless sigma
 * exceeds half of the largest |eigenvalue| of the h_a.  No reference counterpart. */
int exc_ctx_crt_form(const exc_ctx *ctx, int *form, int *nmod, double *sigma);
/* The first n moduli and beta of the modular split (for a caller's error estimate). */
int exc_crt_moduli(int n, int *moduli, int *beta);

/* Tile-shape variant of a stage's kernel (EXC_STAGE_YSTAGE, EXC_STAGE_MT, EXC_STAGE_BMT,
 * EXC_STAGE_BSPLIT, EXC_STAGE_CRTGEMM, EXC_STAGE_OLP), and for EXC_STAGE_GRAPH the shape of
 * the grouped hipGraph (0: the overlap and lo kernels on a branch of their own beside the
 * Y and muffin-tin stages, 1: all stages in one chain), as chosen by
 * the autotuner (exciting_amd/autotune.py); -1 restores the built-in choice by basis size.
 * exc_ctx_num_variants returns how many variants a stage has.  No reference counterpart. */
int exc_ctx_set_variant(exc_ctx *ctx, int stage, int variant);
int exc_ctx_num_variants(int stage);
int exc_ctx_read_timing(exc_ctx *ctx, double *ms, int *launches);

#ifdef __cplusplus
}
#endif
#endif /* EXC_LAPW_H */
