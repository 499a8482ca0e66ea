! Fortran driver of the GPU setup with the reference's call signatures.
!
!   match(ngp, gpc, tpgpc, sfacgp, apwalm)                        match.f90:13
!   hamiltonsetup(system, ngp, apwalm, igpig, vgpc)               hamiltonsetup.f90:8
!   overlapsetup(system, ngp, apwalm, igpig, vgpc)                overlapsetup.f90:8
!   hamiltonandoverlapsetup(system, ngp, apwalm, igpig, vgpc)     hamiltonandoverlapsetup.f90:8
!   solvewithgpu(system, nstfv, evecfv, evalfv)                   modfvsystem.f90:585 (solvewithlapack)
!   gendmatmt_gpu(ngp, apwalm, nmatmax, nstfv, evecfv, occ, wkpt, dm [, evecsv])
!                                                                 gendmatmt.f90:13
!   hmlint_gpu(lmaxvr, nrmt, spr, apwfr, lofr, veffmt, haa, loa, lolo, oalo, ololo)
!                                                                 hmlint.f90:14 + olprad.f90:13
!   set_radial_gpu(nrmt, spr, apwfr)                             match.f90:139-148 (zd by polynom)
!   gencfun_gpu(bvec, cfunig, cfunir [, cfdamp, gmaxvr])          gencfun.f90:13
!   genrhomt_gpu(dm, rhomt)                                       genrhomt.f90:13
!   genrhoir_gpu(ngp, igpig, nmatmax, nstfv, evecfv, occ, wkpt, epsocc, rhoir)
!                                                                 genrhoir.f90:13
!   genveffig_gpu(veffir, veffig)                                 genveffig.f90:13
!   rhovalk_gpu(lradstep, zbshtvr, rfshtvr, ngp, apwalm, nmatmax, nstfv, evecfv, occ, wkpt,
!               epsocc, rhomt [, evecsv])                         rhovalk.f90:1
!
! The module state the reference routines read implicitly (mod_atoms, mod_APW_LO,
! mod_muffin_tin, mod_eigensystem's mt_hscf, mod_Gvector, ...) is handed over once per SCF
! iteration by exc_gpu_init, which keeps the device context in this module.  Error
! handling follows the reference: a singular matching matrix is reported and the run
! continues (match.f90:166-175); any other failure stops the program.
!
! The setup calls receive apwalm as the reference's do; when it is the array the last match
! call filled (same address and ngp) and exc_resident_apwalm is .true. (the default), the
! library takes its own device copy of those coefficients instead of copying the host array
! again.  seceqnfv.f90:70-77 leaves apwalm unchanged between match and the setup; a caller
! that changes it in between sets exc_resident_apwalm = .false.
module exc_setup_gpu
  use, intrinsic :: iso_c_binding
  use exc_lapw_c
  use exc_evsystem
  implicit none
  private

  type(c_ptr), save :: ctx = c_null_ptr
  integer, save :: ngkmax_ = 0
  integer, save :: lmaxapw_ = 0, natmtot_ = 0
  ! the array and ngp of the last match call (for the device-resident apwalm)
  type(c_ptr), save :: last_apwalm = c_null_ptr
  integer, save :: last_ngp = -1
  logical, public, save :: exc_resident_apwalm = .true.

  public :: exc_gpu_init, exc_gpu_finalize, match, hamiltonandoverlapsetup, hamiltonsetup, overlapsetup, &
            exc_gpu_nlotot, forcek_gpu, &
            solvewithgpu, gendmatmt_gpu, exc_gpu_dmat_wfsize, hmlint_gpu, set_radial_gpu, gencfun_gpu, &
            genveffig_gpu, genrhomt_gpu, genrhoir_gpu, rhovalk_gpu

contains

  subroutine check(rc, where)
    integer(c_int), intent(in) :: rc
    character(*), intent(in) :: where
    if (rc < 0) then
      write(*, '("Error(", A, "): GPU setup returned ", I0)') where, rc
      error stop 1
    end if
  end subroutine check

  ! Upload the SCF-constant state.  Shapes as in the reference modules:
  ! natoms(nspecies), rmt(nspecies), atposc(3,natmtot), nlorb(nspecies), lorbl(nlomax,nspecies),
  ! idxlo(lolmmax,nlomax,natmtot), dmat(apwordmax,apwordmax,0:lmaxapw,natmtot),
  ! haa(maxaa,maxaa,natmtot), loa(maxnlo,maxaa,natmtot), lolo(maxnlo,maxnlo,natmtot),
  ! oalo(apwordmax,nlomax,natmtot), ololo(nlomax,nlomax,natmtot), ivg(3,ngrtot), ivgig(box),
  ! cfunig(ngrtot), veffig(ngvec), ffacg(ngvec,nspecies)
  subroutine exc_gpu_init(dims, ngkmax, natoms, rmt, atposc, nlorb, lorbl, idxlo, dmat, haa, loa, lolo, &
                          oalo, ololo, ivg, ivgig, cfunig, veffig, ffacg)
    type(exc_dims), intent(in) :: dims
    integer, intent(in) :: ngkmax
    integer(c_int), intent(in) :: natoms(*), nlorb(*), lorbl(*), idxlo(*), ivg(*), ivgig(*)
    real(8), intent(in) :: rmt(*), atposc(*), dmat(*), oalo(*), ololo(*), ffacg(*)
    complex(8), intent(in) :: haa(*), loa(*), lolo(*), cfunig(*), veffig(*)
    integer(c_int) :: rc
    if (c_associated(ctx)) call exc_gpu_finalize()
    call check(exc_ctx_create(dims, ctx), "exc_ctx_create")
    ngkmax_ = ngkmax
    lmaxapw_ = dims%lmaxapw
    natmtot_ = dims%natmtot
    call check(exc_ctx_set_atoms(ctx, natoms, rmt, atposc), "exc_ctx_set_atoms")
    call check(exc_ctx_set_lo(ctx, nlorb, lorbl, idxlo), "exc_ctx_set_lo")
    rc = exc_ctx_set_radial(ctx, dmat)
    call check(rc, "exc_ctx_set_radial")
    if (rc > 0) write(*, '("Warning(exc_gpu_init): singular matching matrix, info = ", I0)') rc
    call check(exc_ctx_set_mt(ctx, haa, loa, lolo, oalo, ololo), "exc_ctx_set_mt")
    call check(exc_ctx_set_gvec(ctx, ivg, ivgig, cfunig, veffig, ffacg), "exc_ctx_set_gvec")
  end subroutine exc_gpu_init

  subroutine exc_gpu_finalize()
    integer(c_int) :: rc
    if (c_associated(ctx)) rc = exc_ctx_destroy(ctx)
    ctx = c_null_ptr
  end subroutine exc_gpu_finalize

  integer function exc_gpu_nlotot()
    exc_gpu_nlotot = exc_ctx_nlotot(ctx)
  end function exc_gpu_nlotot

  ! match.f90:13 — same arguments; ngkmax is the leading dimension set at init
  subroutine match(ngp, gpc, tpgpc, sfacgp, apwalm)
    integer, intent(in) :: ngp
    real(8), intent(in) :: gpc(ngkmax_), tpgpc(2, ngkmax_)
    complex(8), intent(in) :: sfacgp(ngkmax_, *)
    complex(8), intent(inout), target :: apwalm(ngkmax_, *)
    integer(c_int) :: rc, l, ias
    last_apwalm = c_null_ptr
    rc = exc_match(ctx, int(ngp, c_int), int(ngkmax_, c_int), gpc, tpgpc, sfacgp, apwalm)
    call check(rc, "match")
    last_apwalm = c_loc(apwalm(1, 1))
    last_ngp = ngp
    if (rc > 0) then
      l = mod(rc - 1, lmaxapw_ + 1)
      ias = (rc - 1) / (lmaxapw_ + 1) + 1
      write(*,*)
      write(*, '("Error(match): could not find APW matching coefficients")')
      write(*, '(" for atom ", I4, " and l ", I4)') ias, l
      write(*, '(" singular matching matrix, INFO = ", I8)') rc
      write(*,*)
    end if
  end subroutine match

  ! hamiltonandoverlapsetup.f90:8 — H and O of one k-point into system%hamilton%za and
  ! system%overlap%za (rank = nmat).  The matrices are overwritten, which equals the
  ! reference's accumulation into the zeroed matrices that newsystem provides.
  subroutine hamiltonandoverlapsetup(system, ngp, apwalm, igpig, vgpc)
    type(evsystem), intent(inout) :: system
    integer, intent(in) :: ngp
    complex(8), intent(in), target :: apwalm(ngkmax_, *)
    integer, intent(in) :: igpig(ngkmax_)
    real(8), intent(in) :: vgpc(3, ngkmax_)
    integer(c_int) :: ldh
    ldh = int(system%hamilton%rank, c_int)
    if (system%overlap%rank /= ldh) error stop "hamiltonandoverlapsetup: H and O ranks differ"
    call check(exc_hamiltonandoverlapsetup(ctx, int(ngp, c_int), int(ngkmax_, c_int), apw_arg(apwalm, ngp), &
                                           igpig, vgpc, ldh, system%hamilton%za, system%overlap%za), &
               "hamiltonandoverlapsetup")
  end subroutine hamiltonandoverlapsetup

  ! hamiltonsetup.f90:8 — H of one k-point into system%hamilton%za (rank = nmat), both
  ! triangles, overwritten; seceqnfv.f90:76 calls it right before overlapsetup
  subroutine hamiltonsetup(system, ngp, apwalm, igpig, vgpc)
    type(evsystem), intent(inout) :: system
    integer, intent(in) :: ngp
    complex(8), intent(in), target :: apwalm(ngkmax_, *)
    integer, intent(in) :: igpig(ngkmax_)
    real(8), intent(in) :: vgpc(3, ngkmax_)
    call check(exc_hamiltonsetup(ctx, int(ngp, c_int), int(ngkmax_, c_int), apw_arg(apwalm, ngp), igpig, vgpc, &
                                 int(system%hamilton%rank, c_int), system%hamilton%za), "hamiltonsetup")
  end subroutine hamiltonsetup

  ! overlapsetup.f90:8 — O of one k-point into system%overlap%za, both triangles, overwritten
  subroutine overlapsetup(system, ngp, apwalm, igpig, vgpc)
    type(evsystem), intent(inout) :: system
    integer, intent(in) :: ngp
    complex(8), intent(in), target :: apwalm(ngkmax_, *)
    integer, intent(in) :: igpig(ngkmax_)
    real(8), intent(in) :: vgpc(3, ngkmax_)
    call check(exc_overlapsetup(ctx, int(ngp, c_int), int(ngkmax_, c_int), apw_arg(apwalm, ngp), igpig, vgpc, &
                                int(system%overlap%rank, c_int), system%overlap%za), "overlapsetup")
  end subroutine overlapsetup

  ! the apwalm argument of a setup call: c_null_ptr (the library's device copy) when the array
  ! is the one the last match call filled for this ngp, else its address
  type(c_ptr) function apw_arg(apwalm, ngp)
    complex(8), intent(in), target :: apwalm(ngkmax_, *)
    integer, intent(in) :: ngp
    apw_arg = c_loc(apwalm(1, 1))
    if (exc_resident_apwalm .and. ngp == last_ngp .and. c_associated(apw_arg, last_apwalm)) apw_arg = c_null_ptr
  end function apw_arg

  ! forcek.f90:13 — the IBS force contribution of one k-point (spin-unpolarised,
  ! first-variational states), added to forceibs(3, natmtot) with the k-point weight:
  ! forceibs -= d/dtau sum_n wkpt occ_n e_n  (F = -dE/dtau, forcek.f90:371-383)
  subroutine forcek_gpu(ngp, apwalm, igpig, vgpc, sfacgp, nmatmax, nstfv, evecfv, evalfv, occ, wkpt, forceibs)
    integer, intent(in) :: ngp, nmatmax, nstfv
    complex(8), intent(in) :: apwalm(ngkmax_, *), sfacgp(ngkmax_, *), evecfv(nmatmax, nstfv)
    integer, intent(in) :: igpig(ngkmax_)
    real(8), intent(in) :: vgpc(3, ngkmax_), evalfv(nstfv), occ(nstfv), wkpt
    real(8), intent(inout) :: forceibs(3, natmtot_)
    real(8) :: wh(nstfv), wo(nstfv), dedtau(3, natmtot_)
    wh = wkpt * occ
    wo = -wkpt * occ * evalfv
    dedtau = 0.d0
    call check(exc_ho_vjp(ctx, int(ngp, c_int), int(ngkmax_, c_int), apwalm, igpig, vgpc, sfacgp, &
                          int(nstfv, c_int), int(nmatmax, c_int), evecfv, wh, wo, dedtau, c_null_ptr), "forcek")
    forceibs = forceibs - dedtau
  end subroutine forcek_gpu

  ! modfvsystem.f90:585 (solvewithlapack) — the lowest nstfv eigenpairs of the first-variational
  ! system, the counterpart of its ZHEGVX(1, 'V', 'I', 'U', ..., 1, nstfv, ...) (:691-693), on the
  ! GPU.  H and O are overwritten as ZHEGVX leaves them; a failure stops the program as the
  ! reference does (:698-711).
  subroutine solvewithgpu(system, nstfv, evecfv, evalfv)
    type(evsystem), intent(inout) :: system
    integer, intent(in) :: nstfv
    complex(8), intent(out) :: evecfv(:, :)
    real(8), intent(out) :: evalfv(nstfv)
    integer(c_int) :: rc, n
    n = int(system%hamilton%rank, c_int)
    if (size(evecfv, 1) < n .or. size(evecfv, 2) < nstfv) error stop "solvewithgpu: evecfv is too small"
    rc = exc_eigen(ctx, n, n, system%hamilton%za, system%overlap%za, int(nstfv, c_int), evalfv, evecfv, &
                   int(size(evecfv, 1), c_int))
    if (rc /= 0) then
      write(*,*)
      write(*, '("Error(solvewithgpu): diagonalisation failed")')
      write(*, '(" ZHEGVX-style INFO = ", I8)') rc
      write(*,*)
      error stop 1
    end if
  end subroutine solvewithgpu

  ! Rows (and columns) of the muffin-tin density matrix, mt_dm%maxaa + mt_dm%maxnlo
  integer function exc_gpu_dmat_wfsize()
    exc_gpu_dmat_wfsize = int(exc_dmat_wfsize(ctx))
    if (exc_gpu_dmat_wfsize < 0) call check(int(exc_gpu_dmat_wfsize, c_int), "exc_dmat_wfsize")
  end function exc_gpu_dmat_wfsize

  ! gendmatmt.f90:13 (gendmatmt(ik, evecfv, evecsv), the spinless part) with the module state it
  ! reads passed in: apwalm of the k-point (match), wkpt(ik), occsv(:, ik).  Adds to
  ! dm(wfsize, wfsize, natmtot) (mt_dm%alpha%ff), per atom a,
  !   dm(:, :, a) += wkpt * wf diag(occ) wf^H,  wf = [A_a^T evecfv; evecfv(lo rows of a)] evecsv
  ! with evecsv(nstfv, nstfv) the identity when absent.
  subroutine gendmatmt_gpu(ngp, apwalm, nmatmax, nstfv, evecfv, occ, wkpt, dm, evecsv)
    integer, intent(in) :: ngp, nmatmax, nstfv
    complex(8), intent(in) :: apwalm(ngkmax_, *), evecfv(nmatmax, nstfv)
    real(8), intent(in) :: occ(nstfv), wkpt
    complex(8), intent(inout) :: dm(:, :, :)
    complex(8), intent(in), target, optional :: evecsv(nstfv, nstfv)
    type(c_ptr) :: psv
    if (size(dm, 1) /= exc_gpu_dmat_wfsize() .or. size(dm, 2) /= size(dm, 1) .or. size(dm, 3) /= natmtot_) &
      error stop "gendmatmt_gpu: dm must be (wfsize, wfsize, natmtot)"
    psv = c_null_ptr
    if (present(evecsv)) psv = c_loc(evecsv)
    call check(exc_gendmatmt(ctx, int(ngp, c_int), int(ngkmax_, c_int), apwalm, int(nstfv, c_int), &
                             int(nmatmax, c_int), evecfv, psv, occ, wkpt, dm), "gendmatmt")
  end subroutine gendmatmt_gpu

  ! hmlint.f90:14 (mt_kin + mt_pot, ValenceRelativity 'none', spinless) and olprad.f90:13 with
  ! the module state they read passed in: the radial meshes nrmt(nspecies), spr(nrmtmax,
  ! nspecies), apwfr(nrmtmax, 2, apwordmax, 0:lmaxapw, natmtot), lofr(nrmtmax, 2, nlomax,
  ! natmtot), veffmt(lmmaxvr, nrmtmax, natmtot).  The matrices replace the context's (the
  ! setup uses them from here on) and are returned in mt_hscf%main%aa / loa / lolo and oalo /
  ! ololo shapes.  Called once per SCF iteration, after the radial functions are updated.
  subroutine hmlint_gpu(lmaxvr, nrmt, spr, apwfr, lofr, veffmt, haa, loa, lolo, oalo, ololo)
    integer, intent(in) :: lmaxvr
    integer(c_int), intent(in) :: nrmt(:)
    real(8), intent(in) :: spr(:, :), apwfr(*), lofr(*), veffmt(*)
    complex(8), intent(out) :: haa(*), loa(*), lolo(*)
    real(8), intent(out) :: oalo(*), ololo(*)
    call check(exc_ctx_hmlint(ctx, int(lmaxvr, c_int), int(size(spr, 1), c_int), nrmt, spr, apwfr, lofr, veffmt, &
                              haa, loa, lolo, oalo, ololo), "hmlint")
  end subroutine hmlint_gpu

  ! match.f90:139-148 — the boundary values zd that match builds with polynom from apwfr at R,
  ! installed once per SCF iteration (after the radial functions change) instead of dmat.
  ! A singular matching matrix is reported as match.f90:166-175 does; the run continues.
  subroutine set_radial_gpu(nrmt, spr, apwfr)
    integer(c_int), intent(in) :: nrmt(:)
    real(8), intent(in) :: spr(:, :), apwfr(*)
    integer(c_int) :: rc
    rc = exc_ctx_set_radial_fr(ctx, int(size(spr, 1), c_int), nrmt, spr, apwfr)
    call check(rc, "set_radial_gpu")
    if (rc > 0) write(*, '("Warning(set_radial_gpu): singular matching matrix, info = ", I0)') rc
  end subroutine set_radial_gpu

  ! gencfun.f90:13: Theta(G) (cfunig(ngrtot)) and Theta(r) (cfunir(ngrtot)) on the GPU from
  ! the atoms of the context; bvec(3,3) as mod_lattice.  With cfdamp > 0 (and gmaxvr, both as
  ! input%groundstate) Theta(G) is damped by exp(-(cfdamp |G| / gmaxvr)^2) after Theta(r) is
  ! taken (gencfun.f90:108-114); absent, no damping.  Theta(G) replaces the context's cfunig;
  ! Theta(r) stays on the device for genveffig_gpu.
  subroutine gencfun_gpu(bvec, cfunig, cfunir, cfdamp, gmaxvr)
    real(8), intent(in) :: bvec(3, 3)
    complex(8), intent(out) :: cfunig(*)
    real(8), intent(out) :: cfunir(*)
    real(8), intent(in), optional :: cfdamp, gmaxvr
    if (present(cfdamp) .and. present(gmaxvr)) then
      call check(exc_ctx_set_cfdamp(ctx, cfdamp, gmaxvr), "gencfun (cfdamp)")
    else
      call check(exc_ctx_set_cfdamp(ctx, 0.d0, 1.d0), "gencfun (cfdamp)")
    end if
    call check(exc_ctx_gencfun(ctx, bvec, cfunig, cfunir), "gencfun")
  end subroutine gencfun_gpu

  ! genveffig.f90:13: V(G) = forward FFT of veffir(ngrtot) Theta(r) (veffig(ngvec)), which
  ! replaces the context's veffig.  After gencfun_gpu.
  subroutine genveffig_gpu(veffir, veffig)
    real(8), intent(in) :: veffir(*)
    complex(8), intent(out) :: veffig(*)
    call check(exc_ctx_genveffig(ctx, veffir, veffig), "genveffig")
  end subroutine genveffig_gpu

  ! genrhomt.f90:13 (genrhomt(basis, basis, rhomt), scf_cycle.f90:365-366): the muffin-tin
  ! density rhomt(lmmaxvr, nrmtmax, natmtot) from the density matrix dm (mt_dm%alpha%ff, as
  ! gendmatmt_gpu accumulates it) and the radial functions of the last hmlint_gpu call.
  subroutine genrhomt_gpu(dm, rhomt)
    complex(8), intent(in) :: dm(:, :, :)
    real(8), intent(out) :: rhomt(*)
    call check(exc_genrhomt(ctx, dm, rhomt), "genrhomt")
  end subroutine genrhomt_gpu

  ! genrhoir.f90:13 (genrhoir(ik, evecfv, evecsv), spinless first-variational states) with the
  ! module state it reads passed in (igkig(:, 1, ik), occsv(:, ik), wkpt(ik), epsocc): adds
  ! wkpt occ / Omega |psi|^2 of each state with |wkpt occ| > epsocc to rhoir(ngrtot).
  subroutine genrhoir_gpu(ngp, igpig, nmatmax, nstfv, evecfv, occ, wkpt, epsocc, rhoir)
    integer, intent(in) :: ngp, nmatmax, nstfv
    integer(c_int), intent(in) :: igpig(*)
    complex(8), intent(in) :: evecfv(nmatmax, nstfv)
    real(8), intent(in) :: occ(nstfv), wkpt, epsocc
    real(8), intent(inout) :: rhoir(*)
    call check(exc_genrhoir(ctx, int(ngp, c_int), igpig, int(nstfv, c_int), int(nmatmax, c_int), evecfv, occ, wkpt, &
                            epsocc, rhoir), "genrhoir")
  end subroutine genrhoir_gpu

  ! rhovalk.f90:1 (rhovalk(ik, evecfv, evecsv), useDensityMatrix false, scf_cycle.f90:410;
  ! spinless) with the module state it reads passed in (input%groundstate%lradstep, zbshtvr,
  ! rfshtvr of genshtmat, occsv(:, ik), wkpt(ik), epsocc): adds the k-point's muffin-tin
  ! density to rhomt(lmmaxvr, nrmtmax, natmtot) on the radial functions and lmaxvr of the last
  ! hmlint_gpu call; the states are evecfv evecsv, evecsv(nstfv, nstfv) the identity when absent.
  subroutine rhovalk_gpu(lradstep, zbshtvr, rfshtvr, ngp, apwalm, nmatmax, nstfv, evecfv, occ, wkpt, epsocc, rhomt, &
                         evecsv)
    integer, intent(in) :: lradstep, ngp, nmatmax, nstfv
    complex(8), intent(in) :: zbshtvr(:, :), apwalm(:, :, :, :), evecfv(nmatmax, nstfv)
    real(8), intent(in) :: rfshtvr(:, :), occ(nstfv), wkpt, epsocc
    real(8), intent(inout) :: rhomt(*)
    complex(8), intent(in), target, optional :: evecsv(nstfv, nstfv)
    type(c_ptr) :: psv
    psv = c_null_ptr
    if (present(evecsv)) psv = c_loc(evecsv)
    call check(exc_rhovalk(ctx, int(lradstep, c_int), zbshtvr, rfshtvr, int(ngp, c_int), int(size(apwalm, 1), c_int), &
                           apwalm, int(nstfv, c_int), int(nmatmax, c_int), evecfv, psv, occ, wkpt, epsocc, rhomt), &
               "rhovalk")
  end subroutine rhovalk_gpu

end module exc_setup_gpu
