! iso_c_binding interfaces of libexcgpu.so (include/exc_lapw.h).
! Every routine returns the C-ABI status: 0 ok, > 0 LAPACK-style info, < 0 error.
module exc_lapw_c
  use, intrinsic :: iso_c_binding
  implicit none

  integer(c_int), parameter :: EXC_OK = 0, EXC_ERR_ARG = -1, EXC_ERR_HIP = -2, &
                               EXC_ERR_UNSUPPORTED = -3, EXC_ERR_STATE = -4

  type, bind(c) :: exc_dims
    integer(c_int) :: natmtot, nspecies, lmaxapw, lmaxmat, apwordmax, nlomax, lolmmax, maxnlo
    integer(c_int) :: ngrtot, ngvec
    integer(c_int) :: intgv(6)
    real(c_double) :: omega
  end type exc_dims

  interface
    integer(c_int) function exc_ctx_create(dims, ctx) bind(c, name="exc_ctx_create")
      import :: c_int, c_ptr, exc_dims
      type(exc_dims), intent(in) :: dims
      type(c_ptr), intent(out) :: ctx
    end function

    integer(c_int) function exc_ctx_destroy(ctx) bind(c, name="exc_ctx_destroy")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
    end function

    integer(c_int) function exc_ctx_set_atoms(ctx, natoms, rmt, atposc) bind(c, name="exc_ctx_set_atoms")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      integer(c_int), intent(in) :: natoms(*)
      real(c_double), intent(in) :: rmt(*), atposc(*)
    end function

    integer(c_int) function exc_ctx_set_lo(ctx, nlorb, lorbl, idxlo) bind(c, name="exc_ctx_set_lo")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
      integer(c_int), intent(in) :: nlorb(*), lorbl(*), idxlo(*)
    end function

    integer(c_int) function exc_ctx_set_radial(ctx, dmat) bind(c, name="exc_ctx_set_radial")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      real(c_double), intent(in) :: dmat(*)
    end function

    integer(c_int) function exc_ctx_set_mt(ctx, haa, loa, lolo, oalo, ololo) bind(c, name="exc_ctx_set_mt")
      import :: c_int, c_ptr, c_double, c_double_complex
      type(c_ptr), value :: ctx
      complex(c_double_complex), intent(in) :: haa(*), loa(*), lolo(*)
      real(c_double), intent(in) :: oalo(*), ololo(*)
    end function

    integer(c_int) function exc_ctx_set_meffig(ctx, meffig) bind(c, name="exc_ctx_set_meffig")
      import :: c_int, c_ptr, c_double_complex
      type(c_ptr), value :: ctx
      complex(c_double_complex), intent(in) :: meffig(*)
    end function

    integer(c_int) function exc_ctx_set_iora(ctx, h1aa, h1loa, h1lolo, m2effig) bind(c, name="exc_ctx_set_iora")
      import :: c_int, c_ptr, c_double, c_double_complex
      type(c_ptr), value :: ctx
      real(c_double), intent(in) :: h1aa(*), h1loa(*), h1lolo(*)
      complex(c_double_complex), intent(in) :: m2effig(*)
    end function

    integer(c_int) function exc_ctx_set_cfdamp(ctx, cfdamp, gmaxvr) bind(c, name="exc_ctx_set_cfdamp")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      real(c_double), value :: cfdamp, gmaxvr
    end function

    integer(c_int) function exc_ctx_gencfun(ctx, bvec, cfunig, cfunir) bind(c, name="exc_ctx_gencfun")
      import :: c_int, c_ptr, c_double, c_double_complex
      type(c_ptr), value :: ctx
      real(c_double), intent(in) :: bvec(*)
      complex(c_double_complex), intent(inout) :: cfunig(*)
      real(c_double), intent(inout) :: cfunir(*)
    end function

    integer(c_int) function exc_ctx_genveffig(ctx, veffir, veffig) bind(c, name="exc_ctx_genveffig")
      import :: c_int, c_ptr, c_double, c_double_complex
      type(c_ptr), value :: ctx
      real(c_double), intent(in) :: veffir(*)
      complex(c_double_complex), intent(inout) :: veffig(*)
    end function

    integer(c_int) function exc_ctx_set_radial_fr(ctx, nrmtmax, nrmt, spr, apwfr) bind(c, name="exc_ctx_set_radial_fr")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      integer(c_int), value :: nrmtmax
      integer(c_int), intent(in) :: nrmt(*)
      real(c_double), intent(in) :: spr(*), apwfr(*)
    end function

    integer(c_int) function exc_ctx_hmlint(ctx, lmaxvr, nrmtmax, nrmt, spr, apwfr, lofr, veffmt, haa, loa, lolo, &
                                           oalo, ololo) bind(c, name="exc_ctx_hmlint")
      import :: c_int, c_ptr, c_double, c_double_complex
      type(c_ptr), value :: ctx
      integer(c_int), value :: lmaxvr, nrmtmax
      integer(c_int), intent(in) :: nrmt(*)
      real(c_double), intent(in) :: spr(*), apwfr(*), lofr(*), veffmt(*)
      complex(c_double_complex), intent(inout) :: haa(*), loa(*), lolo(*)
      real(c_double), intent(inout) :: oalo(*), ololo(*)
    end function

    integer(c_int) function exc_ctx_set_gvec(ctx, ivg, ivgig, cfunig, veffig, ffacg) bind(c, name="exc_ctx_set_gvec")
      import :: c_int, c_ptr, c_double, c_double_complex
      type(c_ptr), value :: ctx
      integer(c_int), intent(in) :: ivg(*), ivgig(*)
      complex(c_double_complex), intent(in) :: cfunig(*), veffig(*)
      real(c_double), intent(in) :: ffacg(*)
    end function

    integer(c_int) function exc_match(ctx, ngp, ngkmax, gpc, tpgpc, sfacgp, apwalm) bind(c, name="exc_match")
      import :: c_int, c_ptr, c_double, c_double_complex
      type(c_ptr), value :: ctx
      integer(c_int), value :: ngp, ngkmax
      real(c_double), intent(in) :: gpc(*), tpgpc(*)
      complex(c_double_complex), intent(in) :: sfacgp(*)
      complex(c_double_complex), intent(inout) :: apwalm(*)
    end function

    ! apwalm is passed by address (c_loc of its first element), or c_null_ptr for the device
    ! copy of the last exc_match call
    integer(c_int) function exc_hamiltonandoverlapsetup(ctx, ngp, ngkmax, apwalm, igpig, vgpc, ldh, h, o) &
        bind(c, name="exc_hamiltonandoverlapsetup")
      import :: c_int, c_ptr, c_double, c_double_complex
      type(c_ptr), value :: ctx, apwalm
      integer(c_int), value :: ngp, ngkmax, ldh
      integer(c_int), intent(in) :: igpig(*)
      real(c_double), intent(in) :: vgpc(*)
      complex(c_double_complex), intent(inout) :: h(*), o(*)
    end function

    integer(c_int) function exc_hamiltonsetup(ctx, ngp, ngkmax, apwalm, igpig, vgpc, ldh, h) &
        bind(c, name="exc_hamiltonsetup")
      import :: c_int, c_ptr, c_double, c_double_complex
      type(c_ptr), value :: ctx, apwalm
      integer(c_int), value :: ngp, ngkmax, ldh
      integer(c_int), intent(in) :: igpig(*)
      real(c_double), intent(in) :: vgpc(*)
      complex(c_double_complex), intent(inout) :: h(*)
    end function

    integer(c_int) function exc_overlapsetup(ctx, ngp, ngkmax, apwalm, igpig, vgpc, ldh, o) &
        bind(c, name="exc_overlapsetup")
      import :: c_int, c_ptr, c_double, c_double_complex
      type(c_ptr), value :: ctx, apwalm
      integer(c_int), value :: ngp, ngkmax, ldh
      integer(c_int), intent(in) :: igpig(*)
      real(c_double), intent(in) :: vgpc(*)
      complex(c_double_complex), intent(inout) :: o(*)
    end function

    integer(c_int) function exc_ho_vjp(ctx, ngp, ngkmax, apwalm, igpig, vgpc, sfacgp, nst, ldc, evec, wh, wo, &
                                       dedtau, dedv) bind(c, name="exc_ho_vjp")
      import :: c_int, c_ptr, c_double, c_double_complex
      type(c_ptr), value :: ctx
      integer(c_int), value :: ngp, ngkmax, nst, ldc
      complex(c_double_complex), intent(in) :: apwalm(*), sfacgp(*), evec(*)
      integer(c_int), intent(in) :: igpig(*)
      real(c_double), intent(in) :: vgpc(*), wh(*), wo(*)
      real(c_double), intent(inout) :: dedtau(*)
      type(c_ptr), value :: dedv          ! c_loc of a complex(8) dedv(ngvec), or c_null_ptr
    end function

    integer(c_int) function exc_eigen(ctx, nmat, ldh, h, o, nev, eval, evec, ldv) bind(c, name="exc_eigen")
      import :: c_int, c_ptr, c_double, c_double_complex
      type(c_ptr), value :: ctx
      integer(c_int), value :: nmat, ldh, nev, ldv
      complex(c_double_complex), intent(inout) :: h(*), o(*), evec(*)
      real(c_double), intent(inout) :: eval(*)
    end function

    integer(c_int) function exc_dmat_wfsize(ctx) bind(c, name="exc_dmat_wfsize")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
    end function

    integer(c_int) function exc_gendmatmt(ctx, ngp, ngkmax, apwalm, nst, ldv, evecfv, evecsv, occ, wkpt, dm) &
        bind(c, name="exc_gendmatmt")
      import :: c_int, c_ptr, c_double, c_double_complex
      type(c_ptr), value :: ctx, evecsv
      integer(c_int), value :: ngp, ngkmax, nst, ldv
      complex(c_double_complex), intent(in) :: apwalm(*), evecfv(*)
      real(c_double), intent(in) :: occ(*)
      real(c_double), value :: wkpt
      complex(c_double_complex), intent(inout) :: dm(*)
    end function

    integer(c_int) function exc_genrhomt(ctx, dm, rhomt) bind(c, name="exc_genrhomt")
      import :: c_int, c_ptr, c_double, c_double_complex
      type(c_ptr), value :: ctx
      complex(c_double_complex), intent(in) :: dm(*)
      real(c_double), intent(inout) :: rhomt(*)
    end function

    integer(c_int) function exc_rhovalk(ctx, lradstep, zbshtvr, rfshtvr, ngp, ngkmax, apwalm, nst, ldv, evecfv, &
        evecsv, occ, wkpt, epsocc, rhomt) bind(c, name="exc_rhovalk")
      import :: c_int, c_ptr, c_double, c_double_complex
      type(c_ptr), value :: ctx, evecsv
      integer(c_int), value :: lradstep, ngp, ngkmax, nst, ldv
      complex(c_double_complex), intent(in) :: zbshtvr(*), apwalm(*), evecfv(*)
      real(c_double), intent(in) :: rfshtvr(*), occ(*)
      real(c_double), value :: wkpt, epsocc
      real(c_double), intent(inout) :: rhomt(*)
    end function

    integer(c_int) function exc_genrhoir(ctx, ngp, igpig, nst, ldv, evecfv, occ, wkpt, epsocc, rhoir) &
        bind(c, name="exc_genrhoir")
      import :: c_int, c_ptr, c_double, c_double_complex
      type(c_ptr), value :: ctx
      integer(c_int), value :: ngp, nst, ldv
      integer(c_int), intent(in) :: igpig(*)
      complex(c_double_complex), intent(in) :: evecfv(*)
      real(c_double), intent(in) :: occ(*)
      real(c_double), value :: wkpt, epsocc
      real(c_double), intent(inout) :: rhoir(*)
    end function

    integer(c_int) function exc_ctx_nlotot(ctx) bind(c, name="exc_ctx_nlotot")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
    end function

    integer(c_int) function exc_ctx_status(ctx) bind(c, name="exc_ctx_status")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
    end function
  end interface
end module exc_lapw_c
