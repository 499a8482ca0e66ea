﻿!mod$ v1 sum:afc1cc0deae3bf01
!need$ f71f194d1991c68b n exc_lapw_c
!need$ 4e11352129309ed7 n exc_evsystem
!need$ bb381bf46e508468 i __fortran_builtins
!need$ 0bde2ac47243ead2 i iso_c_binding
module exc_setup_gpu
use,intrinsic::iso_c_binding,only:c_associated
use,intrinsic::iso_c_binding,only:c_funloc
use,intrinsic::iso_c_binding,only:c_funptr
use,intrinsic::iso_c_binding,only:c_f_pointer
use,intrinsic::iso_c_binding,only:c_loc
use,intrinsic::iso_c_binding,only:c_null_funptr
use,intrinsic::iso_c_binding,only:c_null_ptr
use,intrinsic::iso_c_binding,only:c_ptr
use,intrinsic::iso_c_binding,only:c_sizeof
use,intrinsic::iso_c_binding,only:operator(==)
use,intrinsic::iso_c_binding,only:operator(/=)
use,intrinsic::iso_c_binding,only:c_int8_t
use,intrinsic::iso_c_binding,only:c_int16_t
use,intrinsic::iso_c_binding,only:c_int32_t
use,intrinsic::iso_c_binding,only:c_int64_t
use,intrinsic::iso_c_binding,only:c_int128_t
use,intrinsic::iso_c_binding,only:c_int
use,intrinsic::iso_c_binding,only:c_short
use,intrinsic::iso_c_binding,only:c_long
use,intrinsic::iso_c_binding,only:c_long_long
use,intrinsic::iso_c_binding,only:c_signed_char
use,intrinsic::iso_c_binding,only:c_size_t
use,intrinsic::iso_c_binding,only:c_intmax_t
use,intrinsic::iso_c_binding,only:c_intptr_t
use,intrinsic::iso_c_binding,only:c_ptrdiff_t
use,intrinsic::iso_c_binding,only:c_int_least8_t
use,intrinsic::iso_c_binding,only:c_int_fast8_t
use,intrinsic::iso_c_binding,only:c_int_least16_t
use,intrinsic::iso_c_binding,only:c_int_fast16_t
use,intrinsic::iso_c_binding,only:c_int_least32_t
use,intrinsic::iso_c_binding,only:c_int_fast32_t
use,intrinsic::iso_c_binding,only:c_int_least64_t
use,intrinsic::iso_c_binding,only:c_int_fast64_t
use,intrinsic::iso_c_binding,only:c_int_least128_t
use,intrinsic::iso_c_binding,only:c_int_fast128_t
use,intrinsic::iso_c_binding,only:c_float
use,intrinsic::iso_c_binding,only:c_double
use,intrinsic::iso_c_binding,only:c_long_double
use,intrinsic::iso_c_binding,only:c_float_complex
use,intrinsic::iso_c_binding,only:c_double_complex
use,intrinsic::iso_c_binding,only:c_long_double_complex
use,intrinsic::iso_c_binding,only:c_bool
use,intrinsic::iso_c_binding,only:c_char
use,intrinsic::iso_c_binding,only:c_null_char
use,intrinsic::iso_c_binding,only:c_alert
use,intrinsic::iso_c_binding,only:c_backspace
use,intrinsic::iso_c_binding,only:c_form_feed
use,intrinsic::iso_c_binding,only:c_new_line
use,intrinsic::iso_c_binding,only:c_carriage_return
use,intrinsic::iso_c_binding,only:c_horizontal_tab
use,intrinsic::iso_c_binding,only:c_vertical_tab
use,intrinsic::iso_c_binding,only:c_float128
use,intrinsic::iso_c_binding,only:c_float128_complex
use,intrinsic::iso_c_binding,only:c_uint8_t
use,intrinsic::iso_c_binding,only:c_uint16_t
use,intrinsic::iso_c_binding,only:c_uint32_t
use,intrinsic::iso_c_binding,only:c_uint64_t
use,intrinsic::iso_c_binding,only:c_uint128_t
use,intrinsic::iso_c_binding,only:c_unsigned_char
use,intrinsic::iso_c_binding,only:c_unsigned_short
use,intrinsic::iso_c_binding,only:c_unsigned
use,intrinsic::iso_c_binding,only:c_unsigned_long
use,intrinsic::iso_c_binding,only:c_unsigned_long_long
use,intrinsic::iso_c_binding,only:c_uintmax_t
use,intrinsic::iso_c_binding,only:c_uint_fast8_t
use,intrinsic::iso_c_binding,only:c_uint_fast16_t
use,intrinsic::iso_c_binding,only:c_uint_fast32_t
use,intrinsic::iso_c_binding,only:c_uint_fast64_t
use,intrinsic::iso_c_binding,only:c_uint_fast128_t
use,intrinsic::iso_c_binding,only:c_uint_least8_t
use,intrinsic::iso_c_binding,only:c_uint_least16_t
use,intrinsic::iso_c_binding,only:c_uint_least32_t
use,intrinsic::iso_c_binding,only:c_uint_least64_t
use,intrinsic::iso_c_binding,only:c_uint_least128_t
use,intrinsic::iso_c_binding,only:c_f_procpointer
use exc_lapw_c,only:exc_ok
use exc_lapw_c,only:exc_err_arg
use exc_lapw_c,only:exc_err_hip
use exc_lapw_c,only:exc_err_unsupported
use exc_lapw_c,only:exc_err_state
use exc_lapw_c,only:exc_dims
use exc_lapw_c,only:exc_ctx_create
use exc_lapw_c,only:exc_ctx_destroy
use exc_lapw_c,only:exc_ctx_set_atoms
use exc_lapw_c,only:exc_ctx_set_lo
use exc_lapw_c,only:exc_ctx_set_radial
use exc_lapw_c,only:exc_ctx_set_mt
use exc_lapw_c,only:exc_ctx_set_gvec
use exc_lapw_c,only:exc_match
use exc_lapw_c,only:exc_hamiltonandoverlapsetup
use exc_lapw_c,only:exc_ho_vjp
use exc_lapw_c,only:exc_eigen
use exc_lapw_c,only:exc_ctx_nlotot
use exc_evsystem,only:hermitianmatrix
use exc_evsystem,only:null
use exc_evsystem,only:evsystem
use exc_evsystem,only:newsystem
use exc_evsystem,only:deletesystem
use,intrinsic::__fortran_builtins,only:iso_c_binding$__fortran_builtins$c_associated_c_ptr=>c_associated_c_ptr
private::c_associated
private::c_funloc
private::c_funptr
private::c_f_pointer
private::c_loc
private::c_null_funptr
private::c_null_ptr
private::c_ptr
private::c_sizeof
private::operator(==)
private::operator(/=)
private::c_int8_t
private::c_int16_t
private::c_int32_t
private::c_int64_t
private::c_int128_t
private::c_int
private::c_short
private::c_long
private::c_long_long
private::c_signed_char
private::c_size_t
private::c_intmax_t
private::c_intptr_t
private::c_ptrdiff_t
private::c_int_least8_t
private::c_int_fast8_t
private::c_int_least16_t
private::c_int_fast16_t
private::c_int_least32_t
private::c_int_fast32_t
private::c_int_least64_t
private::c_int_fast64_t
private::c_int_least128_t
private::c_int_fast128_t
private::c_float
private::c_double
private::c_long_double
private::c_float_complex
private::c_double_complex
private::c_long_double_complex
private::c_bool
private::c_char
private::c_null_char
private::c_alert
private::c_backspace
private::c_form_feed
private::c_new_line
private::c_carriage_return
private::c_horizontal_tab
private::c_vertical_tab
private::c_float128
private::c_float128_complex
private::c_uint8_t
private::c_uint16_t
private::c_uint32_t
private::c_uint64_t
private::c_uint128_t
private::c_unsigned_char
private::c_unsigned_short
private::c_unsigned
private::c_unsigned_long
private::c_unsigned_long_long
private::c_uintmax_t
private::c_uint_fast8_t
private::c_uint_fast16_t
private::c_uint_fast32_t
private::c_uint_fast64_t
private::c_uint_fast128_t
private::c_uint_least8_t
private::c_uint_least16_t
private::c_uint_least32_t
private::c_uint_least64_t
private::c_uint_least128_t
private::c_f_procpointer
private::exc_ok
private::exc_err_arg
private::exc_err_hip
private::exc_err_unsupported
private::exc_err_state
private::exc_dims
private::exc_ctx_create
private::exc_ctx_destroy
private::exc_ctx_set_atoms
private::exc_ctx_set_lo
private::exc_ctx_set_radial
private::exc_ctx_set_mt
private::exc_ctx_set_gvec
private::exc_match
private::exc_hamiltonandoverlapsetup
private::exc_ho_vjp
private::exc_eigen
private::exc_ctx_nlotot
private::hermitianmatrix
private::null
private::evsystem
private::newsystem
private::deletesystem
private::iso_c_binding$__fortran_builtins$c_associated_c_ptr
type(c_ptr),private,save::ctx
integer(4),private,save::ngkmax_
integer(4),private,save::lmaxapw_
integer(4),private,save::natmtot_
private::check
contains
subroutine check(rc,where)
integer(4),intent(in)::rc
character(*,1),intent(in)::where
end
subroutine exc_gpu_init(dims,ngkmax,natoms,rmt,atposc,nlorb,lorbl,idxlo,dmat,haa,loa,lolo,oalo,ololo,ivg,ivgig,cfunig,veffig,ffacg)
type(exc_dims),intent(in)::dims
integer(4),intent(in)::ngkmax
integer(4),intent(in)::natoms(1_8:*)
real(8),intent(in)::rmt(1_8:*)
real(8),intent(in)::atposc(1_8:*)
integer(4),intent(in)::nlorb(1_8:*)
integer(4),intent(in)::lorbl(1_8:*)
integer(4),intent(in)::idxlo(1_8:*)
real(8),intent(in)::dmat(1_8:*)
complex(8),intent(in)::haa(1_8:*)
complex(8),intent(in)::loa(1_8:*)
complex(8),intent(in)::lolo(1_8:*)
real(8),intent(in)::oalo(1_8:*)
real(8),intent(in)::ololo(1_8:*)
integer(4),intent(in)::ivg(1_8:*)
integer(4),intent(in)::ivgig(1_8:*)
complex(8),intent(in)::cfunig(1_8:*)
complex(8),intent(in)::veffig(1_8:*)
real(8),intent(in)::ffacg(1_8:*)
end
subroutine exc_gpu_finalize()
end
function exc_gpu_nlotot()
integer(4)::exc_gpu_nlotot
end
subroutine match(ngp,gpc,tpgpc,sfacgp,apwalm)
integer(4),intent(in)::ngp
real(8),intent(in)::gpc(1_8:int(ngkmax_,kind=8))
real(8),intent(in)::tpgpc(1_8:2_8,1_8:int(ngkmax_,kind=8))
complex(8),intent(in)::sfacgp(1_8:int(ngkmax_,kind=8),1_8:*)
complex(8),intent(inout)::apwalm(1_8:int(ngkmax_,kind=8),1_8:*)
end
subroutine hamiltonandoverlapsetup(system,ngp,apwalm,igpig,vgpc)
type(evsystem),intent(inout)::system
integer(4),intent(in)::ngp
complex(8),intent(in)::apwalm(1_8:int(ngkmax_,kind=8),1_8:*)
integer(4),intent(in)::igpig(1_8:int(ngkmax_,kind=8))
real(8),intent(in)::vgpc(1_8:3_8,1_8:int(ngkmax_,kind=8))
end
subroutine forcek_gpu(ngp,apwalm,igpig,vgpc,sfacgp,nmatmax,nstfv,evecfv,evalfv,occ,wkpt,forceibs)
integer(4),intent(in)::ngp
complex(8),intent(in)::apwalm(1_8:int(ngkmax_,kind=8),1_8:*)
integer(4),intent(in)::igpig(1_8:int(ngkmax_,kind=8))
real(8),intent(in)::vgpc(1_8:3_8,1_8:int(ngkmax_,kind=8))
complex(8),intent(in)::sfacgp(1_8:int(ngkmax_,kind=8),1_8:*)
integer(4),intent(in)::nmatmax
integer(4),intent(in)::nstfv
complex(8),intent(in)::evecfv(1_8:int(nmatmax,kind=8),1_8:int(nstfv,kind=8))
real(8),intent(in)::evalfv(1_8:int(nstfv,kind=8))
real(8),intent(in)::occ(1_8:int(nstfv,kind=8))
real(8),intent(in)::wkpt
real(8),intent(inout)::forceibs(1_8:3_8,1_8:int(natmtot_,kind=8))
end
subroutine solvewithgpu(system,nstfv,evecfv,evalfv)
type(evsystem),intent(inout)::system
integer(4),intent(in)::nstfv
complex(8),intent(out)::evecfv(:,:)
real(8),intent(out)::evalfv(1_8:int(nstfv,kind=8))
end
end
