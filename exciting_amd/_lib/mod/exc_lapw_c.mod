﻿!mod$ v1 sum:f71f194d1991c68b
!need$ 0bde2ac47243ead2 i iso_c_binding
module exc_lapw_c
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
integer(4),parameter::exc_ok=0_4
integer(4),parameter::exc_err_arg=-1_4
integer(4),parameter::exc_err_hip=-2_4
integer(4),parameter::exc_err_unsupported=-3_4
integer(4),parameter::exc_err_state=-4_4
type,bind(c)::exc_dims
integer(4)::natmtot
integer(4)::nspecies
integer(4)::lmaxapw
integer(4)::lmaxmat
integer(4)::apwordmax
integer(4)::nlomax
integer(4)::lolmmax
integer(4)::maxnlo
integer(4)::ngrtot
integer(4)::ngvec
integer(4)::intgv(1_8:6_8)
real(8)::omega
end type
interface
function exc_ctx_create(dims,ctx) bind(c,name="exc_ctx_create")
import::c_ptr
import::exc_dims
type(exc_dims),intent(in)::dims
type(c_ptr),intent(out)::ctx
integer(4)::exc_ctx_create
end
end interface
interface
function exc_ctx_destroy(ctx) bind(c,name="exc_ctx_destroy")
import::c_ptr
type(c_ptr),value::ctx
integer(4)::exc_ctx_destroy
end
end interface
interface
function exc_ctx_set_atoms(ctx,natoms,rmt,atposc) bind(c,name="exc_ctx_set_atoms")
import::c_ptr
type(c_ptr),value::ctx
integer(4),intent(in)::natoms(1_8:*)
real(8),intent(in)::rmt(1_8:*)
real(8),intent(in)::atposc(1_8:*)
integer(4)::exc_ctx_set_atoms
end
end interface
interface
function exc_ctx_set_lo(ctx,nlorb,lorbl,idxlo) bind(c,name="exc_ctx_set_lo")
import::c_ptr
type(c_ptr),value::ctx
integer(4),intent(in)::nlorb(1_8:*)
integer(4),intent(in)::lorbl(1_8:*)
integer(4),intent(in)::idxlo(1_8:*)
integer(4)::exc_ctx_set_lo
end
end interface
interface
function exc_ctx_set_radial(ctx,dmat) bind(c,name="exc_ctx_set_radial")
import::c_ptr
type(c_ptr),value::ctx
real(8),intent(in)::dmat(1_8:*)
integer(4)::exc_ctx_set_radial
end
end interface
interface
function exc_ctx_set_mt(ctx,haa,loa,lolo,oalo,ololo) bind(c,name="exc_ctx_set_mt")
import::c_ptr
type(c_ptr),value::ctx
complex(8),intent(in)::haa(1_8:*)
complex(8),intent(in)::loa(1_8:*)
complex(8),intent(in)::lolo(1_8:*)
real(8),intent(in)::oalo(1_8:*)
real(8),intent(in)::ololo(1_8:*)
integer(4)::exc_ctx_set_mt
end
end interface
interface
function exc_ctx_set_gvec(ctx,ivg,ivgig,cfunig,veffig,ffacg) bind(c,name="exc_ctx_set_gvec")
import::c_ptr
type(c_ptr),value::ctx
integer(4),intent(in)::ivg(1_8:*)
integer(4),intent(in)::ivgig(1_8:*)
complex(8),intent(in)::cfunig(1_8:*)
complex(8),intent(in)::veffig(1_8:*)
real(8),intent(in)::ffacg(1_8:*)
integer(4)::exc_ctx_set_gvec
end
end interface
interface
function exc_match(ctx,ngp,ngkmax,gpc,tpgpc,sfacgp,apwalm) bind(c,name="exc_match")
import::c_ptr
type(c_ptr),value::ctx
integer(4),value::ngp
integer(4),value::ngkmax
real(8),intent(in)::gpc(1_8:*)
real(8),intent(in)::tpgpc(1_8:*)
complex(8),intent(in)::sfacgp(1_8:*)
complex(8),intent(inout)::apwalm(1_8:*)
integer(4)::exc_match
end
end interface
interface
function exc_hamiltonandoverlapsetup(ctx,ngp,ngkmax,apwalm,igpig,vgpc,ldh,h,o) bind(c,name="exc_hamiltonandoverlapsetup")
import::c_ptr
type(c_ptr),value::ctx
integer(4),value::ngp
integer(4),value::ngkmax
complex(8),intent(in)::apwalm(1_8:*)
integer(4),intent(in)::igpig(1_8:*)
real(8),intent(in)::vgpc(1_8:*)
integer(4),value::ldh
complex(8),intent(inout)::h(1_8:*)
complex(8),intent(inout)::o(1_8:*)
integer(4)::exc_hamiltonandoverlapsetup
end
end interface
interface
function exc_ho_vjp(ctx,ngp,ngkmax,apwalm,igpig,vgpc,sfacgp,nst,ldc,evec,wh,wo,dedtau,dedv) bind(c,name="exc_ho_vjp")
import::c_ptr
type(c_ptr),value::ctx
integer(4),value::ngp
integer(4),value::ngkmax
complex(8),intent(in)::apwalm(1_8:*)
integer(4),intent(in)::igpig(1_8:*)
real(8),intent(in)::vgpc(1_8:*)
complex(8),intent(in)::sfacgp(1_8:*)
integer(4),value::nst
integer(4),value::ldc
complex(8),intent(in)::evec(1_8:*)
real(8),intent(in)::wh(1_8:*)
real(8),intent(in)::wo(1_8:*)
real(8),intent(inout)::dedtau(1_8:*)
type(c_ptr),value::dedv
integer(4)::exc_ho_vjp
end
end interface
interface
function exc_eigen(ctx,nmat,ldh,h,o,nev,eval,evec,ldv) bind(c,name="exc_eigen")
import::c_ptr
type(c_ptr),value::ctx
integer(4),value::nmat
integer(4),value::ldh
complex(8),intent(inout)::h(1_8:*)
complex(8),intent(inout)::o(1_8:*)
integer(4),value::nev
real(8),intent(inout)::eval(1_8:*)
complex(8),intent(inout)::evec(1_8:*)
integer(4),value::ldv
integer(4)::exc_eigen
end
end interface
interface
function exc_ctx_nlotot(ctx) bind(c,name="exc_ctx_nlotot")
import::c_ptr
type(c_ptr),value::ctx
integer(4)::exc_ctx_nlotot
end
end interface
end
