﻿!mod$ v1 sum:52b4a804ac1435a3
!need$ 0bde2ac47243ead2 i iso_c_binding
!need$ bb381bf46e508468 i __fortran_builtins
module hm_bem
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
private::iso_c_binding$__fortran_builtins$c_associated_c_ptr
integer(4),parameter::hm_fp64=0_4
integer(4),parameter::hm_fp32_ir=1_4
interface
function hm_create(ctx) bind(c,name="hm_create")
import::c_ptr
type(c_ptr),intent(out)::ctx
integer(4)::hm_create
end
end interface
interface
function hm_destroy(ctx) bind(c,name="hm_destroy")
import::c_ptr
type(c_ptr),value::ctx
integer(4)::hm_destroy
end
end interface
interface
function hm_set_constants(ctx,rho,g,g_disp,amp,xw,yw) bind(c,name="hm_set_constants")
import::c_ptr
type(c_ptr),value::ctx
real(8),value::rho
real(8),value::g
real(8),value::g_disp
real(8),value::amp
real(8),value::xw
real(8),value::yw
integer(4)::hm_set_constants
end
end interface
interface
function hm_set_geometry(ctx,nbody,body_start,ntnd,xyz,ncn,ncon,xyz_p,ds,pnsz,dxyz_p,ni,ixyz_p) bind(c,name="hm_set_geometry")
import::c_ptr
type(c_ptr),value::ctx
integer(4),value::nbody
integer(4),intent(in)::body_start(1_8:*)
integer(4),value::ntnd
real(8),intent(in)::xyz(1_8:*)
integer(4),intent(in)::ncn(1_8:*)
integer(4),intent(in)::ncon(1_8:*)
real(8),intent(in)::xyz_p(1_8:*)
real(8),intent(in)::ds(1_8:*)
real(8),intent(in)::pnsz(1_8:*)
real(8),intent(in)::dxyz_p(1_8:*)
integer(4),value::ni
real(8),intent(in)::ixyz_p(1_8:*)
integer(4)::hm_set_geometry
end
end interface
interface
function hm_set_depth(ctx,h) bind(c,name="hm_set_depth")
import::c_ptr
type(c_ptr),value::ctx
real(8),value::h
integer(4)::hm_set_depth
end
end interface
interface
function hm_set_headings(ctx,nbeta,beta) bind(c,name="hm_set_headings")
import::c_ptr
type(c_ptr),value::ctx
integer(4),value::nbeta
real(8),intent(in)::beta(1_8:*)
integer(4)::hm_set_headings
end
end interface
interface
function hm_set_frequency(ctx,w1) bind(c,name="hm_set_frequency")
import::c_ptr
type(c_ptr),value::ctx
real(8),value::w1
integer(4)::hm_set_frequency
end
end interface
interface
function hm_assemble(ctx) bind(c,name="hm_assemble")
import::c_ptr
type(c_ptr),value::ctx
integer(4)::hm_assemble
end
end interface
interface
function hm_factor(ctx,precision) bind(c,name="hm_factor")
import::c_ptr
type(c_ptr),value::ctx
integer(4),value::precision
integer(4)::hm_factor
end
end interface
interface
function hm_solve(ctx,iters) bind(c,name="hm_solve")
import::c_ptr
type(c_ptr),value::ctx
integer(4),intent(out)::iters
integer(4)::hm_solve
end
end interface
interface
function hm_forces(ctx,amas,bdmp,exfc) bind(c,name="hm_forces")
import::c_ptr
type(c_ptr),value::ctx
real(8),intent(out)::amas(1_8:*)
real(8),intent(out)::bdmp(1_8:*)
complex(8),intent(out)::exfc(1_8:*)
integer(4)::hm_forces
end
end interface
interface
function hm_get_system(ctx,amat,smat,rhs) bind(c,name="hm_get_system")
import::c_ptr
type(c_ptr),value::ctx
type(c_ptr),value::amat
type(c_ptr),value::smat
type(c_ptr),value::rhs
integer(4)::hm_get_system
end
end interface
interface
function hm_get_potentials(ctx,mxpot) bind(c,name="hm_get_potentials")
import::c_ptr
type(c_ptr),value::ctx
complex(8),intent(out)::mxpot(1_8:*)
integer(4)::hm_get_potentials
end
end interface
interface
function hm_fingreen3d_batch(ctx,n,xf,xp,yf,yp,zf,zp,v,wvn,nk,h,flag,grn) bind(c,name="hm_fingreen3d_batch")
import::c_ptr
type(c_ptr),value::ctx
integer(4),value::n
real(8),intent(in)::xf(1_8:*)
real(8),intent(in)::xp(1_8:*)
real(8),intent(in)::yf(1_8:*)
real(8),intent(in)::yp(1_8:*)
real(8),intent(in)::zf(1_8:*)
real(8),intent(in)::zp(1_8:*)
real(8),value::v
real(8),intent(in)::wvn(1_8:*)
integer(4),value::nk
real(8),value::h
integer(4),intent(in)::flag(1_8:*)
complex(8),intent(out)::grn(1_8:*)
integer(4)::hm_fingreen3d_batch
end
end interface
interface
function hm_infgreen3d_batch(ctx,n,xf,xp,yf,yp,zf,zp,v,flag,grn) bind(c,name="hm_infgreen3d_batch")
import::c_ptr
type(c_ptr),value::ctx
integer(4),value::n
real(8),intent(in)::xf(1_8:*)
real(8),intent(in)::xp(1_8:*)
real(8),intent(in)::yf(1_8:*)
real(8),intent(in)::yp(1_8:*)
real(8),intent(in)::zf(1_8:*)
real(8),intent(in)::zp(1_8:*)
real(8),value::v
integer(4),intent(in)::flag(1_8:*)
complex(8),intent(out)::grn(1_8:*)
integer(4)::hm_infgreen3d_batch
end
end interface
interface
function hm_zgetrs(ctx,trans,n,nrhs,a,lda,ipiv,b,ldb,info) bind(c,name="hm_zgetrs")
import::c_ptr
type(c_ptr),value::ctx
character(1_8,1),value::trans
integer(4),value::n
integer(4),value::nrhs
complex(8),intent(in)::a(1_8:*)
integer(4),value::lda
integer(4),intent(in)::ipiv(1_8:*)
complex(8),intent(inout)::b(1_8:*)
integer(4),value::ldb
integer(4),intent(out)::info
integer(4)::hm_zgetrs
end
end interface
interface
function hm_zgetrf(ctx,m,n,a,lda,ipiv,info) bind(c,name="hm_zgetrf")
import::c_ptr
type(c_ptr),value::ctx
integer(4),value::m
integer(4),value::n
complex(8),intent(inout)::a(1_8:*)
integer(4),value::lda
integer(4),intent(out)::ipiv(1_8:*)
integer(4),intent(out)::info
integer(4)::hm_zgetrf
end
end interface
type(c_ptr),private,save::ctx_
integer(4),private,save::n_
integer(4),private,save::nbody_
integer(4),private,save::nbeta_
integer(4),allocatable,private,save::body_start_(:)
complex(8),allocatable,private,save,target::rhs_(:,:)
contains
function hm_context() result(ctx)
type(c_ptr)::ctx
end
subroutine hm_check(rc,what)
integer(4),intent(in)::rc
character(*,1),intent(in)::what
end
subroutine hm_setup(nbody,body_start,ntnd,xyz,ncn,ncon,xyz_p,ds,pnsz,dxyz_p,ni,ixyz_p,h,nbeta,beta,rho,g,g_disp,amp,xw)
integer(4),intent(in)::nbody
integer(4),intent(in)::body_start(1_8:int(nbody+1_4,kind=8))
integer(4),intent(in)::ntnd
real(8),intent(in)::xyz(:,:)
integer(4),intent(in)::ncn(:)
integer(4),intent(in)::ncon(:,:)
real(8),intent(in)::xyz_p(:,:)
real(8),intent(in)::ds(:)
real(8),intent(in)::pnsz(:)
real(8),intent(in)::dxyz_p(:,:)
integer(4),intent(in)::ni
real(8),intent(in)::ixyz_p(:,:)
real(8),intent(in)::h
integer(4),intent(in)::nbeta
real(8),intent(in)::beta(1_8:int(nbeta,kind=8))
real(8),intent(in)::rho
real(8),intent(in)::g
real(8),intent(in)::g_disp
real(8),intent(in)::amp
real(8),intent(in)::xw(1_8:2_8)
end
subroutine hm_set_frequency_f(w1)
real(8),intent(in)::w1
end
subroutine fingreen3d(xf,xp,yf,yp,zf,zp,v,wvn,nk,h,grn,flag)
real(8),intent(in)::xf
real(8),intent(in)::xp
real(8),intent(in)::yf
real(8),intent(in)::yp
real(8),intent(in)::zf
real(8),intent(in)::zp
real(8),intent(in)::v
integer(4),intent(in)::nk
real(8),intent(in)::wvn(1_8:int(nk,kind=8))
real(8),intent(in)::h
complex(8),intent(out)::grn(1_8:4_8)
integer(4),intent(in)::flag
end
subroutine infgreen3d(xf,xp,yf,yp,zf,zp,v,grn,flag)
real(8),intent(in)::xf
real(8),intent(in)::xp
real(8),intent(in)::yf
real(8),intent(in)::yp
real(8),intent(in)::zf
real(8),intent(in)::zp
real(8),intent(in)::v
complex(8),intent(out)::grn(1_8:4_8)
integer(4),intent(in)::flag
end
subroutine assb_left_multi(amat,ipiv,nelem,nsys)
integer(4),intent(in)::nelem
integer(4),intent(in)::nsys
complex(8),intent(out),target::amat(1_8:int(nelem,kind=8),1_8:int(nelem,kind=8),1_8:int(nsys,kind=8))
integer(4),intent(out)::ipiv(1_8:int(nelem,kind=8),1_8:int(nsys,kind=8))
end
subroutine assb_rbc_multi(brmat,nelem,nsys,nelem_start,nelem_end)
integer(4),intent(in)::nelem
integer(4),intent(in)::nsys
complex(8),intent(inout)::brmat(1_8:int(nelem,kind=8),1_8:6_8,1_8:int(nsys,kind=8))
integer(4),intent(in)::nelem_start
integer(4),intent(in)::nelem_end
end
subroutine assb_dbc_multi(bdmat,nelem,nsys,heading)
integer(4),intent(in)::nelem
integer(4),intent(in)::nsys
complex(8),intent(out)::bdmat(1_8:int(nelem,kind=8),1_8:int(nsys,kind=8))
integer(4),intent(in)::heading
end
subroutine radiation_solver_multi(amat,brmat,ipiv,mxpot,nelem,nsys)
integer(4),intent(in)::nelem
integer(4),intent(in)::nsys
complex(8),intent(in)::amat(1_8:int(nelem,kind=8),1_8:int(nelem,kind=8),1_8:int(nsys,kind=8))
complex(8),intent(in)::brmat(1_8:int(nelem,kind=8),1_8:int(6_4*nbody_,kind=8),1_8:int(nsys,kind=8))
integer(4),intent(in)::ipiv(1_8:int(nelem,kind=8),1_8:int(nsys,kind=8))
complex(8),intent(out)::mxpot(1_8:int(nelem,kind=8),1_8:int(6_4*nbody_+1_4,kind=8),1_8:int(nsys,kind=8))
end
subroutine diffraction_solver_multi(amat,bdmat,ipiv,mxpot,nelem,nsys)
integer(4),intent(in)::nelem
integer(4),intent(in)::nsys
complex(8),intent(in)::amat(1_8:int(nelem,kind=8),1_8:int(nelem,kind=8),1_8:int(nsys,kind=8))
complex(8),intent(in)::bdmat(1_8:int(nelem,kind=8),1_8:int(nsys,kind=8))
integer(4),intent(in)::ipiv(1_8:int(nelem,kind=8),1_8:int(nsys,kind=8))
complex(8),intent(inout)::mxpot(1_8:int(nelem,kind=8),1_8:int(6_4*nbody_+1_4,kind=8),1_8:int(nsys,kind=8))
end
subroutine rforce_eforce_multi(precision,amas,bdmp,exfc,iters)
integer(4),intent(in)::precision
real(8),intent(out)::amas(1_8:int(6_4*nbody_,kind=8),1_8:int(6_4*nbody_,kind=8))
real(8),intent(out)::bdmp(1_8:int(6_4*nbody_,kind=8),1_8:int(6_4*nbody_,kind=8))
complex(8),intent(out)::exfc(1_8:int(6_4*nbody_,kind=8),1_8:int(nbeta_,kind=8))
integer(4),intent(out)::iters
end
end
