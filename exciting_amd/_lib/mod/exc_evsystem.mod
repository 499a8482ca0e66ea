﻿!mod$ v1 sum:4e11352129309ed7
module exc_evsystem
type::hermitianmatrix
integer(4)::rank=0_4
logical(4)::packed=.false._4
complex(8),pointer::za(:,:)=>NULL()
end type
intrinsic::null
type::evsystem
type(hermitianmatrix)::hamilton
type(hermitianmatrix)::overlap
end type
contains
subroutine newsystem(self,packed,rank)
type(evsystem),intent(out)::self
logical(4),intent(in)::packed
integer(4),intent(in)::rank
end
subroutine deletesystem(self)
type(evsystem),intent(inout)::self
end
end
