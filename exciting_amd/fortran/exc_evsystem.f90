! The eigensystem container the setup writes into, with the components the setup uses of
! the reference's HermitianMatrix / evsystem types (src_eigensystem/modfvsystem.f90:15-29):
! a full (non-packed) complex(8) matrix za(rank, rank) for H and for O.
module exc_evsystem
  implicit none

  type HermitianMatrix
    integer :: rank = 0
    logical :: packed = .false.
    complex(8), pointer :: za(:, :) => null()
  end type HermitianMatrix

  type evsystem
    type(HermitianMatrix) :: hamilton, overlap
  end type evsystem

contains

  ! newsystem (modfvsystem.f90:92-101): allocate and zero H and O of rank nmat
  subroutine newsystem(self, packed, rank)
    type(evsystem), intent(out) :: self
    logical, intent(in) :: packed
    integer, intent(in) :: rank
    if (packed) error stop "exc_evsystem: packed storage is not produced by the GPU setup"
    self%hamilton%rank = rank
    self%overlap%rank = rank
    allocate(self%hamilton%za(rank, rank), self%overlap%za(rank, rank))
    self%hamilton%za = (0.d0, 0.d0)
    self%overlap%za = (0.d0, 0.d0)
  end subroutine newsystem

  subroutine deletesystem(self)
    type(evsystem), intent(inout) :: self
    if (associated(self%hamilton%za)) deallocate(self%hamilton%za)
    if (associated(self%overlap%za)) deallocate(self%overlap%za)
  end subroutine deletesystem
end module exc_evsystem
