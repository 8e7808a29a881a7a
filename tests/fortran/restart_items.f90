! Writes two files in the item layout of FESOM's t_partit and t_tracer
! derived-type restarts (one parent unformatted write per file; the type's
! user-defined output procedure emits the items: scalars, fixed-length
! integer arrays with their extent, allocatable arrays as extents then data,
! unallocated ones as zero extents, logical and character(20) scalars), so a
! test can check that the library's restart codec reads and rewrites what a
! Fortran compiler writes, byte for byte.
module restart_items
  implicit none
  type partit_like
  contains
    procedure :: put_partit
    generic :: write(unformatted) => put_partit
  end type
  type tracer_like
  contains
    procedure :: put_tracer
    generic :: write(unformatted) => put_tracer
  end type
contains
  subroutine ints1(u, a)
    integer, intent(in) :: u, a(:)
    write(u) size(a)
    write(u) a
  end subroutine
  subroutine none(u, rank)
    integer, intent(in) :: u, rank
    integer :: k
    do k = 1, rank
      write(u) 0
    end do
  end subroutine
  subroutine com(u, with_list)
    integer, intent(in) :: u
    logical, intent(in) :: with_list
    integer :: fixed32(32), fixed33(33)
    fixed32 = 0; fixed33 = 0; fixed33(1) = 1
    write(u) 2
    call ints1(u, fixed32)
    call ints1(u, fixed33)
    if (with_list) then
      call ints1(u, [4, 7, 9])
    else
      call none(u, 1)
    end if
    write(u) 0
    call ints1(u, fixed32)
    call ints1(u, fixed32(1:32))
    call none(u, 1)
    write(u) 3
  end subroutine
  subroutine put_partit(x, unit, iostat, iomsg)
    class(partit_like), intent(in) :: x
    integer, intent(in) :: unit
    integer, intent(out) :: iostat
    character(*), intent(inout) :: iomsg
    integer :: k
    call com(unit, .true.)
    call com(unit, .false.)
    call com(unit, .false.)
    write(unit) 1, 0, 100
    call ints1(unit, [0, 0, 0, 0, 0])
    write(unit) 5, 0
    call ints1(unit, [(k, k = 1, 5)])
    write(unit) 4, 0, 0
    call ints1(unit, [(k, k = 1, 4)])
    write(unit) 8, 0
    call ints1(unit, [(k, k = 1, 8)])
    write(unit) 0
    iostat = 0
  end subroutine
  subroutine put_tracer(x, unit, iostat, iomsg)
    class(tracer_like), intent(in) :: x
    integer, intent(in) :: unit
    integer, intent(out) :: iostat
    character(*), intent(inout) :: iomsg
    real(8) :: v(3, 5)
    integer :: t, k, e(2, 8)
    character(20) :: hor, ver, lim
    write(unit) 2
    do t = 1, 2
      v = reshape([(0.25d0 * k + t, k = 1, 15)], [3, 5])
      write(unit) 3, 5
      write(unit) v
      call none(unit, 3)
      write(unit) 3, 5
      write(unit) v - 1.0d0
      write(unit) .false., 0.5d0, 1.5d0, -2.5d0, (t == 2)
      hor = 'MUSCL'; ver = 'QR4C'; lim = 'FCT'
      write(unit) hor, ver, lim
      write(unit) 1.0d0, 0.75d0, t
    end do
    do k = 1, 13
      if (k == 6) then
        write(unit) 3, 5
        write(unit) v * 2.0d0
      else if (k == 4 .or. k == 5 .or. k == 13) then
        call none(unit, 3)
      else
        call none(unit, 2)
      end if
    end do
    call ints1(unit, [3, 3, 2, 3, 1])
    e = reshape([(mod(k, 4), k = 1, 16)], [2, 8])
    write(unit) 2, 8
    write(unit) e
    iostat = 0
  end subroutine
end module

program write_items
  use restart_items
  implicit none
  type(partit_like) :: p
  type(tracer_like) :: t
  integer :: u
  open(newunit=u, file='t_partit.0', form='unformatted', status='replace')
  write(u) p
  close(u)
  open(newunit=u, file='t_tracer.0', form='unformatted', status='replace')
  write(u) t
  close(u)
end program
