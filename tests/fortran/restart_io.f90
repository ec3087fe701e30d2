! Test helper: writes or reads a velocity restart file with the record layout
! of SAVE_FLOW / READ_FLOW (src/diablo.F90:811-819, 946-966): a header record
! NX, NY, NZ, TIME, TIME_STEP and one record of CU1, CU2, CU3, CP as
! COMPLEX*16 over (0:NX/2, 0:NZ+1, 0:NY+1), kx fastest.  The field values
! follow a closed formula so the other side can check them.
!   restart_io w FILE NX NY NZ     write the formula fields, TIME 1.25, TIME_STEP 42
!   restart_io r FILE NX NY NZ     read and compare; exit status 1 on mismatch
program restart_io
  implicit none
  character(len=256) :: mode, fname, arg
  integer :: nx, ny, nz, nx2p, i, k, j, f, ts, nxt, nyt, nzt, bad
  real(8) :: t
  complex(8), allocatable :: cu(:,:,:,:)
  call get_command_argument(1, mode)
  call get_command_argument(2, fname)
  call get_command_argument(3, arg); read(arg, *) nx
  call get_command_argument(4, arg); read(arg, *) ny
  call get_command_argument(5, arg); read(arg, *) nz
  nx2p = nx / 2
  allocate(cu(0:nx2p, 0:nz+1, 0:ny+1, 4))
  if (mode(1:1) == 'w') then
    do f = 1, 4
      do j = 0, ny + 1
        do k = 0, nz + 1
          do i = 0, nx2p
            cu(i, k, j, f) = expected(i, k, j, f)
          end do
        end do
      end do
    end do
    open(unit=10, file=trim(fname), status='unknown', form='unformatted')
    write(10) nx, ny, nz, 1.25d0, 42
    write(10) (((cu(i, k, j, 1), i=0,nx2p), k=0,nz+1), j=0,ny+1), &
              (((cu(i, k, j, 2), i=0,nx2p), k=0,nz+1), j=0,ny+1), &
              (((cu(i, k, j, 3), i=0,nx2p), k=0,nz+1), j=0,ny+1), &
              (((cu(i, k, j, 4), i=0,nx2p), k=0,nz+1), j=0,ny+1)
    close(10)
  else
    open(unit=10, file=trim(fname), status='old', form='unformatted')
    read(10) nxt, nyt, nzt, t, ts
    if (nxt /= nx .or. nyt /= ny .or. nzt /= nz .or. t /= 1.25d0 .or. ts /= 42) then
      write(*, *) 'header mismatch', nxt, nyt, nzt, t, ts
      stop 1
    end if
    read(10) (((cu(i, k, j, 1), i=0,nx2p), k=0,nz+1), j=0,ny+1), &
             (((cu(i, k, j, 2), i=0,nx2p), k=0,nz+1), j=0,ny+1), &
             (((cu(i, k, j, 3), i=0,nx2p), k=0,nz+1), j=0,ny+1), &
             (((cu(i, k, j, 4), i=0,nx2p), k=0,nz+1), j=0,ny+1)
    close(10)
    bad = 0
    do f = 1, 4
      do j = 0, ny + 1
        do k = 0, nz + 1
          do i = 0, nx2p
            if (cu(i, k, j, f) /= expected(i, k, j, f)) bad = bad + 1
          end do
        end do
      end do
    end do
    if (bad /= 0) then
      write(*, *) 'field mismatch at', bad, 'points'
      stop 1
    end if
    write(*, *) 'ok'
  end if
contains
  complex(8) function expected(i, k, j, f)
    integer, intent(in) :: i, k, j, f
    expected = cmplx(dble(i) + 1000.d0 * k + 1.d6 * j + 0.25d0 * f, &
                     -(dble(f) + 0.5d0 * i - 7.d0 * k + 3.d0 * j), kind=8)
  end function expected
end program restart_io
