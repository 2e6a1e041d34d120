!=======================================================================
! dac_io -- reading and writing CANS DAC files (the format of
! cansio/dacdef*.f / dacopnr*.f) for the Fortran main programs.
! Records: 1, 0, mtype (6 = REAL*8), ndim, dims (-1 = time axis), then
! one record per snapshot.
!=======================================================================
module dac_io
  implicit none
contains

  ! header of a DAC file; returns the spatial dims (time axis dropped)
  subroutine dac_header(iu, file, dims, nd)
    integer, intent(in) :: iu
    character(*), intent(in) :: file
    integer, intent(out) :: dims(4), nd
    integer :: one, ver, mtype, ndim, raw(5), ios
    open(iu, file=file, form='unformatted', status='old', iostat=ios)
    if (ios /= 0) then
       write(6, *) 'dac_io: cannot open ', trim(file)
       error stop 10001
    end if
    read(iu) one
    read(iu) ver
    read(iu) mtype
    read(iu) ndim
    if (one /= 1 .or. mtype /= 6 .or. ndim < 1 .or. ndim > 5) then
       write(6, *) 'dac_io: unsupported header in ', trim(file)
       error stop 10001
    end if
    read(iu) raw(1:ndim)
    nd = ndim
    if (raw(ndim) == -1) nd = ndim - 1
    dims = 1
    dims(1:nd) = raw(1:nd)
  end subroutine dac_header

  ! last snapshot of a REAL*8 DAC file into a (n doubles)
  subroutine dac_read_last(iu, file, a, n)
    integer, intent(in) :: iu, n
    character(*), intent(in) :: file
    double precision, intent(out) :: a(n)
    integer :: dims(4), nd, ios, got
    call dac_header(iu, file, dims, nd)
    if (product(dims(1:max(nd, 1))) /= n) then
       write(6, *) 'dac_io: size mismatch in ', trim(file)
       error stop 10001
    end if
    got = 0
    do
       read(iu, iostat=ios) a
       if (ios /= 0) exit
       got = got + 1
    end do
    close(iu)
    if (got == 0) then
       write(6, *) 'dac_io: no snapshot in ', trim(file)
       error stop 10001
    end if
    ! the failed read at the end of file leaves a undefined: read again
    call dac_header(iu, file, dims, nd)
    do while (got > 0)
       read(iu) a
       got = got - 1
    end do
    close(iu)
  end subroutine dac_read_last

  subroutine dac_dims(iu, file, dims, nd)
    integer, intent(in) :: iu
    character(*), intent(in) :: file
    integer, intent(out) :: dims(4), nd
    call dac_header(iu, file, dims, nd)
    close(iu)
  end subroutine dac_dims

  ! a time-series ("s") file with one snapshot
  subroutine dac_write_series(iu, file, a, dims, nd)
    integer, intent(in) :: iu, nd, dims(nd)
    character(*), intent(in) :: file
    double precision, intent(in) :: a(*)
    integer :: n
    n = product(dims)
    open(iu, file=file, form='unformatted', status='replace')
    write(iu) 1
    write(iu) 0
    write(iu) 6
    write(iu) nd + 1
    write(iu) dims, -1
    write(iu) a(1:n)
    close(iu)
  end subroutine dac_write_series

  subroutine dac_write_scalar_series(iu, file, v)
    integer, intent(in) :: iu
    character(*), intent(in) :: file
    double precision, intent(in) :: v
    open(iu, file=file, form='unformatted', status='replace')
    write(iu) 1
    write(iu) 0
    write(iu) 6
    write(iu) 1
    write(iu) -1
    write(iu) v
    close(iu)
  end subroutine dac_write_scalar_series

end module dac_io
