! Round trip of rb_records: writes the records the test asks for with a given subrecord limit,
! then reads back a file written by the test's Python writer and prints its checksums.
!   records_check write <file> <max_sub> <n1> <n2> ...   (record k: n_k doubles, value i + 1000 k)
!   records_check read <file> <n1> <n2> ...
program records_check
  use iso_c_binding
  use rb_records
  implicit none
  character(len=256) :: mode, path, arg
  integer :: nrec, k, i, narg
  integer(8) :: max_sub, n
  real(8), allocatable, target :: a(:)
  type(rec_writer) :: w
  type(rec_reader) :: r
  narg = command_argument_count()
  call get_command_argument(1, mode)
  call get_command_argument(2, path)
  if (trim(mode) == 'write') then
    call get_command_argument(3, arg)
    read(arg, *) max_sub
    open(unit=11, file=trim(path), access='stream', form='unformatted', status='replace')
    do k = 4, narg
      call get_command_argument(k, arg)
      read(arg, *) n
      allocate(a(n))
      do i = 1, int(n)
        a(i) = dble(i) + 1000.0d0 * dble(k - 3)
      enddo
      ! two puts per record, split unevenly, so that the subrecord boundaries fall inside puts
      call rec_begin(w, 11, 8_8 * n, max_sub)
      call rec_put_f64(w, a(1:n / 3))
      call rec_put_f64(w, a(n / 3 + 1:n))
      call rec_end(w)
      deallocate(a)
    enddo
    close(11)
  else
    open(unit=11, file=trim(path), access='stream', form='unformatted', status='old')
    do k = 3, narg
      call get_command_argument(k, arg)
      read(arg, *) n
      allocate(a(n))
      call rec_open(r, 11)
      call rec_get_f64(r, a(1:n / 2))
      call rec_get_f64(r, a(n / 2 + 1:n))
      call rec_close(r)
      write(*, '(i0, 1x, es25.17)') k - 2, sum(a)
      deallocate(a)
    enddo
    close(11)
  endif
end program records_check
