!==================================================================================================
! rb_records -- Fortran unformatted sequential records in GNU Fortran's on-disk convention, written
! and read through stream I/O so that records of any length keep the layout the reference's own
! toolchain (gfortran, 2DRB/build_in_ascii_path.ps1) gives the restart and snapshot files of
! backupData / output_binary (2DRB/3DRB.F90:1599-1665, 2DRB/2DRB.F90:1357-1430).
!
! A logical record is one or more subrecords.  Each subrecord is a 4-byte marker, at most
! max_sub bytes of data (2147483639 = 2^31 - 9 by default), and a 4-byte marker; both markers
! hold the subrecord's data length.  The leading marker is negated when another subrecord
! follows, the trailing marker when one precedes.  A record of up to max_sub bytes is therefore
! the plain "length, data, length" form every Fortran compiler writes; longer records (a 512^3
! FP64 f record is 20.4 GB) split into subrecords, which a compiler writing one 4-byte or 8-byte
! marker per record would not produce.  Markers are little-endian int32 (the native order).
!==================================================================================================
module rb_records
  use iso_c_binding
  implicit none
  private
  public :: rec_writer, rec_reader, rec_begin, rec_put, rec_end, rec_open, rec_get, rec_close, &
            gfortran_max_subrecord, rec_put_f64, rec_get_f64

  integer(8), parameter :: gfortran_max_subrecord = 2147483639_8

  type rec_writer
    integer :: unit = -1
    integer(8) :: max_sub = gfortran_max_subrecord
    integer(8) :: total = 0, done = 0      ! record bytes, bytes written so far
    integer(8) :: sub_len = 0, sub_left = 0 ! current subrecord's length and bytes still to write
    logical :: first = .true.
  end type rec_writer

  type rec_reader
    integer :: unit = -1
    integer(8) :: sub_len = 0, sub_left = 0
    logical :: more = .false.               ! another subrecord follows the current one
    logical :: first = .true.
    integer(8) :: got = 0
  end type rec_reader

contains

  ! Start a record of `total` bytes on a unit opened with access='stream', form='unformatted'.
  subroutine rec_begin(w, unit, total, max_sub)
    type(rec_writer), intent(out) :: w
    integer, intent(in) :: unit
    integer(8), intent(in) :: total
    integer(8), intent(in), optional :: max_sub
    w%unit = unit
    if (present(max_sub)) w%max_sub = max_sub
    if (w%max_sub < 1 .or. w%max_sub > gfortran_max_subrecord) then
      write(*,*) 'rb_records: subrecord length out of range: ', w%max_sub
      stop 7
    endif
    w%total = total
    w%done = 0
    w%first = .true.
    call start_sub(w)
  end subroutine rec_begin

  subroutine start_sub(w)
    type(rec_writer), intent(inout) :: w
    w%sub_len = min(w%max_sub, w%total - w%done)
    w%sub_left = w%sub_len
    if (w%done + w%sub_len < w%total) then
      write(w%unit) int(-w%sub_len, 4)
    else
      write(w%unit) int(w%sub_len, 4)
    endif
  end subroutine start_sub

  subroutine end_sub(w)
    type(rec_writer), intent(inout) :: w
    if (w%first) then
      write(w%unit) int(w%sub_len, 4)
    else
      write(w%unit) int(-w%sub_len, 4)
    endif
    w%first = .false.
  end subroutine end_sub

  ! Append n bytes at address p to the record.
  subroutine rec_put(w, p, n)
    type(rec_writer), intent(inout) :: w
    type(c_ptr), intent(in) :: p
    integer(8), intent(in) :: n
    character(kind=c_char), pointer :: b(:)
    integer(8) :: pos, k
    if (n <= 0) return
    if (w%done + n > w%total) then
      write(*,*) 'rb_records: more data than the record holds'
      stop 7
    endif
    call c_f_pointer(p, b, [n])
    pos = 1
    do while (pos <= n)
      if (w%sub_left == 0) then
        call end_sub(w)
        call start_sub(w)
      endif
      k = min(w%sub_left, n - pos + 1)
      write(w%unit) b(pos:pos + k - 1)
      pos = pos + k
      w%sub_left = w%sub_left - k
      w%done = w%done + k
    enddo
  end subroutine rec_put

  subroutine rec_put_f64(w, a)
    type(rec_writer), intent(inout) :: w
    real(8), intent(in), target, contiguous :: a(:)
    if (size(a, kind=8) > 0) call rec_put(w, c_loc(a), 8_8 * size(a, kind=8))
  end subroutine rec_put_f64

  subroutine rec_end(w)
    type(rec_writer), intent(inout) :: w
    if (w%done /= w%total .or. w%sub_left /= 0) then
      write(*,*) 'rb_records: record ended after ', w%done, ' of ', w%total, ' bytes'
      stop 7
    endif
    call end_sub(w)
  end subroutine rec_end

  ! Open the next record of a stream unit: reads its first leading marker.
  subroutine rec_open(r, unit)
    type(rec_reader), intent(out) :: r
    integer, intent(in) :: unit
    integer(4) :: m
    r%unit = unit
    read(unit) m
    r%sub_len = abs(int(m, 8))
    r%sub_left = r%sub_len
    r%more = m < 0
    r%first = .true.
    r%got = 0
  end subroutine rec_open

  subroutine close_sub(r)
    type(rec_reader), intent(inout) :: r
    integer(4) :: m
    read(r%unit) m
    if (abs(int(m, 8)) /= r%sub_len .or. ((m < 0) .neqv. (.not. r%first))) then
      write(*,*) 'rb_records: trailing marker ', m, ' does not close a subrecord of ', r%sub_len, ' bytes'
      stop 7
    endif
    r%first = .false.
  end subroutine close_sub

  ! Read the next n bytes of the record to address p.
  subroutine rec_get(r, p, n)
    type(rec_reader), intent(inout) :: r
    type(c_ptr), intent(in) :: p
    integer(8), intent(in) :: n
    character(kind=c_char), pointer :: b(:)
    integer(8) :: pos, k
    integer(4) :: m
    if (n <= 0) return
    call c_f_pointer(p, b, [n])
    pos = 1
    do while (pos <= n)
      if (r%sub_left == 0) then
        if (.not. r%more) then
          write(*,*) 'rb_records: record holds ', r%got, ' bytes, fewer than requested'
          stop 7
        endif
        call close_sub(r)
        read(r%unit) m
        r%sub_len = abs(int(m, 8))
        r%sub_left = r%sub_len
        r%more = m < 0
        cycle
      endif
      k = min(r%sub_left, n - pos + 1)
      read(r%unit) b(pos:pos + k - 1)
      pos = pos + k
      r%sub_left = r%sub_left - k
      r%got = r%got + k
    enddo
  end subroutine rec_get

  subroutine rec_get_f64(r, a)
    type(rec_reader), intent(inout) :: r
    real(8), intent(inout), target, contiguous :: a(:)
    if (size(a, kind=8) > 0) call rec_get(r, c_loc(a), 8_8 * size(a, kind=8))
  end subroutine rec_get_f64

  ! Close the record: every byte read, then its last trailing marker.
  subroutine rec_close(r)
    type(rec_reader), intent(inout) :: r
    if (r%sub_left /= 0 .or. r%more) then
      write(*,*) 'rb_records: record longer than the data read (', r%got, ' bytes read)'
      stop 7
    endif
    call close_sub(r)
  end subroutine rec_close

end module rb_records
