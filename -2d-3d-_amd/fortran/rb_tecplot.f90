!==================================================================================================
! rb_tecplot -- Tecplot binary (.plt, format version 101) files of the fields the driver
! reports: one ordered zone, point packing, every variable double precision.  The byte layout
! is the one the reference's writers produce (output_Tecplot, 2DRB/2DRB.F90:1443-1570;
! write_full_fields_plt / write_slice_fields_plt / write_slice_psi_vort_plt,
! 2DRB/3DRB.F90:3018-3266), so files open in the same tools:
!   header   "#!TDV101", int 1, title, number of variables, variable names,
!            zone marker 299.0, zone name, zone fields (-1, 0, 1, 0, 0), I, J, K, 0,
!            end-of-header marker 357.0
!   data     zone marker 299.0, one format code per variable (2 = double), 0, -1, then the
!            values, all variables of a point together, I fastest
! Strings are written one 4-byte character code per character followed by a 0.
!==================================================================================================
module rb_tecplot
  implicit none
  private
  public :: plt_write

contains

  subroutine put_string(u, s)
    integer, intent(in) :: u
    character(len=*), intent(in) :: s
    integer :: m
    do m = 1, len_trim(s)
      write(u) int(ichar(s(m:m)), 4)
    enddo
    write(u) 0_4
  end subroutine put_string

  ! data(v, p): variable v at point p, p = i + ni (j - 1 + nj (k - 1)) with i fastest
  subroutine plt_write(filename, title, names, ni, nj, nk, data)
    character(len=*), intent(in) :: filename, title
    character(len=*), intent(in) :: names(:)
    integer, intent(in) :: ni, nj, nk
    real(8), intent(in) :: data(:, :)
    integer :: u, v, nv
    nv = size(names)
    open(newunit=u, file=trim(filename), access='stream', form='unformatted', status='replace')
    write(u) '#!TDV101'
    write(u) 1_4
    call put_string(u, title)
    write(u) int(nv, 4)
    do v = 1, nv
      call put_string(u, names(v))
    enddo
    write(u) 299.0_4
    call put_string(u, 'ZONE 001')
    write(u) -1_4, 0_4, 1_4, 0_4, 0_4
    write(u) int(ni, 4), int(nj, 4), int(nk, 4)
    write(u) 0_4
    write(u) 357.0_4
    write(u) 299.0_4
    write(u) (2_4, v = 1, nv)
    write(u) 0_4, -1_4
    write(u) data
    close(u)
  end subroutine plt_write

end module rb_tecplot
