!==================================================================================================
! rb_main -- Rayleigh-Benard driver with the reference's program structure (2DRB/3DRB.F90:198-330,
! 2DRB/2DRB.F90:202-358), stepping the gfx950 solver through the lbm_rb C ABI.
!
!   initial   -> lbm_create + lbm_init_exact / lbm_init_seeded   (or restart: read f, g + scatter)
!   time loop -> lbm_step (the eight sweeps); steadyFlow: check every checkInterval steps and
!                output_Tecplot every backupInterval free-fall times; unsteadyFlow: every
!                outputFrequency free-fall times output_binary, backupData every backupInterval,
!                output_Tecplot (2DRB/3DRB.F90:245-262, 2DRB/2DRB.F90:261-276)
!   end       -> steadyFlow: output_Tecplot + output_binary; lbm_diagnostics (RBcalc_Nu_global,
!                RBcalc_Nu_wall_avg, calNuRe), the mid-plane velocity maxima, MLUPS log; 3D: the
!                final output_Tecplot, 2D: the stream-function report and its Tecplot file
! Files keep the reference's names and formats: backupFile3D-<n>.bin / backupFile-<n>.bin (f, g),
! <binFolderPrefix>-<n>.bin (3D: u, v, w scaled by L/kappa, T, rho; 2D: u, v, T, rho), the
! Tecplot v101 files; records follow GNU Fortran's subrecord convention (rb_records), so restart
! files over 2 GiB read back in the reference's own build.
!
! Parameters keep the reference's names (module commondata3d) and defaults; a namelist file
! given as the first command-line argument overrides them at run time:
!   &rb dims=3, nx=128, ny=128, nz=128, Rayleigh=1.0d6, itc_max=4999 /
! As in the reference the loop runs while (errorU>epsU or errorT>epsT) and itc<=itc_max, i.e.
! itc_max+1 steps when the run does not converge.
!==================================================================================================
program rb_main
  use iso_c_binding
  use lbm_rb
  use rb_records
  use rb_tecplot
  implicit none

  ! ---- commondata3d ------------------------------------------------------------------------
  integer :: dims = 3
  integer :: nx = 120, ny = 120, nz = 120
  real(8) :: Rayleigh = 1.0d7, Prandtl = 0.71d0, Mach = 0.1d0
  real(8) :: Thot = 0.5d0, Tcold = -0.5d0
  integer :: itc_max = 20000000
  integer :: checkInterval = 2000
  real(8) :: epsU = 1.0d-7, epsT = 1.0d-7
  integer :: sideHeated = 0
  integer :: velX = LBM_VEL_NOSLIP, velZ = LBM_VEL_NOSLIP
  integer :: tempX = LBM_T_ADIABATIC, tempY = LBM_T_CONST, tempZ = LBM_T_ADIABATIC
  integer :: useG = 1
  integer :: legacyThermal = 0       ! EnableLegacyThermalScheme (FP64 form only)
  integer :: layout = 0              ! population layout: 0 two arrays, 1 in place (A-A), 2 auto
  real(8) :: outputFrequency = 100.0d0   ! free-fall times between in-loop calNuRe samples
  integer :: reloadDimensionlessTime = 0
  integer :: nuReInLoop = 0          ! 1: the in-loop calNuRe call (commented out in the reference)
  integer :: precision = LBM_FP64
  integer :: device = 0
  integer :: seeded = 1
  integer(8) :: seed = 20261203_8
  real(8) :: noiseAmp = 1.0d-3
  integer :: loadInitField = 0
  character(len=256) :: reloadFile = ''   ! default: backupFile3D-<reloadbinFileNum>.bin (2D: backupFile-)
  integer :: outputBinFile = 0
  integer :: outputPltFile = 0
  character(len=100) :: pltFolderPrefix = ''
  character(len=100) :: binFolderPrefix = ''   ! 3D 'buoyancyCavity3DbinFile', 2D './binFile/buoyancyCavity'
  integer :: backupAtEnd = 0
  integer :: flowMode = 0            ! 0: steadyFlow (the reference's default), 1: unsteadyFlow
  integer :: backupInterval = 1000   ! free-fall times between backups
  integer :: endOutputs = 1          ! the reference's end-of-run Tecplot and binary files
  integer :: reloadbinFileNum = 0    ! unsteadyFlow file numbering continues after a reload
  integer(8) :: maxSubrecord = 2147483639_8   ! bytes of data per subrecord (gfortran's limit)
  character(len=256) :: settingsFile = 'SimulationSettings3D.txt'
  namelist /rb/ dims, nx, ny, nz, Rayleigh, Prandtl, Mach, Thot, Tcold, itc_max, checkInterval, &
       epsU, epsT, sideHeated, velX, velZ, tempX, tempY, tempZ, useG, precision, device, &
       seeded, seed, noiseAmp, loadInitField, reloadFile, outputBinFile, backupAtEnd, settingsFile, &
       outputPltFile, pltFolderPrefix, legacyThermal, layout, outputFrequency, &
       reloadDimensionlessTime, nuReInLoop, binFolderPrefix, flowMode, backupInterval, &
       endOutputs, reloadbinFileNum, maxSubrecord

  type(lbm_config) :: cfg
  type(lbm_params) :: prm
  type(lbm_diag) :: dg
  type(lbm_midplane) :: mp
  type(c_ptr) :: h = c_null_ptr
  integer :: itc, nchunk, qf, qt, ios
  real(8) :: errorU, errorT
  integer :: dimensionlessTime = 0, cadence, target, ns, lead, k
  integer :: binFileNum = 0, tecCadence, bakCadence, lastPlt = -1
  logical :: cadenceIO
  type(lbm_diag), allocatable :: series(:)
  integer(8) :: t0, t1, rate
  real(8) :: seconds
  character(len=256) :: nmlFile
  character(len=32) :: tag0

  if (command_argument_count() >= 1) then
    call get_command_argument(1, nmlFile)
    open(unit=10, file=trim(nmlFile), status='old', action='read', iostat=ios)
    if (ios /= 0) then
      write(*,*) 'cannot open namelist file ', trim(nmlFile)
      stop 1
    endif
    read(10, nml=rb)
    close(10)
  endif
  if (dims == 2) nz = 1
  if (len_trim(pltFolderPrefix) == 0) &
    pltFolderPrefix = merge('buoyancyCavity3DTecplot', 'buoyancyCavityTecplot  ', dims == 3)
  if (len_trim(binFolderPrefix) == 0) &
    binFolderPrefix = merge('buoyancyCavity3DbinFile ', './binFile/buoyancyCavity', dims == 3)
  if (len_trim(reloadFile) == 0) then   ! reloadFilePrefix-<reloadbinFileNum>.bin, 2DRB/3DRB.F90:609-611
    write(tag0, '(i0)') reloadbinFileNum
    reloadFile = trim(merge('backupFile3D', 'backupFile  ', dims == 3))//'-'//trim(tag0)//'.bin'
  endif
  qf = merge(19, 9, dims == 3)
  qt = merge(7, 5, dims == 3)

  open(unit=00, file=trim(settingsFile), status='unknown')
  write(00,*) 'Start: gfx950 MRT-LBM Rayleigh-Benard driver'
  close(00)

  call initial()

  errorU = 100.0d0
  errorT = 100.0d0
  itc = 0
  call system_clock(t0, rate)
  cadence = max(1, int(outputFrequency * prm%time_unit))              ! 2DRB/3DRB.F90:249
  tecCadence = max(1, backupInterval * int(prm%time_unit))            ! :253 (steadyFlow)
  bakCadence = max(1, int(dble(backupInterval) / outputFrequency)) * cadence   ! :259 (unsteadyFlow)
  cadenceIO = (flowMode == 0 .and. outputPltFile == 1) .or. &
              (flowMode == 1 .and. (outputBinFile == 1 .or. outputPltFile == 1))
  do while (((errorU > epsU) .or. (errorT > epsT)) .and. (itc <= itc_max))
    ! advance to the next check point, output point or the last step
    target = itc_max + 1
    if (flowMode == 0) target = min(target, itc + checkInterval - mod(itc, checkInterval))
    if (cadenceIO) target = min(target, itc + cadence - mod(itc, cadence))
    call advance(target)
    if (flowMode == 0 .and. mod(itc, checkInterval) == 0) call check()
    if (cadenceIO .and. mod(itc, cadence) == 0) then
      if (flowMode == 0) then
        if (outputPltFile == 1 .and. mod(itc, tecCadence) == 0) call output_Tecplot()
      else
        if (outputBinFile == 1) then
          call output_binary()
          if (mod(itc, bakCadence) == 0) call backupData()
        endif
        if (outputPltFile == 1) call output_Tecplot()
      endif
    endif
  enddo
  call lbm_require(lbm_sync(h), 'lbm_sync', h, 6)
  call system_clock(t1)
  seconds = dble(t1 - t0) / dble(rate)

  if (flowMode == 0 .and. endOutputs == 1) then       ! 2DRB/3DRB.F90:269-272
    call output_Tecplot()
    call output_binary()
  endif
  call lbm_require(lbm_diagnostics(h, dg), 'lbm_diagnostics', h, 6)
  call calNuRe(dg)                                     ! the end-of-run call, 2DRB/3DRB.F90:297
  call lbm_require(lbm_midplane_extrema(h, mp), 'lbm_midplane_extrema', h, 6)
  if (dims == 3 .and. endOutputs == 1 .and. lastPlt /= itc) call output_Tecplot()   ! :298
  if (backupAtEnd == 1) call backupData()

  write(*,'(a,1x,es16.8)') 'Nu_global =', dg%nu_global
  write(*,'(a,1x,es16.8)') 'Nu_hot(bottom) =', dg%nu_hot
  write(*,'(a,1x,es16.8)') 'Nu_cold(top)   =', dg%nu_cold
  write(*,'(a,1x,es16.8)') 'Nu_middle      =', dg%nu_middle
  write(*,'(a,1x,es16.8,2x,a,1x,es16.8)') 'Nu_hot_max =', dg%nu_hot_max, 'x_max =', dg%nu_hot_max_pos
  write(*,'(a,1x,es16.8,2x,a,1x,es16.8)') 'Nu_hot_min =', dg%nu_hot_min, 'x_min =', dg%nu_hot_min_pos
  write(*,'(a,1x,es16.8)') 'NuVolAvg =', dg%nu_vol
  write(*,'(a,1x,es16.8)') 'ReVolAvg =', dg%re_vol
  call report_midplane()
  if (dims == 2) call report_psi_vort()
  write(*,'(a,1x,i0)') 'itc =', itc

  open(unit=00, file=trim(settingsFile), status='unknown', position='append')
  write(00,*) '======================================================================'
  write(00,*) 'Time (device) = ', seconds, 's'
  write(00,*) 'MLUPS = ', dble(nx) * dble(ny) * dble(nz) * dble(itc) / max(seconds, 1.0d-12) / 1.0d6
  write(00,*) 'Nu_global =', dg%nu_global
  write(00,*) 'Nu_hot    =', dg%nu_hot
  write(00,*) 'Nu_cold   =', dg%nu_cold
  write(00,*) 'Nu_middle =', dg%nu_middle
  write(00,*) 'NuVolAvg  =', dg%nu_vol
  write(00,*) 'ReVolAvg  =', dg%re_vol
  write(00,*) 'Successfully: DNS completed!'
  close(00)

  call lbm_require(lbm_destroy(h), 'lbm_destroy', c_null_ptr, 6)

contains

  !------------------------------------------------------------------------------------------
  ! Steps from itc to target.  With nuReInLoop, calNuRe at every multiple of the cadence on the
  ! way: the first sample point eagerly when it is not a whole cadence away, the following ones
  ! from lbm_run_diagnostics (graph replays).
  !------------------------------------------------------------------------------------------
  subroutine advance(target)
    integer, intent(in) :: target
    if (nuReInLoop == 1) then
      ns = target / cadence - itc / cadence
      if (ns > 0) then
        lead = cadence - mod(itc, cadence)
        if (lead < cadence) then
          call lbm_require(lbm_step(h, int(lead, c_int32_t)), 'lbm_step', h, 6)
          itc = itc + lead
          call lbm_require(lbm_diagnostics(h, dg), 'lbm_diagnostics', h, 6)
          call calNuRe(dg)
          ns = ns - 1
        endif
        if (ns > 0) then
          allocate(series(ns))
          call lbm_require(lbm_run_diagnostics(h, int(ns * cadence, c_int32_t), &
               int(cadence, c_int32_t), series), 'lbm_run_diagnostics', h, 6)
          itc = itc + ns * cadence
          do k = 1, ns
            call calNuRe(series(k))
          enddo
          deallocate(series)
        endif
      endif
    endif
    if (target > itc) call lbm_require(lbm_step(h, int(target - itc, c_int32_t)), 'lbm_step', h, 6)
    itc = target
  end subroutine advance

  ! File-name number: itc in steadyFlow, the output count in unsteadyFlow (continuing from
  ! reloadbinFileNum after a reload), 2DRB/3DRB.F90:1608-1614, :1646-1651
  integer function file_number()
    if (flowMode == 0) then
      file_number = itc
    else
      file_number = binFileNum + merge(reloadbinFileNum, 0, loadInitField == 1)
    endif
  end function file_number

  !------------------------------------------------------------------------------------------
  ! Mid-plane velocity maxima, printed as calc_*mid_max_common (2DRB/3DRB.F90:2399-2518) and
  ! the 2D RBcalc_/SideHeatedcalc_ umid/vmid routines (2DRB/2DRB.F90:2081-2238, :2484-2642)
  ! report them, and appended to the settings log.
  !------------------------------------------------------------------------------------------
  subroutine report_midplane()
    character(len=*), parameter :: fmt3 = '(A,1X,ES16.8,2X,A,1X,ES16.8,2X,A,1X,ES16.8,2X,A,1X,ES16.8)'
    character(len=*), parameter :: fmt2 = '(A,1X,ES16.8,2X,A,1X,ES16.8,2X,A,1X,ES16.8)'
    open(unit=00, file=trim(settingsFile), status='unknown', position='append')
    if (dims == 3) then
      write(*,fmt3) 'u_mid_max =', mp%u_max, 'at y =', mp%u_pos1, 'z =', mp%u_pos2, 'on x_mid =', mp%u_plane
      write(*,fmt3) 'v_mid_max =', mp%v_max, 'at x =', mp%v_pos1, 'z =', mp%v_pos2, 'on y_mid =', mp%v_plane
      write(*,fmt3) 'w_mid_max =', mp%w_max, 'at x =', mp%w_pos1, 'y =', mp%w_pos2, 'on z_mid =', mp%w_plane
      write(00,*) 'u_mid_max =', mp%u_max, ' y_pos =', mp%u_pos1, ' z_pos =', mp%u_pos2
      write(00,*) 'v_mid_max =', mp%v_max, ' x_pos =', mp%v_pos1, ' z_pos =', mp%v_pos2
      write(00,*) 'w_mid_max =', mp%w_max, ' x_pos =', mp%w_pos1, ' y_pos =', mp%w_pos2
    else
      write(*,fmt2) 'u_mid_max* =', mp%u_max, 'y =', mp%u_pos1, 'x_mid =', mp%u_plane
      write(*,fmt2) 'v_mid_max* =', mp%v_max, 'x =', mp%v_pos1, 'y_mid =', mp%v_plane
      write(00,fmt2) 'u_mid_max* =', mp%u_max, 'y =', mp%u_pos1, 'x_mid =', mp%u_plane
      write(00,fmt2) 'v_mid_max* =', mp%v_max, 'x =', mp%v_pos1, 'y_mid =', mp%v_plane
    endif
    close(00)
  end subroutine report_midplane

  !------------------------------------------------------------------------------------------
  ! initial (2DRB/3DRB.F90:342-643): parameters, allocation, exact or restart initialisation
  !------------------------------------------------------------------------------------------
  subroutine initial()
    real(8), allocatable, target :: f(:,:,:,:), g(:,:,:,:)
    integer :: i, j, k, alpha

    cfg%dims = dims; cfg%nx = nx; cfg%ny = ny; cfg%nz = nz
    cfg%rayleigh = Rayleigh; cfg%prandtl = Prandtl; cfg%mach = Mach
    cfg%thot = Thot; cfg%tcold = Tcold
    cfg%side_heated = sideHeated; cfg%vel_x = velX; cfg%vel_z = velZ
    cfg%temp_x = tempX; cfg%temp_y = tempY; cfg%temp_z = tempZ
    cfg%use_g = useG; cfg%precision = precision; cfg%device = device
    cfg%legacy_thermal = legacyThermal
    cfg%layout = layout
    call lbm_require(lbm_create(cfg, h), 'lbm_create', h, 6)
    call lbm_require(lbm_get_params(h, prm), 'lbm_get_params', h, 6)

    open(unit=00, file=trim(settingsFile), status='unknown', position='append')
    write(00,*) '-------------------------------------------------------------------------------'
    write(00,*) 'Mesh:', nx, ny, nz
    write(00,*) 'Rayleigh=', Rayleigh, '; Prandtl =', Prandtl, '; Mach =', Mach
    write(00,*) 'Length unit: L0 =', prm%L
    write(00,*) 'Time unit: Sqrt(L0/(gBeta*DeltaT)) =', prm%time_unit
    write(00,*) 'Velocity unit: Sqrt(gBeta*L0*DeltaT) =', prm%velocity_unit
    write(00,*) 'tauf =', prm%tauf
    write(00,*) 'taug =', prm%taug, '; Qk =', prm%q_k, '; Qnu =', prm%q_nu
    write(00,*) 'viscosity =', prm%viscosity, '; diffusivity =', prm%diffusivity
    write(00,*) 'itc_max =', itc_max
    write(00,*) 'default epsU =', epsU, '; epsT =', epsT
    write(00,*) 'useG =', useG == 1
    write(00,*) 'precision form =', precision

    if (loadInitField == 0) then
      if (seeded == 1) then
        write(00,*) 'Initial field: exact profile + seeded noise, seed =', seed, ' amp =', noiseAmp
        call lbm_require(lbm_init_seeded(h, int(seed, c_int64_t), noiseAmp), 'lbm_init_seeded', h, 6)
      else
        write(00,*) 'Initial field is set exactly'
        call lbm_require(lbm_init_exact(h, c_null_ptr), 'lbm_init_exact', h, 6)
      endif
    else
      ! restart: f then g records, interior nodes, population index outermost (2DRB/3DRB.F90:611-616)
      allocate(f(0:qf-1, nx, ny, nz), g(0:qt-1, nx, ny, nz))
      open(unit=01, file=trim(reloadFile), form='unformatted', access='stream', status='old', &
           iostat=ios)
      if (ios /= 0) then
        write(*,*) 'cannot open restart file ', trim(reloadFile)
        stop 1
      endif
      call read_populations(01, f, qf)
      call read_populations(01, g, qt)
      close(01)
      write(00,*) 'Raw data is loaded from the file: ', trim(reloadFile)
      call lbm_require(lbm_scatter(h, c_loc(f), c_loc(g)), 'lbm_scatter', h, 6)
      deallocate(f, g)
    endif
    if (outputBinFile == 1) then          ! 2DRB/3DRB.F90:364-369
      open(unit=01, file=trim(binFolderPrefix)//'-readme', status='unknown', iostat=ios)
      if (ios /= 0) then
        write(*,*) 'cannot write into the folder of binFolderPrefix ', trim(binFolderPrefix)
        stop 1
      endif
      write(01,*) 'binFile folder exist!'
      close(01)
      write(00,*) 'Data will be stored in ', trim(binFolderPrefix)
    endif
    write(00,*) 'flow mode = ', trim(merge('steadyFlow  ', 'unsteadyFlow', flowMode == 0))
    write(00,*) 'outputFrequency =', outputFrequency, ' free-fall time units'
    write(00,*) 'backupInterval =', backupInterval, ' free-fall time units'
    write(00,*) '-------------------------------------------------------------------------------'
    close(00)
  end subroutine initial

  !------------------------------------------------------------------------------------------
  ! One restart record (2DRB/3DRB.F90:611-616): q planes of nodes, population index outermost,
  ! into p(0:q-1, nx, ny, nz) (population index fastest, the order of lbm_scatter)
  !------------------------------------------------------------------------------------------
  subroutine read_populations(unit, p, q)
    integer, intent(in) :: unit, q
    real(8), intent(inout) :: p(0:, :, :, :)
    real(8), allocatable, target :: plane(:)
    type(rec_reader) :: r
    integer :: alpha, i, j, k, c
    allocate(plane(int(nx, 8) * ny * nz))
    call rec_open(r, unit)
    do alpha = 0, q - 1
      call rec_get_f64(r, plane)
      c = 0
      do k = 1, nz
        do j = 1, ny
          do i = 1, nx
            c = c + 1
            p(alpha, i, j, k) = plane(c)
          enddo
        enddo
      enddo
    enddo
    call rec_close(r)
    deallocate(plane)
  end subroutine read_populations

  subroutine write_populations(unit, p, q)
    integer, intent(in) :: unit, q
    real(8), intent(in) :: p(0:, :, :, :)
    real(8), allocatable, target :: plane(:)
    type(rec_writer) :: w
    integer :: alpha, i, j, k, c
    allocate(plane(int(nx, 8) * ny * nz))
    call rec_begin(w, unit, 8_8 * q * size(plane, kind=8), maxSubrecord)
    do alpha = 0, q - 1
      c = 0
      do k = 1, nz
        do j = 1, ny
          do i = 1, nx
            c = c + 1
            plane(c) = p(alpha, i, j, k)
          enddo
        enddo
      enddo
      call rec_put_f64(w, plane)
    enddo
    call rec_end(w)
    deallocate(plane)
  end subroutine write_populations

  ! One snapshot record: a field of nx * ny * nz values times scale
  subroutine write_field(unit, a, scale)
    integer, intent(in) :: unit
    real(8), intent(in), target, contiguous :: a(:, :, :)
    real(8), intent(in) :: scale
    real(8), allocatable, target :: b(:)
    type(rec_writer) :: w
    b = reshape(a, [size(a, kind=8)]) * scale
    call rec_begin(w, unit, 8_8 * size(b, kind=8), maxSubrecord)
    call rec_put_f64(w, b)
    call rec_end(w)
  end subroutine write_field

  !------------------------------------------------------------------------------------------
  ! check (2DRB/3DRB.F90:1475-1528)
  !------------------------------------------------------------------------------------------
  subroutine check()
    call lbm_require(lbm_check(h, errorU, errorT), 'lbm_check', h, 6)
    write(*,'(I12,1X,ES24.16,1X,ES24.16)') itc, errorU, errorT
  end subroutine check

  !------------------------------------------------------------------------------------------
  ! calNuRe (2DRB/3DRB.F90:1712-1778; 2DRB/2DRB.F90:1610-1675): one sample of the volume
  ! Nusselt and Reynolds time series, appended at time reloadDimensionlessTime +
  ! dimensionlessTime * outputFrequency
  !------------------------------------------------------------------------------------------
  subroutine calNuRe(d)
    type(lbm_diag), intent(in) :: d
    real(8) :: tnow
    dimensionlessTime = dimensionlessTime + 1
    tnow = real(reloadDimensionlessTime + dimensionlessTime * outputFrequency, kind=8)
    open(unit=01, file=trim(merge('Nu_VolAvg_3D.dat', 'Nu_VolAvg.dat   ', dims == 3)), &
         status='unknown', position='append')
    write(01,*) tnow, d%nu_vol
    close(01)
    open(unit=02, file=trim(merge('Re_VolAvg_3D.dat', 'Re_VolAvg.dat   ', dims == 3)), &
         status='unknown', position='append')
    write(02,*) tnow, d%re_vol
    close(02)
  end subroutine calNuRe

  !------------------------------------------------------------------------------------------
  ! output_binary (2DRB/3DRB.F90:1599-1627; 2D 2DRB/2DRB.F90:1357-1388): <binFolderPrefix>-<n>.bin
  ! with u, v, w scaled by L/kappa, then T, rho (2D: u, v, T, rho); n zero-padded to 12 digits
  ! in 3D, plain in 2D
  !------------------------------------------------------------------------------------------
  subroutine output_binary()
    real(8), allocatable, target :: rho(:,:,:), u(:,:,:), v(:,:,:), w(:,:,:), T(:,:,:)
    character(len=100) :: filename
    real(8) :: scale
    allocate(rho(nx,ny,nz), u(nx,ny,nz), v(nx,ny,nz), w(nx,ny,nz), T(nx,ny,nz))
    call lbm_require(lbm_macros(h, c_loc(rho), c_loc(u), c_loc(v), c_loc(w), c_loc(T)), &
         'lbm_macros', h, 6)
    scale = prm%L / prm%diffusivity
    if (flowMode == 1) binFileNum = binFileNum + 1
    if (dims == 3) then
      write(filename,'(i12.12)') file_number()
    else
      write(filename,'(i0)') file_number()
    endif
    open(unit=03, file=trim(binFolderPrefix)//'-'//trim(adjustl(filename))//'.bin', &
         form='unformatted', access='stream', status='replace')
    call write_field(03, u, scale)
    call write_field(03, v, scale)
    if (dims == 3) call write_field(03, w, scale)
    call write_field(03, T, 1.0d0)
    call write_field(03, rho, 1.0d0)
    close(03)
  end subroutine output_binary

  !------------------------------------------------------------------------------------------
  ! 2D stream function and vorticity (calc_psi_vort_and_output, 2DRB/2DRB.F90:2657-2770): the
  ! |psi| statistics in the log and, with outputPltFile, the psi / vorticity Tecplot file
  ! (output_Tecplot_psi_vort, :2782-2862).
  !------------------------------------------------------------------------------------------
  subroutine report_psi_vort()
    real(8), allocatable, target :: psi(:,:), vort(:,:)
    real(8), allocatable :: dat(:,:)
    type(lbm_psi) :: ps
    character(len=12) :: tag
    character(len=*), parameter :: fmt1 = '(a,1x,es16.8)'
    character(len=*), parameter :: fmt3 = '(a,1x,es16.8,2x,a,1x,es16.8,2x,a,1x,es16.8)'
    integer :: i, j
    allocate(psi(nx, ny), vort(nx, ny))
    call lbm_require(lbm_psi_vorticity_2d(h, c_loc(psi), c_loc(vort), ps), 'lbm_psi_vorticity_2d', h, 6)
    write(*,fmt1) 'abs(psi_center_coarse) =', ps%psi_center_coarse
    write(*,fmt1) 'abs(psi_center_fine) =', ps%psi_center_fine
    write(*,fmt3) 'max(|psi|) =', ps%psi_abs_max, 'x* =', ps%x_at_max, 'y* =', ps%y_at_max
    open(unit=00, file=trim(settingsFile), status='unknown', position='append')
    write(00,fmt1) 'abs(psi_center_coarse) =', ps%psi_center_coarse
    write(00,fmt1) 'abs(psi_center_fine) =', ps%psi_center_fine
    write(00,fmt3) 'max(|psi|) =', ps%psi_abs_max, 'x* =', ps%x_at_max, 'y* =', ps%y_at_max
    close(00)
    if (outputPltFile == 1 .or. endOutputs == 1) then
      allocate(dat(4, nx*ny))
      do j = 1, ny
        do i = 1, nx
          dat(:, i + nx * (j - 1)) = (/ coordinate(i), coordinate(j), psi(i,j), vort(i,j) /)
        enddo
      enddo
      write(tag,'(i12.12)') itc
      call plt_write(trim(pltFolderPrefix)//'-psiVort-'//tag//'.plt', 'psi-vort', &
           (/ 'X   ', 'Y   ', 'PSI ', 'VORT' /), nx, ny, 1, dat)
    endif
  end subroutine report_psi_vort

  !------------------------------------------------------------------------------------------
  ! output_Tecplot (2DRB/3DRB.F90:1676-1701; 2D 2DRB/2DRB.F90:1443-1570).  3D: the full field
  ! (X, Y, Z, U, V, W, T with lattice velocities) and, on each mid-plane, the fields (velocities
  ! scaled by L / kappa) and the stream function and vorticity.  2D: X, Y, U_nd, V_nd, T.
  !------------------------------------------------------------------------------------------
  subroutine output_Tecplot()
    real(8), allocatable, target :: rho(:,:,:), u(:,:,:), v(:,:,:), w(:,:,:), T(:,:,:)
    real(8), allocatable, target :: s(:,:,:)
    real(8), allocatable :: dat(:,:), c1(:), c2(:)
    character(len=12) :: tag
    character(len=1), parameter :: ax(3) = (/ 'X', 'Y', 'Z' /)
    character(len=2), parameter :: pl(3) = (/ 'YZ', 'XZ', 'XY' /)
    real(8) :: scale
    integer :: i, j, k, p, a, n1, n2
    allocate(rho(nx,ny,nz), u(nx,ny,nz), v(nx,ny,nz), w(nx,ny,nz), T(nx,ny,nz))
    call lbm_require(lbm_macros(h, c_loc(rho), c_loc(u), c_loc(v), c_loc(w), c_loc(T)), &
         'lbm_macros', h, 6)
    scale = prm%L / prm%diffusivity
    lastPlt = itc
    write(tag,'(i12.12)') itc
    if (dims == 2) then
      allocate(dat(5, nx*ny))
      do j = 1, ny
        do i = 1, nx
          p = i + nx * (j - 1)
          dat(:, p) = (/ coordinate(i), coordinate(j), scale * u(i,j,1), scale * v(i,j,1), &
                         T(i,j,1) /)
        enddo
      enddo
      call plt_write(trim(pltFolderPrefix)//'-'//tag//'.plt', 'MyFirst', &
           (/ 'X   ', 'Y   ', 'U_nd', 'V_nd', 'T   ' /), nx, ny, 1, dat)
      return
    endif
    allocate(dat(7, nx*ny*nz))
    do k = 1, nz
      do j = 1, ny
        do i = 1, nx
          p = i + nx * (j - 1 + ny * (k - 1))
          dat(:, p) = (/ coordinate(i), coordinate(j), coordinate(k), &
                         u(i,j,k), v(i,j,k), w(i,j,k), T(i,j,k) /)
        enddo
      enddo
    enddo
    call plt_write(trim(pltFolderPrefix)//'-fullField-'//tag//'.plt', 'full-field-3d', &
         (/ 'X', 'Y', 'Z', 'U', 'V', 'W', 'T' /), nx, ny, nz, dat)
    deallocate(dat)
    do a = 1, 3
      ! in-plane axes: x-plane (y, z), y-plane (x, z), z-plane (x, y)
      n1 = merge(ny, nx, a == 1)
      n2 = merge(ny, nz, a == 3)
      allocate(s(n1, n2, 6), dat(6, n1*n2), c1(n1), c2(n2))
      call lbm_require(lbm_midplane_slice(h, int(a - 1, c_int32_t), c_loc(s(1,1,1)), &
           c_loc(s(1,1,2)), c_loc(s(1,1,3)), c_loc(s(1,1,4)), c_loc(s(1,1,5)), c_loc(s(1,1,6))), &
           'lbm_midplane_slice', h, 6)
      do i = 1, n1
        c1(i) = coordinate(i)
      enddo
      do j = 1, n2
        c2(j) = coordinate(j)
      enddo
      do j = 1, n2
        do i = 1, n1
          dat(:, i + n1 * (j - 1)) = (/ c1(i), c2(j), s(i,j,1), s(i,j,2), s(i,j,3), s(i,j,4) /)
        enddo
      enddo
      call plt_write(trim(pltFolderPrefix)//'-mid'//ax(a)//'-'//tag//'.plt', 'mid'//ax(a)//'-fields', &
           (/ pl(a)(1:1)//'   ', pl(a)(2:2)//'   ', 'U_nd', 'V_nd', 'W_nd', 'T   ' /), n1, n2, 1, dat)
      do j = 1, n2
        do i = 1, n1
          dat(1:4, i + n1 * (j - 1)) = (/ c1(i), c2(j), s(i,j,5), s(i,j,6) /)
        enddo
      enddo
      call plt_write(trim(pltFolderPrefix)//'-mid'//ax(a)//'psiVort-'//tag//'.plt', &
           'mid'//ax(a)//'-psi-vort', (/ pl(a)(1:1)//'     ', pl(a)(2:2)//'     ', &
           'PSI_'//pl(a), 'VORT_'//ax(a) /), n1, n2, 1, dat(1:4, :))
      deallocate(s, dat, c1, c2)
    enddo
  end subroutine output_Tecplot

  ! node coordinate (i - 1/2) / L of node i (2DRB/3DRB.F90:443-462)
  real(8) function coordinate(i)
    integer, intent(in) :: i
    coordinate = (dble(i) - 0.5d0) / prm%L
  end function coordinate

  !------------------------------------------------------------------------------------------
  ! backupData (2DRB/3DRB.F90:1638-1665; 2D 2DRB/2DRB.F90:1402-1430): f then g, population index
  ! outermost, into backupFile3D-<n>.bin (2D: backupFile-<n>.bin)
  !------------------------------------------------------------------------------------------
  subroutine backupData()
    real(8), allocatable, target :: f(:,:,:,:), g(:,:,:,:)
    character(len=100) :: filename
    allocate(f(0:qf-1, nx, ny, nz), g(0:qt-1, nx, ny, nz))
    call lbm_require(lbm_gather(h, c_loc(f), c_loc(g), c_null_ptr), 'lbm_gather', h, 6)
    write(filename,'(i0)') file_number()
    filename = trim(merge('backupFile3D', 'backupFile  ', dims == 3))//'-'//trim(filename)//'.bin'
    open(unit=05, file=trim(filename), form='unformatted', access='stream', status='replace')
    call write_populations(05, f, qf)
    call write_populations(05, g, qt)
    close(05)
    deallocate(f, g)
    open(unit=00, file=trim(settingsFile), status='unknown', position='append')
    write(00,*) 'Backup f and g to the file: ', trim(filename)
    close(00)
  end subroutine backupData

end program rb_main
