!> umwm_gpu: the UMWM model driver on one MI355X.
!>
!> Keeps the reference's three entry points (umwm_initialize, umwm_run,
!> umwm_finalize; src/umwm_top.F90:7,32,199) and its namelist groups
!> (namelists/main.nml, read in src/umwm_init.F90:67-92), and runs the inner
!> time loop of src/umwm_top.F90:89-150 either operator by operator through
!> the same-named subroutines of umwm_gpu_ops (mode = 'operators') or as one
!> fused substep per iteration (mode = 'fused').
!>
!> Usage: umwm_gpu [namelist_file]   (default namelists/main.nml)
!>
!> Inputs the reference reads from NetCDF are taken as: constant forcing
!> from &FORCING_CONSTANT, bathymetry (topoFromFile) from a raw file of
!> mm*nm float64 values, m fastest (&GPU topo_file).  An optional &GPU group
!> selects the precision, the mode, a JONSWAP start instead of the cold start
!> and a record file, which receives after each dtg the record
!>   int32 im, int32 hour, float64 swh(im), mwp(im), ustar(im).
!> Gridded output and restart files follow &OUTPUT outgrid / outrst as in
!> src/umwm_top.F90:157-195 (output/umwmout_<time>.bin after stress('ocn'),
!> restart/umwmrst_<time>.bin, and with restart = .true. the start reads
!> restart/umwmrst_<startTimeStr>.bin), in the raw container of
!> umwm_amd/csrc/rawio.hpp instead of NetCDF.

!> &STOKES lives in its own scope: its group name is also the name of the
!> &OUTPUT switch (src/umwm_stokes.F90:231-240 reads it inside umwm_stokes).
module umwm_gpu_stokes_nml
  implicit none
contains
  subroutine read_stokes(u, out, ios)
    integer, intent(in) :: u
    real(8), intent(out) :: out(100)
    integer, intent(out) :: ios
    real(8) :: depths(100)
    namelist /STOKES/ depths
    depths = -1
    rewind(u)
    read(u, nml=STOKES, iostat=ios)
    out = depths
  end subroutine read_stokes
end module umwm_gpu_stokes_nml


module umwm_gpu_top
  use iso_c_binding
  use umwm_gpu_stokes_nml
  use umwm_hip_capi
  use umwm_gpu_ops
  implicit none

  private
  public :: umwm_initialize, umwm_run, umwm_finalize

  ! namelist variables (src/umwm_module.F90 names)
  logical :: isGlobal = .false., restart = .false.
  integer :: mm = 21, nm = 11, om = 37, pm = 36
  real(8) :: fmin = 0.0313d0, fmax = 2.0d0, fprog = 2.0d0, dtg = 3600
  character(19) :: startTimeStr = '2012-01-01 00:00:00', stopTimeStr = '2012-01-01 06:00:00'
  real(8) :: g = 9.80665d0, nu_air = 1.56d-5, nu_water = 0.90d-6, sfct = 0.07d0, kappa = 0.4d0
  real(8) :: z = 10.d0, gustiness = 0.d0, dmin = 10.d0, explim = 0.9d0, sin_fac = 0.11d0
  real(8) :: sin_diss1 = 0.10d0, sin_diss2 = 0.001d0, sds_fac = 42.d0, sds_power = 2.4d0
  real(8) :: mss_fac = 360, snl_fac = 5.0d0, sdt_fac = 0.002d0, sbf_fac = 0.003d0, sbp_fac = 0.003d0
  logical :: gridFromFile = .false., topoFromFile = .false., fillEstuaries = .false.
  logical :: fillLakes = .false.
  real(8) :: delx = 10000, dely = 10000, dpt = 1000
  logical :: winds = .false., currents = .false., air_density = .false.
  logical :: water_density = .false., seaice = .false.
  real(8) :: wspd0 = 10, wdir0 = 0, uc0 = 0, vc0 = 0, rhoa0 = 1.2d0, rhow0 = 1030, fice0 = 0
  real(8) :: fice_lth = 0.30d0, fice_uth = 0.75d0
  integer :: outgrid = 1, outspec = 0, outrst = 6, xpl = 20, ypl = 6
  logical :: stokes = .true.
  ! &GPU (no counterpart in the reference)
  integer :: precision = 8
  character(16) :: mode = 'fused'
  logical :: jonswap = .false.
  real(8) :: js_alpha = 0.01d0, js_gamma = 3.3d0, js_fp = 0.3d0
  character(256) :: topo_file = 'input/umwm_topo.bin', out_file = ''

  namelist /DOMAIN/ isGlobal, mm, nm, om, pm, fmin, fmax, fprog, startTimeStr, stopTimeStr, &
                    dtg, restart
  namelist /PHYSICS/ g, nu_air, nu_water, sfct, kappa, z, gustiness, dmin, explim, sin_fac, &
                     sin_diss1, sin_diss2, sds_fac, sds_power, mss_fac, snl_fac, sdt_fac, &
                     sbf_fac, sbp_fac
  namelist /GRID/ gridFromFile, delx, dely, topoFromFile, dpt, fillEstuaries, fillLakes
  namelist /FORCING/ winds, currents, air_density, water_density, seaice
  namelist /FORCING_CONSTANT/ wspd0, wdir0, uc0, vc0, rhoa0, rhow0, fice0, fice_lth, fice_uth
  namelist /OUTPUT/ outgrid, outspec, outrst, xpl, ypl, stokes
  namelist /GPU/ precision, mode, jonswap, js_alpha, js_gamma, js_fp, topo_file, out_file

  real(8) :: depths(100) = -1
  integer :: im = 0, nsteps = 0, iout = 0
  logical :: have_out = .false.  ! newunit= numbers are negative
  real(c_double), allocatable, target :: topo(:)

contains

  integer function b2i(l)
    logical, intent(in) :: l
    b2i = merge(1, 0, l)
  end function b2i

  !> 'YYYY-MM-DD_HH:MM:SS' of seconds since 1970-01-01 (inverse of
  !> epoch_seconds), the time stamp of the output file names.
  function time_string(t) result(s)
    integer(8), intent(in) :: t
    character(19) :: s
    integer(8) :: days, sod, z, era, doe, yoe, y, doy, mp, d, m
    days = t / 86400
    sod = t - days * 86400
    if (sod < 0) then
      sod = sod + 86400
      days = days - 1
    end if
    z = days + 719468
    era = z / 146097
    if (z < 0 .and. mod(z, 146097_8) /= 0) era = era - 1
    doe = z - era * 146097
    yoe = (doe - doe / 1460 + doe / 36524 - doe / 146096) / 365
    y = yoe + era * 400
    doy = doe - (365 * yoe + yoe / 4 - yoe / 100)
    mp = (5 * doy + 2) / 153
    d = doy - (153 * mp + 2) / 5 + 1
    m = mp + merge(3_8, -9_8, mp < 10)
    if (m <= 2) y = y + 1
    write(s, '(i4.4,a,i2.2,a,i2.2,a,i2.2,a,i2.2,a,i2.2)') y, '-', m, '-', d, '_', &
      sod / 3600, ':', mod(sod, 3600_8) / 60, ':', mod(sod, 60_8)
  end function time_string

  !> Seconds since 1970-01-01 of 'YYYY-MM-DD HH:MM:SS' (proleptic Gregorian).
  integer(8) function epoch_seconds(s)
    character(*), intent(in) :: s
    integer :: y, mo, d, h, mi, se
    integer(8) :: era, yoe, doy, doe, yy, days
    read(s(1:4), *) y
    read(s(6:7), *) mo
    read(s(9:10), *) d
    read(s(12:13), *) h
    read(s(15:16), *) mi
    read(s(18:19), *) se
    yy = y
    if (mo <= 2) yy = yy - 1
    era = yy / 400
    if (yy < 0 .and. mod(yy, 400_8) /= 0) era = era - 1
    yoe = yy - era * 400
    doy = (153 * (mo + merge(-3, 9, mo > 2)) + 2) / 5 + d - 1
    doe = yoe * 365 + yoe / 4 - yoe / 100 + doy
    days = era * 146097 + doe - 719468
    epoch_seconds = days * 86400 + h * 3600 + mi * 60 + se
  end function epoch_seconds

  subroutine read_group(u, name)
    integer, intent(in) :: u
    character(*), intent(in) :: name
    integer :: ios
    rewind(u)
    select case (name)
    case ('DOMAIN'); read(u, nml=DOMAIN, iostat=ios)
    case ('PHYSICS'); read(u, nml=PHYSICS, iostat=ios)
    case ('GRID'); read(u, nml=GRID, iostat=ios)
    case ('FORCING'); read(u, nml=FORCING, iostat=ios)
    case ('FORCING_CONSTANT'); read(u, nml=FORCING_CONSTANT, iostat=ios)
    case ('OUTPUT'); read(u, nml=OUTPUT, iostat=ios)
    case ('STOKES'); call read_stokes(u, depths, ios)
    case ('GPU'); read(u, nml=GPU, iostat=ios)
    end select
    if (ios /= 0 .and. name /= 'GPU') then
      call raiseexception('error', 'nmlread', 'cannot read namelist group '//name)
      stop 1
    end if
  end subroutine read_group

  !> src/umwm_top.F90:7-29: namelist, grid, masks, maps, tables, initial
  !> state and stokes_drift('init').
  subroutine umwm_initialize(nmlfile)
    character(*), intent(in) :: nmlfile
    type(t_umwm_params) :: p
    type(t_umwm_grid_input), target :: gi
    type(t_umwm_jonswap) :: js
    integer(c_int32_t) :: dims(8)
    integer :: u, ios, n

    open(newunit=u, file=nmlfile, status='old', action='read', iostat=ios)
    if (ios /= 0) then
      call raiseexception('error', 'nmlread', 'cannot open '//trim(nmlfile))
      stop 1
    end if
    call read_group(u, 'DOMAIN')
    call read_group(u, 'PHYSICS')
    call read_group(u, 'GRID')
    call read_group(u, 'FORCING')
    call read_group(u, 'FORCING_CONSTANT')
    call read_group(u, 'OUTPUT')
    call read_group(u, 'STOKES')
    call read_group(u, 'GPU')
    close(u)
    if (winds .or. currents .or. air_density .or. water_density .or. seaice .or. gridFromFile) then
      call raiseexception('error', 'umwm_initialize', &
                          'forcing and grid files are not read by this driver')
      stop 1
    end if

    p%mm = mm; p%nm = nm; p%om = om; p%pm = pm
    p%isglobal = b2i(isGlobal); p%restart = b2i(restart)
    p%gridfromfile = 0; p%topofromfile = b2i(topoFromFile)
    p%fillestuaries = b2i(fillEstuaries); p%filllakes = 0
    p%winds = 0; p%currents = 0; p%air_density = 0; p%water_density = 0; p%seaice = 0
    p%stokes = b2i(stokes); p%lm = 0; p%reserved0 = 0
    p%fmin = fmin; p%fmax = fmax; p%fprog = fprog; p%dtg = dtg
    p%g = g; p%nu_air = nu_air; p%nu_water = nu_water; p%sfct = sfct; p%kappa = kappa
    p%z = z; p%gustiness = gustiness; p%dmin = dmin; p%explim = explim
    p%sin_fac = sin_fac; p%sin_diss1 = sin_diss1; p%sin_diss2 = sin_diss2
    p%sds_fac = sds_fac; p%sds_power = sds_power; p%mss_fac = mss_fac
    p%snl_fac = snl_fac; p%sdt_fac = sdt_fac; p%sbf_fac = sbf_fac; p%sbp_fac = sbp_fac
    p%delx = delx; p%dely = dely; p%dpt = dpt
    p%wspd0 = wspd0; p%wdir0 = wdir0; p%uc0 = uc0; p%vc0 = vc0
    p%rhoa0 = rhoa0; p%rhow0 = rhow0; p%fice0 = fice0
    p%fice_lth = fice_lth; p%fice_uth = fice_uth
    p%depths = depths

    if (topoFromFile) then
      allocate(topo(mm * nm))
      open(newunit=u, file=trim(topo_file), access='stream', form='unformatted', &
           status='old', iostat=ios)
      if (ios /= 0) then
        call raiseexception('error', 'umwm_initialize', 'cannot open '//trim(topo_file))
        stop 1
      end if
      read(u) topo
      close(u)
      gi%z = c_loc(topo)
    end if

    call check(umwm_hip_create(p, c_loc(gi), c_null_ptr, int(precision, c_int), dom), &
               'umwm_initialize')
    call check(umwm_hip_dims(dom, dims), 'umwm_initialize')
    im = dims(1)
    if (jonswap) then
      js%alpha = js_alpha; js%gamma = js_gamma; js%sigma_a = 0.07d0; js%sigma_b = 0.09d0
      js%fp = js_fp; js%fp_from_wind = 0; js%dir_from_wind = 1; js%thw = 0
      call check(umwm_hip_init_jonswap(dom, js), 'umwm_initialize')
    end if
    call stokes_drift('init')
    ! gridded output calls stress('ocn'), which after fused substeps needs
    ! the sources of the substep that reaches dtg kept
    if (outgrid /= 0 .and. trim(mode) == 'fused') &
      call check(umwm_hip_set_scalar(dom, c_str('keep_sources'), 1d0), 'umwm_initialize')
    ! src/umwm_top.F90:66: restart_read(starttimestr)
    if (restart) call check(umwm_hip_read_restart(dom, &
      c_str('restart/umwmrst_'//time_string(epoch_seconds(startTimeStr))//'.bin')), 'restart_read')
    nsteps = int((epoch_seconds(stopTimeStr) - epoch_seconds(startTimeStr)) / nint(dtg))
    n = len_trim(out_file)
    if (n > 0) then
      open(newunit=iout, file=out_file(1:n), access='stream', form='unformatted', &
           status='replace')
      have_out = .true.
    end if
    write(*, '(a,i9,a,i5,a,a,a,i2)') 'umwm_gpu: sea points', im, ', dtg steps', nsteps, &
      ', mode ', trim(mode), ', precision', precision
  end subroutine umwm_initialize

  subroutine write_fields(hour)
    integer, intent(in) :: hour
    real(c_double), allocatable :: swh(:), mwp(:), ust(:)
    if (.not. have_out) return
    allocate(swh(im), mwp(im), ust(im))
    call check(umwm_hip_get_field(dom, c_str('ht'), swh, int(im, c_int64_t)), 'output')
    call check(umwm_hip_get_field(dom, c_str('mwp'), mwp, int(im, c_int64_t)), 'output')
    call check(umwm_hip_get_field(dom, c_str('ustar'), ust, int(im, c_int64_t)), 'output')
    write(iout) int(im, 4), int(hour, 4), swh, mwp, ust
  end subroutine write_fields

  !> src/umwm_top.F90:32-196 for constant forcing: outer loop over dtg,
  !> inner loop to sumt >= dtg.
  subroutine umwm_run()
    real(c_double) :: sumt, dts, v
    integer(c_int32_t) :: nsub
    logical :: first
    integer :: hour, k
    real(c_double), allocatable :: swh(:)

    allocate(swh(im))
    first = .not. restart
    do hour = 1, nsteps
      if (trim(mode) == 'fused') then
        call check(umwm_hip_run_dtg(dom, nsub), 'umwm_run')
        call check(umwm_hip_get_scalar(dom, c_str('dts'), dts), 'umwm_run')
      else
        call check(umwm_hip_set_scalar(dom, c_str('sumt'), 0d0), 'umwm_run')
        sumt = 0
        nsub = 0
        do while (sumt < dtg)
          call forcinginterpolate()
          call sin_d12()
          call sds_d12()
          call snl_d12()
          call s_ice()
          call source()
          call propagation()
          call refraction()
          call update()
          call stress('atm')
          if (first) then
            call diag()
            call end_first()
            first = .false.
          end if
          call check(umwm_hip_get_scalar(dom, c_str('sumt'), sumt), 'umwm_run')
          nsub = nsub + 1
        end do
        call check(umwm_hip_get_scalar(dom, c_str('dts'), dts), 'umwm_run')
        if (stokes) call stokes_drift()
        call diag()
      end if
      call check(umwm_hip_get_field(dom, c_str('ht'), swh, int(im, c_int64_t)), 'umwm_run')
      v = 0
      do k = 1, im
        v = max(v, swh(k))
      end do
      write(*, '(a,i4,a,i6,a,f9.3,a,f9.5,a,f9.5)') 'umwm_gpu: hour', hour, ' substeps', nsub, &
        ' last dts', dts, ' mean swh', sum(swh) / im, ' max swh', v
      call write_fields(hour)
      call output_and_restart(epoch_seconds(startTimeStr) + int(hour, 8) * nint(dtg, 8))
    end do
  end subroutine umwm_run

  !> src/umwm_top.F90:157-195: gridded output every outgrid hours (every
  !> dtg for -1) after stress('ocn'), restart every outrst hours, on full
  !> hours of the model clock.
  subroutine output_and_restart(t)
    integer(8), intent(in) :: t
    integer(8) :: hod
    logical :: fullhour
    character(19) :: ts
    ts = time_string(t)
    hod = mod(t, 86400_8) / 3600
    fullhour = mod(t, 3600_8) == 0
    if ((outgrid > 0 .and. fullhour .and. mod(hod, int(max(outgrid, 1), 8)) == 0) .or. &
        outgrid == -1) then
      call stress('ocn')
      call check(umwm_hip_write_output(dom, c_str('output/umwmout_'//ts//'.bin')), &
                 'output_grid')
    end if
    if (outrst > 0 .and. fullhour) then
      if (mod(hod, int(outrst, 8)) == 0) &
        call check(umwm_hip_write_restart(dom, c_str('restart/umwmrst_'//ts//'.bin')), &
                   'restart_write')
    end if
  end subroutine output_and_restart

  !> src/umwm_top.F90:199-206
  subroutine umwm_finalize()
    call check(umwm_hip_destroy(dom), 'umwm_finalize')
    if (have_out) close(iout)
  end subroutine umwm_finalize

end module umwm_gpu_top


program umwm_gpu
  use umwm_gpu_top
  implicit none
  character(512) :: nml
  nml = 'namelists/main.nml'
  if (command_argument_count() >= 1) call get_command_argument(1, nml)
  call umwm_initialize(trim(nml))
  call umwm_run()
  call umwm_finalize()
end program umwm_gpu
