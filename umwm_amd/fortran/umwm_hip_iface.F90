!> iso_c_binding layer over the C-ABI of include/umwm_hip.h.
!>
!> umwm_hip_capi   bind(C) mirrors of the C structs and interface blocks of
!>                 the entry points.
!> umwm_gpu_ops    the reference's operator subroutines under their own
!>                 names and argument lists (src/umwm_source_functions.F90:12,
!>                 src/umwm_physics.F90:9,89, src/umwm_advection.F90:18,102,
!>                 src/umwm_stress.F90:17, src/umwm_stokes.F90:13), acting on
!>                 the module's current domain.  A failed call ends the run
!>                 the reference's way: raiseexception, then stop
!>                 (src/umwm_util.F90:129-153).
module umwm_hip_capi
  use iso_c_binding
  implicit none

  integer(c_int), parameter :: UMWM_FP32 = 4, UMWM_FP64 = 8

  type, bind(C) :: t_umwm_params
    integer(c_int32_t) :: mm, nm, om, pm
    integer(c_int32_t) :: isglobal, restart
    integer(c_int32_t) :: gridfromfile, topofromfile, fillestuaries, filllakes
    integer(c_int32_t) :: winds, currents, air_density, water_density, seaice
    integer(c_int32_t) :: stokes, lm, reserved0
    real(c_double) :: fmin, fmax, fprog, dtg
    real(c_double) :: g, nu_air, nu_water, sfct, kappa, z, gustiness, dmin, explim
    real(c_double) :: sin_fac, sin_diss1, sin_diss2, sds_fac, sds_power, mss_fac
    real(c_double) :: snl_fac, sdt_fac, sbf_fac, sbp_fac
    real(c_double) :: delx, dely, dpt
    real(c_double) :: wspd0, wdir0, uc0, vc0, rhoa0, rhow0, fice0, fice_lth, fice_uth
    real(c_double) :: depths(100)
  end type t_umwm_params

  type, bind(C) :: t_umwm_grid_input
    type(c_ptr) :: lon = c_null_ptr, lat = c_null_ptr, z = c_null_ptr
    type(c_ptr) :: exclude = c_null_ptr
    integer(c_int32_t) :: nexclude = 0, reserved0 = 0
  end type t_umwm_grid_input

  type, bind(C) :: t_umwm_jonswap
    real(c_double) :: alpha, gamma, sigma_a, sigma_b, fp
    integer(c_int32_t) :: fp_from_wind, dir_from_wind
    real(c_double) :: thw
  end type t_umwm_jonswap

  interface
    function umwm_hip_error_string(status) result(s) bind(C, name='umwm_hip_error_string')
      import :: c_int, c_ptr
      integer(c_int), value :: status
      type(c_ptr) :: s
    end function
    function umwm_hip_last_error() result(s) bind(C, name='umwm_hip_last_error')
      import :: c_ptr
      type(c_ptr) :: s
    end function
    function umwm_hip_create(params, grid, forcing0, precision, out) result(rc) &
        bind(C, name='umwm_hip_create')
      import :: c_int, c_ptr, t_umwm_params
      type(t_umwm_params), intent(in) :: params
      type(c_ptr), value :: grid, forcing0
      integer(c_int), value :: precision
      type(c_ptr), intent(out) :: out
      integer(c_int) :: rc
    end function
    function umwm_hip_destroy(dom) result(rc) bind(C, name='umwm_hip_destroy')
      import :: c_int, c_ptr
      type(c_ptr), value :: dom
      integer(c_int) :: rc
    end function
    function umwm_hip_dims(dom, out) result(rc) bind(C, name='umwm_hip_dims')
      import :: c_int, c_ptr, c_int32_t
      type(c_ptr), value :: dom
      integer(c_int32_t), intent(out) :: out(8)
      integer(c_int) :: rc
    end function
    function umwm_hip_init_jonswap(dom, js) result(rc) bind(C, name='umwm_hip_init_jonswap')
      import :: c_int, c_ptr, t_umwm_jonswap
      type(c_ptr), value :: dom
      type(t_umwm_jonswap), intent(in) :: js
      integer(c_int) :: rc
    end function
    function umwm_hip_stokes_drift(dom, init) result(rc) bind(C, name='umwm_hip_stokes_drift')
      import :: c_int, c_ptr
      type(c_ptr), value :: dom
      integer(c_int), value :: init
      integer(c_int) :: rc
    end function
    function umwm_hip_stress(dom, option) result(rc) bind(C, name='umwm_hip_stress')
      import :: c_int, c_ptr, c_char
      type(c_ptr), value :: dom
      character(kind=c_char), intent(in) :: option(*)
      integer(c_int) :: rc
    end function
    function umwm_hip_substep(dom, sumt, dts) result(rc) bind(C, name='umwm_hip_substep')
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: dom
      real(c_double), intent(out) :: sumt, dts
      integer(c_int) :: rc
    end function
    function umwm_hip_write_output(dom, path) result(rc) bind(C, name='umwm_hip_write_output')
      import :: c_int, c_ptr, c_char
      type(c_ptr), value :: dom
      character(kind=c_char), intent(in) :: path(*)
      integer(c_int) :: rc
    end function
    function umwm_hip_write_restart(dom, path) result(rc) bind(C, name='umwm_hip_write_restart')
      import :: c_int, c_ptr, c_char
      type(c_ptr), value :: dom
      character(kind=c_char), intent(in) :: path(*)
      integer(c_int) :: rc
    end function
    function umwm_hip_read_restart(dom, path) result(rc) bind(C, name='umwm_hip_read_restart')
      import :: c_int, c_ptr, c_char
      type(c_ptr), value :: dom
      character(kind=c_char), intent(in) :: path(*)
      integer(c_int) :: rc
    end function
    function umwm_hip_substeps(dom, n, sumt, dts, nsub) result(rc) &
        bind(C, name='umwm_hip_substeps')
      import :: c_int, c_ptr, c_double, c_int32_t
      type(c_ptr), value :: dom
      integer(c_int32_t), value :: n
      real(c_double), intent(out) :: sumt, dts
      integer(c_int32_t), intent(out) :: nsub
      integer(c_int) :: rc
    end function
    function umwm_hip_run_dtg(dom, nsub) result(rc) bind(C, name='umwm_hip_run_dtg')
      import :: c_int, c_ptr, c_int32_t
      type(c_ptr), value :: dom
      integer(c_int32_t), intent(out) :: nsub
      integer(c_int) :: rc
    end function
    function umwm_hip_get_field(dom, name, buf, n) result(rc) bind(C, name='umwm_hip_get_field')
      import :: c_int, c_ptr, c_char, c_double, c_int64_t
      type(c_ptr), value :: dom
      character(kind=c_char), intent(in) :: name(*)
      real(c_double), intent(out) :: buf(*)
      integer(c_int64_t), value :: n
      integer(c_int) :: rc
    end function
    function umwm_hip_get_scalar(dom, name, v) result(rc) bind(C, name='umwm_hip_get_scalar')
      import :: c_int, c_ptr, c_char, c_double
      type(c_ptr), value :: dom
      character(kind=c_char), intent(in) :: name(*)
      real(c_double), intent(out) :: v
      integer(c_int) :: rc
    end function
    function umwm_hip_set_scalar(dom, name, v) result(rc) bind(C, name='umwm_hip_set_scalar')
      import :: c_int, c_ptr, c_char, c_double
      type(c_ptr), value :: dom
      character(kind=c_char), intent(in) :: name(*)
      real(c_double), value :: v
      integer(c_int) :: rc
    end function
    function umwm_hip_field_size(dom, name) result(n) bind(C, name='umwm_hip_field_size')
      import :: c_ptr, c_char, c_int64_t
      type(c_ptr), value :: dom
      character(kind=c_char), intent(in) :: name(*)
      integer(c_int64_t) :: n
    end function
  end interface

  ! the reference's argument-less operators: one domain handle each
  interface
    function c_sin_d12(dom) result(rc) bind(C, name='umwm_hip_sin_d12')
      import :: c_int, c_ptr
      type(c_ptr), value :: dom
      integer(c_int) :: rc
    end function
    function c_sds_d12(dom) result(rc) bind(C, name='umwm_hip_sds_d12')
      import :: c_int, c_ptr
      type(c_ptr), value :: dom
      integer(c_int) :: rc
    end function
    function c_snl_d12(dom) result(rc) bind(C, name='umwm_hip_snl_d12')
      import :: c_int, c_ptr
      type(c_ptr), value :: dom
      integer(c_int) :: rc
    end function
    function c_s_ice(dom) result(rc) bind(C, name='umwm_hip_s_ice')
      import :: c_int, c_ptr
      type(c_ptr), value :: dom
      integer(c_int) :: rc
    end function
    function c_source(dom) result(rc) bind(C, name='umwm_hip_source')
      import :: c_int, c_ptr
      type(c_ptr), value :: dom
      integer(c_int) :: rc
    end function
    function c_propagation(dom) result(rc) bind(C, name='umwm_hip_propagation')
      import :: c_int, c_ptr
      type(c_ptr), value :: dom
      integer(c_int) :: rc
    end function
    function c_refraction(dom) result(rc) bind(C, name='umwm_hip_refraction')
      import :: c_int, c_ptr
      type(c_ptr), value :: dom
      integer(c_int) :: rc
    end function
    function c_update(dom) result(rc) bind(C, name='umwm_hip_update')
      import :: c_int, c_ptr
      type(c_ptr), value :: dom
      integer(c_int) :: rc
    end function
    function c_diag(dom) result(rc) bind(C, name='umwm_hip_diag')
      import :: c_int, c_ptr
      type(c_ptr), value :: dom
      integer(c_int) :: rc
    end function
    function c_forcinginterpolate(dom) result(rc) bind(C, name='umwm_hip_forcinginterpolate')
      import :: c_int, c_ptr
      type(c_ptr), value :: dom
      integer(c_int) :: rc
    end function
    function c_end_first(dom) result(rc) bind(C, name='umwm_hip_end_first')
      import :: c_int, c_ptr
      type(c_ptr), value :: dom
      integer(c_int) :: rc
    end function
  end interface

contains

  !> Fortran string of a C string returned by the library.
  function c_string(p) result(s)
    type(c_ptr), intent(in) :: p
    character(:), allocatable :: s
    character(kind=c_char), pointer :: a(:)
    integer :: n
    if (.not. c_associated(p)) then
      s = ''
      return
    end if
    call c_f_pointer(p, a, [4096])
    n = 0
    do while (n < 4096)
      if (a(n + 1) == c_null_char) exit
      n = n + 1
    end do
    allocate(character(n) :: s)
    s = transfer(a(1:n), s)
  end function c_string

  !> NUL-terminated copy of a Fortran string.
  function c_str(s) result(a)
    character(*), intent(in) :: s
    character(kind=c_char) :: a(len_trim(s) + 1)
    integer :: i
    do i = 1, len_trim(s)
      a(i) = s(i:i)
    end do
    a(len_trim(s) + 1) = c_null_char
  end function c_str

end module umwm_hip_capi


module umwm_gpu_ops
  use iso_c_binding
  use umwm_hip_capi
  implicit none

  private
  public :: dom, check, raiseexception
  public :: forcinginterpolate, sin_d12, sds_d12, snl_d12, s_ice, source
  public :: propagation, refraction, update, stress, diag, stokes_drift, end_first

  type(c_ptr), save :: dom = c_null_ptr  ! the current domain

contains

  !> src/umwm_util.F90:129-153: report an exception on stdout.
  subroutine raiseexception(typ, sub, msg)
    character(*), intent(in) :: typ, sub, msg
    write(*, '(a)') 'umwm: '//trim(typ)//': '//trim(sub)//': '//trim(msg)
  end subroutine raiseexception

  !> Fatal on a failed library call (the reference stops on fatal errors).
  subroutine check(rc, where)
    integer(c_int), intent(in) :: rc
    character(*), intent(in) :: where
    if (rc /= 0) then
      call raiseexception('error', where, c_string(umwm_hip_error_string(rc))//': '// &
                          c_string(umwm_hip_last_error()))
      stop 1
    end if
  end subroutine check

  subroutine forcinginterpolate()
    call check(c_forcinginterpolate(dom), 'forcinginterpolate')
  end subroutine
  subroutine sin_d12()
    call check(c_sin_d12(dom), 'sin_d12')
  end subroutine
  subroutine sds_d12()
    call check(c_sds_d12(dom), 'sds_d12')
  end subroutine
  subroutine snl_d12()
    call check(c_snl_d12(dom), 'snl_d12')
  end subroutine
  subroutine s_ice()
    call check(c_s_ice(dom), 's_ice')
  end subroutine
  subroutine source()
    call check(c_source(dom), 'source')
  end subroutine
  subroutine propagation()
    call check(c_propagation(dom), 'propagation')
  end subroutine
  subroutine refraction()
    call check(c_refraction(dom), 'refraction')
  end subroutine
  !> e = ef (src/umwm_top.F90:113)
  subroutine update()
    call check(c_update(dom), 'update')
  end subroutine
  subroutine diag()
    call check(c_diag(dom), 'diag')
  end subroutine
  subroutine end_first()
    call check(c_end_first(dom), 'end_first')
  end subroutine

  !> stress(option), option 'atm' or 'ocn' (src/umwm_stress.F90:17-19)
  subroutine stress(option)
    character(3), intent(in) :: option
    call check(umwm_hip_stress(dom, c_str(option)), 'stress')
  end subroutine stress

  !> stokes_drift(option), option optional 'init' (src/umwm_stokes.F90:13-20)
  subroutine stokes_drift(option)
    character(4), intent(in), optional :: option
    integer(c_int) :: init
    init = 0
    if (present(option)) then
      if (option == 'init') init = 1
    end if
    call check(umwm_hip_stokes_drift(dom, init), 'stokes_drift')
  end subroutine stokes_drift

end module umwm_gpu_ops
