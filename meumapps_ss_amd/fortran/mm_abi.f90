! mm_abi.f90 -- ISO_C_BINDING interface to include/meumapps_ss.h.
!
! A Fortran driver that keeps the reference's Input.txt parsing, set-up,
! initial conditions and dumps "use"s this module and replaces the loop body
! of Multi_MEUMAPPS.f:548-1106 by mm_step.  The derived types are
! interoperable mirrors of mm_params / mm_diag; array members keep the
! Fortran index order of ModGparams (e.g. cela(i,j,p) <-> CEla(i,j,p)).
module mm_abi
  use, intrinsic :: iso_c_binding
  implicit none

  integer(c_int32_t), parameter :: MM_MAX_VAR = 16, MM_MAX_PH = 4, MM_MAX_EL = 4
  integer(c_int32_t), parameter :: MM_GP_STRIDE = 2 * MM_MAX_EL + 1
  integer(c_int32_t), parameter :: MM_FP64 = 0, MM_MIXED = 1
  integer(c_int32_t), parameter :: MM_OK = 0, MM_ERR_ZSOLVE = 2, MM_ERR_QINV = 3, &
       MM_ERR_CINV_ITER = 4, MM_ERR_CINV_ENER = 6, MM_ERR_NEWFAC = 7, MM_ERR_ARG = 20001

  type, bind(c) :: mm_params
     integer(c_int32_t) :: nx, ny, nz
     integer(c_int32_t) :: vartot, numel, numph
     integer(c_int32_t) :: nvar(MM_MAX_PH)
     integer(c_int32_t) :: step, elastep0, noise_step, pad0
     real(c_double) :: t_step, b_mob
     real(c_double) :: cmobs(MM_MAX_EL)
     real(c_double) :: grad_coeff(3, 3, MM_MAX_VAR)
     real(c_double) :: trans_ppt(6, MM_MAX_VAR)
     real(c_double) :: cela(6, 6, MM_MAX_PH)
     real(c_double) :: sela(6, 6, MM_MAX_PH)
     real(c_double) :: gparams(MM_GP_STRIDE, MM_MAX_PH)
     real(c_double) :: wellph(MM_MAX_PH)
     real(c_double) :: e_mean(6)
  end type mm_params

  type, bind(c) :: mm_diag
     integer(c_int32_t) :: step, n_elastic
     real(c_double) :: fint, fw, fch, elen, etot
     real(c_double) :: glob_min, glob_max, phi_max, solid_tot
     real(c_double) :: caviter(MM_MAX_EL)
     real(c_double) :: cfac
  end type mm_diag

  interface
     integer(c_int) function mm_create(p, precision, ctx) bind(c, name="mm_create")
       import :: c_int, c_ptr, mm_params
       type(mm_params), intent(in) :: p
       integer(c_int), value :: precision
       type(c_ptr), intent(out) :: ctx
     end function mm_create

     subroutine mm_destroy(ctx) bind(c, name="mm_destroy")
       import :: c_ptr
       type(c_ptr), value :: ctx
     end subroutine mm_destroy

     ! phi(V,Nx,Ny,Nz), cons(E,Nx,Ny,Nz), el_en(Nx,Ny,Nz), u(Nx,Ny,Nz,3), gradu(Nx,Ny,Nz,9)
     integer(c_int) function mm_upload(ctx, phi, cons, el_en, u, gradu) bind(c, name="mm_upload")
       import :: c_int, c_ptr, c_double
       type(c_ptr), value :: ctx
       real(c_double), intent(in) :: phi(*), cons(*), el_en(*), u(*), gradu(*)
     end function mm_upload

     integer(c_int) function mm_download(ctx, phi, cons, el_en, u, gradu) bind(c, name="mm_download")
       import :: c_int, c_ptr, c_double
       type(c_ptr), value :: ctx
       real(c_double), intent(out) :: phi(*), cons(*), el_en(*), u(*), gradu(*)
     end function mm_download

     integer(c_int) function mm_step(ctx, nsteps, diag) bind(c, name="mm_step")
       import :: c_int, c_ptr, c_int32_t, mm_diag
       type(c_ptr), value :: ctx
       integer(c_int32_t), value :: nsteps
       type(mm_diag), intent(out) :: diag(*)
     end function mm_step

     integer(c_int) function mm_get_step(ctx) bind(c, name="mm_get_step")
       import :: c_int, c_ptr
       type(c_ptr), value :: ctx
     end function mm_get_step

     integer(c_int) function mm_set_step(ctx, step) bind(c, name="mm_set_step")
       import :: c_int, c_ptr, c_int32_t
       type(c_ptr), value :: ctx
       integer(c_int32_t), value :: step
     end function mm_set_step

     type(c_ptr) function mm_errstr_c(code) bind(c, name="mm_errstr")
       import :: c_ptr, c_int
       integer(c_int), value :: code
     end function mm_errstr_c

     integer(c_int) function mm_ic_uniform(nx, ny, nz, nvar, nel, iseed, amp, cav, phi, cons) &
          bind(c, name="mm_ic_uniform")
       import :: c_int, c_int32_t, c_double
       integer(c_int32_t), value :: nx, ny, nz, nvar, nel, iseed
       real(c_double), value :: amp
       real(c_double), intent(in) :: cav(*)
       real(c_double), intent(out) :: phi(*), cons(*)
     end function mm_ic_uniform
  end interface

contains

  ! Message of a return code as a Fortran string (for the STOP text).
  function mm_errstr(code) result(msg)
    integer(c_int), intent(in) :: code
    character(len=:), allocatable :: msg
    type(c_ptr) :: p
    character(kind=c_char), pointer :: s(:)
    integer :: n
    interface
       integer(c_size_t) function c_strlen(str) bind(c, name="strlen")
         import :: c_size_t, c_ptr
         type(c_ptr), value :: str
       end function c_strlen
    end interface
    p = mm_errstr_c(code)
    n = int(c_strlen(p))
    call c_f_pointer(p, s, [n])
    allocate(character(len=n) :: msg)
    msg = transfer(s(1:n), msg)
  end function mm_errstr

end module mm_abi
