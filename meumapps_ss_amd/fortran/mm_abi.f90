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

  ! dL/d of the mm_params arrays of the same names (adjoint)
  type, bind(c) :: mm_adj_grads
     real(c_double) :: grad_coeff(3, 3, MM_MAX_VAR)
     real(c_double) :: wellph(MM_MAX_PH)
     real(c_double) :: cmobs(MM_MAX_EL)
     real(c_double) :: b_mob
     real(c_double) :: trans_ppt(6, MM_MAX_VAR)
  end type mm_adj_grads

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

     ! ran_2.f:1-42 with its SAVE state in state(35) (all zero at program start)
     real(c_double) function mm_ran2(idum, state) bind(c, name="mm_ran2")
       import :: c_double, c_int32_t
       integer(c_int32_t), intent(inout) :: idum
       integer(c_int32_t), intent(inout) :: state(35)
     end function mm_ran2

     ! n consecutive ran_2 draws
     integer(c_int) function mm_ran2_fill(idum, state, n, out) bind(c, name="mm_ran2_fill")
       import :: c_int, c_double, c_int32_t, c_int64_t
       integer(c_int32_t), intent(inout) :: idum
       integer(c_int32_t), intent(inout) :: state(35)
       integer(c_int64_t), value :: n
       real(c_double), intent(out) :: out(*)
     end function mm_ran2_fill

     ! ModFunc.f:77-100 on the ran_2 stream (idum, state(35))
     real(c_double) function mm_langevin(idum, state, sd) bind(c, name="mm_langevin")
       import :: c_double, c_int32_t
       integer(c_int32_t), intent(inout) :: idum
       integer(c_int32_t), intent(inout) :: state(35)
       real(c_double), value :: sd
     end function mm_langevin

     ! the noise loop of one step for one process (Multi_MEUMAPPS.f:554-562)
     integer(c_int) function mm_noise_accumulate(idum, state, sd, nvar, n3, phi_noise) &
          bind(c, name="mm_noise_accumulate")
       import :: c_int, c_double, c_int32_t, c_int64_t
       integer(c_int32_t), intent(inout) :: idum
       integer(c_int32_t), intent(inout) :: state(35)
       real(c_double), value :: sd
       integer(c_int32_t), value :: nvar
       integer(c_int64_t), value :: n3
       real(c_double), intent(inout) :: phi_noise(*)
     end function mm_noise_accumulate

     ! noise of the steps step <= noise_step: a restart run passes
     ! idum = iseed - myid and c_null_ptr for the program-start state
     integer(c_int) function mm_set_noise(ctx, sd, idum, state) bind(c, name="mm_set_noise")
       import :: c_int, c_ptr, c_double, c_int32_t
       type(c_ptr), value :: ctx
       real(c_double), value :: sd
       integer(c_int32_t), value :: idum
       type(c_ptr), value :: state
     end function mm_set_noise

     integer(c_int) function mm_get_noise_state(ctx, idum, state) bind(c, name="mm_get_noise_state")
       import :: c_int, c_ptr, c_int32_t
       type(c_ptr), value :: ctx
       integer(c_int32_t), intent(out) :: idum
       integer(c_int32_t), intent(out) :: state(35)
     end function mm_get_noise_state

     integer(c_int) function mm_set_graph(ctx, on, iters) bind(c, name="mm_set_graph")
       import :: c_int, c_ptr, c_int32_t
       type(c_ptr), value :: ctx
       integer(c_int32_t), value :: on, iters
     end function mm_set_graph

     ! ensembles: params(n); diag(n, nsteps)
     integer(c_int) function mm_batch_create(n, params, precision, batch) bind(c, name="mm_batch_create")
       import :: c_int, c_ptr, c_int32_t, mm_params
       integer(c_int32_t), value :: n
       type(mm_params), intent(in) :: params(*)
       integer(c_int32_t), value :: precision
       type(c_ptr), intent(out) :: batch
     end function mm_batch_create

     integer(c_int) function mm_batch_member(batch, i, ctx) bind(c, name="mm_batch_member")
       import :: c_int, c_ptr, c_int32_t
       type(c_ptr), value :: batch
       integer(c_int32_t), value :: i
       type(c_ptr), intent(out) :: ctx
     end function mm_batch_member

     integer(c_int) function mm_batch_step(batch, nsteps, diag) bind(c, name="mm_batch_step")
       import :: c_int, c_ptr, c_int32_t, mm_diag
       type(c_ptr), value :: batch
       integer(c_int32_t), value :: nsteps
       type(mm_diag), intent(out) :: diag(*)
     end function mm_batch_step

     subroutine mm_batch_destroy(batch) bind(c, name="mm_batch_destroy")
       import :: c_ptr
       type(c_ptr), value :: batch
     end subroutine mm_batch_destroy

     ! adjoint and device checkpoints
     integer(c_int) function mm_adj_enable(ctx) bind(c, name="mm_adj_enable")
       import :: c_int, c_ptr
       type(c_ptr), value :: ctx
     end function mm_adj_enable

     integer(c_int) function mm_adj_zero(ctx) bind(c, name="mm_adj_zero")
       import :: c_int, c_ptr
       type(c_ptr), value :: ctx
     end function mm_adj_zero

     ! lam_phi(Nx,Ny,Nz,V), lam_cons(Nx,Ny,Nz,E), lam_strain(Nx,Ny,Nz,6); c_null_ptr skips
     integer(c_int) function mm_adj_set(ctx, lam_phi, lam_cons, lam_strain) bind(c, name="mm_adj_set")
       import :: c_int, c_ptr
       type(c_ptr), value :: ctx, lam_phi, lam_cons, lam_strain
     end function mm_adj_set

     integer(c_int) function mm_adj_get(ctx, lam_phi, lam_cons, lam_strain) bind(c, name="mm_adj_get")
       import :: c_int, c_ptr
       type(c_ptr), value :: ctx, lam_phi, lam_cons, lam_strain
     end function mm_adj_get

     integer(c_int) function mm_state_bytes(ctx, bytes) bind(c, name="mm_state_bytes")
       import :: c_int, c_ptr, c_int64_t
       type(c_ptr), value :: ctx
       integer(c_int64_t), intent(out) :: bytes
     end function mm_state_bytes

     integer(c_int) function mm_state_save(ctx, dev_dst) bind(c, name="mm_state_save")
       import :: c_int, c_ptr
       type(c_ptr), value :: ctx, dev_dst
     end function mm_state_save

     integer(c_int) function mm_state_load(ctx, dev_src) bind(c, name="mm_state_load")
       import :: c_int, c_ptr
       type(c_ptr), value :: ctx, dev_src
     end function mm_state_load

     integer(c_int) function mm_adj_tape_step(ctx, diag) bind(c, name="mm_adj_tape_step")
       import :: c_int, c_ptr, mm_diag
       type(c_ptr), value :: ctx
       type(mm_diag), intent(out) :: diag
     end function mm_adj_tape_step

     integer(c_int) function mm_adj_step(ctx, grads) bind(c, name="mm_adj_step")
       import :: c_int, c_ptr, mm_adj_grads
       type(c_ptr), value :: ctx
       type(mm_adj_grads), intent(inout) :: grads
     end function mm_adj_step
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
