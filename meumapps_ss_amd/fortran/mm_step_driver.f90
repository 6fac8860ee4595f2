! mm_step_driver.f90 -- the loop body of Multi_MEUMAPPS.f:548-1106 driven from
! Fortran through mm_abi, as a maintainer's driver would: the parameter block
! (the ModGparams state after set-up) and the fields come in as files, the
! step runs on the GPU, the fields and the fort.98 rows go back out.
!
!   mm_step_driver params.bin fields.bin nsteps precision out.bin
!
! params.bin: the bytes of one mm_params; fields.bin: phi(V,Nx,Ny,Nz),
! Cons(E,Nx,Ny,Nz) as binary64 in Fortran order, no record markers (stream
! access); out.bin: phi and Cons after nsteps steps, the same layout.  The
! diagnostics rows go to standard output in the fort.98 format (i8,10e15.7)
! of Multi_MEUMAPPS.f:905-906.
program mm_step_driver
  use mm_abi
  implicit none
  type(mm_params) :: p
  type(mm_diag), allocatable :: d(:)
  type(c_ptr) :: ctx
  real(c_double), allocatable :: phi(:), cons(:), el_en(:), u(:), gradu(:)
  character(len=512) :: a_par, a_fld, a_n, a_prec, a_out
  integer :: nsteps, prec, rc, n3, i, u_io
  call get_command_argument(1, a_par)
  call get_command_argument(2, a_fld)
  call get_command_argument(3, a_n)
  call get_command_argument(4, a_prec)
  call get_command_argument(5, a_out)
  read(a_n, *) nsteps
  read(a_prec, *) prec
  open(newunit=u_io, file=trim(a_par), access='stream', form='unformatted', status='old')
  read(u_io) p
  close(u_io)
  n3 = p%nx * p%ny * p%nz
  allocate(phi(p%vartot * n3), cons(p%numel * n3), el_en(n3), u(3 * n3), gradu(9 * n3))
  open(newunit=u_io, file=trim(a_fld), access='stream', form='unformatted', status='old')
  read(u_io) phi, cons
  close(u_io)
  el_en = 0d0
  u = 0d0
  gradu = 0d0
  rc = mm_create(p, prec, ctx)
  call check(rc, 'mm_create')
  rc = mm_upload(ctx, phi, cons, el_en, u, gradu)
  call check(rc, 'mm_upload')
  allocate(d(nsteps))
  rc = mm_step(ctx, nsteps, d)
  call check(rc, 'mm_step')
  rc = mm_download(ctx, phi, cons, el_en, u, gradu)
  call check(rc, 'mm_download')
  call mm_destroy(ctx)
  do i = 1, nsteps
     write(*, '(i8,10e15.7)') d(i)%step, d(i)%fint, d(i)%fw, d(i)%fch, d(i)%elen, d(i)%etot, &
          d(i)%glob_min, d(i)%glob_max, d(i)%phi_max, d(i)%solid_tot
  end do
  open(newunit=u_io, file=trim(a_out), access='stream', form='unformatted', status='replace')
  write(u_io) phi, cons
  close(u_io)
contains
  subroutine check(code, what)
    integer, intent(in) :: code
    character(len=*), intent(in) :: what
    if (code /= MM_OK) then
       write(*, '(a,a,a)') what, ': ', mm_errstr(code)
       stop 1
    end if
  end subroutine check
end program mm_step_driver
