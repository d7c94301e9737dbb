! Driver of the Fortran binding (melt_ibm_f) used by tests/test_gpu_fortran.py.
! It calls the library the way the reference's time loop does (gcurv.f90:62,
! 172; invtr1.f90:66-68): module-shaped arrays q(1:n1,1:n2,0:n3m+1) with
! n1 = n1m+1 (the reference's padding column and ghost planes) pushed and
! pulled through melt_push / melt_pull, the initial tagCells, tschem steps,
! and the reference's own solxi / solxj / solxk(q(1:n1,1:n2,1:n3m), betadx)
! sequence on an rhs.  Inputs and outputs are raw stream files in the
! directory named by the first argument.
program drive_tschem
  use iso_c_binding
  use melt_ibm_f
  implicit none
  type(melt_config_t) :: cfg
  integer :: u, i, n1, n3m, nv, ne, nf, nsteps
  integer, parameter :: nin = 5, nout = 6
  integer :: ids_in(nin) = [MELT_VX, MELT_VY, MELT_VZ, MELT_TEMP, MELT_FORCX]
  integer :: ids_out(nout) = [MELT_VX, MELT_VY, MELT_VZ, MELT_TEMP, MELT_PR, MELT_VOFP]
  real(c_double) :: rad, dt, bx, by, bz
  real(c_double), allocatable :: q(:, :, :), xyz0(:, :)
  integer(c_int32_t), allocatable :: voe(:, :), eof(:, :)
  character(len=512) :: dir

  call get_command_argument(1, dir)
  open(newunit=u, file=trim(dir)//'/case.bin', access='stream', form='unformatted', status='old')
  read(u) cfg
  read(u) nv, ne, nf, nsteps
  read(u) rad, dt, bx, by, bz
  allocate(xyz0(3, nv), voe(2, ne), eof(3, nf))
  read(u) xyz0, voe, eof
  n1 = cfg%n1m + 1
  n3m = cfg%n3m
  allocate(q(1:n1, 1:n1, 0:n3m + 1))

  call melt_open(cfg)
  do i = 1, nin
    read(u) q
    call melt_push(ids_in(i), q, n1, n1)
  end do
  call melt_geometry(nv, ne, nf, xyz0, voe, eof, rad)
  call melt_dt(dt)
  call tagCells()
  do i = 1, nsteps
    call tschem()
  end do

  open(newunit=i, file=trim(dir)//'/out.bin', access='stream', form='unformatted', status='replace')
  do nv = 1, nout
    q = -7.0d0
    call melt_pull(ids_out(nv), q, n1, n1)
    write(i) q
  end do

  ! the reference's line-solve sequence on a given rhs, q updated in place
  read(u) q
  call melt_push(MELT_RHS, q, n1, n1)
  read(u) q
  call melt_stage(1)
  call solxi(bx)
  call solxj(by)
  call solxk(q(1:n1, 1:n1, 1:n3m), bz)
  write(i) q
  close(i)
  close(u)
  call melt_close()
end program drive_tschem
