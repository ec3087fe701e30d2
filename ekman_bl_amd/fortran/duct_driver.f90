! duct_driver.f90 -- a Fortran host that drives the GPU duct step through the
! C ABI, in the shape of the reference's main program (PROGRAM DIABLO,
! src/diablo.F90:58-190): read the run directory, build the grid metrics,
! set the initial field, then DO TIME_STEP ... DO RK_STEP=1,3: CALL RK_DUCT.
!
! Usage: duct_driver <run_dir>
!   <run_dir>/duct.nml     namelist &duct: p (type ekm_params), n_steps, and
!                          use_graph (1: each substep replayed from a HIP graph)
!   <run_dir>/GRID_YZ.dat  z y per line, K fastest (the reference's grid file,
!                          read as at src/diablo.F90:1120-1128)
!   <run_dir>/ic.bin       U1, U2, U3, TH physical fields, raw float64,
!                          (0:NX+1,0:NZ+1,0:NY+1) each
! Writes <run_dir>/state.bin (CU1, CU2, CU3, CP, CTH, complex128 Fourier
! fields) and prints one line per step like the reference's screen output.
program duct_driver
  use iso_c_binding
  use ekman_bl_amd
  implicit none
  type(ekm_params) :: p
  type(ekm_diag) :: d
  character(len=1024) :: dir
  integer :: n_steps, step, rk, u, ios, j, k, use_graph
  integer(c_int) :: st
  integer(kind=8) :: nphys, nfour, t0, t1, rate
  real(c_double), allocatable :: xpoint(:,:), ypoint(:,:)
  real(c_double), allocatable :: u1(:), u2(:), u3(:), th(:)
  real(c_double), allocatable :: cu1(:), cu2(:), cu3(:), cp(:), cth(:)
  namelist /duct/ p, n_steps, use_graph

  call get_command_argument(1, dir)
  if (len_trim(dir) == 0) dir = '.'
  call ekm_params_default(p)
  n_steps = 1
  use_graph = 0
  open(newunit=u, file=trim(dir)//'/duct.nml', status='old', action='read', iostat=ios)
  if (ios /= 0) stop 'duct_driver: cannot open duct.nml'
  read(u, nml=duct)
  close(u)
  ! this driver has no TH_BAR file: a scalar needs the constant stratification
  ! (src/diablo.F90:1152-1160)
  if (p%n_th > 0 .and. p%cont_strat == 0) stop 'duct_driver: N_TH = 1 needs CONSTANT_STRATIFICATION'

  allocate(xpoint(0:p%nz+1, 0:p%ny+1), ypoint(0:p%nz+1, 0:p%ny+1))
  open(newunit=u, file=trim(dir)//'/GRID_YZ.dat', status='old', action='read', iostat=ios)
  if (ios /= 0) stop 'duct_driver: cannot open GRID_YZ.dat'
  do j = 0, p%ny + 1
     do k = 0, p%nz + 1
        read(u, *) xpoint(k, j), ypoint(k, j)
     end do
  end do
  close(u)

  nphys = int(p%nx + 2, 8) * int(p%nz + 2, 8) * int(p%ny + 2, 8)
  nfour = 2_8 * int(p%nx / 2 + 1, 8) * int(p%nz + 2, 8) * int(p%ny + 2, 8)
  allocate(u1(nphys), u2(nphys), u3(nphys), th(nphys))
  open(newunit=u, file=trim(dir)//'/ic.bin', access='stream', form='unformatted', status='old', &
       action='read', iostat=ios)
  if (ios /= 0) stop 'duct_driver: cannot open ic.bin'
  read(u) u1, u2, u3, th
  close(u)

  call ekm_check(ekm_create(p, xpoint, ypoint, c_null_ptr, ekm_handle), 'ekm_create')
  call ekm_check(ekm_set_initial_physical(ekm_handle, u1, u2, u3, th), 'CREATE_FLOW_DUCT')
  call ekm_check(ekm_get_diag(ekm_handle, d), 'ekm_get_diag')
  write(6, '(a,es12.4)') ' Divergence after the initial projection: ', d%div_ref

  call system_clock(t0, rate)
  do step = 1, n_steps
     do rk = 1, 3
        if (use_graph /= 0) then
           call RK_DUCT_GRAPH_GPU()
        else
           call RK_DUCT_GPU()
        end if
     end do
     call ekm_check(ekm_get_diag(ekm_handle, d), 'ekm_get_diag')
     write(6, '(i6,a9,f15.8,a7,f15.8,a7,es12.4,a9,i6)') d%time_step, '    time=', d%time, '    dt=', &
          d%delta_t, '   div=', d%div_ref, '   cycles=', d%mg_cycles
     if (d%mg_unconverged > 0) write(6, '(a,i6,a)') ' WARNING: ', d%mg_unconverged, &
          ' pressure solves stopped at proj_maxit above proj_tol'
  end do
  call ekm_check(ekm_sync(ekm_handle), 'ekm_sync')
  call system_clock(t1)
  write(6, '(a,f12.6)') ' Seconds per Iteration: ', real(t1 - t0, 8) / real(rate, 8) / max(n_steps, 1)

  allocate(cu1(nfour), cu2(nfour), cu3(nfour), cp(nfour), cth(nfour))
  call ekm_check(ekm_get_state(ekm_handle, cu1, cu2, cu3, cp, cth), 'ekm_get_state')
  open(newunit=u, file=trim(dir)//'/state.bin', access='stream', form='unformatted', status='replace', &
       action='write')
  write(u) cu1, cu2, cu3, cp, cth
  close(u)
  call ekm_check(ekm_destroy(ekm_handle), 'ekm_destroy')
end program duct_driver
