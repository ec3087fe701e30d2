! ekman_bl_amd_mod.f90 -- ISO_C_BINDING interfaces of libekman_bl_amd.so and
! Fortran wrappers with the reference's subroutine surface.
!
! A DIABLO build keeps its own driver and calls these in place of its CPU
! routines:
!   RK_DUCT                          (src/duct.F90:237)      -> ekm_rk_duct, ekm_rk_duct_graph_launch
!   REM_DIV_CURV                     (src/duct.F90:2472)     -> ekm_rem_div_curv
!   FFT_X_TO_FOURIER_OP / _PHYSICAL_OP (calls src/mpi_duct.F90:290,308)
!                                                            -> ekm_fft_x_to_fourier / _physical
!   THOMAS_REAL(A,B,C,G,NY,NX)       (src/boundary.F90:53)   -> ekm_thomas_real
!   THOMAS_REAL_SP(A,B,C,G,D,K,NY,NX) (src/boundary.F90:4)   -> ekm_thomas_real_sp
!   MULTIGRID_MUDPACK(phi)           (src/mg_solver2d.F90:75) -> ekm_multigrid_mudpack
! The duct state lives on the GPU behind one opaque handle (ekm_handle); the
! module-global arrays of the reference are uploaded once (ekm_set_state /
! ekm_set_initial_physical) and read back only for I/O (ekm_get_state).
module ekman_bl_amd
  use iso_c_binding
  implicit none
  private :: c_ptr, c_int, c_double

  integer(c_int), parameter :: EKM_OK = 0, EKM_ERR_ARG = -1, EKM_ERR_HIP = -2, &
       EKM_ERR_UNSUPPORTED = -3, EKM_ERR_DIV_NAN = -4, EKM_ERR_DT_SMALL = -5, EKM_ERR_NODEV = -6
  integer, parameter :: EKM_YMIN = 1, EKM_YMAX = 2, EKM_ZMIN = 3, EKM_ZMAX = 4

  ! struct ekm_params (include/ekman_bl_amd.h); C arrays [3][4] appear as (4,3)
  type, bind(c) :: ekm_params
     integer(c_int) :: nx, ny, nz, n_th
     integer(c_int) :: mg_ixp, mg_jyq, mg_iex, mg_jey
     real(c_double) :: nu, lx, ly, lz
     real(c_double) :: delta_t
     integer(c_int) :: variable_dt
     real(c_double) :: cfl
     integer(c_int) :: update_dt
     integer(c_int) :: f_type
     real(c_double) :: wbulk0, px0, omega0, amp_omega0, f0
     integer(c_int) :: u_bc(4, 3)
     real(c_double) :: u_bc_c1(4, 3)
     real(c_double) :: ri_tau, pr
     integer(c_int) :: cont_strat
     integer(c_int) :: th_bc(4)
     real(c_double) :: th_bc_c1(4)
     integer(c_int) :: mudpack_bc(4)
     integer(c_int) :: mg_maxcy
     real(c_double) :: mg_tolmax
     integer(c_int) :: cmplx_fp32
     integer(c_int) :: coriolis_mean_only
     integer(c_int) :: proj_mode
     integer(c_int) :: proj_maxit
     real(c_double) :: proj_tol
     integer(c_int) :: proj_guess
     integer(c_int) :: proj_krylov
     integer(c_int) :: les_model
     integer(c_int) :: sponge_type
     integer(c_int) :: ny_s_t
     integer(c_int) :: nz_s_r
     integer(c_int) :: save_stats_int
     integer(c_int) :: nz_s_l
     integer(c_int) :: wall_model_type
  end type ekm_params

  type, bind(c) :: ekm_diag
     real(c_double) :: div_ref, div_max, delta_t, time
     integer(c_int) :: time_step, rk_step, mg_solves, mg_cycles, mg_max_cycles, status
     integer(c_int) :: mg_unconverged
  end type ekm_diag

  ! ekm_tuning: kernel launch and memory-placement parameters
  type, bind(c) :: ekm_tuning
     integer(c_int) :: adi_rows, adi_block, adi_les_rows, mg_block, tight_lds_kib
  end type ekm_tuning

  type(c_ptr), save :: ekm_handle = c_null_ptr

  interface
     subroutine ekm_params_default(p) bind(c, name="ekm_params_default")
       import :: ekm_params
       type(ekm_params), intent(out) :: p
     end subroutine ekm_params_default

     integer(c_int) function ekm_create(p, xpoint, ypoint, th_bar, s) bind(c, name="ekm_create")
       import :: ekm_params, c_double, c_ptr, c_int
       type(ekm_params), intent(in) :: p
       real(c_double), intent(in) :: xpoint(*), ypoint(*)
       type(c_ptr), value :: th_bar
       type(c_ptr), intent(out) :: s
     end function ekm_create

     integer(c_int) function ekm_destroy(s) bind(c, name="ekm_destroy")
       import :: c_ptr, c_int
       type(c_ptr), value :: s
     end function ekm_destroy

     integer(c_int) function ekm_set_state(s, cu1, cu2, cu3, cp, cth, time, time_step) bind(c, name="ekm_set_state")
       import :: c_ptr, c_int, c_double
       type(c_ptr), value :: s
       real(c_double), intent(in) :: cu1(*), cu2(*), cu3(*), cp(*), cth(*)
       real(c_double), value :: time
       integer(c_int), value :: time_step
     end function ekm_set_state

     integer(c_int) function ekm_get_state(s, cu1, cu2, cu3, cp, cth) bind(c, name="ekm_get_state")
       import :: c_ptr, c_int, c_double
       type(c_ptr), value :: s
       real(c_double), intent(out) :: cu1(*), cu2(*), cu3(*), cp(*), cth(*)
     end function ekm_get_state

     ! the subgrid models' NU_T / C_DYN and KAPPA_T / C_DYN_TH (les_model = 2)
     integer(c_int) function ekm_get_les(s, nut, cdyn) bind(c, name="ekm_get_les")
       import :: c_ptr, c_int, c_double
       type(c_ptr), value :: s
       real(c_double), intent(out) :: nut(*), cdyn(*)
     end function ekm_get_les

     integer(c_int) function ekm_get_les_th(s, kappa, cdyn_th) bind(c, name="ekm_get_les_th")
       import :: c_ptr, c_int, c_double
       type(c_ptr), value :: s
       real(c_double), intent(out) :: kappa(*), cdyn_th(*)
     end function ekm_get_les_th

     ! x statistics of SAVE_STATS_CURVI (16 planes of (NZ+2)*(NY+2))
     integer(c_int) function ekm_get_xstats(s, out) bind(c, name="ekm_get_xstats")
       import :: c_ptr, c_int, c_double
       type(c_ptr), value :: s
       real(c_double), intent(out) :: out(*)
     end function ekm_get_xstats

     ! the mean-gradient statistics of SAVE_STATS_CURVI: 8 planes (dudy, dwdy,
     ! dudz, dwdz, dthdy, dthdz, Rig, TKE) of (NZ+2)*(NY+2), then THBULK
     integer(c_int) function ekm_get_mean_grads(s, out) bind(c, name="ekm_get_mean_grads")
       import :: c_ptr, c_int, c_double
       type(c_ptr), value :: s
       real(c_double), intent(out) :: out(*)
     end function ekm_get_mean_grads

     ! the higher-order x moments of high_oder_momentums_curvi: 13 planes
     ! (UU VV WW UUU VVV WWW UUUU VVVV WWWW UUV WWV UUUV WWWV)
     integer(c_int) function ekm_get_omega_x(s, out) bind(c, name="ekm_get_omega_x")
       import :: c_ptr, c_int, c_double
       type(c_ptr), value :: s
       real(c_double), intent(out) :: out(*)
     end function ekm_get_omega_x

     integer(c_int) function ekm_get_physical(s, f, out) bind(c, name="ekm_get_physical")
       import :: c_ptr, c_int, c_double
       type(c_ptr), value :: s
       integer(c_int), value :: f
       real(c_double), intent(out) :: out(*)
     end function ekm_get_physical

     integer(c_int) function ekm_get_plane(s, axis, index, out) bind(c, name="ekm_get_plane")
       import :: c_ptr, c_int, c_double
       type(c_ptr), value :: s
       integer(c_int), value :: axis, index
       real(c_double), intent(out) :: out(*)
     end function ekm_get_plane

     integer(c_int) function ekm_get_xmoments(s, out) bind(c, name="ekm_get_xmoments")
       import :: c_ptr, c_int, c_double
       type(c_ptr), value :: s
       real(c_double), intent(out) :: out(*)
     end function ekm_get_xmoments

     ! SPONGE_SIGMA(0:NY+1) and SPONGE_SIGMA_OUT(0:NZ+1) of sponge_th's profiles
     integer(c_int) function ekm_host_sponge_sigma(stype, nz, ny, ny_s_t, ypoint, sig) &
         bind(c, name="ekm_host_sponge_sigma")
       import :: c_int, c_double
       integer(c_int), value :: stype, nz, ny, ny_s_t
       real(c_double), intent(in) :: ypoint(*)
       real(c_double), intent(out) :: sig(*)
     end function ekm_host_sponge_sigma

     integer(c_int) function ekm_host_sponge_sigma_out(stype, nz, ny, nz_s_r, nz_s_l, xpoint, sig) &
         bind(c, name="ekm_host_sponge_sigma_out")
       import :: c_int, c_double
       integer(c_int), value :: stype, nz, ny, nz_s_r, nz_s_l
       real(c_double), intent(in) :: xpoint(*)
       real(c_double), intent(out) :: sig(*)
     end function ekm_host_sponge_sigma_out

     ! PX0 and Qn of the F_TYPE 0 flow-rate controller
     integer(c_int) function ekm_get_forcing(s, px0, qn) bind(c, name="ekm_get_forcing")
       import :: c_ptr, c_int, c_double
       type(c_ptr), value :: s
       real(c_double), intent(out) :: px0, qn
     end function ekm_get_forcing

     ! cycles of every real solve of the last projection (MUDPACK's iparm(17))
     integer(c_int) function ekm_get_solve_cycles(s, cycles, n) bind(c, name="ekm_get_solve_cycles")
       import :: c_ptr, c_int
       type(c_ptr), value :: s
       integer(c_int), intent(out) :: cycles(*)
       integer(c_int), value :: n
     end function ekm_get_solve_cycles

     integer(c_int) function ekm_set_initial_physical(s, u1, u2, u3, th) bind(c, name="ekm_set_initial_physical")
       import :: c_ptr, c_int, c_double
       type(c_ptr), value :: s
       real(c_double), intent(in) :: u1(*), u2(*), u3(*), th(*)
     end function ekm_set_initial_physical

     ! kernel tuning (launch and memory placement; results do not depend on it)
     subroutine ekm_tuning_default(t) bind(c, name="ekm_tuning_default")
       import :: ekm_tuning
       type(ekm_tuning), intent(out) :: t
     end subroutine ekm_tuning_default

     integer(c_int) function ekm_create_tuned(p, xpoint, ypoint, th_bar, t, s) bind(c, name="ekm_create_tuned")
       import :: c_ptr, c_int, c_double, ekm_params, ekm_tuning
       type(ekm_params), intent(in) :: p
       real(c_double), intent(in) :: xpoint(*), ypoint(*)
       type(c_ptr), value :: th_bar
       type(ekm_tuning), intent(in) :: t
       type(c_ptr), intent(out) :: s
     end function ekm_create_tuned

     integer(c_int) function ekm_get_tuning(s, t) bind(c, name="ekm_get_tuning")
       import :: c_ptr, c_int, ekm_tuning
       type(c_ptr), value :: s
       type(ekm_tuning), intent(out) :: t
     end function ekm_get_tuning

     integer(c_int) function ekm_rk_duct(s) bind(c, name="ekm_rk_duct")
       import :: c_ptr, c_int
       type(c_ptr), value :: s
     end function ekm_rk_duct

     integer(c_int) function ekm_time_step(s, nsteps) bind(c, name="ekm_time_step")
       import :: c_ptr, c_int
       type(c_ptr), value :: s
       integer(c_int), value :: nsteps
     end function ekm_time_step

     ! the substep replayed from a captured HIP graph (same results)
     integer(c_int) function ekm_rk_duct_graph_launch(s) bind(c, name="ekm_rk_duct_graph_launch")
       import :: c_ptr, c_int
       type(c_ptr), value :: s
     end function ekm_rk_duct_graph_launch

     integer(c_int) function ekm_time_step_ex(s, nsteps, graph) bind(c, name="ekm_time_step_ex")
       import :: c_ptr, c_int
       type(c_ptr), value :: s
       integer(c_int), value :: nsteps, graph
     end function ekm_time_step_ex

     integer(c_int) function ekm_rem_div_curv(s) bind(c, name="ekm_rem_div_curv")
       import :: c_ptr, c_int
       type(c_ptr), value :: s
     end function ekm_rem_div_curv

     integer(c_int) function ekm_get_diag(s, d) bind(c, name="ekm_get_diag")
       import :: c_ptr, c_int, ekm_diag
       type(c_ptr), value :: s
       type(ekm_diag), intent(out) :: d
     end function ekm_get_diag

     integer(c_int) function ekm_sync(s) bind(c, name="ekm_sync")
       import :: c_ptr, c_int
       type(c_ptr), value :: s
     end function ekm_sync

     integer(c_int) function ekm_fft_x_to_fourier(varp, nx, npencil) bind(c, name="ekm_fft_x_to_fourier")
       import :: c_int, c_double
       real(c_double), intent(inout) :: varp(*)
       integer(c_int), value :: nx, npencil
     end function ekm_fft_x_to_fourier

     integer(c_int) function ekm_fft_x_to_physical(varp, nx, npencil) bind(c, name="ekm_fft_x_to_physical")
       import :: c_int, c_double
       real(c_double), intent(inout) :: varp(*)
       integer(c_int), value :: nx, npencil
     end function ekm_fft_x_to_physical

     integer(c_int) function ekm_thomas_real(a, b, c, g, ny, nx) bind(c, name="ekm_thomas_real")
       import :: c_int, c_double
       real(c_double), intent(inout) :: a(*), b(*), c(*), g(*)
       integer(c_int), value :: ny, nx
     end function ekm_thomas_real

     integer(c_int) function ekm_thomas_real_sp(a, b, c, g, d, k, ny, nx) bind(c, name="ekm_thomas_real_sp")
       import :: c_int, c_double
       real(c_double), intent(inout) :: a(*), b(*), c(*), g(*)
       real(c_double), value :: d
       integer(c_int), value :: k, ny, nx
     end function ekm_thomas_real_sp

     integer(c_int) function ekm_multigrid_mudpack(s, nsolve, kx2, rhs_tmp, phi, cycles) &
          bind(c, name="ekm_multigrid_mudpack")
       import :: c_ptr, c_int, c_double
       type(c_ptr), value :: s
       integer(c_int), value :: nsolve
       real(c_double), intent(in) :: kx2(*), rhs_tmp(*)
       real(c_double), intent(out) :: phi(*)
       integer(c_int), intent(out) :: cycles(*)
     end function ekm_multigrid_mudpack

     ! the projection's correction PHI (physical layout, (NX+2)*(NZ+2)*(NY+2))
     integer(c_int) function ekm_get_phi(s, phi) bind(c, name="ekm_get_phi")
       import :: c_ptr, c_int, c_double
       type(c_ptr), value :: s
       real(c_double), intent(out) :: phi(*)
     end function ekm_get_phi

     ! device address of state field f (0..4: CU1, CU2, CU3, CP, CTH)
     integer(c_int) function ekm_field_device_ptr(s, f, dptr) bind(c, name="ekm_field_device_ptr")
       import :: c_ptr, c_int
       type(c_ptr), value :: s
       integer(c_int), value :: f
       type(c_ptr), intent(out) :: dptr
     end function ekm_field_device_ptr

     ! device memory and the device-pointer component drop-ins
     integer(c_int) function ekm_dev_alloc(p, bytes) bind(c, name="ekm_dev_alloc")
       import :: c_ptr, c_int, c_size_t
       type(c_ptr), intent(out) :: p
       integer(c_size_t), value :: bytes
     end function ekm_dev_alloc

     integer(c_int) function ekm_dev_free(p) bind(c, name="ekm_dev_free")
       import :: c_ptr, c_int
       type(c_ptr), value :: p
     end function ekm_dev_free

     integer(c_int) function ekm_thomas_real_sp_dev(a, b, c, g, d, k, ny, nx, stream) &
         bind(c, name="ekm_thomas_real_sp_dev")
       import :: c_ptr, c_int, c_double
       type(c_ptr), value :: a, b, c, g, stream
       real(c_double), value :: d
       integer(c_int), value :: k, ny, nx
     end function ekm_thomas_real_sp_dev

     integer(c_int) function ekm_multigrid_mudpack_ex(s, nsolve, kx2, rhs_tmp, phi, cycles, iguess) &
         bind(c, name="ekm_multigrid_mudpack_ex")
       import :: c_ptr, c_int, c_double
       type(c_ptr), value :: s
       integer(c_int), value :: nsolve, iguess
       real(c_double), intent(in) :: kx2(*), rhs_tmp(*)
       real(c_double), intent(inout) :: phi(*)
       integer(c_int), intent(out) :: cycles(*)
     end function ekm_multigrid_mudpack_ex

     integer(c_int) function ekm_memcpy_h2d(dst, src, bytes) bind(c, name="ekm_memcpy_h2d")
       import :: c_ptr, c_int, c_size_t
       type(c_ptr), value :: dst, src
       integer(c_size_t), value :: bytes
     end function ekm_memcpy_h2d

     integer(c_int) function ekm_memcpy_d2h(dst, src, bytes) bind(c, name="ekm_memcpy_d2h")
       import :: c_ptr, c_int, c_size_t
       type(c_ptr), value :: dst, src
       integer(c_size_t), value :: bytes
     end function ekm_memcpy_d2h

     integer(c_int) function ekm_gather_rows(dst, src, idx, n, row_bytes, stream) bind(c, name="ekm_gather_rows")
       import :: c_ptr, c_int, c_long, c_size_t
       type(c_ptr), value :: dst, src, stream
       integer(c_long), intent(in) :: idx(*)
       integer(c_int), value :: n
       integer(c_size_t), value :: row_bytes
     end function ekm_gather_rows

     ! ---- learned initial guess of the pressure solves (no counterpart in the
     !      reference, whose MUDPACK solves start from zero) -------------------
     integer(c_int) function ekm_wsnet_create(nz2, ny2, bmax, trainable, net) bind(c, name="ekm_wsnet_create")
       import :: c_ptr, c_int
       integer(c_int), value :: nz2, ny2, bmax, trainable
       type(c_ptr), intent(out) :: net
     end function ekm_wsnet_create

     integer(c_int) function ekm_wsnet_destroy(net) bind(c, name="ekm_wsnet_destroy")
       import :: c_ptr, c_int
       type(c_ptr), value :: net
     end function ekm_wsnet_destroy

     integer(c_long) function ekm_wsnet_nparams(net) bind(c, name="ekm_wsnet_nparams")
       import :: c_ptr, c_long
       type(c_ptr), value :: net
     end function ekm_wsnet_nparams

     integer(c_int) function ekm_wsnet_set_params(net, theta) bind(c, name="ekm_wsnet_set_params")
       import :: c_ptr, c_int, c_float
       type(c_ptr), value :: net
       real(c_float), intent(in) :: theta(*)
     end function ekm_wsnet_set_params

     integer(c_int) function ekm_wsnet_get_params(net, theta) bind(c, name="ekm_wsnet_get_params")
       import :: c_ptr, c_int, c_float
       type(c_ptr), value :: net
       real(c_float), intent(out) :: theta(*)
     end function ekm_wsnet_get_params

     integer(c_int) function ekm_wsnet_set_norm(net, sigma, kx_max) bind(c, name="ekm_wsnet_set_norm")
       import :: c_ptr, c_int, c_double
       type(c_ptr), value :: net
       real(c_double), value :: sigma, kx_max
     end function ekm_wsnet_set_norm

     integer(c_int) function ekm_wsnet_get_norm(net, sigma, kx_max) bind(c, name="ekm_wsnet_get_norm")
       import :: c_ptr, c_int, c_double
       type(c_ptr), value :: net
       real(c_double), intent(out) :: sigma, kx_max
     end function ekm_wsnet_get_norm

     integer(c_int) function ekm_wsnet_prepare(net, rhs, prev, ratio, kx2, nb, x1, scale, stream) &
         bind(c, name="ekm_wsnet_prepare")
       import :: c_ptr, c_int
       type(c_ptr), value :: net, rhs, prev, ratio, kx2, x1, scale, stream
       integer(c_int), value :: nb
     end function ekm_wsnet_prepare

     integer(c_int) function ekm_wsnet_target(net, phi, prev, ratio, scale, nb, t0, stream) &
         bind(c, name="ekm_wsnet_target")
       import :: c_ptr, c_int
       type(c_ptr), value :: net, phi, prev, ratio, scale, t0, stream
       integer(c_int), value :: nb
     end function ekm_wsnet_target

     integer(c_int) function ekm_wsnet_forward(net, x1, nb, y1, stream) bind(c, name="ekm_wsnet_forward")
       import :: c_ptr, c_int
       type(c_ptr), value :: net, x1, y1, stream
       integer(c_int), value :: nb
     end function ekm_wsnet_forward

     integer(c_int) function ekm_wsnet_gradient(net, x1, t0, nb, grad, loss) bind(c, name="ekm_wsnet_gradient")
       import :: c_ptr, c_int, c_float
       type(c_ptr), value :: net, x1, t0
       integer(c_int), value :: nb
       real(c_float), intent(out) :: grad(*), loss
     end function ekm_wsnet_gradient

     integer(c_int) function ekm_wsnet_train_step(net, x1, t0, nb, lr, loss, stream) &
         bind(c, name="ekm_wsnet_train_step")
       import :: c_ptr, c_int, c_float
       type(c_ptr), value :: net, x1, t0, stream
       integer(c_int), value :: nb
       real(c_float), value :: lr
       real(c_float), intent(out) :: loss
     end function ekm_wsnet_train_step

     integer(c_int) function ekm_wsnet_guess(net, rhs, prev, ratio, kx2, nb, out, stream) &
         bind(c, name="ekm_wsnet_guess")
       import :: c_ptr, c_int
       type(c_ptr), value :: net, rhs, prev, ratio, kx2, out, stream
       integer(c_int), value :: nb
     end function ekm_wsnet_guess

     integer(c_int) function ekm_set_warm_start_net(s, net) bind(c, name="ekm_set_warm_start_net")
       import :: c_ptr, c_int
       type(c_ptr), value :: s, net
     end function ekm_set_warm_start_net

     integer(c_int) function ekm_solve_planes(s, rhs, phi, kx2, ratio, nsolve) bind(c, name="ekm_solve_planes")
       import :: c_ptr, c_int
       type(c_ptr), value :: s
       type(c_ptr), intent(out) :: rhs, phi, kx2, ratio
       integer(c_int), intent(out) :: nsolve
     end function ekm_solve_planes

     integer(c_int) function ekm_copy_dev(dst, src, n, stream) bind(c, name="ekm_copy_dev")
       import :: c_ptr, c_int, c_size_t
       type(c_ptr), value :: dst, src, stream
       integer(c_size_t), value :: n
     end function ekm_copy_dev

     integer(c_int) function ekm_fft_x_to_fourier_dev(varp, nx, npencil, stream) &
         bind(c, name="ekm_fft_x_to_fourier_dev")
       import :: c_ptr, c_int
       type(c_ptr), value :: varp, stream
       integer(c_int), value :: nx, npencil
     end function ekm_fft_x_to_fourier_dev

     ! the forward transform without device allocation (NX <= 1024)
     integer(c_int) function ekm_fft_x_to_fourier_dev_noalloc(varp, nx, npencil, stream) &
         bind(c, name="ekm_fft_x_to_fourier_dev_noalloc")
       import :: c_ptr, c_int
       type(c_ptr), value :: varp, stream
       integer(c_int), value :: nx, npencil
     end function ekm_fft_x_to_fourier_dev_noalloc

     integer(c_int) function ekm_fft_x_to_physical_dev(varp, nx, npencil, stream) &
         bind(c, name="ekm_fft_x_to_physical_dev")
       import :: c_ptr, c_int
       type(c_ptr), value :: varp, stream
       integer(c_int), value :: nx, npencil
     end function ekm_fft_x_to_physical_dev

     integer(c_int) function ekm_thomas_real_dev(a, b, c, g, ny, nx, stream) bind(c, name="ekm_thomas_real_dev")
       import :: c_ptr, c_int
       type(c_ptr), value :: a, b, c, g, stream
       integer(c_int), value :: ny, nx
     end function ekm_thomas_real_dev

     ! host-side tables (no device needed)
     integer(c_int) function ekm_host_metrics(nz, ny, xpoint, ypoint, out) bind(c, name="ekm_host_metrics")
       import :: c_int, c_double
       integer(c_int), value :: nz, ny
       real(c_double), intent(in) :: xpoint(*), ypoint(*)
       real(c_double), intent(out) :: out(*)
     end function ekm_host_metrics

     integer(c_int) function ekm_host_les_metrics(nz, ny, xpoint, ypoint, out) bind(c, name="ekm_host_les_metrics")
       import :: c_int, c_double
       integer(c_int), value :: nz, ny
       real(c_double), intent(in) :: xpoint(*), ypoint(*)
       real(c_double), intent(out) :: out(*)
     end function ekm_host_les_metrics

     integer(c_int) function ekm_host_mg_coefficients(p, xpoint, ypoint, lev, kx2, cf, nx, ny) &
         bind(c, name="ekm_host_mg_coefficients")
       import :: c_int, c_double, ekm_params
       type(ekm_params), intent(in) :: p
       real(c_double), intent(in) :: xpoint(*), ypoint(*)
       integer(c_int), value :: lev
       real(c_double), value :: kx2
       real(c_double), intent(out) :: cf(*)
       integer(c_int), intent(out) :: nx, ny
     end function ekm_host_mg_coefficients

     ! timing of the projection launches (HIP events on the solver's stream)
     integer(c_int) function ekm_prof_enable(s, on) bind(c, name="ekm_prof_enable")
       import :: c_ptr, c_int
       type(c_ptr), value :: s
       integer(c_int), value :: on
     end function ekm_prof_enable

     integer(c_int) function ekm_prof_get(s, ms, launches, solves, cycles) bind(c, name="ekm_prof_get")
       import :: c_ptr, c_int, c_double, c_long
       type(c_ptr), value :: s
       real(c_double), intent(out) :: ms
       integer(c_long), intent(out) :: launches, solves, cycles
     end function ekm_prof_get

     ! C strings: the last error message and the version (convert with
     ! c_f_pointer and the terminating NUL)
     type(c_ptr) function ekm_last_error() bind(c, name="ekm_last_error")
       import :: c_ptr
     end function ekm_last_error

     integer(c_int) function ekm_device_name(buf, n) bind(c, name="ekm_device_name")
       import :: c_char, c_int
       character(kind=c_char), intent(out) :: buf(*)
       integer(c_int), value :: n
     end function ekm_device_name

     type(c_ptr) function ekm_version() bind(c, name="ekm_version")
       import :: c_ptr
     end function ekm_version
  end interface

contains

  ! Status check in place of the reference's STOP statements.
  subroutine ekm_check(st, where)
    integer(c_int), intent(in) :: st
    character(len=*), intent(in) :: where
    if (st /= EKM_OK) then
       write(6, '(a,a,a,i0)') 'ekman_bl_amd: ', where, ' failed with status ', st
       if (st == EKM_ERR_DIV_NAN) write(6, *) 'DIV is a NaN'
       if (st == EKM_ERR_DT_SMALL) write(6, *) 'DELTA_T is zero'
       stop 1
    end if
  end subroutine ekm_check

  ! RK_DUCT (src/duct.F90:237): one RKW3+CN substep of the device state.
  subroutine RK_DUCT_GPU()
    call ekm_check(ekm_rk_duct(ekm_handle), 'RK_DUCT')
  end subroutine RK_DUCT_GPU

  ! RK_DUCT as one HIP-graph launch; the projection's checks are read back by
  ! ekm_get_diag / ekm_sync (sticky status)
  subroutine RK_DUCT_GRAPH_GPU()
    call ekm_check(ekm_rk_duct_graph_launch(ekm_handle), 'RK_DUCT (graph)')
  end subroutine RK_DUCT_GRAPH_GPU

  ! REM_DIV_CURV (src/duct.F90:2472): projection of the device state.
  subroutine REM_DIV_CURV_GPU()
    call ekm_check(ekm_rem_div_curv(ekm_handle), 'REM_DIV_CURV')
  end subroutine REM_DIV_CURV_GPU

  ! FFT_X_TO_FOURIER_OP(varp,cvarp,0,NY+1,0,NZP) on an EQUIVALENCEd array:
  ! varp(0:NX+1,0:NZ+1,0:NY+1), transformed in place.
  subroutine FFT_X_TO_FOURIER_OP_GPU(varp, nx, nz, ny)
    integer, intent(in) :: nx, nz, ny
    real(c_double), intent(inout) :: varp(0:nx+1, 0:nz+1, 0:ny+1)
    call ekm_check(ekm_fft_x_to_fourier(varp, int(nx, c_int), int((nz+2)*(ny+2), c_int)), 'FFT_X_TO_FOURIER_OP')
  end subroutine FFT_X_TO_FOURIER_OP_GPU

  subroutine FFT_X_TO_PHYSICAL_OP_GPU(varp, nx, nz, ny)
    integer, intent(in) :: nx, nz, ny
    real(c_double), intent(inout) :: varp(0:nx+1, 0:nz+1, 0:ny+1)
    call ekm_check(ekm_fft_x_to_physical(varp, int(nx, c_int), int((nz+2)*(ny+2), c_int)), 'FFT_X_TO_PHYSICAL_OP')
  end subroutine FFT_X_TO_PHYSICAL_OP_GPU

  ! THOMAS_REAL(A,B,C,G,NY,NX) (src/boundary.F90:53), A..G(0:NX,0:NY)
  subroutine THOMAS_REAL_GPU(A, B, C, G, NY, NX)
    integer, intent(in) :: NY, NX
    real(c_double), intent(inout) :: A(0:NX, 0:NY), B(0:NX, 0:NY), C(0:NX, 0:NY), G(0:NX, 0:NY)
    call ekm_check(ekm_thomas_real(A, B, C, G, int(NY, c_int), int(NX, c_int)), 'THOMAS_REAL')
  end subroutine THOMAS_REAL_GPU

  ! THOMAS_REAL_SP(A,B,C,G,D,K,NY,NX) (src/boundary.F90:4)
  subroutine THOMAS_REAL_SP_GPU(A, B, C, G, D, K, NY, NX)
    integer, intent(in) :: K, NY, NX
    real(c_double), intent(in) :: D
    real(c_double), intent(inout) :: A(0:NX, 0:NY), B(0:NX, 0:NY), C(0:NX, 0:NY), G(0:NX, 0:NY)
    call ekm_check(ekm_thomas_real_sp(A, B, C, G, D, int(K, c_int), int(NY, c_int), int(NX, c_int)), 'THOMAS_REAL_SP')
  end subroutine THOMAS_REAL_SP_GPU

  ! MULTIGRID_MUDPACK(phi) (src/mg_solver2d.F90:75) with its implicit inputs
  ! made explicit: RHS_TMP(0:NZ+1,0:NY+1) and kx2_c; returns iparm(17).
  subroutine MULTIGRID_MUDPACK_GPU(phi, rhs_tmp, kx2, nnx, nny, cycles)
    integer, intent(in) :: nnx, nny
    real(c_double), intent(in) :: rhs_tmp(nnx, nny), kx2
    real(c_double), intent(out) :: phi(nnx, nny)
    integer, intent(out) :: cycles
    real(c_double) :: k2(1)
    integer(c_int) :: cyc(1)
    k2(1) = kx2
    call ekm_check(ekm_multigrid_mudpack(ekm_handle, 1_c_int, k2, rhs_tmp, phi, cyc), 'MULTIGRID_MUDPACK')
    cycles = int(cyc(1))
  end subroutine MULTIGRID_MUDPACK_GPU

end module ekman_bl_amd
