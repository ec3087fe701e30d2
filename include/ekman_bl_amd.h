/*
 * ekman_bl_amd.h -- C ABI of the MI355X (gfx950) DIABLO duct time step.
 *
 * Every entry point takes plain pointers and sizes; no torch or HIP types
 * appear in the signatures (streams are passed as `void*`, a hipStream_t).
 * A Fortran driver binds these through ISO_C_BINDING (see the interfaces in
 * ekman_bl_amd/fortran/ekman_bl_amd_mod.f90); Python binds them with ctypes.
 *
 * Array layout (the reference's NP=1 layout, src/modules.F90:265-279):
 *   physical field  U(0:NX+1, 0:NZ+1, 0:NY+1)  double, x fastest
 *   Fourier field   CU(0:NX/2, 0:NZ+1, 0:NY+1) complex double (re,im
 *                   interleaved), kx fastest; same bytes as the physical
 *                   pencil, so transforms run in place
 *   element (i,k,j) sits at i + (NX+2)*(k + (NZ+2)*j) (physical) and
 *   m + (NX/2+1)*(k + (NZ+2)*j) complex words (Fourier).
 *
 * Status codes replace the reference's STOP statements.
 */
#ifndef EKMAN_BL_AMD_H
#define EKMAN_BL_AMD_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

#define EKM_OK               0
#define EKM_ERR_ARG         -1  /* bad argument / size mismatch (src/mg_solver2d.F90:138-140) */
#define EKM_ERR_HIP         -2  /* a HIP runtime call failed */
#define EKM_ERR_UNSUPPORTED -3  /* boundary-condition type or option outside the supported set */
#define EKM_ERR_DIV_NAN     -4  /* "DIV is a NaN" (src/duct.F90:2811) */
#define EKM_ERR_DT_SMALL    -5  /* "DELTA_T is zero": DELTA_T < 1e-4 (src/duct.F90:2812) */
#define EKM_ERR_NODEV       -6  /* no HIP device visible */

/* Boundary-condition slots, in the order the reference reads them from
 * input.dat (src/diablo.F90:284-326): lower (y=0), upper (y=top),
 * left (z=0), right (z=end). */
#define EKM_YMIN 0
#define EKM_YMAX 1
#define EKM_ZMIN 2
#define EKM_ZMAX 3

/* Run parameters: the subset of grid_def + input.dat that the RK3 substep
 * reads (src/diablo.F90:242-337; rundir_unst/grid_def:6-44). */
typedef struct ekm_params {
  /* grid_def */
  int nx, ny, nz;          /* NX (Fourier, periodic), NY, NZ interior counts */
  int n_th;                /* N_TH: 0 or 1 scalar */
  int mg_ixp, mg_jyq, mg_iex, mg_jey; /* iixp, jjyq, iiex, jjey: NZ+2 = ixp*2^(iex-1)+1, NY+2 = jyq*2^(jey-1)+1 */
  /* input.dat: parameters and time march */
  double nu, lx, ly, lz;
  double delta_t;          /* DELTA_T as read, also DELTA_T_SET (the CFL cap) */
  int variable_dt;         /* VARIABLE_DT */
  double cfl;              /* CFL */
  int update_dt;           /* UPDATE_DT */
  /* forcing */
  int f_type;              /* F_TYPE: 0 constant flow rate, 1 constant and 2 oscillating
                              pressure gradient on mode 0, 3 none (src/duct.F90:411-455) */
  double wbulk0, px0, omega0, amp_omega0, f0;
  /* velocity BCs: [component 0=U1,1=U2(V),2=U3(W)][slot] type and C1 */
  int u_bc[3][4];
  double u_bc_c1[3][4];
  /* scalar (N_TH = 1) */
  double ri_tau, pr;
  int cont_strat;          /* CONSTANT_STRATIFICATION */
  int th_bc[4];
  double th_bc_c1[4];
  /* pressure solver: MUDPACK_BC(1..4) = nxa, nxb, nyc, nyd */
  int mudpack_bc[4];
  int mg_maxcy;            /* maxcy, 6 in the reference (src/mg_solver2d.F90:158) */
  double mg_tolmax;        /* tolmax, 1e-6 in the reference (src/mg_solver2d.F90:188) */
  /* options that reproduce or remove reference quirks */
  int cmplx_fp32;          /* 1: round phi through default-kind CMPLX as the reference does (src/duct.F90:2664) */
  int coriolis_mean_only;  /* 1: subtract the geostrophic offsets from mode 0 only; 0 = reference (all modes, src/duct.F90:471-473) */
  /* pressure-projection form (no counterpart in the reference):
   * 0 = MUDPACK replica, W(2,1) cycles to relmax <= mg_tolmax (the drop-in);
   * 1 = consistent projection: on the cells the divergence diagnostic covers
   *     (J=2..NY-1, K=2..NZ-1) the Poisson rows are the divergence of the
   *     projection's own face gradient, so J div u after the projection equals
   *     the solve's residual; solved by FP32 alternating zebra-line V(1,1)
   *     cycles inside FP64 iterative refinement until max|residual| <= proj_tol.
   *     Periodic z ends (mudpack_bc 0 0) are supported, periodic y ends are
   *     not (EKM_ERR_UNSUPPORTED); with no specified side the kx = 0 solve
   *     returns the solution with phi = 0 at (K = 0, J = NY+1) that satisfies
   *     every other row (its left null vector is computed once at creation) */
  int proj_mode;
  int proj_maxit;          /* cycle cap of proj_mode 1, 2 */
  double proj_tol;         /* absolute residual tolerance of proj_mode 1, 2 */
  int proj_guess;          /* proj_mode 1, 2: 1 = start each solve from the previous substep's phi
                            * scaled by the ratio of the RK weights h_bar; 2 = from the guess of the
                            * network attached with ekm_set_warm_start_net (0 = from zero) */
  int proj_krylov;         /* proj_mode 1, 2: 1 = the outer iteration is GCR(1) (each cycle's
                            * correction made A-orthogonal to the previous one, then a minimal
                            * residual step); 0 = plain iterative refinement */
  int les_model;           /* LES_MODEL_TYPE of the velocity subgrid model (src/les_chan.F90):
                            * 0 = none (DNS), 2 = dynamic Smagorinsky */
  int sponge_type;         /* SPONGE_TYPE (input.dat, src/diablo.F90:278): 0 = none, 1 = parabolic,
                            * 2 = exponential Rayleigh damping (sponge_th, src/duct.F90:3196-3362,
                            * applied when N_TH = 1) */
  int ny_s_t;              /* NY_S_T (grid_def): first row J of the vertical sponge */
  int nz_s_r;              /* NZ_S_R (grid_def): first K of the right streamwise sponge; the
                            * streamwise profile is on when NZ_S_R < NZ+1 */
  int save_stats_int;      /* SAVE_STATS_INT (input.dat, src/diablo.F90:274); the sponge's
                            * stratified type-6 branch zeroes CU2X(0,:,:) when
                            * MOD(2*TIME_STEP, SAVE_STATS_INT) = 0 */
  int nz_s_l;              /* NZ_S_L (grid_def, NZ+1-NZ_S_R in the shipped file): last K of the
                            * left streamwise sponge */
  int wall_model_type;     /* WALL_MODEL_TYPE (input.dat, src/diablo.F90:255): the model
                            * LES_WALL_MODEL (src/les_chan.F90:1914-2144) evaluates in every
                            * substep when W_BC_YMIN = 3; type 3 (the stress model the
                            * reference attributes to Taylor) is supported */
} ekm_params;

/* Fill with the reference defaults that are not grid-size dependent. */
void ekm_params_default(ekm_params* p);

/* Diagnostics of the last substep. */
typedef struct ekm_diag {
  double div_ref;       /* reference metric: sum |J div u| / ((NX+1)*NY*NZ) over J=2..NY-1, K=2..NZ-1, I=1..NX+1 (src/duct.F90:2790-2800) */
  double div_max;       /* max |J div u| over the same points */
  double delta_t;       /* DELTA_T after the last substep */
  double time;          /* TIME */
  int time_step;        /* TIME_STEP */
  int rk_step;          /* next RK_STEP to run (1..3) */
  int mg_solves;        /* real Poisson solves in the last projection */
  int mg_cycles;        /* W(2,1) cycles summed over those solves (reference iparm(17)) */
  int mg_max_cycles;    /* largest cycle count of a single solve */
  int status;           /* status of the last substep */
  int mg_unconverged;   /* proj_mode 1/2: solves of the last projection that stopped at proj_maxit
                         * with the residual above proj_tol (0 for the MUDPACK replica, which
                         * stops at MAXCY without a flag as the reference's solver does) */
} ekm_diag;

typedef struct ekm_solver ekm_solver;  /* opaque: device-resident duct state */

/* Launch and memory-placement parameters of the solver's kernels (no
 * counterpart in the reference).  None changes the arithmetic: a solver gives
 * bit-identical results for every allowed value; the project's autotuner
 * (ekman_bl_amd/autotune.py) picks the fastest per configuration. */
typedef struct ekm_tuning {
  int adi_rows;       /* rows per batch of the plain ADI sweeps: 4, 8 or 12 (default 8) */
  int adi_block;      /* threads per workgroup of the plain ADI sweeps: 64, 128 or 256 (default 256) */
  int adi_les_rows;   /* rows per batch of the ADI sweeps with a subgrid viscosity: 4, 5, 6 or 8 (default 6) */
  int mg_block;       /* threads per workgroup of the MUDPACK replica: 256, 512 or 1024 (default 512) */
  int tight_lds_kib;  /* LDS cap of the consistent projection's coarse levels in KiB (0: the device's maximum) */
} ekm_tuning;
void ekm_tuning_default(ekm_tuning* t);

/* ---- whole-substep drop-in (replaces RK_DUCT, src/duct.F90:237) ---------- */

/* Create a solver on the current HIP device. xpoint/ypoint are GRID_YZ.dat
 * (z and y of every (K,J), K fastest, (NZ+2)*(NY+2) values; read at
 * src/diablo.F90:1120-1128). th_bar: (NZ+2)*(NY+2) background scalar when
 * cont_strat == 0, else NULL (src/diablo.F90:1152-1160). */
int ekm_create(const ekm_params* p, const double* xpoint, const double* ypoint,
               const double* th_bar, ekm_solver** out);
/* The same with the kernels' tuning (NULL: the defaults). */
int ekm_create_tuned(const ekm_params* p, const double* xpoint, const double* ypoint, const double* th_bar,
                     const ekm_tuning* t, ekm_solver** out);
int ekm_get_tuning(ekm_solver* s, ekm_tuning* t);
int ekm_destroy(ekm_solver* s);

/* Upload / download the Fourier state CU1, CU2, CU3, CP, CTH (host arrays,
 * complex interleaved, (NX/2+1)*(NZ+2)*(NY+2) complex each; cth ignored when
 * N_TH = 0). Replaces READ_FLOW/SAVE_FLOW's record 2 (src/diablo.F90:818-831). */
int ekm_set_state(ekm_solver* s, const double* cu1, const double* cu2, const double* cu3,
                  const double* cp, const double* cth, double time, int time_step);
int ekm_get_state(ekm_solver* s, double* cu1, double* cu2, double* cu3, double* cp, double* cth);

/* The subgrid model's eddy viscosity NU_T ((NX+2)*(NZ+2)*(NY+2) doubles,
 * physical layout) and dynamic coefficient C_DYN ((NZ+2)*(NY+2), K fastest)
 * after the last substep with les_model = 2 (module les_chan_var,
 * src/modules.F90:431-483, set by les_chan, src/les_chan.F90:332-831).
 * set_state counts as a restart (the model runs from the first step, with the
 * coefficient computed in every substep of that step); set_initial_physical
 * as a new flow (the model starts at the second step). */
int ekm_get_les(ekm_solver* s, double* nut, double* cdyn);
/* The scalar subgrid model's eddy diffusivity KAPPA_T ((NX+2)*(NZ+2)*(NY+2),
 * physical layout) and dynamic coefficient C_DYN_TH ((NZ+2)*(NY+2)) after the
 * last substep with les_model = 2 and a scalar (KAPPA_T and C_DYN_TH of
 * module run_variable / les_chan_var, set by les_chan_th,
 * src/les_chan_th.F90:146-334; the coefficient is recomputed every tenth
 * step).  Requires cont_strat = 1 with a scalar. */
int ekm_get_les_th(ekm_solver* s, double* kappa, double* cdyn_th);

/* Physical-space initial fields (U1,U2,U3,TH on (NX+2)*(NZ+2)*(NY+2) doubles):
 * forward x-FFT with 2/3 truncation, CP = 0, then the initial projection
 * (CREATE_FLOW_DUCT, src/duct.F90:3089-3189). */
int ekm_set_initial_physical(ekm_solver* s, const double* u1, const double* u2,
                             const double* u3, const double* th);

/* One RKW3+CN substep (RK_DUCT, src/duct.F90:237-2469); advances RK_STEP and,
 * after the third substep, TIME and TIME_STEP (src/diablo.F90:152-157).
 * The substep is enqueued on the solver's stream without a host round trip:
 * DELTA_T (CFL), the mean-pressure controller and the projection's checks run
 * on the device, and the checks' outcome (EKM_ERR_DIV_NAN, EKM_ERR_DT_SMALL,
 * the reference's STOPs at src/duct.F90:2811-2812) is kept as a sticky status
 * that ekm_get_diag, ekm_sync and ekm_time_step read back.  Returns launch
 * errors, or the sticky status as last read back. */
int ekm_rk_duct(ekm_solver* s);
/* The same substep replayed from a HIP graph: each variant of the substep
 * (RK stage, subgrid-model refresh, CFL update, ...) is captured on first use
 * and then launched as one graph; results are bit-identical to ekm_rk_duct.
 * Same return as ekm_rk_duct. */
int ekm_rk_duct_graph_launch(ekm_solver* s);
/* n full time steps (three substeps each); waits for the device and returns
 * the run's status.  ekm_time_step_ex with graph = 1 replays captured substeps. */
int ekm_time_step(ekm_solver* s, int nsteps);
int ekm_time_step_ex(ekm_solver* s, int nsteps, int graph);
/* Pressure projection alone (REM_DIV_CURV, src/duct.F90:2472-2818): the
 * correction phi is left in the solver for ekm_get_phi. */
int ekm_rem_div_curv(ekm_solver* s);
int ekm_get_phi(ekm_solver* s, double* phi);
int ekm_get_diag(ekm_solver* s, ekm_diag* d);
/* Cycle count of every real solve of the last projection (MUDPACK's iparm(17)
 * per MULTIGRID_MUDPACK call, src/mg_solver2d.F90:195-199), n >= mg_solves
 * entries; solves 2m and 2m+1 belong to kx mode m (imaginary, then real part
 * of the right-hand side, src/duct.F90:2641-2660). */
int ekm_get_solve_cycles(ekm_solver* s, int* cycles, int n);
/* Mean pressure gradient PX0 and last flow rate Qn of the F_TYPE 0 flow-rate
 * controller (PX0, Qn of module run_variable, updated every substep at
 * src/duct.F90:414-436; both start from PX0 and Q0 = WBULK0,
 * src/diablo.F90:359-361); for F_TYPE 1, 2, 3 px0 is the fixed PX0. */
int ekm_get_forcing(ekm_solver* s, double* px0, double* qn);
/* x statistics of the current state, the moments of SAVE_STATS_CURVI
 * (src/flow_statistics.F90:163-215, :420-495), computed on the device from the
 * Fourier fields: 16 planes of (NZ+2)*(NY+2) (K fastest) for every K, J:
 * 0-4 the x means of U1, U2, U3, P, TH; 5-8 urms, vrms, wrms, thrms as the
 * reference forms them (sums over x of the squared fluctuations, not divided
 * by NX); 9-15 uv, uw, wv, pu (= w'p'), pv, thv, thw (sums over x divided by
 * NX).  The sums run over the NX physical points (the reference's loops also
 * visit the two padding entries of the physical array, whose contents its FFT
 * leaves undefined). */
int ekm_get_xstats(ekm_solver* s, double* out);
/* The mean-gradient statistics of SAVE_STATS_CURVI from the x means of the
 * current state (src/flow_statistics.F90:164-175, 228-275, 455-553): 8 planes
 * of (NZ+2)*(NY+2) doubles (K fastest) -- 0 dudy, 1 dwdy, 2 dudz, 3 dwdz,
 * 4 dthdy (with the background gradient when cont_strat = 0), 5 dthdz, 6 the
 * gradient Richardson number Rig, 7 the mean-flow kinetic energy TKE -- then
 * THBULK (1 double).  The reference's row/column fills are applied; points it
 * never writes are zero; the scalar planes are zero when N_TH = 0. */
int ekm_get_mean_grads(ekm_solver* s, double* out);
/* The higher-order x moments of high_oder_momentums_curvi
 * (src/flow_statistics.F90:1026-1095): 13 planes of (NZ+2)*(NY+2) doubles
 * (K fastest) -- UU, VV, WW, UUU, VVV, WWW, UUUU, VVVV, WWWW, UUV, WWV, UUUV,
 * WWWV -- the x averages over the NX physical points of powers of the velocity
 * fluctuations about the mode-0 means. */
int ekm_get_xmoments(ekm_solver* s, double* out);
/* The mean streamwise vorticity of SAVE_STATS_CURVI (src/flow_statistics.F90:
 * 301-327), omega_x = dW/dy - dV/dz of the x means: one plane of (NZ+2)*(NY+2)
 * doubles (K fastest); rows 0 and NY+1 copy rows 1 and NY, columns 0 and NZ+1
 * are zero. */
int ekm_get_omega_x(ekm_solver* s, double* out);
/* One plane of the physical U1, U2, U3, P and TH (five planes in that order,
 * TH zero when N_TH = 0), the data the reference's plane writers take from
 * U1X..THX (src/flow_statistics.F90:2004-2629):
 *   axis 0: x-y plane at K = index (0..NZ+1), each (NY+2) x NX, I fastest
 *   axis 1: z-y plane at I = index (0..NX-1), each (NY+2) x (NZ+2), K fastest
 *   axis 2: x-z plane at J = index (0..NY+1), each (NZ+2) x NX, I fastest */
int ekm_get_plane(ekm_solver* s, int axis, int index, double* out);
/* Field f (0..4: U1, U2, U3, P, TH) of the current state in physical space,
 * the physical layout above ((NX+2)*(NZ+2)*(NY+2) doubles; the two padding
 * entries of each x pencil are undefined); zeros for TH when N_TH = 0.  The
 * data of the reference's full 3D writer (plot_3D, src/flow_statistics.F90:
 * 2004-2040). */
int ekm_get_physical(ekm_solver* s, int f, double* out);
/* Blocks until queued work on the solver's stream is done; returns the run's
 * sticky status. */
int ekm_sync(ekm_solver* s);
/* Device pointer of state field f (0=U1,1=U2,2=U3,3=P,4=TH); the field's
 * representation is Fourier between substeps. */
int ekm_field_device_ptr(ekm_solver* s, int f, void** dptr);

/* ---- component drop-ins (host pointers: copy in, compute on the GPU, copy out) */

/* FFT_X_TO_FOURIER_OP(varp,cvarp,0,NY+1,0,NZP) (call sites src/mpi_duct.F90:290;
 * routine in the missing fft.F90): in-place batched real->complex transform
 * along x of `npencil` pencils of NX reals laid out with stride NX+2, scaled
 * by 1/NX, modes NX/3+1..NX/2 zeroed (src/mpi_duct.F90:291-294). */
int ekm_fft_x_to_fourier(double* varp, int nx, int npencil);
/* FFT_X_TO_PHYSICAL_OP (call sites src/mpi_duct.F90:308): modes 0..NX/3 are
 * used, the rest ignored; padding entries NX, NX+1 set to 0. */
int ekm_fft_x_to_physical(double* varp, int nx, int npencil);
/* THOMAS_REAL(A,B,C,G,NY,NX) (src/boundary.F90:53-84): NX+1 systems of NY+1
 * unknowns, batch index fastest, A(0:NX,0:NY); solution returned in G; A and B
 * are left unchanged (the reference clobbers them). */
int ekm_thomas_real(double* a, double* b, double* c, double* g, int ny, int nx);
/* THOMAS_REAL_SP(A,B,C,G,D,K,NY,NX) (src/boundary.F90:4-47). */
int ekm_thomas_real_sp(double* a, double* b, double* c, double* g, double d, int k, int ny, int nx);
/* MULTIGRID_MUDPACK for `nsolve` right-hand sides at once (src/mg_solver2d.F90:75-201):
 * rhs_tmp[s] and phi[s] are (NZ+2)*(NY+2) planes (K fastest) of RHS_TMP and
 * phi, kx2[s] the squared wavenumber each solve uses in cofcr
 * (src/duct.F90:3900-3934). cycles[s] receives iparm(17). The metrics and BC
 * flags are the solver's. */
int ekm_multigrid_mudpack(ekm_solver* s, int nsolve, const double* kx2,
                          const double* rhs_tmp, double* phi, int* cycles);
/* The same with MUDPACK's iguess (iparm(4), src/mud2cr.f:93,238): 1 = phi
 * holds the initial guesses on entry, 0 = start from zero as
 * MULTIGRID_MUDPACK does (src/mg_solver2d.F90:157,190). */
int ekm_multigrid_mudpack_ex(ekm_solver* s, int nsolve, const double* kx2, const double* rhs_tmp, double* phi,
                             int* cycles, int iguess);

/* ---- device-pointer variants (stream = hipStream_t or NULL) -------------- */
int ekm_fft_x_to_fourier_dev(double* varp, int nx, int npencil, void* stream);
int ekm_fft_x_to_physical_dev(double* varp, int nx, int npencil, void* stream);
/* ekm_fft_x_to_fourier_dev without any device allocation.  No counterpart in
 * the reference: for callers that cannot allocate device memory (code under
 * HIP graph capture, tight serving loops, hosts that forbid hipMalloc).  The
 * plan's twiddle tables live in a file-scope device array of 1538 doubles
 * (12 KB on each device, held until the process ends), so NX <= 1024; a
 * larger NX returns EKM_ERR_ARG with a message (ekm_last_error).  Every call
 * on a device shares that array: calls that overlap on one device race.
 * Each call rewrites the array and waits for its transform to end. */
int ekm_fft_x_to_fourier_dev_noalloc(double* varp, int nx, int npencil, void* stream);
int ekm_thomas_real_dev(const double* a, const double* b, const double* c, double* g,
                        int ny, int nx, void* stream);
/* THOMAS_REAL_SP(A,B,C,G,D,K,NY,NX) (src/boundary.F90:4-47) on device arrays;
 * the elimination of row K+1 by row K is applied on the fly (A, B, C are not
 * modified). */
int ekm_thomas_real_sp_dev(const double* a, const double* b, const double* c, double* g, double d, int k,
                           int ny, int nx, void* stream);

/* ---- learned initial guess of the pressure solves ------------------------
 * No counterpart in the reference, whose MUDPACK solves always start from
 * zero (iguess = 0, phi = 0: src/mg_solver2d.F90:157,190; the iguess branch
 * of src/mud2cr.f:93,238).  A convolutional encoder-decoder (five levels,
 * about 230k parameters, bf16 matrix-core products with fp32 accumulation)
 * maps each real solve's right-hand side RHS_TMP, the previous projection's
 * solution of the same solve and kx to a start; the consistent projection
 * (proj_mode 1, 2 with proj_guess = 2) then iterates to the same tolerance.
 * Planes are (NZ+2)*(NY+2) doubles, K fastest; NZ+2-1 and NY+2-1 must be
 * multiples of 32.  Bulk arguments are device pointers (stream: hipStream_t
 * or NULL); parameter vectors are host float arrays in the layer order of
 * the network (per layer W[cout][cin][taps], then b[cout]). */
typedef struct ekm_wsnet ekm_wsnet;
/* bmax: the largest number of planes per call; trainable: also allocate the
 * gradient, Adam and loss buffers.  Weights start He-initialised (seed 0). */
int ekm_wsnet_create(int nz2, int ny2, int bmax, int trainable, ekm_wsnet** out);
int ekm_wsnet_destroy(ekm_wsnet* n);
long ekm_wsnet_nparams(const ekm_wsnet* n);
int ekm_wsnet_set_params(ekm_wsnet* n, const float* host);
int ekm_wsnet_get_params(ekm_wsnet* n, float* host);
/* input normalisation: the scaled previous solution is divided by s*sigma
 * (s = max|rhs| per plane), kx by kx_max.  The network predicts what it adds
 * to ratio * prev, the previous substep's solution scaled by the ratio of the
 * two substeps' H_BAR (the start proj_guess 1 uses). */
int ekm_wsnet_set_norm(ekm_wsnet* n, double sigma, double kx_max);
int ekm_wsnet_get_norm(const ekm_wsnet* n, double* sigma, double* kx_max);
/* network input x1 (nb x (NY+3)/2 x (NZ+3)/2 x 4 floats) and the scales s
 * (nb doubles) of nb planes; prev may be NULL (zero); ratio: device pointer
 * to one double scaling prev, or NULL (1) */
int ekm_wsnet_prepare(ekm_wsnet* n, const double* rhs, const double* prev, const double* ratio, const double* kx2,
                      int nb, float* x1, double* scale, void* stream);
/* training target (phi - ratio * prev) / (s sigma) at the plane's resolution
 * (nb planes of floats) */
int ekm_wsnet_target(ekm_wsnet* n, const double* phi, const double* prev, const double* ratio, const double* scale,
                     int nb, float* t0, void* stream);
/* prediction at level 1 (nb x (NY+3)/2 x (NZ+3)/2 floats; y1 may be NULL) */
int ekm_wsnet_forward(ekm_wsnet* n, const float* x1, int nb, float* y1, void* stream);
/* gradient of the mean relative loss  sum (up(y) - t)^2 / sum t^2  with
 * respect to the parameters (host floats), and the loss; no update */
int ekm_wsnet_gradient(ekm_wsnet* n, const float* x1, const float* t0, int nb, float* grad_host, float* loss);
/* one Adam step (beta 0.9, 0.999) on that loss; returns the batch loss */
int ekm_wsnet_train_step(ekm_wsnet* n, const float* x1, const float* t0, int nb, float lr, float* loss, void* stream);
/* the guess of nb solves, ratio * prev + s sigma up(y) as double planes; out
 * may alias prev */
int ekm_wsnet_guess(ekm_wsnet* n, const double* rhs, const double* prev, const double* ratio, const double* kx2,
                    int nb, double* out, void* stream);
/* Attach a network to a solver (NULL detaches): with proj_mode 1 or 2 and
 * proj_guess = 2, each real solve after the first projection starts from the
 * network's guess.  The network must match the plane and hold bmax >= the
 * solver's 2 (NX/3+1) solves; the solver does not own it. */
int ekm_set_warm_start_net(ekm_solver* s, ekm_wsnet* net);
/* Device pointers of the last projection's right-hand sides (RHS_TMP) and
 * solutions (nsolve planes each, (NZ+2)*(NY+2) doubles, K fastest; solve 2m
 * and 2m+1 belong to kx mode m), of every solve's kx^2 and of the H_BAR ratio
 * of that projection's substep to the one before (one double): what the
 * warm-start training data is dumped from. */
int ekm_solve_planes(ekm_solver* s, const double** rhs, const double** phi, const double** kx2,
                     const double** ratio, int* nsolve);

/* ---- measurement and utility ------------------------------------------- */
/* Time every multigrid launch of the solver with HIP events on its stream
 * (on = 1 resets the counters). */
int ekm_prof_enable(ekm_solver* s, int on);
int ekm_prof_get(ekm_solver* s, double* ms, long* launches, long* solves, long* cycles);
/* Plain device copy of n doubles (calibration kernel for HBM byte counters). */
int ekm_copy_dev(double* dst, const double* src, size_t n, void* stream);
int ekm_dev_alloc(void** p, size_t bytes);
int ekm_dev_free(void* p);
/* dst row r = src row idx[r] for r < n (device rows of row_bytes; idx is a
 * host array); waits for the copies */
int ekm_gather_rows(void* dst, const void* src, const long* idx, int n, size_t row_bytes, void* stream);
/* host <-> device copies (synchronous) */
int ekm_memcpy_h2d(void* dst, const void* src, size_t bytes);
int ekm_memcpy_d2h(void* dst, const void* src, size_t bytes);
/* Host-side tables, no device needed: metrics (15*(NZ+3)*(NY+3): GMAT_11,12,22,
 * CJOB_11,12,21,22 with two faces each, then INT_JACOB) and the multigrid
 * coefficients cf(i,j,1..9) of level `lev` (1 = coarsest). */
int ekm_host_metrics(int nz, int ny, const double* xpoint, const double* ypoint, double* out);
/* The subgrid-stress metric variants of JACOBI_TRANS (src/diablo.F90:1379-1474):
 * 12*(NZ+3)*(NY+3) = GMAT_11_z,12_z,22_z,11_y,12_y,22_y with two faces each. */
int ekm_host_les_metrics(int nz, int ny, const double* xpoint, const double* ypoint, double* out);
/* SPONGE_SIGMA(J), J = 0..NY+1, of the vertical sponge from ypoint(NZ+1, J)
 * (replaces the profile part of JACOBI_TRANS, src/diablo.F90:1533-1556;
 * type 1 parabolic, 2 exponential, 0 zero). */
int ekm_host_sponge_sigma(int type, int nz, int ny, int ny_s_t, const double* ypoint, double* sig);
/* SPONGE_SIGMA_OUT(K), K = 0..NZ+1, of the streamwise sponge from
 * xpoint(K, NY+1) (src/diablo.F90:1489-1523; zero unless NZ_S_R < NZ+1). */
int ekm_host_sponge_sigma_out(int type, int nz, int ny, int nz_s_r, int nz_s_l, const double* xpoint,
                              double* sig);
int ekm_host_mg_coefficients(const ekm_params* p, const double* xpoint, const double* ypoint,
                             int lev, double kx2, double* cf, int* nx, int* ny);
/* Name and architecture of the current HIP device (NUL-terminated, at most n bytes). */
int ekm_device_name(char* buf, int n);
/* Message of the last HIP error seen by this thread. */
const char* ekm_last_error(void);

/* Library version string. */
const char* ekm_version(void);

#ifdef __cplusplus
}
#endif
#endif /* EKMAN_BL_AMD_H */
