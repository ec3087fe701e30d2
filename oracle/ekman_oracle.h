/*
 * ekman_oracle.h -- CPU FP64 oracle of the DIABLO duct RK3 substep.
 *
 * TEST INFRASTRUCTURE ONLY.  This is a plain-C restatement of the reference
 * algorithm (jortizta/ekman_bl: src/duct.F90, src/boundary.F90,
 * src/mg_solver2d.F90, src/mud2cr.f, src/mudcom.f, src/diablo.F90) against
 * which the HIP path is tested.  Only tests/, __graft_entry__.smoke() and bench.py's
 * cpu_baseline leg may load it.  The product never links it.
 *
 * Pinning (what checks each function; "parity unpinned" marks a function no
 * independent check covers; tests/test_oracle_cpu.py unless noted):
 *   ora_fft_fwd / ora_fft_inv   numpy rfft/irfft with 1/NX scaling and the 2/3
 *                               rule; spectral derivative of sin(m x)
 *   ora_thomas_real(_sp)        dense numpy.linalg.solve (with the extra
 *                               (K+1, K-1) entry of the _SP form)
 *   metrics (ora_get_metric)    closed forms on uniform and stretched
 *                               rectilinear grids, the 1e-8 truncation; the
 *                               shipped rundir_unst/GRID_YZ.dat (fixture) is the
 *                               tanh family the grids come from
 *   ora_mudpack                 mud2cr itself (src/mud2cr.f + src/mudcom.f built
 *                               into oracle/_ref by oracle/Makefile), bit for
 *                               bit with equal cycle counts; convergence to the
 *                               sparse direct solution of the assembled 9-point
 *                               system; exactness on quadratics
 *   ora_tight_solve / _operator the operator rows against the divergence of the
 *                               projection's face gradient computed directly;
 *                               the solve against a sparse direct solve
 *   ora_rem_div_curv            uniform grid: J div u after a tight solve is the
 *                               multigrid residual; proj_mode 1 removes the
 *                               divergence to the solve tolerance
 *   LES metric variants         rectilinear grids: GMAT_11_z = 2 GMAT_11,
 *   (ora_get_metric 8..13)      GMAT_11_y = GMAT_11, GMAT_22_z = GMAT_22,
 *                               GMAT_22_y = 2 GMAT_22, GMAT_12_* = 0; skewed
 *                               grids: X_z + X_y = 3 X for X = 11, 12, 22
 *   les_chan (LES model 2)      x-independent flow: C_DYN = 0 and the substep
 *                               is the plain substep; C_DYN >= 0 after clipping,
 *                               NU_T >= 0 and zero outside the interior
 *   ora_adi_field with NU_T     a constant NU_T on U1 (whose subgrid metrics are
 *                               the molecular ones) is the plain ADI with
 *                               t1 + t2 NU_T; the U2/U3 metric selection and the
 *                               explicit cross terms (les_xterm): parity unpinned
 *                               beyond the reference text (the GPU path is
 *                               checked against them)
 *   ora_xmoments                one x mode about non-zero means: the cos^n
 *                               averages in closed form
 *   ora_omega_x                 sheared affine grid: linear U2, U3 give
 *                               dW/dy - dV/dz exactly, and the row fills
 *   ora_get_physical            the inverse transform pinned above
 *   ora_wall_model              uniform velocities at JSTART: U_1, W_1 in
 *                               closed form, U_STAR a fixed point of the log
 *                               law, the row values in closed form; the
 *                               type-3 rows: U(1) - U(0) equals the model's
 *                               value after the implicit solve
 *   ora_mean_grads              sheared affine grid: linear mean fields give
 *                               their exact derivatives, Rig, TKE, the fills
 *                               and THBULK in closed form
 *   sponge_sigma / sponge_th    SPONGE_SIGMA and SPONGE_SIGMA_OUT closed forms;
 *                               SPONGE_SIGMA: (0 at NY_S_T, 1 at
 *                               NY+1, type 1 on the K = NZ+1 column, type 2 a
 *                               constant growth factor); sigma = 0 and a state
 *                               at the targets leave the substep bit-identical
 *                               to the substep without a sponge; the damping
 *                               lowers the top rows' fluctuation energy.  The
 *                               stratified type-6 branch (whole-plane CF2
 *                               relaxation, CU2X(0,:,:) zeroing): parity
 *                               unpinned beyond the reference text
 *   ora_rk_duct / ora_time_step tests/test_oracle_timestep.py: one step of a
 *                               uniform inviscid flow is the RK3 stability
 *                               polynomial of i f dt (RK weights, Coriolis);
 *                               Crank-Nicolson decay of discrete Dirichlet
 *                               modes; the laminar Ekman spiral steady to
 *                               second order; the CFL closed form; second-order
 *                               scalar transport on a sheared grid; cofcr at
 *                               second order; lagged periodic rows.  The ghost
 *                               and corner updates and the dt-update timing:
 *                               parity unpinned beyond the reference text (the
 *                               GPU path is checked against them)
 *   boundary rows (ora_bc_rows) the reference's own APPLY_BC_* (src/boundary.F90)
 *                               compiled unchanged into oracle/_ref with the
 *                               harness oracle/ref_bc_shim.f90: every supported
 *                               type on every side, bit for bit
 *                               (tests/test_ref_boundary_cpu.py)
 */
#ifndef EKMAN_ORACLE_H
#define EKMAN_ORACLE_H
#include "../include/ekman_bl_amd.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct ora ora;

ora* ora_create(const ekm_params* p, const double* xpoint, const double* ypoint,
                const double* th_bar);
void ora_destroy(ora* o);
int ora_set_state(ora* o, const double* cu1, const double* cu2, const double* cu3,
                  const double* cp, const double* cth, double time, int time_step);
int ora_get_state(ora* o, double* cu1, double* cu2, double* cu3, double* cp, double* cth);
int ora_set_initial_physical(ora* o, const double* u1, const double* u2,
                             const double* u3, const double* th);
int ora_rk_duct(ora* o);
int ora_get_les(ora* o, double* nut, double* cdyn);
/* the scalar model: KAPPA_T (physical P field) and C_DYN_TH ((NZ+2)*(NY+2)) */
int ora_get_les_th(ora* o, double* kappa, double* cdyn_th);
/* the F_TYPE 0 controller: current PX0 and the last flow rate Qn */
int ora_get_forcing(ora* o, double* px0, double* qn);
/* SPONGE_SIGMA(0:NY+1) of the vertical sponge (zero without one) */
int ora_get_sponge(ora* o, double* sigma);
/* SPONGE_SIGMA_OUT(0:NZ+1) of the streamwise sponge (zero without one) */
int ora_get_sponge_out(ora* o, double* sigma_out);
/* x statistics (16 planes of (NZ+2)*(NY+2)), direct sums in physical space */
int ora_xstats(ora* o, double* out);
/* the mean-gradient statistics of SAVE_STATS_CURVI: 8 planes of
 * (NZ+2)*(NY+2) (dudy, dwdy, dudz, dwdz, dthdy, dthdz, Rig, TKE), then THBULK */
int ora_mean_grads(ora* o, double* out);
/* the higher-order x moments of the velocity fluctuations (13 planes of
 * (NZ+2)*(NY+2): UU VV WW UUU VVV WWW UUUU VVVV WWWW UUV WWV UUUV WWWV) */
int ora_xmoments(ora* o, double* out);
/* omega_x of SAVE_STATS_CURVI: one plane of (NZ+2)*(NY+2) */
int ora_omega_x(ora* o, double* out);
/* field f (0..4: U1, U2, U3, P, TH) in physical space, (NX+2)*(NZ+2)*(NY+2) */
int ora_get_physical(ora* o, int f, double* out);
/* LES_WALL_MODEL (type 3) on given physical U1, U3: the lower-row values
 * [2][NZ+2][NX+2] (U1's, U3's) into out and U_1, W_1, U_STAR, TAO_WALL,
 * TAO_WALL_1 into info[5]; EKM_ERR_ARG without type-3 rows */
int ora_wall_model(ora* o, const double* u1, const double* u3, int first, double* out, double* info);
/* the boundary rows one ADI line gets (dir 0: y line at K = line, 1: z line
 * at J = line) from the physical field u, into caller-filled a, b, c, g of
 * (rows x (NX+2)) doubles; only the boundary rows are written */
int ora_bc_rows(ora* o, int comp, int dir, int line, const double* u, double* a, double* b, double* c, double* g);
/* one field's ADI (comp 0..2 velocity, 3 scalar) on physical U, R in place,
 * with the subgrid viscosity nut (physical P field) unless NULL */
int ora_adi_field(ora* o, int comp, double* U, double* R, double t1, const double* nut, double t2);
int ora_time_step(ora* o, int nsteps);
int ora_rem_div_curv(ora* o);
int ora_get_phi(ora* o, double* phi);
int ora_get_diag(ora* o, ekm_diag* d);
/* metrics: which = 0..6 -> GMAT_11,GMAT_12,GMAT_22,CJOB_11,CJOB_12,CJOB_21,CJOB_22
 * ((NZ+3)*(NY+3)*2, K fastest then J then face), 7 -> INT_JACOB ((NZ+3)*(NY+3)),
 * 8..13 -> GMAT_11_z,GMAT_12_z,GMAT_22_z,GMAT_11_y,GMAT_12_y,GMAT_22_y (two faces). */
int ora_get_metric(ora* o, int which, double* out);

/* components */
void ora_fft_fwd(double* varp, int nx, long npencil);
void ora_fft_inv(double* varp, int nx, long npencil);
void ora_thomas_real(double* a, double* b, double* c, double* g, int ny, int nx);
void ora_thomas_real_sp(double* a, double* b, double* c, double* g, double d, int k,
                        int ny, int nx);
/* MULTIGRID_MUDPACK for one right-hand side: rhs_tmp, phi are (NZ+2)*(NY+2)
 * planes, K fastest.  Returns the cycle count (iparm(17)). */
int ora_mudpack(ora* o, double kx2, const double* rhs_tmp, double* phi);
/* Assembled level-`lev` coefficients cf(i,j,1..10) as MUDPACK stores them
 * (lev = 1 coarsest .. ngrid finest), for kx2; returns nx*ny or <0. */
int ora_mg_discretize(ora* o, double kx2, int lev, double* cf, int* nx, int* ny);
int ora_mg_ngrid(ora* o);
/* proj_mode 1: one real solve of the consistent system (returns V cycles),
 * and its fine-level operator cf[9][NY+2][NZ+2] for kx2. */
int ora_tight_solve(ora* o, double kx2, const double* rhs_tmp, double* phi);
int ora_tight_operator(ora* o, double kx2, double* cf);

/* Callbacks with MUDPACK's calling convention (all arguments by reference),
 * bound to the oracle's metrics for the reference build in oracle/_ref. */
void ora_cb_bind(ora* o, double kx2);
void ora_cofcr_cb(const double* x, const double* y, double* cxx, double* cxy, double* cyy,
                  double* cx, double* cy, double* ce);
void ora_bndcr_cb(const int* kbdy, const double* xory, double* alfa, double* beta,
                  double* gama, double* gbdy);

#ifdef __cplusplus
}
#endif
#endif
