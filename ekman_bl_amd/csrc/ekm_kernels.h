// ekm_kernels.h -- launchers of the substep kernels (ekm_kernels.hip).
#pragma once
#include "ekm_internal.h"

struct RhsArgs {
  const double2* cu[3];
  const double2* cp;
  const double2* cth;
  double2* cr[3];
  double2* cf[3];
  double2* crth;
  double2* cfth;
  const double* kx;
  const double* kx2;
  const double* thbar;
  Ranges rv, rt;
  int rk, th, cont_strat, cor_mean_only, px_on;
  double nu, f0, ug, wg, ri_tau, nupr;
  const StepState* ss;  // TEMP3, TEMP4 and the mean pressure gradient of the substep
};

int launch_rhs_fourier(const Dims& d, const Metrics& M, const RhsArgs& a, hipStream_t st);
int launch_contravariant(const Dims& d, const Metrics& M, const Ranges& r, const double* U2,
                         const double* U3, double* U2b, double* U3b, hipStream_t st);
// One field of the batched ADI stage: its index ranges, the boundary rows of
// its y and z systems, the field, its right-hand side, a scratch field S and
// 7 * NZ2*NY2 doubles of line factors.  With a subgrid viscosity (nut not
// null) the systems differ from point to point along x: gz and gy are the
// metrics of its z- and y-direction terms (GMAT_11 / GMAT_22 for U1, the _y
// variants for U2, the _z variants for U3) and pv is a P field of scratch for
// the pivots; lu is then unused.
struct AdiField {
  Ranges r;
  BcTable by, bz;
  double* U;
  double* R;
  double* S;
  double* lu;
  const double* nut;
  const double* gz;
  const double* gy;
  double* pv;
  const double* wall;  // the wall model's row values, [2][NZ2][NX2] (U1, U3), or null
};
constexpr int kAdiMaxFields = 4;
// Implicit weights of the ADI stage: TEMP1 / div[f] per field (div = 1, or
// Pr for the scalar) and the subgrid-viscosity weight H_BAR/2, from the
// substep's TEMP1 and H_BAR in the step state.
struct AdiT1 {
  const StepState* ss;
  double div[kAdiMaxFields];
};
// The ADI sweeps of nf independent fields in one set of launches; fs: nf
// descriptors in device memory, hf the same on the host (for the grid sizes).
// th_corners applies the scalar corner rows to field th_field afterwards.
int launch_adi_batch(const Dims& d, const Metrics& M, const AdiField* fs, const AdiField* hf, int nf,
                     const AdiT1& t1, int th_field, int th_corners, hipStream_t st, const ekm_tuning& tn);
// LES_WALL_MODEL with WALL_MODEL_TYPE 3 (src/les_chan.F90:1914-2144): the
// row values of the lower type-3 rows from the physical U1, U3 at JSTART
constexpr int kWallMaxNz2 = 2048;  // NZ+2 bound of the wall model's per-K means (LDS)
struct WallArgs {
  const double* u1;
  const double* u3;
  double* wall;          // [2][NZ2][NX2] row values (U1, U3)
  Ranges r;
  double z;              // DYY(2,2) = ypoint(2,2) - ypoint(2,1)
  double nu;
  double cos_a, sin_a;   // COS and SIN of the surface stress angle, as the reference forms them
  int first;             // TIME_STEP = 1
};
int launch_wall_model(const Dims& d, const Metrics& M, const WallArgs& a, hipStream_t st);
int launch_cfl(const Dims& d, const Metrics& M, const Ranges& r, const double* U1, const double* U2,
               const double* U3, double cfl, double dx, unsigned long long* out, hipStream_t st);
int launch_faces(const Dims& d, const Metrics& M, const Ranges& r, const double* CU2, const double* CU3,
                 double* F2, double* F3, hipStream_t st);
int launch_mg_gather(const Dims& d, const Metrics& M, const Ranges& r, const double* kx, const double* CU1,
                     const double* F2, const double* F3, double* rhs, int nmode, hipStream_t st);
int launch_mg_scatter(const Dims& d, const double* phi, int nmode, int fp32, double* PHI, hipStream_t st);
int launch_project(const Dims& d, const Metrics& M, const Ranges& r, const double* kx, const double* PHI,
                   const double* F2, const double* F3, double* CU1, double* CU2, double* CU3, double* DIV,
                   hipStream_t st);
int launch_pupdate(const Dims& d, const Ranges& r, const double* PHI, double* CP, const StepState* ss,
                   hipStream_t st);
int launch_div_reduce(const Dims& d, const double* S, double* out, hipStream_t st);
// x statistics (16 planes, see k_xstats) of the Fourier state; TH may be null
int launch_xstats(const Dims& d, const double* U1, const double* U2, const double* U3, const double* P,
                  const double* TH, double* out, hipStream_t st);
// flow-rate integral of the F_TYPE 0 controller into out[0]
int launch_flow_rate(const Dims& d, const Ranges& r, const double* CU3, const double* dy, double* out,
                     hipStream_t st);
// sponge_th (src/duct.F90:3196-3362) with the streamwise profile off; see
// k_sponge_fluct / k_sponge_mean
struct SpongeArgs {
  double2* cf[3];
  double2* cfth;
  double2* cu[3];
  const double2* cth;
  const double* sigma;  // SPONGE_SIGMA(0:NY+1)
  const double* sigma_out;  // SPONGE_SIGMA_OUT(0:NZ+1), null when the streamwise profile is off
  const StepState* ss;   // Sponge_Amp = 0.1 / DELTA_T of the substep
  Ranges r;             // the velocity ranges JSTART..JEND, ZSTART..ZEND
  int j0;               // first row of the fluctuation damping (max(JSTART, NY_S_T), or
                        // JSTART with the streamwise profile)
  int ny_sponge;        // first row of the mean-U3 average
  double t1, t2, t3;    // mode-0 targets V_BC_YMAX_C1, U_BC_YMAX_C1, WBULK0
  int zmax6;            // type-6 z-max velocity rows with RI_TAU != 0
  int zero_cu2;         // ... and MOD(2*TIME_STEP, SAVE_STATS_INT) = 0
  double jst;           // INT_JACOB(ZEND+1, JEND+1)
};
int launch_sponge(const Dims& d, const Metrics& M, const SpongeArgs& a, hipStream_t st);
// the mean-gradient statistics of SAVE_STATS_CURVI; see k_mean_grads
struct MeanGradArgs {
  const double2 *u1, *u2, *u3, *th;  // the Fourier state (th null without a scalar)
  const double* thbar;               // TH_BAR plane, null with CONT_STRAT
  Ranges rv;                         // the velocity ranges
  double ri_tau;                     // RI_TAU(1)
  double ke_scale;                   // 1, or (AMP_OMEGA0/OMEGA0)**(-2) when AMP_OMEGA0 != 0
  int cont_strat;
  double* out;                       // 8 planes of (NZ+2)*(NY+2), then THBULK
};
int launch_mean_grads(const Dims& d, const Metrics& M, const MeanGradArgs& a, hipStream_t st);
// the higher-order x moments (13 planes, see k_xmoments) from the physical
// velocities U1p..U3p and the Fourier state's means
// mean streamwise vorticity plane (NZ2*NY2) from the Fourier U2, U3
int launch_omega_x(const Dims& d, const Metrics& M, const double* U2, const double* U3, double* out, hipStream_t st);
// elements of a plane of axis 0 (x-y), 1 (z-y) or 2 (x-z); the gather of one from a physical field
size_t plane_elems(const Dims& d, int axis);
int launch_plane_gather(const Dims& d, int axis, int index, const double* U, double* out, hipStream_t st);
int launch_xmoments(const Dims& d, const double* U1p, const double* U2p, const double* U3p, const double* U1,
                    const double* U2, const double* U3, double* out, hipStream_t st);
// THOMAS_REAL on device arrays; ksp >= 0 adds THOMAS_REAL_SP's elimination of
// row ksp+1 by row ksp with factor dd
int launch_thomas_generic(const double* A, const double* B, const double* C, double* G, double* W, int ny,
                          int nx, hipStream_t st, int ksp = -1, double dd = 0.0);

// ---- step-state control kernels (one thread each) ---------------------------
// constants of RK substep rk (1..3) from H_BAR: TEMP1, TEMP3, TEMP4, TEMP5
// (src/duct.F90:283-287), the mean pressure gradient (F_TYPE 1, 2;
// src/duct.F90:431-453), Sponge_Amp and the start of the consistent solve
struct StepConsts {
  double nu;
  double beta[3], zeta[3];   // BETA_BAR, ZETA_BAR (src/diablo.F90:568-576)
  int f_type;
  double px0, amp_omega0, omega0;
  int guess_on;              // start the solve from mg.phi scaled by H_BAR / H_BAR_prev (1) or 1 (2)
};
int launch_step_begin(StepState* ss, int rk, const StepConsts& c, hipStream_t st);
// F_TYPE 0 controller update from the flow rate in red[3] (src/duct.F90:414-436)
int launch_flow_ctrl(StepState* ss, double q_den, int nzz, double wbulk0, hipStream_t st);
// DELTA_T = min(CFL minimum in red[2], DELTA_T_SET) and H_BAR (src/duct.F90:2417-2419)
int launch_dt_update(StepState* ss, hipStream_t st);
// diagnostics of a projection: the reference's DIV and the max from red[0..1]
// (src/duct.F90:2790-2800), cycle totals of the nsolve solves, the status
// checks of src/duct.F90:2811-2812 (sticky)
// unconv (may be null): per-solve flags of a consistent projection that stopped at its cycle limit
int launch_div_final(StepState* ss, const int* cycles, const int* unconv, int nsolve, double denom,
                     hipStream_t st);
// end of a substep: H_BAR_prev, and TIME += DELTA_T after the third (src/diablo.F90:152-157)
int launch_step_end(StepState* ss, int rk, hipStream_t st);
