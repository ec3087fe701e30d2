// ekm_internal.h -- shared types of the gfx950 duct-step library.
//
// Device layout (the reference's NP=1 layout, src/modules.F90:265-279):
//   physical field U(0:NX+1, 0:NZ+1, 0:NY+1) double, x fastest;
//   Fourier field CU(0:NX/2, 0:NZ+1, 0:NY+1) double2, kx fastest, in place.
//   metrics (0:NZ+2, 0:NY+2, face 1:2), K fastest (src/ALLOCATION.F90:166-170).
#pragma once
#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>
#include <vector>

#include "../../include/ekman_bl_amd.h"

#define EKM_CHECK(call)                                  \
  do {                                                   \
    hipError_t e__ = (call);                             \
    if (e__ != hipSuccess) {                             \
      ekm_note_hip_error(e__, #call, __FILE__, __LINE__); \
      return EKM_ERR_HIP;                                \
    }                                                    \
  } while (0)

void ekm_note_hip_error(hipError_t e, const char* what, const char* file, int line);
// records a message for ekm_last_error (and prints it)
void ekm_note_error(const char* msg);

struct Dims {
  int NX, NY, NZ;       // grid_def sizes
  int NXC, NKX;         // NX/2+1 complex modes, NX/3 kept by the 2/3 rule
  int NX2, NY2, NZ2;    // padded extents
  int NZ3, NY3;         // metric extents
  size_t npts;          // doubles per 3D field
  size_t plane;         // NZ2*NY2
  size_t mplane;        // NZ3*NY3
};

// Device metrics, each (NZ3*NY3*2) except jac (NZ3*NY3).
struct Metrics {
  const double *g11, *g12, *g22, *c11, *c12, *c21, *c22, *jac;
};

__host__ __device__ inline size_t mf(const Dims& d, int k, int j, int f) {
  return (size_t)k + (size_t)d.NZ3 * ((size_t)j + (size_t)d.NY3 * (size_t)(f - 1));
}
__host__ __device__ inline size_t mi(const Dims& d, int k, int j) {
  return (size_t)k + (size_t)d.NZ3 * (size_t)j;
}

// Index ranges of INIT_DUCT (src/duct.F90:7-232).
struct Ranges {
  int js, je, zs, ze;
};

// Run state of a solver that changes from substep to substep, kept in device
// memory and updated by one-thread control kernels (ekm_kernels.hip), so that
// a substep needs no host round trip and can be replayed from a HIP graph:
// DELTA_T and H_BAR after the CFL update (src/duct.F90:2417-2419), TIME, the
// TEMP1..5 of the substep in flight (src/duct.F90:283-287), the mean pressure
// gradient, the F_TYPE 0 controller state (src/duct.F90:414-436), the
// reduction accumulators and the diagnostics of the last projection with a
// sticky status word (the reference's STOPs, src/duct.F90:2811-2812).
// Every kernel that reads one of these takes a pointer to this block.
struct StepState {
  double delta_t, delta_t_set, time;
  double h_bar[3];
  double hb, temp1, temp3, temp4, temp5;  // of the substep in flight
  double px;         // PX0 (+ AMP_OMEGA0 cos(OMEGA0 TIME)) of the substep in flight
  double amp;        // Sponge_Amp = 0.1 / DELTA_T
  double guess;      // start of the consistent solve: 0 (from zero) or the scale of mg.phi
  double hb_prev;    // H_BAR of the previous substep
  double prev_ratio; // H_BAR of the substep in flight over hb_prev (0 before any substep)
  double px0, qn;    // F_TYPE 0 controller
  // reduction accumulators; each is reset by the kernel that consumes it:
  // 0 sum |J div u|, 1 max |J div u| (bits), 2 CFL minimum (bits), 3 flow rate
  double red[4];
  double div_ref, div_max;
  int mg_solves, mg_cycles, mg_max_cycles, status;
  int mg_unconverged;  // solves of the last projection that stopped at the cycle limit
  long long prof_launches, prof_solves, prof_cycles;  // totals since the last reset
};

// One boundary row of an ADI system, the net effect of the APPLY_BC_*
// routines of src/boundary.F90 (IC_TYPE = 7) after all their assignments.
enum BcKind : int {
  BC_CONST = 0,    // rhs = val
  BC_LAGGED = 1,   // rhs = field on the same line at index src (type 6)
  BC_THFLUX_LO = 2,  // rhs = val / (0.5*(C22(K,1,2)+C22(K+1,1,1))) * J(K,1)
  BC_THFLUX_NY = 3,  // rhs = 0.5*val*(J(K,NY)+J(K,NY-1)) / C22(K,NY-1,1)
  BC_THFLUX_NY1 = 4, // rhs = 0.5*val*(J(K,NY+1)+J(K,NY)) / C22(K,NY,1)
  BC_WALL = 5,       // rhs = the wall model's value at (i,K): array src (0 U1, 1 U3)
};
struct BcRow {
  int row;  // -1 = unused
  int kind;
  int src;
  int pad;
  double a, b, c, val;
};
struct BcTable {
  BcRow r[4];
  int sp;          // 1: THOMAS_REAL_SP elimination of row NZ+1 by row NZ (type 5)
  int sp_metric;   // D = GMAT_11(NZ,J,2) (U1) instead of 1.0
  int pad[2];
};

// ---- FFT ------------------------------------------------------------------
struct FftPlan {
  int nx = 0;
  double* tw = nullptr;   // exp(-2 pi i k / (NX/2)), k < NX/4 (re,im)
  double* tw2 = nullptr;  // exp(-2 pi i k / NX), k <= NX/2 (re,im)
};
int fft_plan_init(FftPlan& p, int nx);
// the plan's twiddle tables on the host: fft_tw_doubles(nx) values of tw,
// then fft_tw2_doubles(nx) of tw2
void fft_twiddles(int nx, double* tw, double* tw2);
size_t fft_tw_doubles(int nx);
size_t fft_tw2_doubles(int nx);
void fft_plan_free(FftPlan& p);
// in-place transforms of npencil pencils of stride NX+2
int fft_forward(const FftPlan& p, double* v, long npencil, hipStream_t st);
int fft_inverse(const FftPlan& p, double* v, long npencil, hipStream_t st);
// fused flux -> forward transform -> CF/CR accumulation of one transported
// field; nut (with its cross-term metric gl) adds the subgrid viscosity's
// cross terms
int launch_flux_accum(const FftPlan& p, const Dims& d, const Metrics& M, const Ranges& r, const double* U1,
                      const double* Q, const double* U2b, const double* U3b, double nu, const double* kx, double* CF,
                      double* CR, const double* temp5, hipStream_t st, const double* nut = nullptr,
                      const double* gl = nullptr);

// ---- multigrid --------------------------------------------------------------
struct HostMetrics;
struct MgLevel {
  int nx = 0, ny = 0;
  int in_lds = 0;                   // level lives in LDS inside the solve kernel
};
struct MgPlan {
  int ngrid = 0;
  int nxa = 0, nxb = 0, nyc = 0, nyd = 0;
  int maxcy = 0;
  double tolmax = 0.0;
  int nsolve_max = 0;
  MgLevel lev[17];
  size_t ws_global = 0;      // doubles per solve in global memory
  size_t ws_lds = 0;         // doubles per solve in LDS
  int lc = 0;                // finest level held in LDS
  int nops_fmg = 0, nops_w = 0;
  double* ws = nullptr;      // global workspace [nsolve_max][ws_global]
  double* coefmem = nullptr; // all level coefficient arrays
  int* cycles = nullptr;     // [2][nsolve_max]: cycle counts, then 1 where a consistent solve stopped unconverged
  void* desc = nullptr;      // device copy of the level descriptor, phase tables and cycle programs
  double* kx2 = nullptr;     // [nsolve_max]
  double* rhs = nullptr;     // [nsolve_max][plane] RHS_TMP
  double* phi = nullptr;     // [nsolve_max][plane]
  int block = 512;
};
int mg_plan_build(MgPlan& mg, const ekm_params& p, const Dims& d, const HostMetrics& hm, int nsolve_max,
                  int block, size_t lds_budget_doubles);
void mg_plan_free(MgPlan& mg);
// solves nsolve planes of mg.rhs (RHS_TMP) into mg.phi (iguess: mg.phi holds
// the initial guesses); cycle counts in mg.cycles
int mg_solve_ex(MgPlan& mg, const Dims& d, int nsolve, int iguess, hipStream_t st);

// ---- consistent projection (ekm_mgt.hip) --------------------------------------
struct TightPlan {
  int fp32 = 0, cm = 0, nsolve_max = 0;
  bool cross = false;                 // operator has cross-derivative terms (9-point)
  bool per = false;                   // periodic z ends
  int lds_bytes = 0;                  // dynamic LDS of the solve kernel (coarse levels)
  long lds_budget = 0;                // cap on it in bytes (0: what the device allows)
  long ws_elems = 0, ws64_elems = 0;  // per solve
  char* mem = nullptr;                // descriptor + coefficient tables
  void* fac = nullptr;                // line factors [live mode][...]
  void* ws = nullptr;                 // [nsolve][ws_elems] (float or double)
  double* ws64 = nullptr;             // [nsolve][ws64_elems]
};
// kx2_mode: kx^2 of modes 0..nmode-1 (solve s uses mode s/2)
int tight_plan_build(TightPlan& tp, const ekm_params& p, const Dims& d, const HostMetrics& hm, const Ranges& rv,
                     const double* kx2_mode, int nmode, int nsolve_max, long lds_budget = 0);
void tight_plan_free(TightPlan& tp);
// solves nsolve planes of mg.rhs (RHS_TMP) into mg.phi; V-cycle counts in
// mg.cycles.  guess (device pointer, may be null): when *guess != 0, start
// from *guess * (mg.phi on entry), else from zero.
int tight_solve_ex(TightPlan& tp, MgPlan& mg, const Dims& d, int nsolve, const double* guess, hipStream_t st);

// ---- solver -------------------------------------------------------------------
struct ekm_solver_impl;
