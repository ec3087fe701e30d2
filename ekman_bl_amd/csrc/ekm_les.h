// ekm_les.h -- the dynamic Smagorinsky subgrid model of the velocity
// (src/les_chan.F90 les_chan, LES_MODEL_TYPE 2).
#pragma once
#include "ekm_internal.h"

// device fields of the model: S[6] the strain rate (Fourier, then physical)
// and afterwards the subgrid stress, |S|, the eddy viscosity NU_T (physical,
// full P fields), the filter width Delta and the dynamic coefficient C_DYN
// ((NZ+2) x (NY+2) planes)
struct LesFields {
  double* S[6];
  double* sabs;
  double* nut;
  double* delta;
  double* cdyn;
  // the scalar model (src/les_chan_th.F90): eddy diffusivity KAPPA_T (P
  // field) and the dynamic coefficient C_DYN_TH ((NZ+2) x (NY+2) plane)
  double* kappa;
  double* cdyn_th;
};

// scratch of the dynamic coefficient: filtered velocities, filtered |S|,
// three P fields of intermediate products and the two plane sums
struct LesScratch {
  double* ub[3];
  double* s2b;
  double* t1;
  double* m0;
  double* fm0;
  double* ft2;
  double* nsum;
  double* dsum;
};

// les_chan: U in Fourier space on entry, physical on exit; NU_T set; the
// divergence of the subgrid stress added to CF (Fourier).  `refresh`
// recomputes C_DYN (every fifth step at the first substep, and on the first
// step of a run).
int les_chan_run(const FftPlan& fft, const Dims& d, const Metrics& M, const Ranges& r, const double* kx,
                 double* const U[3], double* const CF[3], const LesFields& L, const LesScratch& w, bool refresh,
                 hipStream_t st);

// les_chan_th (LES_MODEL_TYPE_TH 2) for the scalar, run after les_chan in the
// same substep: TH in Fourier space on entry, physical on exit; KAPPA_T set;
// the x part of the subgrid scalar flux added to CFTH.  Uses |S| and, on a
// `refresh` (every tenth step at the first substep, and on the first step of
// a run), the filtered velocities and |S| that les_chan left in w; S[0..5]
// serve as scratch.  rv: velocity index ranges, rt: the scalar's.
int les_chan_th_run(const FftPlan& fft, const Dims& d, const Metrics& M, const Ranges& rv, const Ranges& rt,
                    const double* kx, const double* thbar, double* const U[3], double* TH, double* CFTH,
                    const LesFields& L, const LesScratch& w, bool refresh, hipStream_t st);
