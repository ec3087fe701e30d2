// ekm_solver.hip -- device-resident duct solver and the C ABI
// (include/ekman_bl_amd.h).  One solver owns one HIP stream and all of its
// fields; host<->device copies happen only in set/get calls.
#include <algorithm>
#include <cmath>
#include <cstddef>
#include <map>
#include <cstdlib>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "ekm_host.h"
#include "ekm_internal.h"
#include "ekm_kernels.h"
#include "ekm_les.h"
#include "ekm_wsnet.h"

static thread_local std::string g_last_error;

void ekm_note_error(const char* msg) {
  g_last_error = msg;
  std::fprintf(stderr, "ekman_bl_amd: %s\n", msg);
}

void ekm_note_hip_error(hipError_t e, const char* what, const char* file, int line) {
  char buf[512];
  std::snprintf(buf, sizeof(buf), "%s failed: %s (%s:%d)", what, hipGetErrorString(e), file, line);
  g_last_error = buf;
  std::fprintf(stderr, "ekman_bl_amd: %s\n", buf);
}

struct ekm_solver {
  ekm_params p;
  ekm_tuning tn;
  Dims d;
  Ranges rv, rt;
  hipStream_t st = nullptr;
  FftPlan fft;
  MgPlan mg;
  TightPlan tp;  // proj_mode 1, 2
  HostMetrics hm;
  double* met = nullptr;
  Metrics M;
  double *kx = nullptr, *kx2 = nullptr, *kx2s = nullptr, *thbar = nullptr;
  double* U[3] = {nullptr, nullptr, nullptr};
  double *P = nullptr, *TH = nullptr;
  double* CF[3] = {nullptr, nullptr, nullptr};
  double* CFTH = nullptr;
  double* R[3] = {nullptr, nullptr, nullptr};
  double* RTH = nullptr;
  double *S1 = nullptr, *S2 = nullptr, *U2b = nullptr, *U3b = nullptr, *PHI = nullptr;
  double* lu = nullptr;     // line factors of the ADI fields, 7 * plane doubles each
  AdiField adi_h[kAdiMaxFields];
  AdiField* adi_d = nullptr;  // the same descriptors on the device
  int adi_nf = 0;
  // run state that changes from substep to substep (DELTA_T, H_BAR, TIME,
  // TEMP1..5, controller, reductions, diagnostics, status) lives on the
  // device; the host reads it only when a caller asks (ekm_get_diag, ...)
  StepState* ss = nullptr;
  double* red = nullptr;  // ss->red
  BcTable bcy[4], bcz[4];
  double beta_bar[3], zeta_bar[3];
  int rk_step = 1, time_step = 0;
  int status_seen = EKM_OK;  // the sticky status as last read back
  std::vector<int> cyc_host;
  int kx2_dirty = 1;  // mg.kx2 does not hold the solver's wavenumbers
  // captured substeps, one executable graph per host-decided variant
  std::map<unsigned, hipGraphExec_t> graphs;
  // live timing of the multigrid launch with HIP events on the solver stream
  // (eager substeps only): pairs recorded since the last fold
  int prof = 0;
  std::vector<hipEvent_t> evs;
  size_t ev_used = 0;
  double prof_ms = 0.0;
  long prof_timed = 0;
  // LES (src/les_chan.F90): model fields and scratch, and the reference's
  // CREATE_NEW_FLOW / FIRST_TIME, which decide when the model runs and when
  // its coefficient is recomputed
  LesFields les{};
  LesScratch lesw{};
  double* lesmet = nullptr;           // GMAT_11_z,12_z,22_z,11_y,12_y,22_y (two faces each)
  AdiField adi_hl[kAdiMaxFields];     // ADI descriptors with the subgrid viscosity
  AdiField* adi_dl = nullptr;
  int create_new_flow = 1, first_time = 1;
  int have_prev_phi = 0;  // mg.phi holds the previous substep's correction
  ekm_wsnet* net = nullptr;  // learned initial guess (proj_guess = 2), not owned
  // F_TYPE 0 flow-rate controller (src/duct.F90:414-436): its PX0 and Qn are
  // in the step state, started at Q0 = Qn = WBULK0 (src/diablo.F90:359-361);
  // y spacings of its integral
  double* fdy = nullptr;
  double q_den = 1.0;  // (ypoint(1,JEND) - ypoint(1,JSTART)) * (ZEND-ZSTART-1)
  // sponge_th (src/duct.F90:3196-3362): SPONGE_SIGMA(0:NY+1) and
  // INT_JACOB(ZEND+1, JEND+1), the Jacobian its type-6 branch reads
  double* sponge = nullptr;
  double sponge_jst = 0.0;
  double* sponge_out = nullptr;  // SPONGE_SIGMA_OUT(0:NZ+1), when the streamwise profile is on
  double* wall = nullptr;        // the wall model's row values [2][NZ2][NX2] (W_BC_YMIN = 3)
  double wall_z = 0.0;           // DYY(2,2)
};

extern "C" {

const char* ekm_version(void) { return "ekman_bl_amd 0.1.0 gfx950"; }
const char* ekm_last_error(void) { return g_last_error.c_str(); }

void ekm_tuning_default(ekm_tuning* t) {
  t->adi_rows = 8;
  t->adi_block = 256;
  t->adi_les_rows = 6;
  t->mg_block = 512;
  t->tight_lds_kib = 0;
}

static int check_tuning(const ekm_tuning& t) {
  if (t.adi_rows != 4 && t.adi_rows != 8 && t.adi_rows != 12) return EKM_ERR_ARG;
  if (t.adi_block != 64 && t.adi_block != 128 && t.adi_block != 256) return EKM_ERR_ARG;
  if (t.adi_les_rows != 4 && t.adi_les_rows != 5 && t.adi_les_rows != 6 && t.adi_les_rows != 8) return EKM_ERR_ARG;
  if (t.mg_block != 256 && t.mg_block != 512 && t.mg_block != 1024) return EKM_ERR_ARG;
  if (t.tight_lds_kib < 0) return EKM_ERR_ARG;
  return EKM_OK;
}

void ekm_params_default(ekm_params* p) {
  std::memset(p, 0, sizeof(*p));
  p->nu = 1.0e-3;
  p->lx = 4.0 * M_PI;
  p->ly = 20.0;
  p->lz = 2.0 * M_PI;
  p->delta_t = 0.1;
  p->variable_dt = 1;
  p->cfl = 0.5;
  p->update_dt = 1;
  p->f_type = 3;
  p->pr = 1.0;
  p->mg_maxcy = 6;
  p->mg_tolmax = 1.0e-6;
  for (int s = 0; s < 4; ++s) p->mudpack_bc[s] = 2;
  p->proj_mode = 0;
  p->proj_maxit = 50;
  p->proj_tol = 1.0e-12;
  p->proj_krylov = 1;
  p->ny_s_t = p->nz_s_r = p->save_stats_int = 1 << 30;  // no sponge rows
  p->nz_s_l = -(1 << 30);
  p->wall_model_type = 3;  // rundir_unst/input.dat:11
}

}  // extern "C"

// ---------------------------------------------------------------------------
static int check_params(const ekm_params& p) {
  if (p.nx < 8 || p.nx > 1024 || (p.nx & (p.nx - 1))) return EKM_ERR_UNSUPPORTED;
  if (p.ny < 4 || p.nz < 4) return EKM_ERR_ARG;
  if (p.n_th < 0 || p.n_th > 1) return EKM_ERR_UNSUPPORTED;
  if (p.nz + 2 != p.mg_ixp * (1 << (p.mg_iex - 1)) + 1) return EKM_ERR_ARG;   // src/mg_solver2d.F90:138-140
  if (p.ny + 2 != p.mg_jyq * (1 << (p.mg_jey - 1)) + 1) return EKM_ERR_ARG;
  if (p.mg_ixp < 2 || p.mg_jyq < 2 || p.mg_maxcy < 1) return EKM_ERR_ARG;
  if (p.f_type < 0 || p.f_type > 3) return EKM_ERR_UNSUPPORTED;  // 3: no mean forcing
  // MOD(TIME_STEP, UPDATE_DT) decides the CFL update (src/duct.F90:2417)
  if (p.variable_dt && p.update_dt < 1) return EKM_ERR_ARG;
  if (!(p.delta_t > 0.0)) return EKM_ERR_ARG;
  for (int s = 0; s < 4; ++s)
    if (p.mudpack_bc[s] < 0 || p.mudpack_bc[s] > 2) return EKM_ERR_UNSUPPORTED;
  if ((p.mudpack_bc[0] == 0) != (p.mudpack_bc[1] == 0)) return EKM_ERR_ARG;
  if ((p.mudpack_bc[2] == 0) != (p.mudpack_bc[3] == 0)) return EKM_ERR_ARG;
  if (p.proj_mode < 0 || p.proj_mode > 2) return EKM_ERR_UNSUPPORTED;
  if (p.u_bc[2][EKM_YMIN] == 3) {
    // LES_WALL_MODEL: type 3 only (type 1 reads a U_STAR it never sets,
    // type 2 a CU1X_WALL_MODEL it allocates and never assigns, type 4 the scalar's new
    // value, src/les_chan.F90:1914-2144); its z means read tmp_u(1:NZ),
    // which it fills over ZSTART..ZEND only
    if (p.wall_model_type != 3) return EKM_ERR_UNSUPPORTED;
    const int uz0 = p.u_bc[0][EKM_ZMIN], uz1 = p.u_bc[0][EKM_ZMAX];
    if (!(uz0 == 1 || uz0 == 6) || !(uz1 == 0 || uz1 == 1 || uz1 == 6)) return EKM_ERR_UNSUPPORTED;
    if (p.nz + 2 > kWallMaxNz2) return EKM_ERR_UNSUPPORTED;
  }
  if (p.les_model != 0 && p.les_model != 2) return EKM_ERR_UNSUPPORTED;  // dynamic Smagorinsky only
  // the scalar's subgrid model with a non-uniform background (cont_strat 0)
  // adds KAPPA_T indexed by Fourier mode to the background diffusion
  // (src/duct.F90:610-618); that form is not supported
  if (p.les_model && p.n_th && !p.cont_strat) return EKM_ERR_UNSUPPORTED;
  if (p.sponge_type < 0 || p.sponge_type > 2) return EKM_ERR_UNSUPPORTED;
  if (p.sponge_type) {
    // NY_S_T = NY+1 makes the vertical profile 0/0 (src/diablo.F90:1538,1550);
    // the streamwise one (NZ_S_R < NZ+1) reads xpoint(NZ_S_R) and xpoint(NZ_S_L)
    if (p.ny_s_t < 0 || p.ny_s_t == p.ny + 1 || p.save_stats_int < 1) return EKM_ERR_ARG;
    if (p.nz_s_r < p.nz + 1 && (p.nz_s_r < 0 || p.nz_s_l < 0 || p.nz_s_l > p.nz + 1)) return EKM_ERR_ARG;
  }
  if (p.proj_mode > 0) {
    // periodic z ends are wrapped z lines; periodic y ends are not handled
    if (p.mudpack_bc[2] == 0) return EKM_ERR_UNSUPPORTED;
    if (!(p.proj_tol > 0.0) || p.proj_maxit < 1) return EKM_ERR_ARG;
    if (p.proj_guess < 0 || p.proj_guess > 2) return EKM_ERR_ARG;
    if (p.proj_krylov < 0 || p.proj_krylov > 1) return EKM_ERR_ARG;
  }
  return EKM_OK;
}

// INIT_DUCT (src/duct.F90:7-232) for IC_TYPE = 7
static void init_ranges(const ekm_params& p, Ranges& rv, Ranges& rt) {
  int NY = p.ny, NZ = p.nz;
  int wy0 = p.u_bc[2][EKM_YMIN], wy1 = p.u_bc[2][EKM_YMAX];
  int uz0 = p.u_bc[0][EKM_ZMIN], uz1 = p.u_bc[0][EKM_ZMAX];
  rv.js = (wy0 == 0 || wy0 == 1 || wy0 == 6 || wy0 == 7 || wy0 == 3) ? 1 : 2;
  rv.zs = (uz0 == 1 || uz0 == 6) ? 1 : 2;
  rv.je = (wy1 == 0 || wy1 == 1 || wy1 == 6) ? NY : NY - 1;
  rv.ze = (uz1 == 0 || uz1 == 1 || uz1 == 6) ? NZ : NZ - 1;
  int ty0 = p.th_bc[EKM_YMIN], ty1 = p.th_bc[EKM_YMAX];
  int tz0 = p.th_bc[EKM_ZMIN], tz1 = p.th_bc[EKM_ZMAX];
  rt.js = (ty0 == 0 || ty0 == 1 || ty0 == 5 || ty0 == 2 || ty0 == 6 || ty0 == 7 || ty0 == 8) ? 1 : 2;
  rt.zs = (tz0 == 0 || tz0 == 1 || tz0 == 6) ? 1 : 2;
  rt.je = (ty1 == 0 || ty1 == 1 || ty1 == 6 || ty1 == 5) ? NY : NY - 1;
  rt.ze = (tz1 == 0 || tz1 == 1 || tz1 == 5 || tz1 == 6) ? NZ : NZ - 1;
}

// device address of a field of the step state
#define SS_PTR(s, field) \
  (reinterpret_cast<decltype(&(static_cast<StepState*>(nullptr))->field)>(reinterpret_cast<char*>((s)->ss) + \
                                                                         offsetof(StepState, field)))

// host -> step state, on the solver's stream (the caller's value is on its
// stack, so the copy completes before returning)
static int ss_put(ekm_solver* s, size_t off, const void* v, size_t n) {
  EKM_CHECK(hipMemcpyAsync(reinterpret_cast<char*>(s->ss) + off, v, n, hipMemcpyHostToDevice, s->st));
  EKM_CHECK(hipStreamSynchronize(s->st));
  return EKM_OK;
}

// waits for the stream and reads the step state; records its sticky status
static int ss_get(ekm_solver* s, StepState* h) {
  EKM_CHECK(hipMemcpyAsync(h, s->ss, sizeof(StepState), hipMemcpyDeviceToHost, s->st));
  EKM_CHECK(hipStreamSynchronize(s->st));
  s->status_seen = h->status;
  return EKM_OK;
}

// a new run (initial condition or restart): TIME and TIME_STEP as given, the
// status cleared
static int ss_restart(ekm_solver* s, double time) {
  const int ok = EKM_OK;
  if (ss_put(s, offsetof(StepState, time), &time, sizeof(double)) != EKM_OK) return EKM_ERR_HIP;
  if (ss_put(s, offsetof(StepState, status), &ok, sizeof(int)) != EKM_OK) return EKM_ERR_HIP;
  s->status_seen = EKM_OK;
  return EKM_OK;
}

// the solver's wavenumbers back into mg.kx2 after the multigrid drop-in used it
static int restore_kx2(ekm_solver* s) {
  if (!s->kx2_dirty) return EKM_OK;
  EKM_CHECK(hipMemcpyAsync(s->mg.kx2, s->kx2s, 2 * (s->d.NKX + 1) * sizeof(double), hipMemcpyDeviceToDevice, s->st));
  s->kx2_dirty = 0;
  return EKM_OK;
}

// sums the recorded solve timings (waits for the stream)
static int fold_events(ekm_solver* s) {
  if (!s->ev_used) return EKM_OK;
  EKM_CHECK(hipStreamSynchronize(s->st));
  for (size_t q = 0; q < s->ev_used; ++q) {
    float ms = 0.f;
    EKM_CHECK(hipEventElapsedTime(&ms, s->evs[2 * q], s->evs[2 * q + 1]));
    s->prof_ms += ms;
  }
  s->ev_used = 0;
  return EKM_OK;
}

static int free_all(ekm_solver* s) {
  double* f[] = {s->U[0], s->U[1], s->U[2], s->P, s->TH, s->CF[0], s->CF[1], s->CF[2], s->CFTH,
                 s->R[0], s->R[1], s->R[2], s->RTH, s->S1, s->S2, s->U2b, s->U3b, s->PHI, s->lu,
                 s->met, s->kx, s->kx2, s->kx2s, s->thbar, s->lesmet, s->fdy, s->sponge, s->sponge_out, s->wall};
  for (double* q : f)
    if (q) (void)hipFree(q);
  if (s->adi_d) (void)hipFree(s->adi_d);
  if (s->adi_dl) (void)hipFree(s->adi_dl);
  {
    double* lf[] = {s->les.S[0], s->les.S[1], s->les.S[2], s->les.S[3], s->les.S[4], s->les.S[5], s->les.sabs,
                    s->les.nut, s->les.delta, s->les.cdyn, s->les.kappa, s->les.cdyn_th,
                    s->lesw.ub[0], s->lesw.ub[1], s->lesw.ub[2],
                    s->lesw.s2b, s->lesw.t1, s->lesw.m0, s->lesw.fm0, s->lesw.ft2, s->lesw.nsum, s->lesw.dsum};
    for (double* q : lf)
      if (q) (void)hipFree(q);
  }
  fft_plan_free(s->fft);
  mg_plan_free(s->mg);
  tight_plan_free(s->tp);
  if (s->ss) (void)hipFree(s->ss);
  for (auto& g : s->graphs) (void)hipGraphExecDestroy(g.second);
  s->graphs.clear();
  for (hipEvent_t e : s->evs) (void)hipEventDestroy(e);
  if (s->st) (void)hipStreamDestroy(s->st);
  return EKM_OK;
}

// metric of the subgrid-viscosity terms of velocity component c (0..2):
// which 0: z direction (GMAT_11, GMAT_11_y, GMAT_11_z for U1, U2, U3), 1: y
// direction (GMAT_22, GMAT_22_y, GMAT_22_z), 2: the explicit cross terms
// (GMAT_12, GMAT_12_y, GMAT_12_z) (src/duct.F90:990-1120, :1667-2391)
static const double* les_metric(const ekm_solver* s, int c, int which) {
  if (c == 0) return which == 0 ? s->M.g11 : which == 1 ? s->M.g22 : s->M.g12;
  const size_t mn = s->d.mplane;
  const int base = c == 2 ? 0 : 3;  // _z variants first, then _y
  const int w = which == 0 ? 0 : which == 1 ? 2 : 1;
  return s->lesmet + (size_t)(base + w) * 2 * mn;
}

static int fwd(ekm_solver* s, double* f) { return fft_forward(s->fft, f, (long)s->d.plane, s->st); }
static int inv(ekm_solver* s, double* f) { return fft_inverse(s->fft, f, (long)s->d.plane, s->st); }

#define TRY(x)                \
  do {                        \
    int e__ = (x);            \
    if (e__ != EKM_OK) return e__; \
  } while (0)

extern "C" int ekm_create_tuned(const ekm_params* pp, const double* xpoint, const double* ypoint,
                                const double* th_bar, const ekm_tuning* tune, ekm_solver** out) {
  if (!pp || !xpoint || !ypoint || !out) return EKM_ERR_ARG;
  *out = nullptr;
  TRY(check_params(*pp));
  ekm_tuning tn;
  ekm_tuning_default(&tn);
  if (tune) tn = *tune;
  TRY(check_tuning(tn));
  // a scalar without the constant stratification reads its background TH_BAR
  // (src/diablo.F90:1152-1160), which the caller must then pass
  if (pp->n_th && !pp->cont_strat && !th_bar) return EKM_ERR_ARG;
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) return EKM_ERR_NODEV;
  ekm_solver* s = new ekm_solver();
  s->p = *pp;
  s->tn = tn;
  const ekm_params& p = s->p;
  Dims& d = s->d;
  d.NX = p.nx; d.NY = p.ny; d.NZ = p.nz;
  d.NXC = p.nx / 2 + 1; d.NKX = p.nx / 3;
  d.NX2 = p.nx + 2; d.NY2 = p.ny + 2; d.NZ2 = p.nz + 2;
  d.NZ3 = p.nz + 3; d.NY3 = p.ny + 3;
  d.npts = (size_t)d.NX2 * d.NZ2 * d.NY2;
  d.plane = (size_t)d.NZ2 * d.NY2;
  d.mplane = (size_t)d.NZ3 * d.NY3;
  init_ranges(p, s->rv, s->rt);
  int st = EKM_OK;
  for (int c = 0; c < 4 && st == EKM_OK; ++c) {
    if (c == 3 && !p.n_th) break;
    st = host_bc_table(p, c, 0, s->bcy[c]);
    if (st == EKM_OK) st = host_bc_table(p, c, 1, s->bcz[c]);
  }
  if (st != EKM_OK) { delete s; return st; }
  auto fail = [&](int e) { free_all(s); delete s; return e; };
  if (hipStreamCreateWithFlags(&s->st, hipStreamNonBlocking) != hipSuccess) return fail(EKM_ERR_HIP);
  // metrics (JACOBI_TRANS) on the host, then uploaded once
  host_metrics(d.NZ, d.NY, xpoint, ypoint, s->hm);
  size_t mn = d.mplane;
  if (hipMalloc(&s->met, 15 * mn * sizeof(double)) != hipSuccess) return fail(EKM_ERR_HIP);
  for (int w = 0; w < 7; ++w)
    if (hipMemcpy(s->met + (size_t)w * 2 * mn, s->hm.m[w].data(), 2 * mn * sizeof(double), hipMemcpyHostToDevice) != hipSuccess)
      return fail(EKM_ERR_HIP);
  if (hipMemcpy(s->met + 14 * mn, s->hm.jac.data(), mn * sizeof(double), hipMemcpyHostToDevice) != hipSuccess)
    return fail(EKM_ERR_HIP);
  s->M = Metrics{s->met, s->met + 2 * mn, s->met + 4 * mn, s->met + 6 * mn, s->met + 8 * mn,
                 s->met + 10 * mn, s->met + 12 * mn, s->met + 14 * mn};
  // wavenumbers (INIT_FFT: KXP, KX2P; src/ALLOCATION.F90:63-65)
  std::vector<double> kx(d.NXC), kx2(d.NXC), kx2s(2 * (d.NKX + 1));
  for (int m = 0; m < d.NXC; ++m) {
    kx[m] = 2.0 * M_PI * m / p.lx;
    kx2[m] = kx[m] * kx[m];
  }
  for (int m = 0; m <= d.NKX; ++m) kx2s[2 * m] = kx2s[2 * m + 1] = kx2[m];
  if (hipMalloc(&s->kx, d.NXC * sizeof(double)) != hipSuccess || hipMalloc(&s->kx2, d.NXC * sizeof(double)) != hipSuccess ||
      hipMalloc(&s->kx2s, kx2s.size() * sizeof(double)) != hipSuccess)
    return fail(EKM_ERR_HIP);
  (void)hipMemcpy(s->kx, kx.data(), d.NXC * sizeof(double), hipMemcpyHostToDevice);
  (void)hipMemcpy(s->kx2, kx2.data(), d.NXC * sizeof(double), hipMemcpyHostToDevice);
  (void)hipMemcpy(s->kx2s, kx2s.data(), kx2s.size() * sizeof(double), hipMemcpyHostToDevice);
  // background scalar (src/diablo.F90:1152-1160)
  std::vector<double> tb(d.plane, 0.0);
  for (int j = 0; j < d.NY2; ++j)
    for (int k = 0; k < d.NZ2; ++k) {
      size_t q = (size_t)k + (size_t)d.NZ2 * j;
      tb[q] = p.cont_strat ? (-ypoint[q] + ypoint[(size_t)d.NZ2 * (d.NY + 1)]) : (p.n_th ? th_bar[q] : 0.0);
    }
  if (hipMalloc(&s->thbar, d.plane * sizeof(double)) != hipSuccess) return fail(EKM_ERR_HIP);
  (void)hipMemcpy(s->thbar, tb.data(), d.plane * sizeof(double), hipMemcpyHostToDevice);
  // fields
  double** f[] = {&s->U[0], &s->U[1], &s->U[2], &s->P, &s->TH, &s->CF[0], &s->CF[1], &s->CF[2], &s->CFTH,
                  &s->R[0], &s->R[1], &s->R[2], &s->RTH, &s->S1, &s->S2, &s->U2b, &s->U3b, &s->PHI};
  for (double** q : f) {
    if (hipMalloc(q, d.npts * sizeof(double)) != hipSuccess) return fail(EKM_ERR_HIP);
    if (hipMemset(*q, 0, d.npts * sizeof(double)) != hipSuccess) return fail(EKM_ERR_HIP);
  }
  if (hipMalloc(&s->lu, (size_t)kAdiMaxFields * 7 * d.plane * sizeof(double)) != hipSuccess) return fail(EKM_ERR_HIP);
  // ADI descriptors: U1, U2, U3 and theta; their scratch S fields are the
  // buffers that are free between the fluxes and the projection
  {
    double* Sx[kAdiMaxFields] = {s->S1, s->S2, s->U2b, s->U3b};
    const int nf = p.n_th ? 4 : 3;
    bool wall_rows = false;
    for (int c = 0; c < 3; ++c) wall_rows = wall_rows || p.u_bc[c][EKM_YMIN] == 3;
    if (wall_rows) {
      const size_t n = 2 * (size_t)d.NX2 * d.NZ2;
      if (hipMalloc(&s->wall, n * sizeof(double)) != hipSuccess) return fail(EKM_ERR_HIP);
      if (hipMemset(s->wall, 0, n * sizeof(double)) != hipSuccess) return fail(EKM_ERR_HIP);
      s->wall_z = ypoint[2 + (size_t)d.NZ2 * 2] - ypoint[2 + (size_t)d.NZ2 * 1];
    }
    for (int f = 0; f < nf; ++f) {
      AdiField& a = s->adi_h[f];
      const int c = f < 3 ? f : 3;
      a.r = f < 3 ? s->rv : s->rt;
      a.by = s->bcy[c];
      a.bz = s->bcz[c];
      a.U = f < 3 ? s->U[f] : s->TH;
      a.R = f < 3 ? s->R[f] : s->RTH;
      a.S = Sx[f];
      a.lu = s->lu + (size_t)f * 7 * d.plane;
      a.nut = nullptr;
      a.gz = a.gy = nullptr;
      a.pv = nullptr;
      a.wall = f < 3 ? s->wall : nullptr;
    }
    s->adi_nf = nf;
    if (hipMalloc(&s->adi_d, sizeof(s->adi_h)) != hipSuccess) return fail(EKM_ERR_HIP);
    if (hipMemcpy(s->adi_d, s->adi_h, sizeof(s->adi_h), hipMemcpyHostToDevice) != hipSuccess) return fail(EKM_ERR_HIP);
  }
  if (p.les_model) {
    // subgrid model fields and the filter width Delta = (DX J)^(1/3) at the
    // cell centres (src/les_chan.F90:343-348, beta = 1)
    double** pf[] = {&s->les.S[0], &s->les.S[1], &s->les.S[2], &s->les.S[3], &s->les.S[4], &s->les.S[5],
                     &s->les.sabs, &s->les.nut, &s->lesw.ub[0], &s->lesw.ub[1], &s->lesw.ub[2], &s->lesw.s2b,
                     &s->lesw.t1, &s->lesw.m0, &s->lesw.fm0, &s->lesw.ft2};
    for (double** q : pf) {
      if (hipMalloc(q, d.npts * sizeof(double)) != hipSuccess) return fail(EKM_ERR_HIP);
      if (hipMemset(*q, 0, d.npts * sizeof(double)) != hipSuccess) return fail(EKM_ERR_HIP);
    }
    double** pp[] = {&s->les.delta, &s->les.cdyn, &s->lesw.nsum, &s->lesw.dsum};
    for (double** q : pp) {
      if (hipMalloc(q, d.plane * sizeof(double)) != hipSuccess) return fail(EKM_ERR_HIP);
      if (hipMemset(*q, 0, d.plane * sizeof(double)) != hipSuccess) return fail(EKM_ERR_HIP);
    }
    if (p.n_th) {  // the scalar model's KAPPA_T and C_DYN_TH
      if (hipMalloc(&s->les.kappa, d.npts * sizeof(double)) != hipSuccess ||
          hipMalloc(&s->les.cdyn_th, d.plane * sizeof(double)) != hipSuccess)
        return fail(EKM_ERR_HIP);
      if (hipMemset(s->les.kappa, 0, d.npts * sizeof(double)) != hipSuccess ||
          hipMemset(s->les.cdyn_th, 0, d.plane * sizeof(double)) != hipSuccess)
        return fail(EKM_ERR_HIP);
    }
    std::vector<double> dl(d.plane);
    const double dx = p.lx / (double)d.NX;
    for (int j = 0; j < d.NY2; ++j)
      for (int k = 0; k < d.NZ2; ++k)
        dl[(size_t)k + (size_t)d.NZ2 * j] = std::pow(dx * s->hm.jac[(size_t)k + (size_t)d.NZ3 * j], 1.0 / 3.0);
    if (hipMemcpy(s->les.delta, dl.data(), d.plane * sizeof(double), hipMemcpyHostToDevice) != hipSuccess)
      return fail(EKM_ERR_HIP);
    // metrics of the subgrid stress (src/diablo.F90:1379-1474) and the ADI
    // descriptors of the velocities with NU_T; their pivot scratch fields are
    // the coefficient scratch, which is free once the model has run
    if (hipMalloc(&s->lesmet, 12 * mn * sizeof(double)) != hipSuccess) return fail(EKM_ERR_HIP);
    for (int w = 0; w < 6; ++w)
      if (hipMemcpy(s->lesmet + (size_t)w * 2 * mn, s->hm.m[7 + w].data(), 2 * mn * sizeof(double),
                    hipMemcpyHostToDevice) != hipSuccess)
        return fail(EKM_ERR_HIP);
    double* pvs[kAdiMaxFields] = {s->lesw.t1, s->lesw.m0, s->lesw.fm0, s->lesw.ft2};
    for (int f = 0; f < kAdiMaxFields; ++f) {
      s->adi_hl[f] = s->adi_h[f];
      if (f >= s->adi_nf) continue;
      AdiField& a = s->adi_hl[f];
      // the scalar's terms carry KAPPA_T with the molecular metrics
      // (src/duct.F90:1373-1385, :1437-1450, :1489-1500, :1557-1570)
      a.nut = f < 3 ? s->les.nut : s->les.kappa;
      a.gz = f < 3 ? les_metric(s, f, 0) : s->M.g11;
      a.gy = f < 3 ? les_metric(s, f, 1) : s->M.g22;
      a.pv = pvs[f];
    }
    if (hipMalloc(&s->adi_dl, sizeof(s->adi_hl)) != hipSuccess) return fail(EKM_ERR_HIP);
    if (hipMemcpy(s->adi_dl, s->adi_hl, sizeof(s->adi_hl), hipMemcpyHostToDevice) != hipSuccess)
      return fail(EKM_ERR_HIP);
  }
  {
    StepState h;
    std::memset(&h, 0, sizeof(h));
    h.delta_t = h.delta_t_set = p.delta_t;
    h.h_bar[0] = h.delta_t * (8.0 / 15.0);
    h.h_bar[1] = h.delta_t * (2.0 / 15.0);
    h.h_bar[2] = h.delta_t * (5.0 / 15.0);
    h.px0 = p.px0;
    h.qn = p.wbulk0;
    h.red[2] = 999.0;  // the CFL minimum starts above any time step
    if (hipMalloc(&s->ss, sizeof(StepState)) != hipSuccess) return fail(EKM_ERR_HIP);
    if (hipMemcpy(s->ss, &h, sizeof(h), hipMemcpyHostToDevice) != hipSuccess) return fail(EKM_ERR_HIP);
    s->red = SS_PTR(s, red[0]);
  }
  if (p.sponge_type && p.n_th) {
    std::vector<double> sig(d.NY2);
    host_sponge_sigma(p.sponge_type, d.NZ, d.NY, p.ny_s_t, ypoint, sig.data());
    if (hipMalloc(&s->sponge, d.NY2 * sizeof(double)) != hipSuccess) return fail(EKM_ERR_HIP);
    if (hipMemcpy(s->sponge, sig.data(), d.NY2 * sizeof(double), hipMemcpyHostToDevice) != hipSuccess)
      return fail(EKM_ERR_HIP);
    s->sponge_jst = s->hm.jac[(size_t)(s->rv.ze + 1) + (size_t)d.NZ3 * (s->rv.je + 1)];
    if (p.nz_s_r < p.nz + 1) {
      std::vector<double> so(d.NZ2);
      host_sponge_sigma_out(p.sponge_type, d.NZ, d.NY, p.nz_s_r, p.nz_s_l, xpoint, so.data());
      if (hipMalloc(&s->sponge_out, d.NZ2 * sizeof(double)) != hipSuccess) return fail(EKM_ERR_HIP);
      if (hipMemcpy(s->sponge_out, so.data(), d.NZ2 * sizeof(double), hipMemcpyHostToDevice) != hipSuccess)
        return fail(EKM_ERR_HIP);
    }
  }
  if (p.f_type == 0) {
    std::vector<double> dy(d.plane, 0.0);
    for (int j = 0; j + 1 < d.NY2; ++j)
      for (int k = 0; k < d.NZ2; ++k)
        dy[(size_t)k + (size_t)d.NZ2 * j] = ypoint[(size_t)k + (size_t)d.NZ2 * (j + 1)] - ypoint[(size_t)k + (size_t)d.NZ2 * j];
    const Ranges& r = s->rv;
    const double span = ypoint[1 + (size_t)d.NZ2 * r.je] - ypoint[1 + (size_t)d.NZ2 * r.js];
    if (r.ze - r.zs - 1 < 1 || r.je - r.js < 1 || !(span != 0.0)) return fail(EKM_ERR_UNSUPPORTED);
    s->q_den = span;
    if (hipMalloc(&s->fdy, d.plane * sizeof(double)) != hipSuccess) return fail(EKM_ERR_HIP);
    if (hipMemcpy(s->fdy, dy.data(), d.plane * sizeof(double), hipMemcpyHostToDevice) != hipSuccess)
      return fail(EKM_ERR_HIP);
  }
  if (fft_plan_init(s->fft, p.nx) != EKM_OK) return fail(EKM_ERR_HIP);
  int nsolve = 2 * (d.NKX + 1);
  {
    // one multigrid block per solve; the coarse levels live in LDS, sized by
    // what one workgroup may allocate on this device
    int lds_bytes = 0, dev = 0;
    if (hipGetDevice(&dev) != hipSuccess ||
        hipDeviceGetAttribute(&lds_bytes, hipDeviceAttributeMaxSharedMemoryPerBlock, dev) != hipSuccess)
      return fail(EKM_ERR_HIP);
    const int block = tn.mg_block;  // a power of two: tree reductions
    const size_t budget = (size_t)lds_bytes / sizeof(double);
    st = mg_plan_build(s->mg, p, d, s->hm, nsolve, block, budget);
  }
  if (st != EKM_OK) return fail(st);
  if (p.proj_mode > 0) {
    st = tight_plan_build(s->tp, p, d, s->hm, s->rv, kx2.data(), d.NKX + 1, nsolve, (long)tn.tight_lds_kib * 1024);
    if (st != EKM_OK) return fail(st);
  }
  s->cyc_host.resize(nsolve);
  s->beta_bar[0] = 1.0; s->beta_bar[1] = 25.0 / 8.0; s->beta_bar[2] = 9.0 / 4.0;
  s->zeta_bar[0] = 0.0; s->zeta_bar[1] = -17.0 / 8.0; s->zeta_bar[2] = -5.0 / 4.0;
  if (hipStreamSynchronize(s->st) != hipSuccess) return fail(EKM_ERR_HIP);
  *out = s;
  return EKM_OK;
}

extern "C" int ekm_create(const ekm_params* p, const double* xpoint, const double* ypoint, const double* th_bar,
                          ekm_solver** out) {
  return ekm_create_tuned(p, xpoint, ypoint, th_bar, nullptr, out);
}

extern "C" int ekm_get_tuning(ekm_solver* s, ekm_tuning* t) {
  if (!t) return EKM_ERR_ARG;
  *t = s->tn;
  return EKM_OK;
}

extern "C" int ekm_destroy(ekm_solver* s) {
  if (!s) return EKM_OK;
  (void)hipStreamSynchronize(s->st);
  free_all(s);
  delete s;
  return EKM_OK;
}

extern "C" int ekm_sync(ekm_solver* s) {
  StepState h;
  TRY(ss_get(s, &h));
  return h.status;
}

extern "C" int ekm_field_device_ptr(ekm_solver* s, int f, void** dptr) {
  double* t[5] = {s->U[0], s->U[1], s->U[2], s->P, s->TH};
  if (f < 0 || f > 4) return EKM_ERR_ARG;
  *dptr = t[f];
  return EKM_OK;
}

extern "C" int ekm_set_state(ekm_solver* s, const double* cu1, const double* cu2, const double* cu3,
                             const double* cp, const double* cth, double time, int time_step) {
  s->have_prev_phi = 0;
  s->create_new_flow = 0;  // a restart: the subgrid model runs from the first step
  s->first_time = 1;
  size_t n = s->d.npts * sizeof(double);
  const double* src[4] = {cu1, cu2, cu3, cp};
  double* dst[4] = {s->U[0], s->U[1], s->U[2], s->P};
  for (int q = 0; q < 4; ++q) EKM_CHECK(hipMemcpyAsync(dst[q], src[q], n, hipMemcpyHostToDevice, s->st));
  if (s->p.n_th && cth) EKM_CHECK(hipMemcpyAsync(s->TH, cth, n, hipMemcpyHostToDevice, s->st));
  EKM_CHECK(hipStreamSynchronize(s->st));
  TRY(ss_restart(s, time));
  s->time_step = time_step;
  s->rk_step = 1;
  return EKM_OK;
}

// NU_T (physical, the layout of a field) and C_DYN ((NZ+2) x (NY+2), K
// fastest) of the subgrid model (src/modules.F90:431-483 les_chan_var)
extern "C" int ekm_get_xstats(ekm_solver* s, double* out) {
  if (!out) return EKM_ERR_ARG;
  const size_t n = 16 * s->d.plane * sizeof(double);
  double* dev = nullptr;
  EKM_CHECK(hipMalloc((void**)&dev, n));
  int st = launch_xstats(s->d, s->U[0], s->U[1], s->U[2], s->P, s->p.n_th ? s->TH : nullptr, dev, s->st);
  if (st == EKM_OK && hipMemcpyAsync(out, dev, n, hipMemcpyDeviceToHost, s->st) != hipSuccess) st = EKM_ERR_HIP;
  if (hipStreamSynchronize(s->st) != hipSuccess) st = EKM_ERR_HIP;
  (void)hipFree(dev);
  return st;
}

extern "C" int ekm_get_mean_grads(ekm_solver* s, double* out) {
  if (!out) return EKM_ERR_ARG;
  const ekm_params& p = s->p;
  const size_t n = (8 * s->d.plane + 1) * sizeof(double);
  MeanGradArgs a;
  a.u1 = (const double2*)s->U[0];
  a.u2 = (const double2*)s->U[1];
  a.u3 = (const double2*)s->U[2];
  a.th = p.n_th ? (const double2*)s->TH : nullptr;
  a.thbar = (p.n_th && !p.cont_strat) ? s->thbar : nullptr;
  a.rv = s->rv;
  a.ri_tau = p.ri_tau;
  a.ke_scale = p.amp_omega0 != 0.0 ? std::pow(p.amp_omega0 / p.omega0, -2.0) : 1.0;
  a.cont_strat = p.cont_strat;
  EKM_CHECK(hipMalloc((void**)&a.out, n));
  int st = launch_mean_grads(s->d, s->M, a, s->st);
  if (st == EKM_OK && hipMemcpyAsync(out, a.out, n, hipMemcpyDeviceToHost, s->st) != hipSuccess) st = EKM_ERR_HIP;
  if (hipStreamSynchronize(s->st) != hipSuccess) st = EKM_ERR_HIP;
  (void)hipFree(a.out);
  return st;
}

extern "C" int ekm_get_omega_x(ekm_solver* s, double* out) {
  if (!out) return EKM_ERR_ARG;
  const size_t n = s->d.plane * sizeof(double);
  double* dev = nullptr;
  EKM_CHECK(hipMalloc((void**)&dev, n));
  int st = launch_omega_x(s->d, s->M, s->U[1], s->U[2], dev, s->st);
  if (st == EKM_OK && hipMemcpyAsync(out, dev, n, hipMemcpyDeviceToHost, s->st) != hipSuccess) st = EKM_ERR_HIP;
  if (hipStreamSynchronize(s->st) != hipSuccess) st = EKM_ERR_HIP;
  (void)hipFree(dev);
  return st;
}

extern "C" int ekm_get_physical(ekm_solver* s, int f, double* out) {
  const Dims& d = s->d;
  if (!out || f < 0 || f > 4) return EKM_ERR_ARG;
  const double* src = f < 3 ? s->U[f] : f == 3 ? s->P : s->p.n_th ? s->TH : nullptr;
  if (!src) {
    std::memset(out, 0, d.npts * sizeof(double));
    return EKM_OK;
  }
  double* tmp = nullptr;
  EKM_CHECK(hipMalloc((void**)&tmp, d.npts * sizeof(double)));
  int st = EKM_OK;
  if (hipMemcpyAsync(tmp, src, d.npts * sizeof(double), hipMemcpyDeviceToDevice, s->st) != hipSuccess) st = EKM_ERR_HIP;
  if (st == EKM_OK) st = inv(s, tmp);
  if (st == EKM_OK && hipMemcpyAsync(out, tmp, d.npts * sizeof(double), hipMemcpyDeviceToHost, s->st) != hipSuccess)
    st = EKM_ERR_HIP;
  if (hipStreamSynchronize(s->st) != hipSuccess) st = EKM_ERR_HIP;
  (void)hipFree(tmp);
  return st;
}

extern "C" int ekm_get_plane(ekm_solver* s, int axis, int index, double* out) {
  const Dims& d = s->d;
  if (!out || axis < 0 || axis > 2) return EKM_ERR_ARG;
  const int hi = axis == 0 ? d.NZ + 1 : axis == 1 ? d.NX - 1 : d.NY + 1;
  if (index < 0 || index > hi) return EKM_ERR_ARG;
  const size_t m = plane_elems(d, axis);
  double* tmp = nullptr;  // one physical field, then the five planes
  EKM_CHECK(hipMalloc((void**)&tmp, (d.npts + 5 * m) * sizeof(double)));
  double* o = tmp + d.npts;
  const double* src[5] = {s->U[0], s->U[1], s->U[2], s->P, s->p.n_th ? s->TH : nullptr};
  int st = EKM_OK;
  for (int f = 0; f < 5 && st == EKM_OK; ++f) {
    if (!src[f]) {
      if (hipMemsetAsync(o + f * m, 0, m * sizeof(double), s->st) != hipSuccess) st = EKM_ERR_HIP;
      continue;
    }
    if (hipMemcpyAsync(tmp, src[f], d.npts * sizeof(double), hipMemcpyDeviceToDevice, s->st) != hipSuccess)
      st = EKM_ERR_HIP;
    if (st == EKM_OK) st = inv(s, tmp);
    if (st == EKM_OK) st = launch_plane_gather(d, axis, index, tmp, o + f * m, s->st);
  }
  if (st == EKM_OK && hipMemcpyAsync(out, o, 5 * m * sizeof(double), hipMemcpyDeviceToHost, s->st) != hipSuccess)
    st = EKM_ERR_HIP;
  if (hipStreamSynchronize(s->st) != hipSuccess) st = EKM_ERR_HIP;
  (void)hipFree(tmp);
  return st;
}

extern "C" int ekm_get_xmoments(ekm_solver* s, double* out) {
  if (!out) return EKM_ERR_ARG;
  const Dims& d = s->d;
  const size_t n = 13 * d.plane * sizeof(double);
  double* tmp = nullptr;  // three physical velocity fields, then the planes
  EKM_CHECK(hipMalloc((void**)&tmp, 3 * d.npts * sizeof(double) + n));
  double* o = tmp + 3 * d.npts;
  int st = EKM_OK;
  for (int f = 0; f < 3 && st == EKM_OK; ++f) {
    if (hipMemcpyAsync(tmp + f * d.npts, s->U[f], d.npts * sizeof(double), hipMemcpyDeviceToDevice, s->st) != hipSuccess)
      st = EKM_ERR_HIP;
    else
      st = inv(s, tmp + f * d.npts);
  }
  if (st == EKM_OK) st = launch_xmoments(d, tmp, tmp + d.npts, tmp + 2 * d.npts, s->U[0], s->U[1], s->U[2], o, s->st);
  if (st == EKM_OK && hipMemcpyAsync(out, o, n, hipMemcpyDeviceToHost, s->st) != hipSuccess) st = EKM_ERR_HIP;
  if (hipStreamSynchronize(s->st) != hipSuccess) st = EKM_ERR_HIP;
  (void)hipFree(tmp);
  return st;
}

extern "C" int ekm_get_forcing(ekm_solver* s, double* px0, double* qn) {
  StepState h;
  TRY(ss_get(s, &h));
  if (px0) *px0 = s->p.f_type == 0 ? h.px0 : s->p.px0;
  if (qn) *qn = h.qn;
  return EKM_OK;
}

extern "C" int ekm_get_les_th(ekm_solver* s, double* kappa, double* cdyn_th) {
  if (!s->p.les_model || !s->p.n_th) return EKM_ERR_ARG;
  if (kappa) EKM_CHECK(hipMemcpyAsync(kappa, s->les.kappa, s->d.npts * sizeof(double), hipMemcpyDeviceToHost, s->st));
  if (cdyn_th)
    EKM_CHECK(hipMemcpyAsync(cdyn_th, s->les.cdyn_th, s->d.plane * sizeof(double), hipMemcpyDeviceToHost, s->st));
  EKM_CHECK(hipStreamSynchronize(s->st));
  return EKM_OK;
}

extern "C" int ekm_get_les(ekm_solver* s, double* nut, double* cdyn) {
  if (!s->p.les_model) return EKM_ERR_ARG;
  if (nut) EKM_CHECK(hipMemcpyAsync(nut, s->les.nut, s->d.npts * sizeof(double), hipMemcpyDeviceToHost, s->st));
  if (cdyn) EKM_CHECK(hipMemcpyAsync(cdyn, s->les.cdyn, s->d.plane * sizeof(double), hipMemcpyDeviceToHost, s->st));
  EKM_CHECK(hipStreamSynchronize(s->st));
  return EKM_OK;
}

extern "C" int ekm_get_state(ekm_solver* s, double* cu1, double* cu2, double* cu3, double* cp, double* cth) {
  size_t n = s->d.npts * sizeof(double);
  double* dst[4] = {cu1, cu2, cu3, cp};
  double* src[4] = {s->U[0], s->U[1], s->U[2], s->P};
  for (int q = 0; q < 4; ++q)
    if (dst[q]) EKM_CHECK(hipMemcpyAsync(dst[q], src[q], n, hipMemcpyDeviceToHost, s->st));
  if (cth && s->p.n_th) EKM_CHECK(hipMemcpyAsync(cth, s->TH, n, hipMemcpyDeviceToHost, s->st));
  EKM_CHECK(hipStreamSynchronize(s->st));
  return EKM_OK;
}

// REM_DIV_CURV (src/duct.F90:2472-2818); phi is left in s->PHI.  The
// diagnostics and the status checks (src/duct.F90:2790-2812) are formed on
// the device into the step state; nothing is read back here.  guess: device
// pointer to the scale of the consistent solve's start (null: from zero).
// timed: HIP events around the solve launch (not inside a graph capture).
static int rem_div(ekm_solver* s, const double* guess, bool timed, bool learned = false) {
  const Dims& d = s->d;
  hipStream_t st = s->st;
  int nmode = d.NKX + 1, nsolve = 2 * nmode;
  TRY(restore_kx2(s));
  TRY(launch_faces(d, s->M, s->rv, s->U[1], s->U[2], s->U2b, s->U3b, st));
  TRY(launch_mg_gather(d, s->M, s->rv, s->kx, s->U[0], s->U2b, s->U3b, s->mg.rhs, nmode, st));
  // the learned start: the network reads this projection's right-hand sides
  // and the previous solutions in mg.phi and writes its guess there
  if (learned) TRY(wsnet_guess(s->net, s->mg.rhs, s->mg.phi, SS_PTR(s, prev_ratio), s->kx2s, nsolve, s->mg.phi, st));
  if (timed) {
    if (s->evs.size() < 2 * (s->ev_used + 1)) {
      hipEvent_t a, b;
      EKM_CHECK(hipEventCreate(&a));
      EKM_CHECK(hipEventCreate(&b));
      s->evs.push_back(a);
      s->evs.push_back(b);
    }
    EKM_CHECK(hipEventRecord(s->evs[2 * s->ev_used], st));
  }
  if (s->p.proj_mode == 0)
    TRY(mg_solve_ex(s->mg, d, nsolve, 0, st));
  else
    TRY(tight_solve_ex(s->tp, s->mg, d, nsolve, guess, st));
  if (timed) {
    EKM_CHECK(hipEventRecord(s->evs[2 * s->ev_used + 1], st));
    s->ev_used += 1;
    s->prof_timed += 1;
    if (s->ev_used >= 64) TRY(fold_events(s));
  }
  TRY(launch_mg_scatter(d, s->mg.phi, nmode, s->p.cmplx_fp32, s->PHI, st));
  TRY(launch_project(d, s->M, s->rv, s->kx, s->PHI, s->U2b, s->U3b, s->U[0], s->U[1], s->U[2], s->S1, st));
  TRY(inv(s, s->S1));
  TRY(launch_div_reduce(d, s->S1, s->red, st));
  // the MUDPACK replica stops at MAXCY as the reference's solver does, with no flag
  const int* unconv = s->p.proj_mode == 0 ? nullptr : s->mg.cycles + s->mg.nsolve_max;
  TRY(launch_div_final(s->ss, s->mg.cycles, unconv, nsolve, (double)(d.NX + 1) * (double)d.NY * (double)d.NZ, st));
  return EKM_OK;
}

extern "C" int ekm_rem_div_curv(ekm_solver* s) {
  s->have_prev_phi = 0;
  TRY(rem_div(s, nullptr, s->prof != 0));
  StepState h;
  TRY(ss_get(s, &h));
  return h.status;
}

extern "C" int ekm_get_phi(ekm_solver* s, double* phi) {
  EKM_CHECK(hipMemcpyAsync(phi, s->PHI, s->d.npts * sizeof(double), hipMemcpyDeviceToHost, s->st));
  EKM_CHECK(hipStreamSynchronize(s->st));
  return EKM_OK;
}

extern "C" int ekm_set_initial_physical(ekm_solver* s, const double* u1, const double* u2, const double* u3,
                                        const double* th) {
  size_t n = s->d.npts * sizeof(double);
  const double* src[3] = {u1, u2, u3};
  for (int q = 0; q < 3; ++q) {
    EKM_CHECK(hipMemcpyAsync(s->U[q], src[q], n, hipMemcpyHostToDevice, s->st));
    TRY(fwd(s, s->U[q]));
  }
  if (s->p.n_th && th) {
    EKM_CHECK(hipMemcpyAsync(s->TH, th, n, hipMemcpyHostToDevice, s->st));
    TRY(fwd(s, s->TH));
  }
  double* z[] = {s->P, s->CF[0], s->CF[1], s->CF[2], s->CFTH};
  for (double* q : z) EKM_CHECK(hipMemsetAsync(q, 0, n, s->st));
  TRY(ss_restart(s, 0.0));
  s->time_step = 0;
  s->rk_step = 1;
  s->have_prev_phi = 0;
  s->create_new_flow = 1;
  s->first_time = 1;
  TRY(rem_div(s, nullptr, false));
  StepState h;
  TRY(ss_get(s, &h));
  return h.status;
}

// What the host decides for a substep from its counters alone: the RK stage,
// when the subgrid models run and refresh their coefficients, the sponge's
// CU2 zeroing, the CFL update and the warm start.  A captured substep graph
// is keyed by it.
struct SubstepPlan {
  int rk, cur_step;
  bool les_on, refresh, refresh_th, zero_cu2, cfl_now;
  int guess_on;  // 0 from zero, 1 the previous correction scaled, 2 the network's guess
  bool wall_on, wall_first;  // LES_WALL_MODEL (W_BC_YMIN = 3), on TIME_STEP 1
  unsigned key() const {
    return (unsigned)rk | (les_on << 2) | (refresh << 3) | (refresh_th << 4) | (zero_cu2 << 5) | (cfl_now << 6) |
           ((unsigned)guess_on << 7) | ((unsigned)wall_on << 9) | ((unsigned)wall_first << 10);
  }
};

static SubstepPlan plan_substep(const ekm_solver* s) {
  const ekm_params& p = s->p;
  SubstepPlan q{};
  q.rk = s->rk_step;
  q.cur_step = s->time_step + 1;
  // the subgrid model is not run on the first step of a new flow (:641-657)
  q.les_on = p.les_model && (!s->create_new_flow || q.cur_step > 1);
  q.refresh = q.les_on && (((q.cur_step % 5) == 0 && q.rk == 1) || s->first_time);
  q.refresh_th = q.les_on && p.n_th && (((q.cur_step % 10) == 0 && q.rk == 1) || s->first_time);
  q.zero_cu2 = p.n_th && p.sponge_type && (2 * q.cur_step) % p.save_stats_int == 0;
  // CFL time step at the third substep (:2417-2419)
  q.cfl_now = p.variable_dt && q.rk == 3 && (q.cur_step % p.update_dt) == 0;
  // the wall model before the velocities' implicit solves (:1687-1691)
  q.wall_on = p.u_bc[2][EKM_YMIN] == 3;
  q.wall_first = q.wall_on && q.cur_step == 1;
  // proj_guess: the previous substep's correction, scaled by the ratio of the
  // RK weights, or the network's guess from it, starts the consistent solve
  q.guess_on = 0;
  if (p.proj_mode > 0 && p.proj_guess && s->have_prev_phi) {
    if (p.proj_guess == 1) q.guess_on = 1;
    else if (s->net) q.guess_on = 2;
  }
  return q;
}

// RK_DUCT (src/duct.F90:237-2469): enqueues one substep on the solver's
// stream.  Only kernels are enqueued (no host round trip, no memset), so the
// same sequence runs eagerly or is captured into a graph.
static int enqueue_substep(ekm_solver* s, const SubstepPlan& q, bool timed) {
  const ekm_params& p = s->p;
  const Dims& d = s->d;
  hipStream_t st = s->st;
  const int rk = q.rk;
  const int th = p.n_th > 0;
  {
    StepConsts c;
    c.nu = p.nu;
    for (int r = 0; r < 3; ++r) {
      c.beta[r] = s->beta_bar[r];
      c.zeta[r] = s->zeta_bar[r];
    }
    c.f_type = p.f_type;
    c.px0 = p.px0;
    c.amp_omega0 = p.amp_omega0;
    c.omega0 = p.omega0;
    c.guess_on = q.guess_on;
    TRY(launch_step_begin(s->ss, rk, c, st));
  }
  RhsArgs a;
  for (int f = 0; f < 3; ++f) {
    a.cu[f] = (const double2*)s->U[f];
    a.cr[f] = (double2*)s->R[f];
    a.cf[f] = (double2*)s->CF[f];
  }
  a.cp = (const double2*)s->P;
  a.cth = (const double2*)s->TH;
  a.crth = (double2*)s->RTH;
  a.cfth = (double2*)s->CFTH;
  a.kx = s->kx;
  a.kx2 = s->kx2;
  a.thbar = s->thbar;
  a.rv = s->rv;
  a.rt = s->rt;
  a.rk = rk;
  a.th = th;
  a.cont_strat = p.cont_strat;
  a.cor_mean_only = p.coriolis_mean_only;
  a.px_on = (p.f_type == 0 || p.f_type == 1 || p.f_type == 2);
  a.nu = p.nu;
  a.f0 = p.f0;
  a.ug = p.u_bc_c1[0][EKM_YMAX];
  a.wg = p.wbulk0;
  a.ri_tau = p.ri_tau;
  a.nupr = p.nu / p.pr;
  a.ss = s->ss;
  if (p.f_type == 0) {
    // the flow-rate controller (:414-436) on the state entering the substep
    TRY(launch_flow_rate(d, s->rv, s->U[2], s->fdy, s->red + 3, st));
    TRY(launch_flow_ctrl(s->ss, s->q_den, s->rv.ze - s->rv.zs - 1, p.wbulk0, st));
  }
  TRY(launch_rhs_fourier(d, s->M, a, st));
  // sponge_th, called after the scalar's explicit terms (:635)
  if (th && p.sponge_type) {
    SpongeArgs g;
    for (int f = 0; f < 3; ++f) {
      g.cf[f] = (double2*)s->CF[f];
      g.cu[f] = (double2*)s->U[f];
    }
    g.cfth = (double2*)s->CFTH;
    g.cth = (const double2*)s->TH;
    g.sigma = s->sponge;
    g.sigma_out = s->sponge_out;
    g.ss = s->ss;
    g.r = s->rv;
    g.j0 = s->sponge_out ? s->rv.js : std::max(s->rv.js, p.ny_s_t);
    g.ny_sponge = p.ny_s_t > p.ny ? 2 * p.ny / 3 : p.ny_s_t;
    g.t1 = p.u_bc_c1[1][EKM_YMAX];
    g.t2 = p.u_bc_c1[0][EKM_YMAX];
    g.t3 = p.wbulk0;
    g.zmax6 = p.u_bc[0][EKM_ZMAX] == 6 && p.u_bc[1][EKM_ZMAX] == 6 && p.u_bc[2][EKM_ZMAX] == 6 && p.ri_tau != 0.0;
    g.zero_cu2 = q.zero_cu2;
    g.jst = s->sponge_jst;
    TRY(launch_sponge(d, s->M, g, st));
  }
  // the subgrid model (:641-657), which leaves the velocities in physical
  // space, or the velocities to physical space (:658-671)
  if (q.les_on) {
    TRY(les_chan_run(s->fft, d, s->M, s->rv, s->kx, s->U, s->CF, s->les, s->lesw, q.refresh, st));
  } else {
    for (int f = 0; f < 3; ++f) TRY(inv(s, s->U[f]));
  }
  // the scalar's subgrid model (:650-654), which leaves the scalar in
  // physical space, or the scalar to physical space
  if (th && q.les_on) {
    TRY(les_chan_th_run(s->fft, d, s->M, s->rv, s->rt, s->kx, s->thbar, s->U, s->TH, s->CFTH, s->les, s->lesw,
                        q.refresh_th, st));
  } else if (th) {
    TRY(inv(s, s->TH));
  }
  // momentum fluxes (:683-1161)
  TRY(launch_contravariant(d, s->M, s->rv, s->U[1], s->U[2], s->U2b, s->U3b, st));
  // (with the model: its eddy-viscosity cross terms, :990-1120)
  const double* temp5 = SS_PTR(s, temp5);
  for (int f = 0; f < 3; ++f) {
    TRY(launch_flux_accum(s->fft, d, s->M, s->rv, s->U[0], s->U[f], s->U2b, s->U3b, p.nu, s->kx, s->CF[f],
                          s->R[f], temp5, st, q.les_on ? s->les.nut : nullptr,
                          q.les_on ? les_metric(s, f, 2) : nullptr));
  }
  for (int f = 0; f < 3; ++f) TRY(inv(s, s->R[f]));
  if (th) {  // (:1169-1625)
    TRY(launch_flux_accum(s->fft, d, s->M, s->rt, s->U[0], s->TH, s->U2b, s->U3b, p.nu / p.pr, s->kx, s->CFTH,
                          s->RTH, temp5, st, q.les_on ? s->les.kappa : nullptr, q.les_on ? s->M.g12 : nullptr));
    TRY(inv(s, s->RTH));
  }
  // LES_WALL_MODEL (:1687-1691) from the velocities entering the implicit
  // solves, in physical space here as there
  if (q.wall_on) {
    WallArgs w;
    w.u1 = s->U[0];
    w.u3 = s->U[2];
    w.wall = s->wall;
    w.r = s->rv;
    w.z = s->wall_z;
    w.nu = p.nu;
    // the stress angle alpha = 20.6 is an implicitly typed default real there
    w.cos_a = (double)std::cos(20.6f * 3.1415f / 180.0f);
    w.sin_a = (double)std::sin(20.6f * 3.1415f / 180.0f);
    w.first = q.wall_first;
    TRY(launch_wall_model(d, s->M, w, st));
  }
  // ADI of the scalar (:1334-1625) and the velocities (:1640-2391): the four
  // implicit solves are independent, so they share one set of launches
  {
    const int corners = th ? ((p.th_bc[EKM_ZMAX] == 1 ? 1 : 0) | (p.th_bc[EKM_ZMIN] == 6 ? 2 : 0)) : 0;
    AdiT1 t1{s->ss, {1.0, 1.0, 1.0, p.pr}};
    TRY(launch_adi_batch(d, s->M, q.les_on ? s->adi_dl : s->adi_d, q.les_on ? s->adi_hl : s->adi_h, s->adi_nf, t1,
                         3, corners, st, s->tn));
  }
  // CFL time step at the third substep (:2417-2419)
  if (q.cfl_now) {
    TRY(launch_cfl(d, s->M, s->rv, s->U[0], s->U[1], s->U[2], p.cfl, p.lx / (double)d.NX,
                   reinterpret_cast<unsigned long long*>(s->red + 2), st));
    TRY(launch_dt_update(s->ss, st));
  }
  // back to Fourier space and project (:2421-2455)
  for (int f = 0; f < 3; ++f) TRY(fwd(s, s->U[f]));
  if (th) TRY(fwd(s, s->TH));
  TRY(rem_div(s, q.guess_on ? SS_PTR(s, guess) : nullptr, timed, q.guess_on == 2));
  // pressure update (:2460-2466)
  TRY(launch_pupdate(d, s->rv, s->PHI, s->P, s->ss, st));
  TRY(launch_step_end(s->ss, rk, st));
  return EKM_OK;
}

// the host counters after a substep (src/diablo.F90:152-157)
static void advance_counters(ekm_solver* s, const SubstepPlan& q) {
  s->have_prev_phi = 1;
  if (q.rk == 3) {
    s->time_step = q.cur_step;
    s->rk_step = 1;
    s->first_time = 0;
  } else {
    s->rk_step = q.rk + 1;
  }
}

// One substep, launched eagerly.  Returns launch errors, or the sticky status
// of the run as last read back (the status of this substep's projection is
// formed on the device and read by ekm_get_diag, ekm_time_step, ekm_sync).
extern "C" int ekm_rk_duct(ekm_solver* s) {
  const SubstepPlan q = plan_substep(s);
  TRY(enqueue_substep(s, q, s->prof != 0));
  advance_counters(s, q);
  return s->status_seen;
}

// One substep replayed from a HIP graph: the substep's variant is captured on
// first use (the kernels and their order are those of ekm_rk_duct) and
// launched as one graph after that.
extern "C" int ekm_rk_duct_graph_launch(ekm_solver* s) {
  const SubstepPlan q = plan_substep(s);
  TRY(restore_kx2(s));
  auto it = s->graphs.find(q.key());
  if (it == s->graphs.end()) {
    EKM_CHECK(hipStreamBeginCapture(s->st, hipStreamCaptureModeRelaxed));
    const int e = enqueue_substep(s, q, false);
    hipGraph_t g = nullptr;
    const hipError_t ce = hipStreamEndCapture(s->st, &g);
    if (e != EKM_OK || ce != hipSuccess) {
      if (g) (void)hipGraphDestroy(g);
      if (e != EKM_OK) return e;
      EKM_CHECK(ce);
    }
    hipGraphExec_t ex = nullptr;
    const hipError_t ie = hipGraphInstantiate(&ex, g, nullptr, nullptr, 0);
    (void)hipGraphDestroy(g);
    EKM_CHECK(ie);
    it = s->graphs.emplace(q.key(), ex).first;
  }
  EKM_CHECK(hipGraphLaunch(it->second, s->st));
  advance_counters(s, q);
  return s->status_seen;
}

// n full time steps (three substeps each); graph = 1 replays captured
// substeps.  Waits for the device and returns the run's status.
extern "C" int ekm_time_step_ex(ekm_solver* s, int nsteps, int graph) {
  for (int n = 0; n < nsteps; ++n)
    for (int q = 0; q < 3; ++q) TRY(graph ? ekm_rk_duct_graph_launch(s) : ekm_rk_duct(s));
  StepState h;
  TRY(ss_get(s, &h));
  return h.status;
}

extern "C" int ekm_time_step(ekm_solver* s, int nsteps) { return ekm_time_step_ex(s, nsteps, 0); }

extern "C" int ekm_set_warm_start_net(ekm_solver* s, ekm_wsnet* net) {
  if (net) {
    int nz2, ny2, bmax;
    wsnet_dims(net, &nz2, &ny2, &bmax);
    if (nz2 != s->d.NZ2 || ny2 != s->d.NY2 || bmax < 2 * (s->d.NKX + 1)) return EKM_ERR_ARG;
  }
  EKM_CHECK(hipStreamSynchronize(s->st));
  // captured substeps with the learned start launch the previous network
  for (auto& g : s->graphs) (void)hipGraphExecDestroy(g.second);
  s->graphs.clear();
  s->net = net;
  return EKM_OK;
}

extern "C" int ekm_solve_planes(ekm_solver* s, const double** rhs, const double** phi, const double** kx2,
                                const double** ratio, int* nsolve) {
  if (ratio) *ratio = SS_PTR(s, prev_ratio);
  if (rhs) *rhs = s->mg.rhs;
  if (phi) *phi = s->mg.phi;
  if (kx2) *kx2 = s->kx2s;
  if (nsolve) *nsolve = 2 * (s->d.NKX + 1);
  return EKM_OK;
}

extern "C" int ekm_get_solve_cycles(ekm_solver* s, int* cycles, int n) {
  const int ns = (int)s->cyc_host.size();
  if (!cycles || n < ns) return EKM_ERR_ARG;
  EKM_CHECK(hipMemcpyAsync(s->cyc_host.data(), s->mg.cycles, ns * sizeof(int), hipMemcpyDeviceToHost, s->st));
  EKM_CHECK(hipStreamSynchronize(s->st));
  for (int q = 0; q < ns; ++q) cycles[q] = s->cyc_host[q];
  return EKM_OK;
}

extern "C" int ekm_get_diag(ekm_solver* s, ekm_diag* dg) {
  StepState h;
  TRY(ss_get(s, &h));
  dg->div_ref = h.div_ref;
  dg->div_max = h.div_max;
  dg->delta_t = h.delta_t;
  dg->time = h.time;
  dg->time_step = s->time_step;
  dg->rk_step = s->rk_step;
  dg->mg_solves = h.mg_solves;
  dg->mg_cycles = h.mg_cycles;
  dg->mg_max_cycles = h.mg_max_cycles;
  dg->status = h.status;
  dg->mg_unconverged = h.mg_unconverged;
  return EKM_OK;
}

extern "C" int ekm_multigrid_mudpack_ex(ekm_solver* s, int nsolve, const double* kx2, const double* rhs_tmp,
                                        double* phi, int* cycles, int iguess) {
  if (nsolve < 1 || nsolve > s->mg.nsolve_max || iguess < 0 || iguess > 1) return EKM_ERR_ARG;
  s->have_prev_phi = 0;
  size_t pl = s->d.plane;
  s->kx2_dirty = 1;
  EKM_CHECK(hipMemcpyAsync(s->mg.kx2, kx2, nsolve * sizeof(double), hipMemcpyHostToDevice, s->st));
  EKM_CHECK(hipMemcpyAsync(s->mg.rhs, rhs_tmp, nsolve * pl * sizeof(double), hipMemcpyHostToDevice, s->st));
  // iguess = 1: phi holds the initial guesses (src/mud2cr.f:93,238)
  if (iguess) EKM_CHECK(hipMemcpyAsync(s->mg.phi, phi, nsolve * pl * sizeof(double), hipMemcpyHostToDevice, s->st));
  TRY(mg_solve_ex(s->mg, s->d, nsolve, iguess, s->st));
  EKM_CHECK(hipMemcpyAsync(phi, s->mg.phi, nsolve * pl * sizeof(double), hipMemcpyDeviceToHost, s->st));
  if (cycles) EKM_CHECK(hipMemcpyAsync(cycles, s->mg.cycles, nsolve * sizeof(int), hipMemcpyDeviceToHost, s->st));
  EKM_CHECK(hipStreamSynchronize(s->st));
  return EKM_OK;
}

extern "C" int ekm_multigrid_mudpack(ekm_solver* s, int nsolve, const double* kx2, const double* rhs_tmp, double* phi,
                                     int* cycles) {
  return ekm_multigrid_mudpack_ex(s, nsolve, kx2, rhs_tmp, phi, cycles, 0);
}

// ---- component drop-ins -----------------------------------------------------
static int fft_host(double* varp, int nx, int npencil, int inverse) {
  if (nx < 8 || npencil < 0) return EKM_ERR_ARG;
  size_t n = (size_t)(nx + 2) * npencil * sizeof(double);
  double* dv = nullptr;
  EKM_CHECK(hipMalloc(&dv, n));
  int st = EKM_OK;
  if (hipMemcpy(dv, varp, n, hipMemcpyHostToDevice) != hipSuccess) st = EKM_ERR_HIP;
  if (st == EKM_OK) st = inverse ? ekm_fft_x_to_physical_dev(dv, nx, npencil, nullptr) : ekm_fft_x_to_fourier_dev(dv, nx, npencil, nullptr);
  if (st == EKM_OK && hipMemcpy(varp, dv, n, hipMemcpyDeviceToHost) != hipSuccess) st = EKM_ERR_HIP;
  (void)hipFree(dv);
  return st;
}
extern "C" int ekm_fft_x_to_fourier(double* varp, int nx, int npencil) { return fft_host(varp, nx, npencil, 0); }
extern "C" int ekm_fft_x_to_physical(double* varp, int nx, int npencil) { return fft_host(varp, nx, npencil, 1); }

static int thomas_dev(const double* a, const double* b, const double* c, double* g, int ny, int nx, int ksp,
                      double dd, void* stream) {
  if (nx < 0 || ny < 1 || ksp >= ny) return EKM_ERR_ARG;
  double* w = nullptr;
  size_t n = (size_t)(nx + 1) * (ny + 1);
  EKM_CHECK(hipMalloc(&w, n * sizeof(double)));
  int st = launch_thomas_generic(a, b, c, g, w, ny, nx, (hipStream_t)stream, ksp, dd);
  if (hipStreamSynchronize((hipStream_t)stream) != hipSuccess) st = EKM_ERR_HIP;
  (void)hipFree(w);
  return st;
}

extern "C" int ekm_thomas_real_dev(const double* a, const double* b, const double* c, double* g, int ny, int nx,
                                   void* stream) {
  return thomas_dev(a, b, c, g, ny, nx, -1, 0.0, stream);
}

extern "C" int ekm_thomas_real_sp_dev(const double* a, const double* b, const double* c, double* g, double d, int k,
                                      int ny, int nx, void* stream) {
  if (k < 0) return EKM_ERR_ARG;
  return thomas_dev(a, b, c, g, ny, nx, k, d, stream);
}

static int thomas_host(double* a, double* b, double* c, double* g, int ny, int nx) {
  if (nx < 0 || ny < 1) return EKM_ERR_ARG;
  size_t n = (size_t)(nx + 1) * (ny + 1) * sizeof(double);
  double* dv = nullptr;
  EKM_CHECK(hipMalloc(&dv, 4 * n));
  double *da = dv, *db = (double*)((char*)dv + n), *dc = (double*)((char*)dv + 2 * n), *dg = (double*)((char*)dv + 3 * n);
  int st = EKM_OK;
  if (hipMemcpy(da, a, n, hipMemcpyHostToDevice) != hipSuccess || hipMemcpy(db, b, n, hipMemcpyHostToDevice) != hipSuccess ||
      hipMemcpy(dc, c, n, hipMemcpyHostToDevice) != hipSuccess || hipMemcpy(dg, g, n, hipMemcpyHostToDevice) != hipSuccess)
    st = EKM_ERR_HIP;
  if (st == EKM_OK) st = ekm_thomas_real_dev(da, db, dc, dg, ny, nx, nullptr);
  if (st == EKM_OK && hipMemcpy(g, dg, n, hipMemcpyDeviceToHost) != hipSuccess) st = EKM_ERR_HIP;
  (void)hipFree(dv);
  return st;
}

extern "C" int ekm_thomas_real(double* a, double* b, double* c, double* g, int ny, int nx) {
  return thomas_host(a, b, c, g, ny, nx);
}

// THOMAS_REAL_SP (src/boundary.F90:4-47): the extra elimination of row K+1
// by row K is applied on the host copy, then the standard sweep runs on the GPU.
extern "C" int ekm_thomas_real_sp(double* a, double* b, double* c, double* g, double dd, int k, int ny, int nx) {
  if (k < 0 || k >= ny) return EKM_ERR_ARG;
  size_t ld = (size_t)nx + 1;
  std::vector<double> a2(a, a + ld * (ny + 1)), b2(b, b + ld * (ny + 1)), g2(g, g + ld * (ny + 1));
  for (int i = 0; i <= nx; ++i) {
    size_t p = i + ld * k, q = p + ld;
    a2[q] = a2[q] - b2[p] * dd / a2[p];
    b2[q] = b2[q] - c[p] * dd / a2[p];
    g2[q] = g2[q] - g2[p] * dd / a2[p];
  }
  int st = thomas_host(a2.data(), b2.data(), c, g2.data(), ny, nx);
  if (st == EKM_OK) std::memcpy(g, g2.data(), g2.size() * sizeof(double));
  return st;
}

// ---- measurement hooks ----------------------------------------------------------
extern "C" int ekm_prof_enable(ekm_solver* s, int on) {
  TRY(fold_events(s));
  EKM_CHECK(hipStreamSynchronize(s->st));
  const long long zero[3] = {0, 0, 0};
  TRY(ss_put(s, offsetof(StepState, prof_launches), zero, sizeof(zero)));
  s->prof = on;
  s->prof_ms = 0.0;
  s->prof_timed = 0;
  return EKM_OK;
}

// total milliseconds of the timed (eager) multigrid launches since
// ekm_prof_enable and their count, and the real solves and cycles of every
// projection since then
extern "C" int ekm_prof_get(ekm_solver* s, double* ms, long* launches, long* solves, long* cycles) {
  TRY(fold_events(s));
  StepState h;
  TRY(ss_get(s, &h));
  *ms = s->prof_ms;
  *launches = s->prof ? s->prof_timed : (long)h.prof_launches;
  *solves = (long)h.prof_solves;
  *cycles = (long)h.prof_cycles;
  return EKM_OK;
}

namespace {
__global__ void k_copy(double* __restrict__ dst, const double* __restrict__ src, size_t n) {
  size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x, stride = (size_t)gridDim.x * blockDim.x;
  for (; t < n; t += stride) dst[t] = src[t];
}
}  // namespace

// plain device copy of n doubles: the calibration kernel for HBM byte counters
extern "C" int ekm_copy_dev(double* dst, const double* src, size_t n, void* stream) {
  hipLaunchKernelGGL(k_copy, dim3(4096), dim3(256), 0, (hipStream_t)stream, dst, src, n);
  EKM_CHECK(hipGetLastError());
  EKM_CHECK(hipStreamSynchronize((hipStream_t)stream));
  return EKM_OK;
}

// allocate / free device memory for callers without a device allocator
extern "C" int ekm_dev_alloc(void** p, size_t bytes) {
  EKM_CHECK(hipMalloc(p, bytes));
  EKM_CHECK(hipMemset(*p, 0, bytes));
  return EKM_OK;
}
extern "C" int ekm_dev_free(void* p) {
  EKM_CHECK(hipFree(p));
  return EKM_OK;
}
// dst row r = src row idx[r] (rows of row_bytes, device memory; idx on the host)
extern "C" int ekm_gather_rows(void* dst, const void* src, const long* idx, int n, size_t row_bytes, void* stream) {
  hipStream_t st = (hipStream_t)stream;
  for (int r = 0; r < n; ++r)
    EKM_CHECK(hipMemcpyAsync((char*)dst + (size_t)r * row_bytes, (const char*)src + (size_t)idx[r] * row_bytes,
                             row_bytes, hipMemcpyDeviceToDevice, st));
  EKM_CHECK(hipStreamSynchronize(st));
  return EKM_OK;
}
// name and architecture of the current device, e.g. "AMD Instinct MI355X (gfx950:sramecc+:xnack-)"
extern "C" int ekm_device_name(char* buf, int n) {
  if (!buf || n < 1) return EKM_ERR_ARG;
  int dev = 0;
  EKM_CHECK(hipGetDevice(&dev));
  hipDeviceProp_t pr;
  EKM_CHECK(hipGetDeviceProperties(&pr, dev));
  std::snprintf(buf, (size_t)n, "%s (%s)", pr.name, pr.gcnArchName);
  return EKM_OK;
}

extern "C" int ekm_memcpy_h2d(void* dst, const void* src, size_t bytes) {
  EKM_CHECK(hipMemcpy(dst, src, bytes, hipMemcpyHostToDevice));
  return EKM_OK;
}
extern "C" int ekm_memcpy_d2h(void* dst, const void* src, size_t bytes) {
  EKM_CHECK(hipDeviceSynchronize());
  EKM_CHECK(hipMemcpy(dst, src, bytes, hipMemcpyDeviceToHost));
  return EKM_OK;
}
