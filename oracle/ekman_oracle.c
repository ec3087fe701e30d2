/*
 * ekman_oracle.c -- CPU FP64 oracle of the DIABLO duct RK3 substep.
 *
 * TEST INFRASTRUCTURE ONLY (see ekman_oracle.h for the pinning list).  A
 * plain-C restatement of the reference algorithm; each function cites the
 * reference file:line it follows.  Loops keep the reference's index ranges
 * and operation order so that the GPU path can be compared against it at
 * round-off level.  Built with -ffp-contract=off (no fused multiply-add).
 */
#include "ekman_oracle.h"

#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

/* ------------------------------------------------------------------------ */
/* context                                                                   */
/* ------------------------------------------------------------------------ */

typedef struct mg_level {
  int nx, ny;
  double* phi; /* (0:nx+1, 0:ny+1) */
  double* cf;  /* cf(nx,ny,10), Fortran order */
} mg_level;

typedef struct mg_ctx {
  int nxa, nxb, nyc, nyd;
  int ixp, jyq, iex, jey, ngrid, nfx, nfy, maxcy;
  double xa, xb, yc, yd, tolmax, relmax;
  int kcycle, iprer, ipost, intpol, itero;
  mg_level lev[51]; /* 1..ngrid */
  double* resf;     /* nfx*nfy scratch (wk(kps) in the reference) */
  double* phcor;    /* (nfx+2)*(nfy+2) scratch */
  const ora* o;
  double kx2;
} mg_ctx;

struct ora {
  ekm_params p;
  int NX, NY, NZ, NXC, NKX, NX2, NY2, NZ2, NZ3, NY3;
  size_t npts; /* doubles per 3D field */
  /* grid and metrics (JACOBI_TRANS) */
  double *xp, *yp;               /* (0:NZ+1,0:NY+1) */
  double* met[13];               /* (0:NZ+2,0:NY+2,1:2): G11 G12 G22 C11 C12 C21 C22, then the
                                  * LES variants G11_z G12_z G22_z G11_y G12_y G22_y */
  double* jac;                   /* INT_JACOB (0:NZ+2,0:NY+2) */
  double* thbar;                 /* TH_BAR (0:NZ+1,0:NY+1) */
  double *kx, *kx2;              /* KXP, KX2P (0:NX/2) */
  /* fields: Fourier between substeps, in place with the physical pencil */
  double *U[3], *P, *TH;
  double *CF[3], *CFTH, *R[3], *RTH;
  double *S1, *S2, *U2b, *U3b, *CS1, *PHI;
  int jstart, jend, zstart, zend, jstart_th, jend_th, zstart_th, zend_th;
  double time, delta_t, delta_t_set;
  double h_bar[3], beta_bar[3], zeta_bar[3];
  int rk_step, time_step;
  ekm_diag diag;
  /* LES (src/les_chan.F90 model 2): strain / stress Sij (physical, 6 P
   * fields), |S|, filtered velocity, eddy viscosity, filter width and the
   * dynamic coefficient; CREATE_NEW_FLOW and FIRST_TIME of the reference */
  double *Sij[6], *SABS, *UBT[3], *NUT, *DELTA, *CDYN;
  /* scalar model (src/les_chan_th.F90 model 2): the test-filtered |S| kept
   * from the velocity model's refresh (S_2BAR), the eddy diffusivity
   * KAPPA_T and the dynamic coefficient C_DYN_TH */
  double *S2B, *KAPT, *CDYNT;
  int create_new_flow, first_time;
  /* F_TYPE 0 flow-rate controller (src/duct.F90:414-436): the running mean
   * pressure gradient PX0 and the previous flow rate Qn, set at start-up as
   * Q0 = Qn = WBULK0 (src/diablo.F90:359-361) */
  double px0, qn;
  /* SPONGE_SIGMA (0:NY+1) and SPONGE_SIGMA_OUT (0:NZ+1) of sponge_th */
  double *sigma, *sigma_out;
  /* LES_WALL_MODEL (W_BC_YMIN = 3): U_BC_LOWER_WALLMODEL and
   * W_BC_LOWER_WALLMODEL as [2][NZ2][NX2], DYY(2,2), and U_1, W_1, U_STAR,
   * TAO_WALL, TAO_WALL_1 of the last evaluation */
  double* wall;
  double wall_z, wall_info[5];
};

#define IDX(o, i, k, j) ((size_t)(i) + (size_t)(o)->NX2 * ((size_t)(k) + (size_t)(o)->NZ2 * (size_t)(j)))
#define CRE(o, m, k, j) (2 * (size_t)(m) + (size_t)(o)->NX2 * ((size_t)(k) + (size_t)(o)->NZ2 * (size_t)(j)))
#define MI(o, k, j) ((size_t)(k) + (size_t)(o)->NZ3 * (size_t)(j))
#define MF(o, k, j, f) (MI(o, k, j) + (size_t)(o)->NZ3 * (size_t)(o)->NY3 * (size_t)((f) - 1))
#define PL(o, k, j) ((size_t)(k) + (size_t)(o)->NZ2 * (size_t)(j))

#define G11(o, k, j, f) ((o)->met[0][MF(o, k, j, f)])
#define G12(o, k, j, f) ((o)->met[1][MF(o, k, j, f)])
#define G22(o, k, j, f) ((o)->met[2][MF(o, k, j, f)])
#define C11(o, k, j, f) ((o)->met[3][MF(o, k, j, f)])
#define C12(o, k, j, f) ((o)->met[4][MF(o, k, j, f)])
#define C21(o, k, j, f) ((o)->met[5][MF(o, k, j, f)])
#define C22(o, k, j, f) ((o)->met[6][MF(o, k, j, f)])
#define JAC(o, k, j) ((o)->jac[MI(o, k, j)])
#define NMET 13

static double* dalloc(size_t n) {
  double* a = (double*)malloc(n * sizeof(double));
  if (a) memset(a, 0, n * sizeof(double));
  return a;
}

/* ------------------------------------------------------------------------ */
/* x-FFT (the missing fft.F90; conventions from the call sites)              */
/* ------------------------------------------------------------------------ */
/*
 * Forward: c_m = (1/NX) sum_i u_i exp(-2 pi i m i / NX), m = 0..NX/2, modes
 * NKX+1..NX/2 set to zero (src/mpi_duct.F90:290-294).  Inverse: u_i =
 * Re c_0 + 2 Re sum_{m=1}^{NKX} c_m exp(+2 pi i m i / NX), padding I = NX,
 * NX+1 zeroed (src/mpi_duct.F90:303-313).  A length-NX/2 complex radix-2 FFT
 * plus the usual even/odd split.
 */
typedef struct fft_plan {
  int n, m, nkx;
  double *wr, *wi;   /* exp(-2 pi i k / m), k < m/2 */
  double *tr, *ti;   /* exp(-2 pi i k / n), k <= m */
  int* rev;
} fft_plan;

static fft_plan* plan_make(int n) {
  fft_plan* pl = (fft_plan*)calloc(1, sizeof(fft_plan));
  int m = n / 2, lg = 0;
  pl->n = n;
  pl->m = m;
  pl->nkx = n / 3;
  while ((1 << lg) < m) lg++;
  pl->wr = dalloc(m);
  pl->wi = dalloc(m);
  pl->tr = dalloc(m + 1);
  pl->ti = dalloc(m + 1);
  pl->rev = (int*)malloc(sizeof(int) * m);
  for (int k = 0; k < m; ++k) {
    pl->wr[k] = cos(2.0 * M_PI * k / m);
    pl->wi[k] = -sin(2.0 * M_PI * k / m);
    int r = 0;
    for (int b = 0; b < lg; ++b)
      if (k & (1 << b)) r |= 1 << (lg - 1 - b);
    pl->rev[k] = r;
  }
  for (int k = 0; k <= m; ++k) {
    pl->tr[k] = cos(2.0 * M_PI * k / n);
    pl->ti[k] = -sin(2.0 * M_PI * k / n);
  }
  return pl;
}

static void plan_free(fft_plan* pl) {
  free(pl->wr); free(pl->wi); free(pl->tr); free(pl->ti); free(pl->rev); free(pl);
}

/* in-place complex FFT of length m (sign -1 if inv == 0, +1 otherwise), unscaled */
static void cfft(const fft_plan* pl, double* re, double* im, int inv) {
  int m = pl->m;
  for (int k = 0; k < m; ++k) {
    int r = pl->rev[k];
    if (r > k) {
      double t = re[k]; re[k] = re[r]; re[r] = t;
      t = im[k]; im[k] = im[r]; im[r] = t;
    }
  }
  for (int len = 2; len <= m; len <<= 1) {
    int half = len >> 1, step = m / len;
    for (int s = 0; s < m; s += len)
      for (int q = 0; q < half; ++q) {
        double wr = pl->wr[q * step], wi = inv ? -pl->wi[q * step] : pl->wi[q * step];
        int a = s + q, b = a + half;
        double xr = re[b] * wr - im[b] * wi;
        double xi = re[b] * wi + im[b] * wr;
        re[b] = re[a] - xr; im[b] = im[a] - xi;
        re[a] = re[a] + xr; im[a] = im[a] + xi;
      }
  }
}

static void fft_fwd_pencil(const fft_plan* pl, double* v, double* re, double* im) {
  int n = pl->n, m = pl->m;
  for (int q = 0; q < m; ++q) { re[q] = v[2 * q]; im[q] = v[2 * q + 1]; }
  cfft(pl, re, im, 0);
  double sc = 1.0 / (double)n;
  for (int k = 0; k <= m; ++k) {
    int a = k % m, b = (m - k) % m;
    double zr = re[a], zi = im[a], yr = re[b], yi = -im[b]; /* conj(Z_{m-k}) */
    double er = 0.5 * (zr + yr), ei = 0.5 * (zi + yi);
    double orr = 0.5 * (zi - yi), oi = -0.5 * (zr - yr); /* (Z - conj)/(2i) */
    double xr = er + (pl->tr[k] * orr - pl->ti[k] * oi);
    double xi = ei + (pl->tr[k] * oi + pl->ti[k] * orr);
    if (k <= pl->nkx) { v[2 * k] = xr * sc; v[2 * k + 1] = xi * sc; }
    else { v[2 * k] = 0.0; v[2 * k + 1] = 0.0; }
  }
  (void)n;
}

static void fft_inv_pencil(const fft_plan* pl, double* v, double* re, double* im) {
  int m = pl->m, nkx = pl->nkx;
  /* X_k for k=0..m with X_k = 0 above NKX and Im X_0 ignored */
  for (int k = 0; k < m; ++k) {
    int b = m - k;
    double xr = (k <= nkx) ? v[2 * k] : 0.0, xi = (k <= nkx && k > 0) ? v[2 * k + 1] : 0.0;
    double yr = (b <= nkx) ? v[2 * b] : 0.0, yi = (b <= nkx && b > 0) ? -v[2 * b + 1] : 0.0;
    if (b == m) { yr = 0.0; yi = 0.0; } /* Nyquist: zero after the 2/3 rule */
    double er = xr + yr, ei = xi + yi;           /* X_k + conj(X_{m-k}) */
    double dr = xr - yr, di = xi - yi;           /* X_k - conj(X_{m-k}) */
    double cr = pl->tr[k], ci = -pl->ti[k];      /* exp(+2 pi i k / n) */
    double orr = dr * cr - di * ci, oi = dr * ci + di * cr;
    re[k] = er - oi; /* E + i O */
    im[k] = ei + orr;
  }
  cfft(pl, re, im, 1);
  for (int q = 0; q < m; ++q) { v[2 * q] = re[q]; v[2 * q + 1] = im[q]; }
  v[2 * m] = 0.0;
  v[2 * m + 1] = 0.0;
}

static void fft_many(double* varp, int nx, long npencil, int inv) {
  fft_plan* pl = plan_make(nx);
#pragma omp parallel
  {
    double* re = dalloc(nx / 2 + 1);
    double* im = dalloc(nx / 2 + 1);
#pragma omp for schedule(static)
    for (long p = 0; p < npencil; ++p) {
      double* v = varp + (size_t)p * (size_t)(nx + 2);
      if (inv) fft_inv_pencil(pl, v, re, im);
      else fft_fwd_pencil(pl, v, re, im);
    }
    free(re);
    free(im);
  }
  plan_free(pl);
}

void ora_fft_fwd(double* varp, int nx, long npencil) { fft_many(varp, nx, npencil, 0); }
void ora_fft_inv(double* varp, int nx, long npencil) { fft_many(varp, nx, npencil, 1); }

static void fwd(ora* o, double* f) { ora_fft_fwd(f, o->NX, (long)o->NZ2 * o->NY2); }
static void inv(ora* o, double* f) { ora_fft_inv(f, o->NX, (long)o->NZ2 * o->NY2); }

/* ------------------------------------------------------------------------ */
/* Thomas solvers (src/boundary.F90:4-84)                                    */
/* ------------------------------------------------------------------------ */
/* A,B,C,G(0:NX,0:NY), batch index first.  A and B are clobbered as in the
 * reference. */
void ora_thomas_real(double* a, double* b, double* c, double* g, int ny, int nx) {
  size_t ld = (size_t)nx + 1;
  for (int j = 0; j < ny; ++j)
    for (int i = 0; i <= nx; ++i) {
      size_t p = i + ld * j, q = p + ld;
      a[q] = -a[q] / b[p];
      b[q] = b[q] + a[q] * c[p];
      g[q] = g[q] + a[q] * g[p];
    }
  for (int i = 0; i <= nx; ++i) g[i + ld * ny] = g[i + ld * ny] / b[i + ld * ny];
  for (int j = ny - 1; j >= 0; --j)
    for (int i = 0; i <= nx; ++i) {
      size_t p = i + ld * j;
      g[p] = (g[p] - c[p] * g[p + ld]) / b[p];
    }
}

/* src/boundary.F90:4-47: one extra elimination of row K+1 by row K with
 * factor D/A(K) before the standard sweep. */
void ora_thomas_real_sp(double* a, double* b, double* c, double* g, double d, int k,
                        int ny, int nx) {
  size_t ld = (size_t)nx + 1;
  for (int i = 0; i <= nx; ++i) {
    size_t p = i + ld * k, q = p + ld;
    a[q] = a[q] - b[p] * d / a[p];
    b[q] = b[q] - c[p] * d / a[p];
    g[q] = g[q] - g[p] * d / a[p];
  }
  ora_thomas_real(a, b, c, g, ny, nx);
}

/* ------------------------------------------------------------------------ */
/* metrics: JACOBI_TRANS (src/diablo.F90:1088-1376)                          */
/* ------------------------------------------------------------------------ */
static double trunc8(double v) { return trunc(v * 100000000.0) * 1e-8; } /* :1220-1230 */

static void metrics(ora* o) {
  int NZ = o->NZ, NY = o->NY;
  size_t n = (size_t)(NZ + 2) * (NY + 2) * 2; /* (1:NZ+1,1:NY+1,2) stored with 0-based pad */
  double *xz = dalloc(n), *yz = dalloc(n), *xe = dalloc(n), *ye = dalloc(n);
  double* ij = dalloc(n);
#define L(k, j, f) ((size_t)(k) + (size_t)(NZ + 2) * ((size_t)(j) + (size_t)(NY + 2) * (size_t)((f) - 1)))
#define XP(k, j) (o->xp[PL(o, k, j)])
#define YP(k, j) (o->yp[PL(o, k, j)])
  for (size_t q = 0; q < n; ++q) xz[q] = yz[q] = xe[q] = ye[q] = NAN;
  /* :1166-1190 */
  for (int j = 1; j <= NY; ++j)
    for (int k = 1; k <= NZ; ++k) {
      xz[L(k, j, 1)] = XP(k + 1, j) - XP(k, j);
      xz[L(k, j, 2)] = 0.250 * (XP(k + 1, j + 1) + XP(k + 1, j) - XP(k - 1, j + 1) - XP(k - 1, j));
      yz[L(k, j, 1)] = YP(k + 1, j) - YP(k, j);
      yz[L(k, j, 2)] = 0.250 * (YP(k + 1, j + 1) + YP(k + 1, j) - YP(k - 1, j + 1) - YP(k - 1, j));
      xe[L(k, j, 2)] = XP(k, j + 1) - XP(k, j);
      xe[L(k, j, 1)] = 0.250 * (XP(k, j + 1) + XP(k + 1, j + 1) - XP(k, j - 1) - XP(k + 1, j - 1));
      ye[L(k, j, 1)] = 0.250 * (YP(k, j + 1) + YP(k + 1, j + 1) - YP(k, j - 1) - YP(k + 1, j - 1));
      ye[L(k, j, 2)] = YP(k, j + 1) - YP(k, j);
    }
  /* :1193-1217 linear extrapolation, same statement order */
  for (int k = 1; k <= NZ + 1; ++k) {
    xe[L(k, 1, 1)] = 2.0 * xe[L(k, 2, 1)] - xe[L(k, 3, 1)];
    for (int f = 1; f <= 2; ++f) {
      xe[L(k, NY + 1, f)] = 2.0 * xe[L(k, NY, f)] - xe[L(k, NY - 1, f)];
      xz[L(k, NY + 1, f)] = 2.0 * xz[L(k, NY, f)] - xz[L(k, NY - 1, f)];
    }
    ye[L(k, 1, 1)] = 2.0 * ye[L(k, 2, 1)] - ye[L(k, 3, 1)];
    for (int f = 1; f <= 2; ++f) {
      ye[L(k, NY + 1, f)] = 2.0 * ye[L(k, NY, f)] - ye[L(k, NY - 1, f)];
      yz[L(k, NY + 1, f)] = 2.0 * yz[L(k, NY, f)] - yz[L(k, NY - 1, f)];
    }
  }
  for (int j = 1; j <= NY + 1; ++j) {
    xz[L(1, j, 2)] = 2.0 * xz[L(2, j, 2)] - xz[L(3, j, 2)];
    for (int f = 1; f <= 2; ++f) {
      xz[L(NZ + 1, j, f)] = 2.0 * xz[L(NZ, j, f)] - xz[L(NZ - 1, j, f)];
      xe[L(NZ + 1, j, f)] = 2.0 * xe[L(NZ, j, f)] - xe[L(NZ - 1, j, f)];
    }
    yz[L(1, j, 2)] = 2.0 * yz[L(2, j, 2)] - yz[L(3, j, 2)];
    for (int f = 1; f <= 2; ++f) {
      yz[L(NZ + 1, j, f)] = 2.0 * yz[L(NZ, j, f)] - yz[L(NZ - 1, j, f)];
      ye[L(NZ + 1, j, f)] = 2.0 * ye[L(NZ, j, f)] - ye[L(NZ - 1, j, f)];
    }
  }
  /* :1220-1230 truncation to 1e-8 */
  for (int j = 1; j <= NY + 1; ++j)
    for (int k = 1; k <= NZ + 1; ++k)
      for (int f = 1; f <= 2; ++f) {
        xz[L(k, j, f)] = trunc8(xz[L(k, j, f)]);
        xe[L(k, j, f)] = trunc8(xe[L(k, j, f)]);
        yz[L(k, j, f)] = trunc8(yz[L(k, j, f)]);
        ye[L(k, j, f)] = trunc8(ye[L(k, j, f)]);
      }
  /* :1233-1264 */
  for (int q = 0; q < NMET; ++q)
    for (size_t t = 0; t < (size_t)o->NZ3 * o->NY3 * 2; ++t) o->met[q][t] = NAN;
  for (size_t t = 0; t < (size_t)o->NZ3 * o->NY3; ++t) o->jac[t] = NAN;
  for (int j = 1; j <= NY + 1; ++j)
    for (int k = 1; k <= NZ + 1; ++k)
      for (int f = 1; f <= 2; ++f) {
        size_t q = L(k, j, f);
        double J = xz[q] * ye[q] - xe[q] * yz[q];
        ij[q] = J;
        double zx = ye[q] / J, ex = -yz[q] / J, zy = -xe[q] / J, ey = xz[q] / J;
        G11(o, k, j, f) = (zx * zx + zy * zy) * J;
        G12(o, k, j, f) = (zx * ex + zy * ey) * J;
        G22(o, k, j, f) = (ex * ex + ey * ey) * J;
        C11(o, k, j, f) = zx * J;
        C12(o, k, j, f) = zy * J;
        C21(o, k, j, f) = ex * J;
        C22(o, k, j, f) = ey * J;
        /* :1379-1403 LES variants: the stress of U3 doubles the zeta_x and
         * eta_x products, that of U2 the zeta_y and eta_y products */
        o->met[7][MF(o, k, j, f)] = (2.0 * zx * zx + zy * zy) * J;
        o->met[8][MF(o, k, j, f)] = (2.0 * zx * ex + zy * ey) * J;
        o->met[9][MF(o, k, j, f)] = (2.0 * ex * ex + ey * ey) * J;
        o->met[10][MF(o, k, j, f)] = (zx * zx + 2.0 * zy * zy) * J;
        o->met[11][MF(o, k, j, f)] = (zx * ex + 2.0 * zy * ey) * J;
        o->met[12][MF(o, k, j, f)] = (ex * ex + 2.0 * ey * ey) * J;
      }
  /* :1268-1292 cell Jacobian */
  for (int j = 1; j <= NY; ++j)
    for (int k = 1; k <= NZ; ++k)
      JAC(o, k, j) = 0.250 * (ij[L(k, j, 1)] + ij[L(k, j, 2)] + ij[L(k, j + 1, 1)] + ij[L(k + 1, j, 2)]);
  for (int j = 0; j <= NY + 1; ++j) {
    JAC(o, NZ + 1, j) = 2.0 * JAC(o, NZ, j) - JAC(o, NZ - 1, j);
    JAC(o, 0, j) = 2.0 * JAC(o, 1, j) - JAC(o, 2, j);
  }
  for (int k = 0; k <= NZ + 1; ++k) {
    JAC(o, k, NY + 1) = 2.0 * JAC(o, k, NY) - JAC(o, k, NY - 1);
    JAC(o, k, 0) = 2.0 * JAC(o, k, 1) - JAC(o, k, 2);
  }
  JAC(o, 0, 0) = 0.5 * (2.0 * JAC(o, 1, 0) - JAC(o, 2, 0) + 2.0 * JAC(o, 0, 1) - JAC(o, 0, 2));
  JAC(o, NZ + 1, NY + 1) = 0.5 * (2.0 * JAC(o, NZ, NY + 1) - JAC(o, NZ - 1, NY + 1) +
                                  2.0 * JAC(o, NZ + 1, NY) - JAC(o, NZ + 1, NY - 1));
  /* :1294-1340 ghost metrics, same statement order (GMAT_12 corner is not
   * re-set: the reference assigns GMAT_22 twice at :1322-1326); the LES
   * variants follow the same rules (:1405-1474, GMAT_22_z and GMAT_22_y are
   * the ones assigned twice there) */
  for (int j = 0; j <= NY + 1; ++j)
    for (int f = 1; f <= 2; ++f)
      for (int t = 0; t < NMET; ++t) {
        double* a = o->met[t];
        a[MF(o, 0, j, f)] = 2.0 * a[MF(o, 1, j, f)] - a[MF(o, 2, j, f)];
      }
  for (int k = 0; k <= NZ + 1; ++k)
    for (int f = 1; f <= 2; ++f)
      for (int t = 0; t < NMET; ++t) {
        double* a = o->met[t];
        a[MF(o, k, 0, f)] = 2.0 * a[MF(o, k, 1, f)] - a[MF(o, k, 2, f)];
      }
  for (int f = 1; f <= 2; ++f) {
    /* GMAT_11, GMAT_22, CJOB_11, CJOB_12, CJOB_21, CJOB_22, GMAT_11_z,
     * GMAT_22_z, GMAT_11_y, GMAT_22_y */
    int list[10] = {0, 2, 3, 4, 5, 6, 7, 9, 10, 12};
    for (int t = 0; t < 10; ++t) {
      double* a = o->met[list[t]];
      a[MF(o, 0, 0, f)] = 0.5 * (2.0 * a[MF(o, 0, 1, f)] - a[MF(o, 0, 2, f)] +
                                 2.0 * a[MF(o, 1, 0, f)] - a[MF(o, 2, 0, f)]);
    }
  }
  /* :1343-1376 copies for the multigrid coefficient callback */
  for (int f = 1; f <= 2; ++f) {
    for (int q = 0; q < NMET; ++q) {
      double* a = o->met[q];
      for (int k = 0; k <= NZ + 1; ++k) a[MF(o, k, NY + 2, f)] = a[MF(o, k, NY + 1, f)];
      for (int j = 0; j <= NY + 1; ++j) a[MF(o, NZ + 2, j, f)] = a[MF(o, NZ + 1, j, f)];
      a[MF(o, NZ + 2, NY + 2, f)] = a[MF(o, NZ + 1, NY + 1, f)];
    }
  }
#undef L
#undef XP
#undef YP
  free(xz); free(yz); free(xe); free(ye); free(ij);
}

int ora_get_metric(ora* o, int which, double* out) {
  if (which < 0 || which > 13) return EKM_ERR_ARG;
  size_t n = (size_t)o->NZ3 * o->NY3 * (which == 7 ? 1 : 2);
  memcpy(out, which < 7 ? o->met[which] : which == 7 ? o->jac : o->met[which - 1], n * sizeof(double));
  return EKM_OK;
}

/* ------------------------------------------------------------------------ */
/* MULTIGRID_MUDPACK (src/mg_solver2d.F90:75-255) and the mud2cr solver     */
/* (src/mud2cr.f, src/mudcom.f), restated for method=0, kcycle=2, iprer=2,  */
/* ipost=1, intpol=3 (src/mg_solver2d.F90:146-165).                          */
/* ------------------------------------------------------------------------ */

/* Boundary callback.  The reference bndcr (src/mg_solver2d.F90:206-255)
 * implements kbdy=3,4 (y boundaries) as dphi/dy = 0: alfa=0, beta=1, gama=0,
 * gbdy=0.  kbdy=1,2 (mixed at the z ends, which the reference leaves out) are
 * generalised here for free-slip z ends: dphi/dz = 0, alfa=1, beta=0, gama=0,
 * gbdy=0. */
static void bndyc(int kbdy, double xory, double* alfa, double* beta, double* gama,
                  double* gbdy) {
  (void)xory;
  if (kbdy == 1 || kbdy == 2) { *alfa = 1.0; *beta = 0.0; }
  else { *alfa = 0.0; *beta = 1.0; }
  *gama = 0.0;
  *gbdy = 0.0;
}

/* cofcr (src/duct.F90:3900-3934) */
static void cofcr(const ora* o, double kx2, double x, double y, double* cxx, double* cxy,
                  double* cyy, double* cx, double* cy, double* ce) {
  int k = (int)x, j = (int)y;
  *cxx = 0.25 * (G11(o, k, j, 2) + G11(o, k + 1, j, 2) + G11(o, k, j, 1) + G11(o, k, j + 1, 1));
  *cxy = 0.5 * (G12(o, k, j, 2) + G12(o, k + 1, j, 2) + G12(o, k, j, 1) + G12(o, k, j + 1, 1));
  *cyy = 0.25 * (G22(o, k, j, 2) + G22(o, k + 1, j, 2) + G22(o, k, j, 1) + G22(o, k, j + 1, 1));
  *cx = G11(o, k + 1, j, 2) - G11(o, k, j, 2) + G12(o, k, j + 1, 1) - G12(o, k, j, 1);
  *cy = G22(o, k, j + 1, 1) - G22(o, k, j, 1) + G12(o, k + 1, j, 2) - G12(o, k, j, 2);
  *ce = -(kx2 * JAC(o, k, j));
}

#define COF(x, y) cofcr(c->o, c->kx2, (x), (y), &cxx, &cxy, &cyy, &cx, &cy, &ce)
#define CF(i, j, l) cf[((i)-1) + (size_t)nx * (((j)-1) + (size_t)ny * ((l)-1))]
#define PH(i, j) phi[(i) + (size_t)(nx + 2) * (j)]

/* dismd2cr (src/mud2cr.f:501-1270), method 0: no line-relaxation factors */
static void mg_discretize(mg_ctx* c, int lev) {
  mg_level* L = &c->lev[lev];
  int nx = L->nx, ny = L->ny;
  double* cf = L->cf;
  double dlx = (c->xb - c->xa) / (nx - 1), dlx2 = dlx + dlx, dlxx = dlx * dlx;
  double dly = (c->yd - c->yc) / (ny - 1), dly2 = dly + dly, dlyy = dly * dly;
  double dlxy = dlx * dly, dlxy2 = dlxy + dlxy, dlxy4 = dlxy2 + dlxy2;
  double dyox = dly / dlx, dxoy = dlx / dly;
  double cxx, cxy, cyy, cx, cy, ce, c1, c2, c3, c4, c5, c6, c7, c8;
  double x, y;
  /* interior and nonspecified boundaries (:560-590) */
  for (int j = 1; j <= ny; ++j) {
    y = c->yc + (j - 1) * dly;
    for (int i = 1; i <= nx; ++i) {
      x = c->xa + (i - 1) * dlx;
      COF(x, y);
      cxx = fmax(cxx, fabs(cx) * dlx * 0.5);
      cyy = fmax(cyy, fabs(cy) * dly * 0.5);
      c1 = cxx / dlxx + cx / dlx2;
      c2 = cxy / dlxy4;
      c3 = cyy / dlyy + cy / dly2;
      c4 = -c2;
      c5 = cxx / dlxx - cx / dlx2;
      c6 = c2;
      c7 = cyy / dlyy - cy / dly2;
      c8 = -c2;
      CF(i, j, 1) = c1; CF(i, j, 2) = c2; CF(i, j, 3) = c3; CF(i, j, 4) = c4;
      CF(i, j, 5) = c5; CF(i, j, 6) = c6; CF(i, j, 7) = c7; CF(i, j, 8) = c8;
      CF(i, j, 9) = ce - (c1 + c3 + c5 + c7);
    }
  }
  /* corners: virtual-virtual points not used (:594-637) */
  COF(c->xa, c->yc);
  c4 = -cxy / dlxy2;
  CF(1, 1, 1) -= c4; CF(1, 1, 2) = 0.0; CF(1, 1, 3) -= c4; CF(1, 1, 4) = c4;
  CF(1, 1, 5) -= c4; CF(1, 1, 6) = 0.0; CF(1, 1, 7) -= c4; CF(1, 1, 8) = c4;
  CF(1, 1, 9) += 2.0 * c4;
  COF(c->xb, c->yc);
  c2 = cxy / dlxy2;
  CF(nx, 1, 1) -= c2; CF(nx, 1, 2) = c2; CF(nx, 1, 3) -= c2; CF(nx, 1, 4) = 0.0;
  CF(nx, 1, 5) -= c2; CF(nx, 1, 6) = c2; CF(nx, 1, 7) -= c2; CF(nx, 1, 8) = 0.0;
  CF(nx, 1, 9) += 2.0 * c2;
  COF(c->xa, c->yd);
  c2 = cxy / dlxy2;
  CF(1, ny, 1) -= c2; CF(1, ny, 2) = c2; CF(1, ny, 3) -= c2; CF(1, ny, 4) = 0.0;
  CF(1, ny, 5) -= c2; CF(1, ny, 6) = c2; CF(1, ny, 7) -= c2; CF(1, ny, 8) = 0.0;
  CF(1, ny, 9) += 2.0 * c2;
  COF(c->xb, c->yd);
  c4 = -cxy / dlxy2;
  CF(nx, ny, 1) -= c4; CF(nx, ny, 2) = 0.0; CF(nx, ny, 3) -= c4; CF(nx, ny, 4) = c4;
  CF(nx, ny, 5) -= c4; CF(nx, ny, 6) = 0.0; CF(nx, ny, 7) -= c4; CF(nx, ny, 8) = c4;
  CF(nx, ny, 9) += 2.0 * c4;

  double am1, bm1, gm1, dm1, a0, b0, g0, d0, ap1, bp1, gp1, dp1;
  /* mixed derivative b.c. along x=xa (:639-679) */
  if (c->nxa == 2) {
    bndyc(1, c->yc, &am1, &bm1, &gm1, &dm1);
    bndyc(1, c->yc + dly, &a0, &b0, &g0, &d0);
    for (int j = 2; j <= ny - 1; ++j) {
      y = c->yc + j * dly;
      bndyc(1, y, &ap1, &bp1, &gp1, &dp1);
      c4 = CF(1, j, 4); c5 = CF(1, j, 5); c6 = CF(1, j, 6);
      CF(1, j, 1) = CF(1, j, 1) + c5;
      CF(1, j, 2) = CF(1, j, 2) + c4;
      CF(1, j, 3) = CF(1, j, 3) + c6 * (-bm1 / am1 * dxoy) + c5 * (b0 / a0 * dxoy) +
                    c4 * (3. * bp1 / ap1 * dxoy + dlx2 * gp1 / ap1);
      CF(1, j, 4) = 0.0; CF(1, j, 5) = 0.0; CF(1, j, 6) = 0.0;
      CF(1, j, 7) = CF(1, j, 7) + c6 * (-3. * bm1 / am1 * dxoy + dlx2 * gm1 / am1) +
                    c5 * (-b0 / a0 * dxoy) + c4 * (bp1 / ap1 * dxoy);
      CF(1, j, 8) = CF(1, j, 8) + c6;
      CF(1, j, 9) = CF(1, j, 9) + c6 * (4. * bm1 / am1 * dxoy) + c5 * (dlx2 * g0 / a0) +
                    c4 * (-4. * bp1 / ap1 * dxoy);
      am1 = a0; bm1 = b0; gm1 = g0; dm1 = d0;
      a0 = ap1; b0 = bp1; g0 = gp1; d0 = dp1;
    }
  }
  /* along x=xb (:681-720) */
  if (c->nxb == 2) {
    bndyc(2, c->yc, &am1, &bm1, &gm1, &dm1);
    bndyc(2, c->yc + dly, &a0, &b0, &g0, &d0);
    for (int j = 2; j <= ny - 1; ++j) {
      y = c->yc + j * dly;
      bndyc(2, y, &ap1, &bp1, &gp1, &dp1);
      c1 = CF(nx, j, 1); c2 = CF(nx, j, 2); c8 = CF(nx, j, 8);
      CF(nx, j, 1) = 0.0; CF(nx, j, 2) = 0.0;
      CF(nx, j, 3) = CF(nx, j, 3) - c8 * (-bm1 / am1 * dxoy) - c1 * (b0 / a0 * dxoy) -
                     c2 * (3.0 * bp1 / ap1 * dxoy + dlx2 * gp1 / ap1);
      CF(nx, j, 4) = CF(nx, j, 4) + c2;
      CF(nx, j, 5) = CF(nx, j, 5) + c1;
      CF(nx, j, 6) = CF(nx, j, 6) + c8;
      CF(nx, j, 7) = CF(nx, j, 7) + c8 * (3.0 * bm1 / am1 * dxoy - dlx2 * gm1 / am1) +
                     c1 * (b0 / a0 * dxoy) + c2 * (-bp1 / ap1 * dxoy);
      CF(nx, j, 8) = 0.0;
      CF(nx, j, 9) = CF(nx, j, 9) - c8 * (4.0 * bm1 / am1 * dxoy) - c1 * (dlx2 * g0 / a0) -
                     c2 * (-4.0 * bp1 / ap1 * dxoy);
      am1 = a0; bm1 = b0; gm1 = g0; dm1 = d0;
      a0 = ap1; b0 = bp1; g0 = gp1; d0 = dp1;
    }
  }
  /* along y=yc (:721-759) */
  if (c->nyc == 2) {
    bndyc(3, c->xa, &am1, &bm1, &gm1, &dm1);
    bndyc(3, c->xa + dlx, &a0, &b0, &g0, &d0);
    for (int i = 2; i <= nx - 1; ++i) {
      x = c->xa + i * dlx;
      bndyc(3, x, &ap1, &bp1, &gp1, &dp1);
      c6 = CF(i, 1, 6); c7 = CF(i, 1, 7); c8 = CF(i, 1, 8);
      CF(i, 1, 1) = CF(i, 1, 1) + c6 * (-am1 / bm1 * dyox) + c7 * (a0 / b0 * dyox) +
                    c8 * (3.0 * ap1 / bp1 * dyox + dly2 * gp1 / bp1);
      CF(i, 1, 2) = CF(i, 1, 2) + c8;
      CF(i, 1, 3) = CF(i, 1, 3) + c7;
      CF(i, 1, 4) = CF(i, 1, 4) + c6;
      CF(i, 1, 5) = CF(i, 1, 5) + c6 * (-3.0 * am1 / bm1 * dyox + dly2 * gm1 / bm1) +
                    c7 * (-a0 / b0 * dyox) + c8 * (ap1 / bp1 * dyox);
      CF(i, 1, 6) = 0.0; CF(i, 1, 7) = 0.0; CF(i, 1, 8) = 0.0;
      CF(i, 1, 9) = CF(i, 1, 9) + c6 * (4.0 * am1 / bm1 * dyox) + c7 * dly2 * g0 / b0 +
                    c8 * (-4.0 * ap1 / bp1 * dyox);
      am1 = a0; bm1 = b0; gm1 = g0; dm1 = d0;
      a0 = ap1; b0 = bp1; g0 = gp1; d0 = dp1;
    }
  }
  /* along y=yd (:761-801) */
  if (c->nyd == 2) {
    bndyc(4, c->xa, &am1, &bm1, &gm1, &dm1);
    bndyc(4, c->xa + dlx, &a0, &b0, &g0, &d0);
    for (int i = 2; i <= nx - 1; ++i) {
      x = c->xa + i * dlx;
      bndyc(4, x, &ap1, &bp1, &gp1, &dp1);
      c2 = CF(i, ny, 2); c3 = CF(i, ny, 3); c4 = CF(i, ny, 4);
      CF(i, ny, 1) = CF(i, ny, 1) + c4 * (am1 / bm1 * dyox) + c3 * (-a0 / b0 * dyox) +
                     c2 * (-3. * ap1 / bp1 * dyox - dly2 * gp1 / bp1);
      CF(i, ny, 2) = 0.0; CF(i, ny, 3) = 0.0; CF(i, ny, 4) = 0.0;
      CF(i, ny, 5) = CF(i, ny, 5) + c4 * (3.0 * am1 / bm1 * dyox - dly2 * gm1 / bm1) +
                     c3 * (a0 / b0 * dyox) + c2 * (-ap1 / bp1 * dyox);
      CF(i, ny, 6) = CF(i, ny, 6) + c4;
      CF(i, ny, 7) = CF(i, ny, 7) + c3;
      CF(i, ny, 8) = CF(i, ny, 8) + c2;
      CF(i, ny, 9) = CF(i, ny, 9) + c4 * (-4. * am1 / bm1 * dyox) + c3 * (-dly2 * g0 / b0) +
                     c2 * (4. * ap1 / bp1 * dyox);
      am1 = a0; bm1 = b0; gm1 = g0; dm1 = d0;
      a0 = ap1; b0 = bp1; g0 = gp1; d0 = dp1;
    }
  }
  (void)dm1; (void)d0; (void)dp1;
  double aa, ba, ga, da, ab, bb, gb, db, ac, bc, gc, dc, ad, bd, gd, dd, det;
  /* corner adjustments at y=yc (:817-960) */
  if (c->nyc == 2) {
    if (c->nxa == 0) {
      bndyc(3, c->xa, &a0, &b0, &g0, &d0);
      bndyc(3, c->xa + dlx, &ap1, &bp1, &gp1, &dp1);
      c7 = CF(1, 1, 7); c8 = CF(1, 1, 8);
      CF(1, 1, 1) = CF(1, 1, 1) + c7 * (a0 / b0 * dyox) +
                    c8 * (3.0 * ap1 / bp1 * dyox + dly2 * gp1 / bp1);
      CF(1, 1, 2) = CF(1, 1, 2) + c8;
      CF(1, 1, 3) = CF(1, 1, 3) + c7;
      CF(1, 1, 5) = CF(1, 1, 5) + c7 * (-a0 / b0 * dyox) + c8 * (ap1 / bp1 * dyox);
      CF(1, 1, 6) = 0.0; CF(1, 1, 7) = 0.0; CF(1, 1, 8) = 0.0;
      CF(1, 1, 9) = CF(1, 1, 9) + c8 * (-4.0 * ap1 / bp1 * dyox) + c7 * (dly2 * g0 / b0);
      bndyc(3, c->xb - dlx, &am1, &bm1, &gm1, &dm1);
      bndyc(3, c->xb, &a0, &b0, &g0, &d0);
      c6 = CF(nx, 1, 6); c7 = CF(nx, 1, 7);
      CF(nx, 1, 1) = CF(nx, 1, 1) + c6 * (-am1 / bm1 * dyox) + c7 * (a0 / b0 * dyox);
      CF(nx, 1, 3) = CF(nx, 1, 3) + c7;
      CF(nx, 1, 4) = CF(nx, 1, 4) + c6;
      CF(nx, 1, 5) = CF(nx, 1, 5) + c6 * (-3.0 * am1 / bm1 * dyox + dly2 * gm1 / bm1) +
                     c7 * (-a0 / b0 * dyox);
      CF(nx, 1, 6) = 0.0; CF(nx, 1, 7) = 0.0; CF(nx, 1, 8) = 0.0;
      CF(nx, 1, 9) = CF(nx, 1, 9) + c6 * (4.0 * am1 / bm1 * dyox) + c7 * (dly2 * g0 / b0);
    } else if (c->nxa == 2) {
      bndyc(3, c->xa + dlx, &ac, &bc, &gc, &dc);
      bndyc(1, c->yc + dly, &aa, &ba, &ga, &da);
      c4 = CF(1, 1, 4); c8 = CF(1, 1, 8);
      CF(1, 1, 1) = CF(1, 1, 1) + c8 * (3.0 * ac / bc * dyox + dly2 * gc / bc);
      CF(1, 1, 2) = CF(1, 1, 2) + c8 + c4;
      CF(1, 1, 3) = CF(1, 1, 3) + c4 * (3.0 * ba / aa * dxoy + dlx2 * ga / aa);
      CF(1, 1, 4) = 0.0;
      CF(1, 1, 5) = CF(1, 1, 5) + c8 * (ac / bc * dyox);
      CF(1, 1, 6) = 0.0;
      CF(1, 1, 7) = CF(1, 1, 7) + c4 * (ba / aa * dxoy);
      CF(1, 1, 8) = 0.0;
      CF(1, 1, 9) = CF(1, 1, 9) + c8 * (-4.0 * ac / bc * dyox) + c4 * (-4.0 * ba / aa * dxoy);
      c5 = CF(1, 1, 5); c7 = CF(1, 1, 7);
      bndyc(3, c->xa, &ac, &bc, &gc, &dc);
      bndyc(1, c->yc, &aa, &ba, &ga, &da);
      det = aa * bc - ba * ac;
      CF(1, 1, 1) = CF(1, 1, 1) + c5;
      CF(1, 1, 3) = CF(1, 1, 3) + c7;
      CF(1, 1, 5) = 0.0; CF(1, 1, 7) = 0.0;
      CF(1, 1, 9) = CF(1, 1, 9) + c5 * (dlx2 * (ga * bc - ba * gc) / det) +
                    c7 * (dly2 * (aa * gc - ga * ac) / det);
    }
    if (c->nxb == 2) {
      bndyc(2, c->yc + dly, &ab, &bb, &gb, &db);
      bndyc(3, c->xb - dlx, &ac, &bc, &gc, &dc);
      c2 = CF(nx, 1, 2); c6 = CF(nx, 1, 6);
      CF(nx, 1, 1) = CF(nx, 1, 1) + c6 * (-ac / bc * dyox);
      CF(nx, 1, 2) = 0.0;
      CF(nx, 1, 3) = CF(nx, 1, 3) + c2 * (-3.0 * bb / ab * dxoy - dlx2 * gb / ab);
      CF(nx, 1, 4) = CF(nx, 1, 4) + c6 + c2;
      CF(nx, 1, 5) = CF(nx, 1, 5) + c6 * (-3.0 * ac / bc * dyox + dly2 * gc / bc);
      CF(nx, 1, 6) = 0.0;
      CF(nx, 1, 7) = CF(nx, 1, 7) - c2 * (bb / ab * dxoy);
      CF(nx, 1, 8) = 0.0;
      CF(nx, 1, 9) = CF(nx, 1, 9) + c6 * (4.0 * ac / bc * dyox) + c2 * (4.0 * bb / ab * dxoy);
      bndyc(3, c->xb, &ac, &bc, &gc, &dc);
      bndyc(2, c->yc, &ab, &bb, &gb, &db);
      det = bc * ab - ac * bb;
      c1 = CF(nx, 1, 1); c7 = CF(nx, 1, 7);
      CF(nx, 1, 1) = 0.0;
      CF(nx, 1, 3) = CF(nx, 1, 3) + c7;
      CF(nx, 1, 5) = CF(nx, 1, 5) + c1;
      CF(nx, 1, 7) = 0.0;
      CF(nx, 1, 9) = CF(nx, 1, 9) + c1 * (dlx2 * (bb * gc - gb * bc) / det) +
                     c7 * (dly2 * (ab * gc - gb * ac) / det);
    }
  }
  /* corner adjustments at y=yd (:963-1087) */
  if (c->nyd == 2) {
    if (c->nxa == 0) {
      bndyc(4, c->xa, &a0, &b0, &g0, &d0);
      bndyc(4, c->xa + dlx, &ap1, &bp1, &gp1, &dp1);
      c2 = CF(1, ny, 2); c3 = CF(1, ny, 3);
      CF(1, ny, 1) = CF(1, ny, 1) + c3 * (-a0 / b0 * dyox) +
                     c2 * (-3.0 * ap1 / bp1 * dyox - dly2 * gp1 / bp1);
      CF(1, ny, 2) = 0.0; CF(1, ny, 3) = 0.0; CF(1, ny, 4) = 0.0;
      CF(1, ny, 5) = CF(1, ny, 5) + c3 * (a0 / b0 * dyox) + c2 * (-ap1 / bp1 * dyox);
      CF(1, ny, 7) = CF(1, ny, 7) + c3;
      CF(1, ny, 8) = CF(1, ny, 8) + c2;
      CF(1, ny, 9) = CF(1, ny, 9) + c3 * (-dly2 * g0 / b0) + c2 * (4.0 * ap1 / bp1 * dyox);
      bndyc(4, c->xb - dlx, &am1, &bm1, &gm1, &dm1);
      bndyc(4, c->xb, &a0, &b0, &g0, &d0);
      c3 = CF(nx, ny, 3); c4 = CF(nx, ny, 4);
      CF(nx, ny, 1) = CF(nx, ny, 1) + c4 * (am1 / bm1 * dyox) + c3 * (-a0 / b0 * dyox);
      CF(nx, ny, 2) = 0.0; CF(nx, ny, 3) = 0.0; CF(nx, ny, 4) = 0.0;
      CF(nx, ny, 5) = CF(nx, ny, 5) + c4 * (3.0 * am1 / bm1 * dyox - dly2 * gm1 / bm1) +
                      c3 * (a0 / b0 * dyox);
      CF(nx, ny, 6) = CF(nx, ny, 6) + c4;
      CF(nx, ny, 7) = CF(nx, ny, 7) + c3;
      CF(nx, ny, 9) = CF(nx, ny, 9) + c4 * (-4.0 * am1 / bm1 * dyox) + c3 * (-dly2 * g0 / b0);
    } else if (c->nxa == 2) {
      bndyc(4, c->xa + dlx, &ad, &bd, &gd, &dd);
      bndyc(1, c->yd - dly, &aa, &ba, &ga, &da);
      c2 = CF(1, ny, 2); c6 = CF(1, ny, 6);
      CF(1, ny, 1) = CF(1, ny, 1) + c2 * (-3.0 * ad / bd * dyox - dly2 * gd / bd);
      CF(1, ny, 2) = 0.0;
      CF(1, ny, 3) = CF(1, ny, 3) + c6 * (-ba / aa * dxoy);
      CF(1, ny, 4) = 0.0;
      CF(1, ny, 5) = CF(1, ny, 5) + c2 * (-ad / bd * dyox);
      CF(1, ny, 6) = 0.0;
      CF(1, ny, 7) = CF(1, ny, 7) + c6 * (-3.0 * ba / aa * dxoy + dlx2 * ga / aa);
      CF(1, ny, 8) = CF(1, ny, 8) + c6 + c2;
      CF(1, ny, 9) = CF(1, ny, 9) + c2 * (4.0 * ad / bd * dyox) + c6 * (4.0 * ba / aa * dxoy);
      bndyc(1, c->yd, &aa, &ba, &ga, &da);
      bndyc(4, c->xa, &ad, &bd, &gd, &dd);
      det = ad * ba - bd * aa;
      c3 = CF(1, ny, 3); c5 = CF(1, ny, 5);
      CF(1, ny, 1) = CF(1, ny, 1) + c5;
      CF(1, ny, 3) = 0.0; CF(1, ny, 5) = 0.0;
      CF(1, ny, 7) = CF(1, ny, 7) + c3;
      CF(1, ny, 9) = CF(1, ny, 9) + c5 * (dlx2 * (ba * gd - ga * bd) / det) +
                     c3 * (dly2 * (aa * gd - ga * ad) / det);
    }
    if (c->nxb == 2) {
      bndyc(4, c->xb - dlx, &ad, &bd, &gd, &dd);
      bndyc(2, c->yd - dly, &ab, &bb, &gb, &db);
      c4 = CF(nx, ny, 4); c8 = CF(nx, ny, 8);
      CF(nx, ny, 1) = CF(nx, ny, 1) + c4 * (ad / bd * dyox);
      CF(nx, ny, 2) = 0.0;
      CF(nx, ny, 3) = CF(nx, ny, 3) + c8 * (bb / ab * dxoy);
      CF(nx, ny, 4) = 0.0;
      CF(nx, ny, 5) = CF(nx, ny, 5) + c4 * (3.0 * ad / bd * dyox - dly2 * gd / bd);
      CF(nx, ny, 6) = CF(nx, ny, 6) + c4 + c8;
      CF(nx, ny, 7) = CF(nx, ny, 7) + c8 * (3.0 * bb / ab * dxoy - dlx2 * gb / ab);
      CF(nx, ny, 8) = 0.0;
      CF(nx, ny, 9) = CF(nx, ny, 9) + c4 * (-4.0 * ad / bd * dyox) + c8 * (-4.0 * bb / ab * dxoy);
      bndyc(4, c->xb, &ad, &bd, &gd, &dd);
      bndyc(2, c->yd, &ab, &bb, &gb, &db);
      det = ad * bb - bd * ab;
      c1 = CF(nx, ny, 1); c3 = CF(nx, ny, 3);
      CF(nx, ny, 1) = 0.0; CF(nx, ny, 3) = 0.0;
      CF(nx, ny, 5) = CF(nx, ny, 5) + c1;
      CF(nx, ny, 7) = CF(nx, ny, 7) + c3;
      CF(nx, ny, 9) = CF(nx, ny, 9) + c1 * (dlx2 * (bd * gb - gd * bb) / det) +
                      c3 * (dly2 * (ab * gd - gb * ad) / det);
    }
  }
  /* mixed-periodic corners (:1089-1166) */
  if (c->nxa == 2 && c->nyc == 0) {
    bndyc(1, c->yc, &a0, &b0, &g0, &d0);
    bndyc(1, c->yc + dly, &ap1, &bp1, &gp1, &dp1);
    c5 = CF(1, 1, 5); c4 = CF(1, 1, 4);
    CF(1, 1, 3) = CF(1, 1, 3) + c5 * (b0 / a0 * dxoy) +
                  c4 * (3. * bp1 / ap1 * dxoy + dlx2 * gp1 / ap1);
    CF(1, 1, 2) = CF(1, 1, 2) + c4;
    CF(1, 1, 1) = CF(1, 1, 1) + c5;
    CF(1, 1, 7) = CF(1, 1, 7) + c5 * (-b0 / a0 * dxoy) + c4 * (bp1 / ap1 * dxoy);
    CF(1, 1, 6) = 0.0; CF(1, 1, 5) = 0.0; CF(1, 1, 4) = 0.0;
    CF(1, 1, 9) = CF(1, 1, 9) + c4 * (-4. * bp1 / ap1 * dxoy) + c5 * (dlx2 * g0 / a0);
    bndyc(1, c->yd - dly, &am1, &bm1, &gm1, &dm1);
    bndyc(1, c->yd, &a0, &b0, &g0, &d0);
    c6 = CF(1, ny, 6); c5 = CF(1, ny, 5);
    CF(1, ny, 3) = CF(1, ny, 3) + c6 * (-bm1 / am1 * dxoy) + c5 * (b0 / a0 * dxoy);
    CF(1, ny, 1) = CF(1, ny, 1) + c5;
    CF(1, ny, 8) = CF(1, ny, 8) + c6;
    CF(1, ny, 7) = CF(1, ny, 7) + c6 * (-3. * bm1 / am1 * dxoy + dlx2 * gm1 / am1) +
                   c5 * (-b0 / a0 * dxoy);
    CF(1, ny, 6) = 0.0; CF(1, ny, 5) = 0.0; CF(1, ny, 4) = 0.0;
    CF(1, ny, 9) = CF(1, ny, 9) + c6 * (4. * bm1 / am1 * dxoy) + c5 * (dlx2 * g0 / a0);
  }
  if (c->nxb == 2 && c->nyc == 0) {
    bndyc(2, c->yc, &a0, &b0, &g0, &d0);
    bndyc(2, c->yc + dly, &ap1, &bp1, &gp1, &dp1);
    c2 = CF(nx, 1, 2); c1 = CF(nx, 1, 1);
    CF(nx, 1, 3) = CF(nx, 1, 3) + c1 * (-b0 / a0 * dxoy) +
                   c2 * (-3. * bp1 / ap1 * dxoy - dlx2 * gp1 / ap1);
    CF(nx, 1, 2) = 0.0; CF(nx, 1, 1) = 0.0; CF(nx, 1, 8) = 0.0;
    CF(nx, 1, 7) = CF(nx, 1, 7) + c1 * (b0 / a0 * dxoy) + c2 * (-bp1 / ap1 * dxoy);
    CF(nx, 1, 5) = CF(nx, 1, 5) + c1;
    CF(nx, 1, 4) = CF(nx, 1, 4) + c2;
    CF(nx, 1, 9) = CF(nx, 1, 9) + c1 * (-dlx2 * g0 / a0) + c2 * (4. * bp1 / ap1 * dxoy);
    bndyc(2, c->yd - dly, &am1, &bm1, &gm1, &dm1);
    bndyc(2, c->yd, &a0, &b0, &g0, &d0);
    c1 = CF(nx, ny, 1); c8 = CF(nx, ny, 8);
    CF(nx, ny, 3) = CF(nx, ny, 3) + c8 * (bm1 / am1 * dxoy) + c1 * (-b0 / a0 * dxoy);
    CF(nx, ny, 2) = 0.0; CF(nx, ny, 1) = 0.0; CF(nx, ny, 8) = 0.0;
    CF(nx, ny, 7) = CF(nx, ny, 7) + c8 * (3. * bm1 / am1 * dxoy - dlx2 * gm1 / am1) +
                    c1 * (b0 / a0 * dxoy);
    CF(nx, ny, 6) = CF(nx, ny, 6) + c8;
    CF(nx, ny, 5) = CF(nx, ny, 5) + c1;
    CF(nx, ny, 9) = CF(nx, ny, 9) + c8 * (-4. * bm1 / am1 * dxoy) + c1 * (-dlx2 * g0 / a0);
  }
  /* specified (Dirichlet) boundaries (:1167-1205) */
  if (c->nxa == 1)
    for (int j = 1; j <= ny; ++j) { for (int l = 1; l <= 9; ++l) CF(1, j, l) = 0.0; CF(1, j, 9) = 1.0; }
  if (c->nxb == 1)
    for (int j = 1; j <= ny; ++j) { for (int l = 1; l <= 9; ++l) CF(nx, j, l) = 0.0; CF(nx, j, 9) = 1.0; }
  if (c->nyc == 1)
    for (int i = 1; i <= nx; ++i) { for (int l = 1; l <= 9; ++l) CF(i, 1, l) = 0.0; CF(i, 1, 9) = 1.0; }
  if (c->nyd == 1)
    for (int i = 1; i <= nx; ++i) { for (int l = 1; l <= 9; ++l) CF(i, ny, l) = 0.0; CF(i, ny, 9) = 1.0; }
}

/* adjmd2cr (src/mud2cr.f:1271-1616).  Every derivative-boundary correction
 * there is a multiple of gbdy, which bndcr returns as 0 (src/mg_solver2d.F90:
 * 230,248), so those corrections add exact zeros; what remains is setting the
 * right-hand side at specified boundaries from phi (:1580-1615). */
static void mg_adjust(mg_ctx* c, int lev) {
  mg_level* L = &c->lev[lev];
  int nx = L->nx, ny = L->ny;
  double *cf = L->cf, *phi = L->phi;
  if (c->nxa == 1) for (int j = 1; j <= ny; ++j) CF(1, j, 10) = PH(1, j);
  if (c->nxb == 1) for (int j = 1; j <= ny; ++j) CF(nx, j, 10) = PH(nx, j);
  if (c->nyc == 1) for (int i = 1; i <= nx; ++i) CF(i, 1, 10) = PH(i, 1);
  if (c->nyd == 1) for (int i = 1; i <= nx; ++i) CF(i, ny, 10) = PH(i, ny);
}

/* relmd2crp (src/mud2cr.f:1693-1795): four colours, each swept as four
 * row-phase loops (j = 1,2,3,4 mod 4), periodic virtual points refreshed
 * after every colour. */
static void mg_relax(mg_ctx* c, int lev) {
  mg_level* L = &c->lev[lev];
  int nx = L->nx, ny = L->ny;
  double *cf = L->cf, *phi = L->phi;
  int ist[4] = {1, 4, 3, 2};
  for (int color = 0; color < 4; ++color) {
    for (int ph = 0; ph < 4; ++ph) {
      for (int j = 1 + ph; j <= ny; j += 4)
        for (int i = ist[ph]; i <= nx; i += 4)
          PH(i, j) = (CF(i, j, 10) - (CF(i, j, 1) * PH(i + 1, j) + CF(i, j, 2) * PH(i + 1, j + 1) +
                                      CF(i, j, 3) * PH(i, j + 1) + CF(i, j, 4) * PH(i - 1, j + 1) +
                                      CF(i, j, 5) * PH(i - 1, j) + CF(i, j, 6) * PH(i - 1, j - 1) +
                                      CF(i, j, 7) * PH(i, j - 1) + CF(i, j, 8) * PH(i + 1, j - 1))) /
                     CF(i, j, 9);
    }
    if (c->nxa == 0)
      for (int j = 1; j <= ny; ++j) { PH(0, j) = PH(nx - 1, j); PH(nx + 1, j) = PH(2, j); }
    if (c->nyc == 0)
      for (int i = 1; i <= nx; ++i) { PH(i, 0) = PH(i, ny - 1); PH(i, ny + 1) = PH(i, 2); }
    int t = ist[3];
    ist[3] = ist[2]; ist[2] = ist[1]; ist[1] = ist[0]; ist[0] = t;
  }
}

/* res2 (src/mudcom.f:120-258): full weighting of resf into rhsc */
static void mg_res2(const mg_ctx* c, int nx, int ny, const double* resf, int ncx, int ncy,
                    double* rhsc) {
#define RF(i, j) resf[((i)-1) + (size_t)nx * ((j)-1)]
#define RC(i, j) rhsc[((i)-1) + (size_t)ncx * ((j)-1)]
#define FW(im, ip, jm, jp, i, j) \
  ((RF(im, jm) + RF(ip, jm) + RF(im, jp) + RF(ip, jp) + 2. * (RF(im, j) + RF(ip, j) + RF(i, jm) + RF(i, jp)) + 4. * RF(i, j)) * .0625)
  int ix = (ncx == nx) ? 0 : 1, jy = (ncy == ny) ? 0 : 1;
  for (int jc = 2; jc <= ncy - 1; ++jc) {
    int j = jy ? jc + jc - 1 : jc;
    for (int ic = 2; ic <= ncx - 1; ++ic) {
      int i = ix ? ic + ic - 1 : ic;
      RC(ic, jc) = FW(i - 1, i + 1, j - 1, j + 1, i, j);
    }
  }
  for (int jc = 1; jc <= ncy; jc += ncy - 1) {
    int j = jc + jy * (jc - 1);
    int jm1 = j - 1 > 2 ? j - 1 : 2, jp1 = j + 1 < ny - 1 ? j + 1 : ny - 1;
    if (j == 1 && c->nyc == 0) jm1 = ny - 1;
    if (j == ny && c->nyc == 0) jp1 = 2;
    for (int ic = 1; ic <= ncx; ic += ncx - 1) {
      int i = ic + ix * (ic - 1);
      int im1 = i - 1 > 2 ? i - 1 : 2, ip1 = i + 1 < nx - 1 ? i + 1 : nx - 1;
      if (i == 1 && c->nxa == 0) im1 = nx - 1;
      if (i == nx && c->nxa == 0) ip1 = 2;
      RC(ic, jc) = FW(im1, ip1, jm1, jp1, i, j);
    }
    for (int ic = 2; ic <= ncx - 1; ++ic) {
      int i = ic + ix * (ic - 1);
      RC(ic, jc) = FW(i - 1, i + 1, jm1, jp1, i, j);
    }
  }
  for (int ic = 1; ic <= ncx; ic += ncx - 1) {
    int i = ic + ix * (ic - 1);
    int im1 = i - 1 > 2 ? i - 1 : 2, ip1 = i + 1 < nx - 1 ? i + 1 : nx - 1;
    if (i == 1 && c->nxa == 0) im1 = nx - 1;
    if (i == nx && c->nxa == 0) ip1 = 2;
    for (int jc = 2; jc <= ncy - 1; ++jc) {
      int j = jc + jy * (jc - 1);
      RC(ic, jc) = FW(im1, ip1, j - 1, j + 1, i, j);
    }
  }
  if (c->nxa == 1) for (int jc = 1; jc <= ncy; ++jc) RC(1, jc) = 0.0;
  if (c->nxb == 1) for (int jc = 1; jc <= ncy; ++jc) RC(ncx, jc) = 0.0;
  if (c->nyc == 1) for (int ic = 1; ic <= ncx; ++ic) RC(ic, 1) = 0.0;
  if (c->nyd == 1) for (int ic = 1; ic <= ncx; ++ic) RC(ic, ncy) = 0.0;
#undef RF
#undef RC
#undef FW
}

/* resmd2cr (src/mud2cr.f:1618-1665) */
static void mg_restrict(mg_ctx* c, int lev) {
  mg_level *F = &c->lev[lev], *C = &c->lev[lev - 1];
  int nx = F->nx, ny = F->ny;
  double *cf = F->cf, *phi = F->phi;
  for (size_t t = 0; t < (size_t)(C->nx + 2) * (C->ny + 2); ++t) C->phi[t] = 0.0;
  for (int j = 1; j <= ny; ++j)
    for (int i = 1; i <= nx; ++i)
      c->resf[(i - 1) + (size_t)nx * (j - 1)] =
          CF(i, j, 10) - (CF(i, j, 1) * PH(i + 1, j) + CF(i, j, 2) * PH(i + 1, j + 1) +
                          CF(i, j, 3) * PH(i, j + 1) + CF(i, j, 4) * PH(i - 1, j + 1) +
                          CF(i, j, 5) * PH(i - 1, j) + CF(i, j, 6) * PH(i - 1, j - 1) +
                          CF(i, j, 7) * PH(i, j - 1) + CF(i, j, 8) * PH(i + 1, j - 1) +
                          CF(i, j, 9) * PH(i, j));
  mg_res2(c, nx, ny, c->resf, C->nx, C->ny, C->cf + (size_t)9 * C->nx * C->ny);
}

/* prolon1 (src/mudcom.f:404-500), p(0:ncx+1) -> q(0:nx+1) */
static void mg_prolon1(int ncx, const double* p, int nx, double* q, int nxa, int nxb, int intpol) {
  int ist = 1, ioddst = 1, ifn = nx, ioddfn = nx;
  if (nxa == 1) { ist = 2; ioddst = 3; }
  if (nxb == 1) { ifn = nx - 1; ioddfn = nx - 2; }
  if (intpol == 1 || ncx < 4) {
    if (ncx < nx) {
      for (int i = ioddst; i <= ioddfn; i += 2) q[i] = p[(i + 1) / 2];
      for (int i = 2; i <= ifn; i += 2) q[i] = 0.5 * (q[i - 1] + q[i + 1]);
    } else {
      for (int i = ist; i <= ifn; ++i) q[i] = p[i];
    }
    if (nxa == 0) { q[0] = q[nx - 1]; q[nx + 1] = q[2]; }
    return;
  }
  if (ncx < nx) {
    for (int i = ioddst; i <= ioddfn; i += 2) q[i] = p[(i + 1) / 2];
    for (int i = 4; i <= nx - 3; i += 2) q[i] = (-q[i - 3] + 9. * (q[i - 1] + q[i + 1]) - q[i + 3]) * .0625;
    if (nxa != 0) {
      q[2] = (5. * q[1] + 15. * q[3] - 5. * q[5] + q[7]) * .0625;
      q[nx - 1] = (5. * q[nx] + 15. * q[nx - 2] - 5. * q[nx - 4] + q[nx - 6]) * .0625;
    } else {
      q[2] = (-q[nx - 2] + 9. * (q[1] + q[3]) - q[5]) * .0625;
      q[nx - 1] = (-q[nx - 4] + 9. * (q[nx - 2] + q[nx]) - q[3]) * .0625;
      q[nx + 1] = q[2];
      q[0] = q[nx - 1];
    }
  } else {
    for (int i = ist; i <= ifn; ++i) q[i] = p[i];
    if (nxa == 0) { q[0] = q[nx - 1]; q[nx + 1] = q[2]; }
  }
}

/* prolon2 (src/mudcom.f:261-402), p(0:ncx+1,0:ncy+1) -> q(0:nx+1,0:ny+1) */
static void mg_prolon2(const mg_ctx* c, int ncx, int ncy, const double* p, int nx, int ny,
                       double* q) {
  int nxa = c->nxa, nxb = c->nxb, nyc = c->nyc, nyd = c->nyd, intpol = c->intpol;
  int ist = 1, ifn = nx, jst = 1, jfn = ny, joddst = 1, joddfn = ny;
  size_t lq = (size_t)nx + 2, lp = (size_t)ncx + 2;
#define Q(i, j) q[(i) + lq * (j)]
  if (nxa == 1) ist = 2;
  if (nxb == 1) ifn = nx - 1;
  if (nyc == 1) { jst = 2; joddst = 3; }
  if (nyd == 1) { jfn = ny - 1; joddfn = ny - 2; }
  if (intpol == 1 || ncy < 4) {
    if (ncy < ny) {
      for (int j = joddst; j <= joddfn; j += 2)
        mg_prolon1(ncx, p + lp * (j / 2 + 1), nx, q + lq * j, nxa, nxb, intpol);
      for (int j = 2; j <= jfn; j += 2)
        for (int i = ist; i <= ifn; ++i) Q(i, j) = 0.5 * (Q(i, j - 1) + Q(i, j + 1));
    } else {
      for (int j = jst; j <= jfn; ++j) mg_prolon1(ncx, p + lp * j, nx, q + lq * j, nxa, nxb, intpol);
    }
    if (nyc == 0)
      for (int i = ist; i <= ifn; ++i) { Q(i, 0) = Q(i, ny - 1); Q(i, ny + 1) = Q(i, 2); }
    return;
  }
  if (ncy < ny) {
    for (int j = joddst; j <= joddfn; j += 2)
      mg_prolon1(ncx, p + lp * (j / 2 + 1), nx, q + lq * j, nxa, nxb, intpol);
    for (int j = 4; j <= ny - 3; j += 2)
      for (int i = ist; i <= ifn; ++i)
        Q(i, j) = (-Q(i, j - 3) + 9. * (Q(i, j - 1) + Q(i, j + 1)) - Q(i, j + 3)) * .0625;
    if (nyc != 0) {
      for (int i = ist; i <= ifn; ++i) {
        Q(i, 2) = (5. * Q(i, 1) + 15. * Q(i, 3) - 5. * Q(i, 5) + Q(i, 7)) * .0625;
        Q(i, ny - 1) = (5. * Q(i, ny) + 15. * Q(i, ny - 2) - 5. * Q(i, ny - 4) + Q(i, ny - 6)) * .0625;
      }
    } else {
      for (int i = ist; i <= ifn; ++i) {
        Q(i, 2) = (-Q(i, ny - 2) + 9. * (Q(i, 1) + Q(i, 3)) - Q(i, 5)) * .0625;
        Q(i, ny - 1) = (-Q(i, ny - 4) + 9. * (Q(i, ny - 2) + Q(i, ny)) - Q(i, 3)) * .0625;
        Q(i, ny + 1) = Q(i, 2);
        Q(i, 0) = Q(i, ny - 1);
      }
    }
  } else {
    for (int j = jst; j <= jfn; ++j) mg_prolon1(ncx, p + lp * j, nx, q + lq * j, nxa, nxb, intpol);
    if (nyc == 0)
      for (int i = ist; i <= ifn; ++i) { Q(i, 0) = Q(i, ny - 1); Q(i, ny + 1) = Q(i, 2); }
  }
#undef Q
}

/* cor2 (src/mudcom.f:502-553) */
static void mg_correct(mg_ctx* c, int lev) {
  mg_level *F = &c->lev[lev + 1], *C = &c->lev[lev];
  int nx = F->nx, ny = F->ny;
  double* phi = F->phi;
  size_t n = (size_t)(nx + 2) * (ny + 2);
  for (size_t t = 0; t < n; ++t) c->phcor[t] = 0.0;
  mg_prolon2(c, C->nx, C->ny, C->phi, nx, ny, c->phcor);
  int ist = 1, ifn = nx, jst = 1, jfn = ny;
  if (c->nxa == 1) ist = 2;
  if (c->nxb == 1) ifn = nx - 1;
  if (c->nyc == 1) jst = 2;
  if (c->nyd == 1) jfn = ny - 1;
  for (int j = jst; j <= jfn; ++j)
    for (int i = ist; i <= ifn; ++i) PH(i, j) = PH(i, j) + c->phcor[i + (size_t)(nx + 2) * j];
  if (c->nyc == 0)
    for (int i = ist; i <= ifn; ++i) { PH(i, 0) = PH(i, ny - 1); PH(i, ny + 1) = PH(i, 2); }
  if (c->nxa == 0)
    for (int j = jst; j <= jfn; ++j) { PH(0, j) = PH(nx - 1, j); PH(nx + 1, j) = PH(2, j); }
}

/* kcymd2cr (src/mud2cr.f:340-499): one k-cycle from level kcur */
static void mg_kcycle(mg_ctx* c, int kcur) {
  int kount[51];
  for (int l = 1; l <= kcur; ++l) kount[l] = 0;
  for (int l = 0; l < c->iprer; ++l) mg_relax(c, kcur);
  if (kcur != 1) {
    mg_restrict(c, kcur);
    int klevel = kcur - 1, nrel = c->iprer;
    for (;;) {
      if (klevel == kcur) break;
      kount[klevel]++;
      for (int l = 0; l < nrel; ++l) mg_relax(c, klevel);
      if (kount[klevel] == c->kcycle + 1) {
        mg_correct(c, klevel);
        kount[klevel] = 0;
        klevel++;
        nrel = c->ipost;
      } else if (klevel > 1) {
        mg_restrict(c, klevel);
        klevel--;
        nrel = c->iprer;
      } else {
        for (int l = 0; l < c->ipost; ++l) mg_relax(c, klevel);
        mg_correct(c, 1);
        nrel = c->ipost;
        klevel = 2;
      }
    }
  }
  for (int l = 0; l < c->ipost; ++l) mg_relax(c, kcur);
}

/* trsfc2 (src/mudcom.f:61-118): inject phi and rhs from level lev+1 to lev */
static void mg_transfer(mg_ctx* c, int lev) {
  mg_level *F = &c->lev[lev + 1], *C = &c->lev[lev];
  int nx = F->nx, ny = F->ny, ncx = C->nx, ncy = C->ny;
  const double *phi = F->phi, *rhs = F->cf + (size_t)9 * nx * ny;
  double *phic = C->phi, *rhsc = C->cf + (size_t)9 * ncx * ncy;
  size_t lc = (size_t)ncx + 2;
  for (int jc = 0; jc <= ncy + 1; ++jc) { phic[0 + lc * jc] = 0.0; phic[ncx + 1 + lc * jc] = 0.0; }
  for (int ic = 0; ic <= ncx + 1; ++ic) { phic[ic] = 0.0; phic[ic + lc * (ncy + 1)] = 0.0; }
  for (int jc = 1; jc <= ncy; ++jc) {
    int j = (ncy < ny) ? jc + jc - 1 : jc;
    for (int ic = 1; ic <= ncx; ++ic) {
      int i = (ncx < nx) ? ic + ic - 1 : ic;
      phic[ic + lc * jc] = PH(i, j);
      rhsc[(ic - 1) + (size_t)ncx * (jc - 1)] = rhs[(i - 1) + (size_t)nx * (j - 1)];
    }
  }
}

static void mg_setup(mg_ctx* c, const ora* o) {
  memset(c, 0, sizeof(*c));
  const ekm_params* p = &o->p;
  c->o = o;
  c->nxa = p->mudpack_bc[0]; c->nxb = p->mudpack_bc[1];
  c->nyc = p->mudpack_bc[2]; c->nyd = p->mudpack_bc[3];
  c->ixp = p->mg_ixp; c->jyq = p->mg_jyq; c->iex = p->mg_iex; c->jey = p->mg_jey;
  c->ngrid = c->iex > c->jey ? c->iex : c->jey;
  c->nfx = o->NZ2; c->nfy = o->NY2;
  c->maxcy = p->mg_maxcy; c->tolmax = p->mg_tolmax;
  c->kcycle = 2; c->iprer = 2; c->ipost = 1; c->intpol = 3;
  c->xa = 0.0; c->xb = (double)(c->nfx - 1); c->yc = 0.0; c->yd = (double)(c->nfy - 1);
  for (int k = 1; k <= c->ngrid; ++k) {
    int ex = k + c->iex - c->ngrid, ey = k + c->jey - c->ngrid;
    mg_level* L = &c->lev[k];
    L->nx = c->ixp * (1 << ((ex > 1 ? ex : 1) - 1)) + 1;
    L->ny = c->jyq * (1 << ((ey > 1 ? ey : 1) - 1)) + 1;
    L->phi = dalloc((size_t)(L->nx + 2) * (L->ny + 2));
    L->cf = dalloc((size_t)L->nx * L->ny * 10);
  }
  c->resf = dalloc((size_t)c->nfx * c->nfy);
  c->phcor = dalloc((size_t)(c->nfx + 2) * (c->nfy + 2));
}

static void mg_free(mg_ctx* c) {
  for (int k = 1; k <= c->ngrid; ++k) { free(c->lev[k].phi); free(c->lev[k].cf); }
  free(c->resf);
  free(c->phcor);
}

/* MULTIGRID_MUDPACK (src/mg_solver2d.F90:75-201): phi = 0, rhs = -RHS_TMP
 * (phi_bc, src/duct.F90:3867-3895), discretise every level (intl=0), then
 * mud2cr1 (src/mud2cr.f:211-338) with iguess = 0. */
static int mg_solve(mg_ctx* c, double kx2, const double* rhs_tmp, double* phiout) {
  int nfx = c->nfx, nfy = c->nfy;
  c->kx2 = kx2;
  for (int k = 1; k <= c->ngrid; ++k) mg_discretize(c, k);
  mg_level* T = &c->lev[c->ngrid];
  /* swk2: phif = 0, rhs = -RHS_TMP, virtual boundaries zero */
  for (size_t t = 0; t < (size_t)(nfx + 2) * (nfy + 2); ++t) T->phi[t] = 0.0;
  double* rhs = T->cf + (size_t)9 * nfx * nfy;
  for (int j = 1; j <= nfy; ++j)
    for (int i = 1; i <= nfx; ++i)
      rhs[(i - 1) + (size_t)nfx * (j - 1)] = -rhs_tmp[(i - 1) + (size_t)nfx * (j - 1)];
  for (int k = c->ngrid - 1; k >= 1; --k) mg_transfer(c, k);
  for (int k = 1; k <= c->ngrid; ++k) mg_adjust(c, k);
  for (int k = 1; k <= c->ngrid - 1; ++k) {
    mg_kcycle(c, k);
    mg_level *C = &c->lev[k], *F = &c->lev[k + 1];
    mg_prolon2(c, C->nx, C->ny, C->phi, F->nx, F->ny, F->phi);
  }
  /* phif holds the previous iterate for the relmax test (zero on entry) */
  for (size_t t = 0; t < (size_t)nfx * nfy; ++t) phiout[t] = 0.0;
  double* phi = T->phi;
  int nx = nfx;
  c->itero = 0;
  for (int iter = 1; iter <= c->maxcy; ++iter) {
    c->itero = iter;
    mg_kcycle(c, c->ngrid);
    if (c->tolmax > 0.0) {
      double relmax = 0.0, phmax = 0.0;
      for (int j = 1; j <= nfy; ++j)
        for (int i = 1; i <= nfx; ++i) {
          double w = PH(i, j), *pf = &phiout[(i - 1) + (size_t)nfx * (j - 1)];
          phmax = fmax(phmax, fabs(w));
          relmax = fmax(relmax, fabs(w - *pf));
          *pf = w;
        }
      if (phmax > 0.0) relmax = relmax / phmax;
      c->relmax = relmax;
      if (relmax <= c->tolmax) return c->itero;
    }
  }
  for (int j = 1; j <= nfy; ++j)
    for (int i = 1; i <= nfx; ++i) phiout[(i - 1) + (size_t)nfx * (j - 1)] = PH(i, j);
  return c->itero;
}
#undef COF
#undef CF
#undef PH

/* ------------------------------------------------------------------------ */
/* Consistent projection (proj_mode 1).  Not in the reference: the rows of   */
/* the cells the divergence diagnostic covers (J=2..NY-1, K=2..NZ-1) are     */
/* replaced by the divergence (src/duct.F90:2754-2760) of the projection's   */
/* face gradient (:2713-2724) and the centre correction of CU1 (:2731), so   */
/* J div u after the projection is the residual of the solve there.  Other   */
/* rows keep MUDPACK's discretisation.  Solved by alternating zebra-line     */
/* V(1,1) cycles (full weighting, bilinear prolongation) inside residual     */
/* iterations until max|residual| <= proj_tol.                               */
/*                                                                           */
/* Periodic z ends (MUDPACK nxa = nxb = 0): as in MUDPACK, every column of a */
/* level is an unknown with its own row, and the virtual neighbours are      */
/* e(-1) = e(nx-2) and e(nx) = e(1) (src/mud2cr.f:1693-1795), held in the    */
/* ring columns; a z line is then the chain 0..nx-1 plus the couplings of    */
/* its first point to nx-2 and of its last to 1, and the restriction wraps   */
/* the same way.  A problem with no specified side is singular at           */
/* kx = 0 (constants are in the null space of every row); its solution is    */
/* pinned to 0 at the point (K = 0, J = NY+1) of every level, a MUDPACK      */
/* boundary row, so that the rows of the divergence cells all hold.          */
/* ------------------------------------------------------------------------ */
typedef struct tlev {
  int nx, ny;
  double *cf, *e, *r, *res; /* cf: 9 planes [q][j][i]; e: (nx+2)(ny+2) with a ring */
  unsigned char* spec;
  int pin;                  /* the pinned point (i + nx j) of a singular solve, or -1 */
} tlev;

typedef struct tctx {
  int ngrid, nxa, nxb, nyc, nyd;
  int per;                  /* periodic z ends */
  tlev lev[17];
  double *lo, *di, *up, *s, *cp, *dp, *phi, *zz, *ys, *zs;
} tctx;

#define TSPEC(L, p) ((L)->spec[p] || (int)(p) == (L)->pin)

/* Row of cell (K,J) of the consistent operator, coefficients in MUDPACK's
 * neighbour order c1 (K+1,J), c2 (K+1,J+1), c3 (K,J+1), c4 (K-1,J+1),
 * c5 (K-1,J), c6 (K-1,J-1), c7 (K,J-1), c8 (K+1,J-1), c9 (K,J).  Faces the
 * projection does not correct (outside J=JSTART..JEND, K=ZSTART..ZEND)
 * contribute nothing. */
static void tight_row(const ora* o, double kx2, int K, int J, double* c) {
  const double tm = (J + 1 <= o->jend) ? 1.0 : 0.0, bm = (J >= o->jstart) ? 1.0 : 0.0;
  const double rm = (K + 1 <= o->zend) ? 1.0 : 0.0, lm = (K >= o->zstart) ? 1.0 : 0.0;
  const double a = G22(o, K, J + 1, 1) * tm, b = G22(o, K, J, 1) * bm;
  const double cc = G11(o, K + 1, J, 2) * rm, d = G11(o, K, J, 2) * lm;
  const double e = 0.25 * G12(o, K, J + 1, 1) * tm, f = 0.25 * G12(o, K, J, 1) * bm;
  const double g = 0.25 * G12(o, K + 1, J, 2) * rm, h = 0.25 * G12(o, K, J, 2) * lm;
  c[0] = (e - f) + cc;
  c[1] = e + g;
  c[2] = (a + g) - h;
  c[3] = -e - h;
  c[4] = (-e + f) + d;
  c[5] = f + h;
  c[6] = (b - g) + h;
  c[7] = -f - g;
  c[8] = (((-a - b) - cc) - d) - kx2 * JAC(o, K, J);
}

static void tight_setup(tctx* t, const mg_ctx* c) {
  memset(t, 0, sizeof(*t));
  t->ngrid = c->ngrid;
  t->nxa = c->nxa; t->nxb = c->nxb; t->nyc = c->nyc; t->nyd = c->nyd;
  t->per = c->nxa == 0;
  size_t mx = 0;
  for (int k = 1; k <= c->ngrid; ++k) {
    tlev* L = &t->lev[k];
    L->nx = c->lev[k].nx; L->ny = c->lev[k].ny;
    size_t n = (size_t)L->nx * L->ny;
    L->cf = dalloc(9 * n);
    L->e = dalloc((size_t)(L->nx + 2) * (L->ny + 2));
    L->r = dalloc(n);
    L->res = dalloc(n);
    L->spec = (unsigned char*)calloc(n, 1);
    L->pin = -1;
    for (int j = 0; j < L->ny; ++j)
      for (int i = 0; i < L->nx; ++i)
        L->spec[i + (size_t)L->nx * j] = (t->nxa == 1 && i == 0) || (t->nxb == 1 && i == L->nx - 1) ||
                                         (t->nyc == 1 && j == 0) || (t->nyd == 1 && j == L->ny - 1);
    if ((size_t)L->nx > mx) mx = L->nx;
    if ((size_t)L->ny > mx) mx = L->ny;
  }
  t->lo = dalloc(mx); t->di = dalloc(mx); t->up = dalloc(mx);
  t->s = dalloc(mx); t->cp = dalloc(mx); t->dp = dalloc(mx); t->zz = dalloc(mx);
  t->ys = dalloc(mx); t->zs = dalloc(2 * mx);
  t->phi = dalloc((size_t)(c->nfx + 2) * (c->nfy + 2));
}

static void tight_free(tctx* t) {
  for (int k = 1; k <= t->ngrid; ++k) {
    tlev* L = &t->lev[k];
    free(L->cf); free(L->e); free(L->r); free(L->res); free(L->spec);
  }
  free(t->lo); free(t->di); free(t->up); free(t->s); free(t->cp); free(t->dp); free(t->zz); free(t->ys); free(t->zs); free(t->phi);
}

/* operator coefficients of every level for kx2: MUDPACK's discretisation,
 * with the consistent rows on the finest level */
static void tight_discretize(tctx* t, mg_ctx* c, const ora* o, double kx2) {
  c->kx2 = kx2;
  for (int k = 1; k <= t->ngrid; ++k) {
    mg_discretize(c, k);
    tlev* L = &t->lev[k];
    size_t n = (size_t)L->nx * L->ny;
    for (int q = 0; q < 9; ++q) memcpy(L->cf + q * n, c->lev[k].cf + q * n, n * sizeof(double));
  }
  /* singular: kx = 0 with no specified side */
  const int sing = kx2 == 0.0 && t->nxa != 1 && t->nxb != 1 && t->nyc != 1 && t->nyd != 1;
  for (int k = 1; k <= t->ngrid; ++k) {
    tlev* L = &t->lev[k];
    L->pin = sing ? L->nx * (L->ny - 1) : -1;
  }
  tlev* F = &t->lev[t->ngrid];
  size_t n = (size_t)F->nx * F->ny;
  double row[9];
  for (int J = 2; J <= o->NY - 1; ++J)
    for (int K = 2; K <= o->NZ - 1; ++K) {
      tight_row(o, kx2, K, J, row);
      for (int q = 0; q < 9; ++q) F->cf[q * n + K + (size_t)F->nx * J] = row[q];
    }
}

#define TE(L, i, j) (L)->e[((i) + 1) + (size_t)((L)->nx + 2) * ((j) + 1)]
#define TC(L, q, i, j) (L)->cf[(size_t)(q) * (L)->nx * (L)->ny + (i) + (size_t)(L)->nx * (j)]

/* res = r - A e on level L (0 at specified points) */
static void tight_residual(const tlev* L, const double* e, const double* r, double* res) {
  const int nx = L->nx, ny = L->ny;
  static const int di[8] = {1, 1, 0, -1, -1, -1, 0, 1}, dj[8] = {0, 1, 1, 1, 0, -1, -1, -1};
  for (int j = 0; j < ny; ++j)
    for (int i = 0; i < nx; ++i) {
      size_t p = i + (size_t)nx * j;
      if (TSPEC(L, p)) { res[p] = 0.0; continue; }
      double acc = 0.0;
      for (int q = 0; q < 8; ++q) acc += TC(L, q, i, j) * e[(i + di[q] + 1) + (size_t)(nx + 2) * (j + dj[q] + 1)];
      acc += TC(L, 8, i, j) * e[(i + 1) + (size_t)(nx + 2) * (j + 1)];
      res[p] = r[p] - acc;
    }
}

static void thomas_line(int n, const double* lo, const double* di, const double* up, const double* s,
                        double* cp, double* dp, double* x, size_t stride) {
  cp[0] = up[0] / di[0];
  dp[0] = s[0] / di[0];
  for (int i = 1; i < n; ++i) {
    double den = di[i] - lo[i] * cp[i - 1];
    cp[i] = up[i] / den;
    dp[i] = (s[i] - lo[i] * dp[i - 1]) / den;
  }
  x[(size_t)(n - 1) * stride] = dp[n - 1];
  for (int i = n - 2; i >= 0; --i) x[(size_t)i * stride] = dp[i] - cp[i] * x[(size_t)(i + 1) * stride];
}

/* periodic ends: the ring columns take the virtual values e(-1) = e(nx-2)
 * and e(nx) = e(1) */
static void tight_wrap(const tlev* L, double* e) {
  const int nx = L->nx, ld = nx + 2;
  for (int j = -1; j <= L->ny; ++j) {
    double* row = e + (size_t)ld * (j + 1) + 1;
    row[-1] = row[nx - 2];
    row[nx] = row[1];
  }
}

/* z line of a periodic level: the tridiagonal chain T (lo, di, up over
 * 0..n-1) plus lo[0] coupling point 0 to n-2 and up[n-1] coupling point n-1
 * to 1.  A = T + U V^T with U = [lo0 e_0, up_{n-1} e_{n-1}], V = [e_{n-2}, e_1];
 * Woodbury: x = y - W (I + V^T W)^-1 V^T y with y = T^-1 s, W = T^-1 U. */
static void thomas_periodic(tctx* t, int n, double* x, size_t stride) {
  const double a0 = t->lo[0], bn = t->up[n - 1];
  double* u = t->zz;
  double* y = t->ys;
  double* w = t->zs; /* [2][n] */
  thomas_line(n, t->lo, t->di, t->up, t->s, t->cp, t->dp, y, 1);
  for (int i = 0; i < n; ++i) u[i] = 0.0;
  u[0] = a0;
  thomas_line(n, t->lo, t->di, t->up, u, t->cp, t->dp, w, 1);
  u[0] = 0.0;
  u[n - 1] = bn;
  thomas_line(n, t->lo, t->di, t->up, u, t->cp, t->dp, w + n, 1);
  /* M = I + V^T W (2x2), rows: components n-2 and 1 */
  const double m00 = 1.0 + w[n - 2], m01 = w[n + n - 2], m10 = w[1], m11 = 1.0 + w[n + 1];
  const double det = m00 * m11 - m01 * m10;
  const double g0 = (m11 * y[n - 2] - m01 * y[1]) / det, g1 = (m00 * y[1] - m10 * y[n - 2]) / det;
  for (int i = 0; i < n; ++i) x[(size_t)i * stride] = y[i] - (w[i] * g0 + w[n + i] * g1);
}

/* zebra line Gauss-Seidel: dir 0 lines along i (z) for rows j of parity par,
 * dir 1 lines along j (y) for columns i of parity par */
static void tight_lines(tctx* t, tlev* L, int dir, int par) {
  const int nx = L->nx, ny = L->ny, ld = nx + 2;
  if (dir == 0) {
    for (int j = par; j < ny; j += 2) {
      for (int i = 0; i < nx; ++i) {
        size_t p = i + (size_t)nx * j;
        if (TSPEC(L, p)) { t->lo[i] = 0.0; t->di[i] = 1.0; t->up[i] = 0.0; t->s[i] = TE(L, i, j); continue; }
        t->s[i] = L->r[p] - (TC(L, 1, i, j) * TE(L, i + 1, j + 1) + TC(L, 2, i, j) * TE(L, i, j + 1) +
                             TC(L, 3, i, j) * TE(L, i - 1, j + 1) + TC(L, 5, i, j) * TE(L, i - 1, j - 1) +
                             TC(L, 6, i, j) * TE(L, i, j - 1) + TC(L, 7, i, j) * TE(L, i + 1, j - 1));
        t->lo[i] = TC(L, 4, i, j); t->di[i] = TC(L, 8, i, j); t->up[i] = TC(L, 0, i, j);
      }
      if (t->per)
        thomas_periodic(t, nx, &TE(L, 0, j), 1);
      else
        thomas_line(nx, t->lo, t->di, t->up, t->s, t->cp, t->dp, &TE(L, 0, j), 1);
    }
  } else {
    for (int i = par; i < nx; i += 2) {
      for (int j = 0; j < ny; ++j) {
        size_t p = i + (size_t)nx * j;
        if (TSPEC(L, p)) { t->lo[j] = 0.0; t->di[j] = 1.0; t->up[j] = 0.0; t->s[j] = TE(L, i, j); continue; }
        t->s[j] = L->r[p] - (TC(L, 0, i, j) * TE(L, i + 1, j) + TC(L, 1, i, j) * TE(L, i + 1, j + 1) +
                             TC(L, 3, i, j) * TE(L, i - 1, j + 1) + TC(L, 4, i, j) * TE(L, i - 1, j) +
                             TC(L, 5, i, j) * TE(L, i - 1, j - 1) + TC(L, 7, i, j) * TE(L, i + 1, j - 1));
        t->lo[j] = TC(L, 6, i, j); t->di[j] = TC(L, 8, i, j); t->up[j] = TC(L, 2, i, j);
      }
      thomas_line(ny, t->lo, t->di, t->up, t->s, t->cp, t->dp, &TE(L, i, 0), (size_t)ld);
    }
  }
  if (t->per) tight_wrap(L, L->e);
}

static void tight_smooth(tctx* t, tlev* L) {
  tight_lines(t, L, 0, 0); tight_lines(t, L, 0, 1);
  tight_lines(t, L, 1, 0); tight_lines(t, L, 1, 1);
}

/* full weighting (1/4,1/2,1/4 per direction, zero outside the grid) of the
 * fine residual onto the coarse points; a direction that is not coarsened
 * keeps its index; periodic ends take the virtual columns -1 -> nx-2 and
 * nx -> 1 */
static void tight_restrict(const tctx* t, const tlev* F, const double* res, tlev* C) {
  const int sx = C->nx < F->nx ? 2 : 1, sy = C->ny < F->ny ? 2 : 1;
  static const double w[3] = {0.25, 0.5, 0.25};
  for (int jc = 0; jc < C->ny; ++jc)
    for (int ic = 0; ic < C->nx; ++ic) {
      size_t pc = ic + (size_t)C->nx * jc;
      if (TSPEC(C, pc)) { C->r[pc] = 0.0; continue; }
      int i = sx * ic, j = sy * jc;
      double acc = 0.0;
      for (int b = -1; b <= 1; ++b)
        for (int a = -1; a <= 1; ++a) {
          int ii = i + a, jj = j + b;
          if (t->per) ii = ii < 0 ? F->nx - 2 : ii >= F->nx ? 1 : ii;
          if (ii < 0 || ii >= F->nx || jj < 0 || jj >= F->ny) continue;
          acc += (w[a + 1] * w[b + 1]) * res[ii + (size_t)F->nx * jj];
        }
      C->r[pc] = acc;
    }
}

/* e_fine += bilinear interpolation of e_coarse (not at specified points) */
static void tight_prolong(const tctx* t, const tlev* C, tlev* F) {
  const int sx = C->nx < F->nx ? 2 : 1, sy = C->ny < F->ny ? 2 : 1;
  for (int j = 0; j < F->ny; ++j)
    for (int i = 0; i < F->nx; ++i) {
      if (TSPEC(F, i + (size_t)F->nx * j)) continue;
      int i0 = i / sx, i1 = (i + sx - 1) / sx, j0 = j / sy, j1 = (j + sy - 1) / sy;
      double v = 0.25 * ((TE(C, i0, j0) + TE(C, i1, j0)) + (TE(C, i0, j1) + TE(C, i1, j1)));
      TE(F, i, j) += v;
    }
  if (t->per) tight_wrap(F, F->e);
}

static void tight_vcycle(tctx* t, int l) {
  tlev* L = &t->lev[l];
  if (l == 1) {
    for (int s = 0; s < 20; ++s) tight_smooth(t, L);
    return;
  }
  tight_smooth(t, L);
  tight_residual(L, L->e, L->r, L->res);
  tlev* C = &t->lev[l - 1];
  tight_restrict(t, L, L->res, C);
  memset(C->e, 0, (size_t)(C->nx + 2) * (C->ny + 2) * sizeof(double));
  tight_vcycle(t, l - 1);
  tight_prolong(t, C, L);
  tight_smooth(t, L);
}

/* Solves the consistent system for one real right-hand side (MUDPACK's
 * convention: rhs = -RHS_TMP); returns the number of V cycles. */
static int tight_solve(tctx* t, mg_ctx* c, const ora* o, double kx2, const double* rhs_tmp, double* phiout) {
  tight_discretize(t, c, o, kx2);
  tlev* F = &t->lev[t->ngrid];
  const int nx = F->nx, ny = F->ny;
  const size_t n = (size_t)nx * ny;
  double* b = dalloc(n);
  double* res = dalloc(n);
  for (size_t p = 0; p < n; ++p) b[p] = TSPEC(F, p) ? 0.0 : -rhs_tmp[p];
  memset(t->phi, 0, (size_t)(nx + 2) * (ny + 2) * sizeof(double));
  int it = 0;
  const int maxit = o->p.proj_maxit > 0 ? o->p.proj_maxit : 50;
  for (;;) {
    tight_residual(F, t->phi, b, res);
    double rmax = 0.0;
    for (size_t p = 0; p < n; ++p) rmax = fmax(rmax, fabs(res[p]));
    if (rmax <= o->p.proj_tol || it >= maxit) break;
    memcpy(F->r, res, n * sizeof(double));
    memset(F->e, 0, (size_t)(nx + 2) * (ny + 2) * sizeof(double));
    tight_vcycle(t, t->ngrid);
    for (size_t q = 0; q < (size_t)(nx + 2) * (ny + 2); ++q) t->phi[q] += F->e[q];
    ++it;
  }
  for (int j = 0; j < ny; ++j)
    for (int i = 0; i < nx; ++i) phiout[i + (size_t)nx * j] = t->phi[(i + 1) + (size_t)(nx + 2) * (j + 1)];
  free(b);
  free(res);
  return it;
}
#undef TE
#undef TC

int ora_tight_solve(ora* o, double kx2, const double* rhs_tmp, double* phi) {
  mg_ctx c;
  tctx t;
  mg_setup(&c, o);
  tight_setup(&t, &c);
  int it = tight_solve(&t, &c, o, kx2, rhs_tmp, phi);
  tight_free(&t);
  mg_free(&c);
  return it;
}

/* Fine-level operator of proj_mode 1 for kx2: cf[9][NY+2][NZ+2]. */
int ora_tight_operator(ora* o, double kx2, double* cf) {
  mg_ctx c;
  tctx t;
  mg_setup(&c, o);
  tight_setup(&t, &c);
  tight_discretize(&t, &c, o, kx2);
  tlev* F = &t.lev[t.ngrid];
  memcpy(cf, F->cf, 9 * (size_t)F->nx * F->ny * sizeof(double));
  tight_free(&t);
  mg_free(&c);
  return EKM_OK;
}

int ora_mg_ngrid(ora* o) {
  int a = o->p.mg_iex, b = o->p.mg_jey;
  return a > b ? a : b;
}

int ora_mudpack(ora* o, double kx2, const double* rhs_tmp, double* phi) {
  mg_ctx c;
  mg_setup(&c, o);
  int cyc = mg_solve(&c, kx2, rhs_tmp, phi);
  mg_free(&c);
  return cyc;
}

int ora_mg_discretize(ora* o, double kx2, int lev, double* cf, int* nx, int* ny) {
  mg_ctx c;
  mg_setup(&c, o);
  if (lev < 1 || lev > c.ngrid) { mg_free(&c); return EKM_ERR_ARG; }
  c.kx2 = kx2;
  mg_discretize(&c, lev);
  *nx = c.lev[lev].nx;
  *ny = c.lev[lev].ny;
  memcpy(cf, c.lev[lev].cf, sizeof(double) * (size_t)(*nx) * (*ny) * 10);
  int n = (*nx) * (*ny);
  mg_free(&c);
  return n;
}

/* callbacks for the reference MUDPACK build (oracle/_ref) */
static const ora* cb_o;
static double cb_kx2;
void ora_cb_bind(ora* o, double kx2) { cb_o = o; cb_kx2 = kx2; }
void ora_cofcr_cb(const double* x, const double* y, double* cxx, double* cxy, double* cyy,
                  double* cx, double* cy, double* ce) {
  cofcr(cb_o, cb_kx2, *x, *y, cxx, cxy, cyy, cx, cy, ce);
}
void ora_bndcr_cb(const int* kbdy, const double* xory, double* alfa, double* beta,
                  double* gama, double* gbdy) {
  bndyc(*kbdy, *xory, alfa, beta, gama, gbdy);
}

/* ------------------------------------------------------------------------ */
/* set-up: INIT_DUCT index ranges (src/duct.F90:7-232), RK constants         */
/* (src/diablo.F90:568-576), wavenumbers (INIT_FFT, called at :563)          */
/* ------------------------------------------------------------------------ */
static int supported(const ekm_params* p) {
  for (int c = 0; c < 3; ++c)
    for (int s = 0; s < 4; ++s) {
      int t = p->u_bc[c][s];
      /* type 3 (wall-model rows) on the y ends, and on the z-max end for V, W */
      if (t != 0 && t != 1 && t != 6 && !(t == 3 && (s == EKM_YMIN || s == EKM_YMAX)) &&
          !(t == 3 && s == EKM_ZMAX && c > 0))
        return 0;
    }
  if (p->u_bc[2][EKM_YMIN] == 3) {
    /* WALL_MODEL_TYPE 3 only; its z means need ZSTART = 1 and ZEND = NZ */
    int uz0 = p->u_bc[0][EKM_ZMIN], uz1 = p->u_bc[0][EKM_ZMAX];
    if (p->wall_model_type != 3) return 0;
    if (!(uz0 == 1 || uz0 == 6) || !(uz1 == 0 || uz1 == 1 || uz1 == 6)) return 0;
  }
  if (p->n_th) {
    int t;
    t = p->th_bc[EKM_YMIN]; if (t != 0 && t != 1 && t != 5 && t != 6) return 0;
    t = p->th_bc[EKM_YMAX]; if (t != 0 && t != 1 && t != 6) return 0;
    t = p->th_bc[EKM_ZMIN]; if (t != 0 && t != 1 && t != 6) return 0;
    t = p->th_bc[EKM_ZMAX]; if (t != 0 && t != 1 && t != 5 && t != 6) return 0;
  }
  if (p->f_type < 0 || p->f_type > 3) return 0; /* 3: no mean forcing */
  for (int s = 0; s < 4; ++s) if (p->mudpack_bc[s] < 0 || p->mudpack_bc[s] > 2) return 0;
  if ((p->mudpack_bc[0] == 0) != (p->mudpack_bc[1] == 0)) return 0;
  if ((p->mudpack_bc[2] == 0) != (p->mudpack_bc[3] == 0)) return 0;
  if (p->nz + 2 != p->mg_ixp * (1 << (p->mg_iex - 1)) + 1) return 0;
  if (p->ny + 2 != p->mg_jyq * (1 << (p->mg_jey - 1)) + 1) return 0;
  if (p->nx < 8 || (p->nx & (p->nx - 1))) return 0;
  if (p->proj_mode < 0 || p->proj_mode > 1) return 0;
  if (p->les_model != 0 && p->les_model != 2) return 0;
  /* the scalar model with a non-uniform background (cont_strat 0) adds
   * KAPPA_T indexed by Fourier mode to the background diffusion
   * (src/duct.F90:610-618); not restated */
  if (p->les_model && p->n_th && !p->cont_strat) return 0;
  /* sponge: the vertical profile only (the streamwise one, NZ_S_R < NZ+1,
   * also scales the scalar's diffusion and is not restated) */
  if (p->sponge_type < 0 || p->sponge_type > 2) return 0;
  if (p->sponge_type && (p->ny_s_t < 0 || p->ny_s_t == p->ny + 1 || p->save_stats_int < 1)) return 0;
  if (p->sponge_type && p->nz_s_r < p->nz + 1 && (p->nz_s_r < 0 || p->nz_s_l < 0 || p->nz_s_l > p->nz + 1)) return 0;
  /* the consistent projection: periodic z ends are cyclic lines, periodic
   * y ends are not handled */
  if (p->proj_mode == 1 && (p->mudpack_bc[2] == 0 || !(p->proj_tol > 0.0))) return 0;
  return 1;
}

static void init_ranges(ora* o) {
  const ekm_params* p = &o->p;
  int NY = o->NY, NZ = o->NZ;
  int wy0 = p->u_bc[2][EKM_YMIN], wy1 = p->u_bc[2][EKM_YMAX];
  int uz0 = p->u_bc[0][EKM_ZMIN], uz1 = p->u_bc[0][EKM_ZMAX];
  o->jstart = (wy0 == 0 || wy0 == 1 || wy0 == 6 || wy0 == 7 || wy0 == 3) ? 1 : 2;
  o->zstart = (uz0 == 1 || uz0 == 6) ? 1 : 2;
  o->jend = (wy1 == 0 || wy1 == 1 || wy1 == 6) ? NY : NY - 1;
  o->zend = (uz1 == 0 || uz1 == 1 || uz1 == 6) ? NZ : NZ - 1; /* IC_TYPE 7 */
  int ty0 = p->th_bc[EKM_YMIN], ty1 = p->th_bc[EKM_YMAX];
  int tz0 = p->th_bc[EKM_ZMIN], tz1 = p->th_bc[EKM_ZMAX];
  o->jstart_th = (ty0 == 0 || ty0 == 1 || ty0 == 5 || ty0 == 2 || ty0 == 6 || ty0 == 7 || ty0 == 8) ? 1 : 2;
  o->zstart_th = (tz0 == 0 || tz0 == 1 || tz0 == 6) ? 1 : 2;
  o->jend_th = (ty1 == 0 || ty1 == 1 || ty1 == 6 || ty1 == 5) ? NY : NY - 1;
  o->zend_th = (tz1 == 0 || tz1 == 1 || tz1 == 5 || tz1 == 6) ? NZ : NZ - 1;
}

static void set_hbar(ora* o) {
  o->h_bar[0] = o->delta_t * (8.0 / 15.0);
  o->h_bar[1] = o->delta_t * (2.0 / 15.0);
  o->h_bar[2] = o->delta_t * (5.0 / 15.0);
}

static void sponge_sigma(ora* o);

ora* ora_create(const ekm_params* p, const double* xpoint, const double* ypoint,
                const double* th_bar) {
  if (!supported(p)) return NULL;
  ora* o = (ora*)calloc(1, sizeof(ora));
  o->p = *p;
  o->NX = p->nx; o->NY = p->ny; o->NZ = p->nz;
  o->NXC = p->nx / 2 + 1; o->NKX = p->nx / 3; o->NX2 = p->nx + 2;
  o->NY2 = p->ny + 2; o->NZ2 = p->nz + 2; o->NZ3 = p->nz + 3; o->NY3 = p->ny + 3;
  o->npts = (size_t)o->NX2 * o->NZ2 * o->NY2;
  size_t pl = (size_t)o->NZ2 * o->NY2, ml = (size_t)o->NZ3 * o->NY3;
  o->xp = dalloc(pl); o->yp = dalloc(pl); o->thbar = dalloc(pl);
  memcpy(o->xp, xpoint, pl * sizeof(double));
  memcpy(o->yp, ypoint, pl * sizeof(double));
  for (int q = 0; q < NMET; ++q) o->met[q] = dalloc(ml * 2);
  o->jac = dalloc(ml);
  metrics(o);
  /* TH_BAR (src/diablo.F90:1152-1160) */
  for (int j = 0; j < o->NY2; ++j)
    for (int k = 0; k < o->NZ2; ++k)
      o->thbar[PL(o, k, j)] = p->cont_strat ? (-o->yp[PL(o, k, j)] + o->yp[PL(o, 0, o->NY + 1)])
                                            : (th_bar ? th_bar[PL(o, k, j)] : 0.0);
  o->kx = dalloc(o->NXC); o->kx2 = dalloc(o->NXC);
  for (int m = 0; m < o->NXC; ++m) {
    o->kx[m] = 2.0 * M_PI * m / p->lx;
    o->kx2[m] = o->kx[m] * o->kx[m];
  }
  double** f[] = {&o->U[0], &o->U[1], &o->U[2], &o->P, &o->TH, &o->CF[0], &o->CF[1], &o->CF[2],
                  &o->CFTH, &o->R[0], &o->R[1], &o->R[2], &o->RTH, &o->S1, &o->S2, &o->U2b,
                  &o->U3b, &o->CS1, &o->PHI};
  for (size_t q = 0; q < sizeof(f) / sizeof(f[0]); ++q) *f[q] = dalloc(o->npts);
  if (p->les_model) {
    for (int q = 0; q < 6; ++q) o->Sij[q] = dalloc(o->npts);
    for (int q = 0; q < 3; ++q) o->UBT[q] = dalloc(o->npts);
    o->S2B = dalloc(o->npts);
    o->KAPT = dalloc(o->npts);
    o->CDYNT = dalloc(pl);
    o->SABS = dalloc(o->npts);
    o->NUT = dalloc(o->npts);
    o->DELTA = dalloc(pl);
    o->CDYN = dalloc(pl);
    /* filter width at cell centres (les_chan.F90:343-348, beta = 1) */
    for (int j = 0; j <= o->NY + 1; ++j)
      for (int k = 0; k <= o->NZ + 1; ++k) o->DELTA[PL(o, k, j)] = pow(p->lx / (double)o->NX * JAC(o, k, j), 1.0 / 3.0);
  }
  o->create_new_flow = 1;
  o->first_time = 1;
  o->px0 = p->px0;
  o->qn = p->wbulk0;
  o->sigma = dalloc((size_t)o->NY + 2);
  if (p->u_bc[0][EKM_YMIN] == 3 || p->u_bc[1][EKM_YMIN] == 3 || p->u_bc[2][EKM_YMIN] == 3) {
    o->wall = dalloc(2 * (size_t)o->NX2 * o->NZ2);
    o->wall_z = o->yp[PL(o, 2, 2)] - o->yp[PL(o, 2, 1)]; /* DYY(2,2), src/diablo.F90:1140 */
  }
  o->sigma_out = dalloc((size_t)o->NZ + 2);
  sponge_sigma(o);
  init_ranges(o);
  o->delta_t = p->delta_t;
  o->delta_t_set = p->delta_t;
  set_hbar(o);
  o->beta_bar[0] = 1.0; o->beta_bar[1] = 25.0 / 8.0; o->beta_bar[2] = 9.0 / 4.0;
  o->zeta_bar[0] = 0.0; o->zeta_bar[1] = -17.0 / 8.0; o->zeta_bar[2] = -5.0 / 4.0;
  o->rk_step = 1;
  return o;
}

void ora_destroy(ora* o) {
  if (!o) return;
  free(o->xp); free(o->yp); free(o->thbar); free(o->jac); free(o->kx); free(o->kx2); free(o->sigma); free(o->sigma_out); free(o->wall);
  for (int q = 0; q < NMET; ++q) free(o->met[q]);
  double* f[] = {o->U[0], o->U[1], o->U[2], o->P, o->TH, o->CF[0], o->CF[1], o->CF[2], o->CFTH,
                 o->R[0], o->R[1], o->R[2], o->RTH, o->S1, o->S2, o->U2b, o->U3b, o->CS1, o->PHI};
  for (size_t q = 0; q < sizeof(f) / sizeof(f[0]); ++q) free(f[q]);
  for (int q = 0; q < 6; ++q) free(o->Sij[q]);
  for (int q = 0; q < 3; ++q) free(o->UBT[q]);
  free(o->SABS); free(o->NUT); free(o->DELTA); free(o->CDYN);
  free(o->S2B); free(o->KAPT); free(o->CDYNT);
  free(o);
}

/* ------------------------------------------------------------------------ */
/* LES_WALL_MODEL, WALL_MODEL_TYPE 3 (src/les_chan.F90:1914-2144) for one   */
/* rank holding every x point (MIN(NXP,NXP_L) = NX-1), on the physical U1,  */
/* U3 entering the implicit solves; loop for loop                           */
/* ------------------------------------------------------------------------ */
static void wall_model(ora* o, const double* U1, const double* U3, int first) {
  const int NX = o->NX, NZ = o->NZ, js = o->jstart;
  const size_t pl = (size_t)o->NX2 * o->NZ2;
  double *wu = o->wall, *ww = o->wall + pl;
  if (first) { /* TIME_STEP = 1: W_BC = U3X(I,K,2), U_BC kept (:2112-2118) */
    for (int k = o->zstart; k <= o->zend; ++k)
      for (int i = 0; i < NX; ++i) ww[(size_t)k * o->NX2 + i] = U3[IDX(o, i, k, 2)];
    return;
  }
  const double KAPA = 0.41, NU_tmp = o->p.nu, Z = o->wall_z;
  double* tmp_w = dalloc((size_t)NZ + 1);
  double* tmp_u = dalloc((size_t)NZ + 1);
  for (int k = o->zstart; k <= o->zend; ++k) { /* :1996-1999 */
    double sw = 0.0, su = 0.0;
    for (int i = 0; i < NX; ++i) { sw += U3[IDX(o, i, k, js)]; su += U1[IDX(o, i, k, js)]; }
    tmp_w[k] = sw / (double)(NX - 1);
    tmp_u[k] = su / (double)(NX - 1);
  }
  double su = 0.0, sw = 0.0;
  for (int k = 1; k <= NZ; ++k) { su += tmp_u[k]; sw += tmp_w[k]; }
  const double U_1 = su / (double)NZ, W_1 = sw / (double)NZ;
  const double M = sqrt(U_1 * U_1 + W_1 * W_1);
  double U_STAR = M;
  for (int it = 1; it <= 30; ++it) U_STAR = M / (1.0 / KAPA * log(Z / NU_tmp * U_STAR) + (double)5.2f);
  const float alpha = 20.6f; /* implicitly typed default real */
  const double TAO_WALL = U_STAR * U_STAR * (double)cosf(alpha * 3.1415f / 180.0f);
  const double TAO_WALL_1 = U_STAR * U_STAR * (double)sinf(alpha * 3.1415f / 180.0f);
  for (int k = o->zstart; k <= o->zend; ++k)
    for (int i = 0; i < NX; ++i) { /* :2099-2102 */
      const double den = 0.5 * (C22(o, k, 1, 2) + C22(o, k + 1, 1, 1)), jk = JAC(o, k, 1);
      const double u = U1[IDX(o, i, k, js)], w = U3[IDX(o, i, k, js)];
      ww[(size_t)k * o->NX2 + i] = w / W_1 * (TAO_WALL / NU_tmp) / den * jk;
      wu[(size_t)k * o->NX2 + i] = fmax(u / W_1 * (TAO_WALL / NU_tmp), u / U_1 * (TAO_WALL_1 / NU_tmp)) / den * jk;
    }
  o->wall_info[0] = U_1; o->wall_info[1] = W_1; o->wall_info[2] = U_STAR;
  o->wall_info[3] = TAO_WALL; o->wall_info[4] = TAO_WALL_1;
  free(tmp_w); free(tmp_u);
}

/* the model on given physical U1, U3 ((NX+2)(NZ+2)(NY+2) each): the row
 * values [2][NZ+2][NX+2] into out, U_1, W_1, U_STAR, TAO_WALL, TAO_WALL_1
 * into info (5) */
int ora_wall_model(ora* o, const double* u1, const double* u3, int first, double* out, double* info) {
  if (!o->wall) return EKM_ERR_ARG;
  wall_model(o, u1, u3, first);
  memcpy(out, o->wall, 2 * (size_t)o->NX2 * o->NZ2 * sizeof(double));
  memcpy(info, o->wall_info, sizeof(o->wall_info));
  return EKM_OK;
}

/* ------------------------------------------------------------------------ */
/* ADI boundary rows: APPLY_BC_{1,2,3,TH}_{LOWER,UPPER,LEFT,RIGHT}           */
/* (src/boundary.F90:124-1760) for IC_TYPE = 7                               */
/* ------------------------------------------------------------------------ */
typedef struct sysb { double *a, *b, *c, *g; size_t ld; } sysb;

static void row(sysb* s, int r, double a, double b, double c) {
  for (size_t i = 0; i < s->ld; ++i) {
    s->a[i + s->ld * r] = a; s->b[i + s->ld * r] = b; s->c[i + s->ld * r] = c;
  }
}
static void rhsc(sysb* s, int r, double v) { for (size_t i = 0; i < s->ld; ++i) s->g[i + s->ld * r] = v; }
/* lagged "periodic" value: the field on the same line at index `src` */
static void rhs_line(sysb* s, int r, const double* u, size_t off, size_t stride, int src) {
  for (size_t i = 0; i < s->ld; ++i) s->g[i + s->ld * r] = u[off + i + stride * (size_t)src];
}

/* y-direction rows for line K.  comp: 0..2 velocity, 3 scalar. off/stride
 * address U(I,K,j) = u[off + I + stride*j]. */
static void bc_y(const ora* o, int comp, int K, sysb* s, const double* u, size_t off, size_t stride) {
  const ekm_params* p = &o->p;
  int NY = o->NY;
  int lo = comp < 3 ? p->u_bc[comp][EKM_YMIN] : p->th_bc[EKM_YMIN];
  int hi = comp < 3 ? p->u_bc[comp][EKM_YMAX] : p->th_bc[EKM_YMAX];
  double c0 = comp < 3 ? p->u_bc_c1[comp][EKM_YMIN] : p->th_bc_c1[EKM_YMIN];
  double c1 = comp < 3 ? p->u_bc_c1[comp][EKM_YMAX] : p->th_bc_c1[EKM_YMAX];
  /* LOWER */
  if (lo == 3) {
    /* IC_TYPE 7, WALL_MODEL_TYPE 3 (:532-578, :894-937, :1338-1380):
     * U(1) - U(0) = U_BC / W_BC_LOWER_WALLMODEL(I,K); V's array is never
     * assigned by the model (zero) */
    row(s, 0, 0., -1., 1.);
    for (size_t i = 0; i < s->ld; ++i)
      s->g[i] = comp == 1 ? 0.0 : o->wall[(comp == 0 ? 0 : (size_t)o->NX2 * o->NZ2) + (size_t)K * o->NX2 + i];
  } else if (lo == 0) { row(s, 0, 0., 1., 0.); rhsc(s, 0, c0); }
  else if (lo == 1 || lo == 5) {
    if (comp == 3) { /* :163-178 */
      row(s, 0, 0., -1., 1.);
      rhsc(s, 0, c0 / (0.5 * (C22(o, K, 1, 2) + C22(o, K + 1, 1, 1))) * JAC(o, K, 1));
    } else if (comp == 2) { row(s, 0, 0., -1., 1.); rhsc(s, 0, c0); } /* :1338-1345 */
    else { row(s, 1, 0., -1., 1.); rhsc(s, 1, c0); }                  /* :500-506, :868-874 */
  } else { row(s, 0, 0., 1., 0.); rhs_line(s, 0, u, off, stride, NY); }
  /* UPPER */
  if (hi == 3) {
    /* :635-650, :1016-1030, :1480-1493; U_BC_UPPER_NWM and W_BC_UPPER_NWM
     * are allocated and never assigned by the reference: zero here */
    if (comp == 1) {
      row(s, NY + 1, 0., 1., 0.); rhsc(s, NY + 1, 0.0);
      row(s, NY, 0., 1., 0.); rhsc(s, NY, 0.0);
    } else {
      row(s, NY, -1., 1., 0.); rhsc(s, NY, 0.0);
      row(s, NY + 1, 0., 1., 0.); rhs_line(s, NY + 1, u, off, stride, NY);
    }
  } else if (hi == 0) {
    row(s, NY + 1, 0., 1., 0.); rhsc(s, NY + 1, c1);
    if (comp == 1 || comp == 2) { row(s, NY, 0., 1., 0.); rhsc(s, NY, c1); } /* :951-960, :1410-1419 */
  } else if (hi == 1) {
    if (comp == 3) { /* :333-346 */
      row(s, NY, -1., 1., 0.);
      rhsc(s, NY, 0.5 * c1 * (JAC(o, K, NY) + JAC(o, K, NY - 1)) / C22(o, K, NY - 1, 1));
      row(s, NY + 1, -1., 1., 0.);
      rhsc(s, NY + 1, 0.5 * c1 * (JAC(o, K, NY + 1) + JAC(o, K, NY)) / C22(o, K, NY, 1));
    } else if (comp == 1) { row(s, NY, -1., 1., 0.); rhsc(s, NY, c1); } /* :962-968 */
    else { row(s, NY + 1, -1., 1., 0.); rhsc(s, NY + 1, c1); }
  } else { row(s, NY + 1, 0., 1., 0.); rhs_line(s, NY + 1, u, off, stride, 1); }
}

/* z-direction rows for line J */
static void bc_z(const ora* o, int comp, int J, sysb* s, const double* u, size_t off, size_t stride) {
  const ekm_params* p = &o->p;
  int NZ = o->NZ;
  int lo = comp < 3 ? p->u_bc[comp][EKM_ZMIN] : p->th_bc[EKM_ZMIN];
  int hi = comp < 3 ? p->u_bc[comp][EKM_ZMAX] : p->th_bc[EKM_ZMAX];
  double c0 = comp < 3 ? p->u_bc_c1[comp][EKM_ZMIN] : p->th_bc_c1[EKM_ZMIN];
  double c1 = comp < 3 ? p->u_bc_c1[comp][EKM_ZMAX] : p->th_bc_c1[EKM_ZMAX];
  (void)J;
  /* LEFT */
  if (lo == 0) {
    /* U1: C1 on rows 0,1 (:668-683); U2,U3 (IC_TYPE 7): U2_BAR/U3_BAR = 0
     * (:1062-1075, :1540-1549; VEL_PROFILE sets them to zero with Q = 0,
     * src/duct.F90:3699-3731); TH: 0 then C1 (:374-384) */
    double v0 = comp == 0 ? c0 : 0.0, v1 = (comp == 0 || comp == 3) ? c0 : 0.0;
    row(s, 0, 0., 1., 0.); rhsc(s, 0, v0);
    row(s, 1, 0., 1., 0.); rhsc(s, 1, v1);
  } else if (lo == 1) { row(s, 0, 0., -1., 1.); rhsc(s, 0, c0); }
  else { row(s, 0, 0., 1., 0.); rhs_line(s, 0, u, off, stride, NZ); }
  /* RIGHT */
  if (hi == 3) { /* V :1249-1262, W :1722-1734 (W_BC_UPPER_NWM never assigned: zero) */
    if (comp == 1) {
      row(s, NZ, -1., 1., 0.); rhsc(s, NZ, 0.0);
      row(s, NZ + 1, 0., 1., 0.); rhs_line(s, NZ + 1, u, off, stride, NZ);
    } else {
      row(s, NZ + 1, -1., 1., 0.); rhsc(s, NZ + 1, 0.0);
      row(s, NZ, 0., 1., 0.); rhsc(s, NZ, 0.0);
    }
  } else if (hi == 0) {
    double vtop = comp == 2 ? c1 : 0.0; /* U3: C1 on NZ+1 and NZ (:1623-1632) */
    row(s, NZ + 1, 0., 1., 0.); rhsc(s, NZ + 1, vtop);
    row(s, NZ, 0., 1., 0.); rhsc(s, NZ, c1);
  } else if (hi == 1) {
    row(s, NZ + 1, -1., 1., 0.); rhsc(s, NZ + 1, c1);
    if (comp == 0) { row(s, NZ + 1, 0., 1., 0.); rhsc(s, NZ + 1, 0.0); } /* :792-797 */
  } else if (hi == 5) { row(s, NZ + 1, -2., 1., 0.); rhsc(s, NZ + 1, 0.0); }   /* :443-449 */
  else { row(s, NZ + 1, 0., 1., 0.); rhs_line(s, NZ + 1, u, off, stride, 1); }
}

/* the boundary rows of one line as the ADI sweeps assemble them, for tests:
 * dir 0 the y line at K = line, dir 1 the z line at J = line; u the
 * physical field the lagged rows read; a, b, c, g of len x NX2 entries
 * (row-major by line row), prefilled by the caller, only the boundary rows
 * written */
int ora_bc_rows(ora* o, int comp, int dir, int line, const double* u, double* a, double* b, double* c, double* g) {
  if (comp < 0 || comp > 3 || dir < 0 || dir > 1) return EKM_ERR_ARG;
  sysb s = {a, b, c, g, (size_t)o->NX2};
  if (dir == 0) bc_y(o, comp, line, &s, u, IDX(o, 0, line, 0), (size_t)o->NX2 * o->NZ2);
  else bc_z(o, comp, line, &s, u, IDX(o, 0, 0, line), (size_t)o->NX2);
  return EKM_OK;
}

/* ------------------------------------------------------------------------ */
/* ADI for one field (theta: :1334-1625; U1: :1640-1835; U2: :1877-2110;     */
/* U3: :2137-2391)                                                           */
/* ------------------------------------------------------------------------ */
/* the z- and y-direction metrics of the subgrid viscosity of velocity
 * component c: GMAT_11 / GMAT_22 for U1 (:1667-1731), the _y variants for U2
 * (:1916-1928, :1963-1975), the _z variants for U3 (:2177-2236) */
static const double* les_gz(const ora* o, int c) { return o->met[c == 0 || c == 3 ? 0 : c == 1 ? 10 : 7]; }
static const double* les_gy(const ora* o, int c) { return o->met[c == 0 || c == 3 ? 2 : c == 1 ? 12 : 9]; }

/* nut != NULL adds the subgrid viscosity (:1667-1680, :1718-1731,
 * :1771-1784, :1819-1832 and their U2, U3 analogues) with weight t2 =
 * H_BAR/2: GZ is the z-direction metric (GMAT_11, GMAT_11_y, GMAT_11_z for
 * U1, U2, U3), GY the y-direction one (GMAT_22, GMAT_22_y, GMAT_22_z). */
static void adi(ora* o, int comp, double* U, double* R, double t1, const double* nut, const double* GZ,
                const double* GY, double t2) {
  int NX2 = o->NX2, NY = o->NY, NZ = o->NZ;
  int js = comp < 3 ? o->jstart : o->jstart_th, je = comp < 3 ? o->jend : o->jend_th;
  int zs = comp < 3 ? o->zstart : o->zstart_th, ze = comp < 3 ? o->zend : o->zend_th;
  double* S1 = o->S1;
#pragma omp parallel for schedule(static)
  for (int j = js; j <= je; ++j)
    for (int k = zs; k <= ze; ++k)
      for (int i = 0; i < NX2; ++i) {
        size_t q = IDX(o, i, k, j);
        S1[q] = R[q] - JAC(o, k, j) * U[q];
      }
#pragma omp parallel for schedule(static)
  for (int j = js; j <= je; ++j)
    for (int k = zs; k <= ze; ++k)
      for (int i = 0; i < NX2; ++i) {
        size_t q = IDX(o, i, k, j), dk = NX2;
        R[q] = 0.5 * S1[q] + JAC(o, k, j) * U[q] +
               t1 * (G11(o, k + 1, j, 2) * (U[q + dk] - U[q]) - G11(o, k, j, 2) * (U[q] - U[q - dk]));
        if (nut)
          R[q] = R[q] + t2 * (0.5 * (nut[q + dk] + nut[q]) * (GZ[MF(o, k + 1, j, 2)] * (U[q + dk] - U[q])) -
                              0.5 * (nut[q] + nut[q - dk]) * (GZ[MF(o, k, j, 2)] * (U[q] - U[q - dk])));
      }
  /* y sweeps, lines K */
  size_t sy = (size_t)NX2 * o->NZ2;
#pragma omp parallel
  {
    size_t n = (size_t)NX2 * (NY + 2);
    sysb s = {dalloc(n), dalloc(n), dalloc(n), dalloc(n), (size_t)NX2};
    for (size_t t = 0; t < n; ++t) s.b[t] = 1.0;
#pragma omp for schedule(static)
    for (int k = zs; k <= ze; ++k) {
      for (int j = js; j <= je; ++j) {
        double l = -(t1 * G22(o, k, j, 1)), d = JAC(o, k, j) + t1 * (G22(o, k, j + 1, 1) + G22(o, k, j, 1));
        double u = -(t1 * G22(o, k, j + 1, 1));
        for (int i = 0; i < NX2; ++i) {
          size_t r = i + (size_t)NX2 * j;
          s.a[r] = l; s.b[r] = d; s.c[r] = u; s.g[r] = R[IDX(o, i, k, j)];
        }
        if (nut)
          for (int i = 0; i < NX2; ++i) {
            size_t r = i + (size_t)NX2 * j, q = IDX(o, i, k, j);
            double wm = t2 * 0.5 * (nut[q] + nut[q - sy]) * GY[MF(o, k, j, 1)];
            double wp = t2 * 0.5 * (nut[q] + nut[q + sy]) * GY[MF(o, k, j + 1, 1)];
            s.a[r] = s.a[r] - wm;
            s.b[r] = s.b[r] + wp + wm;
            s.c[r] = s.c[r] - wp;
          }
      }
      bc_y(o, comp, k, &s, U, IDX(o, 0, k, 0), sy);
      ora_thomas_real(s.a, s.b, s.c, s.g, NY + 1, NX2 - 1);
      for (int j = 0; j <= NY + 1; ++j)
        for (int i = 0; i < NX2; ++i) U[IDX(o, i, k, j)] = s.g[i + (size_t)NX2 * j];
    }
    free(s.a); free(s.b); free(s.c); free(s.g);
  }
#pragma omp parallel for schedule(static)
  for (int j = js; j <= je; ++j)
    for (int k = zs; k <= ze; ++k)
      for (int i = 0; i < NX2; ++i) {
        size_t q = IDX(o, i, k, j), dj = sy;
        R[q] = 0.5 * S1[q] + JAC(o, k, j) * U[q] +
               t1 * (G22(o, k, j + 1, 1) * (U[q + dj] - U[q]) - G22(o, k, j, 1) * (U[q] - U[q - dj]));
        if (nut)
          R[q] = R[q] + t2 * (0.5 * (nut[q + dj] + nut[q]) * (GY[MF(o, k, j + 1, 1)] * (U[q + dj] - U[q])) -
                              0.5 * (nut[q - dj] + nut[q]) * (GY[MF(o, k, j, 1)] * (U[q] - U[q - dj])));
      }
  if (comp < 3) /* IC_TYPE 7: R(I,K,NY+1) = R(I,K,NY) (:1787-1793, :2023-2029) */
    for (int k = zs; k <= ze; ++k)
      for (int i = 0; i < NX2; ++i) R[IDX(o, i, k, NY + 1)] = R[IDX(o, i, k, NY)];
  /* z sweeps, lines J */
  int hi = comp < 3 ? o->p.u_bc[comp][EKM_ZMAX] : o->p.th_bc[EKM_ZMAX];
#pragma omp parallel
  {
    size_t n = (size_t)NX2 * (NZ + 2);
    sysb s = {dalloc(n), dalloc(n), dalloc(n), dalloc(n), (size_t)NX2};
    for (size_t t = 0; t < n; ++t) s.b[t] = 1.0;
#pragma omp for schedule(static)
    for (int j = js; j <= je; ++j) {
      for (int k = zs; k <= ze; ++k) {
        double l = -(t1 * G11(o, k, j, 2)), d = JAC(o, k, j) + t1 * (G11(o, k + 1, j, 2) + G11(o, k, j, 2));
        double u = -(t1 * G11(o, k + 1, j, 2));
        for (int i = 0; i < NX2; ++i) {
          size_t r = i + (size_t)NX2 * k;
          s.a[r] = l; s.b[r] = d; s.c[r] = u; s.g[r] = R[IDX(o, i, k, j)];
        }
        if (nut)
          for (int i = 0; i < NX2; ++i) {
            size_t r = i + (size_t)NX2 * k, q = IDX(o, i, k, j), dk = NX2;
            double wm = t2 * 0.5 * (nut[q] + nut[q - dk]) * GZ[MF(o, k, j, 2)];
            double wp = t2 * 0.5 * (nut[q + dk] + nut[q]) * GZ[MF(o, k + 1, j, 2)];
            s.a[r] = s.a[r] - wm;
            s.b[r] = s.b[r] + wp + wm;
            s.c[r] = s.c[r] - wp;
          }
      }
      bc_z(o, comp, j, &s, U, IDX(o, 0, 0, j), (size_t)NX2);
      if (hi == 5) {
        double D = comp == 0 ? G11(o, NZ, j, 2) : 1.0;
        ora_thomas_real_sp(s.a, s.b, s.c, s.g, D, NZ, NZ + 1, NX2 - 1);
      } else {
        ora_thomas_real(s.a, s.b, s.c, s.g, NZ + 1, NX2 - 1);
      }
      for (int k = 0; k <= NZ + 1; ++k)
        for (int i = 0; i < NX2; ++i) U[IDX(o, i, k, j)] = s.g[i + (size_t)NX2 * k];
    }
    free(s.a); free(s.b); free(s.c); free(s.g);
  }
  if (comp == 3) { /* corners (:1606-1622) */
    if (o->p.th_bc[EKM_ZMAX] == 1)
      for (int i = 0; i < NX2; ++i) {
        U[IDX(o, i, NZ + 1, NY + 1)] = U[IDX(o, i, NZ, NY + 1)];
        U[IDX(o, i, NZ + 1, 0)] = U[IDX(o, i, NZ, 0)];
        U[IDX(o, i, 0, NY + 1)] = U[IDX(o, i, 1, NY + 1)];
        U[IDX(o, i, 0, 0)] = U[IDX(o, i, 1, 0)];
      }
    if (o->p.th_bc[EKM_ZMIN] == 6)
      for (int i = 0; i < NX2; ++i) {
        U[IDX(o, i, NZ + 1, NY + 1)] = U[IDX(o, i, 1, NY + 1)];
        U[IDX(o, i, NZ + 1, 0)] = U[IDX(o, i, 1, 0)];
        U[IDX(o, i, 0, NY + 1)] = U[IDX(o, i, NZ, NY + 1)];
        U[IDX(o, i, 0, 0)] = U[IDX(o, i, NZ, 0)];
      }
  }
}

/* One field's ADI on caller arrays (U, R physical, in place), with the
 * subgrid viscosity nut when it is not NULL: the building block the
 * substep runs, exposed so its eddy-viscosity terms can be checked alone. */
int ora_adi_field(ora* o, int comp, double* U, double* R, double t1, const double* nut, double t2) {
  if (comp < 0 || comp > 3 || (comp == 3 && !o->p.n_th)) return EKM_ERR_ARG;
  adi(o, comp, U, R, t1, nut, les_gz(o, comp), les_gy(o, comp), t2);
  return EKM_OK;
}

/* ------------------------------------------------------------------------ */
/* nonlinear fluxes (src/duct.F90:683-1120, :1169-1263)                      */
/* ------------------------------------------------------------------------ */
static void contravariant(ora* o, const double* U2, const double* U3) {
  int NX2 = o->NX2;
#pragma omp parallel for schedule(static)
  for (int j = o->jstart; j <= o->jend + 1; ++j)
    for (int k = o->zstart; k <= o->zend + 1; ++k)
      for (int i = 0; i < NX2; ++i) {
        size_t q = IDX(o, i, k, j), dj = (size_t)NX2 * o->NZ2, dk = NX2;
        o->U2b[q] = 0.5 * C22(o, k, j, 1) * (U2[q] + U2[q - dj]) + 0.5 * C21(o, k, j, 1) * (U3[q] + U3[q - dj]);
        o->U3b[q] = 0.5 * C11(o, k, j, 2) * (U3[q] + U3[q - dk]) + 0.5 * C12(o, k, j, 2) * (U2[q] + U2[q - dk]);
      }
}

/* S = J * A * B on the range, zero elsewhere */
static void xflux(ora* o, double* S, const double* A, const double* B, int js, int je, int zs, int ze) {
  memset(S, 0, o->npts * sizeof(double));
#pragma omp parallel for schedule(static)
  for (int j = js; j <= je; ++j)
    for (int k = zs; k <= ze; ++k)
      for (int i = 0; i < o->NX2; ++i) {
        size_t q = IDX(o, i, k, j);
        S[q] = JAC(o, k, j) * A[q] * B[q];
      }
}

/* S = -(d(U2b q)/deta) + nu d/dzeta(G12 dq/deta) - (d(U3b q)/dzeta) + nu d/deta(G12 dq/dzeta) */
static void yzflux(ora* o, double* S, const double* Q, double nu, int js, int je, int zs, int ze) {
  memset(S, 0, o->npts * sizeof(double));
  int NX2 = o->NX2;
  size_t dj = (size_t)NX2 * o->NZ2, dk = NX2;
  const double *U2b = o->U2b, *U3b = o->U3b;
#pragma omp parallel for schedule(static)
  for (int j = js; j <= je; ++j)
    for (int k = zs; k <= ze; ++k)
      for (int i = 0; i < NX2; ++i) {
        size_t q = IDX(o, i, k, j);
        double s1 = 0.5 * U2b[q + dj] * (Q[q] + Q[q + dj]) - 0.5 * U2b[q] * (Q[q] + Q[q - dj]);
        double s2 = 0.5 * U3b[q + dk] * (Q[q] + Q[q + dk]) - 0.5 * U3b[q] * (Q[q] + Q[q - dk]);
        s1 = -s1 + 0.25 * nu *
                       (G12(o, k + 1, j, 2) * (Q[q + dk + dj] + Q[q + dj] - Q[q - dj] - Q[q + dk - dj]) -
                        G12(o, k, j, 2) * (Q[q - dk + dj] + Q[q + dj] - Q[q - dj] - Q[q - dk - dj]));
        s2 = -s2 + 0.25 * nu *
                       (G12(o, k, j + 1, 1) * (Q[q + dk + dj] + Q[q + dk] - Q[q - dk + dj] - Q[q - dk]) -
                        G12(o, k, j, 1) * (Q[q + dk] + Q[q + dk - dj] - Q[q - dk] - Q[q - dk - dj]));
        S[q] = s1 + s2;
      }
}

/* Eddy-viscosity cross terms of the subgrid model (:990-1120): with the
 * face averages of NU_T, S = 1/4 [ GL(K+1) nu(K+1/2) (Q_eta at K+1/2)
 * - GL(K) nu(K-1/2) (Q_eta at K-1/2) + GL(J+1) nu(J+1/2) (Q_zeta at J+1/2)
 * - GL(J) nu(J-1/2) (Q_zeta at J-1/2) ]; GL is GMAT_12 for U1, GMAT_12_y for
 * U2 and GMAT_12_z for U3. */
static void les_xterm(ora* o, double* S, const double* Q, const double* N, const double* GL, int js, int je, int zs,
                      int ze) {
  memset(S, 0, o->npts * sizeof(double));
  int NX2 = o->NX2;
  size_t dj = (size_t)NX2 * o->NZ2, dk = NX2;
#pragma omp parallel for schedule(static)
  for (int j = js; j <= je; ++j)
    for (int k = zs; k <= ze; ++k)
      for (int i = 0; i < NX2; ++i) {
        size_t q = IDX(o, i, k, j);
        S[q] = 0.25 * (GL[MF(o, k + 1, j, 2)] * 0.5 * (N[q + dk] + N[q]) *
                           (Q[q + dk + dj] + Q[q + dj] - Q[q - dj] - Q[q + dk - dj]) -
                       GL[MF(o, k, j, 2)] * 0.5 * (N[q] + N[q - dk]) *
                           (Q[q - dk + dj] + Q[q + dj] - Q[q - dj] - Q[q - dk - dj]) +
                       GL[MF(o, k, j + 1, 1)] * 0.5 * (N[q] + N[q + dj]) *
                           (Q[q + dk + dj] + Q[q + dk] - Q[q - dk + dj] - Q[q - dk]) -
                       GL[MF(o, k, j, 1)] * 0.5 * (N[q] + N[q - dj]) *
                           (Q[q + dk] + Q[q + dk - dj] - Q[q - dk] - Q[q - dk - dj]));
      }
}

/* CF += sgn * (i kx)^p * FFT(S) on the range (p = 1: x-derivative flux) */
static void accum(ora* o, double* CF, double* S, int ikx, int js, int je, int zs, int ze) {
  fwd(o, S);
#pragma omp parallel for schedule(static)
  for (int j = js; j <= je; ++j)
    for (int k = zs; k <= ze; ++k)
      for (int m = 0; m < o->NXC; ++m) {
        size_t q = CRE(o, m, k, j);
        if (ikx) { /* CF = CF - CIKXP * CS1 */
          double kx = o->kx[m];
          CF[q] = CF[q] - (0.0 * S[q] - kx * S[q + 1]);
          CF[q + 1] = CF[q + 1] - (0.0 * S[q + 1] + kx * S[q]);
        } else {
          CF[q] += S[q];
          CF[q + 1] += S[q + 1];
        }
      }
}

/* ------------------------------------------------------------------------ */
/* REM_DIV_CURV (src/duct.F90:2472-2818)                                     */
/* ------------------------------------------------------------------------ */
static int rem_div(ora* o) {
  int NXC = o->NXC, NKX = o->NKX, NY = o->NY, NZ = o->NZ;
  int js = o->jstart, je = o->jend, zs = o->zstart, ze = o->zend;
  double *CU1 = o->U[0], *CU2 = o->U[1], *CU3 = o->U[2];
  double *CR1 = o->PHI, *CS1 = o->CS1, *CU2b = o->U2b, *CU3b = o->U3b;
  size_t dj = (size_t)o->NX2 * o->NZ2, dk = o->NX2;
  memset(CR1, 0, o->npts * sizeof(double));
  memset(CS1, 0, o->npts * sizeof(double));
  memset(CU2b, 0, o->npts * sizeof(double));
  memset(CU3b, 0, o->npts * sizeof(double));
  /* :2511-2522 */
#pragma omp parallel for schedule(static)
  for (int j = js; j <= je + 1; ++j)
    for (int k = zs; k <= ze + 1; ++k)
      for (int m = 0; m < NXC; ++m)
        for (int c = 0; c < 2; ++c) {
          size_t q = CRE(o, m, k, j) + c;
          CU2b[q] = 0.5 * C22(o, k, j, 1) * (CU2[q] + CU2[q - dj]) + 0.5 * C21(o, k, j, 1) * (CU3[q] + CU3[q - dj]);
          CU3b[q] = 0.5 * C11(o, k, j, 2) * (CU3[q] + CU3[q - dk]) + 0.5 * C12(o, k, j, 2) * (CU2[q] + CU2[q - dk]);
        }
  /* :2527-2535 */
#pragma omp parallel for schedule(static)
  for (int j = js; j <= je; ++j)
    for (int k = zs - 1; k <= ze; ++k)
      for (int m = 0; m < NXC; ++m) {
        size_t q = CRE(o, m, k, j);
        double kj = o->kx[m] * JAC(o, k, j);
        /* CIKXP*J*CU1: (0 + i kx) J (a + i b) */
        double re = (0.0 * CU1[q] - kj * CU1[q + 1]) + (CU2b[q + dj] - CU2b[q]) + (CU3b[q + dk] - CU3b[q]);
        double im = (0.0 * CU1[q + 1] + kj * CU1[q]) + (CU2b[q + 1 + dj] - CU2b[q + 1]) + (CU3b[q + 1 + dk] - CU3b[q + 1]);
        CR1[q] = re; CR1[q + 1] = im;
        CS1[q] = re; CS1[q + 1] = im;
      }
  /* :2620-2668, modes above NKX have zero right-hand sides and solve to 0 */
  size_t pl = (size_t)o->NZ2 * o->NY2;
  int cyc_sum = 0, cyc_max = 0, nsolve = 0;
#pragma omp parallel reduction(+ : cyc_sum, nsolve) reduction(max : cyc_max)
  {
    double* rhs = dalloc(pl);
    double* pim = dalloc(pl);
    double* pre = dalloc(pl);
    mg_ctx c;
    tctx t;
    mg_setup(&c, o);
    if (o->p.proj_mode == 1) tight_setup(&t, &c);
#pragma omp for schedule(dynamic, 1)
    for (int m = 0; m <= NKX; ++m) {
      for (int part = 0; part < 2; ++part) { /* imaginary part first */
        memset(rhs, 0, pl * sizeof(double));
        for (int k = zs; k <= ze; ++k)
          for (int j = js; j <= je; ++j)
            rhs[PL(o, k, j)] = -CR1[CRE(o, m, k, j) + (part == 0 ? 1 : 0)];
        int cy = o->p.proj_mode == 1 ? tight_solve(&t, &c, o, o->kx2[m], rhs, part == 0 ? pim : pre)
                                     : mg_solve(&c, o->kx2[m], rhs, part == 0 ? pim : pre);
        cyc_sum += cy;
        nsolve += 1;
        if (cy > cyc_max) cyc_max = cy;
      }
      for (int j = 0; j <= NY + 1; ++j)
        for (int k = 0; k <= NZ + 1; ++k) {
          double a = pre[PL(o, k, j)], b = pim[PL(o, k, j)];
          if (o->p.cmplx_fp32) { a = (double)(float)a; b = (double)(float)b; }
          CR1[CRE(o, m, k, j)] = a;
          CR1[CRE(o, m, k, j) + 1] = b;
        }
    }
    if (o->p.proj_mode == 1) tight_free(&t);
    mg_free(&c);
    free(rhs); free(pim); free(pre);
  }
  o->diag.mg_solves = nsolve;
  o->diag.mg_cycles = cyc_sum;
  o->diag.mg_max_cycles = cyc_max;
  /* :2713-2724 face velocities */
#pragma omp parallel for schedule(static)
  for (int j = js; j <= je; ++j)
    for (int k = zs; k <= ze; ++k)
      for (int m = 0; m < NXC; ++m)
        for (int c = 0; c < 2; ++c) {
          size_t q = CRE(o, m, k, j) + c;
          CU2b[q] = CU2b[q] - G22(o, k, j, 1) * (CR1[q] - CR1[q - dj]) -
                    0.25 * G12(o, k, j, 1) * (CR1[q + dk] + CR1[q + dk - dj] - CR1[q - dk] - CR1[q - dk - dj]);
          CU3b[q] = CU3b[q] - G11(o, k, j, 2) * (CR1[q] - CR1[q - dk]) -
                    0.25 * G12(o, k, j, 2) * (CR1[q + dj] + CR1[q - dk + dj] - CR1[q - dj] - CR1[q - dk - dj]);
        }
  /* :2729-2750 cell-centre velocities */
#pragma omp parallel for schedule(static)
  for (int j = js; j <= je; ++j)
    for (int k = zs; k <= ze; ++k)
      for (int m = 0; m < NXC; ++m) {
        size_t q = CRE(o, m, k, j);
        double kx = o->kx[m];
        double pr = CR1[q], pi = CR1[q + 1];
        CU1[q] = CU1[q] - (0.0 * pr - kx * pi);
        CU1[q + 1] = CU1[q + 1] - (0.0 * pi + kx * pr);
        for (int c = 0; c < 2; ++c) {
          size_t r = q + c;
          CU2[r] = CU2[r] - 0.5 * (C22(o, k, j + 1, 1) * (CR1[r] + CR1[r + dj]) - C22(o, k, j, 1) * (CR1[r] + CR1[r - dj]) +
                                   C12(o, k + 1, j, 2) * (CR1[r + dk] + CR1[r]) - C12(o, k, j, 2) * (CR1[r] + CR1[r - dk])) /
                                JAC(o, k, j);
          CU3[r] = CU3[r] - 0.5 * (C11(o, k + 1, j, 2) * (CR1[r] + CR1[r + dk]) - C11(o, k, j, 2) * (CR1[r] + CR1[r - dk]) +
                                   C21(o, k, j + 1, 1) * (CR1[r + dj] + CR1[r]) - C21(o, k, j, 1) * (CR1[r] + CR1[r - dj])) /
                                JAC(o, k, j);
        }
      }
  /* :2754-2800 divergence after the projection */
#pragma omp parallel for schedule(static)
  for (int j = js; j <= je; ++j)
    for (int k = zs; k <= ze; ++k)
      for (int m = 0; m < NXC; ++m) {
        size_t q = CRE(o, m, k, j);
        double kj = o->kx[m] * JAC(o, k, j);
        CS1[q] = (0.0 * CU1[q] - kj * CU1[q + 1]) + (CU2b[q + dj] - CU2b[q]) + (CU3b[q + dk] - CU3b[q]);
        CS1[q + 1] = (0.0 * CU1[q + 1] + kj * CU1[q]) + (CU2b[q + 1 + dj] - CU2b[q + 1]) + (CU3b[q + 1 + dk] - CU3b[q + 1]);
      }
  inv(o, CS1);
  double div = 0.0, dmax = 0.0;
  for (int j = 2; j <= NY - 1; ++j)
    for (int k = 2; k <= NZ - 1; ++k)
      for (int i = 1; i <= o->NX + 1; ++i) {
        double v = fabs(CS1[IDX(o, i, k, j)]);
        div += v;
        if (v > dmax) dmax = v;
      }
  div = div / ((double)(o->NX + 1) * (double)NY * (double)NZ);
  o->diag.div_ref = div;
  o->diag.div_max = dmax;
  if (isnan(div)) return EKM_ERR_DIV_NAN;
  if (o->delta_t < 1.0e-4) return EKM_ERR_DT_SMALL;
  return EKM_OK;
}

int ora_rem_div_curv(ora* o) { return rem_div(o); }
int ora_get_phi(ora* o, double* phi) { memcpy(phi, o->PHI, o->npts * sizeof(double)); return EKM_OK; }

/* ------------------------------------------------------------------------ */
/* courant_curv_mpi (src/duct.F90:3605-3674)                                 */
/* ------------------------------------------------------------------------ */
static void courant(ora* o) {
  contravariant(o, o->U[1], o->U[2]);
  double dx = o->p.lx / (double)o->NX, cfl = o->p.cfl, dt = 999.0;
  for (int j = o->jstart; j <= o->jend; ++j)
    for (int k = o->zstart; k <= o->zend; ++k)
      for (int i = 0; i < o->NX2; ++i) {
        size_t q = IDX(o, i, k, j);
        double dtx = cfl * dx / fabs(o->U[0][q]);
        double dty = cfl * JAC(o, k, j) / fabs(o->U2b[q]);
        double dtz = cfl * JAC(o, k, j) / fabs(o->U3b[q]);
        dt = fmin(dt, fmin(dtx, fmin(dty, dtz)));
      }
  if (dt <= 0) return;
  o->delta_t = dt >= o->delta_t_set ? o->delta_t_set : dt;
  set_hbar(o);
}

/* ------------------------------------------------------------------------ */
/* RK_DUCT (src/duct.F90:237-2469)                                           */
/* ------------------------------------------------------------------------ */
/* ------------------------------------------------------------------------ */
/* LES, dynamic Smagorinsky model of the velocity (src/les_chan.F90 model 2, */
/* LES_IMPLICIT = .FALSE.)                                                   */
/* ------------------------------------------------------------------------ */
#define LES_ALPHA 2.44950 /* test / LES filter width ratio (les_chan.F90:37) */

/* x test filter [1/4 1/2 1/4], periodic over i = 0..NX-1, every (k, j)
 * (les_filter_chan filter_type 2, src/les_chan.F90:1719-1838, through
 * FILTER_VAR, src/mpi_duct.F90:834-852); the padding i = NX, NX+1 is kept */
static void les_filter(const ora* o, const double* a, double* b) {
  const int NX = o->NX;
#pragma omp parallel for schedule(static)
  for (int j = 0; j <= o->NY + 1; ++j)
    for (int k = 0; k <= o->NZ + 1; ++k) {
      const double* A = a + IDX(o, 0, k, j);
      double* B = b + IDX(o, 0, k, j);
      double tmp[4096 + 2];
      for (int i = 0; i < NX; ++i) {
        const int im = i == 0 ? NX - 1 : i - 1, ip = i == NX - 1 ? 0 : i + 1;
        tmp[i] = ((0.0 * A[im] + 0.25 * A[im]) + 0.5 * A[i] + 0.25 * A[ip]) + 0.0 * A[ip];
      }
      for (int i = 0; i < NX; ++i) B[i] = tmp[i];
      if (a != b) { B[NX] = A[NX]; B[NX + 1] = A[NX + 1]; }
    }
}

/* the two face-difference operators of the strain (compute_strain_chan,
 * src/les_chan.F90:1623-1717) at (k, j) on a field with strides dk, dj:
 * D2 uses CJOB_12 / CJOB_22, D3 uses CJOB_11 / CJOB_21 */
static inline double les_d2(const ora* o, const double* f, size_t q, int k, int j, size_t dk, size_t dj) {
  return 0.5 * C12(o, k + 1, j, 2) * (f[q] + f[q + dk]) - 0.5 * C12(o, k, j, 2) * (f[q] + f[q - dk]) +
         0.5 * C22(o, k, j + 1, 1) * (f[q] + f[q + dj]) - 0.5 * C22(o, k, j, 1) * (f[q] + f[q - dj]);
}
static inline double les_d3(const ora* o, const double* f, size_t q, int k, int j, size_t dk, size_t dj) {
  return 0.5 * C11(o, k + 1, j, 2) * (f[q] + f[q + dk]) - 0.5 * C11(o, k, j, 2) * (f[q] + f[q - dk]) +
         0.5 * C21(o, k, j + 1, 1) * (f[q] + f[q + dj]) - 0.5 * C21(o, k, j, 1) * (f[q] + f[q - dj]);
}

static void les_ghosts(const ora* o, double* a) {
  for (int k = 0; k <= o->NZ + 1; ++k)
    for (int i = 0; i < o->NX2; ++i) {
      a[IDX(o, i, k, o->NY + 1)] = a[IDX(o, i, k, o->NY)];
      a[IDX(o, i, k, 0)] = a[IDX(o, i, k, 1)];
    }
  for (int j = 0; j <= o->NY + 1; ++j)
    for (int i = 0; i < o->NX2; ++i) {
      a[IDX(o, i, o->NZ + 1, j)] = a[IDX(o, i, o->NZ, j)];
      a[IDX(o, i, 0, j)] = a[IDX(o, i, 1, j)];
    }
}

/* les_chan (src/les_chan.F90:1-1621, model 2): on entry CU1..3 in Fourier
 * space, CF1..3 holding the explicit terms; on exit U1..3 in physical
 * space, NU_T set and the divergence of the subgrid stress added to CF. */
static void les_chan(ora* o, int cur_step, int rk) {
  const int js = o->jstart, je = o->jend, zs = o->zstart, ze = o->zend, NXC = o->NXC;
  const size_t dj = (size_t)o->NX2 * o->NZ2, dk = o->NX2;
  /* strain rate in Fourier space (compute_strain_chan :1623-1717), interior */
  for (int q = 0; q < 6; ++q) memset(o->Sij[q], 0, o->npts * sizeof(double));
  const double *CU1 = o->U[0], *CU2 = o->U[1], *CU3 = o->U[2];
#pragma omp parallel for schedule(static)
  for (int j = js; j <= je; ++j)
    for (int k = zs; k <= ze; ++k) {
      const double Jc = JAC(o, k, j);
      for (int m = 0; m < NXC; ++m) {
        const double kx = o->kx[m];
        for (int c = 0; c < 2; ++c) {
          const size_t q = CRE(o, m, k, j) + c, qo = CRE(o, m, k, j) + (1 - c);
          const double sg = c == 0 ? -1.0 : 1.0; /* (i kx f)_re = -kx f_im, (i kx f)_im = kx f_re */
          o->Sij[0][q] = sg * kx * CU1[qo];
          o->Sij[1][q] = les_d2(o, CU2, q, k, j, dk, dj) / Jc;
          o->Sij[2][q] = les_d3(o, CU3, q, k, j, dk, dj) / Jc;
          o->Sij[3][q] = 0.5 * (les_d2(o, CU1, q, k, j, dk, dj) / Jc + sg * kx * CU2[qo]);
          o->Sij[4][q] = 0.5 * (les_d3(o, CU1, q, k, j, dk, dj) / Jc + sg * kx * CU3[qo]);
          o->Sij[5][q] = 0.5 * (les_d3(o, CU2, q, k, j, dk, dj) + les_d2(o, CU3, q, k, j, dk, dj)) / Jc;
        }
      }
    }
  for (int q = 0; q < 6; ++q) inv(o, o->Sij[q]);
  /* |S| at the interior points (:351-364), extended to the y ghosts (:366-371) */
  memset(o->SABS, 0, o->npts * sizeof(double));
#pragma omp parallel for schedule(static)
  for (int j = js; j <= je; ++j)
    for (int k = zs; k <= ze; ++k)
      for (int i = 0; i < o->NX2; ++i) {
        const size_t q = IDX(o, i, k, j);
        const double s1 = o->Sij[0][q], s2 = o->Sij[1][q], s3 = o->Sij[2][q];
        const double s4 = o->Sij[3][q], s5 = o->Sij[4][q], s6 = o->Sij[5][q];
        o->SABS[q] = sqrt(2.0 * s1 * s1 + 4.0 * s4 * s4 + 4.0 * s5 * s5 + 2.0 * s2 * s2 + 4.0 * s6 * s6 + 2.0 * s3 * s3);
      }
  for (int k = 0; k <= o->NZ + 1; ++k)
    for (int i = 0; i < o->NX2; ++i) {
      o->SABS[IDX(o, i, k, 0)] = o->SABS[IDX(o, i, k, 1)];
      o->SABS[IDX(o, i, k, o->NY + 1)] = o->SABS[IDX(o, i, k, o->NY)];
    }
  /* velocities to physical space and test-filtered (:412-464) */
  for (int f = 0; f < 3; ++f) inv(o, o->U[f]);
  for (int f = 0; f < 3; ++f) les_filter(o, o->U[f], o->UBT[f]);
  const double *U1 = o->U[0], *U2 = o->U[1], *U3 = o->U[2];
  /* dynamic coefficient every 5th step at the first substep, or on the
   * first step of a run (:475-814) */
  if (((cur_step % 5) == 0 && rk == 1) || o->first_time) {
    const size_t np = o->npts;
    double* S2B = o->S2B;
    double* T1 = dalloc(np);
    double* T2 = dalloc(np);
    double* M = dalloc(np);
    double* nsum = dalloc((size_t)o->NZ2 * o->NY2);
    double* dsum = dalloc((size_t)o->NZ2 * o->NY2);
    les_filter(o, o->SABS, S2B); /* kept for the scalar model */
    static const int ia[6] = {0, 1, 2, 0, 0, 1}, ib[6] = {0, 1, 2, 1, 2, 2};
    for (int ij = 0; ij < 6; ++ij) {
      const double cross = ij < 3 ? 1.0 : 2.0;
      les_filter(o, o->Sij[ij], T1);
      for (int j = 0; j <= o->NY + 1; ++j)
        for (int k = 0; k <= o->NZ + 1; ++k) {
          const double D = o->DELTA[PL(o, k, j)];
          for (int i = 0; i < o->NX2; ++i) {
            const size_t q = IDX(o, i, k, j);
            T1[q] = T1[q] * ((LES_ALPHA * D) * (LES_ALPHA * D)) * S2B[q];
            M[q] = (D * D) * o->SABS[q] * o->Sij[ij][q];
          }
        }
      les_filter(o, M, M);
      for (size_t q = 0; q < np; ++q) M[q] = T1[q] - M[q];
      const double *Ua = o->U[ia[ij]], *Ub = o->U[ib[ij]];
      for (size_t q = 0; q < np; ++q) T2[q] = Ua[q] * Ub[q];
      les_filter(o, T2, T2);
      const double *Ba = o->UBT[ia[ij]], *Bb = o->UBT[ib[ij]];
      for (int j = 0; j <= o->NY + 1; ++j)
        for (int k = 0; k <= o->NZ + 1; ++k) {
          double sn = 0.0, sd = 0.0;
          for (int i = 0; i < o->NX; ++i) {
            const size_t q = IDX(o, i, k, j);
            sn += M[q] * (T2[q] - Ba[q] * Bb[q]);
            sd += M[q] * M[q];
          }
          dsum[PL(o, k, j)] += cross * sd;
          nsum[PL(o, k, j)] += cross * sn;
        }
    }
    memset(o->CDYN, 0, (size_t)o->NZ2 * o->NY2 * sizeof(double));
    for (int j = js; j <= je; ++j)
      for (int k = zs; k <= ze; ++k) {
        const size_t t = PL(o, k, j);
        o->CDYN[t] = dsum[t] != 0.0 ? -0.5 * nsum[t] / dsum[t] : 0.0;
      }
    for (size_t t = 0; t < (size_t)o->NZ2 * o->NY2; ++t)
      if (o->CDYN[t] < 0.0) o->CDYN[t] = 0.0;
    free(T1); free(T2); free(M); free(nsum); free(dsum);
  }
  /* eddy viscosity (:825-831) */
  memset(o->NUT, 0, o->npts * sizeof(double));
  for (int j = js; j <= je; ++j)
    for (int k = zs; k <= ze; ++k) {
      const size_t t = PL(o, k, j);
      const double D = o->DELTA[t];
      for (int i = 0; i < o->NX2; ++i) o->NUT[IDX(o, i, k, j)] = o->CDYN[t] * (D * D) * o->SABS[IDX(o, i, k, j)];
    }
  /* subgrid stress, the implicit diagonal parts removed (:833-972) */
  for (int ij = 0; ij < 6; ++ij) {
    double* S = o->Sij[ij];
    if (ij == 0 || ij == 1 || ij == 2) {
      const int j0 = ij == 2 ? 0 : 1;
      for (int j = j0; j <= o->NY + 1; ++j)
        for (int k = 1; k <= o->NZ + 1; ++k) {
          const size_t t = PL(o, k, j);
          const double D = o->DELTA[t], C = o->CDYN[t];
          for (int i = 0; i < o->NX2; ++i) {
            const size_t q = IDX(o, i, k, j);
            S[q] = ij == 0 ? ((-2.0 * C) * (D * D)) * o->SABS[q] * S[q] : 0.0;
          }
        }
    } else {
      const double* V = ij == 5 ? U2 : U1;
      for (int j = js; j <= je; ++j)
        for (int k = zs; k <= ze; ++k) {
          const size_t t = PL(o, k, j);
          const double D = o->DELTA[t], C = o->CDYN[t], Jc = JAC(o, k, j);
          for (int i = 0; i < o->NX2; ++i) {
            const size_t q = IDX(o, i, k, j);
            const double g = ij == 3 ? les_d2(o, V, q, k, j, dk, dj) : les_d3(o, V, q, k, j, dk, dj);
            S[q] = ((-2.0 * C) * (D * D)) * o->SABS[q] * (S[q] - (0.5 * g) / Jc);
          }
        }
    }
    les_ghosts(o, S);
  }
  (void)U3;
  /* to Fourier space (the stress of ij = 2, 3 is zero) and the explicit
   * divergence into CF (:1169-1246, :1313-1356) */
  fwd(o, o->Sij[0]);
  fwd(o, o->Sij[3]);
  fwd(o, o->Sij[4]);
  fwd(o, o->Sij[5]);
  memset(o->Sij[1], 0, o->npts * sizeof(double));
  memset(o->Sij[2], 0, o->npts * sizeof(double));
#pragma omp parallel for schedule(static)
  for (int j = js; j <= je; ++j)
    for (int k = zs; k <= ze; ++k) {
      const double Jc = JAC(o, k, j);
      for (int m = 0; m < NXC; ++m) {
        const double kx = o->kx[m];
        for (int c = 0; c < 2; ++c) {
          const size_t q = CRE(o, m, k, j) + c, qo = CRE(o, m, k, j) + (1 - c);
          const double sg = c == 0 ? -1.0 : 1.0;
          const double *C1 = o->Sij[0], *C2 = o->Sij[1], *C3 = o->Sij[2], *C4 = o->Sij[3], *C5 = o->Sij[4],
                       *C6 = o->Sij[5];
          o->CF[0][q] = ((o->CF[0][q] - Jc * (sg * kx * C1[qo])) - les_d2(o, C4, q, k, j, dk, dj)) -
                        les_d3(o, C5, q, k, j, dk, dj);
          o->CF[1][q] = ((o->CF[1][q] - Jc * (sg * kx * C4[qo])) - les_d2(o, C2, q, k, j, dk, dj)) -
                        les_d3(o, C6, q, k, j, dk, dj);
          o->CF[2][q] = ((o->CF[2][q] - Jc * (sg * kx * C5[qo])) - les_d2(o, C6, q, k, j, dk, dj)) -
                        les_d3(o, C3, q, k, j, dk, dj);
        }
      }
    }
}

/* les_chan_th (src/les_chan_th.F90 model 2, scalar 1): on entry CTH in
 * Fourier space and U1..3, |S| and (on a refresh) the filtered velocities
 * and |S| from les_chan; on exit TH in physical space, KAPPA_T set and the x
 * part of the subgrid scalar flux added to CFTH (:411-419).  The y and z
 * parts act through KAPPA_T in the cross terms and the ADI. */
static void les_chan_th(ora* o, int cur_step, int rk) {
  const int js = o->jstart, je = o->jend, zs = o->zstart, ze = o->zend, NXC = o->NXC;
  const size_t dj = (size_t)o->NX2 * o->NZ2, dk = o->NX2, np = o->npts;
  double *G0 = o->Sij[0], *G1 = o->Sij[1], *G2 = o->Sij[2], *CTH = o->TH;
  /* scalar gradient (compute_scalar_grad :480-496) in Fourier space at the
   * velocity interior; the other rows carry nothing the model uses (their
   * C_DYN_TH and KAPPA_T are zero), so they are zero here */
  memset(G0, 0, np * sizeof(double));
  memset(G1, 0, np * sizeof(double));
  memset(G2, 0, np * sizeof(double));
#pragma omp parallel for schedule(static)
  for (int j = js; j <= je; ++j)
    for (int k = zs; k <= ze; ++k) {
      const double Jc = JAC(o, k, j);
      for (int m = 0; m < NXC; ++m) {
        const double kx = o->kx[m];
        for (int c = 0; c < 2; ++c) {
          const size_t q = CRE(o, m, k, j) + c, qo = CRE(o, m, k, j) + (1 - c);
          const double sg = c == 0 ? -1.0 : 1.0;
          G0[q] = sg * kx * CTH[qo];
          G1[q] = les_d2(o, CTH, q, k, j, dk, dj) / Jc;
          G2[q] = les_d3(o, CTH, q, k, j, dk, dj) / Jc;
        }
      }
    }
  inv(o, G0);
  inv(o, G1);
  inv(o, G2);
  /* background gradient added to d/dy (:521-531) */
  const double* TB = o->thbar;
  for (int j = js; j <= je; ++j)
    for (int k = zs; k <= ze; ++k) {
      const double b = les_d2(o, TB, PL(o, k, j), k, j, 1, (size_t)o->NZ2) / JAC(o, k, j);
      for (int i = 0; i < o->NX2; ++i) G1[IDX(o, i, k, j)] = G1[IDX(o, i, k, j)] + b;
    }
  inv(o, CTH); /* :144 */
  const double* TH = CTH;
  /* dynamic coefficient every 10th step at the first substep, or on the
   * first step of a run (:146-321); U_BAR_TIL and S_2BAR are the velocity
   * model's, refreshed in the same substep */
  if (((cur_step % 10) == 0 && rk == 1) || o->first_time) {
    double* THB = dalloc(np);
    double* T1 = dalloc(np);
    double* T2 = dalloc(np);
    double* M = dalloc(np);
    double* nsum = dalloc((size_t)o->NZ2 * o->NY2);
    double* dsum = dalloc((size_t)o->NZ2 * o->NY2);
    les_filter(o, TH, THB);
    const double* G[3] = {G0, G1, G2};
    for (int ij = 0; ij < 3; ++ij) {
      les_filter(o, G[ij], T1);
      for (int j = 0; j <= o->NY + 1; ++j)
        for (int k = 0; k <= o->NZ + 1; ++k) {
          const double D = o->DELTA[PL(o, k, j)];
          for (int i = 0; i < o->NX2; ++i) {
            const size_t q = IDX(o, i, k, j);
            T1[q] = T1[q] * ((LES_ALPHA * D) * (LES_ALPHA * D)) * o->S2B[q];
            M[q] = (D * D) * o->SABS[q] * G[ij][q];
          }
        }
      les_filter(o, M, M);
      for (size_t q = 0; q < np; ++q) M[q] = T1[q] - M[q];
      const double* Ui = o->U[ij];
      for (size_t q = 0; q < np; ++q) T2[q] = Ui[q] * TH[q];
      les_filter(o, T2, T2);
      const double* B = o->UBT[ij];
      for (int j = 0; j <= o->NY + 1; ++j)
        for (int k = 0; k <= o->NZ + 1; ++k) {
          double sn = 0.0, sd = 0.0;
          for (int i = 0; i < o->NX; ++i) {
            const size_t q = IDX(o, i, k, j);
            sn += M[q] * (T2[q] - THB[q] * B[q]);
            sd += M[q] * M[q];
          }
          dsum[PL(o, k, j)] += sd;
          nsum[PL(o, k, j)] += sn;
        }
    }
    memset(o->CDYNT, 0, (size_t)o->NZ2 * o->NY2 * sizeof(double));
    for (int j = js; j <= je; ++j)
      for (int k = zs; k <= ze; ++k) {
        const size_t t = PL(o, k, j);
        o->CDYNT[t] = dsum[t] != 0.0 ? -0.5 * nsum[t] / dsum[t] : 0.0;
      }
    for (size_t t = 0; t < (size_t)o->NZ2 * o->NY2; ++t)
      if (o->CDYNT[t] < 0.0) o->CDYNT[t] = 0.0;
    free(THB); free(T1); free(T2); free(M); free(nsum); free(dsum);
  }
  /* eddy diffusivity on the scalar's interior (:326-334) and the x flux
   * -C_DYN_TH Delta^2 |S| dTH/dx (:341-351; the y and z components, which
   * the reference also forms, are not used) */
  memset(o->KAPT, 0, np * sizeof(double));
  for (int j = o->jstart_th; j <= o->jend_th; ++j)
    for (int k = o->zstart_th; k <= o->zend_th; ++k) {
      const size_t t = PL(o, k, j);
      const double D = o->DELTA[t];
      for (int i = 0; i < o->NX2; ++i) o->KAPT[IDX(o, i, k, j)] = o->CDYNT[t] * (D * D) * o->SABS[IDX(o, i, k, j)];
    }
  for (int j = 0; j <= o->NY + 1; ++j)
    for (int k = 0; k <= o->NZ + 1; ++k) {
      const size_t t = PL(o, k, j);
      const double D = o->DELTA[t];
      for (int i = 0; i < o->NX2; ++i) {
        const size_t q = IDX(o, i, k, j);
        G0[q] = -(o->CDYNT[t] * (D * D) * o->SABS[q] * G0[q]);
      }
    }
  fwd(o, G0);
  /* CFTH -= J i kx (flux)^ (:411-419) */
  for (int j = o->jstart_th; j <= o->jend_th; ++j)
    for (int k = o->zstart_th; k <= o->zend_th; ++k) {
      const double Jc = JAC(o, k, j);
      for (int m = 0; m < NXC; ++m) {
        const size_t q = CRE(o, m, k, j);
        const double a = Jc * o->kx[m];
        o->CFTH[q] = o->CFTH[q] - (0.0 * G0[q] - a * G0[q + 1]);
        o->CFTH[q + 1] = o->CFTH[q + 1] - (0.0 * G0[q + 1] + a * G0[q]);
      }
    }
}

/* x statistics of SAVE_STATS_CURVI (src/flow_statistics.F90:163-215,
 * :420-495) by direct sums over the NX physical points of copies of the
 * fields transformed to physical space; same 16-plane layout as the library's
 * ekm_get_xstats (means of U1, U2, U3, P, TH; urms, vrms, wrms, thrms as sums
 * of squared fluctuations; uv, uw, wv, pu = w'p', pv, thv, thw as sums / NX). */
int ora_xstats(ora* o, double* out) {
  const size_t np = o->npts, pl = (size_t)o->NZ2 * o->NY2;
  double* F[5];
  const double* src[5] = {o->U[0], o->U[1], o->U[2], o->P, o->p.n_th ? o->TH : NULL};
  for (int q = 0; q < 5; ++q) {
    F[q] = dalloc(np);
    if (src[q]) memcpy(F[q], src[q], np * sizeof(double));
    else memset(F[q], 0, np * sizeof(double));
    inv(o, F[q]);
  }
  const int NX = o->NX;
  for (int j = 0; j < o->NY2; ++j)
    for (int k = 0; k < o->NZ2; ++k) {
      const size_t t = PL(o, k, j);
      double mean[5];
      for (int q = 0; q < 5; ++q) mean[q] = src[q] ? src[q][CRE(o, 0, k, j)] : 0.0; /* dble(CU(0,K,J)) */
      double s[11] = {0};
      for (int i = 0; i < NX; ++i) {
        const size_t q = IDX(o, i, k, j);
        const double u = F[0][q] - mean[0], v = F[1][q] - mean[1], w = F[2][q] - mean[2];
        const double p = F[3][q] - mean[3], th = F[4][q] - mean[4];
        s[0] += u * u; s[1] += v * v; s[2] += w * w; s[3] += th * th;
        s[4] += u * v; s[5] += u * w; s[6] += w * v; s[7] += w * p; s[8] += v * p;
        s[9] += th * v; s[10] += th * w;
      }
      for (int q = 0; q < 5; ++q) out[q * pl + t] = mean[q];
      for (int q = 0; q < 4; ++q) out[(5 + q) * pl + t] = s[q];
      for (int q = 4; q < 11; ++q) out[(5 + q) * pl + t] = s[q] / (double)NX;
    }
  for (int q = 0; q < 5; ++q) free(F[q]);
  return EKM_OK;
}

/* SAVE_STATS_CURVI's mean gradients, loop by loop (src/flow_statistics.F90:
 * 164-175 TKE, 228-249 dudy/dwdy, 250-272 dudz/dwdz, 464-491 dthdy, 495-513
 * dthdz, 517-553 THBULK and Rig), on the real parts of mode 0 */
int ora_mean_grads(ora* o, double* out) {
  const ekm_params* p = &o->p;
  const int NZ = o->NZ, NY = o->NY;
  const size_t pl = (size_t)o->NZ2 * o->NY2;
  memset(out, 0, (8 * pl + 1) * sizeof(double));
  double *dudy = out, *dwdy = out + pl, *dudz = out + 2 * pl, *dwdz = out + 3 * pl;
  double *dthdy = out + 4 * pl, *dthdz = out + 5 * pl, *rig = out + 6 * pl, *tke = out + 7 * pl;
#define M0(f, k, j) ((f)[CRE(o, 0, k, j)])
#define A4(c, k, j) (0.125 * (c(o, k, j, 1) + c(o, k, j + 1, 1) + c(o, k, j, 2) + c(o, k + 1, j, 2)))
  double *U1 = o->U[0], *U2 = o->U[1], *U3 = o->U[2], *TH = o->TH;
  for (int k = 0; k <= NZ + 1; ++k)
    for (int j = 0; j <= NY + 1; ++j) {
      tke[PL(o, k, j)] = (M0(U1, k, j) * M0(U1, k, j) + M0(U2, k, j) * M0(U2, k, j)) + M0(U3, k, j) * M0(U3, k, j);
      if (p->amp_omega0 != 0.0) tke[PL(o, k, j)] = tke[PL(o, k, j)] * pow(p->amp_omega0 / p->omega0, -2.0);
    }
  for (int k = o->zstart; k <= o->zend; ++k)
    for (int j = o->jstart; j <= o->jend; ++j) {
      dudy[PL(o, k, j)] = (A4(C12, k, j) * (M0(U1, k + 1, j) - M0(U1, k - 1, j)) +
                           A4(C22, k, j) * (M0(U1, k, j + 1) - M0(U1, k, j - 1))) / JAC(o, k, j);
      dwdy[PL(o, k, j)] = (A4(C12, k, j) * (M0(U3, k + 1, j) - M0(U3, k - 1, j)) +
                           A4(C22, k, j) * (M0(U3, k, j + 1) - M0(U3, k, j - 1))) / JAC(o, k, j);
    }
  for (int k = 1; k <= NZ; ++k) {
    dudy[PL(o, k, 0)] = dudy[PL(o, k, 1)];
    dwdy[PL(o, k, 0)] = dwdy[PL(o, k, 1)];
    dudy[PL(o, k, NY + 1)] = dudy[PL(o, k, NY)];
    dwdy[PL(o, k, NY + 1)] = dwdy[PL(o, k, NY)];
  }
  for (int k = 1; k <= NZ; ++k)
    for (int j = 1; j <= NY; ++j) {
      dudz[PL(o, k, j)] = (A4(C21, k, j) * (M0(U1, k, j + 1) - M0(U1, k, j - 1)) +
                           A4(C11, k, j) * (M0(U1, k + 1, j) - M0(U1, k - 1, j))) / JAC(o, k, j);
      dwdz[PL(o, k, j)] = (A4(C21, k, j) * (M0(U3, k, j + 1) - M0(U3, k, j - 1)) +
                           A4(C11, k, j) * (M0(U3, k + 1, j) - M0(U3, k - 1, j))) / JAC(o, k, j);
    }
  for (int k = 1; k <= NZ; ++k) {
    dudz[PL(o, k, 0)] = dudz[PL(o, k, 1)];
    dwdz[PL(o, k, 0)] = dwdz[PL(o, k, 1)];
    dudz[PL(o, k, NY + 1)] = dudz[PL(o, k, NY)];
    dwdz[PL(o, k, NY + 1)] = dwdz[PL(o, k, NY)];
  }
  for (int j = 1; j <= NY; ++j) {
    dwdz[PL(o, 1, j)] = dwdz[PL(o, 2, j)];
    dudz[PL(o, 1, j)] = dudz[PL(o, 2, j)];
  }
  if (p->n_th) {
    const double* TB = o->thbar;
    for (int k = 1; k <= NZ; ++k)
      for (int j = 1; j <= NY; ++j) {
        size_t q = PL(o, k, j);
        dthdy[q] = (A4(C12, k, j) * (M0(TH, k + 1, j) - M0(TH, k - 1, j)) +
                    A4(C22, k, j) * (M0(TH, k, j + 1) - M0(TH, k, j - 1))) / JAC(o, k, j);
        if (!p->cont_strat)
          dthdy[q] = dthdy[q] + (A4(C12, k, j) * (TB[PL(o, k + 1, j)] - TB[PL(o, k - 1, j)]) +
                                 A4(C22, k, j) * (TB[PL(o, k, j + 1)] - TB[PL(o, k, j - 1)])) / JAC(o, k, j);
      }
    for (int j = 1; j <= NY; ++j) {
      dthdy[PL(o, 0, j)] = dthdy[PL(o, 1, j)];
      dthdy[PL(o, NZ + 1, j)] = dthdy[PL(o, NZ, j)];
    }
    for (int k = 0; k <= NZ + 1; ++k) {
      dthdy[PL(o, k, 0)] = 2.0 * dthdy[PL(o, k, 1)] - dthdy[PL(o, k, 2)];
      dthdy[PL(o, k, NY + 1)] = 2.0 * dthdy[PL(o, k, NY)] - dthdy[PL(o, k, NY - 1)];
    }
    for (int k = 1; k <= NZ; ++k)
      for (int j = 1; j <= NY; ++j)
        dthdz[PL(o, k, j)] = (A4(C21, k, j) * (M0(TH, k, j + 1) - M0(TH, k, j - 1)) +
                              A4(C11, k, j) * (M0(TH, k + 1, j) - M0(TH, k - 1, j))) / JAC(o, k, j);
    for (int j = 1; j <= NY; ++j) {
      dthdz[PL(o, 0, j)] = dthdz[PL(o, 1, j)];
      dthdz[PL(o, NZ + 1, j)] = dthdz[PL(o, NZ, j)];
    }
    for (int k = 0; k <= NZ + 1; ++k) {
      dthdz[PL(o, k, 0)] = 2.0 * dthdz[PL(o, k, 1)] - dthdz[PL(o, k, 2)];
      dthdz[PL(o, k, NY + 1)] = 2.0 * dthdz[PL(o, k, NY)] - dthdz[PL(o, k, NY - 1)];
    }
    double ub = 0.0;
    for (int j = 1; j <= NY; ++j)
      for (int k = 1; k <= NZ; ++k) {
        size_t q = PL(o, k, j);
        ub = ub + 0.25 * (M0(TH, k, j) + M0(TH, k - 1, j) + M0(TH, k - 1, j - 1) + M0(TH, k, j - 1));
        double g = p->cont_strat ? dthdy[q] - 1.0 : dthdy[q];
        if (dwdy[q] == 0.0)
          rig[q] = -p->ri_tau * g * 1.0e18;
        else
          rig[q] = -p->ri_tau * g / (dwdy[q] * dwdy[q] + dudy[q] * dudy[q]);
      }
    out[8 * pl] = ub / (double)(float)(NY * NZ);
  }
#undef M0
#undef A4
  return EKM_OK;
}

/* omega_x of SAVE_STATS_CURVI (src/flow_statistics.F90:301-327): for
 * K = 1..NZ, J = 1..NY, (the four-face average of CJOB_12 times the K
 * difference of U3 plus that of CJOB_22 times the J difference) / INT_JACOB,
 * minus (CJOB_21 times the J difference of U2 plus CJOB_11 times the K
 * difference) / INT_JACOB, on the real parts of mode 0; then rows 0 and NY+1
 * copied from rows 1 and NY (:323-326) */
int ora_omega_x(ora* o, double* out) {
  const int NZ = o->NZ, NY = o->NY;
  memset(out, 0, (size_t)o->NZ2 * o->NY2 * sizeof(double));
#define M0(f, k, j) ((f)[CRE(o, 0, k, j)])
#define A4(c, k, j) (0.125 * (c(o, k, j, 1) + c(o, k, j + 1, 1) + c(o, k, j, 2) + c(o, k + 1, j, 2)))
  const double *U2 = o->U[1], *U3 = o->U[2];
  for (int k = 1; k <= NZ; ++k)
    for (int j = 1; j <= NY; ++j)
      out[PL(o, k, j)] = (A4(C12, k, j) * (M0(U3, k + 1, j) - M0(U3, k - 1, j)) +
                          A4(C22, k, j) * (M0(U3, k, j + 1) - M0(U3, k, j - 1))) / JAC(o, k, j) -
                         (A4(C21, k, j) * (M0(U2, k, j + 1) - M0(U2, k, j - 1)) +
                          A4(C11, k, j) * (M0(U2, k + 1, j) - M0(U2, k - 1, j))) / JAC(o, k, j);
  for (int k = 1; k <= NZ; ++k) {
    out[PL(o, k, 0)] = out[PL(o, k, 1)];
    out[PL(o, k, NY + 1)] = out[PL(o, k, NY)];
  }
#undef M0
#undef A4
  return EKM_OK;
}

/* field f (0..4: U1, U2, U3, P, TH) of the current state in physical space:
 * (NX+2)*(NZ+2)*(NY+2) doubles, the inverse transform of a copy */
int ora_get_physical(ora* o, int f, double* out) {
  if (f < 0 || f > 4) return EKM_ERR_ARG;
  const double* src = f < 3 ? o->U[f] : f == 3 ? o->P : o->p.n_th ? o->TH : NULL;
  if (!src) {
    memset(out, 0, o->npts * sizeof(double));
    return EKM_OK;
  }
  memcpy(out, src, o->npts * sizeof(double));
  inv(o, out);
  return EKM_OK;
}

/* high_oder_momentums_curvi (src/flow_statistics.F90:1026-1095): per (K,J)
 * the x averages of powers of the velocity fluctuations about the mode-0
 * means, in the order UU, VV, WW, UUU, VVV, WWW, UUUU, VVVV, WWWW, UUV, WWV,
 * UUUV, WWWV; direct sums over the NX physical points */
int ora_xmoments(ora* o, double* out) {
  const size_t np = o->npts, pl = (size_t)o->NZ2 * o->NY2;
  double* F[3];
  for (int q = 0; q < 3; ++q) {
    F[q] = dalloc(np);
    memcpy(F[q], o->U[q], np * sizeof(double));
    inv(o, F[q]);
  }
  const int NX = o->NX;
  for (int j = 0; j < o->NY2; ++j)
    for (int k = 0; k < o->NZ2; ++k) {
      const double mu = o->U[0][CRE(o, 0, k, j)], mv = o->U[1][CRE(o, 0, k, j)], mw = o->U[2][CRE(o, 0, k, j)];
      double s[13] = {0};
      for (int i = 0; i < NX; ++i) {
        const size_t q = IDX(o, i, k, j);
        const double u = F[0][q] - mu, v = F[1][q] - mv, w = F[2][q] - mw;
        s[0] += u * u; s[1] += v * v; s[2] += w * w;
        s[3] += u * u * u; s[4] += v * v * v; s[5] += w * w * w;
        s[6] += (u * u) * (u * u); s[7] += (v * v) * (v * v); s[8] += (w * w) * (w * w);
        s[9] += u * u * v; s[10] += w * w * v; s[11] += u * u * u * v; s[12] += w * w * w * v;
      }
      for (int q = 0; q < 13; ++q) out[q * pl + PL(o, k, j)] = s[q] / (double)NX;
    }
  for (int q = 0; q < 3; ++q) free(F[q]);
  return EKM_OK;
}

int ora_get_sponge(ora* o, double* sigma) {
  memcpy(sigma, o->sigma, ((size_t)o->NY + 2) * sizeof(double));
  return EKM_OK;
}

int ora_get_sponge_out(ora* o, double* sigma_out) {
  memcpy(sigma_out, o->sigma_out, ((size_t)o->NZ + 2) * sizeof(double));
  return EKM_OK;
}

int ora_get_forcing(ora* o, double* px0, double* qn) {
  if (px0) *px0 = o->p.f_type == 0 ? o->px0 : o->p.px0;
  if (qn) *qn = o->qn;
  return EKM_OK;
}

int ora_get_les_th(ora* o, double* kappa, double* cdyn_th) {
  if (!o->p.les_model || !o->p.n_th) return EKM_ERR_ARG;
  if (kappa) memcpy(kappa, o->KAPT, o->npts * sizeof(double));
  if (cdyn_th) memcpy(cdyn_th, o->CDYNT, (size_t)o->NZ2 * o->NY2 * sizeof(double));
  return EKM_OK;
}

int ora_get_les(ora* o, double* nut, double* cdyn) {
  if (!o->p.les_model) return EKM_ERR_ARG;
  if (nut) memcpy(nut, o->NUT, o->npts * sizeof(double));
  if (cdyn) memcpy(cdyn, o->CDYN, (size_t)o->NZ2 * o->NY2 * sizeof(double));
  return EKM_OK;
}

/* SPONGE_SIGMA(J), J = 0..NY+1 (src/diablo.F90:1533-1556): zero below NY_S_T;
 * type 1: ((ypoint(NZ+1,J) - ypoint(NZ+1,NY_S_T)) b)^2 with b = 1/REAL(span),
 * type 2: (exp(b (J - NY_S_T)) - 1)/(co_b - 1), co_b = 30,
 * b = log(co_b)/REAL(NY+1-NY_S_T).  REAL() is a default (single) real. */
static void sponge_sigma(ora* o) {
  const ekm_params* p = &o->p;
  const int NY = o->NY, nys = p->ny_s_t, kt = o->NZ + 1;
  for (int j = 0; j <= NY + 1; ++j) o->sigma[j] = 0.0;
  if (!p->sponge_type) return;
  if (nys < 0 || nys > NY + 1) {
    /* no vertical rows */
  } else if (p->sponge_type == 1) {
    const double y0 = o->yp[PL(o, kt, nys)];
    const double b = 1.0 / (double)(float)(o->yp[PL(o, kt, NY + 1)] - y0);
    for (int j = nys; j <= NY + 1; ++j) {
      const double xx = (o->yp[PL(o, kt, j)] - y0) * b;
      o->sigma[j] = xx * xx;
    }
  } else {
    const double co_b = 30.0, b = log(co_b) / (double)(float)(NY + 1 - nys);
    for (int j = nys; j <= NY + 1; ++j) o->sigma[j] = (exp(b * (double)(j - nys)) - 1.0) / (co_b - 1.0);
  }
  /* SPONGE_SIGMA_OUT(K) (src/diablo.F90:1489-1523), from xpoint(K, NY+1) */
  const int NZ = o->NZ, zr = p->nz_s_r, zl = p->nz_s_l;
  for (int k = 0; k <= NZ + 1; ++k) o->sigma_out[k] = 0.0;
  if (zr >= NZ + 1) return;
#define XT(k) (o->xp[PL(o, k, NY + 1)])
  if (p->sponge_type == 1) {
    const double b = 1.0 / (double)(float)(XT(NZ + 1) - XT(zr));
    for (int k = 0; k <= NZ + 1; ++k) {
      double xx;
      if (k >= zr) {
        xx = (XT(k) - XT(zr)) * b;
        o->sigma_out[k] = xx * xx;
      } else if (k <= zl) {
        xx = (XT(zl) - XT(k)) * b;
        o->sigma_out[k] = xx * xx;
      }
    }
  } else {
    const double co_b = 25.0, b = log(co_b) / (double)(float)(NZ + 1 - zr);
    for (int k = 0; k <= NZ + 1; ++k) {
      if (k >= zr)
        o->sigma_out[k] = (exp(b * (double)abs(k - zr)) - 1.0) / (co_b - 1.0);
      else if (k <= zl)
        o->sigma_out[k] = (exp(b * (double)abs(zl - k)) - 1.0) / (co_b - 1.0);
    }
  }
#undef XT
}

/* sponge_th (src/duct.F90:3196-3362): DYNAMIC_SPONE is .TRUE. and
 * Sponge_Amp = 0.1/DELTA_T; with the streamwise profile off SPONGE_SIGMA_OUT
 * is zero and its inflow/outflow updates are the identity.  (ART_VISCOSITY
 * enters RK_DUCT only under WAVE_ABS, which is .FALSE., src/diablo.F90:1486.) */
static void sponge_th(ora* o, int cur_step) {
  const ekm_params* p = &o->p;
  const int js = o->jstart, je = o->jend, zs = o->zstart, ze = o->zend;
  const double A = 0.1 / o->delta_t;
  double *CU1 = o->U[0], *CU2 = o->U[1], *CU3 = o->U[2], *CTH = o->TH;
  double *CF1 = o->CF[0], *CF2 = o->CF[1], *CF3 = o->CF[2], *CFT = o->CFTH;
  /* fluctuations, I = 1..NX2P (:3228-3236) */
  for (int j = js; j <= je; ++j)
    for (int k = zs; k <= ze; ++k)
      for (int m = 1; m < o->NXC; ++m)
        for (int c = 0; c < 2; ++c) {
          const size_t q = CRE(o, m, k, j) + c;
          for (int f = 0; f < 3; ++f)
            o->CF[f][q] = o->CF[f][q] - A * o->sigma[j] * (o->U[f][q] - 0.0) * JAC(o, k, j);
          CFT[q] = CFT[q] - A * o->sigma[j] * (CTH[q] - 0.0) * JAC(o, k, j);
        }
  /* inflow/outflow (:3240-3250), a separate loop */
  for (int j = js; j <= je; ++j)
    for (int k = zs; k <= ze; ++k)
      for (int m = 1; m < o->NXC; ++m)
        for (int c = 0; c < 2; ++c) {
          const size_t q = CRE(o, m, k, j) + c;
          for (int f = 0; f < 3; ++f)
            o->CF[f][q] = o->CF[f][q] - A * o->sigma_out[k] * (o->U[f][q] - 0.0) * JAC(o, k, j);
          CFT[q] = CFT[q] - A * o->sigma_out[k] * (CTH[q] - 0.0) * JAC(o, k, j);
        }
  /* mode 0 (rank 0): the mean of CU3 over the top rows (:3261-3276) */
  const int nysp = p->ny_s_t > o->NY ? 2 * o->NY / 3 : p->ny_s_t;
  double s3 = 0.0;
  for (int J = nysp; J <= je; ++J)
    for (int K = zs; K <= ze; ++K) s3 = s3 + CU3[CRE(o, 0, K, J)];
  s3 = s3 / (double)((ze - zs + 1) * (je - nysp + 1));
  const double t1 = p->u_bc_c1[1][EKM_YMAX], t2 = p->u_bc_c1[0][EKM_YMAX], t3 = p->wbulk0;
  const int dyn = fabs(s3 - t3) > t3 * 0.001;
  for (int j = js; j <= je; ++j)
    for (int k = zs; k <= ze; ++k) {
      const size_t q = CRE(o, 0, k, j);
      const double Jc = JAC(o, k, j);
      /* a complex minus a real target changes the real part only */
      if (dyn) { /* :3283-3292 */
        CF1[q] = CF1[q] - A * o->sigma[j] * (CU1[q] - t1) * Jc;
        CF1[q + 1] = CF1[q + 1] - A * o->sigma[j] * CU1[q + 1] * Jc;
        CF2[q] = CF2[q] - A * o->sigma[j] * (CU2[q] - t2) * Jc;
        CF2[q + 1] = CF2[q + 1] - A * o->sigma[j] * CU2[q + 1] * Jc;
        CF3[q] = CF3[q] - A * o->sigma[j] * (CU3[q] - t3) * Jc;
        CF3[q + 1] = CF3[q + 1] - A * o->sigma[j] * CU3[q + 1] * Jc;
        CFT[q] = CFT[q] - A * o->sigma[j] * (CTH[q] - 0.0) * Jc;
        CFT[q + 1] = CFT[q + 1] - A * o->sigma[j] * (CTH[q + 1] - 0.0) * Jc;
      } else { /* :3294-3301 */
        CF2[q] = CF2[q] - A * o->sigma[j] * (CU2[q] - t2) * Jc;
        CF2[q + 1] = CF2[q + 1] - A * o->sigma[j] * CU2[q + 1] * Jc;
        CF1[q] = CF1[q] - A * o->sigma[j] * (CU1[q] - t1) * Jc;
        CF1[q + 1] = CF1[q + 1] - A * o->sigma[j] * CU1[q + 1] * Jc;
      }
    }
  /* :3324-3332: the whole mode-0 plane of CF2, with INT_JACOB(K,J) read
   * after the loops above (K = ZEND+1, J = JEND+1); CU2X(0,:,:) zeroed on
   * the steps where MOD(2*TIME_STEP, SAVE_STATS_INT) = 0 */
  if (p->u_bc[0][EKM_ZMAX] == 6 && p->u_bc[1][EKM_ZMAX] == 6 && p->u_bc[2][EKM_ZMAX] == 6 && p->ri_tau != 0.0) {
    const double Jst = JAC(o, ze + 1, je + 1);
    for (int J = 0; J <= o->NY + 1; ++J)
      for (int K = 0; K <= o->NZ + 1; ++K) {
        const size_t q = CRE(o, 0, K, J);
        CF2[q] = CF2[q] - A * (CU2[q] - t2) * Jst;
        CF2[q + 1] = CF2[q + 1] - A * CU2[q + 1] * Jst;
      }
    if ((2 * cur_step) % p->save_stats_int == 0)
      for (int J = 0; J <= o->NY + 1; ++J)
        for (int K = 0; K <= o->NZ + 1; ++K) {
          const size_t q = CRE(o, 0, K, J);
          CU2[q] = 0.0;
          CU2[q + 1] = 0.0;
        }
  }
  /* inflow/outflow of mode 0 (:3350-3356) */
  for (int j = js; j <= je; ++j)
    for (int k = zs; k <= ze; ++k) {
      const size_t q = CRE(o, 0, k, j);
      const double Jc = JAC(o, k, j);
      CF1[q] = CF1[q] - 2 * A * o->sigma_out[k] * (CU1[q] - 0.0) * Jc;
      CF1[q + 1] = CF1[q + 1] - 2 * A * o->sigma_out[k] * CU1[q + 1] * Jc;
      CF2[q] = CF2[q] - 2 * A * o->sigma_out[k] * (CU2[q] - 0.0) * Jc;
      CF2[q + 1] = CF2[q + 1] - 2 * A * o->sigma_out[k] * CU2[q + 1] * Jc;
      CF3[q] = CF3[q] - 2 * A * o->sigma_out[k] * (CU3[q] - t3) * Jc;
      CF3[q + 1] = CF3[q + 1] - 2 * A * o->sigma_out[k] * CU3[q + 1] * Jc;
      CFT[q] = CFT[q] - 2 * A * o->sigma_out[k] * (CTH[q] - 0.0) * Jc;
      CFT[q + 1] = CFT[q + 1] - 2 * A * o->sigma_out[k] * (CTH[q + 1] - 0.0) * Jc;
    }
}

int ora_rk_duct(ora* o) {
  const ekm_params* p = &o->p;
  int rk = o->rk_step, NXC = o->NXC;
  int js = o->jstart, je = o->jend, zs = o->zstart, ze = o->zend;
  int cur_step = o->time_step + 1;
  double hb = o->h_bar[rk - 1];
  double TEMP1 = p->nu * hb / 2.0, TEMP3 = o->zeta_bar[rk - 1] * hb, TEMP4 = hb;
  double TEMP5 = o->beta_bar[rk - 1] * hb;
  double *CU1 = o->U[0], *CU2 = o->U[1], *CU3 = o->U[2], *CP = o->P, *CTH = o->TH;
  size_t dj = (size_t)o->NX2 * o->NZ2, dk = o->NX2;
  int th = p->n_th > 0;
  /* :298-455 explicit Fourier-space RHS */
  for (int c = 0; c < 3; ++c) memset(o->R[c], 0, o->npts * sizeof(double));
#pragma omp parallel for schedule(static)
  for (int j = js; j <= je; ++j)
    for (int k = zs; k <= ze; ++k)
      for (int m = 0; m < NXC; ++m)
        for (int c = 0; c < 2; ++c) {
          size_t q = CRE(o, m, k, j) + c;
          double Jc = JAC(o, k, j);
          for (int f = 0; f < 3; ++f) {
            o->R[f][q] = Jc * o->U[f][q];
            if (rk > 1) o->R[f][q] = o->R[f][q] + TEMP3 * o->CF[f][q];
          }
          double cs1 = 0.5 * C11(o, k + 1, j, 2) * (CP[q] + CP[q + dk]) - 0.5 * C11(o, k, j, 2) * (CP[q] + CP[q - dk]) +
                       0.5 * C21(o, k, j + 1, 1) * (CP[q] + CP[q + dj]) - 0.5 * C21(o, k, j, 1) * (CP[q] + CP[q - dj]);
          double cs2 = 0.5 * C12(o, k + 1, j, 2) * (CP[q] + CP[q + dk]) - 0.5 * C12(o, k, j, 2) * (CP[q] + CP[q - dk]) +
                       0.5 * C22(o, k, j + 1, 1) * (CP[q] + CP[q + dj]) - 0.5 * C22(o, k, j, 1) * (CP[q] + CP[q - dj]);
          o->PHI[q] = cs2; /* scratch: CS2X */
          o->CS1[q] = cs1;
        }
#pragma omp parallel for schedule(static)
  for (int j = js; j <= je; ++j)
    for (int k = zs; k <= ze; ++k)
      for (int m = 0; m < NXC; ++m) {
        size_t q = CRE(o, m, k, j);
        double b = TEMP4 * o->kx[m] * JAC(o, k, j);
        o->R[0][q] = o->R[0][q] - (0.0 * CP[q] - b * CP[q + 1]);
        o->R[0][q + 1] = o->R[0][q + 1] - (0.0 * CP[q + 1] + b * CP[q]);
        for (int c = 0; c < 2; ++c) {
          o->R[1][q + c] = o->R[1][q + c] - TEMP4 * o->PHI[q + c];
          o->R[2][q + c] = o->R[2][q + c] - TEMP4 * o->CS1[q + c];
        }
      }
  if (p->f_type == 0) { /* constant flow rate (:414-436) */
    double qn1 = 0.0;
    for (int j = js; j <= je - 1; ++j)
      for (int k = zs + 1; k <= ze - 1; ++k)
        qn1 = qn1 + (CU3[CRE(o, 0, k, j + 1)] + CU3[CRE(o, 0, k, j)]) * (o->yp[PL(o, k, j + 1)] - o->yp[PL(o, k, j)]) / 2;
    qn1 = qn1 / (o->yp[PL(o, 1, je)] - o->yp[PL(o, 1, js)]) / (double)(float)(ze - zs - 1);
    /* WEIGHT = 0.1 is assigned from a default-real literal (src/diablo.F90:361) */
    const double weight = (double)0.1f, q0 = p->wbulk0;
    o->px0 = o->px0 - (weight / TEMP4) * (o->qn + q0 - 2 * qn1);
    o->qn = qn1;
  }
  if (p->f_type == 0 || p->f_type == 1 || p->f_type == 2) { /* rank 0 owns mode 0 (:411-455) */
    double px = p->f_type == 0 ? o->px0
              : p->f_type == 1 ? p->px0 : (p->px0 + p->amp_omega0 * cos(p->omega0 * o->time));
    for (int j = js; j <= je; ++j)
      for (int k = zs; k <= ze; ++k) {
        size_t q = CRE(o, 0, k, j);
        o->R[2][q] = o->R[2][q] - TEMP4 * px * JAC(o, k, j);
      }
  }
  /* :461-475 horizontal viscous terms and Coriolis */
  double ug = p->u_bc_c1[0][EKM_YMAX], wg = p->wbulk0;
#pragma omp parallel for schedule(static)
  for (int j = js; j <= je; ++j)
    for (int k = zs; k <= ze; ++k)
      for (int m = 0; m < NXC; ++m) {
        size_t q = CRE(o, m, k, j);
        double Jc = JAC(o, k, j), f = p->nu * o->kx2[m] * Jc;
        for (int c = 0; c < 2; ++c)
          for (int fld = 0; fld < 3; ++fld) o->CF[fld][q + c] = -(f * o->U[fld][q + c]);
        double og = (p->coriolis_mean_only && m > 0) ? 0.0 : ug;
        double ow = (p->coriolis_mean_only && m > 0) ? 0.0 : wg;
        o->CF[2][q] = o->CF[2][q] + p->f0 * (CU1[q] - og) * Jc;
        o->CF[2][q + 1] = o->CF[2][q + 1] + p->f0 * CU1[q + 1] * Jc;
        o->CF[0][q] = o->CF[0][q] - p->f0 * (CU3[q] - ow) * Jc;
        o->CF[0][q + 1] = o->CF[0][q + 1] - p->f0 * CU3[q + 1] * Jc;
      }
  if (th) {
    double nupr = p->nu / p->pr;
    /* :517-523 buoyancy */
#pragma omp parallel for schedule(static)
    for (int j = js; j <= je; ++j)
      for (int k = zs; k <= ze; ++k)
        for (int m = 0; m < NXC; ++m)
          for (int c = 0; c < 2; ++c) {
            size_t q = CRE(o, m, k, j) + c;
            o->CF[1][q] = o->CF[1][q] - JAC(o, k, j) * p->ri_tau * CTH[q];
          }
    /* :532-632 scalar RHS */
    memset(o->RTH, 0, o->npts * sizeof(double));
    const double* TB = o->thbar;
    size_t tk = 1, tj = o->NZ2;
#pragma omp parallel for schedule(static)
    for (int j = o->jstart_th; j <= o->jend_th; ++j)
      for (int k = o->zstart_th; k <= o->zend_th; ++k)
        for (int m = 0; m < NXC; ++m) {
          size_t q = CRE(o, m, k, j);
          double Jc = JAC(o, k, j);
          for (int c = 0; c < 2; ++c) {
            o->RTH[q + c] = Jc * CTH[q + c];
            if (rk > 1) o->RTH[q + c] = o->RTH[q + c] + TEMP3 * o->CFTH[q + c];
            o->CFTH[q + c] = -nupr * Jc * o->kx2[m] * CTH[q + c];
          }
          if (p->cont_strat) {
            for (int c = 0; c < 2; ++c) o->CFTH[q + c] = o->CFTH[q + c] + Jc * CU2[q + c];
          } else {
            size_t t = PL(o, k, j);
            double gb = 0.5 * C12(o, k + 1, j, 2) * (TB[t] + TB[t + tk]) - 0.5 * C12(o, k, j, 2) * (TB[t] + TB[t - tk]) +
                        0.5 * C22(o, k, j + 1, 1) * (TB[t] + TB[t + tj]) - 0.5 * C22(o, k, j, 1) * (TB[t] + TB[t - tj]);
            for (int c = 0; c < 2; ++c) o->CFTH[q + c] = o->CFTH[q + c] - CU2[q + c] * gb;
            double dif = 0.25 * nupr *
                         (+G12(o, k + 1, j, 2) * (TB[t + tk + tj] + TB[t + tj] - TB[t - tj] - TB[t + tk - tj]) -
                          G12(o, k, j, 2) * (TB[t - tk + tj] + TB[t + tj] - TB[t - tj] - TB[t - tk - tj]) +
                          G12(o, k, j + 1, 1) * (TB[t + tk + tj] + TB[t + tk] - TB[t - tk + tj] - TB[t - tk]) -
                          G12(o, k, j, 1) * (TB[t + tk] + TB[t + tk - tj] - TB[t - tk] - TB[t - tk - tj]) +
                          4.0 * (G11(o, k + 1, j, 2) * (TB[t + tk] - TB[t]) - G11(o, k, j, 2) * (TB[t] - TB[t - tk])) +
                          4.0 * (G22(o, k, j + 1, 1) * (TB[t + tj] - TB[t]) - G22(o, k, j, 1) * (TB[t] - TB[t - tj])));
            o->CFTH[q] = o->CFTH[q] + dif;
          }
        }
  }
  if (th && p->sponge_type) sponge_th(o, cur_step); /* :635, after the scalar's explicit terms */
  /* :641-671 the subgrid model (which leaves the velocities in physical
   * space), or the velocities to physical space; the scalar to physical space */
  const int les = p->les_model && (!o->create_new_flow || cur_step > 1);
  if (les)
    les_chan(o, cur_step, rk);
  else
    for (int f = 0; f < 3; ++f) inv(o, o->U[f]);
  if (th && les)
    les_chan_th(o, cur_step, rk); /* :650-654, leaves TH physical */
  else if (th)
    inv(o, CTH);
  double *U1 = CU1, *U2 = CU2, *U3 = CU3;
  /* :683-1120 momentum fluxes */
  contravariant(o, U2, U3);
  for (int f = 0; f < 3; ++f) {
    xflux(o, o->S1, U1, o->U[f], js, je, zs, ze);
    accum(o, o->CF[f], o->S1, 1, js, je, zs, ze);
    yzflux(o, o->S1, o->U[f], p->nu, js, je, zs, ze);
    accum(o, o->CF[f], o->S1, 0, js, je, zs, ze);
  }
  if (les) /* :990-1120 */
    for (int f = 0; f < 3; ++f) {
      les_xterm(o, o->S1, o->U[f], o->NUT, o->met[f == 0 ? 1 : f == 1 ? 11 : 8], js, je, zs, ze);
      accum(o, o->CF[f], o->S1, 0, js, je, zs, ze);
    }
  /* :1127-1161 */
#pragma omp parallel for schedule(static)
  for (int j = js; j <= je; ++j)
    for (int k = zs; k <= ze; ++k)
      for (int m = 0; m < NXC; ++m)
        for (int c = 0; c < 2; ++c)
          for (int f = 0; f < 3; ++f) {
            size_t q = CRE(o, m, k, j) + c;
            o->R[f][q] = o->R[f][q] + TEMP5 * o->CF[f][q];
          }
  for (int f = 0; f < 3; ++f) inv(o, o->R[f]);
  if (th) { /* :1169-1625 */
    int tjs = o->jstart_th, tje = o->jend_th, tzs = o->zstart_th, tze = o->zend_th;
    xflux(o, o->S1, CTH, U1, tjs, tje, tzs, tze);
    accum(o, o->CFTH, o->S1, 1, tjs, tje, tzs, tze);
    yzflux(o, o->S1, CTH, p->nu / p->pr, tjs, tje, tzs, tze);
    accum(o, o->CFTH, o->S1, 0, tjs, tje, tzs, tze);
    if (les) { /* KAPPA_T cross terms (:1266-1311) */
      les_xterm(o, o->S1, CTH, o->KAPT, o->met[1], tjs, tje, tzs, tze);
      accum(o, o->CFTH, o->S1, 0, tjs, tje, tzs, tze);
    }
#pragma omp parallel for schedule(static)
    for (int j = tjs; j <= tje; ++j)
      for (int k = tzs; k <= tze; ++k)
        for (int m = 0; m < NXC; ++m)
          for (int c = 0; c < 2; ++c) {
            size_t q = CRE(o, m, k, j) + c;
            o->RTH[q] = o->RTH[q] + TEMP5 * o->CFTH[q];
          }
    inv(o, o->RTH);
    adi(o, 3, CTH, o->RTH, TEMP1 / p->pr, les ? o->KAPT : NULL, les_gz(o, 3), les_gy(o, 3), hb / 2.0);
  }
  /* :1687-1691 the wall model, before the velocities' implicit solves */
  if (p->u_bc[2][EKM_YMIN] == 3) wall_model(o, o->U[0], o->U[2], cur_step == 1);
  /* :1640-2391 velocity ADI */
  for (int f = 0; f < 3; ++f)
    adi(o, f, o->U[f], o->R[f], TEMP1, les ? o->NUT : NULL, les_gz(o, f), les_gy(o, f), hb / 2.0);
  /* :2417-2419 */
  if (p->variable_dt && rk == 3 && (cur_step % p->update_dt) == 0) courant(o);
  /* :2421-2451 */
  for (int f = 0; f < 3; ++f) fwd(o, o->U[f]);
  if (th) fwd(o, CTH);
  int st = rem_div(o);
  /* :2460-2466 pressure update */
  double *PHI = o->PHI;
  for (int j = js - 1; j <= je + 1; ++j)
    for (int k = zs - 1; k <= ze + 1; ++k)
      for (int m = 0; m < NXC; ++m)
        for (int c = 0; c < 2; ++c) {
          size_t q = CRE(o, m, k, j) + c;
          CP[q] = CP[q] + PHI[q] / TEMP4;
        }
  (void)U1;
  o->diag.status = st;
  if (rk == 3) {
    o->time = o->time + o->delta_t;
    o->time_step = cur_step;
    o->rk_step = 1;
    o->first_time = 0;
  } else {
    o->rk_step = rk + 1;
  }
  return st;
}

int ora_time_step(ora* o, int nsteps) {
  for (int n = 0; n < nsteps; ++n)
    for (int s = 0; s < 3; ++s) {
      int st = ora_rk_duct(o);
      if (st != EKM_OK) return st;
    }
  return EKM_OK;
}

/* ------------------------------------------------------------------------ */
/* state in/out                                                              */
/* ------------------------------------------------------------------------ */
int ora_set_state(ora* o, const double* cu1, const double* cu2, const double* cu3,
                  const double* cp, const double* cth, double time, int time_step) {
  size_t n = o->npts * sizeof(double);
  memcpy(o->U[0], cu1, n); memcpy(o->U[1], cu2, n); memcpy(o->U[2], cu3, n);
  memcpy(o->P, cp, n);
  if (o->p.n_th && cth) memcpy(o->TH, cth, n);
  o->time = time;
  o->time_step = time_step;
  o->rk_step = 1;
  o->create_new_flow = 0; /* a restart: the subgrid model runs from the first step */
  o->first_time = 1;
  return EKM_OK;
}

int ora_get_state(ora* o, double* cu1, double* cu2, double* cu3, double* cp, double* cth) {
  size_t n = o->npts * sizeof(double);
  if (cu1) memcpy(cu1, o->U[0], n);
  if (cu2) memcpy(cu2, o->U[1], n);
  if (cu3) memcpy(cu3, o->U[2], n);
  if (cp) memcpy(cp, o->P, n);
  if (cth && o->p.n_th) memcpy(cth, o->TH, n);
  return EKM_OK;
}

/* CREATE_FLOW_DUCT tail (src/duct.F90:3089-3189): forward transforms of the
 * physical initial fields, then the initial projection; CP starts at zero. */
int ora_set_initial_physical(ora* o, const double* u1, const double* u2, const double* u3,
                             const double* th) {
  size_t n = o->npts * sizeof(double);
  memcpy(o->U[0], u1, n); memcpy(o->U[1], u2, n); memcpy(o->U[2], u3, n);
  for (int f = 0; f < 3; ++f) fwd(o, o->U[f]);
  if (o->p.n_th && th) { memcpy(o->TH, th, n); fwd(o, o->TH); }
  memset(o->P, 0, n);
  for (int f = 0; f < 3; ++f) memset(o->CF[f], 0, n);
  memset(o->CFTH, 0, n);
  o->time = 0.0; o->time_step = 0; o->rk_step = 1;
  o->create_new_flow = 1;
  o->first_time = 1;
  return rem_div(o);
}

int ora_get_diag(ora* o, ekm_diag* d) {
  *d = o->diag;
  d->delta_t = o->delta_t;
  d->time = o->time;
  d->time_step = o->time_step;
  d->rk_step = o->rk_step;
  return EKM_OK;
}
