// ekm_host.cpp -- host-side set-up of the duct step: curvilinear metrics,
// multigrid coefficient tables and ADI boundary rows.  No device code; the
// C entry points at the bottom let CPU tests compare these tables with the
// oracle without a GPU.
#include <cmath>
#include <cstring>
#include <vector>

#include "ekm_host.h"

// ---------------------------------------------------------------------------
// JACOBI_TRANS (src/diablo.F90:1088-1376): face metrics GMAT_*, CJOB_* and
// the cell Jacobian INT_JACOB from GRID_YZ.dat coordinates.
// ---------------------------------------------------------------------------
namespace {
struct Face {
  int nz, ny;  // interior sizes
  std::vector<double> v;
  Face(int nz_, int ny_) : nz(nz_), ny(ny_), v((size_t)(nz_ + 2) * (ny_ + 2) * 2, NAN) {}
  double& operator()(int k, int j, int f) { return v[(size_t)k + (size_t)(nz + 2) * ((size_t)j + (size_t)(ny + 2) * (f - 1))]; }
};
}  // namespace

int host_metrics(int NZ, int NY, const double* xp, const double* yp, HostMetrics& out) {
  const int NZ2 = NZ + 2, NZ3 = NZ + 3, NY3 = NY + 3;
  auto X = [&](int k, int j) { return xp[(size_t)k + (size_t)NZ2 * j]; };
  auto Y = [&](int k, int j) { return yp[(size_t)k + (size_t)NZ2 * j]; };
  Face xz(NZ, NY), yz(NZ, NY), xe(NZ, NY), ye(NZ, NY), jf(NZ, NY);
  // face derivatives (:1166-1190): face 1 differences K forward, face 2 J forward
  for (int j = 1; j <= NY; ++j)
    for (int k = 1; k <= NZ; ++k) {
      xz(k, j, 1) = X(k + 1, j) - X(k, j);
      yz(k, j, 1) = Y(k + 1, j) - Y(k, j);
      xz(k, j, 2) = 0.250 * (X(k + 1, j + 1) + X(k + 1, j) - X(k - 1, j + 1) - X(k - 1, j));
      yz(k, j, 2) = 0.250 * (Y(k + 1, j + 1) + Y(k + 1, j) - Y(k - 1, j + 1) - Y(k - 1, j));
      xe(k, j, 1) = 0.250 * (X(k, j + 1) + X(k + 1, j + 1) - X(k, j - 1) - X(k + 1, j - 1));
      ye(k, j, 1) = 0.250 * (Y(k, j + 1) + Y(k + 1, j + 1) - Y(k, j - 1) - Y(k + 1, j - 1));
      xe(k, j, 2) = X(k, j + 1) - X(k, j);
      ye(k, j, 2) = Y(k, j + 1) - Y(k, j);
    }
  auto ext = [](Face& a, int k, int j, int f, int k1, int j1, int k2, int j2) {
    a(k, j, f) = 2.0 * a(k1, j1, f) - a(k2, j2, f);
  };
  // linear extrapolation to the last face row and column (:1193-1217); the
  // statement order matters where a later line reads an earlier result
  for (int k = 1; k <= NZ + 1; ++k) {
    ext(xe, k, 1, 1, k, 2, k, 3);
    for (int f = 1; f <= 2; ++f) { ext(xe, k, NY + 1, f, k, NY, k, NY - 1); ext(xz, k, NY + 1, f, k, NY, k, NY - 1); }
    ext(ye, k, 1, 1, k, 2, k, 3);
    for (int f = 1; f <= 2; ++f) { ext(ye, k, NY + 1, f, k, NY, k, NY - 1); ext(yz, k, NY + 1, f, k, NY, k, NY - 1); }
  }
  for (int j = 1; j <= NY + 1; ++j) {
    ext(xz, 1, j, 2, 2, j, 3, j);
    for (int f = 1; f <= 2; ++f) { ext(xz, NZ + 1, j, f, NZ, j, NZ - 1, j); ext(xe, NZ + 1, j, f, NZ, j, NZ - 1, j); }
    ext(yz, 1, j, 2, 2, j, 3, j);
    for (int f = 1; f <= 2; ++f) { ext(yz, NZ + 1, j, f, NZ, j, NZ - 1, j); ext(ye, NZ + 1, j, f, NZ, j, NZ - 1, j); }
  }
  // 1e-8 truncation "smooth transformation" (:1220-1230)
  auto tr = [](double v) { return std::trunc(v * 100000000.0) * 1e-8; };
  for (int f = 1; f <= 2; ++f)
    for (int j = 1; j <= NY + 1; ++j)
      for (int k = 1; k <= NZ + 1; ++k) {
        xz(k, j, f) = tr(xz(k, j, f));
        xe(k, j, f) = tr(xe(k, j, f));
        yz(k, j, f) = tr(yz(k, j, f));
        ye(k, j, f) = tr(ye(k, j, f));
      }
  size_t mn = (size_t)NZ3 * NY3;
  for (int w = 0; w < 13; ++w) out.m[w].assign(mn * 2, NAN);
  out.jac.assign(mn, NAN);
  auto M = [&](int w, int k, int j, int f) -> double& { return out.m[w][(size_t)k + (size_t)NZ3 * ((size_t)j + (size_t)NY3 * (f - 1))]; };
  auto JC = [&](int k, int j) -> double& { return out.jac[(size_t)k + (size_t)NZ3 * j]; };
  // (:1233-1264)
  for (int f = 1; f <= 2; ++f)
    for (int j = 1; j <= NY + 1; ++j)
      for (int k = 1; k <= NZ + 1; ++k) {
        double J = xz(k, j, f) * ye(k, j, f) - xe(k, j, f) * yz(k, j, f);
        jf(k, j, f) = J;
        double zx = ye(k, j, f) / J, ex = -yz(k, j, f) / J;
        double zy = -xe(k, j, f) / J, ey = xz(k, j, f) / J;
        M(0, k, j, f) = (zx * zx + zy * zy) * J;
        M(1, k, j, f) = (zx * ex + zy * ey) * J;
        M(2, k, j, f) = (ex * ex + ey * ey) * J;
        M(3, k, j, f) = zx * J;
        M(4, k, j, f) = zy * J;
        M(5, k, j, f) = ex * J;
        M(6, k, j, f) = ey * J;
        // subgrid-stress variants (:1379-1403): U3's stress doubles the
        // zeta_x / eta_x products, U2's the zeta_y / eta_y products
        M(7, k, j, f) = (2.0 * zx * zx + zy * zy) * J;
        M(8, k, j, f) = (2.0 * zx * ex + zy * ey) * J;
        M(9, k, j, f) = (2.0 * ex * ex + ey * ey) * J;
        M(10, k, j, f) = (zx * zx + 2.0 * zy * zy) * J;
        M(11, k, j, f) = (zx * ex + 2.0 * zy * ey) * J;
        M(12, k, j, f) = (ex * ex + 2.0 * ey * ey) * J;
      }
  // cell Jacobian and its extrapolation (:1268-1292)
  for (int j = 1; j <= NY; ++j)
    for (int k = 1; k <= NZ; ++k)
      JC(k, j) = 0.250 * (jf(k, j, 1) + jf(k, j, 2) + jf(k, j + 1, 1) + jf(k + 1, j, 2));
  for (int j = 0; j <= NY + 1; ++j) {
    JC(NZ + 1, j) = 2.0 * JC(NZ, j) - JC(NZ - 1, j);
    JC(0, j) = 2.0 * JC(1, j) - JC(2, j);
  }
  for (int k = 0; k <= NZ + 1; ++k) {
    JC(k, NY + 1) = 2.0 * JC(k, NY) - JC(k, NY - 1);
    JC(k, 0) = 2.0 * JC(k, 1) - JC(k, 2);
  }
  JC(0, 0) = 0.5 * (2.0 * JC(1, 0) - JC(2, 0) + 2.0 * JC(0, 1) - JC(0, 2));
  JC(NZ + 1, NY + 1) = 0.5 * (2.0 * JC(NZ, NY + 1) - JC(NZ - 1, NY + 1) + 2.0 * JC(NZ + 1, NY) - JC(NZ + 1, NY - 1));
  // ghost faces (:1294-1340, variants :1405-1448); GMAT_12 and its variants
  // keep their extrapolated corner because the reference's corner formulas
  // assign GMAT_22 (GMAT_22_z, GMAT_22_y) twice
  for (int w = 0; w < 13; ++w)
    for (int f = 1; f <= 2; ++f) {
      for (int j = 0; j <= NY + 1; ++j) M(w, 0, j, f) = 2.0 * M(w, 1, j, f) - M(w, 2, j, f);
      for (int k = 0; k <= NZ + 1; ++k) M(w, k, 0, f) = 2.0 * M(w, k, 1, f) - M(w, k, 2, f);
    }
  for (int w : {0, 2, 3, 4, 5, 6, 7, 9, 10, 12})
    for (int f = 1; f <= 2; ++f)
      M(w, 0, 0, f) = 0.5 * (2.0 * M(w, 0, 1, f) - M(w, 0, 2, f) + 2.0 * M(w, 1, 0, f) - M(w, 2, 0, f));
  // copies for the coefficient callback (:1343-1376, variants :1450-1474)
  for (int w = 0; w < 13; ++w)
    for (int f = 1; f <= 2; ++f) {
      for (int k = 0; k <= NZ + 1; ++k) M(w, k, NY + 2, f) = M(w, k, NY + 1, f);
      for (int j = 0; j <= NY + 1; ++j) M(w, NZ + 2, j, f) = M(w, NZ + 1, j, f);
      M(w, NZ + 2, NY + 2, f) = M(w, NZ + 1, NY + 1, f);
    }
  return EKM_OK;
}

// ---------------------------------------------------------------------------
// Multigrid coefficients of one level: MUDPACK's dismd2cr (src/mud2cr.f:
// 501-1270) for the coefficient callback cofcr (src/duct.F90:3900-3934) and
// homogeneous derivative conditions (bndcr returns gbdy = 0; y sides
// alfa=0, beta=1; z sides alfa=1, beta=0).  With those values every
// mixed-boundary adjustment reduces to folding the virtual-point
// coefficients onto their mirror images (phi(0)=phi(2)); the terms that carry
// alfa/beta ratios or gama are exact zeros and are not formed.
// c9 is split as ((ce - s9) + corr9) with ce = -(kx2*jl) supplied per solve.
// ---------------------------------------------------------------------------
int host_mg_level(const HostMetrics& hm, int NZ, int NY, const int bc[4], int lx_n, int ly_n,
                  HostMgLevel& L) {
  const int NZ3 = NZ + 3, NY3 = NY + 3;
  const int nx = lx_n, ny = ly_n;
  L.nx = nx;
  L.ny = ny;
  size_t n = (size_t)nx * ny;
  L.c.assign(8 * n, 0.0);
  L.s9.assign(n, 0.0);
  L.corr9.assign(n, 0.0);
  L.jl.assign(n, 0.0);
  auto G = [&](int w, int k, int j, int f) { return hm.m[w][(size_t)k + (size_t)NZ3 * ((size_t)j + (size_t)NY3 * (f - 1))]; };
  auto C = [&](int l, int i, int j) -> double& { return L.c[(size_t)(l - 1) * n + (size_t)(i - 1) + (size_t)nx * (j - 1)]; };
  const double xa = 0.0, xb = (double)(NZ + 1), yc = 0.0, yd = (double)(NY + 1);
  const double dlx = (xb - xa) / (nx - 1), dly = (yd - yc) / (ny - 1);
  const double dlx2 = dlx + dlx, dlxx = dlx * dlx, dly2 = dly + dly, dlyy = dly * dly;
  const double dlxy = dlx * dly, dlxy2 = dlxy + dlxy, dlxy4 = dlxy2 + dlxy2;
  std::vector<double> cxyv(n);
  for (int j = 1; j <= ny; ++j) {
    double y = yc + (j - 1) * dly;
    for (int i = 1; i <= nx; ++i) {
      double x = xa + (i - 1) * dlx;
      int k = (int)x, jj = (int)y;
      double cxx = 0.25 * (G(0, k, jj, 2) + G(0, k + 1, jj, 2) + G(0, k, jj, 1) + G(0, k, jj + 1, 1));
      double cxy = 0.5 * (G(1, k, jj, 2) + G(1, k + 1, jj, 2) + G(1, k, jj, 1) + G(1, k, jj + 1, 1));
      double cyy = 0.25 * (G(2, k, jj, 2) + G(2, k + 1, jj, 2) + G(2, k, jj, 1) + G(2, k, jj + 1, 1));
      double cx = G(0, k + 1, jj, 2) - G(0, k, jj, 2) + G(1, k, jj + 1, 1) - G(1, k, jj, 1);
      double cy = G(2, k, jj + 1, 1) - G(2, k, jj, 1) + G(1, k + 1, jj, 2) - G(1, k, jj, 2);
      cxx = std::fmax(cxx, std::fabs(cx) * dlx * 0.5);  // upwind clamp (:570-571)
      cyy = std::fmax(cyy, std::fabs(cy) * dly * 0.5);
      double c1 = cxx / dlxx + cx / dlx2, c2 = cxy / dlxy4, c3 = cyy / dlyy + cy / dly2;
      double c5 = cxx / dlxx - cx / dlx2, c7 = cyy / dlyy - cy / dly2;
      C(1, i, j) = c1; C(2, i, j) = c2; C(3, i, j) = c3; C(4, i, j) = -c2;
      C(5, i, j) = c5; C(6, i, j) = c2; C(7, i, j) = c7; C(8, i, j) = -c2;
      size_t p = (size_t)(i - 1) + (size_t)nx * (j - 1);
      L.s9[p] = c1 + c3 + c5 + c7;
      L.jl[p] = hm.jac[(size_t)k + (size_t)NZ3 * jj];
      cxyv[p] = cxy;
    }
  }
  auto corr = [&](int i, int j) -> double& { return L.corr9[(size_t)(i - 1) + (size_t)nx * (j - 1)]; };
  auto cxyat = [&](int i, int j) { return cxyv[(size_t)(i - 1) + (size_t)nx * (j - 1)]; };
  // corners, virtual-virtual points not used (:594-637)
  {
    double v = -cxyat(1, 1) / dlxy2;
    C(1, 1, 1) -= v; C(2, 1, 1) = 0.0; C(3, 1, 1) -= v; C(4, 1, 1) = v;
    C(5, 1, 1) -= v; C(6, 1, 1) = 0.0; C(7, 1, 1) -= v; C(8, 1, 1) = v; corr(1, 1) = 2.0 * v;
    v = cxyat(nx, 1) / dlxy2;
    C(1, nx, 1) -= v; C(2, nx, 1) = v; C(3, nx, 1) -= v; C(4, nx, 1) = 0.0;
    C(5, nx, 1) -= v; C(6, nx, 1) = v; C(7, nx, 1) -= v; C(8, nx, 1) = 0.0; corr(nx, 1) = 2.0 * v;
    v = cxyat(1, ny) / dlxy2;
    C(1, 1, ny) -= v; C(2, 1, ny) = v; C(3, 1, ny) -= v; C(4, 1, ny) = 0.0;
    C(5, 1, ny) -= v; C(6, 1, ny) = v; C(7, 1, ny) -= v; C(8, 1, ny) = 0.0; corr(1, ny) = 2.0 * v;
    v = -cxyat(nx, ny) / dlxy2;
    C(1, nx, ny) -= v; C(2, nx, ny) = 0.0; C(3, nx, ny) -= v; C(4, nx, ny) = v;
    C(5, nx, ny) -= v; C(6, nx, ny) = 0.0; C(7, nx, ny) -= v; C(8, nx, ny) = v; corr(nx, ny) = 2.0 * v;
  }
  const int nxa = bc[0], nxb = bc[1], nyc = bc[2], nyd = bc[3];
  // derivative boundaries: fold the virtual column/row onto its mirror
  if (nxa == 2)
    for (int j = 2; j <= ny - 1; ++j) {
      double c4 = C(4, 1, j), c5 = C(5, 1, j), c6 = C(6, 1, j);
      C(1, 1, j) += c5; C(2, 1, j) += c4; C(8, 1, j) += c6;
      C(4, 1, j) = 0.0; C(5, 1, j) = 0.0; C(6, 1, j) = 0.0;
    }
  if (nxb == 2)
    for (int j = 2; j <= ny - 1; ++j) {
      double c1 = C(1, nx, j), c2 = C(2, nx, j), c8 = C(8, nx, j);
      C(4, nx, j) += c2; C(5, nx, j) += c1; C(6, nx, j) += c8;
      C(1, nx, j) = 0.0; C(2, nx, j) = 0.0; C(8, nx, j) = 0.0;
    }
  if (nyc == 2)
    for (int i = 2; i <= nx - 1; ++i) {
      double c6 = C(6, i, 1), c7 = C(7, i, 1), c8 = C(8, i, 1);
      C(2, i, 1) += c8; C(3, i, 1) += c7; C(4, i, 1) += c6;
      C(6, i, 1) = 0.0; C(7, i, 1) = 0.0; C(8, i, 1) = 0.0;
    }
  if (nyd == 2)
    for (int i = 2; i <= nx - 1; ++i) {
      double c2 = C(2, i, ny), c3 = C(3, i, ny), c4 = C(4, i, ny);
      C(6, i, ny) += c4; C(7, i, ny) += c3; C(8, i, ny) += c2;
      C(2, i, ny) = 0.0; C(3, i, ny) = 0.0; C(4, i, ny) = 0.0;
    }
  // corners of derivative sides (:817-1087)
  if (nyc == 2) {
    if (nxa == 0) {
      double c7 = C(7, 1, 1), c8 = C(8, 1, 1);
      C(2, 1, 1) += c8; C(3, 1, 1) += c7; C(6, 1, 1) = 0.0; C(7, 1, 1) = 0.0; C(8, 1, 1) = 0.0;
      double c6 = C(6, nx, 1);
      c7 = C(7, nx, 1);
      C(3, nx, 1) += c7; C(4, nx, 1) += c6; C(6, nx, 1) = 0.0; C(7, nx, 1) = 0.0; C(8, nx, 1) = 0.0;
    } else if (nxa == 2) {
      double c4 = C(4, 1, 1), c8 = C(8, 1, 1);
      C(2, 1, 1) = C(2, 1, 1) + c8 + c4; C(4, 1, 1) = 0.0; C(6, 1, 1) = 0.0; C(8, 1, 1) = 0.0;
      double c5 = C(5, 1, 1), c7 = C(7, 1, 1);
      C(1, 1, 1) += c5; C(3, 1, 1) += c7; C(5, 1, 1) = 0.0; C(7, 1, 1) = 0.0;
    }
    if (nxb == 2) {
      double c2 = C(2, nx, 1), c6 = C(6, nx, 1);
      C(2, nx, 1) = 0.0; C(4, nx, 1) = C(4, nx, 1) + c6 + c2; C(6, nx, 1) = 0.0; C(8, nx, 1) = 0.0;
      double c1 = C(1, nx, 1), c7 = C(7, nx, 1);
      C(1, nx, 1) = 0.0; C(3, nx, 1) += c7; C(5, nx, 1) += c1; C(7, nx, 1) = 0.0;
    }
  }
  if (nyd == 2) {
    if (nxa == 0) {
      double c2 = C(2, 1, ny), c3 = C(3, 1, ny);
      C(2, 1, ny) = 0.0; C(3, 1, ny) = 0.0; C(4, 1, ny) = 0.0; C(7, 1, ny) += c3; C(8, 1, ny) += c2;
      c3 = C(3, nx, ny);
      double c4 = C(4, nx, ny);
      C(2, nx, ny) = 0.0; C(3, nx, ny) = 0.0; C(4, nx, ny) = 0.0; C(6, nx, ny) += c4; C(7, nx, ny) += c3;
    } else if (nxa == 2) {
      double c2 = C(2, 1, ny), c6 = C(6, 1, ny);
      C(2, 1, ny) = 0.0; C(4, 1, ny) = 0.0; C(6, 1, ny) = 0.0; C(8, 1, ny) = C(8, 1, ny) + c6 + c2;
      double c3 = C(3, 1, ny), c5 = C(5, 1, ny);
      C(1, 1, ny) += c5; C(3, 1, ny) = 0.0; C(5, 1, ny) = 0.0; C(7, 1, ny) += c3;
    }
    if (nxb == 2) {
      double c4 = C(4, nx, ny), c8 = C(8, nx, ny);
      C(2, nx, ny) = 0.0; C(4, nx, ny) = 0.0; C(6, nx, ny) = C(6, nx, ny) + c4 + c8; C(8, nx, ny) = 0.0;
      double c1 = C(1, nx, ny), c3 = C(3, nx, ny);
      C(1, nx, ny) = 0.0; C(3, nx, ny) = 0.0; C(5, nx, ny) += c1; C(7, nx, ny) += c3;
    }
  }
  if (nxa == 2 && nyc == 0) {
    double c5 = C(5, 1, 1), c4 = C(4, 1, 1);
    C(2, 1, 1) += c4; C(1, 1, 1) += c5; C(6, 1, 1) = 0.0; C(5, 1, 1) = 0.0; C(4, 1, 1) = 0.0;
    double c6 = C(6, 1, ny);
    c5 = C(5, 1, ny);
    C(1, 1, ny) += c5; C(8, 1, ny) += c6; C(6, 1, ny) = 0.0; C(5, 1, ny) = 0.0; C(4, 1, ny) = 0.0;
  }
  if (nxb == 2 && nyc == 0) {
    double c2 = C(2, nx, 1), c1 = C(1, nx, 1);
    C(2, nx, 1) = 0.0; C(1, nx, 1) = 0.0; C(8, nx, 1) = 0.0; C(5, nx, 1) += c1; C(4, nx, 1) += c2;
    c1 = C(1, nx, ny);
    double c8 = C(8, nx, ny);
    C(2, nx, ny) = 0.0; C(1, nx, ny) = 0.0; C(8, nx, ny) = 0.0; C(6, nx, ny) += c8; C(5, nx, ny) += c1;
  }
  // specified boundaries: identity rows, c9 = 1 via jl = 0, s9 = -1, corr9 = 0
  auto spec = [&](int i, int j) {
    for (int l = 1; l <= 8; ++l) C(l, i, j) = 0.0;
    size_t p = (size_t)(i - 1) + (size_t)nx * (j - 1);
    L.jl[p] = 0.0;
    L.s9[p] = -1.0;
    L.corr9[p] = 0.0;
  };
  if (nxa == 1) for (int j = 1; j <= ny; ++j) spec(1, j);
  if (nxb == 1) for (int j = 1; j <= ny; ++j) spec(nx, j);
  if (nyc == 1) for (int i = 1; i <= nx; ++i) spec(i, 1);
  if (nyd == 1) for (int i = 1; i <= nx; ++i) spec(i, ny);
  return EKM_OK;
}

// ---------------------------------------------------------------------------
// ADI boundary rows: the net effect of APPLY_BC_{1,2,3,TH}_* (src/boundary.F90
// :124-1760) for IC_TYPE = 7 and BC codes 0 (Dirichlet), 1 (Neumann),
// 5 (scalar special), 6 (lagged "periodic").  Later assignments of a routine
// win, as in the reference (e.g. U1 right Neumann ends as [0 1 0 | 0], :775-797).
// comp: 0=U1, 1=U2, 2=U3, 3=theta.  dir 0: y (LOWER/UPPER), 1: z (LEFT/RIGHT).
// ---------------------------------------------------------------------------
static void put(BcTable& t, int row, double a, double b, double c, int kind, double val, int src) {
  for (auto& e : t.r)
    if (e.row == row) { e = BcRow{row, kind, src, 0, a, b, c, val}; return; }
  for (auto& e : t.r)
    if (e.row < 0) { e = BcRow{row, kind, src, 0, a, b, c, val}; return; }
}

int host_bc_table(const ekm_params& p, int comp, int dir, BcTable& t) {
  for (auto& e : t.r) e = BcRow{-1, 0, 0, 0, 0.0, 1.0, 0.0, 0.0};
  t.sp = 0;
  t.sp_metric = 0;
  const int NY = p.ny, NZ = p.nz;
  int lo = dir == 0 ? EKM_YMIN : EKM_ZMIN, hi = dir == 0 ? EKM_YMAX : EKM_ZMAX;
  int tl = comp < 3 ? p.u_bc[comp][lo] : p.th_bc[lo];
  int th_ = comp < 3 ? p.u_bc[comp][hi] : p.th_bc[hi];
  double cl = comp < 3 ? p.u_bc_c1[comp][lo] : p.th_bc_c1[lo];
  double ch = comp < 3 ? p.u_bc_c1[comp][hi] : p.th_bc_c1[hi];
  if (dir == 0) {
    const int top = NY + 1;
    switch (tl) {
      case 0: put(t, 0, 0., 1., 0., BC_CONST, cl, 0); break;
      case 1:
      case 5:
        if (comp == 3) put(t, 0, 0., -1., 1., BC_THFLUX_LO, cl, 0);      // :163-178
        else if (comp == 2) put(t, 0, 0., -1., 1., BC_CONST, cl, 0);     // :1338-1345
        else if (tl == 1) put(t, 1, 0., -1., 1., BC_CONST, cl, 0);       // :500-506, :868-874
        else return EKM_ERR_UNSUPPORTED;
        break;
      case 6: put(t, 0, 0., 1., 0., BC_LAGGED, 0.0, NY); break;
      case 3:
        // wall-model rows of IC_TYPE 7 with WALL_MODEL_TYPE 3 (:532-578,
        // :894-937, :1338-1380): U(1) - U(0) = the model's value, V's never
        // set by the model (zero)
        if (comp == 3 || p.wall_model_type != 3) return EKM_ERR_UNSUPPORTED;
        if (comp == 1) put(t, 0, 0., -1., 1., BC_CONST, 0.0, 0);
        else put(t, 0, 0., -1., 1., BC_WALL, 0.0, comp == 0 ? 0 : 1);
        break;
      default: return EKM_ERR_UNSUPPORTED;
    }
    switch (th_) {
      case 3:
        // the upper type-3 rows (:635-650, :1016-1030, :1480-1493): their
        // values come from U_BC_UPPER_NWM / W_BC_UPPER_NWM, which the
        // reference allocates and never assigns; taken as zero here
        if (comp == 3) return EKM_ERR_UNSUPPORTED;
        if (comp == 1) {
          put(t, top, 0., 1., 0., BC_CONST, 0.0, 0);
          put(t, NY, 0., 1., 0., BC_CONST, 0.0, 0);
        } else {
          put(t, NY, -1., 1., 0., BC_CONST, 0.0, 0);
          put(t, top, 0., 1., 0., BC_LAGGED, 0.0, NY);
        }
        break;
      case 0:
        put(t, top, 0., 1., 0., BC_CONST, ch, 0);
        if (comp == 1 || comp == 2) put(t, NY, 0., 1., 0., BC_CONST, ch, 0);  // :951-960, :1410-1419
        break;
      case 1:
        if (comp == 3) {  // :333-346
          put(t, NY, -1., 1., 0., BC_THFLUX_NY, ch, 0);
          put(t, top, -1., 1., 0., BC_THFLUX_NY1, ch, 0);
        } else if (comp == 1) {
          put(t, NY, -1., 1., 0., BC_CONST, ch, 0);  // :962-968
        } else {
          put(t, top, -1., 1., 0., BC_CONST, ch, 0);
        }
        break;
      case 6: put(t, top, 0., 1., 0., BC_LAGGED, 0.0, 1); break;
      default: return EKM_ERR_UNSUPPORTED;
    }
  } else {
    const int top = NZ + 1;
    switch (tl) {
      case 0: {
        double v0 = comp == 0 ? cl : 0.0, v1 = (comp == 0 || comp == 3) ? cl : 0.0;
        put(t, 0, 0., 1., 0., BC_CONST, v0, 0);
        put(t, 1, 0., 1., 0., BC_CONST, v1, 0);
        break;
      }
      case 1: put(t, 0, 0., -1., 1., BC_CONST, cl, 0); break;
      case 6: put(t, 0, 0., 1., 0., BC_LAGGED, 0.0, NZ); break;
      default: return EKM_ERR_UNSUPPORTED;
    }
    switch (th_) {
      case 0:
        put(t, top, 0., 1., 0., BC_CONST, comp == 2 ? ch : 0.0, 0);
        put(t, NZ, 0., 1., 0., BC_CONST, ch, 0);
        break;
      case 1:
        if (comp == 0) put(t, top, 0., 1., 0., BC_CONST, 0.0, 0);  // :775-797
        else put(t, top, -1., 1., 0., BC_CONST, ch, 0);
        break;
      case 5:
        if (comp != 3) return EKM_ERR_UNSUPPORTED;
        put(t, top, -2., 1., 0., BC_CONST, 0.0, 0);  // :443-449
        t.sp = 1;
        break;
      case 6: put(t, top, 0., 1., 0., BC_LAGGED, 0.0, 1); break;
      case 3:
        // V (:1249-1262): U2X(I,NZ,J) on row NZ+1; W (:1722-1734): zero rows,
        // W_BC_UPPER_NWM never assigned.  U1's rows (:802-815) read U1X at
        // the swapped index (I,J,NZ) and an unassigned DZ: not supported
        if (comp == 1) {
          put(t, NZ, -1., 1., 0., BC_CONST, 0.0, 0);
          put(t, top, 0., 1., 0., BC_LAGGED, 0.0, NZ);
        } else if (comp == 2) {
          put(t, top, -1., 1., 0., BC_CONST, 0.0, 0);
          put(t, NZ, 0., 1., 0., BC_CONST, 0.0, 0);
        } else {
          return EKM_ERR_UNSUPPORTED;
        }
        break;
      default: return EKM_ERR_UNSUPPORTED;
    }
  }
  return EKM_OK;
}

// ---------------------------------------------------------------------------
// C entry points for CPU-side checks of the tables (no device needed)
// ---------------------------------------------------------------------------
extern "C" int ekm_host_metrics(int nz, int ny, const double* xpoint, const double* ypoint, double* out) {
  HostMetrics hm;
  int st = host_metrics(nz, ny, xpoint, ypoint, hm);
  if (st) return st;
  size_t mn = (size_t)(nz + 3) * (ny + 3);
  for (int w = 0; w < 7; ++w) std::memcpy(out + (size_t)w * 2 * mn, hm.m[w].data(), 2 * mn * sizeof(double));
  std::memcpy(out + 14 * mn, hm.jac.data(), mn * sizeof(double));
  return EKM_OK;
}

// SPONGE_SIGMA(J), J = 0..NY+1 (src/diablo.F90:1533-1556): zero below
// NY_S_T; from NY_S_T up, parabolic (type 1) in ypoint(NZ+1, J) or
// exponential (type 2, co_b = 30) in the row index, both 0 at NY_S_T and 1 at
// NY+1.  The spans pass through REAL(), a default (single precision) real.
void host_sponge_sigma(int type, int nz, int ny, int ny_s_t, const double* ypoint, double* sig) {
  for (int j = 0; j <= ny + 1; ++j) sig[j] = 0.0;
  if (ny_s_t < 0 || ny_s_t > ny + 1) return;
  auto y = [&](int j) { return ypoint[(size_t)(nz + 1) + (size_t)(nz + 2) * j]; };
  if (type == 1) {
    const double b = 1.0 / (double)(float)(y(ny + 1) - y(ny_s_t));
    for (int j = ny_s_t; j <= ny + 1; ++j) {
      const double xx = (y(j) - y(ny_s_t)) * b;
      sig[j] = xx * xx;
    }
  } else if (type == 2) {
    const double co_b = 30.0;
    const double b = std::log(co_b) / (double)(float)(ny + 1 - ny_s_t);
    for (int j = ny_s_t; j <= ny + 1; ++j) sig[j] = (std::exp(b * (double)(j - ny_s_t)) - 1.0) / (co_b - 1.0);
  }
}

// SPONGE_SIGMA_OUT(K), K = 0..NZ+1 (src/diablo.F90:1489-1523): zero unless
// NZ_S_R < NZ+1; from NZ_S_R up and from NZ_S_L down, parabolic (type 1) in
// xpoint(K, NY+1) scaled by the right span, or exponential (type 2, co_b = 25)
// in the index distance; the spans pass through REAL() (single precision).
void host_sponge_sigma_out(int type, int nz, int ny, int nz_s_r, int nz_s_l, const double* xpoint, double* sig) {
  for (int k = 0; k <= nz + 1; ++k) sig[k] = 0.0;
  if (nz_s_r >= nz + 1 || nz_s_r < 0 || nz_s_l < 0 || nz_s_l > nz + 1) return;
  auto x = [&](int k) { return xpoint[(size_t)k + (size_t)(nz + 2) * (ny + 1)]; };
  if (type == 1) {
    const double b = 1.0 / (double)(float)(x(nz + 1) - x(nz_s_r));
    for (int k = 0; k <= nz + 1; ++k) {
      if (k >= nz_s_r) {
        const double xx = (x(k) - x(nz_s_r)) * b;
        sig[k] = xx * xx;
      } else if (k <= nz_s_l) {
        const double xx = (x(nz_s_l) - x(k)) * b;
        sig[k] = xx * xx;
      }
    }
  } else if (type == 2) {
    const double co_b = 25.0;
    const double b = std::log(co_b) / (double)(float)(nz + 1 - nz_s_r);
    for (int k = 0; k <= nz + 1; ++k) {
      if (k >= nz_s_r)
        sig[k] = (std::exp(b * (double)std::abs(k - nz_s_r)) - 1.0) / (co_b - 1.0);
      else if (k <= nz_s_l)
        sig[k] = (std::exp(b * (double)std::abs(nz_s_l - k)) - 1.0) / (co_b - 1.0);
    }
  }
}

extern "C" int ekm_host_sponge_sigma_out(int type, int nz, int ny, int nz_s_r, int nz_s_l, const double* xpoint,
                                         double* sig) {
  if (type < 0 || type > 2 || nz < 1 || ny < 1) return EKM_ERR_ARG;
  host_sponge_sigma_out(type, nz, ny, nz_s_r, nz_s_l, xpoint, sig);
  return EKM_OK;
}

extern "C" int ekm_host_sponge_sigma(int type, int nz, int ny, int ny_s_t, const double* ypoint, double* sig) {
  if (type < 0 || type > 2 || nz < 1 || ny < 1) return EKM_ERR_ARG;
  host_sponge_sigma(type, nz, ny, ny_s_t, ypoint, sig);
  return EKM_OK;
}

extern "C" int ekm_host_les_metrics(int nz, int ny, const double* xpoint, const double* ypoint, double* out) {
  HostMetrics hm;
  int st = host_metrics(nz, ny, xpoint, ypoint, hm);
  if (st) return st;
  size_t mn = (size_t)(nz + 3) * (ny + 3);
  for (int w = 0; w < 6; ++w) std::memcpy(out + (size_t)w * 2 * mn, hm.m[7 + w].data(), 2 * mn * sizeof(double));
  return EKM_OK;
}

// cf(i,j,1..9) of level `lev` (1 = coarsest) for wavenumber^2 kx2, in
// MUDPACK's cf(nx,ny,10) order (the tenth plane is left untouched).
extern "C" int ekm_host_mg_coefficients(const ekm_params* p, const double* xpoint, const double* ypoint,
                                        int lev, double kx2, double* cf, int* nx, int* ny) {
  HostMetrics hm;
  int st = host_metrics(p->nz, p->ny, xpoint, ypoint, hm);
  if (st) return st;
  int ngrid = p->mg_iex > p->mg_jey ? p->mg_iex : p->mg_jey;
  if (lev < 1 || lev > ngrid) return EKM_ERR_ARG;
  int ex = lev + p->mg_iex - ngrid, ey = lev + p->mg_jey - ngrid;
  int lnx = p->mg_ixp * (1 << ((ex > 1 ? ex : 1) - 1)) + 1;
  int lny = p->mg_jyq * (1 << ((ey > 1 ? ey : 1) - 1)) + 1;
  HostMgLevel L;
  st = host_mg_level(hm, p->nz, p->ny, p->mudpack_bc, lnx, lny, L);
  if (st) return st;
  size_t n = (size_t)lnx * lny;
  for (int l = 0; l < 8; ++l) std::memcpy(cf + (size_t)l * n, L.c.data() + (size_t)l * n, n * sizeof(double));
  for (size_t q = 0; q < n; ++q) cf[8 * n + q] = ((-(kx2 * L.jl[q])) - L.s9[q]) + L.corr9[q];
  *nx = lnx;
  *ny = lny;
  return EKM_OK;
}
