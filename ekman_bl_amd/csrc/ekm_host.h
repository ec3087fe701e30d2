// ekm_host.h -- host-side set-up tables (ekm_host.cpp).
#pragma once
#include <vector>

#include "ekm_internal.h"

struct HostMetrics {
  // GMAT_11,12,22, CJOB_11,12,21,22, then the subgrid-stress variants
  // GMAT_11_z,12_z,22_z,11_y,12_y,22_y: (0:NZ+2,0:NY+2,1:2)
  std::vector<double> m[13];
  std::vector<double> jac;   // INT_JACOB (0:NZ+2,0:NY+2)
};

struct HostMgLevel {
  int nx = 0, ny = 0;
  std::vector<double> c;      // c1..c8, [8][ny][nx]
  std::vector<double> s9;     // c1+c3+c5+c7 before boundary folding
  std::vector<double> corr9;  // corner correction of c9
  std::vector<double> jl;     // INT_JACOB at the level's points (0 at specified points)
};

int host_metrics(int NZ, int NY, const double* xp, const double* yp, HostMetrics& out);
int host_mg_level(const HostMetrics& hm, int NZ, int NY, const int bc[4], int nx, int ny, HostMgLevel& L);
int host_bc_table(const ekm_params& p, int comp, int dir, BcTable& t);
// SPONGE_SIGMA(0:NY+1) of the vertical sponge (src/diablo.F90:1533-1556)
void host_sponge_sigma(int type, int nz, int ny, int ny_s_t, const double* ypoint, double* sig);
// SPONGE_SIGMA_OUT(0:NZ+1) of the streamwise sponge (src/diablo.F90:1489-1523)
void host_sponge_sigma_out(int type, int nz, int ny, int nz_s_r, int nz_s_l, const double* xpoint, double* sig);
