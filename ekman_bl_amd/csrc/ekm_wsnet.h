// ekm_wsnet.h -- the learned initial guess of the per-mode pressure solves.
#pragma once
#include "ekm_internal.h"

struct ekm_wsnet;

// The network's guess for n real solves: from the right-hand sides rhs and
// the previous solutions prev (n planes of (NZ+2)*(NY+2) doubles, K fastest)
// and each solve's kx^2, writes the guess planes to out (may alias prev: the
// inputs are read first).  Device pointers, enqueued on st.
int wsnet_guess(ekm_wsnet* net, const double* rhs, const double* prev, const double* kx2, int n, double* out,
                hipStream_t st);
// plane extents and the largest batch the network was created for
void wsnet_dims(const ekm_wsnet* net, int* nz2, int* ny2, int* bmax);
