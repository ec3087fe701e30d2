// ekm_wsnet.h -- the learned initial guess of the per-mode pressure solves.
#pragma once
#include "ekm_internal.h"

struct ekm_wsnet;

// The network's guess for n real solves: from the right-hand sides rhs, the
// previous solutions prev (n planes of (NZ+2)*(NY+2) doubles, K fastest),
// the ratio that scales prev (one double: H_BAR of this substep over the
// previous one; null: 1) and each solve's kx^2, writes ratio prev plus the
// network's correction to out (may alias prev: each point reads its own
// previous value first).  Device pointers, enqueued on st.
int wsnet_guess(ekm_wsnet* net, const double* rhs, const double* prev, const double* ratio, const double* kx2, int n,
                double* out, hipStream_t st);
// plane extents and the largest batch the network was created for
void wsnet_dims(const ekm_wsnet* net, int* nz2, int* ny2, int* bmax);
