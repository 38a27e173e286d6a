// Learned preconditioner interface used by the Krylov solver.
#pragma once

#include <cstdint>

#include "flgpu_internal.h"
#include "krylov_common.h"

namespace flgpu {

// Predict G for the device-resident reduced system (values in the solver's
// working precision).  Called once per solve.
int learned_prepare(flgpu_ctx* c, bool mixed, bool fsai = false);

// The diagonal scaling D = diag(A)^(-1/2) of the preconditioner's scaled
// space (valid after learned_prepare).
const float* learned_scale(flgpu_ctx* c);

// z = G^T (G r) = D Gs^T Gs D r with y as scratch; rt = D r is the scaled
// residual (formed by the solver's vector kernels).  The per-block sums
// r.z, z.z, z.w (w = A p of the iteration) are written to
// partials[k * kPartStride + b] for the solver to reduce; *nparts receives
// the number of blocks.  With done_flag set the kernels return at once.
template <typename V>
int learned_precond_apply(flgpu_ctx* c, const V* r, const V* rt, V* z, V* y, const V* w, double* partials,
                          int* nparts, const int* done_flag);

}  // namespace flgpu
