// Learned preconditioner interface used by the Krylov solver.
#pragma once

#include <cstdint>

#include "flgpu_internal.h"
#include "krylov_common.h"

namespace flgpu {

// Predict G for the device-resident reduced system (values in the solver's
// working precision).  Called once per solve.
int learned_prepare(flgpu_ctx* c, bool mixed);

// z = G^T (G r) with y as scratch; then the fixed-order sums r.z, r.r,
// z.z, z.w are finalised into the solver scalars (scal points at the
// solver's Scal: rz_old <- rz, rz, zz, zw; and the initial norm when
// set_norm0).
template <typename V>
int learned_precond_apply(flgpu_ctx* c, const V* r, V* z, V* y, const V* w, double* partials,
                          unsigned* counter, Scal* scal, bool set_norm0);

}  // namespace flgpu
