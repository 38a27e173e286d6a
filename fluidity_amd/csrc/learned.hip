// Learned preconditioner (placeholder until the network kernels land).
#include <hip/hip_runtime.h>

#include "flgpu_internal.h"
#include "learned.h"

namespace flgpu {

struct LearnedState {};

int learned_prepare(flgpu_ctx* c, bool mixed) {
    (void)c;
    (void)mixed;
    return set_error(FLGPU_EUNSUPPORTED, "learned preconditioner not built");
}

template <typename V>
int learned_precond_apply(flgpu_ctx* c, const V* r, V* z, V* y, const V* w, double* partials,
                          unsigned* counter, Scal* scal, bool set_norm0) {
    (void)c; (void)r; (void)z; (void)y; (void)w; (void)partials; (void)counter; (void)scal; (void)set_norm0;
    return set_error(FLGPU_EUNSUPPORTED, "learned preconditioner not built");
}
template int learned_precond_apply<double>(flgpu_ctx*, const double*, double*, double*, const double*,
                                           double*, unsigned*, Scal*, bool);
template int learned_precond_apply<float>(flgpu_ctx*, const float*, float*, float*, const float*, double*,
                                          unsigned*, Scal*, bool);

void learned_free(flgpu_ctx* c) {
    delete c->ls;
    c->ls = nullptr;
}

}  // namespace flgpu

using namespace flgpu;

extern "C" {
int flgpu_learned_param_count(void) { return 0; }
int flgpu_learned_set_params(flgpu_ctx*, const float*) { return set_error(FLGPU_EUNSUPPORTED, "not built"); }
int flgpu_learned_get_params(flgpu_ctx*, float*) { return set_error(FLGPU_EUNSUPPORTED, "not built"); }
int flgpu_learned_init_params(flgpu_ctx*, uint64_t) { return set_error(FLGPU_EUNSUPPORTED, "not built"); }
int flgpu_learned_train_step(flgpu_ctx*, int32_t, uint64_t, float, float*) {
    return set_error(FLGPU_EUNSUPPORTED, "not built");
}
int flgpu_learned_loss_grad(flgpu_ctx*, int32_t, uint64_t, float*, float*) {
    return set_error(FLGPU_EUNSUPPORTED, "not built");
}
int flgpu_learned_factor(flgpu_ctx*, float*, int64_t*) { return set_error(FLGPU_EUNSUPPORTED, "not built"); }
}
