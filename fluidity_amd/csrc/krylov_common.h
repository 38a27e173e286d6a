// Device helpers shared by the Krylov and learned-preconditioner kernels:
// deterministic block reductions and CSR products.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace flgpu {

enum {
    KSP_CONVERGED_RTOL = 2,
    KSP_CONVERGED_ATOL = 3,
    KSP_DIVERGED_ITS = -3,
    KSP_DIVERGED_DTOL = -4,
    KSP_DIVERGED_INDEFINITE_PC = -8,
    KSP_DIVERGED_NANORINF = -9,
    KSP_DIVERGED_INDEFINITE_MAT = -10
};

// Solver scalars, device resident.
struct Scal {
    double rz, rz_old, pAp, alpha, beta;
    double rr, zz, zw;
    double rnorm0, last_norm, true_rr, bb;
    double pivot, dmin, dmax;
    int32_t its, reason, done, max_its;
    double rtol, atol;
    int32_t norm_type, flexible, pad0, pad1;
};

constexpr int kVecBlock = 256;    // threads per block of the vector kernels
constexpr int kMaxVecGrid = 2048; // fixed grid: partial sums have a fixed order

// Sum over a 256-thread block in a fixed order; result valid in thread 0.
template <int NV>
__device__ __forceinline__ void block_sum(double (&v)[NV], double (*sh)[NV]) {
#pragma unroll
    for (int k = 0; k < NV; k++)
        for (int off = 32; off > 0; off >>= 1) v[k] += __shfl_xor(v[k], off, 64);
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    if (lane == 0)
#pragma unroll
        for (int k = 0; k < NV; k++) sh[wave][k] = v[k];
    __syncthreads();
    if (threadIdx.x == 0) {
        const int nw = blockDim.x >> 6;
#pragma unroll
        for (int k = 0; k < NV; k++) {
            double s = sh[0][k];
            for (int q = 1; q < nw; q++) s += sh[q][k];
            v[k] = s;
        }
    }
    __syncthreads();
}

__device__ __forceinline__ double load_coherent(const double* p) {
    unsigned long long u = __hip_atomic_load(reinterpret_cast<const unsigned long long*>(p),
                                             __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    return __longlong_as_double((long long)u);
}

// Every block writes its NV partial sums; the last block to finish adds
// all partials in block order (so the result does not depend on which
// block finished last) and returns true with the totals in out (thread 0).
template <int NV>
__device__ bool grid_sum(double (&v)[NV], double* __restrict__ partials, unsigned* counter,
                         double (&out)[NV]) {
    __shared__ double sh[kVecBlock / 64][NV];
    __shared__ unsigned ticket;
    block_sum<NV>(v, sh);
    if (threadIdx.x == 0) {
#pragma unroll
        for (int k = 0; k < NV; k++) partials[(size_t)blockIdx.x * NV + k] = v[k];
        __threadfence();
        ticket = atomicAdd(counter, 1u);
    }
    __syncthreads();
    if (ticket != gridDim.x - 1) return false;
    // acquire (agent-scope fence: stale L1 lines are dropped), then the
    // partials are read with ordinary, pipelined loads
    __threadfence();
    const double* vp = partials;
    double acc[NV];
#pragma unroll
    for (int k = 0; k < NV; k++) acc[k] = 0.0;
    for (int b = threadIdx.x; b < (int)gridDim.x; b += blockDim.x)
#pragma unroll
        for (int k = 0; k < NV; k++) acc[k] += vp[(size_t)b * NV + k];
    block_sum<NV>(acc, sh);
    if (threadIdx.x == 0) {
#pragma unroll
        for (int k = 0; k < NV; k++) out[k] = acc[k];
        *counter = 0u;
    }
    return true;
}

}  // namespace flgpu
