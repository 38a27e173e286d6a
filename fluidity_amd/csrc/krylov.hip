// Krylov solve of the assembled system on gfx950.
//
// The hand-off reproduces petsc_solve (femtools/Solvers.F90:111-187):
//   - inactive (strong Dirichlet) rows become "ghost" rows holding only the
//     diagonal ghost_pivot = sqrt((maxdiag+mindiag)*maxdiag/2) over all
//     |A_ii| (femtools/Petsc_Tools.F90:959-984, :1045-1057);
//   - inactive columns are dropped and lifted into the rhs,
//     b = rhs - A*ghost_rhs (femtools/Solvers.F90:1072-1099), and ghost rows
//     get b = 0;
//   - after the solve the ghost entries of x take the rhs values
//     (femtools/Petsc_Tools.F90:724-731).
// Methods: KSPCG-style preconditioned CG, flexible CG (Polak-Ribiere beta)
// for the learned preconditioner, restarted GMRES with left Jacobi.
// Precision: FP64, or FP32 inner iterations with FP64 dot-product
// accumulation and FP64 true-residual replacement ("reliable updates",
// iterative refinement without restarting CG); the FP32 iterations add
// their steps alpha p to an FP32 increment that is added to the FP64
// solution at every replacement (and at the end), so x itself is FP64.
// Every reduction has a fixed order: results repeat bit for bit.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "flgpu_internal.h"
#include "krylov_common.h"
#include "learned.h"

namespace flgpu {


struct KrylovWork {
    int n = 0;
    int64_t nnz = 0;
    DevBuf<double> val64, b, x, r, z, p, w, dinv, y, rt;
    DevBuf<float> val32, r32, z32, p32, w32, dinv32, y32, rt32;
    DevBuf<double> s;        // single-reduction CG: s = A p by recurrence
    DevBuf<float> s32;
    DevBuf<float> dx32;      // mixed CG: the FP32 increment of x since the last replacement
    bool dx_on = false;      // dx32 holds steps not yet added to x
    int npc = 0;             // partial sums of the last preconditioner apply
    DevBuf<double> partials;
    DevBuf<Scal> scal;       // two slots, see carry_if_done
    int cur = 0;             // slot holding the current state
    DevBuf<int32_t> rb, rbk; // row blocks of the streamed CSR product
    DevBuf<int16_t> col16;   // int16 column offsets (when use16)
    bool use16 = false;
    int nrb = 0;
    int64_t rb_key = -1;     // sparsity generation rb was built for
    // sliced ELL copy of the solver matrix, used by spmv when its padding is small
    bool use_sell = false, sell16 = false;
    int nsl = 0;
    int64_t nslot = 0;
    DevBuf<int32_t> sptr, ssrc, scol32;
    DevBuf<int16_t> scol16;
    DevBuf<double> sval64;
    DevBuf<float> sval32;
    Scal* h_scal = nullptr;  // host copy of the current slot (pinned when h_scal_owned)
    bool h_scal_owned = false;
    int nblk = 1;
    Scal* S() { return scal.p + cur; }
    double* part(int region) { return partials.p + (size_t)region * kPartStride; }
    // multicolour SSOR: rows grouped by colour (ascending inside a colour)
    DevBuf<int32_t> mc_rows;
    std::vector<int32_t> mc_ptr;  // colour c owns mc_rows[mc_ptr[c] .. mc_ptr[c+1])
    int64_t mc_key = -1;          // sparsity generation of the colouring
    // GMRES
    DevBuf<double> V, H, gv, cs, sn, hcol;
    DevBuf<float> V32;
    int gm = 0;
    bool have_x = false;
    bool prepared = false, prepared_mixed = false;
    ~KrylovWork() {
        if (h_scal && h_scal_owned) (void)hipHostFree(h_scal);
    }
};

// ---------------------------------------------------------------------
// Reduced system
// ---------------------------------------------------------------------

// Per-block min and max of |A_ii| (min/max do not depend on the order).
__global__ void __launch_bounds__(kVecBlock)
k_diag_minmax(int n, const int32_t* __restrict__ diag, const double* __restrict__ A,
              double* __restrict__ partials) {
    double mn = 1.79769313486231570e308, mx = -1.79769313486231570e308;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        const int d = diag[i];
        const double v = d >= 0 ? fabs(A[d]) : 0.0;
        mn = fmin(mn, v);
        mx = fmax(mx, v);
    }
    __shared__ double sh[kVecBlock / 64][2];
    for (int off = 32; off > 0; off >>= 1) {
        mn = fmin(mn, __shfl_xor(mn, off, 64));
        mx = fmax(mx, __shfl_xor(mx, off, 64));
    }
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    if (lane == 0) {
        sh[wave][0] = mn;
        sh[wave][1] = mx;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int q = 1; q < (int)(blockDim.x >> 6); q++) {
            sh[0][0] = fmin(sh[0][0], sh[q][0]);
            sh[0][1] = fmax(sh[0][1], sh[q][1]);
        }
        partials[blockIdx.x] = sh[0][0];
        partials[kPartStride + blockIdx.x] = sh[0][1];
    }
}

// One wave: ghost pivot sqrt((max+min)*max/2) from the per-block extrema.
__global__ void __launch_bounds__(64)
k_pivot(int nparts, const double* __restrict__ partials, Scal* sc) {
    double a = 1.79769313486231570e308, b = -1.79769313486231570e308;
    for (int q = threadIdx.x; q < nparts; q += 64) {
        a = fmin(a, partials[q]);
        b = fmax(b, partials[kPartStride + q]);
    }
    for (int off = 32; off > 0; off >>= 1) {
        a = fmin(a, __shfl_xor(a, off, 64));
        b = fmax(b, __shfl_xor(b, off, 64));
    }
    if (threadIdx.x == 0) {
        sc->dmin = a;
        sc->dmax = b;
        sc->pivot = sqrt((b + a) * b / 2.0);
    }
}

// The reduced (PETSc-side) system, per row block: the block's entries are
// read once, coalesced, and one thread per row then finishes the row from
// LDS.  Ghost rows keep only the pivot on the diagonal, inactive columns
// are dropped and lifted into the rhs -- b_i = rhs_i - sum_k A_ik rhs_k over
// the inactive columns k, in entry order with the product rounded first,
// as the reference's csr_mult of the ghost values does -- and the Jacobi
// inverse diagonal and the initial x are set.
__global__ void __launch_bounds__(kRowBlock)
k_build(int nrb, const int32_t* __restrict__ rb, const int32_t* __restrict__ rowptr,
        const int32_t* __restrict__ col, const double* __restrict__ A, const double* __restrict__ rhs,
        const int32_t* __restrict__ inactive, int have_inactive, const Scal* __restrict__ sc,
        double* __restrict__ v64, float* __restrict__ v32, double* __restrict__ b, double* __restrict__ dinv,
        float* __restrict__ dinv32, double* __restrict__ x, int zero_x) {
    __shared__ double slift[kStreamNnz];
    __shared__ double sdiag[kRowBlock];
    __shared__ uint8_t serow[kStreamNnz];
    const double pivot = sc->pivot;
    const RowBlocks it(nrb);
    for (int blk = it.b; blk < it.hi; blk += it.step) {
        const int r0 = rb[blk], r1 = rb[blk + 1];
        const int k0 = rowptr[r0], k1 = rowptr[r1];
        const int row = r0 + threadIdx.x;
        const bool staged = k1 - k0 <= kStreamNnz;
        if (staged) {
            if (row < r1) {
                sdiag[threadIdx.x] = 0.0;
                for (int k = rowptr[row]; k < rowptr[row + 1]; k++) serow[k - k0] = (uint8_t)threadIdx.x;
            }
            __syncthreads();
#pragma unroll
            for (int j = 0; j < kStreamPerThread; j++) {
                const int k = k0 + threadIdx.x + j * kRowBlock;
                if (k < k1) {
                    const int l = serow[k - k0], i = r0 + l, cj = col[k];
                    const double a = A[k];
                    const bool ghost = have_inactive && inactive[i];
                    const bool off = have_inactive && inactive[cj];
                    double v = a;
                    if (ghost) v = (cj == i) ? pivot : 0.0;
                    else if (off) v = 0.0;
                    v64[k] = v;
                    if (v32) v32[k] = (float)v;
                    if (cj == i) sdiag[l] = v;
                    slift[k - k0] = (!ghost && off) ? a * rhs[cj] : 0.0;
                }
            }
            __syncthreads();
        }
        if (row < r1) {
            const bool ghost = have_inactive && inactive[row];
            double lift = 0.0, dval = 0.0;
            if (staged) {
                for (int k = rowptr[row]; k < rowptr[row + 1]; k++) lift = lift + slift[k - k0];
                dval = sdiag[threadIdx.x];
            } else {  // a row longer than the staging buffer (alone in its block)
                for (int k = rowptr[row]; k < rowptr[row + 1]; k++) {
                    const int cj = col[k];
                    const double a = A[k];
                    const bool off = have_inactive && inactive[cj];
                    double v = a;
                    if (ghost) v = (cj == row) ? pivot : 0.0;
                    else if (off) v = 0.0;
                    v64[k] = v;
                    if (v32) v32[k] = (float)v;
                    if (cj == row) dval = v;
                    if (!ghost && off) lift = lift + a * rhs[cj];
                }
            }
            b[row] = ghost ? 0.0 : rhs[row] - lift;
            const double di = dval != 0.0 ? 1.0 / dval : 1.0;
            dinv[row] = di;
            if (dinv32) dinv32[row] = (float)di;
            if (zero_x || ghost) x[row] = 0.0;
        }
        __syncthreads();
    }
}

// ---------------------------------------------------------------------
// PCG / FCG vector kernels
// ---------------------------------------------------------------------

// Convergence test of KSPConvergedDefault on the norm chosen by norm_type,
// then beta (Fletcher-Reeves, or Polak-Ribiere when flexible).
__device__ void cg_test(Scal* sc) {
    sc->its += 1;
    const double nrm = sc->norm_type == FLGPU_NORM_UNPRECONDITIONED ? sqrt(sc->rr) : sqrt(sc->zz);
    sc->last_norm = nrm;
    if (nrm != nrm || isinf(nrm)) {
        sc->reason = KSP_DIVERGED_NANORINF;
        sc->done = 1;
    } else if (nrm <= sc->atol) {
        sc->reason = KSP_CONVERGED_ATOL;
        sc->done = 1;
    } else if (nrm <= sc->rtol * sc->rnorm0) {
        sc->reason = KSP_CONVERGED_RTOL;
        sc->done = 1;
    } else if (sc->rnorm0 > 0 && nrm >= 1e5 * sc->rnorm0) {
        sc->reason = KSP_DIVERGED_DTOL;
        sc->done = 1;
    } else if (sc->its >= sc->max_its) {
        sc->reason = KSP_DIVERGED_ITS;
        sc->done = 1;
    }
}

__device__ void cg_test_and_beta(Scal* sc) {
    cg_test(sc);
    if (sc->flexible)
        sc->beta = -sc->alpha * sc->zw / sc->rz_old;  // z.(r_new - r_old) / (z_old.r_old)
    else
        sc->beta = sc->rz / sc->rz_old;
    if (!sc->done && sc->rz < 0.0) {
        sc->reason = KSP_DIVERGED_INDEFINITE_PC;
        sc->done = 1;
    }
}

// The solver scalars advance through two slots: a kernel that changes
// them reads slot "in" in every block and block 0 writes the next state
// to slot "out", so no block can see a value written during its own
// kernel.  A finished solve carries the state over unchanged.
__device__ __forceinline__ bool carry_if_done(const Scal* __restrict__ in, Scal* __restrict__ out) {
    if (!in->done) return false;
    if (blockIdx.x == 0 && threadIdx.x == 0 && out != in) *out = *in;
    return true;
}

// R consecutive entries as one vector load / store (R * sizeof(T) bytes;
// the caller checks the alignment with rows_aligned)
template <typename T, int R>
using Rows = T __attribute__((ext_vector_type(R)));

template <int R, typename T>
__device__ __forceinline__ Rows<T, R> ld_rows(const T* __restrict__ a, int i) {
    return *reinterpret_cast<const Rows<T, R>*>(a + i);
}
template <int R, typename T>
__device__ __forceinline__ void st_rows(T* __restrict__ a, int i, const Rows<T, R>& v) {
    *reinterpret_cast<Rows<T, R>*>(a + i) = v;
}

// pAp from the SpMV partials, alpha = rz / pAp; x += alpha p;
// r -= alpha w; [Jacobi: z = D^-1 r]; [learned: rt = D_s r, the residual in
// the preconditioner's scaled space]; partials r.z, r.r, z.z, z.w.  Each
// thread takes R consecutive rows per step with vector loads (the rows past
// the last whole group go one at a time), so several loads of every vector
// are in flight per thread.
template <typename V, int R>
__global__ void __launch_bounds__(kVecBlock)
k_cg_update(int n, V* __restrict__ x, const V* __restrict__ p, V* __restrict__ r,
            const V* __restrict__ w, V* __restrict__ z, const V* __restrict__ dinv, int jacobi,
            const float* __restrict__ dsc, V* __restrict__ rt,
            const double* __restrict__ pspmv, int nspmv, double* __restrict__ pvec,
            const Scal* __restrict__ in, Scal* __restrict__ out) {
    __shared__ double s_alpha;
    __shared__ int s_stop;
    if (carry_if_done(in, out)) return;
    double pap[1];
    sum_partials<1>(pspmv, nspmv, pap);
    if (threadIdx.x == 0) {
        Scal s = *in;
        s.pAp = pap[0];
        s.rz_old = s.rz;
        if (!(pap[0] > 0.0)) {
            s.reason = (pap[0] != pap[0]) ? KSP_DIVERGED_NANORINF : KSP_DIVERGED_INDEFINITE_MAT;
            s.done = 1;
        } else {
            s.alpha = s.rz / pap[0];
        }
        s_alpha = s.alpha;
        s_stop = s.done;
        if (blockIdx.x == 0) *out = s;
    }
    __syncthreads();
    if (s_stop) return;
    const double alpha = s_alpha;
    const V a = (V)alpha;
    double s_rz = 0.0, s_rr = 0.0, s_zz = 0.0, s_zw = 0.0;
    auto row = [&](V& xi, V pi, V wi, V& ri, V di, float si, V& zi, V& ti) {
        xi = fma((V)alpha, pi, xi);
        ri = ri - a * wi;
        if (dsc) ti = ri * (V)si;
        if (jacobi) {
            zi = di * ri;
            s_rz += (double)ri * (double)zi;
            s_zz += (double)zi * (double)zi;
            s_zw += (double)zi * (double)wi;
        }
        s_rr += (double)ri * (double)ri;
    };
    const int ng = n / R, tid = blockIdx.x * blockDim.x + threadIdx.x, nt = gridDim.x * blockDim.x;
    for (int g = tid; g < ng; g += nt) {
        const int i = g * R;
        const Rows<V, R> pv = ld_rows<R>(p, i), wv = ld_rows<R>(w, i);
        Rows<V, R> rv = ld_rows<R>(r, i), dv = {}, zv = {}, tv = {};
        Rows<V, R> xv = ld_rows<R>((const V*)x, i);
        Rows<float, R> sv = {};
        if (jacobi) dv = ld_rows<R>(dinv, i);
        if (dsc) sv = ld_rows<R>(dsc, i);
#pragma unroll
        for (int u = 0; u < R; u++) {
            V xi = xv[u];
            V ri = rv[u], zi = (V)0, ti = (V)0;
            row(xi, pv[u], wv[u], ri, dv[u], sv[u], zi, ti);
            xv[u] = xi;
            rv[u] = ri;
            zv[u] = zi;
            tv[u] = ti;
        }
        st_rows<R>(x, i, xv);
        st_rows<R>(r, i, rv);
        if (jacobi) st_rows<R>(z, i, zv);
        if (dsc) st_rows<R>(rt, i, tv);
    }
    for (int i = ng * R + tid; i < n; i += nt) {
        V xi = x[i];
        V ri = r[i], zi = (V)0, ti = (V)0;
        row(xi, p[i], w[i], ri, jacobi ? dinv[i] : (V)0, dsc ? dsc[i] : 0.f, zi, ti);
        x[i] = xi;
        r[i] = ri;
        if (jacobi) z[i] = zi;
        if (dsc) rt[i] = ti;
    }
    double v[4] = {s_rz, s_rr, s_zz, s_zw};
    block_partials<4>(v, pvec);
}

// r.z, r.r, z.z, z.w from the update partials (or r.z, z.z, z.w from the
// learned apply), convergence test and beta in every block; p = z + beta p
// (R rows per thread and step, as k_cg_update).
template <typename V, int R>
__global__ void __launch_bounds__(kVecBlock)
k_cg_p(int n, const V* __restrict__ z, V* __restrict__ p, const double* __restrict__ pvec, int nvec,
       const double* __restrict__ ppc, int npc, const Scal* __restrict__ in, Scal* __restrict__ out) {
    __shared__ double s_beta;
    __shared__ int s_stop;
    // The thread's first group of rows does not depend on the scalars: its
    // loads are issued before the done flag's and the partial sums', so the
    // three are in flight together (one memory round trip, not two).
    const int ng = n / R, tid = blockIdx.x * blockDim.x + threadIdx.x, nt = gridDim.x * blockDim.x;
    Rows<V, R> zv = {}, pv = {};
    auto load = [&](int i) {
        zv = ld_rows<R>(z, i);
        pv = ld_rows<R>((const V*)p, i);
    };
    if (tid < ng) load(tid * R);
    if (carry_if_done(in, out)) return;
    double u[4], q[3] = {0.0, 0.0, 0.0};
    sum_partials<4>(pvec, nvec, u);
    if (ppc) sum_partials<3>(ppc, npc, q);
    if (threadIdx.x == 0) {
        Scal s = *in;
        s.rr = u[1];
        if (ppc) {
            s.rz = q[0];
            s.zz = q[1];
            s.zw = q[2];
        } else {
            s.rz = u[0];
            s.zz = u[2];
            s.zw = u[3];
        }
        cg_test_and_beta(&s);
        s_beta = s.beta;
        s_stop = s.done;
        if (blockIdx.x == 0) *out = s;
    }
    __syncthreads();
    if (s_stop) return;
    const V beta = (V)s_beta;
    for (int g = tid; g < ng; g += nt) {
        const int i = g * R;
        if (g != tid) load(i);
#pragma unroll
        for (int k = 0; k < R; k++) pv[k] = fma(beta, pv[k], zv[k]);
        st_rows<R>(p, i, pv);
    }
    for (int i = ng * R + tid; i < n; i += nt) p[i] = fma(beta, p[i], z[i]);
}

// Single-reduction CG (Chronopoulos & Gear 1989), the recurrence PETSc's
// KSPPIPECG builds on: with u = M r, w = A u, gamma = r.u and delta = u.w
// from one reduction phase,
//   beta = gamma / gamma_old, alpha = gamma / (delta - beta gamma / alpha_old),
//   p = u + beta p, s = w + beta s (s = A p), x += alpha p, r -= alpha s,
// then u = M r (Jacobi here; a factor by its own kernels) and w = A u (the
// next product).  One vector kernel an iteration instead of k_cg_update
// and k_cg_p.  The previous iteration's sums: delta (the product's u.w,
// pspmv), gamma, u.u and r.r (pg: this kernel's own region from the
// previous iteration, or the factor apply's r.z, z.z with pc_layout set,
// then r.r from prr); first: the iteration after a restart (gamma, r.r, u.u
// already in the scalars, p = u, s = w).  Partials r.u, r.r, u.u, 0 to
// pvec (a region other than pg and prr).
template <typename V, int R>
__global__ void __launch_bounds__(kVecBlock)
k_cg1_vec(int n, double* __restrict__ x, V* __restrict__ p, V* __restrict__ s, V* __restrict__ r,
          V* __restrict__ u, const V* __restrict__ w, const V* __restrict__ dinv, int jacobi,
          const float* __restrict__ dsc, V* __restrict__ rt, const double* __restrict__ pspmv, int nspmv,
          const double* __restrict__ pg, int ng, int pc_layout, const double* __restrict__ prr, int nrr, int first,
          double* __restrict__ pvec, const Scal* __restrict__ in, Scal* __restrict__ out) {
    __shared__ double s_ab[2];
    __shared__ int s_stop;
    // the first group's loads in flight with the scalars' (as k_cg_p)
    const int ng_ = n / R, tid = blockIdx.x * blockDim.x + threadIdx.x, nt = gridDim.x * blockDim.x;
    Rows<V, R> pv = {}, sv = {}, rv = {}, uv = {}, wv = {}, dv = {};
    Rows<double, R> xv = {};
    Rows<float, R> fv = {};
    auto load = [&](int i) {
        pv = ld_rows<R>((const V*)p, i);
        sv = ld_rows<R>((const V*)s, i);
        rv = ld_rows<R>((const V*)r, i);
        uv = ld_rows<R>((const V*)u, i);
        wv = ld_rows<R>(w, i);
        xv = ld_rows<R>((const double*)x, i);
        if (jacobi) dv = ld_rows<R>(dinv, i);
        if (dsc) fv = ld_rows<R>(dsc, i);
    };
    if (tid < ng_) load(tid * R);
    if (carry_if_done(in, out)) return;
    double dl[1], g[3] = {0.0, 0.0, 0.0}, q[3] = {0.0, 0.0, 0.0};
    sum_partials<1>(pspmv, nspmv, dl);
    if (!first) {
        sum_partials<3>(pg, ng, g);
        if (pc_layout) sum_partials<3>(prr, nrr, q);
    }
    if (threadIdx.x == 0) {
        Scal sc = *in;
        if (!first) {
            sc.rz_old = sc.rz;
            sc.rz = g[0];
            sc.rr = pc_layout ? q[1] : g[1];
            sc.zz = pc_layout ? g[1] : g[2];
            cg_test(&sc);
        }
        const double delta = dl[0];
        sc.pAp = delta;
        if (!sc.done) {
            const double beta = first ? 0.0 : sc.rz / sc.rz_old;
            const double denom = first ? delta : delta - beta * sc.rz / sc.alpha;
            if (sc.rz < 0.0) {
                sc.reason = KSP_DIVERGED_INDEFINITE_PC;
                sc.done = 1;
            } else if (!(denom > 0.0)) {
                sc.reason = (denom != denom) ? KSP_DIVERGED_NANORINF : KSP_DIVERGED_INDEFINITE_MAT;
                sc.done = 1;
            } else {
                sc.beta = beta;
                sc.alpha = sc.rz / denom;
            }
        }
        s_ab[0] = sc.alpha;
        s_ab[1] = sc.beta;
        s_stop = sc.done;
        if (blockIdx.x == 0) *out = sc;
    }
    __syncthreads();
    if (s_stop) return;
    const double alpha = s_ab[0];
    const V a = (V)alpha, b = (V)s_ab[1];
    double s_g = 0.0, s_rr = 0.0, s_uu = 0.0;
    auto row = [&](double& xi, V& pi, V& si, V ui, V wi, V& ri, V di, float dsi, V& uo, V& ti) {
        pi = first ? ui : fma(b, pi, ui);
        si = first ? wi : fma(b, si, wi);
        xi = fma(alpha, (double)pi, xi);
        ri = ri - a * si;
        if (dsc) ti = ri * (V)dsi;
        if (jacobi) {
            uo = di * ri;
            s_g += (double)ri * (double)uo;
            s_uu += (double)uo * (double)uo;
        }
        s_rr += (double)ri * (double)ri;
    };
    for (int gi = tid; gi < ng_; gi += nt) {
        const int i = gi * R;
        if (gi != tid) load(i);
        Rows<V, R> uo = {}, tv = {};
#pragma unroll
        for (int k = 0; k < R; k++) {
            double xi = xv[k];
            V pi = pv[k], si = sv[k], ri = rv[k], uk = (V)0, ti = (V)0;
            row(xi, pi, si, uv[k], wv[k], ri, dv[k], fv[k], uk, ti);
            xv[k] = xi;
            pv[k] = pi;
            sv[k] = si;
            rv[k] = ri;
            uo[k] = uk;
            tv[k] = ti;
        }
        st_rows<R>(x, i, xv);
        st_rows<R>(p, i, pv);
        st_rows<R>(s, i, sv);
        st_rows<R>(r, i, rv);
        if (jacobi) st_rows<R>(u, i, uo);
        if (dsc) st_rows<R>(rt, i, tv);
    }
    for (int i = ng_ * R + tid; i < n; i += nt) {
        double xi = x[i];
        V pi = p[i], si = s[i], ri = r[i], uk = (V)0, ti = (V)0;
        row(xi, pi, si, u[i], w[i], ri, jacobi ? dinv[i] : (V)0, dsc ? dsc[i] : 0.f, uk, ti);
        x[i] = xi;
        p[i] = pi;
        s[i] = si;
        r[i] = ri;
        if (jacobi) u[i] = uk;
        if (dsc) rt[i] = ti;
    }
    double v[4] = {s_g, s_rr, s_uu, 0.0};
    block_partials<4>(v, pvec);
}

template <typename V>
__global__ void __launch_bounds__(kVecBlock)
k_copy(int n, const V* __restrict__ a, V* __restrict__ b) {
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) b[i] = a[i];
}

template <typename V>
__global__ void __launch_bounds__(kVecBlock)
k_fill(int n, V v, V* __restrict__ a) {
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) a[i] = v;
}

// r_V = (V) r64 ; z = D^-1 r ; partials r.z, r.r, z.z
template <typename V>
__global__ void __launch_bounds__(kVecBlock)
k_restart_vecs(int n, const double* __restrict__ r64, V* __restrict__ r, V* __restrict__ z,
               const V* __restrict__ dinv, int jacobi, const float* __restrict__ dsc, V* __restrict__ rt,
               double* __restrict__ pvec) {
    double s_rz = 0.0, s_rr = 0.0, s_zz = 0.0;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        const V ri = (V)r64[i];
        r[i] = ri;
        if (dsc) rt[i] = ri * (V)dsc[i];
        s_rr += (double)ri * (double)ri;
        if (jacobi) {
            const V zi = dinv[i] * ri;
            z[i] = zi;
            s_rz += (double)ri * (double)zi;
            s_zz += (double)zi * (double)zi;
        }
    }
    double v[3] = {s_rz, s_rr, s_zz};
    block_partials<3>(v, pvec);
}

__global__ void __launch_bounds__(kVecBlock)
k_dot2(int n, const double* __restrict__ a, double* __restrict__ partials) {
    double s = 0.0;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
        s += a[i] * a[i];
    double v[1] = {s};
    block_partials<1>(v, partials);
}

// One-block finalisation of the rarely used reductions, in place on the
// current scalar slot.
enum FinOp { FIN_TRUE_RR = 0, FIN_BB = 1, FIN_RESTART = 2, FIN_LEARNED = 3 };

__device__ __forceinline__ void set_norm0(Scal& s) {
    const double nrm = s.norm_type == FLGPU_NORM_UNPRECONDITIONED ? sqrt(s.rr) : sqrt(s.zz);
    s.rnorm0 = nrm;
    s.last_norm = nrm;
}

__global__ void __launch_bounds__(kVecBlock)
k_finalize(int op, const double* __restrict__ partials, int nparts, int jacobi, int norm0, Scal* sc) {
    double v[3];
    if (op == FIN_TRUE_RR || op == FIN_BB) {
        double one[1];
        sum_partials<1>(partials + (op == FIN_TRUE_RR ? kPartStride : 0), nparts, one);
        if (threadIdx.x == 0) {
            if (op == FIN_TRUE_RR)
                sc->true_rr = one[0];
            else
                sc->bb = one[0];
        }
        return;
    }
    sum_partials<3>(partials, nparts, v);
    if (threadIdx.x != 0) return;
    Scal s = *sc;
    if (op == FIN_RESTART) {  // r.z, r.r, z.z
        s.rr = v[1];
        if (jacobi) {
            s.rz = v[0];
            s.zz = v[2];
        }
    } else {  // learned apply: r.z, z.z, z.w
        s.rz = v[0];
        s.zz = v[1];
        s.zw = v[2];
    }
    if (norm0) set_norm0(s);
    *sc = s;
}

__global__ void __launch_bounds__(kVecBlock)
k_finish_x(int n, double* __restrict__ x, const double* __restrict__ rhs,
           const int32_t* __restrict__ inactive, int have_inactive) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n && have_inactive && inactive[i]) x[i] = rhs[i];
}

// ---------------------------------------------------------------------
// Multicolour SSOR (the "sor" preconditioner)
// ---------------------------------------------------------------------
//
// PETSc's PCSOR, which femtools/Solvers.F90:1844-1854 selects for "sor",
// applies one local symmetric Gauss-Seidel sweep (omega = 1) in the natural
// row order -- a sequential recurrence.  Here the rows are taken colour by
// colour, ascending inside a colour, where the colours are a first-fit
// colouring of the matrix graph in row order (rows of one colour share no
// entry, so a colour's rows update independently): the same sweep on the
// symmetrically permuted matrix, again symmetric positive definite for an
// SPD system, one launch per colour.
//   forward:  z = 0; for c = 0 .. C-1:  z_i = (r_i - sum_{j != i} a_ij z_j) / a_ii
//   backward:        for c = C-2 .. 0:  the same with the current z
// (the last colour's backward step would recompute its forward values).
// Each row is summed by one thread in entry order with explicit fused
// multiply-adds, the order of the oracle's restatement.

template <typename V>
__global__ void __launch_bounds__(kVecBlock)
k_fill_done(int n, V* __restrict__ a, const int* __restrict__ done_flag) {
    if (done_flag && *done_flag) return;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) a[i] = (V)0;
}

template <typename V>
__global__ void __launch_bounds__(256)
k_sor_colour(int nr, const int32_t* __restrict__ rows, const int32_t* __restrict__ rowptr,
             const int32_t* __restrict__ col, const V* __restrict__ val, const V* __restrict__ r, V* __restrict__ z,
             const int* __restrict__ done_flag) {
    if (done_flag && *done_flag) return;
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= nr) return;
    const int i = rows[t];
    V acc = r[i], aii = (V)0;
    for (int k = rowptr[i]; k < rowptr[i + 1]; k++) {
        const int j = col[k];
        if (j == i)
            aii = val[k];
        else
            acc = fma(-val[k], z[j], acc);
    }
    z[i] = acc / aii;
}

// Per-block sums r.z, z.z, z.w (w may be null: z.w = 0).
template <typename V>
__global__ void __launch_bounds__(kVecBlock)
k_pc_dots(int n, const V* __restrict__ r, const V* __restrict__ z, const V* __restrict__ w,
          double* __restrict__ partials, const int* __restrict__ done_flag) {
    if (done_flag && *done_flag) return;
    double s_rz = 0.0, s_zz = 0.0, s_zw = 0.0;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        const double zi = (double)z[i];
        s_rz += (double)r[i] * zi;
        s_zz += zi * zi;
        if (w) s_zw += zi * (double)w[i];
    }
    double v[3] = {s_rz, s_zz, s_zw};
    block_partials<3>(v, partials);
}

// ---------------------------------------------------------------------
// host side
// ---------------------------------------------------------------------

static int lanes_for(const flgpu_ctx* c) {
    const double avg = (double)c->nnz / std::max(1, c->rows);
    if (avg <= 6) return 4;
    if (avg <= 12) return 8;
    if (avg <= 24) return 16;
    if (avg <= 48) return 32;
    return 64;
}

static int spmv_grid(const flgpu_ctx* c) {
    return c->kw->use_sell ? sell_grid(c->kw->nsl) : stream_grid(c->kw->nrb);
}

// The sliced ELL values matching a CSR value array of the solver matrix
// (null for any other array, which takes the CSR product).
static const double* sell_values(const KrylovWork* k, const double* val) {
    return k->use_sell && val == k->val64.p ? k->sval64.p : nullptr;
}
static const float* sell_values(const KrylovWork* k, const float* val) {
    return k->use_sell && val == k->val32.p ? k->sval32.p : nullptr;
}

template <typename V, typename CT, typename XV, typename YV>
static void sell_launch(flgpu_ctx* c, KrylovWork* k, const V* sval, const CT* col, const XV* x, YV* y,
                        const double* bsub, int want_dot, const int* done_flag) {
    const SellOp<V, CT> m{k->nsl, c->rows, k->sptr.p, c->rowptr.p, col, sval, nullptr};
    const int grid = sell_grid(k->nsl);
    if (want_dot)
        hipLaunchKernelGGL((k_sell<V, V, CT, XV, YV, DOT_XY>), dim3(grid), dim3(kRowBlock), 0, c->stream, m, x, y,
                           bsub, (const V*)nullptr, (const V*)nullptr, k->part(PART_SPMV), done_flag);
    else
        hipLaunchKernelGGL((k_sell<V, V, CT, XV, YV, DOT_NONE>), dim3(grid), dim3(kRowBlock), 0, c->stream, m, x,
                           y, bsub, (const V*)nullptr, (const V*)nullptr, k->part(PART_SPMV), done_flag);
}

template <typename AV, typename CT>
static CsrOp<AV, CT> matrix_op(flgpu_ctx* c, KrylovWork* k, const AV* val, const CT* col) {
    return CsrOp<AV, CT>{k->nrb, k->rb.p, k->rbk.p, c->rowptr.p, col, val, nullptr, nullptr};
}

template <typename V, typename CT, typename XV, typename YV>
static void spmv_launch(flgpu_ctx* c, KrylovWork* k, const V* val, const CT* col, const XV* x, YV* y,
                        const double* bsub, int want_dot, const int* done_flag) {
    const int grid = stream_grid(k->nrb);
    const CsrOp<V, CT> m = matrix_op<V, CT>(c, k, val, col);
    if (want_dot)
        hipLaunchKernelGGL((k_csr_stream<V, V, CT, XV, YV, DOT_XY>), dim3(grid), dim3(kRowBlock), 0, c->stream, m,
                           x, y, bsub, (const V*)nullptr, (const V*)nullptr, k->part(PART_SPMV), done_flag);
    else
        hipLaunchKernelGGL((k_csr_stream<V, V, CT, XV, YV, DOT_NONE>), dim3(grid), dim3(kRowBlock), 0, c->stream,
                           m, x, y, bsub, (const V*)nullptr, (const V*)nullptr, k->part(PART_SPMV), done_flag);
}

// y = A x (or y = bsub - A x); with want_dot the per-block sums x.y, y.y
// go to the PART_SPMV region, spmv_grid(c) of them.  The column indices
// are the int16 offsets when they fit.
template <typename V, typename XV, typename YV>
static int spmv(flgpu_ctx* c, KrylovWork* k, const V* val, const XV* x, YV* y, const double* bsub,
                int want_dot, const int* done_flag) {
    if (const V* sv = sell_values(k, val)) {
        if (k->sell16)
            sell_launch<V, int16_t, XV, YV>(c, k, sv, k->scol16.p, x, y, bsub, want_dot, done_flag);
        else
            sell_launch<V, int32_t, XV, YV>(c, k, sv, k->scol32.p, x, y, bsub, want_dot, done_flag);
    } else if (k->use16)
        spmv_launch<V, int16_t, XV, YV>(c, k, val, k->col16.p, x, y, bsub, want_dot, done_flag);
    else
        spmv_launch<V, int32_t, XV, YV>(c, k, val, c->col.p, x, y, bsub, want_dot, done_flag);
    FLGPU_HIP(hipGetLastError());
    return FLGPU_OK;
}

// The sliced ELL plan of the solver matrix, kept when its padding is at
// most kSellMaxFill times the entries (FLGPU_SPMV=csr keeps the CSR
// stream product throughout).
constexpr double kSellMaxFill = 1.2;

static int sell_setup(flgpu_ctx* c, KrylovWork* k) {
    const char* e = getenv("FLGPU_SPMV");
    k->use_sell = false;
    k->sptr.release();
    k->ssrc.release();
    k->scol16.release();
    k->scol32.release();
    k->sval64.release();
    k->sval32.release();
    if (e && std::string(e) == "csr") return FLGPU_OK;
    SellHost h;
    sell_plan(c->h_rowptr.data(), c->h_col.data(), c->rows, nullptr, h);
    if (h.fill(c->nnz) > kSellMaxFill) return FLGPU_OK;
    FLGPU_TRY(k->sptr.upload(h.sptr.data(), h.sptr.size(), c->stream));
    FLGPU_TRY(k->ssrc.upload(h.src.data(), h.src.size(), c->stream));
    if (h.use16) FLGPU_TRY(k->scol16.upload(h.c16.data(), h.c16.size(), c->stream));
    else FLGPU_TRY(k->scol32.upload(h.c32.data(), h.c32.size(), c->stream));
    FLGPU_HIP(hipStreamSynchronize(c->stream));  // h is freed on return
    k->nsl = h.nsl;
    k->nslot = h.nslot;
    k->sell16 = h.use16;
    k->use_sell = true;
    return FLGPU_OK;
}

static int ensure_work(flgpu_ctx* c, bool mixed) {
    if (!c->kw) c->kw = new KrylovWork();
    KrylovWork* k = c->kw;
    const int n = c->rows;
    if (k->n != n || k->nnz != c->nnz) {
        k->n = n;
        k->nnz = c->nnz;
        k->have_x = false;
    }
    FLGPU_TRY(k->val64.alloc(c->nnz));
    FLGPU_TRY(k->b.alloc(n));
    FLGPU_TRY(k->x.alloc(n));
    FLGPU_TRY(k->r.alloc(n));
    FLGPU_TRY(k->z.alloc(n));
    FLGPU_TRY(k->p.alloc(n));
    FLGPU_TRY(k->w.alloc(n));
    FLGPU_TRY(k->dinv.alloc(n));
    FLGPU_TRY(k->y.alloc(n));
    FLGPU_TRY(k->rt.alloc(n));
    if (mixed) {
        FLGPU_TRY(k->rt32.alloc(n));
        FLGPU_TRY(k->val32.alloc(c->nnz));
        FLGPU_TRY(k->r32.alloc(n));
        FLGPU_TRY(k->z32.alloc(n));
        FLGPU_TRY(k->p32.alloc(n));
        FLGPU_TRY(k->w32.alloc(n));
        FLGPU_TRY(k->dinv32.alloc(n));
        FLGPU_TRY(k->y32.alloc(n));
    }
    FLGPU_TRY(k->partials.reserve((size_t)kPartStride * PART_COMPONENTS));
    if (k->rb_key != c->sparsity_gen) {
        std::vector<int32_t> rb, rbk;
        stream_row_blocks(c->h_rowptr.data(), n, rb, rbk);
        FLGPU_TRY(k->rb.upload(rb.data(), rb.size(), c->stream));
        FLGPU_TRY(k->rbk.upload(rbk.data(), rbk.size(), c->stream));
        std::vector<int16_t> c16;
        k->use16 = compress_columns(c->h_rowptr.data(), c->h_col.data(), rb, c16);
        if (k->use16) FLGPU_TRY(k->col16.upload(c16.data(), c16.size(), c->stream));
        else k->col16.release();
        FLGPU_HIP(hipStreamSynchronize(c->stream));  // rb and c16 are freed on return
        k->nrb = (int)rb.size() - 1;
        FLGPU_TRY(sell_setup(c, k));
        k->rb_key = c->sparsity_gen;
    }
    if (k->use_sell) {
        FLGPU_TRY(k->sval64.alloc((size_t)k->nslot));
        if (mixed) FLGPU_TRY(k->sval32.alloc((size_t)k->nslot));
    }
    FLGPU_TRY(k->scal.alloc(2));
    if (!k->h_scal) {
        FLGPU_HIP(hipHostMalloc((void**)&k->h_scal, sizeof(Scal), hipHostMallocDefault));
        k->h_scal_owned = true;
    }
    k->nblk = (int)std::min<int64_t>(kMaxVecGrid, (n + kVecBlock - 1) / kVecBlock);
    k->nblk = std::max(k->nblk, 1);
    return FLGPU_OK;
}

static int read_scal(flgpu_ctx* c, KrylovWork* k) {
    FLGPU_HIP(hipMemcpyAsync(k->h_scal, k->S(), sizeof(Scal), hipMemcpyDeviceToHost, c->stream));
    FLGPU_HIP(hipStreamSynchronize(c->stream));
    return FLGPU_OK;
}

static int write_scal(flgpu_ctx* c, KrylovWork* k) {
    FLGPU_HIP(hipMemcpyAsync(k->S(), k->h_scal, sizeof(Scal), hipMemcpyHostToDevice, c->stream));
    // a pageable host copy is read by the copy after the call returns
    if (!k->h_scal_owned) FLGPU_HIP(hipStreamSynchronize(c->stream));
    return FLGPU_OK;
}

static int finalize(flgpu_ctx* c, KrylovWork* k, int op, int region, int nparts, bool jacobi, bool norm0) {
    hipLaunchKernelGGL(k_finalize, dim3(1), dim3(kVecBlock), 0, c->stream, op, k->part(region), nparts,
                       jacobi ? 1 : 0, norm0 ? 1 : 0, k->S());
    FLGPU_HIP(hipGetLastError());
    return FLGPU_OK;
}

// Build the reduced (PETSc-side) system from the assembled one.
static int build_reduced(flgpu_ctx* c, KrylovWork* k, bool mixed, bool zero_x) {
    const int n = c->rows;
    hipLaunchKernelGGL(k_diag_minmax, dim3(k->nblk), dim3(kVecBlock), 0, c->stream, n, c->diag.p, c->A.p,
                       k->part(PART_MISC));
    hipLaunchKernelGGL(k_pivot, dim3(1), dim3(64), 0, c->stream, k->nblk, k->part(PART_MISC), k->S());
    hipLaunchKernelGGL(k_build, dim3(stream_grid(k->nrb)), dim3(kRowBlock), 0, c->stream, k->nrb, k->rb.p,
                       c->rowptr.p, c->col.p, c->A.p, c->rhs.p, c->inactive.p, c->have_inactive ? 1 : 0, k->S(),
                       k->val64.p, mixed ? k->val32.p : nullptr, k->b.p, k->dinv.p, mixed ? k->dinv32.p : nullptr,
                       k->x.p, zero_x ? 1 : 0);
    if (k->use_sell) {
        const unsigned g = grid_for(k->nslot, 256);
        hipLaunchKernelGGL((k_sell_gather<double, double>), dim3(g), dim3(256), 0, c->stream, k->nslot, k->ssrc.p,
                           k->val64.p, k->sval64.p);
        if (mixed)
            hipLaunchKernelGGL((k_sell_gather<float, float>), dim3(g), dim3(256), 0, c->stream, k->nslot,
                               k->ssrc.p, k->val32.p, k->sval32.p);
    }
    FLGPU_HIP(hipGetLastError());
    return FLGPU_OK;
}

// x += dx; dx = 0 (the mixed solves' FP32 increment into the FP64 solution)
__global__ void __launch_bounds__(kVecBlock)
k_fold_x(int n, double* __restrict__ x, float* __restrict__ dx) {
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        x[i] = x[i] + (double)dx[i];
        dx[i] = 0.f;
    }
}

// True residual r64 = b - A x (FP64) and its squared norm in scal.true_rr.
static int true_residual(flgpu_ctx* c, KrylovWork* k) {
    if (k->dx_on) {
        hipLaunchKernelGGL(k_fold_x, dim3(k->nblk), dim3(kVecBlock), 0, c->stream, c->rows, k->x.p, k->dx32.p);
        FLGPU_HIP(hipGetLastError());
    }
    FLGPU_TRY((spmv<double, double, double>(c, k, k->val64.p, k->x.p, k->r.p, k->b.p, 1, nullptr)));
    return finalize(c, k, FIN_TRUE_RR, PART_SPMV, spmv_grid(c), false, false);
}

static int norm_b(flgpu_ctx* c, KrylovWork* k) {
    hipLaunchKernelGGL(k_dot2, dim3(k->nblk), dim3(kVecBlock), 0, c->stream, c->rows, k->b.p, k->part(PART_MISC));
    return finalize(c, k, FIN_BB, PART_MISC, k->nblk, false, false);
}

// ------------------------- multicolour SSOR ------------------------------

// First-fit colouring of the rows of a CSR pattern in row order: row i
// takes the smallest colour that none of its columns j < i holds.  Rows are
// then listed colour by colour, ascending inside a colour.
static void colour_rows(int n, const int32_t* rowptr, const int32_t* col, std::vector<int32_t>& ptr,
                 std::vector<int32_t>& rows) {
    std::vector<int32_t> colour(n, -1), mark;  // mark[q] == i: colour q is taken next to row i
    int ncol = 0;
    for (int i = 0; i < n; i++) {
        for (int k = rowptr[i]; k < rowptr[i + 1]; k++) {
            const int j = col[k];
            if (j < 0 || j >= i) continue;
            const int q = colour[j];
            if (q >= (int)mark.size()) mark.resize(q + 1, -1);
            mark[q] = i;
        }
        int q = 0;
        while (q < (int)mark.size() && mark[q] == i) q++;
        colour[i] = q;
        ncol = std::max(ncol, q + 1);
    }
    ptr.assign(ncol + 1, 0);
    for (int i = 0; i < n; i++) ptr[colour[i] + 1]++;
    for (int c = 0; c < ncol; c++) ptr[c + 1] += ptr[c];
    rows.resize(n);
    std::vector<int32_t> pos(ptr.begin(), ptr.end() - 1);
    for (int i = 0; i < n; i++) rows[pos[colour[i]]++] = i;
}

static int ssor_prepare(flgpu_ctx* c, KrylovWork* k) {
    if (k->mc_key == c->sparsity_gen) return FLGPU_OK;
    if (!c->all_diag) return set_error(FLGPU_EUNSUPPORTED, "sor preconditioner: a row has no diagonal entry");
    std::vector<int32_t> rows;
    colour_rows(c->rows, c->h_rowptr.data(), c->h_col.data(), k->mc_ptr, rows);
    FLGPU_TRY(k->mc_rows.upload(rows.data(), rows.size(), c->stream));
    FLGPU_HIP(hipStreamSynchronize(c->stream));  // rows is freed on return
    k->mc_key = c->sparsity_gen;
    return FLGPU_OK;
}

// z = M^-1 r (the forward and backward sweeps); with done_flag set the
// kernels return at once.
template <typename V>
static int ssor_sweeps(flgpu_ctx* c, KrylovWork* k, const V* val, const V* r, V* z, const int* done_flag) {
    const int ncol = (int)k->mc_ptr.size() - 1;
    hipLaunchKernelGGL((k_fill_done<V>), dim3(k->nblk), dim3(kVecBlock), 0, c->stream, c->rows, z, done_flag);
    auto sweep = [&](int q) {
        const int nr = k->mc_ptr[q + 1] - k->mc_ptr[q];
        if (nr > 0)
            hipLaunchKernelGGL((k_sor_colour<V>), dim3(grid_for(nr, 256)), dim3(256), 0, c->stream, nr,
                               k->mc_rows.p + k->mc_ptr[q], c->rowptr.p, c->col.p, val, r, z, done_flag);
    };
    for (int q = 0; q < ncol; q++) sweep(q);
    for (int q = ncol - 2; q >= 0; q--) sweep(q);
    FLGPU_HIP(hipGetLastError());
    return FLGPU_OK;
}

template <typename V>
static int ssor_apply(flgpu_ctx* c, KrylovWork* k, const V* val, const V* r, V* z, const V* w, double* partials,
                      int* nparts, const int* done_flag) {
    FLGPU_TRY(ssor_sweeps<V>(c, k, val, r, z, done_flag));
    hipLaunchKernelGGL((k_pc_dots<V>), dim3(k->nblk), dim3(kVecBlock), 0, c->stream, c->rows, r, z, w, partials,
                       done_flag);
    FLGPU_HIP(hipGetLastError());
    *nparts = k->nblk;
    return FLGPU_OK;
}

// ------------------------- (F)PCG ----------------------------------------

template <typename V>
struct CgVecs {
    V *r, *z, *p, *w, *dinv, *y, *rt;
    const V* val;
    V* s = nullptr;  // single-reduction CG only
    V* xw = nullptr; // what the vector kernels add alpha p to: x, or the FP32 increment
};

template <typename V>
static CgVecs<V> cg_vecs(KrylovWork* k);
template <>
CgVecs<double> cg_vecs<double>(KrylovWork* k) {
    return {k->r.p, k->z.p, k->p.p, k->w.p, k->dinv.p, k->y.p, k->rt.p, k->val64.p};
}
template <>
CgVecs<float> cg_vecs<float>(KrylovWork* k) {
    return {k->r32.p, k->z32.p, k->p32.p, k->w32.p, k->dinv32.p, k->y32.p, k->rt32.p, k->val32.p};
}

// The preconditioner of a (F)PCG iteration: the diagonal (Jacobi, or the
// identity as a unit diagonal), applied inside the vector kernels; or the
// learned factor or the multicolour SSOR sweeps, applied by kernels of their
// own after the residual update.
enum PcMode { PCM_DIAG = 0, PCM_LEARNED = 1, PCM_SSOR = 2 };

template <typename V>
static int ssor_apply(flgpu_ctx* c, KrylovWork* k, const V* val, const V* r, V* z, const V* w, double* partials,
                      int* nparts, const int* done_flag);

// z = M^-1 r for the learned factor (z = G^T (G r)) or SSOR; the per-block
// sums r.z, z.z, z.w (w = A p of this iteration) go to the PART_PC region.
// Returns the number of partial sums in *nparts.
template <typename V>
static int pc_apply(flgpu_ctx* c, KrylovWork* k, const CgVecs<V>& v, PcMode mode, const int* done_flag,
                    int* nparts) {
    if (mode == PCM_SSOR) return ssor_apply<V>(c, k, v.val, v.r, v.z, v.w, k->part(PART_PC), nparts, done_flag);
    return learned_precond_apply<V>(c, v.r, v.rt, v.z, v.y, v.w, k->part(PART_PC), nparts, done_flag);
}

// r = b - A x (FP64), r_V = r, z = M^-1 r with the sums r.z, r.r, z.z
// (and the initial norm when norm0) in the current scalar slot.
template <typename V>
static int restart_residual(flgpu_ctx* c, KrylovWork* k, const CgVecs<V>& v, PcMode mode, bool norm0,
                            bool recompute) {
    if (recompute) FLGPU_TRY(true_residual(c, k));
    const bool diag = mode == PCM_DIAG;
    const float* dsc = mode == PCM_LEARNED ? learned_scale(c) : nullptr;
    hipLaunchKernelGGL((k_restart_vecs<V>), dim3(k->nblk), dim3(kVecBlock), 0, c->stream, c->rows, k->r.p,
                       v.r, v.z, v.dinv, diag ? 1 : 0, dsc, v.rt, k->part(PART_VEC));
    FLGPU_HIP(hipGetLastError());
    FLGPU_TRY(finalize(c, k, FIN_RESTART, PART_VEC, k->nblk, diag, norm0 && diag));
    if (!diag) {
        int np = 0;
        FLGPU_TRY(pc_apply<V>(c, k, v, mode, nullptr, &np));
        k->npc = np;
        FLGPU_TRY(finalize(c, k, FIN_LEARNED, PART_PC, np, false, norm0));
    }
    return FLGPU_OK;
}

// The x the (F)PCG vector kernels update: FP64 x itself, or (mixed) the
// FP32 increment dx, zeroed here and added to x by every true residual
// (true_residual folds it first), so x is whole at each replacement and
// at the end of the solve.  (The single-reduction CG keeps its updates in
// FP64: with the FP32 increment its counts rose by 5-9% on the 3-D
// systems, against about 1 for (F)PCG.)
template <typename V>
static int x_work(flgpu_ctx* c, KrylovWork* k, CgVecs<V>& v) {
    if constexpr (std::is_same<V, float>::value) {
        FLGPU_TRY(k->dx32.alloc(c->rows));
        hipLaunchKernelGGL((k_fill<float>), dim3(k->nblk), dim3(kVecBlock), 0, c->stream, c->rows, 0.0f, k->dx32.p);
        FLGPU_HIP(hipGetLastError());
        v.xw = k->dx32.p;
        k->dx_on = true;
    } else {
        v.xw = k->x.p;
        k->dx_on = false;
    }
    return FLGPU_OK;
}

template <typename V>
static int cg_restart(flgpu_ctx* c, KrylovWork* k, const CgVecs<V>& v, PcMode mode, bool set_norm0) {
    FLGPU_TRY(restart_residual<V>(c, k, v, mode, set_norm0, true));
    hipLaunchKernelGGL((k_copy<V>), dim3(k->nblk), dim3(kVecBlock), 0, c->stream, c->rows, (const V*)v.z, v.p);
    FLGPU_HIP(hipGetLastError());
    return FLGPU_OK;
}

// True when every given array (null ones aside) starts on a multiple of R
// of its entries: the vector kernels then load R rows at once.
template <int R, typename... P>
static bool rows_aligned(const P*... ptrs) {
    return ((ptrs == nullptr || reinterpret_cast<uintptr_t>(ptrs) % (R * sizeof(P)) == 0) && ...);
}

// One (F)PCG iteration: w = A p; update x, r, z; [z = M^-1 r]; test,
// beta, p.  Four kernels (diagonal), six (learned) or 2C + 3 (SSOR over C
// colours), no host round trip.
template <typename V>
static int cg_iteration(flgpu_ctx* c, KrylovWork* k, const CgVecs<V>& v, PcMode mode) {
    const bool diag = mode == PCM_DIAG;
    FLGPU_TRY((spmv<V, V, V>(c, k, v.val, v.p, v.w, nullptr, 1, &k->S()->done)));
    Scal* in = k->S();
    k->cur ^= 1;
    const float* dsc = mode == PCM_LEARNED ? learned_scale(c) : nullptr;
    constexpr int R = 16 / sizeof(V);
    if (rows_aligned<R>(v.xw, v.p, v.r, v.w, v.z, v.dinv, dsc, v.rt))
        hipLaunchKernelGGL((k_cg_update<V, R>), dim3(k->nblk), dim3(kVecBlock), 0, c->stream, c->rows, v.xw, v.p,
                           v.r, v.w, v.z, v.dinv, diag ? 1 : 0, dsc, v.rt, k->part(PART_SPMV), spmv_grid(c),
                           k->part(PART_VEC), in, k->S());
    else
        hipLaunchKernelGGL((k_cg_update<V, 1>), dim3(k->nblk), dim3(kVecBlock), 0, c->stream, c->rows, v.xw, v.p,
                           v.r, v.w, v.z, v.dinv, diag ? 1 : 0, dsc, v.rt, k->part(PART_SPMV), spmv_grid(c),
                           k->part(PART_VEC), in, k->S());
    int npc = 0;
    if (!diag) FLGPU_TRY(pc_apply<V>(c, k, v, mode, &k->S()->done, &npc));
    in = k->S();
    k->cur ^= 1;
    if (rows_aligned<R>(v.z, v.p))
        hipLaunchKernelGGL((k_cg_p<V, R>), dim3(k->nblk), dim3(kVecBlock), 0, c->stream, c->rows, v.z, v.p,
                           k->part(PART_VEC), k->nblk, diag ? nullptr : k->part(PART_PC), npc, in, k->S());
    else
        hipLaunchKernelGGL((k_cg_p<V, 1>), dim3(k->nblk), dim3(kVecBlock), 0, c->stream, c->rows, v.z, v.p,
                           k->part(PART_VEC), k->nblk, diag ? nullptr : k->part(PART_PC), npc, in, k->S());
    FLGPU_HIP(hipGetLastError());
    return FLGPU_OK;
}

template <typename V>
static int run_cg(flgpu_ctx* c, const flgpu_solver_params* sp, flgpu_solve_info* info, bool learned) {
    KrylovWork* k = c->kw;
    const bool mixed = std::is_same<V, float>::value;
    CgVecs<V> v = cg_vecs<V>(k);
    FLGPU_TRY(x_work<V>(c, k, v));
    PcMode mode = learned ? PCM_LEARNED : sp->pc == FLGPU_PC_SOR ? PCM_SSOR : PCM_DIAG;
    bool use_learned = learned;
    const int check_every = sp->check_every > 0 ? sp->check_every : 16;
    const double stall_factor = sp->stall_factor > 0 ? sp->stall_factor : 0.9;
    const int stall_window = sp->stall_window > 0 ? sp->stall_window : 50;

    // identity preconditioner = Jacobi with dinv = 1: handled by the
    // "jacobi" kernels when pc == NONE by filling dinv with ones
    if (!learned && sp->pc == FLGPU_PC_NONE) {
        hipLaunchKernelGGL((k_fill<double>), dim3(k->nblk), dim3(kVecBlock), 0, c->stream, c->rows, 1.0, k->dinv.p);
        if (mixed)
            hipLaunchKernelGGL((k_fill<float>), dim3(k->nblk), dim3(kVecBlock), 0, c->stream, c->rows, 1.0f,
                               k->dinv32.p);
        FLGPU_HIP(hipGetLastError());
    }
    if (mode == PCM_SSOR) FLGPU_TRY(ssor_prepare(c, k));
    Scal& h = *k->h_scal;
    // keep the pivot computed by build_reduced
    FLGPU_TRY(read_scal(c, k));
    {
        const double pivot = h.pivot, dmin = h.dmin, dmax = h.dmax;
        std::memset(&h, 0, sizeof h);
        h.pivot = pivot;
        h.dmin = dmin;
        h.dmax = dmax;
    }
    h.max_its = sp->max_its > 0 ? sp->max_its : 10000;
    h.rtol = sp->rtol;
    h.atol = sp->atol;
    h.norm_type = sp->norm_type;
    h.flexible = (use_learned || sp->ksp == FLGPU_KSP_FCG) ? 1 : 0;
    FLGPU_TRY(write_scal(c, k));
    FLGPU_TRY(norm_b(c, k));
    FLGPU_TRY(cg_restart<V>(c, k, v, mode, true));
    FLGPU_TRY(read_scal(c, k));
    info->refinements = 0;
    info->fallback = 0;
    const double bnorm = std::sqrt(h.bb);
    if (h.last_norm == 0.0 || h.last_norm <= sp->atol) {
        // converged before the first iteration: the device copy of the
        // scalars is what the solve reports from, so it gets the reason too
        h.reason = KSP_CONVERGED_ATOL;
        h.done = 1;
        FLGPU_TRY(write_scal(c, k));
    }
    double rmax = std::sqrt(h.rr);          // reliable-update reference
    const double delta = 0.1;
    int window_its = 0;
    double window_norm = h.last_norm;
    auto t0 = std::chrono::steady_clock::now();
    while (!h.done) {
        for (int q = 0; q < check_every; q++) FLGPU_TRY(cg_iteration<V>(c, k, v, mode));
        FLGPU_TRY(read_scal(c, k));
        const double rnow = std::sqrt(h.rr);
        if (mixed && (h.done ? h.reason > 0 : rnow < delta * rmax)) {
            // reliable update: replace the FP32 residual by the FP64 one
            const int reason = h.reason, done = h.done;
            FLGPU_TRY(true_residual(c, k));
            FLGPU_TRY(read_scal(c, k));
            const double tr = std::sqrt(h.true_rr);
            info->refinements++;
            const double target = std::max(sp->rtol * (sp->norm_type == FLGPU_NORM_UNPRECONDITIONED ? h.rnorm0 : bnorm),
                                           sp->atol);
            // The FP32 recurrence said converged: the FP64 true residual
            // decides (for either norm type it is the unpreconditioned
            // ||b - A x|| against rtol times ||r0|| or ||b||, the
            // criterion the solve reports in true_relres).  Short of it,
            // the iteration goes on, or at max_its the solve has diverged.
            if (done && reason > 0 && tr > target) {
                if (h.its < h.max_its) {
                    h.done = 0;
                    h.reason = 0;
                } else {
                    h.reason = KSP_DIVERGED_ITS;
                }
                FLGPU_TRY(write_scal(c, k));
            }
            if (!h.done) {
                // The FP32 recurrence drifts from the true residual by about
                // eps32 times the residual of the last replacement.  After a
                // drop of more than 1e3 since then (a fast preconditioner
                // between two host checks) that drift is no longer small
                // against r, and the old direction p is not conjugate to the
                // replaced residual: the direction restarts from z.
                const bool restart_p = rnow < 1e-3 * rmax;
                FLGPU_TRY(restart_residual<V>(c, k, v, mode, false, false));
                if (restart_p) {
                    hipLaunchKernelGGL((k_copy<V>), dim3(k->nblk), dim3(kVecBlock), 0, c->stream, c->rows,
                                       (const V*)v.z, v.p);
                    FLGPU_HIP(hipGetLastError());
                }
                FLGPU_TRY(read_scal(c, k));
                rmax = std::sqrt(h.rr);
            }
        }
        if (use_learned && !h.done) {
            window_its += check_every;
            if (window_its >= stall_window) {
                if (h.last_norm > stall_factor * window_norm) {
                    // stalled: fall back to Jacobi-PCG from the current x
                    use_learned = false;
                    mode = PCM_DIAG;
                    info->fallback = 1;
                    h.flexible = sp->ksp == FLGPU_KSP_FCG ? 1 : 0;
                    FLGPU_TRY(write_scal(c, k));
                    FLGPU_TRY(cg_restart<V>(c, k, v, PCM_DIAG, false));
                    FLGPU_TRY(read_scal(c, k));
                    rmax = std::sqrt(h.rr);
                }
                window_its = 0;
                window_norm = h.last_norm;
            }
        }
    }
    FLGPU_TRY(true_residual(c, k));
    FLGPU_TRY(read_scal(c, k));
    auto t1 = std::chrono::steady_clock::now();
    info->seconds = std::chrono::duration<double>(t1 - t0).count();
    info->its = h.its;
    info->reason = h.reason;
    info->true_relres = bnorm > 0 ? std::sqrt(h.true_rr) / bnorm : std::sqrt(h.true_rr);
    info->precond_relres = h.rnorm0 > 0 ? h.last_norm / h.rnorm0 : h.last_norm;
    return FLGPU_OK;
}

// ------------------------- single-reduction CG ---------------------------

// The restart of the single-reduction recurrence: r = b - A x (FP64), r_V,
// u = M r with gamma, r.r, u.u in the scalars (restart_residual), then
// w = A u with u.w in the PART_SPMV partials; the next iteration takes
// p = u, s = w.
template <typename V>
static int cg1_restart(flgpu_ctx* c, KrylovWork* k, const CgVecs<V>& v, PcMode mode, bool norm0, bool recompute) {
    FLGPU_TRY(restart_residual<V>(c, k, v, mode, norm0, recompute));
    return spmv<V, V, V>(c, k, v.val, v.z, v.w, nullptr, 1, &k->S()->done);
}

// One iteration: the vector update (k_cg1_vec), [the factor or SSOR
// apply], the product w = A u.  Two kernels with the diagonal
// preconditioner, four with a factor.  reg: the partial-sum region the
// previous update wrote (read here), the other one is written.
template <typename V>
static int cg1_iteration(flgpu_ctx* c, KrylovWork* k, const CgVecs<V>& v, PcMode mode, bool first, int& reg) {
    const bool diag = mode == PCM_DIAG;
    const int reg_out = reg == PART_VEC ? PART_VEC2 : PART_VEC;
    Scal* in = k->S();
    k->cur ^= 1;
    const float* dsc = mode == PCM_LEARNED ? learned_scale(c) : nullptr;
    const double* pg = diag ? k->part(reg) : k->part(PART_PC);
    const int ng = diag ? k->nblk : k->npc;
    constexpr int R = 16 / sizeof(V);
    auto launch = [&](auto kern) {
        hipLaunchKernelGGL(kern, dim3(k->nblk), dim3(kVecBlock), 0, c->stream, c->rows, k->x.p, v.p, v.s, v.r, v.z,
                           (const V*)v.w, (const V*)v.dinv, diag ? 1 : 0, dsc, v.rt, (const double*)k->part(PART_SPMV),
                           spmv_grid(c), pg, ng, diag ? 0 : 1, (const double*)k->part(reg), k->nblk, first ? 1 : 0,
                           k->part(reg_out), (const Scal*)in, k->S());
    };
    if (rows_aligned<R>(k->x.p, v.p, v.s, v.r, v.z, v.w, v.dinv, dsc, v.rt))
        launch(&k_cg1_vec<V, R>);
    else
        launch(&k_cg1_vec<V, 1>);
    FLGPU_HIP(hipGetLastError());
    reg = reg_out;
    if (!diag) {
        int npc = 0;
        FLGPU_TRY(pc_apply<V>(c, k, v, mode, &k->S()->done, &npc));
        k->npc = npc;
    }
    return spmv<V, V, V>(c, k, v.val, v.z, v.w, nullptr, 1, &k->S()->done);
}

template <typename V>
static int run_cg1(flgpu_ctx* c, const flgpu_solver_params* sp, flgpu_solve_info* info, bool factor) {
    KrylovWork* k = c->kw;
    const bool mixed = std::is_same<V, float>::value;
    CgVecs<V> v = cg_vecs<V>(k);
    if (mixed) {
        FLGPU_TRY(k->s32.alloc(c->rows));
        v.s = reinterpret_cast<V*>(k->s32.p);
    } else {
        FLGPU_TRY(k->s.alloc(c->rows));
        v.s = reinterpret_cast<V*>(k->s.p);
    }
    PcMode mode = factor ? PCM_LEARNED : sp->pc == FLGPU_PC_SOR ? PCM_SSOR : PCM_DIAG;
    bool use_factor = factor;
    const int check_every = sp->check_every > 0 ? sp->check_every : 16;
    const double stall_factor = sp->stall_factor > 0 ? sp->stall_factor : 0.9;
    const int stall_window = sp->stall_window > 0 ? sp->stall_window : 50;
    if (!factor && sp->pc == FLGPU_PC_NONE) {
        hipLaunchKernelGGL((k_fill<double>), dim3(k->nblk), dim3(kVecBlock), 0, c->stream, c->rows, 1.0, k->dinv.p);
        if (mixed)
            hipLaunchKernelGGL((k_fill<float>), dim3(k->nblk), dim3(kVecBlock), 0, c->stream, c->rows, 1.0f,
                               k->dinv32.p);
        FLGPU_HIP(hipGetLastError());
    }
    if (mode == PCM_SSOR) FLGPU_TRY(ssor_prepare(c, k));
    Scal& h = *k->h_scal;
    FLGPU_TRY(read_scal(c, k));
    {
        const double pivot = h.pivot, dmin = h.dmin, dmax = h.dmax;
        std::memset(&h, 0, sizeof h);
        h.pivot = pivot;
        h.dmin = dmin;
        h.dmax = dmax;
    }
    h.max_its = sp->max_its > 0 ? sp->max_its : 10000;
    h.rtol = sp->rtol;
    h.atol = sp->atol;
    h.norm_type = sp->norm_type;
    FLGPU_TRY(write_scal(c, k));
    FLGPU_TRY(norm_b(c, k));
    FLGPU_TRY(cg1_restart<V>(c, k, v, mode, true, true));
    FLGPU_TRY(read_scal(c, k));
    info->refinements = 0;
    info->fallback = 0;
    const double bnorm = std::sqrt(h.bb);
    if (h.last_norm == 0.0 || h.last_norm <= sp->atol) {
        h.reason = KSP_CONVERGED_ATOL;
        h.done = 1;
        FLGPU_TRY(write_scal(c, k));
    }
    double rmax = std::sqrt(h.rr);
    const double delta = 0.1;
    int window_its = 0;
    double window_norm = h.last_norm;
    bool first = true;
    int reg = PART_VEC;  // restart_residual's partials
    auto t0 = std::chrono::steady_clock::now();
    while (!h.done) {
        for (int q = 0; q < check_every; q++) {
            FLGPU_TRY(cg1_iteration<V>(c, k, v, mode, first, reg));
            first = false;
        }
        FLGPU_TRY(read_scal(c, k));
        const double rnow = std::sqrt(h.rr);
        if (mixed && (h.done ? h.reason > 0 : rnow < delta * rmax)) {
            // reliable update, as in the (F)PCG loop: the FP64 true residual
            // replaces the FP32 one and decides convergence
            const int reason = h.reason, done = h.done;
            FLGPU_TRY(true_residual(c, k));
            FLGPU_TRY(read_scal(c, k));
            const double tr = std::sqrt(h.true_rr);
            info->refinements++;
            const double target = std::max(sp->rtol * (sp->norm_type == FLGPU_NORM_UNPRECONDITIONED ? h.rnorm0 : bnorm),
                                           sp->atol);
            if (done && reason > 0 && tr > target) {
                if (h.its < h.max_its) {
                    h.done = 0;
                    h.reason = 0;
                } else {
                    h.reason = KSP_DIVERGED_ITS;
                }
                FLGPU_TRY(write_scal(c, k));
            }
            if (!h.done) {
                // As in the (F)PCG loop, the direction survives a drop of
                // up to 1e3 since the last replacement: s = A p is then
                // recomputed (its FP32 recurrence drifts like r's) and the
                // next update forms beta from the replaced gamma and the
                // gamma that formed p.  After a larger drop the recurrence
                // restarts from the replaced residual (p = u, s = w); the
                // update that follows tests no residual, so the count takes
                // the replaced one here.
                const bool restart_p = rnow < 1e-3 * rmax;
                const double rz_p = h.rz;
                if (!restart_p) FLGPU_TRY((spmv<V, V, V>(c, k, v.val, v.p, v.s, nullptr, 0, &k->S()->done)));
                FLGPU_TRY(cg1_restart<V>(c, k, v, mode, false, false));
                FLGPU_TRY(read_scal(c, k));
                if (restart_p)
                    h.its += 1;
                else
                    h.rz = rz_p;
                FLGPU_TRY(write_scal(c, k));
                first = restart_p;
                reg = PART_VEC;
                rmax = std::sqrt(h.rr);
            }
        }
        if (use_factor && !h.done) {
            window_its += check_every;
            if (window_its >= stall_window) {
                if (h.last_norm > stall_factor * window_norm) {
                    use_factor = false;
                    mode = PCM_DIAG;
                    info->fallback = 1;
                    FLGPU_TRY(write_scal(c, k));
                    FLGPU_TRY(cg1_restart<V>(c, k, v, PCM_DIAG, false, true));
                    FLGPU_TRY(read_scal(c, k));
                    h.its += 1;
                    FLGPU_TRY(write_scal(c, k));
                    first = true;
                    reg = PART_VEC;
                    rmax = std::sqrt(h.rr);
                }
                window_its = 0;
                window_norm = h.last_norm;
            }
        }
    }
    FLGPU_TRY(true_residual(c, k));
    FLGPU_TRY(read_scal(c, k));
    auto t1 = std::chrono::steady_clock::now();
    info->seconds = std::chrono::duration<double>(t1 - t0).count();
    info->its = h.its;
    info->reason = h.reason;
    info->true_relres = bnorm > 0 ? std::sqrt(h.true_rr) / bnorm : std::sqrt(h.true_rr);
    info->precond_relres = h.rnorm0 > 0 ? h.last_norm / h.rnorm0 : h.last_norm;
    return FLGPU_OK;
}

// ------------------------- GMRES ------------------------------------------

template <typename T>
__device__ __forceinline__ T lane_sum(T v, int L) {
    for (int off = L >> 1; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    return v;
}

// w = D^-1 (A v)  [left Jacobi]
template <int L, typename V>
__global__ void __launch_bounds__(kRowBlock)
k_gm_matvec(int n, const int32_t* __restrict__ rowptr, const int32_t* __restrict__ col,
            const V* __restrict__ val, const V* __restrict__ dinv, const V* __restrict__ x,
            V* __restrict__ y) {
    constexpr int ROWS = kRowBlock / L;
    const int lane = threadIdx.x % L, rloc = threadIdx.x / L;
    for (int grp = blockIdx.x; grp * ROWS < n; grp += gridDim.x) {
        const int row = grp * ROWS + rloc;
        V acc = (V)0;
        if (row < n)
            for (int k = rowptr[row] + lane; k < rowptr[row + 1]; k += L) acc += val[k] * x[col[k]];
        acc = lane_sum<V>(acc, L);
        if (row < n && lane == 0) y[row] = dinv ? dinv[row] * acc : acc;
    }
}

// h[j] = w . V_j for j <= kk (one pass over w), partials per block
template <typename V>
__global__ void __launch_bounds__(kVecBlock)
k_gm_dots(int n, int kk, const V* __restrict__ Vb, const V* __restrict__ w,
          double* __restrict__ partials) {
    __shared__ double sh[kVecBlock / 64][1];
    for (int j = 0; j <= kk; j++) {
        const V* vj = Vb + (size_t)n * j;
        double s = 0.0;
        for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
            s += (double)w[i] * (double)vj[i];
        double a[1] = {s};
        block_sum<1>(a, sh);
        if (threadIdx.x == 0) partials[(size_t)j * gridDim.x + blockIdx.x] = a[0];
    }
}

// hcol[j] (+)= sum over blocks (fixed order); one block per j
__global__ void k_gm_reduce(int kk, int nblk, const double* __restrict__ partials,
                            double* __restrict__ hcol, int accumulate) {
    __shared__ double sh[kVecBlock / 64][1];
    const int j = blockIdx.x;
    if (j > kk) return;
    double s = 0.0;
    for (int b = threadIdx.x; b < nblk; b += blockDim.x) s += partials[(size_t)j * nblk + b];
    double a[1] = {s};
    block_sum<1>(a, sh);
    if (threadIdx.x == 0) hcol[j] = accumulate ? hcol[j] + a[0] : a[0];
}

// w -= sum_j h[j] V_j  (h from hcol, or its second-pass increment)
template <typename V>
__global__ void __launch_bounds__(kVecBlock)
k_gm_orth(int n, int kk, const V* __restrict__ Vb, const double* __restrict__ h, V* __restrict__ w) {
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        double acc = (double)w[i];
        for (int j = 0; j <= kk; j++) acc -= h[j] * (double)Vb[(size_t)n * j + i];
        w[i] = (V)acc;
    }
}

template <typename V>
__global__ void __launch_bounds__(kVecBlock)
k_gm_scale(int n, const V* __restrict__ w, V* __restrict__ out, const double* __restrict__ nrm) {
    const double s = nrm[0] > 0 ? 1.0 / nrm[0] : 0.0;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
        out[i] = (V)((double)w[i] * s);
}

template <typename V>
__global__ void __launch_bounds__(kVecBlock)
k_gm_norm(int n, const V* __restrict__ w, double* __restrict__ partials) {
    __shared__ double sh[kVecBlock / 64][1];
    double s = 0.0;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
        s += (double)w[i] * (double)w[i];
    double a[1] = {s};
    block_sum<1>(a, sh);
    if (threadIdx.x == 0) partials[blockIdx.x] = a[0];
}

__global__ void k_gm_sqrt_reduce(int nblk, const double* __restrict__ partials, double* __restrict__ out) {
    __shared__ double sh[kVecBlock / 64][1];
    double s = 0.0;
    for (int b = threadIdx.x; b < nblk; b += blockDim.x) s += partials[b];
    double a[1] = {s};
    block_sum<1>(a, sh);
    if (threadIdx.x == 0) out[0] = sqrt(a[0]);
}

// x += sum_j y_j V_j
template <typename V>
__global__ void __launch_bounds__(kVecBlock)
k_gm_update_x(int n, int kk, const V* __restrict__ Vb, const double* __restrict__ y, double* __restrict__ x) {
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        double acc = x[i];
        for (int j = 0; j < kk; j++) acc += y[j] * (double)Vb[(size_t)n * j + i];
        x[i] = acc;
    }
}

// r_V = D^-1 r64  (left-preconditioned residual of the FP64 residual)
template <typename V>
__global__ void __launch_bounds__(kVecBlock)
k_gm_precres(int n, const double* __restrict__ r64, const double* __restrict__ dinv, V* __restrict__ out) {
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
        out[i] = dinv ? (V)(dinv[i] * r64[i]) : (V)r64[i];
}

template <typename V>
static int run_gmres(flgpu_ctx* c, const flgpu_solver_params* sp, flgpu_solve_info* info) {
    KrylovWork* k = c->kw;
    const int n = c->rows;
    const int m = sp->restart > 0 ? sp->restart : 30;
    const bool mixed = std::is_same<V, float>::value;
    const V* val = mixed ? (const V*)k->val32.p : (const V*)k->val64.p;
    const V* dinv = mixed ? (const V*)k->dinv32.p : (const V*)k->dinv.p;
    V* Vb;
    if (mixed) {
        FLGPU_TRY(k->V32.alloc((size_t)n * (m + 1)));
        Vb = (V*)k->V32.p;
    } else {
        FLGPU_TRY(k->V.alloc((size_t)n * (m + 1)));
        Vb = (V*)k->V.p;
    }
    if (sp->pc == FLGPU_PC_NONE) {
        hipLaunchKernelGGL((k_fill<double>), dim3(k->nblk), dim3(kVecBlock), 0, c->stream, n, 1.0, k->dinv.p);
        if (mixed)
            hipLaunchKernelGGL((k_fill<float>), dim3(k->nblk), dim3(kVecBlock), 0, c->stream, n, 1.0f, k->dinv32.p);
        FLGPU_HIP(hipGetLastError());
    }
    // left SSOR: the product and the residual go through tmp, then the sweeps
    const bool ssor = sp->pc == FLGPU_PC_SOR;
    if (ssor) FLGPU_TRY(ssor_prepare(c, k));
    V* tmp = mixed ? (V*)k->rt32.p : (V*)k->rt.p;
    FLGPU_TRY(k->hcol.alloc(m + 2));
    FLGPU_TRY(k->partials.reserve((size_t)kMaxVecGrid * (m + 2)));
    const int nb = k->nblk;
    const int L = lanes_for(c);
    const int rows_per_blk = kRowBlock / L;
    int mgrid = (int)std::min<int64_t>(kMaxRowGrid, (n + rows_per_blk - 1) / rows_per_blk);
    mgrid = std::max(mgrid, 1);
    auto matvec = [&](const V* xin, V* yout) -> int {
#define FL_GMV(LL)                                                                                         \
    case LL:                                                                                               \
        hipLaunchKernelGGL((k_gm_matvec<LL, V>), dim3(mgrid), dim3(kRowBlock), 0, c->stream, n, c->rowptr.p, \
                           c->col.p, val, ssor ? nullptr : dinv, xin, ssor ? tmp : yout);                  \
        break;
        switch (L) {
            FL_GMV(4)
            FL_GMV(8)
            FL_GMV(16)
            FL_GMV(32)
            FL_GMV(64)
        }
#undef FL_GMV
        if (ssor) FLGPU_TRY(ssor_sweeps<V>(c, k, val, tmp, yout, nullptr));
        return FLGPU_OK;
    };
    std::vector<double> H((size_t)(m + 1) * m), g(m + 1), cs(m), sn(m), y(m), hc(m + 2);
    double hn = 0.0;
    int total = 0, reason = 0;
    double rnorm0 = -1.0, last = 0.0;
    Scal& h = *k->h_scal;
    std::memset(&h, 0, sizeof h);
    FLGPU_TRY(write_scal(c, k));
    FLGPU_TRY(norm_b(c, k));
    FLGPU_TRY(read_scal(c, k));
    const double bnorm = std::sqrt(h.bb);
    auto t0 = std::chrono::steady_clock::now();
    int refinements = 0;
    while (!reason) {
        // residual in FP64, preconditioned, in the working precision
        FLGPU_TRY(true_residual(c, k));
        hipLaunchKernelGGL((k_gm_precres<V>), dim3(nb), dim3(kVecBlock), 0, c->stream, n, k->r.p,
                           ssor ? nullptr : k->dinv.p, ssor ? tmp : Vb);
        if (ssor) FLGPU_TRY(ssor_sweeps<V>(c, k, val, tmp, Vb, nullptr));
        hipLaunchKernelGGL((k_gm_norm<V>), dim3(nb), dim3(kVecBlock), 0, c->stream, n, Vb, k->partials.p);
        hipLaunchKernelGGL(k_gm_sqrt_reduce, dim3(1), dim3(kVecBlock), 0, c->stream, nb, k->partials.p, k->hcol.p);
        FLGPU_HIP(hipMemcpyAsync(&hn, k->hcol.p, sizeof(double), hipMemcpyDeviceToHost, c->stream));
        FLGPU_HIP(hipStreamSynchronize(c->stream));
        refinements++;
        const double beta = hn;
        if (rnorm0 < 0) {
            rnorm0 = beta;
            last = beta;
            if (beta == 0.0 || beta <= sp->atol) {
                reason = KSP_CONVERGED_ATOL;
                break;
            }
        }
        hipLaunchKernelGGL((k_gm_scale<V>), dim3(nb), dim3(kVecBlock), 0, c->stream, n, Vb, Vb, k->hcol.p);
        std::fill(g.begin(), g.end(), 0.0);
        g[0] = beta;
        int kk;
        for (kk = 0; kk < m && !reason; kk++) {
            V* vk = Vb + (size_t)n * kk;
            V* vn = Vb + (size_t)n * (kk + 1);
            FLGPU_TRY(matvec(vk, vn));
            // classical Gram-Schmidt, twice
            for (int pass = 0; pass < 2; pass++) {
                hipLaunchKernelGGL((k_gm_dots<V>), dim3(nb), dim3(kVecBlock), 0, c->stream, n, kk, Vb, vn,
                                   k->partials.p);
                hipLaunchKernelGGL(k_gm_reduce, dim3(kk + 1), dim3(kVecBlock), 0, c->stream, kk, nb, k->partials.p,
                                   k->hcol.p + 1, 0);
                hipLaunchKernelGGL((k_gm_orth<V>), dim3(nb), dim3(kVecBlock), 0, c->stream, n, kk, Vb,
                                   k->hcol.p + 1, vn);
                FLGPU_HIP(hipMemcpyAsync(hc.data(), k->hcol.p + 1, sizeof(double) * (kk + 1),
                                         hipMemcpyDeviceToHost, c->stream));
                FLGPU_HIP(hipStreamSynchronize(c->stream));
                for (int j = 0; j <= kk; j++) H[j + (m + 1) * kk] = (pass == 0 ? 0.0 : H[j + (m + 1) * kk]) + hc[j];
            }
            hipLaunchKernelGGL((k_gm_norm<V>), dim3(nb), dim3(kVecBlock), 0, c->stream, n, vn, k->partials.p);
            hipLaunchKernelGGL(k_gm_sqrt_reduce, dim3(1), dim3(kVecBlock), 0, c->stream, nb, k->partials.p, k->hcol.p);
            hipLaunchKernelGGL((k_gm_scale<V>), dim3(nb), dim3(kVecBlock), 0, c->stream, n, vn, vn, k->hcol.p);
            FLGPU_HIP(hipMemcpyAsync(&hn, k->hcol.p, sizeof(double), hipMemcpyDeviceToHost, c->stream));
            FLGPU_HIP(hipStreamSynchronize(c->stream));
            H[kk + 1 + (m + 1) * kk] = hn;
            for (int j = 0; j < kk; j++) {
                const double a = H[j + (m + 1) * kk], b2 = H[j + 1 + (m + 1) * kk];
                H[j + (m + 1) * kk] = cs[j] * a + sn[j] * b2;
                H[j + 1 + (m + 1) * kk] = -sn[j] * a + cs[j] * b2;
            }
            const double a = H[kk + (m + 1) * kk], b2 = H[kk + 1 + (m + 1) * kk];
            const double den = std::sqrt(a * a + b2 * b2);
            cs[kk] = den == 0.0 ? 1.0 : a / den;
            sn[kk] = den == 0.0 ? 0.0 : b2 / den;
            H[kk + (m + 1) * kk] = cs[kk] * a + sn[kk] * b2;
            H[kk + 1 + (m + 1) * kk] = 0.0;
            g[kk + 1] = -sn[kk] * g[kk];
            g[kk] = cs[kk] * g[kk];
            total++;
            last = std::fabs(g[kk + 1]);
            if (last != last) reason = KSP_DIVERGED_NANORINF;
            else if (last <= sp->atol) reason = KSP_CONVERGED_ATOL;
            else if (last <= sp->rtol * rnorm0) reason = KSP_CONVERGED_RTOL;
            else if (total >= (sp->max_its > 0 ? sp->max_its : 10000)) reason = KSP_DIVERGED_ITS;
        }
        for (int i = kk - 1; i >= 0; i--) {
            double acc = g[i];
            for (int j = i + 1; j < kk; j++) acc -= H[i + (m + 1) * j] * y[j];
            y[i] = acc / H[i + (m + 1) * i];
        }
        FLGPU_HIP(hipMemcpyAsync(k->hcol.p, y.data(), sizeof(double) * kk, hipMemcpyHostToDevice, c->stream));
        hipLaunchKernelGGL((k_gm_update_x<V>), dim3(nb), dim3(kVecBlock), 0, c->stream, n, kk, Vb, k->hcol.p,
                           k->x.p);
        FLGPU_HIP(hipGetLastError());
        if (mixed && reason > 0) {
            // the FP32 estimate converged: test the FP64 residual
            FLGPU_TRY(true_residual(c, k));
            FLGPU_TRY(read_scal(c, k));
            if (std::sqrt(h.true_rr) > std::max(sp->rtol * bnorm, sp->atol) &&
                total < (sp->max_its > 0 ? sp->max_its : 10000))
                reason = 0;
        }
    }
    FLGPU_TRY(true_residual(c, k));
    FLGPU_TRY(read_scal(c, k));
    auto t1 = std::chrono::steady_clock::now();
    info->seconds = std::chrono::duration<double>(t1 - t0).count();
    info->its = total;
    info->reason = reason;
    info->true_relres = bnorm > 0 ? std::sqrt(h.true_rr) / bnorm : std::sqrt(h.true_rr);
    info->precond_relres = rnorm0 > 0 ? last / rnorm0 : last;
    info->refinements = mixed ? refinements : 0;
    info->fallback = 0;
    return FLGPU_OK;
}

int solve(flgpu_ctx* c, const flgpu_solver_params* sp, double* x_host, flgpu_solve_info* info_out) {
    flgpu_solve_info info{};
    const bool mixed = sp->precision == FLGPU_MIXED;
    if (sp->precision != FLGPU_FP64 && sp->precision != FLGPU_MIXED)
        return set_error(FLGPU_EINVAL, "flgpu_solve: precision %d", sp->precision);
    if (sp->ksp != FLGPU_KSP_CG && sp->ksp != FLGPU_KSP_GMRES && sp->ksp != FLGPU_KSP_FCG &&
        sp->ksp != FLGPU_KSP_PIPECG)
        return set_error(FLGPU_EINVAL, "flgpu_solve: ksp %d", sp->ksp);
    if (sp->pc != FLGPU_PC_NONE && sp->pc != FLGPU_PC_JACOBI && sp->pc != FLGPU_PC_LEARNED && sp->pc != FLGPU_PC_SOR &&
        sp->pc != FLGPU_PC_FSAI)
        return set_error(FLGPU_EINVAL, "flgpu_solve: pc %d", sp->pc);
    const bool factor = sp->pc == FLGPU_PC_LEARNED || sp->pc == FLGPU_PC_FSAI;
    if (factor && sp->ksp == FLGPU_KSP_GMRES)
        return set_error(FLGPU_EUNSUPPORTED, "flgpu_solve: the learned and fsai preconditioners run with cg/fcg/pipecg");
    FLGPU_TRY(ensure_work(c, mixed));
    KrylovWork* k = c->kw;
    const bool zero_x = sp->start_from_zero || !x_host;
    if (x_host && !sp->start_from_zero)
        FLGPU_HIP(hipMemcpyAsync(k->x.p, x_host, sizeof(double) * c->rows, hipMemcpyHostToDevice, c->stream));
    FLGPU_TRY(build_reduced(c, k, mixed, zero_x));
    k->prepared = true;
    k->prepared_mixed = mixed;
    if (factor) FLGPU_TRY(learned_prepare(c, mixed, sp->pc == FLGPU_PC_FSAI));
    k->dx_on = false;  // set by the mixed CG loops for their FP32 increment
    if (sp->ksp == FLGPU_KSP_GMRES) {
        FLGPU_TRY(mixed ? run_gmres<float>(c, sp, &info) : run_gmres<double>(c, sp, &info));
    } else if (sp->ksp == FLGPU_KSP_PIPECG) {
        FLGPU_TRY(mixed ? run_cg1<float>(c, sp, &info, factor) : run_cg1<double>(c, sp, &info, factor));
    } else {
        FLGPU_TRY(mixed ? run_cg<float>(c, sp, &info, factor) : run_cg<double>(c, sp, &info, factor));
    }
    k->dx_on = false;  // the loops end with a true residual, which adds the increment to x
    hipLaunchKernelGGL(k_finish_x, dim3(grid_for(c->rows, kVecBlock)), dim3(kVecBlock), 0, c->stream, c->rows,
                       k->x.p, c->rhs.p, c->inactive.p, c->have_inactive ? 1 : 0);
    FLGPU_HIP(hipGetLastError());
    if (x_host)
        FLGPU_HIP(hipMemcpyAsync(x_host, k->x.p, sizeof(double) * c->rows, hipMemcpyDeviceToHost, c->stream));
    FLGPU_HIP(hipStreamSynchronize(c->stream));
    k->have_x = true;
    if (info_out) *info_out = info;
    return FLGPU_OK;
}

void krylov_invalidate(flgpu_ctx* c) {
    if (c->kw) c->kw->prepared = c->kw->prepared_mixed = false;
}

void krylov_free(flgpu_ctx* c) {
    delete c->kw;
    c->kw = nullptr;
}

// ---------------------------------------------------------------------
// The solve without allocation
// ---------------------------------------------------------------------
//
// Every buffer of the FP64 (F)PCG -- reduced matrix values, the vectors,
// per-block partial sums, the two solver-scalar slots and the row blocks of
// the streamed product -- is carved out of one file-scope device array.
// The host code never names the array itself (in host code its name is a
// host-side stand-in): it takes the array's device address once per GPU
// from hipGetSymbolAddress (noalloc_setup, called by flgpu_init with the
// device current) and passes that address on.  Every context on a GPU
// shares that GPU's array, so solves that overlap on one GPU race.

constexpr size_t kNoAllocBytes = (size_t)64 << 20;
constexpr int kNoAllocDevices = 64;
static __device__ __align__(256) unsigned char scratch[kNoAllocBytes];
static std::mutex g_scratch_mu;
static unsigned char* g_scratch_ptr[kNoAllocDevices] = {};

int noalloc_setup(int device) {
    if (device < 0 || device >= kNoAllocDevices)
        return set_error(FLGPU_EUNSUPPORTED, "flgpu_init: device %d beyond the %d the library tracks", device,
                         kNoAllocDevices);
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (g_scratch_ptr[device]) return FLGPU_OK;
    void* p = nullptr;
    FLGPU_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    if (!p) return set_error(FLGPU_EHIP, "flgpu_init: no device address for the solver work array");
    g_scratch_ptr[device] = static_cast<unsigned char*>(p);
    return FLGPU_OK;
}

static unsigned char* noalloc_ptr(int device) {
    if (device < 0 || device >= kNoAllocDevices) return nullptr;
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    return g_scratch_ptr[device];
}

// Row blocks of a fixed number of rows, rows_per * maxlen <= kStreamNnz
// (or one row per block): computed on the device, no host table.
__global__ void __launch_bounds__(256)
k_fixed_row_blocks(int n, int rows_per, int nrb, const int32_t* __restrict__ rowptr, int32_t* __restrict__ rb,
                   int32_t* __restrict__ rbk) {
    const int b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b > nrb) return;
    const int r = (int)std::min<int64_t>((int64_t)b * rows_per, n);
    rb[b] = r;
    rbk[b] = rowptr[r];
}

// Carves the work of an n-row, nnz-entry FP64 (F)PCG out of the array;
// false when it does not fit (*need gets the bytes it would take).
static bool noalloc_carve(unsigned char* base, int n, int64_t nnz, int nrb, KrylovWork& k, size_t* need) {
    size_t off = 0;
    auto take = [&](size_t bytes) {
        const size_t at = off;
        off = (off + bytes + 255) & ~(size_t)255;
        return at;
    };
    const size_t o_val = take(sizeof(double) * (size_t)nnz);
    size_t o_vec[7];
    for (size_t& o : o_vec) o = take(sizeof(double) * (size_t)n);
    const size_t o_part = take(sizeof(double) * (size_t)kPartStride * PART_COMPONENTS);
    const size_t o_scal = take(sizeof(Scal) * 2);
    const size_t o_rb = take(sizeof(int32_t) * ((size_t)nrb + 1));
    const size_t o_rbk = take(sizeof(int32_t) * ((size_t)nrb + 1));
    *need = off;
    if (off > kNoAllocBytes) return false;
    k.val64.borrow(reinterpret_cast<double*>(base + o_val), (size_t)nnz);
    DevBuf<double>* vecs[7] = {&k.b, &k.x, &k.r, &k.z, &k.p, &k.w, &k.dinv};
    for (int q = 0; q < 7; q++) vecs[q]->borrow(reinterpret_cast<double*>(base + o_vec[q]), (size_t)n);
    k.partials.borrow(reinterpret_cast<double*>(base + o_part), (size_t)kPartStride * PART_COMPONENTS);
    k.scal.borrow(reinterpret_cast<Scal*>(base + o_scal), 2);
    k.rb.borrow(reinterpret_cast<int32_t*>(base + o_rb), (size_t)nrb + 1);
    k.rbk.borrow(reinterpret_cast<int32_t*>(base + o_rbk), (size_t)nrb + 1);
    return true;
}

int solve_noalloc(flgpu_ctx* c, const flgpu_solver_params* sp, double* x_host, flgpu_solve_info* info_out) {
    if (!x_host) return set_error(FLGPU_EINVAL, "flgpu_solve_noalloc: x is required (host, rows values)");
    if (sp->precision != FLGPU_FP64)
        return set_error(FLGPU_EUNSUPPORTED, "flgpu_solve_noalloc: FP64 only (precision %d)", sp->precision);
    if (sp->ksp != FLGPU_KSP_CG && sp->ksp != FLGPU_KSP_FCG)
        return set_error(FLGPU_EUNSUPPORTED, "flgpu_solve_noalloc: cg or fcg only (ksp %d)", sp->ksp);
    if (sp->pc != FLGPU_PC_NONE && sp->pc != FLGPU_PC_JACOBI)
        return set_error(FLGPU_EUNSUPPORTED, "flgpu_solve_noalloc: pc none or jacobi only (pc %d)", sp->pc);
    unsigned char* base = noalloc_ptr(c->device);
    if (!base) return set_error(FLGPU_EINVAL, "flgpu_solve_noalloc: device %d was not set up by flgpu_init", c->device);
    const int n = c->rows;
    int maxlen = 0;
    for (int i = 0; i < n; i++) maxlen = std::max(maxlen, c->h_rowptr[i + 1] - c->h_rowptr[i]);
    const int rows_per = std::max(1, std::min(kRowBlock, kStreamNnz / std::max(1, maxlen)));
    const int nrb = (int)(((int64_t)n + rows_per - 1) / rows_per);
    KrylovWork k;  // every buffer a view of the array; nothing allocated or freed
    size_t need = 0;
    if (!noalloc_carve(base, n, c->nnz, nrb, k, &need))
        return set_error(FLGPU_EUNSUPPORTED,
                         "flgpu_solve_noalloc: a system of %d rows and %lld entries needs %zu bytes of solver work, "
                         "more than the %zu bytes of its fixed array",
                         n, (long long)c->nnz, need, kNoAllocBytes);
    Scal h_scal;  // pageable: read_scal / write_scal synchronise
    k.n = n;
    k.nnz = c->nnz;
    k.use16 = false;
    k.nrb = nrb;
    k.h_scal = &h_scal;
    k.h_scal_owned = false;
    k.nblk = std::max(1, (int)std::min<int64_t>(kMaxVecGrid, ((int64_t)n + kVecBlock - 1) / kVecBlock));
    flgpu_solve_info info{};
    KrylovWork* saved = c->kw;
    c->kw = &k;
    const int rc = [&]() -> int {
        hipLaunchKernelGGL(k_fixed_row_blocks, dim3(grid_for((int64_t)nrb + 1, 256)), dim3(256), 0, c->stream, n,
                           rows_per, nrb, c->rowptr.p, k.rb.p, k.rbk.p);
        FLGPU_HIP(hipGetLastError());
        if (!sp->start_from_zero)
            FLGPU_HIP(hipMemcpyAsync(k.x.p, x_host, sizeof(double) * n, hipMemcpyHostToDevice, c->stream));
        FLGPU_TRY(build_reduced(c, &k, false, sp->start_from_zero != 0));
        FLGPU_TRY(run_cg<double>(c, sp, &info, false));
        hipLaunchKernelGGL(k_finish_x, dim3(grid_for(n, kVecBlock)), dim3(kVecBlock), 0, c->stream, n, k.x.p,
                           c->rhs.p, c->inactive.p, c->have_inactive ? 1 : 0);
        FLGPU_HIP(hipGetLastError());
        FLGPU_HIP(hipMemcpyAsync(x_host, k.x.p, sizeof(double) * n, hipMemcpyDeviceToHost, c->stream));
        FLGPU_HIP(hipStreamSynchronize(c->stream));
        return FLGPU_OK;
    }();
    c->kw = saved;
    if (rc != FLGPU_OK) {
        (void)hipStreamSynchronize(c->stream);  // no kernel still reads the views
        return rc;
    }
    if (info_out) *info_out = info;
    return FLGPU_OK;
}

int solution_get(flgpu_ctx* c, double* x) {
    if (!c->kw || !c->kw->have_x) return set_error(FLGPU_EINVAL, "flgpu_solution_get: no solution");
    FLGPU_HIP(hipMemcpyAsync(x, c->kw->x.p, sizeof(double) * c->rows, hipMemcpyDeviceToHost, c->stream));
    FLGPU_HIP(hipStreamSynchronize(c->stream));
    return FLGPU_OK;
}

int csr_mult(flgpu_ctx* c, const double* x, double* y) {
    FLGPU_TRY(ensure_work(c, false));
    KrylovWork* k = c->kw;
    DevBuf<double> dx, dy;
    FLGPU_TRY(dx.upload(x, c->cols, c->stream));
    FLGPU_TRY(dy.alloc(c->rows));
    FLGPU_TRY((spmv<double, double, double>(c, k, c->A.p, dx.p, dy.p, nullptr, 0, nullptr)));
    FLGPU_HIP(hipMemcpyAsync(y, dy.p, sizeof(double) * c->rows, hipMemcpyDeviceToHost, c->stream));
    FLGPU_HIP(hipStreamSynchronize(c->stream));
    return FLGPU_OK;
}

// rt = D r (the residual in the factor's scaled space)
__global__ void __launch_bounds__(kVecBlock)
k_scale_res(int n, const double* __restrict__ r, const float* __restrict__ dsq, double* __restrict__ rt) {
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
        rt[i] = r[i] * (double)dsq[i];
}

int pc_apply_host(flgpu_ctx* c, const flgpu_solver_params* sp, const double* r, double* z) {
    if (sp->pc != FLGPU_PC_NONE && sp->pc != FLGPU_PC_JACOBI && sp->pc != FLGPU_PC_SOR &&
        sp->pc != FLGPU_PC_LEARNED && sp->pc != FLGPU_PC_FSAI)
        return set_error(FLGPU_EUNSUPPORTED, "flgpu_pc_apply: pc %d", sp->pc);
    FLGPU_TRY(ensure_work(c, false));
    KrylovWork* k = c->kw;
    FLGPU_TRY(build_reduced(c, k, false, true));
    k->prepared = true;
    k->prepared_mixed = false;
    const int n = c->rows;
    FLGPU_HIP(hipMemcpyAsync(k->r.p, r, sizeof(double) * n, hipMemcpyHostToDevice, c->stream));
    if (sp->pc == FLGPU_PC_SOR) {
        FLGPU_TRY(ssor_prepare(c, k));
        FLGPU_TRY(ssor_sweeps<double>(c, k, k->val64.p, k->r.p, k->z.p, nullptr));
    } else if (sp->pc == FLGPU_PC_LEARNED || sp->pc == FLGPU_PC_FSAI) {
        FLGPU_TRY(learned_prepare(c, false, sp->pc == FLGPU_PC_FSAI));
        hipLaunchKernelGGL(k_scale_res, dim3(k->nblk), dim3(kVecBlock), 0, c->stream, n, k->r.p, learned_scale(c),
                           k->rt.p);
        FLGPU_HIP(hipGetLastError());
        int np = 0;
        FLGPU_TRY(learned_precond_apply<double>(c, k->r.p, k->rt.p, k->z.p, k->y.p, k->r.p, k->part(PART_PC), &np,
                                                nullptr));
    } else {
        hipLaunchKernelGGL((k_gm_precres<double>), dim3(k->nblk), dim3(kVecBlock), 0, c->stream, n, k->r.p,
                           sp->pc == FLGPU_PC_JACOBI ? k->dinv.p : nullptr, k->z.p);
        FLGPU_HIP(hipGetLastError());
    }
    FLGPU_HIP(hipMemcpyAsync(z, k->z.p, sizeof(double) * n, hipMemcpyDeviceToHost, c->stream));
    FLGPU_HIP(hipStreamSynchronize(c->stream));
    return FLGPU_OK;
}

const double* reduced_values(flgpu_ctx* c) { return c->kw ? c->kw->val64.p : nullptr; }

int reduced_prepare(flgpu_ctx* c) {
    if (c->kw && c->kw->prepared && c->kw->prepared_mixed) return FLGPU_OK;
    return solver_prepare(c, FLGPU_MIXED);
}

int solver_prepare(flgpu_ctx* c, int precision) {
    const bool mixed = precision == FLGPU_MIXED;
    FLGPU_TRY(ensure_work(c, mixed));
    FLGPU_TRY(build_reduced(c, c->kw, mixed, true));
    FLGPU_HIP(hipMemsetAsync(c->kw->p.p, 0, sizeof(double) * c->rows, c->stream));
    if (mixed) FLGPU_HIP(hipMemsetAsync(c->kw->p32.p, 0, sizeof(float) * c->rows, c->stream));
    FLGPU_HIP(hipStreamSynchronize(c->stream));
    c->kw->prepared_mixed = mixed;
    c->kw->prepared = true;
    return FLGPU_OK;
}

// Streams through a buffer (read and write) to evict the SpMV's operands
// from the L2 caches and the Infinity Cache before a cold timing.
__global__ void __launch_bounds__(256) k_flush(uint4* __restrict__ buf, int64_t n) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        uint4 v = buf[i];
        v.x += 1u;
        buf[i] = v;
    }
}

// reps products, each after a flush of flush_mib MiB and timed alone
// between its own events; *ms = their mean.
int time_spmv_cold(flgpu_ctx* c, int precision, int reps, int flush_mib, double* ms) {
    const bool mixed = precision == FLGPU_MIXED;
    if (!c->kw || !c->kw->prepared || (mixed && !c->kw->prepared_mixed))
        FLGPU_TRY(solver_prepare(c, precision));
    KrylovWork* k = c->kw;
    const int64_t n16 = (int64_t)flush_mib * (1 << 20) / 16;
    DevBuf<uint4> fb;
    FLGPU_TRY(fb.alloc((size_t)n16));
    FLGPU_HIP(hipMemsetAsync(fb.p, 0, sizeof(uint4) * n16, c->stream));
    std::vector<hipEvent_t> ev(2 * (size_t)reps, nullptr);
    int rc = FLGPU_OK;
    for (auto& e : ev)
        if (hipEventCreate(&e) != hipSuccess) rc = set_error(FLGPU_EHIP, "flgpu_time_spmv_cold: hipEventCreate");
    for (int r = 0; r < reps && rc == FLGPU_OK; r++) {
        hipLaunchKernelGGL(k_flush, dim3(2048), dim3(256), 0, c->stream, fb.p, n16);
        if (hipEventRecord(ev[2 * r], c->stream) != hipSuccess) rc = set_error(FLGPU_EHIP, "hipEventRecord");
        if (rc == FLGPU_OK)
            rc = mixed ? spmv<float, float, float>(c, k, k->val32.p, k->p32.p, k->w32.p, nullptr, 0, nullptr)
                       : spmv<double, double, double>(c, k, k->val64.p, k->p.p, k->w.p, nullptr, 0, nullptr);
        if (rc == FLGPU_OK && hipEventRecord(ev[2 * r + 1], c->stream) != hipSuccess)
            rc = set_error(FLGPU_EHIP, "hipEventRecord");
    }
    if (rc == FLGPU_OK && hipStreamSynchronize(c->stream) != hipSuccess)
        rc = set_error(FLGPU_EHIP, "flgpu_time_spmv_cold: hipStreamSynchronize");
    double sum = 0.0;
    for (int r = 0; r < reps && rc == FLGPU_OK; r++) {
        float t = 0.f;
        if (hipEventElapsedTime(&t, ev[2 * r], ev[2 * r + 1]) != hipSuccess)
            rc = set_error(FLGPU_EHIP, "hipEventElapsedTime");
        sum += t;
    }
    for (auto& e : ev)
        if (e) (void)hipEventDestroy(e);
    if (rc == FLGPU_OK) *ms = sum / reps;
    return rc;
}

// Bare solver-matrix SpMV w = A p, timed with HIP events on the stream.
int time_spmv(flgpu_ctx* c, int precision, int reps, double* ms) {
    const bool mixed = precision == FLGPU_MIXED;
    if (!c->kw || !c->kw->prepared || (mixed && !c->kw->prepared_mixed))
        FLGPU_TRY(solver_prepare(c, precision));
    KrylovWork* k = c->kw;
    hipEvent_t e0, e1;
    FLGPU_HIP(hipEventCreate(&e0));
    FLGPU_HIP(hipEventCreate(&e1));
    FLGPU_HIP(hipEventRecord(e0, c->stream));
    for (int r = 0; r < reps; r++) {
        if (mixed)
            FLGPU_TRY((spmv<float, float, float>(c, k, k->val32.p, k->p32.p, k->w32.p, nullptr, 0, nullptr)));
        else
            FLGPU_TRY((spmv<double, double, double>(c, k, k->val64.p, k->p.p, k->w.p, nullptr, 0, nullptr)));
    }
    FLGPU_HIP(hipEventRecord(e1, c->stream));
    FLGPU_HIP(hipEventSynchronize(e1));
    float t = 0.f;
    FLGPU_HIP(hipEventElapsedTime(&t, e0, e1));
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    *ms = (double)t / reps;
    return FLGPU_OK;
}

}  // namespace flgpu
