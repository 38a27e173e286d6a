// Element assembly of the CG advection-diffusion equation on gfx950.
//
// Equation (assemble/Advection_Diffusion_CG.F90:702-1202, "delta" form):
//   matrix = M + dt*theta*(A_adv + A_abs + K),  rhs = -(A_adv + A_abs + K) T + S
// Per element the row i of the local matrix is produced by row_contrib();
// every kernel in this file calls it with the same arguments, so all
// schemes give the same bits, and each CSR entry receives its element
// contributions colour by colour, ascending element id inside a colour
// (the order of the reference's OpenMP colour loop, :574-598).
//
// Arithmetic: explicit fused multiply-adds, no other contraction
// (-ffp-contract=off for this file).  When the physical gradients are
// constant over the element (linear T on linear simplices, the
// .not.m_nonlinear branch of femtools/Transform_elements.F90:777-780) the
// quadrature sums are contracted into the tables what, mhat and that
// (host-side, exact algebra); quadratic T keeps the per-point loops.
#include <hip/hip_runtime.h>

#include <cmath>
#include <vector>

#include "flgpu_internal.h"
#include "element_kernels.h"

namespace flgpu {

template <typename R>
__device__ void stage_tables(const ElemTables* __restrict__ t, LTab<R>& s) {
    const int tid = threadIdx.x, nt = blockDim.x;
    for (int k = tid; k < kMaxLoc * kMaxGi; k += nt) s.n[k] = (R)t->n[k];
    for (int k = tid; k < kMaxLoc * kMaxGi * kMaxDim; k += nt) s.dn[k] = (R)t->dn[k];
    for (int k = tid; k < kMaxGi; k += nt) s.w[k] = (R)t->w[k];
    for (int k = tid; k < kMaxLoc; k += nt) s.what[k] = (R)t->what[k];
    for (int k = tid; k < kMaxLoc * kMaxLoc; k += nt) s.mhat[k] = (R)t->mhat[k];
    for (int k = tid; k < 64; k += nt) s.that[k] = (R)t->that[k];
    for (int k = tid; k < 4 * kMaxDim; k += nt) s.xdn[k] = (R)t->xdn[k];
    if (tid == 0) {
        s.loc = t->loc;
        s.ngi = t->ngi;
        s.lx = t->loc_x;
    }
}

template <int DIM, int LOC, typename R>
__device__ __forceinline__ void gather_elem(int ele, const int32_t* __restrict__ conn,
                                            const int32_t* __restrict__ conn_x,
                                            const double* __restrict__ X4,
                                            const double* __restrict__ T, Elem<DIM, LOC, R>& e) {
#pragma unroll
    for (int k = 0; k < LOC; k++) e.nodes[k] = conn[(int64_t)ele * LOC + k];
#pragma unroll
    for (int nn = 0; nn < DIM + 1; nn++) {
        const int xn = conn_x[(int64_t)ele * (DIM + 1) + nn];
#pragma unroll
        for (int a = 0; a < DIM; a++) e.X[a][nn] = (R)X4[(int64_t)xn * 4 + a];
    }
#pragma unroll
    for (int k = 0; k < LOC; k++) e.T[k] = (R)T[e.nodes[k]];
}

// Position of column cj in the sorted row (bisection of csr_sparsity_pos,
// femtools/Sparse_Tools.F90:2411-2517); -1 when absent.
__device__ __forceinline__ int64_t row_pos(const int32_t* __restrict__ rowptr,
                                           const int32_t* __restrict__ col, int r, int cj) {
    int lo = rowptr[r], hi = rowptr[r + 1] - 1;
    while (lo <= hi) {
        int mid = (lo + hi) >> 1;
        int v = col[mid];
        if (v == cj) return mid;
        if (v < cj) lo = mid + 1;
        else hi = mid - 1;
    }
    return -1;
}

// One launch per colour; thread = (element of the colour, local row i).
// Elements of one colour share no node, so the read-modify-writes of a
// launch never meet.
template <int DIM, int LOC, typename R>
__global__ void __launch_bounds__(256)
k_asm_colour(const ElemTables* __restrict__ tab, const int32_t* __restrict__ elems, int count,
             const int32_t* __restrict__ conn, const int32_t* __restrict__ conn_x,
             const double* __restrict__ X4, AsmFields fd, AsmFlags fl,
             const int32_t* __restrict__ rowptr, const int32_t* __restrict__ col,
             double* __restrict__ A, double* __restrict__ rhs, int* __restrict__ err) {
    __shared__ LTab<R> tb;
    stage_tables<R>(tab, tb);
    __syncthreads();
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (int64_t)count * LOC) return;
    const int q = (int)(t / LOC), i = (int)(t % LOC);
    const int ele = elems[q];
    Elem<DIM, LOC, R> e;
    gather_elem<DIM, LOC, R>(ele, conn, conn_x, X4, fd.T, e);
    R vals[LOC], r;
    row_contrib<DIM, LOC, R>(e, tb, fl, fd, i, vals, r);
    const int row = e.nodes[i];
#pragma unroll
    for (int j = 0; j < LOC; j++) {
        const double v = (double)vals[j];
        if (v == 0.0) continue;  // csr_addto skips zeros (Sparse_Tools.F90:2640)
        const int64_t pos = row_pos(rowptr, col, row, e.nodes[j]);
        if (pos < 0) {
            atomicOr(err, 1);
            continue;
        }
        A[pos] = A[pos] + v;
    }
    rhs[row] = rhs[row] + (double)r;
}

template <int DIM, int LOC, typename R>
static int launch_colour(flgpu_ctx* c, const AsmFields& fd, const AsmFlags& fl, int* d_err) {
    for (int k = 0; k < c->ncol; k++) {
        const int off = c->h_col_ptr[k], cnt = c->h_col_ptr[k + 1] - off;
        if (cnt == 0) continue;
        hipLaunchKernelGGL((k_asm_colour<DIM, LOC, R>), dim3(grid_for((int64_t)cnt * LOC, 256)), dim3(256),
                           0, c->stream, c->d_tab.p, c->col_elems.p + off, cnt, c->conn.p, c->conn_x.p,
                           c->X4.p, fd, fl, c->rowptr.p, c->col.p, c->A.p, c->rhs.p, d_err);
    }
    FLGPU_HIP(hipGetLastError());
    return FLGPU_OK;
}

static int assemble_launch(flgpu_ctx* c, const flgpu_advdif_params* p, int* d_err) {
    if (c->nele <= 0 || c->loc <= 0 || c->nnz <= 0 || c->ncol <= 0)
        return set_error(FLGPU_EINVAL, "flgpu assemble: bind mesh, field mesh, sparsity and colouring first");
    if (c->rows != c->nnodes)
        return set_error(FLGPU_EINVAL, "flgpu assemble: sparsity rows %d != field nodes %d", c->rows, c->nnodes);
    if (!c->fT.set) return set_error(FLGPU_EINVAL, "flgpu assemble: field T not set");
    if (p->have_advection && !c->fU.set) return set_error(FLGPU_EINVAL, "flgpu assemble: velocity not set");
    if (p->have_diffusivity && !c->fK.set) return set_error(FLGPU_EINVAL, "flgpu assemble: diffusivity not set");
    if (p->have_source && !c->fS.set) return set_error(FLGPU_EINVAL, "flgpu assemble: source not set");
    if (p->have_absorption && !c->fA.set) return set_error(FLGPU_EINVAL, "flgpu assemble: absorption not set");
    if (p->precision != FLGPU_FP64 && p->precision != FLGPU_FP32)
        return set_error(FLGPU_EINVAL, "flgpu assemble: precision %d", p->precision);

    const double eps = 2.220446049250313e-16;  // epsilon(0.0) for real*8
    AsmFlags fl;
    fl.dt_theta = p->dt * p->theta;
    fl.beta = p->beta;
    fl.have_mass = p->have_mass != 0;
    fl.lump = p->lump_mass != 0;
    fl.adv = p->have_advection != 0;
    fl.ibp = p->integrate_by_parts != 0;
    fl.diff = p->have_diffusivity != 0;
    fl.iso = p->isotropic_diffusivity != 0;
    fl.src = p->have_source != 0;
    fl.absorp = p->have_absorption != 0;
    fl.need_div = fl.adv && (fl.ibp ? (std::fabs(1.0 - p->beta) > eps) : (std::fabs(p->beta) > eps));
    fl.dtt_on = std::fabs(fl.dt_theta) > eps;
    AsmFields fd;
    fd.T = c->fT.v.p;
    fd.U = c->fU.v.p; fd.u_c = c->fU.constant;
    fd.K = c->fK.v.p; fd.k_c = c->fK.constant;
    fd.S = c->fS.v.p; fd.s_c = c->fS.constant;
    fd.B = c->fA.v.p; fd.b_c = c->fA.constant;

    FLGPU_HIP(hipMemsetAsync(c->A.p, 0, sizeof(double) * c->nnz, c->stream));
    FLGPU_HIP(hipMemsetAsync(c->rhs.p, 0, sizeof(double) * c->rows, c->stream));
    FLGPU_HIP(hipMemsetAsync(c->inactive.p, 0, sizeof(int32_t) * c->rows, c->stream));
    c->have_inactive = false;

    const int dim = c->dim, loc = c->loc;
    const bool fp32 = p->precision == FLGPU_FP32;
    int rc = FLGPU_OK;
    int scheme = p->scheme;
    if (scheme == FLGPU_ASM_BLOCKED && !(c->linear && !fp32))
        return set_error(FLGPU_EUNSUPPORTED, "flgpu assemble: the blocked scheme covers linear T in FP64");
#define FL_DISPATCH(D, L)                                                                    \
    if (dim == D && loc == L) {                                                              \
        rc = fp32 ? launch_colour<D, L, float>(c, fd, fl, d_err)                           \
                  : launch_colour<D, L, double>(c, fd, fl, d_err);                         \
    } else
    FL_DISPATCH(2, 3)
    FL_DISPATCH(2, 6)
    FL_DISPATCH(3, 4)
    FL_DISPATCH(3, 10) { return set_error(FLGPU_EUNSUPPORTED, "flgpu assemble: dim %d loc %d", dim, loc); }
#undef FL_DISPATCH
    return rc;
}

static int assemble_check(flgpu_ctx* c, int* d_err) {
    int h_err = 0;
    FLGPU_HIP(hipMemcpyAsync(&h_err, d_err, sizeof(int), hipMemcpyDeviceToHost, c->stream));
    FLGPU_HIP(hipStreamSynchronize(c->stream));
    if (h_err) {
        c->assembled = false;
        return set_error(FLGPU_EPATTERN, "Attempting to set value in matrix outside sparsity pattern.");
    }
    c->assembled = true;
    return FLGPU_OK;
}

int assemble(flgpu_ctx* c, const flgpu_advdif_params* p) {
    DevBuf<int> d_err;
    FLGPU_TRY(d_err.alloc(1));
    FLGPU_HIP(hipMemsetAsync(d_err.p, 0, sizeof(int), c->stream));
    FLGPU_TRY(assemble_launch(c, p, d_err.p));
    return assemble_check(c, d_err.p);
}

int time_assemble(flgpu_ctx* c, const flgpu_advdif_params* p, int reps, double* ms) {
    if (reps < 1) return set_error(FLGPU_EINVAL, "flgpu_time_assemble: reps < 1");
    DevBuf<int> d_err;
    FLGPU_TRY(d_err.alloc(1));
    FLGPU_HIP(hipMemsetAsync(d_err.p, 0, sizeof(int), c->stream));
    hipEvent_t e0, e1;
    FLGPU_HIP(hipEventCreate(&e0));
    FLGPU_HIP(hipEventCreate(&e1));
    FLGPU_HIP(hipEventRecord(e0, c->stream));
    int rc = FLGPU_OK;
    for (int r = 0; r < reps && rc == FLGPU_OK; r++) rc = assemble_launch(c, p, d_err.p);
    FLGPU_HIP(hipEventRecord(e1, c->stream));
    FLGPU_HIP(hipEventSynchronize(e1));
    float t = 0.f;
    FLGPU_HIP(hipEventElapsedTime(&t, e0, e1));
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    if (rc != FLGPU_OK) return rc;
    *ms = (double)t / reps;
    return assemble_check(c, d_err.p);
}

}  // namespace flgpu
