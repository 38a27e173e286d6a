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

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <vector>

#include "flgpu_internal.h"
#include "element_kernels.h"

namespace flgpu {

// Workgroup of the node-staged slab kernel and the most row entries and
// slots each of its threads stages (the slab plan keeps slabs within them).
constexpr int kStagedThreads = 256;
constexpr int kStagedColsPerThread = 4;
constexpr int kStagedSlotsPerThread = 3;
static_assert(kStagedSlotsPerThread == 3, "phase A picks a thread's slot words by index 0..2");

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

// The linear-element tables into LDS (dn at the first quadrature point,
// re-indexed for ngi = 1).
template <typename R>
__device__ void stage_tables_lin(const ElemTables* __restrict__ t, LTabLin<R>& s) {
    const int tid = threadIdx.x, nt = blockDim.x;
    const int loc = t->loc, ngi = t->ngi;
    for (int k = tid; k < loc * kMaxDim && k < 4 * kMaxDim; k += nt) {
        const int nn = k % loc, d = k / loc;
        s.dn[nn + loc * d] = (R)t->dn[nn + loc * (0 + ngi * d)];
    }
    for (int k = tid; k < 4; k += nt) s.what[k] = (R)t->what[k];
    for (int k = tid; k < 16; k += nt) s.mhat[k] = (R)t->mhat[k];
    for (int k = tid; k < 64; k += nt) s.that[k] = (R)t->that[k];
    for (int k = tid; k < 4 * kMaxDim; k += nt) s.xdn[k] = (R)t->xdn[k];
    if (tid == 0) {
        s.loc = loc;
        s.ngi = 1;
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
    // vertices relative to the last one, differenced in FP64 (exact
    // geometry for either precision: an FP32 difference of two rounded
    // coordinates loses the digits a graded mesh's small elements need)
    const int xl = conn_x[(int64_t)ele * (DIM + 1) + DIM];
    double xo[DIM];
#pragma unroll
    for (int a = 0; a < DIM; a++) {
        xo[a] = X4[(int64_t)xl * 4 + a];
        e.X[a][DIM] = (R)0;
    }
#pragma unroll
    for (int nn = 0; nn < DIM; nn++) {
        const int xn = conn_x[(int64_t)ele * (DIM + 1) + nn];
#pragma unroll
        for (int a = 0; a < DIM; a++) e.X[a][nn] = (R)(X4[(int64_t)xn * 4 + a] - xo[a]);
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

// The element tables in LDS from their flat image (table_image): every
// load of a thread issued before its first store.
template <typename R, int NT>
__device__ __forceinline__ void stage_table_image(const uint32_t* __restrict__ ltab, LTab<R>& tb) {
    constexpr int TWORDS = (int)(sizeof(LTab<R>) / sizeof(uint32_t));
    const int t = threadIdx.x;
    uint32_t tw[(TWORDS + NT - 1) / NT];
#pragma unroll
    for (int u = 0; u < (TWORDS + NT - 1) / NT; u++) tw[u] = t + NT * u < TWORDS ? ltab[t + NT * u] : 0u;
#pragma unroll
    for (int u = 0; u < (TWORDS + NT - 1) / NT; u++)
        if (t + NT * u < TWORDS) reinterpret_cast<uint32_t*>(&tb)[t + NT * u] = tw[u];
}

// The CSR position of every element entry (i, j), by the bisection of
// csr_sparsity_pos (Sparse_Tools.F90:2411-2517), once per mesh and
// sparsity; -1 when the pattern lacks it.
__global__ void __launch_bounds__(256)
k_entry_pos(int64_t n, int loc, const int32_t* __restrict__ conn, const int32_t* __restrict__ rowptr,
            const int32_t* __restrict__ col, int32_t* __restrict__ epos) {
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n) return;
    const int64_t ele = t / ((int64_t)loc * loc);
    const int rr = (int)(t - ele * loc * loc), i = rr / loc, j = rr - i * loc;
    epos[t] = (int32_t)row_pos(rowptr, col, conn[ele * loc + i], conn[ele * loc + j]);
}

static int ensure_entry_positions(flgpu_ctx* c) {
    BlockPlan& P = c->plan;
    if (P.epos_built) return FLGPU_OK;
    const int64_t n = (int64_t)c->nele * c->loc * c->loc;
    if (n > INT32_MAX) return set_error(FLGPU_EUNSUPPORTED, "colour assembly: too many element entries");
    FLGPU_TRY(P.epos.alloc((size_t)n));
    hipLaunchKernelGGL(k_entry_pos, dim3(grid_for(n, 256)), dim3(256), 0, c->stream, n, c->loc, c->conn.p,
                       c->rowptr.p, c->col.p, P.epos.p);
    FLGPU_HIP(hipGetLastError());
    P.epos_built = true;
    P.gather_built = false;  // built from epos
    return FLGPU_OK;
}

// The element tables in the kernels' LDS layout (LTab<R>) as 32-bit words,
// uploaded when they change.
template <typename R>
static int table_image(flgpu_ctx* c, const uint32_t** out) {
    BlockPlan& P = c->plan;
    LTab<R> lt;
    std::memset(&lt, 0, sizeof lt);
    const ElemTables& tt = c->tab;
    for (int k = 0; k < kMaxLoc * kMaxGi; k++) lt.n[k] = (R)tt.n[k];
    for (int k = 0; k < kMaxLoc * kMaxGi * kMaxDim; k++) lt.dn[k] = (R)tt.dn[k];
    for (int k = 0; k < kMaxGi; k++) lt.w[k] = (R)tt.w[k];
    for (int k = 0; k < kMaxLoc; k++) lt.what[k] = (R)tt.what[k];
    for (int k = 0; k < kMaxLoc * kMaxLoc; k++) lt.mhat[k] = (R)tt.mhat[k];
    for (int k = 0; k < 64; k++) lt.that[k] = (R)tt.that[k];
    for (int k = 0; k < 4 * kMaxDim; k++) lt.xdn[k] = (R)tt.xdn[k];
    lt.loc = tt.loc;
    lt.ngi = tt.ngi;
    lt.lx = tt.loc_x;
    DevBuf<uint32_t>& img = sizeof(R) == 8 ? P.ltab64 : P.ltab32;
    std::vector<uint32_t>& held = sizeof(R) == 8 ? P.h_ltab64 : P.h_ltab32;
    std::vector<uint32_t> words(sizeof lt / sizeof(uint32_t));
    std::memcpy(words.data(), &lt, sizeof lt);
    if (words != held || !img.p) {  // (the host copy outlives the asynchronous upload)
        held.swap(words);
        FLGPU_TRY(img.upload(held.data(), held.size(), c->stream));
    }
    *out = img.p;
    return FLGPU_OK;
}

// Stabilised assembly (streamline upwind / SUPG), any degree: one launch
// per colour, EPB elements per workgroup, thread = (element, local row i).
// The velocity and nu_bar at the quadrature points are computed once per
// element (the threads of an element share the points) into LDS; then each
// thread forms its row by the per-quadrature-point loop with the stabilised
// test function (row_stab_q) and adds it at the CSR positions of the
// per-mesh table.  The values are those of row_stab, bit for bit.
template <int DIM, int LOC, typename R>
struct StabLdsQuad : NodalQuad<DIM, LOC, R> {
    static constexpr bool nu_reads_k = false;
    const R* su;   // u at the points, [g][d]
    const R* snu;  // nu_bar at the points
    FLGPU_HD_INLINE void u(int g, R (&uq)[DIM]) const {
#pragma unroll
        for (int d = 0; d < DIM; d++) uq[d] = su[g * DIM + d];
    }
    FLGPU_HD_INLINE R nu(int g, const R (&)[DIM], const R (&)[DIM][DIM], const R (&)[DIM][DIM],
                         const AsmFlags&) const {
        return snu[g];
    }
};

template <int LOC>
constexpr int stab_epb() { return 256 / LOC; }

template <int DIM, int LOC, typename R>
__global__ void __launch_bounds__(256)
k_asm_colour_stab(const uint32_t* __restrict__ ltab, const int32_t* __restrict__ elems, int count,
                  const int32_t* __restrict__ conn, const int32_t* __restrict__ conn_x,
                  const double* __restrict__ X4, AsmFields fd, AsmFlags fl, const int32_t* __restrict__ epos,
                  double* __restrict__ A, double* __restrict__ rhs, int* __restrict__ err) {
    constexpr int EPB = stab_epb<LOC>(), NT = EPB * LOC;
    __shared__ LTab<R> tb;
    __shared__ R s_u[EPB][kElemMaxGi * DIM];
    __shared__ R s_nu[EPB][kElemMaxGi];
    stage_table_image<R, NT>(ltab, tb);
    __syncthreads();
    const int ql = threadIdx.x / LOC, i = threadIdx.x % LOC;
    const int64_t q = (int64_t)blockIdx.x * EPB + ql;
    const bool valid = q < count;
    Elem<DIM, LOC, R> e;
    R JT[DIM][DIM], iJ[DIM][DIM], adet = (R)0;
    int ele = 0;
    if (valid) {
        ele = elems[q];
        gather_elem<DIM, LOC, R>(ele, conn, conn_x, X4, fd.T, e);
        geometry_jt<DIM, R>(e.X, tb, JT, iJ, adet);
        const GlobalNodal<DIM, R> na{fd, e.nodes};
        const NodalQuad<DIM, LOC, R> nq{na, tb, iJ};
        for (int g = i; g < tb.ngi; g += LOC) {
            R uq[DIM], kt[DIM][DIM];
            nq.u(g, uq);
#pragma unroll
            for (int a = 0; a < DIM; a++)
#pragma unroll
                for (int b = 0; b < DIM; b++) kt[a][b] = (R)0;
            if (fl.diff) nq.k_tensor(g, kt);
            s_nu[ql][g] = nq.nu(g, uq, JT, kt, fl);
#pragma unroll
            for (int d = 0; d < DIM; d++) s_u[ql][g * DIM + d] = uq[d];
        }
    }
    __syncthreads();
    if (!valid) return;
    const GlobalNodal<DIM, R> na{fd, e.nodes};
    const StabLdsQuad<DIM, LOC, R> qv{{na, tb, iJ}, &s_u[ql][0], &s_nu[ql][0]};
    R vals[LOC], r;
    row_stab_q<DIM, LOC, R>(tb, fl, JT, iJ, adet, qv, e.T, i, vals, r);
    const int row = conn[(int64_t)ele * LOC + i];
#pragma unroll
    for (int j = 0; j < LOC; j++) {
        const double v = (double)vals[j];
        if (v == 0.0) continue;  // csr_addto skips zeros (Sparse_Tools.F90:2640)
        const int pos = epos[(int64_t)ele * LOC * LOC + i * LOC + j];
        if (pos < 0) {
            atomicOr(err, 1);
            continue;
        }
        A[pos] = A[pos] + v;
    }
    rhs[row] = rhs[row] + (double)r;
}

template <int DIM, int LOC, typename R>
static int launch_colour_stab(flgpu_ctx* c, const AsmFields& fd, const AsmFlags& fl, int* d_err) {
    FLGPU_TRY(ensure_entry_positions(c));
    const uint32_t* img = nullptr;
    FLGPU_TRY(table_image<R>(c, &img));
    for (int k = 0; k < c->ncol; k++) {
        const int off = c->h_col_ptr[k], cnt = c->h_col_ptr[k + 1] - off;
        if (cnt == 0) continue;
        constexpr int EPB = stab_epb<LOC>();
        hipLaunchKernelGGL((k_asm_colour_stab<DIM, LOC, R>), dim3((cnt + EPB - 1) / EPB), dim3(EPB * LOC), 0,
                           c->stream, img, c->col_elems.p + off, cnt, c->conn.p, c->conn_x.p, c->X4.p, fd, fl,
                           c->plan.epos.p, c->A.p, c->rhs.p, d_err);
    }
    FLGPU_HIP(hipGetLastError());
    return FLGPU_OK;
}

// Row i of an element by the general per-point row, with the equation
// types' fields and, for linear T, the mesh movement: |detJ| of the old and
// new positions (vertices relative to the last one, as gather_elem) and
// the grid velocity at the vertices.
template <int DIM, int LOC, typename R, class NA>
__device__ __forceinline__ void eq_row(int ele, const int32_t* __restrict__ conn_x, const Elem<DIM, LOC, R>& e,
                                       const LTab<R>& tb, const AsmFlags& fl, const AsmFields& fd, const NA& na,
                                       int i, R (&vals)[LOC], R& r) {
    if constexpr (LOC == DIM + 1) {
        if (fl.move) {
            MoveData<DIM, R> mv;
            const int32_t* cx = conn_x + (int64_t)ele * (DIM + 1);
            for (int w = 0; w < 2; w++) {
                const double* Xw = w ? fd.Xn4 : fd.Xo4;
                R Xr[DIM][DIM + 1], iJ[DIM][DIM], ad;
#pragma unroll
                for (int nn = 0; nn <= DIM; nn++)
#pragma unroll
                    for (int d = 0; d < DIM; d++)
                        Xr[d][nn] = (R)(Xw[(int64_t)cx[nn] * 4 + d] - Xw[(int64_t)cx[DIM] * 4 + d]);
                geometry<DIM, R>(Xr, tb, iJ, ad);
                if (w)
                    mv.adet_new = ad;
                else
                    mv.adet_old = ad;
            }
#pragma unroll
            for (int nn = 0; nn <= DIM; nn++)
#pragma unroll
                for (int d = 0; d < DIM; d++) mv.ug[nn][d] = (R)fd.UG4[(int64_t)cx[nn] * 4 + d];
            row_general_na<DIM, LOC, R>(e, tb, fl, na, i, vals, r, &mv);
            return;
        }
    }
    row_general_na<DIM, LOC, R>(e, tb, fl, na, i, vals, r);
}

// The first pass of the two-pass scheme for the equation types and mesh
// movement: a thread per (element, row) in element order, one launch (the
// element's LOC row threads adjacent: each stages its node's field values
// in shared memory for all of them), each row's values stored by element id
// at gval[e*LOC^2 + i*LOC + j] and its rhs at grv[e*LOC + i] (zeros
// included) for k_gather_sum's colour-order sums.
constexpr int kRowsEqThreads = 256;

template <int LOC>
__host__ __device__ constexpr int rows_eq_epb() {
    return kRowsEqThreads / LOC;
}

template <int DIM, int LOC, typename R>
__global__ void __launch_bounds__(kRowsEqThreads)
k_asm_rows_eq(const uint32_t* __restrict__ ltab, int nele, const int32_t* __restrict__ conn,
              const int32_t* __restrict__ conn_x, const double* __restrict__ X4, AsmFields fd, AsmFlags fl,
              double* __restrict__ gval, double* __restrict__ grv) {
    constexpr int EPB = rows_eq_epb<LOC>();
    constexpr int NF = shared_nodal_fields<DIM>();
    __shared__ LTab<R> tb;
    __shared__ R s_nodal[EPB * LOC * NF];
    stage_table_image<R, kRowsEqThreads>(ltab, tb);
    const int el = threadIdx.x / LOC, i = threadIdx.x - el * LOC;
    const int64_t ele64 = (int64_t)blockIdx.x * EPB + el;
    const bool live = el < EPB && ele64 < nele;
    const int ele = live ? (int)ele64 : 0;
    if (live) stage_shared_nodal<DIM, R>(fd, fl, conn[(int64_t)ele * LOC + i], s_nodal + (el * LOC + i) * NF);
    __syncthreads();
    if (!live) return;
    Elem<DIM, LOC, R> e;
    gather_elem<DIM, LOC, R>(ele, conn, conn_x, X4, fd.T, e);
    R vals[LOC], r;
    eq_row<DIM, LOC, R>(ele, conn_x, e, tb, fl, fd, SharedNodal<DIM, R>{s_nodal + el * LOC * NF}, i, vals, r);
    double* gv = gval + (int64_t)ele * LOC * LOC + i * LOC;
#pragma unroll
    for (int j = 0; j < LOC; j++) gv[j] = (double)vals[j];
    grv[(int64_t)ele * LOC + i] = (double)r;
}

// One launch per colour; thread = (element of the colour, local row i).
// Elements of one colour share no node, so the read-modify-writes of a
// launch never meet.  EQ: the internal-energy / k-epsilon equation types
// and mesh movement, every row by the general per-point row (eq_row).
template <int DIM, int LOC, typename R, bool EQ = false>
__global__ void __launch_bounds__(256)
k_asm_colour(const uint32_t* __restrict__ ltab, const int32_t* __restrict__ elems, int count,
             const int32_t* __restrict__ conn, const int32_t* __restrict__ conn_x,
             const double* __restrict__ X4, AsmFields fd, AsmFlags fl, const int32_t* __restrict__ epos,
             double* __restrict__ A, double* __restrict__ rhs, int* __restrict__ err) {
    __shared__ LTab<R> tb;
    stage_table_image<R, 256>(ltab, tb);
    __syncthreads();
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (int64_t)count * LOC) return;
    const int q = (int)(t / LOC), i = (int)(t % LOC);
    const int ele = elems[q];
    Elem<DIM, LOC, R> e;
    gather_elem<DIM, LOC, R>(ele, conn, conn_x, X4, fd.T, e);
    R vals[LOC], r;
    if constexpr (EQ)
        eq_row<DIM, LOC, R>(ele, conn_x, e, tb, fl, fd, GlobalNodal<DIM, R>{fd, e.nodes}, i, vals, r);
    else
        row_contrib<DIM, LOC, R>(e, tb, fl, fd, i, vals, r);
    const int row = conn[(int64_t)ele * LOC + i];  // (not e.nodes[i]: a run-time index would put e in scratch)
#pragma unroll
    for (int j = 0; j < LOC; j++) {
        const double v = (double)vals[j];
        if (v == 0.0) continue;  // csr_addto skips zeros (Sparse_Tools.F90:2640)
        const int pos = epos[(int64_t)ele * LOC * LOC + i * LOC + j];
        if (pos < 0) {
            atomicOr(err, 1);
            continue;
        }
        A[pos] = A[pos] + v;
    }
    rhs[row] = rhs[row] + (double)r;
}

template <int DIM, int LOC, typename R, bool EQ = false>
static int launch_colour(flgpu_ctx* c, const AsmFields& fd, const AsmFlags& fl, int* d_err) {
    FLGPU_TRY(ensure_entry_positions(c));
    const uint32_t* img = nullptr;
    FLGPU_TRY(table_image<R>(c, &img));
    for (int k = 0; k < c->ncol; k++) {
        const int off = c->h_col_ptr[k], cnt = c->h_col_ptr[k + 1] - off;
        if (cnt == 0) continue;
        hipLaunchKernelGGL((k_asm_colour<DIM, LOC, R, EQ>), dim3(grid_for((int64_t)cnt * LOC, 256)), dim3(256),
                           0, c->stream, img, c->col_elems.p + off, cnt, c->conn.p, c->conn_x.p, c->X4.p, fd, fl,
                           c->plan.epos.p, c->A.p, c->rhs.p, d_err);
    }
    FLGPU_HIP(hipGetLastError());
    return FLGPU_OK;
}

template <int DIM, int LOC, typename R>
static int launch_colour_eq(flgpu_ctx* c, const AsmFields& fd, const AsmFlags& fl, int* d_err) {
    return launch_colour<DIM, LOC, R, true>(c, fd, fl, d_err);
}

// Quadratic T: the fields' values at the quadrature points of one element,
// computed once per element into LDS (the element's threads split the
// points) instead of once per row from the nodal arrays.
template <int DIM, typename R>
struct LdsQuad {
    const R *uq, *dv, *kq, *sq, *fq;  // per point: DIM, 1, kst, 1, 1 values
    int kst;
    __device__ __forceinline__ void u(int g, R (&out)[DIM]) const {
#pragma unroll
        for (int d = 0; d < DIM; d++) out[d] = uq[g * DIM + d];
    }
    __device__ __forceinline__ R div(int g) const { return dv[g]; }
    __device__ __forceinline__ R k_iso(int g) const { return kq[g]; }
    __device__ __forceinline__ void k_tensor(int g, R (&kt)[DIM][DIM]) const {
#pragma unroll
        for (int a = 0; a < DIM; a++)
#pragma unroll
            for (int b = 0; b < DIM; b++) kt[a][b] = kq[g * DIM * DIM + a + DIM * b];
    }
    __device__ __forceinline__ R absorp(int g) const { return sq[g]; }
    __device__ __forceinline__ R src(int g) const { return fq[g]; }
};

// Nodal values of an element staged in LDS by its row threads (thread i
// holds node i): velocity, diffusivity (kst components), source,
// absorption.
template <int DIM, int LOC, typename R>
struct LdsNodal {
    const R *u, *k, *s, *b;
    int kst;
    __device__ __forceinline__ R U(int n, int d) const { return u[n * DIM + d]; }
    __device__ __forceinline__ R K(int n, int comp) const { return k[n * kst + comp]; }
    __device__ __forceinline__ R S(int n) const { return s[n]; }
    __device__ __forceinline__ R B(int n) const { return b[n]; }
};

// LDS values per element: per quadrature point DIM + 3 + kst, per node
// DIM + kst + 3 (velocity, diffusivity, source, absorption, T).
template <int DIM, int LOC>
__host__ __device__ constexpr int quad_stage_per_element(int kst) {
    return kElemMaxGi * (DIM + 3 + kst) + LOC * (DIM + kst + 3);
}

// Quadratic T, one thread per JB consecutive entries of an element row
// (LOC^2 / JB threads per element): the element's nodal values and
// coordinates are staged in LDS by the threads of each row's first part,
// the physical shape gradients at every point once per element, the
// fields' point values by the first ngi threads; the thread of entries
// (i, j0..j0+JB-1) forms them with entry_general_q, which reads row i's
// point values once for all of them, and row i's lumped mass and rhs are
// summed over the row's entries in column order, the row's two threads
// handing the partial sum on -- the arithmetic and order of row_general.  The CSR positions of the
// entries come from a table built once per mesh and sparsity, the element
// tables from a flat image (one memory round trip).
constexpr int kEntryThreads = 256;

// entries per thread: half a P2 row (tets: 5, triangles: 3)
template <int LOC>
__host__ __device__ constexpr int entry_jb() {
    return LOC == 10 ? 5 : LOC == 6 ? 3 : 1;
}

template <int DIM, int LOC>
__host__ __device__ constexpr int entry_stage_per_element(int kst) {
    return quad_stage_per_element<DIM, LOC>(kst) + DIM * (DIM + 1) + kElemMaxGi * LOC * DIM;
}

// elements per workgroup
template <int LOC>
__host__ __device__ constexpr int entry_epb() {
    return kEntryThreads / (LOC * (LOC / entry_jb<LOC>()));
}

// GATHER: the first pass of the two-pass scheme -- elems is null and the
// workgroups take the elements in id order (neighbours share nodes, so the
// staging gathers hit in L2); element e's entries and row values are stored
// at A[e*LOC^2 + entry] and rhs[e*LOC + row] (zeros included) for the
// per-entry sums of k_gather_sum; epos is not read.
// FEAT: F_ALL for isotropic (or no) diffusivity, F_ALL | F_ANISO for a
// tensor, so that the isotropic kernel holds no tensor in registers.
template <int DIM, int LOC, typename R, bool GATHER = false, int FEAT = F_ALL | F_ANISO>
__global__ void __launch_bounds__(kEntryThreads)
k_asm_colour_entry(const uint32_t* __restrict__ ltab, const int32_t* __restrict__ elems, int count,
                   const int32_t* __restrict__ conn, const int32_t* __restrict__ conn_x,
                   const double* __restrict__ X4, AsmFields fd, AsmFlags fl_in, const int32_t* __restrict__ epos,
                   double* __restrict__ A, double* __restrict__ rhs, int* __restrict__ err) {
    AsmFlags fl = fl_in;
    fl.iso = (FEAT & F_ANISO) ? fl.iso : 1;
    constexpr int L2 = LOC * LOC;
    constexpr int NT = kEntryThreads;
    constexpr int JB = entry_jb<LOC>();
    constexpr int RPL = LOC / JB;     // parts of a row
    constexpr int NTE = LOC * RPL;    // threads per element
    constexpr int EPB = NT / NTE;     // elements per workgroup
    constexpr int NG = kElemMaxGi;
    static_assert(LOC % JB == 0, "whole row parts");
    __shared__ LTab<R> tb;
    extern __shared__ double dynq[];
    const int kst = fl.iso ? 1 : DIM * DIM;
    const int t = threadIdx.x, el = t / NTE, te = t - el * NTE, i = te / RPL, jb = te - i * RPL, j0 = jb * JB;
    stage_table_image<R, NT>(ltab, tb);
    const int q = (int)blockIdx.x * EPB + el;
    const bool live = el < EPB && q < count;
    R* me = reinterpret_cast<R*>(dynq) + (size_t)(live ? el : 0) * entry_stage_per_element<DIM, LOC>(kst);
    R *s_uq = me, *s_dv = s_uq + NG * DIM, *s_kq = s_dv + NG, *s_sq = s_kq + NG * kst, *s_fq = s_sq + NG;
    R *n_u = s_fq + NG, *n_k = n_u + LOC * DIM, *n_s = n_k + LOC * kst, *n_b = n_s + LOC, *n_t = n_b + LOC;
    R* s_x = n_t + LOC;
    R* s_gr = s_x + DIM * (DIM + 1);  // physical shape gradients, [(g * LOC + i) * DIM + d]
    int ele = 0;
    if (live) {
        ele = elems ? elems[q] : q;
        if (jb == 0) {  // the row's first thread stages node i (and vertex i's coordinates)
            const int nd = conn[(int64_t)ele * LOC + i];
            n_t[i] = (R)fd.T[nd];
            if (fl.adv)
#pragma unroll
                for (int d = 0; d < DIM; d++) n_u[i * DIM + d] = fld<R>(fd.U, fd.u_c, nd, d, DIM);
            if (fl.diff)
                for (int cpt = 0; cpt < kst; cpt++) n_k[i * kst + cpt] = fld<R>(fd.K, fd.k_c, nd, cpt, DIM * DIM);
            if (fl.src) n_s[i] = fld<R>(fd.S, fd.s_c, nd, 0, 1);
            if (fl.absorp) n_b[i] = fld<R>(fd.B, fd.b_c, nd, 0, 1);
            if (i < DIM + 1) {  // relative to the last vertex, differenced in FP64 (as gather_elem)
                const int xn = conn_x[(int64_t)ele * (DIM + 1) + i];
                const int xl = conn_x[(int64_t)ele * (DIM + 1) + DIM];
#pragma unroll
                for (int a = 0; a < DIM; a++)
                    s_x[a * (DIM + 1) + i] = (R)(X4[(int64_t)xn * 4 + a] - X4[(int64_t)xl * 4 + a]);
            }
        }
    }
    __syncthreads();  // tables, nodal values, coordinates
    R iJ[DIM][DIM], adet = (R)0;
    if (live) {
        R X[DIM][DIM + 1];
#pragma unroll
        for (int a = 0; a < DIM; a++)
#pragma unroll
            for (int nn = 0; nn < DIM + 1; nn++) X[a][nn] = s_x[a * (DIM + 1) + nn];
        geometry<DIM, R>(X, tb, iJ, adet);
        // every (node, point) gradient once per element instead of twice per
        // entry and point (the same operations, so the same bits)
        for (int u = te; u < LOC * tb.ngi; u += NTE) {
            const int g = u / LOC, nn = u - g * LOC;
#pragma unroll
            for (int d = 0; d < DIM; d++) s_gr[u * DIM + d] = phys_grad<DIM, LOC, R>(tb, iJ, nn, g, d);
        }
        for (int g = te; g < tb.ngi; g += NTE) {  // the values at point g
            const LdsNodal<DIM, LOC, R> ln{n_u, n_k, n_s, n_b, kst};
            const NodalQuad<DIM, LOC, R, LdsNodal<DIM, LOC, R>> nq{ln, tb, iJ};
            if (fl.adv) {
                R u[DIM];
                nq.u(g, u);
#pragma unroll
                for (int d = 0; d < DIM; d++) s_uq[g * DIM + d] = u[d];
                if (fl.need_div) s_dv[g] = nq.div(g);
            }
            if (fl.diff) {
                if (fl.iso) {
                    s_kq[g] = nq.k_iso(g);
                } else {
                    R kt[DIM][DIM];
                    nq.k_tensor(g, kt);
#pragma unroll
                    for (int a = 0; a < DIM; a++)
#pragma unroll
                        for (int b = 0; b < DIM; b++) s_kq[g * DIM * DIM + a + DIM * b] = kt[a][b];
                }
            }
            if (fl.absorp) s_sq[g] = nq.absorp(g);
            if (fl.src) s_fq[g] = nq.src(g);
        }
    }
    __syncthreads();  // point values
    R m_ij[JB], s_ij[JB], src_i = (R)0;
#pragma unroll
    for (int b = 0; b < JB; b++) m_ij[b] = s_ij[b] = (R)0;
    if (live)
        entry_general_q<DIM, LOC, R, LdsQuad<DIM, R>, FEAT, JB>(
            tb, fl, iJ, adet, LdsQuad<DIM, R>{s_uq, s_dv, s_kq, s_sq, s_fq, kst}, i, j0, m_ij, s_ij, src_i, s_gr);
    // Row i's lumped mass and rhs: one chain over the row's entries in
    // column order, the first part's partial handed to the second part
    // (the adjacent lane: a row's two threads never straddle a wavefront).
    static_assert(RPL == 2, "a row in two parts");
    const int lane = threadIdx.x & 63;
    R lump = (R)0, r = (R)0;
    if (jb == 0) {
#pragma unroll
        for (int b = 0; b < JB; b++) {
            lump += m_ij[b];
            r = FMA(-s_ij[b], n_t[j0 + b], r);
        }
    }
    const R lump_in = __shfl(lump, (lane + 63) & 63), r_in = __shfl(r, (lane + 63) & 63);
    if (jb == 1) {
        lump = lump_in;
        r = r_in;
#pragma unroll
        for (int b = 0; b < JB; b++) {
            lump += m_ij[b];
            r = FMA(-s_ij[b], n_t[j0 + b], r);
        }
        if (fl.src) r = r + src_i;
    }
    const R lump_back = __shfl(lump, (lane + 1) & 63);
    if (jb == 0) lump = lump_back;  // the row's whole lumped mass in both parts
    if (!live) return;
    const R dtte = fl.dtt_on ? (R)fl.dt_theta : (R)0;
    if (GATHER) {
#pragma unroll
        for (int b = 0; b < JB; b++) {
            const int j = j0 + b;
            R m = (fl.have_mass && !fl.lump) ? m_ij[b] : (R)0;
            if (fl.have_mass && fl.lump && j == i) m = lump;
            A[(int64_t)ele * L2 + i * LOC + j] = (double)FMA(dtte, s_ij[b], m);
        }
        if (jb == 1) rhs[(int64_t)ele * LOC + i] = (double)r;
        return;
    }
#pragma unroll
    for (int b = 0; b < JB; b++) {
        const int j = j0 + b;
        R m = (fl.have_mass && !fl.lump) ? m_ij[b] : (R)0;
        if (fl.have_mass && fl.lump && j == i) m = lump;
        const double v = (double)FMA(dtte, s_ij[b], m);
        if (v != 0.0) {  // csr_addto skips zeros (Sparse_Tools.F90:2640)
            const int pos = epos[(int64_t)ele * L2 + i * LOC + j];
            if (pos < 0)
                atomicOr(err, 1);
            else
                A[pos] = A[pos] + v;
        }
    }
    if (jb == 1) {
        const int row = conn[(int64_t)ele * LOC + i];
        rhs[row] = rhs[row] + (double)r;
    }
}

template <int DIM, int LOC, typename R, int FEAT>
static int launch_colour_entry_feat(flgpu_ctx* c, const AsmFields& fd, const AsmFlags& fl, int* d_err) {
    constexpr int EPB = entry_epb<LOC>();
    FLGPU_TRY(ensure_entry_positions(c));
    const uint32_t* img = nullptr;
    FLGPU_TRY(table_image<R>(c, &img));
    const int kst = fl.iso ? 1 : DIM * DIM;
    const size_t lds = sizeof(R) * (size_t)EPB * entry_stage_per_element<DIM, LOC>(kst);
    auto kern = k_asm_colour_entry<DIM, LOC, R, false, FEAT>;
    FLGPU_HIP(hipFuncSetAttribute(reinterpret_cast<const void*>(kern), hipFuncAttributeMaxDynamicSharedMemorySize,
                                  (int)lds));
    for (int k = 0; k < c->ncol; k++) {
        const int off = c->h_col_ptr[k], cnt = c->h_col_ptr[k + 1] - off;
        if (cnt == 0) continue;
        hipLaunchKernelGGL(kern, dim3((cnt + EPB - 1) / EPB), dim3(kEntryThreads), lds, c->stream, img,
                           c->col_elems.p + off, cnt, c->conn.p, c->conn_x.p, c->X4.p, fd, fl, c->plan.epos.p,
                           c->A.p, c->rhs.p, d_err);
    }
    FLGPU_HIP(hipGetLastError());
    return FLGPU_OK;
}

template <int DIM, int LOC, typename R>
static int launch_colour_entry(flgpu_ctx* c, const AsmFields& fd, const AsmFlags& fl, int* d_err) {
    if (fl.diff && !fl.iso) return launch_colour_entry_feat<DIM, LOC, R, F_ALL | F_ANISO>(c, fd, fl, d_err);
    return launch_colour_entry_feat<DIM, LOC, R, F_ALL>(c, fd, fl, d_err);
}

// ---------------------------------------------------------------------
// Two-pass scheme (quadratic T).  Pass 1 (k_asm_colour_entry<..., true>)
// forms every element's entries and row values in one launch over all
// elements and stores them by element id; pass 2 (k_gather_sum) gives each CSR
// entry and rhs row one thread that adds its element values in colour
// order, ascending element id inside a colour -- the order in which the
// colour scheme's launches add them, and adding a zero leaves a sum's bits
// unchanged as csr_addto's skip of zeros does, so both schemes give the
// same bits.  The per-entry lists are built once per mesh, sparsity and
// colouring from the entry positions (counted, scanned, filled, and each
// list sorted by colour-order place, then rewritten as element ids).
// ---------------------------------------------------------------------

__global__ void __launch_bounds__(256)
k_gather_count(int64_t n, int loc, const int32_t* __restrict__ elems, const int32_t* __restrict__ epos,
               const int32_t* __restrict__ conn, int32_t* __restrict__ ecnt, int32_t* __restrict__ rcnt,
               int32_t* __restrict__ nbad) {
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n) return;
    const int l2 = loc * loc;
    const int64_t q = t / l2;
    const int rr = (int)(t - q * l2);
    const int64_t ele = elems[q];
    const int pos = epos[ele * l2 + rr];
    if (pos >= 0)
        atomicAdd(&ecnt[pos], 1);
    else
        atomicAdd(nbad, 1);
    if (rr % loc == 0) atomicAdd(&rcnt[conn[ele * loc + rr / loc]], 1);
}

__global__ void __launch_bounds__(256)
k_gather_fill(int64_t n, int loc, const int32_t* __restrict__ elems, const int32_t* __restrict__ epos,
              const int32_t* __restrict__ conn, const int32_t* __restrict__ eptr, const int32_t* __restrict__ rptr,
              int32_t* __restrict__ ecur, int32_t* __restrict__ rcur, int32_t* __restrict__ elist,
              int32_t* __restrict__ rlist, int32_t* __restrict__ bad, int32_t* __restrict__ nbad) {
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n) return;
    const int l2 = loc * loc;
    const int64_t q = t / l2;
    const int rr = (int)(t - q * l2);
    const int64_t ele = elems[q];
    const int pos = epos[ele * l2 + rr];
    if (pos >= 0)
        elist[eptr[pos] + atomicAdd(&ecur[pos], 1)] = (int32_t)t;
    else
        bad[atomicAdd(nbad, 1)] = (int32_t)t;
    if (rr % loc == 0) {
        const int row = conn[ele * loc + rr / loc];
        rlist[rptr[row] + atomicAdd(&rcur[row], 1)] = (int32_t)(q * loc + rr / loc);
    }
}

// Each list ascending (insertion sort; a list holds one value per element
// containing the entry's nodes, a few dozen at most).
__global__ void __launch_bounds__(256)
k_gather_sort(int n, const int32_t* __restrict__ ptr, int32_t* __restrict__ list) {
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= n) return;
    const int k0 = ptr[e], k1 = ptr[e + 1];
    for (int k = k0 + 1; k < k1; k++) {
        const int v = list[k];
        int m = k - 1;
        while (m >= k0 && list[m] > v) {
            list[m + 1] = list[m];
            m--;
        }
        list[m + 1] = v;
    }
}

// A list sorted by colour-order place (q*per + k) rewritten as element
// ids (elems[q]*per + k), the indexing of the first pass's output.
__global__ void __launch_bounds__(256)
k_gather_remap(int64_t n, int per, const int32_t* __restrict__ elems, int32_t* __restrict__ list) {
    const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    const int t = list[k], q = t / per;
    list[k] = elems[q] * per + (t - q * per);
}

__global__ void __launch_bounds__(256)
k_ptr32(int64_t n, const int64_t* __restrict__ in, int32_t* __restrict__ out) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i <= n) out[i] = (int32_t)in[i];
}

// Pass 2: thread per CSR entry (then per rhs row) summing its values.
__global__ void __launch_bounds__(256)
k_gather_sum(int nnz, int rows, const int32_t* __restrict__ eptr, const int32_t* __restrict__ elist,
             const double* __restrict__ gval, const int32_t* __restrict__ rptr, const int32_t* __restrict__ rlist,
             const double* __restrict__ grv, double* __restrict__ A, double* __restrict__ rhs) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    // U list items a step: their index loads, then their value gathers, all
    // in flight before the additions (in list order; a step past the list's
    // end re-reads its last item and adds nothing)
    constexpr int U = 4;
    auto run = [&](int k0, int k1, const int32_t* __restrict__ list, const double* __restrict__ val) {
        double s = 0.0;
        for (int kb = k0; kb < k1; kb += U) {
            int ix[U];
            double v[U];
#pragma unroll
            for (int u = 0; u < U; u++) ix[u] = list[min(kb + u, k1 - 1)];
#pragma unroll
            for (int u = 0; u < U; u++) v[u] = val[ix[u]];
#pragma unroll
            for (int u = 0; u < U; u++)
                if (kb + u < k1) s = s + v[u];
        }
        return s;
    };
    if (t < nnz) {
        A[t] = run(eptr[t], eptr[t + 1], elist, gval);
    } else if (t - nnz < rows) {
        const int r = t - nnz;
        rhs[r] = run(rptr[r], rptr[r + 1], rlist, grv);
    }
}

// An element entry outside the pattern is an error when its value is not
// zero (csr_addto skips zeros before it looks for the entry).
__global__ void __launch_bounds__(256)
k_gather_bad(int nbad, const int32_t* __restrict__ bad, const double* __restrict__ gval, int* __restrict__ err) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t < nbad && gval[bad[t]] != 0.0) atomicOr(err, 1);
}

static int build_gather_plan(flgpu_ctx* c) {
    BlockPlan& P = c->plan;
    FLGPU_TRY(ensure_entry_positions(c));
    if (P.gather_built) return FLGPU_OK;
    const int loc = c->loc;
    const int64_t n = (int64_t)c->nele * loc * loc;
    if (c->nnz >= INT32_MAX || n >= INT32_MAX)
        return set_error(FLGPU_EUNSUPPORTED, "two-pass assembly: more than 2^31 element entries");
    const hipStream_t s = c->stream;
    DevBuf<int32_t> ecnt, rcnt, nbad;
    FLGPU_TRY(ecnt.alloc((size_t)c->nnz));
    FLGPU_TRY(rcnt.alloc((size_t)c->rows));
    FLGPU_TRY(nbad.alloc(1));
    FLGPU_HIP(hipMemsetAsync(ecnt.p, 0, sizeof(int32_t) * c->nnz, s));
    FLGPU_HIP(hipMemsetAsync(rcnt.p, 0, sizeof(int32_t) * c->rows, s));
    FLGPU_HIP(hipMemsetAsync(nbad.p, 0, sizeof(int32_t), s));
    hipLaunchKernelGGL(k_gather_count, dim3(grid_for(n, 256)), dim3(256), 0, s, n, loc, c->col_elems.p, P.epos.p,
                       c->conn.p, ecnt.p, rcnt.p, nbad.p);
    FLGPU_HIP(hipGetLastError());
    DevBuf<int64_t> e64, r64;
    FLGPU_TRY(exclusive_scan(s, ecnt.p, c->nnz, e64));
    FLGPU_TRY(exclusive_scan(s, rcnt.p, c->rows, r64));
    int32_t h_nbad = 0;
    int64_t h_ne = 0, h_nr = 0;
    FLGPU_HIP(hipMemcpyAsync(&h_nbad, nbad.p, sizeof(int32_t), hipMemcpyDeviceToHost, s));
    FLGPU_HIP(hipMemcpyAsync(&h_ne, e64.p + c->nnz, sizeof(int64_t), hipMemcpyDeviceToHost, s));
    FLGPU_HIP(hipMemcpyAsync(&h_nr, r64.p + c->rows, sizeof(int64_t), hipMemcpyDeviceToHost, s));
    FLGPU_HIP(hipStreamSynchronize(s));
    FLGPU_TRY(P.gptr.alloc((size_t)c->nnz + 1));
    FLGPU_TRY(P.rptr.alloc((size_t)c->rows + 1));
    hipLaunchKernelGGL(k_ptr32, dim3(grid_for(c->nnz + 1, 256)), dim3(256), 0, s, c->nnz, e64.p, P.gptr.p);
    hipLaunchKernelGGL(k_ptr32, dim3(grid_for((int64_t)c->rows + 1, 256)), dim3(256), 0, s, (int64_t)c->rows,
                       r64.p, P.rptr.p);
    FLGPU_TRY(P.glist.alloc((size_t)std::max<int64_t>(h_ne, 1)));
    FLGPU_TRY(P.rlist.alloc((size_t)std::max<int64_t>(h_nr, 1)));
    FLGPU_TRY(P.gbad.alloc((size_t)std::max(h_nbad, 1)));
    FLGPU_HIP(hipMemsetAsync(ecnt.p, 0, sizeof(int32_t) * c->nnz, s));
    FLGPU_HIP(hipMemsetAsync(rcnt.p, 0, sizeof(int32_t) * c->rows, s));
    FLGPU_HIP(hipMemsetAsync(nbad.p, 0, sizeof(int32_t), s));
    hipLaunchKernelGGL(k_gather_fill, dim3(grid_for(n, 256)), dim3(256), 0, s, n, loc, c->col_elems.p, P.epos.p,
                       c->conn.p, P.gptr.p, P.rptr.p, ecnt.p, rcnt.p, P.glist.p, P.rlist.p, P.gbad.p, nbad.p);
    hipLaunchKernelGGL(k_gather_sort, dim3(grid_for(c->nnz, 256)), dim3(256), 0, s, (int)c->nnz, P.gptr.p,
                       P.glist.p);
    hipLaunchKernelGGL(k_gather_sort, dim3(grid_for(c->rows, 256)), dim3(256), 0, s, c->rows, P.rptr.p, P.rlist.p);
    if (h_ne > 0)
        hipLaunchKernelGGL(k_gather_remap, dim3(grid_for(h_ne, 256)), dim3(256), 0, s, h_ne, loc * loc,
                           c->col_elems.p, P.glist.p);
    if (h_nr > 0)
        hipLaunchKernelGGL(k_gather_remap, dim3(grid_for(h_nr, 256)), dim3(256), 0, s, h_nr, loc, c->col_elems.p,
                           P.rlist.p);
    if (h_nbad > 0)
        hipLaunchKernelGGL(k_gather_remap, dim3(grid_for(h_nbad, 256)), dim3(256), 0, s, (int64_t)h_nbad, loc * loc,
                           c->col_elems.p, P.gbad.p);
    FLGPU_HIP(hipGetLastError());
    FLGPU_TRY(P.gval.alloc((size_t)n));
    FLGPU_TRY(P.grv.alloc((size_t)c->nele * loc));
    FLGPU_HIP(hipStreamSynchronize(s));  // the counters are freed on return
    P.ngbad = h_nbad;
    P.gather_built = true;
    return FLGPU_OK;
}

template <int DIM, int LOC, typename R>
static int launch_gather(flgpu_ctx* c, const AsmFields& fd, const AsmFlags& fl, int* d_err) {
    constexpr int EPB = entry_epb<LOC>();
    FLGPU_TRY(build_gather_plan(c));
    BlockPlan& P = c->plan;
    const uint32_t* img = nullptr;
    FLGPU_TRY(table_image<R>(c, &img));
    const int kst = fl.iso ? 1 : DIM * DIM;
    const size_t lds = sizeof(R) * (size_t)EPB * entry_stage_per_element<DIM, LOC>(kst);
    auto kern = (fl.diff && !fl.iso) ? k_asm_colour_entry<DIM, LOC, R, true, F_ALL | F_ANISO>
                                     : k_asm_colour_entry<DIM, LOC, R, true, F_ALL>;
    FLGPU_HIP(hipFuncSetAttribute(reinterpret_cast<const void*>(kern), hipFuncAttributeMaxDynamicSharedMemorySize,
                                  (int)lds));
    hipLaunchKernelGGL(kern, dim3((c->nele + EPB - 1) / EPB), dim3(kEntryThreads), lds, c->stream, img, nullptr,
                       c->nele, c->conn.p, c->conn_x.p, c->X4.p, fd, fl, nullptr, P.gval.p, P.grv.p, d_err);
    hipLaunchKernelGGL(k_gather_sum, dim3(grid_for(c->nnz + c->rows, 256)), dim3(256), 0, c->stream, (int)c->nnz,
                       c->rows, P.gptr.p, P.glist.p, P.gval.p, P.rptr.p, P.rlist.p, P.grv.p, c->A.p, c->rhs.p);
    if (P.ngbad > 0)
        hipLaunchKernelGGL(k_gather_bad, dim3(grid_for(P.ngbad, 256)), dim3(256), 0, c->stream, P.ngbad, P.gbad.p,
                           P.gval.p, d_err);
    FLGPU_HIP(hipGetLastError());
    return FLGPU_OK;
}

// The equation types and mesh movement in two passes (the default for them):
// every row by eq_row in one launch, then the per-entry colour-order sums of
// the quadratic scheme (the same bits as their colour scheme).
template <int DIM, int LOC, typename R>
static int launch_gather_eq(flgpu_ctx* c, const AsmFields& fd, const AsmFlags& fl, int* d_err) {
    FLGPU_TRY(build_gather_plan(c));
    BlockPlan& P = c->plan;
    const uint32_t* img = nullptr;
    FLGPU_TRY(table_image<R>(c, &img));
    constexpr int EPB = rows_eq_epb<LOC>();
    hipLaunchKernelGGL((k_asm_rows_eq<DIM, LOC, R>), dim3((c->nele + EPB - 1) / EPB), dim3(kRowsEqThreads), 0,
                       c->stream, img, c->nele, c->conn.p, c->conn_x.p, c->X4.p, fd, fl, P.gval.p, P.grv.p);
    hipLaunchKernelGGL(k_gather_sum, dim3(grid_for(c->nnz + c->rows, 256)), dim3(256), 0, c->stream, (int)c->nnz,
                       c->rows, P.gptr.p, P.glist.p, P.gval.p, P.rptr.p, P.rlist.p, P.grv.p, c->A.p, c->rhs.p);
    if (P.ngbad > 0)
        hipLaunchKernelGGL(k_gather_bad, dim3(grid_for(P.ngbad, 256)), dim3(256), 0, c->stream, P.ngbad, P.gbad.p,
                           P.gval.p, d_err);
    FLGPU_HIP(hipGetLastError());
    return FLGPU_OK;
}

// ---------------------------------------------------------------------
// Blocked scheme (linear T, FP64).  The CSR rows are cut into slabs of
// consecutive rows, small enough that every element contribution to the
// slab fits in LDS; one workgroup assembles one slab.  A slot is a pair
// (row, element containing the row's node); the slots of a row follow the
// node's element list in colour order, ascending element id inside a
// colour (fixed by the colouring, built once per mesh):
//   phase A: one thread per slot computes the element's geometry and the
//     row of its element matrix and rhs, and stores them in the slot;
//   phase B: one thread per CSR entry of the slab adds the entry's
//     contributions slot by slot, i.e. in colour order -- the order of the
//     reference's colour loop (:574-598) and of the colour scheme above,
//     so both schemes give the same bits.  An element contributes to an
//     entry at most once, and an addend of zero leaves the sum unchanged,
//     exactly as csr_addto's skip of zeros (Sparse_Tools.F90:2640) does.
// The slab's entries and slots are contiguous in the CSR and incidence
// arrays, so they are read and written once, coalesced, and the matrix
// needs no zeroing pass.
// ---------------------------------------------------------------------

// Position of column cj in lcol[lo, hi) (sorted); -1 when absent.
__device__ __forceinline__ int lds_pos(const int32_t* lcol, int lo, int hi, int cj) {
    hi -= 1;
    while (lo <= hi) {
        const int mid = (lo + hi) >> 1;
        const int v = lcol[mid];
        if (v == cj) return mid;
        if (v < cj) lo = mid + 1;
        else hi = mid - 1;
    }
    return -1;
}

// Five waves per SIMD: the register cap (102) spills a few values of the
// FP64 element arithmetic, which costs less than the latency it hides.
template <int DIM, int FEAT>
__global__ void __launch_bounds__(256, 5)
k_asm_blocked(const ElemTables* __restrict__ tab, int nslab, int rows_cap, int nnz_cap, int slot_cap,
              const int32_t* __restrict__ brow, const int32_t* __restrict__ iptr, const int32_t* __restrict__ inc,
              const int32_t* __restrict__ conn, const int32_t* __restrict__ conn_x, const double* __restrict__ X4,
              AsmFields fd, AsmFlags fl, const int32_t* __restrict__ rowptr, const int32_t* __restrict__ col,
              double* __restrict__ A, double* __restrict__ rhs, int* __restrict__ err) {
    constexpr int LOC = DIM + 1;
    __shared__ LTab<double> tb;
    extern __shared__ double dyn[];
    double* sval = dyn;                                            // LOC * slot_cap, [j][slot]
    double* srhs = sval + (size_t)slot_cap * LOC;                  // slot_cap
    uint32_t* spos = reinterpret_cast<uint32_t*>(srhs + slot_cap); // slot_cap
    int32_t* lcol = reinterpret_cast<int32_t*>(spos + slot_cap);   // nnz_cap
    int32_t* lrp = lcol + nnz_cap;                                 // rows_cap + 1
    int32_t* soff = lrp + rows_cap + 1;                            // rows_cap + 1
    uint16_t* srow = reinterpret_cast<uint16_t*>(soff + rows_cap + 1);  // slot_cap: row of each slot
    uint16_t* erow = srow + slot_cap;                                   // nnz_cap: row of each entry
    stage_tables<double>(tab, tb);

    // Workgroups go to the 8 XCDs in turn (workgroup w to XCD w % 8); each
    // XCD takes a contiguous run of slabs so that neighbouring slabs, which
    // share elements and nodes, meet in the same L2.
    const int per = (nslab + 7) >> 3;
    const int b = (int)(blockIdx.x & 7) * per + (int)(blockIdx.x >> 3);
    if (b >= nslab) return;
    const int tid = threadIdx.x, nt = blockDim.x;
    const int r0 = brow[b], r1 = brow[b + 1], R = r1 - r0;
    const int k0 = rowptr[r0], E = rowptr[r1] - k0;
    const int s0 = iptr[r0], S = iptr[r1] - s0;
    for (int l = tid; l <= R; l += nt) {
        lrp[l] = rowptr[r0 + l] - k0;
        soff[l] = iptr[r0 + l] - s0;
    }
    for (int k = tid; k < E; k += nt) lcol[k] = col[k0 + k];
    __syncthreads();
    for (int l = tid; l < R; l += nt) {
        for (int k = soff[l]; k < soff[l + 1]; k++) srow[k] = (uint16_t)l;
        for (int k = lrp[l]; k < lrp[l + 1]; k++) erow[k] = (uint16_t)l;
    }
    __syncthreads();

    // phase A: one element row per slot
    for (int sl = tid; sl < S; sl += nt) {
        // the element tables are re-read from LDS in every iteration rather
        // than held in ~50 registers across the loop
        __asm__ __volatile__("" ::: "memory");
        const int ele = inc[s0 + sl];
        const int l = srow[sl], row = r0 + l;
        Elem<DIM, LOC, double> e;
        gather_elem<DIM, LOC, double>(ele, conn, conn_x, X4, fd.T, e);
        int i = 0;
#pragma unroll
        for (int k = 1; k < LOC; k++)
            if (e.nodes[k] == row) i = k;
        GeomLin<DIM, LOC, double> gm;
        geom_linear<DIM, LOC, double>(e, tb, gm);
        double vals[LOC], r;
        row_linear_g<DIM, LOC, double, FEAT>(e, gm, tb, fl, GlobalNodal<DIM, double>{fd, e.nodes}, i, vals, r);
        const int lo = lrp[l], hi = lrp[l + 1];
        uint32_t packed = 0;
#pragma unroll
        for (int j = 0; j < LOC; j++) {
            const int q = lds_pos(lcol, lo, hi, e.nodes[j]);
            if (q < 0) atomicOr(err, 1);
            packed |= (uint32_t)(q < 0 ? 255 : q - lo) << (8 * j);
            sval[j * slot_cap + sl] = vals[j];
        }
        spos[sl] = packed;
        srhs[sl] = r;
    }
    __syncthreads();

    // phase B: one thread per entry, contributions in slot (= colour) order
    for (int t = tid; t < E; t += nt) {
        const int l = erow[t];
        const uint32_t p = (uint32_t)(t - lrp[l]);
        double acc = 0.0;
#pragma unroll 4
        for (int k = soff[l]; k < soff[l + 1]; k++) {
            const uint32_t pk = spos[k];
            int j = -1;
#pragma unroll
            for (int jj = 0; jj < LOC; jj++)
                if (((pk >> (8 * jj)) & 255u) == p) j = jj;
            const double vv = sval[(j < 0 ? 0 : j) * slot_cap + k];
            acc = acc + (j < 0 ? 0.0 : vv);
        }
        A[k0 + t] = acc;
    }
    for (int l = tid; l < R; l += nt) {
        double acc = 0.0;
        for (int k = soff[l]; k < soff[l + 1]; k++) acc = acc + srhs[k];
        rhs[r0 + l] = acc;
    }
}

// Nodal field values from the slab's LDS copy (a constant field is staged
// as its one value at every node).  The diffusivity is staged as its first
// component (isotropic: every component reads it, so a row compiled for
// both cases reads one value) or, for the stabilised kernel with
// anisotropic diffusivity, as the whole tensor (TENSOR: component c at
// c * kstride; kstride 0 when only the first is staged).
template <int DIM, typename R, bool TENSOR = false>
struct StagedNodal {
    static constexpr bool eq_fields = false;  // the slabs stage no density
    const int* lid;
    const R *sU, *sK, *sS, *sB;
    int ncap;
    int kstride = 0;
    __device__ __forceinline__ R U(int k, int d) const { return (R)sU[d * ncap + lid[k]]; }
    __device__ __forceinline__ R K(int k, int comp) const {
        if constexpr (TENSOR) return (R)sK[comp * kstride + lid[k]];
        return (R)sK[lid[k]];
    }
    __device__ __forceinline__ R S(int k) const { return (R)sS[lid[k]]; }
    __device__ __forceinline__ R B(int k) const { return (R)sB[lid[k]]; }
};

// The blocked scheme with the slab's nodes staged in LDS: the slab's node
// list (its rows and every node of their elements) is read once, the
// coordinates and nodal fields of those nodes are gathered once into LDS,
// and each slot names its element's vertices by slab-local number, so the
// element rows read no global memory but their slot words.
//
// Contributions: a row with n slots receives 4n values (3n for triangles),
// which the plan orders by entry and, within an entry, by slot (colour
// order).  Each slot's second word says where its values go in that order
// (the row-relative place of each off-diagonal value, 9 bits each, and the
// number of the row's vertex in the element, 2 bits); the diagonal values
// of a row are consecutive from the row's diagonal offset.  Phase B then
// sums every entry's run of values in order: the same values in the same
// order as the colour scheme, so the same bits.

// The plan arrays and capacities the node-staged kernels read.
struct StagedPlan {
    const uint32_t* tabw;  // the element tables (LTabLin<R>) as 32-bit words
    const int32_t *meta, *iptr, *nodes;
    const uint16_t *rdg, *ecp;
    const uint32_t *slotw, *slotd;
    const uint8_t* srb;
    int nslab, rows_cap, nnz_cap, slot_cap, node_cap;
    int stage_u, stage_k, stage_s, stage_b;
    int kcomp;  // diffusivity components staged per node: 1 or DIM*DIM
};

// A slab's ranges (its 32-byte record) and its entry offsets' whole words.
struct SlabMeta {
    int r0, R, k0, E, s0, S, n0, NN, ce0, nwe;
};

__device__ __forceinline__ SlabMeta slab_meta(const int32_t* m) {
    SlabMeta s;
    s.r0 = m[0];
    s.R = m[1] - m[0];
    s.k0 = m[2];
    s.E = m[3] - m[2];
    s.s0 = m[4];
    s.S = m[5] - m[4];
    s.n0 = m[6];
    s.NN = m[7] - m[6];
    s.ce0 = s.k0 & ~1;                            // half-words as whole words
    s.nwe = (s.k0 + s.E - s.ce0 + 1) >> 1;
    return s;
}

// What a slab's record locates, per thread: its row's slot and diagonal
// offsets, its node id, its entry-offset words, and the vertex words,
// destination words and row bytes of the slots it computes.
// (Scalar members, not arrays: phase A picks a slot's words by its loop
// counter, which on an array would be a scratch-memory index.)
struct SlabWords {
    int vip, vdg, vn;
    uint32_t vs0, vs1, vs2, vd0, vd1, vd2, vr0, vr1, vr2, vcp0, vcp1;
};
static_assert(kStagedSlotsPerThread == 3 && kStagedColsPerThread == 4, "SlabWords holds 3 slots, 2 offset words");

__device__ __forceinline__ void load_slab_words(const StagedPlan& sp, const SlabMeta& m, int tid, SlabWords& w) {
    constexpr int NT = kStagedThreads;
    w.vip = tid <= m.R ? sp.iptr[m.r0 + tid] : 0;
    w.vdg = tid < m.R ? sp.rdg[m.r0 + tid] : 0;
    const uint32_t* sw = sp.slotw + m.s0 + tid;
    const uint32_t* sd = sp.slotd + m.s0 + tid;
    const uint8_t* sr = sp.srb + m.s0 + tid;
    const bool in0 = tid < m.S, in1 = tid + NT < m.S, in2 = tid + 2 * NT < m.S;
    w.vs0 = in0 ? sw[0] : 0u;
    w.vd0 = in0 ? sd[0] : 0u;
    w.vr0 = in0 ? sr[0] : 0u;
    w.vs1 = in1 ? sw[NT] : 0u;
    w.vd1 = in1 ? sd[NT] : 0u;
    w.vr1 = in1 ? sr[NT] : 0u;
    w.vs2 = in2 ? sw[2 * NT] : 0u;
    w.vd2 = in2 ? sd[2 * NT] : 0u;
    w.vr2 = in2 ? sr[2 * NT] : 0u;
    const uint32_t* cp = reinterpret_cast<const uint32_t*>(sp.ecp + m.ce0) + tid;
    w.vcp0 = tid < m.nwe ? cp[0] : 0u;
    w.vcp1 = tid + NT < m.nwe ? cp[NT] : 0u;
    w.vn = tid < m.NN ? sp.nodes[m.n0 + tid] : 0;
}

// One staged node's coordinates and fields (thread q holds node q); KC
// diffusivity components (sp.kcomp of them are staged).
template <int DIM, int KC = 1>
struct NodeVals {
    double x[DIM], t, k[KC], s, b, u[DIM];
};

template <int DIM, int KC>
__device__ __forceinline__ void load_node_vals(const StagedPlan& sp, const double* __restrict__ X4,
                                               const AsmFields& fd, int vn, bool have, NodeVals<DIM, KC>& v) {
#pragma unroll
    for (int a = 0; a < DIM; a++) v.x[a] = v.u[a] = 0.0;
#pragma unroll
    for (int a = 0; a < KC; a++) v.k[a] = 0.0;
    v.t = v.s = v.b = 0.0;
    if (have) {
#pragma unroll
        for (int a = 0; a < DIM; a++) v.x[a] = X4[(int64_t)vn * 4 + a];
        v.t = fd.T[vn];
        if (sp.stage_u)
#pragma unroll
            for (int a = 0; a < DIM; a++) v.u[a] = fd.U[fd.u_c ? a : (int64_t)vn * DIM + a];
        if (sp.stage_k)
#pragma unroll
            for (int a = 0; a < KC; a++)
                if (KC == 1 || a < sp.kcomp) v.k[a] = fd.K[(fd.k_c ? 0 : (int64_t)vn * DIM * DIM) + a];
        if (sp.stage_s) v.s = fd.S[fd.s_c ? 0 : vn];
        if (sp.stage_b) v.b = fd.B[fd.b_c ? 0 : vn];
    }
}

// A slab's LDS copy: node data (struct of arrays, in the arithmetic type
// R), row slot offsets, row diagonal offsets, entry offsets.
template <typename R>
struct SlabLds {
    double* sX;  // coordinates in FP64 for either precision (differenced before rounding)
    R *sT, *sU, *sK, *sS, *sB;
    int32_t *soff, *sdg;
    uint32_t* scp;
};

template <int DIM, typename R>
__host__ __device__ __forceinline__ size_t slab_lds_bytes(const StagedPlan& sp) {
    const size_t nf = (size_t)1 + (sp.stage_u ? DIM : 0) + (sp.stage_k ? sp.kcomp : 0) + sp.stage_s + sp.stage_b;
    const size_t b = sizeof(double) * DIM * sp.node_cap + sizeof(R) * nf * sp.node_cap +
                     sizeof(int32_t) * ((size_t)(sp.rows_cap + 1) + sp.rows_cap + ((sp.nnz_cap + 3) >> 1));
    return (b + 15) & ~(size_t)15;
}

template <int DIM, typename R>
__device__ __forceinline__ SlabLds<R> slab_lds(char* base, const StagedPlan& sp) {
    SlabLds<R> L;
    const int nc = sp.node_cap;
    L.sX = reinterpret_cast<double*>(base);
    L.sT = reinterpret_cast<R*>(L.sX + (size_t)DIM * nc);
    L.sU = L.sT + nc;
    L.sK = L.sU + (sp.stage_u ? (size_t)DIM * nc : 0);
    L.sS = L.sK + (sp.stage_k ? (size_t)sp.kcomp * nc : 0);
    L.sB = L.sS + (sp.stage_s ? nc : 0);
    L.soff = reinterpret_cast<int32_t*>(L.sB + (sp.stage_b ? nc : 0));
    L.sdg = L.soff + sp.rows_cap + 1;
    L.scp = reinterpret_cast<uint32_t*>(L.sdg + sp.rows_cap);
    return L;
}

template <typename R>
__device__ __forceinline__ void store_slab_words(const SlabWords& w, const SlabMeta& m, const SlabLds<R>& L,
                                                 int tid) {
    constexpr int NT = kStagedThreads;
    if (tid <= m.R) L.soff[tid] = w.vip - m.s0;
    if (tid < m.R) L.sdg[tid] = w.vdg;
    if (tid < m.nwe) L.scp[tid] = w.vcp0;
    if (tid + NT < m.nwe) L.scp[tid + NT] = w.vcp1;
}

template <int DIM, int KC, typename R>
__device__ __forceinline__ void store_node_vals(const StagedPlan& sp, const NodeVals<DIM, KC>& v, int NN,
                                                const SlabLds<R>& L, int q) {
    if (q >= NN) return;
    const int nc = sp.node_cap;
#pragma unroll
    for (int a = 0; a < DIM; a++) L.sX[a * nc + q] = v.x[a];
    L.sT[q] = (R)v.t;
    if (sp.stage_u)
#pragma unroll
        for (int a = 0; a < DIM; a++) L.sU[a * nc + q] = (R)v.u[a];
    if (sp.stage_k)
#pragma unroll
        for (int a = 0; a < KC; a++)
            if (KC == 1 || a < sp.kcomp) L.sK[a * nc + q] = (R)v.k[a];
    if (sp.stage_s) L.sS[q] = (R)v.s;
    if (sp.stage_b) L.sB[q] = (R)v.b;
}

// The element row of the unstabilised kernel: geometry by differences of
// the vertex coordinates, the compiled term set FEAT.
template <int DIM, int FEAT, typename R>
struct StagedRowLinear {
    const LTabLin<R>& tb;
    const AsmFlags& fl;
    const SlabLds<R>& L;
    int nc;
    __device__ __forceinline__ void operator()(const Elem<DIM, DIM + 1, R>& e, const int (&lid)[DIM + 1], int i, int,
                                               R (&vals)[DIM + 1], R& r) const {
        GeomLin<DIM, DIM + 1, R> gm;
        geom_linear_p1<DIM, DIM + 1, R>(e, gm);
        row_linear_g<DIM, DIM + 1, R, FEAT, StagedNodal<DIM, R>, LTabLin<R>>(
            e, gm, tb, fl, StagedNodal<DIM, R>{lid, L.sU, L.sK, L.sS, L.sB, nc}, i, vals, r);
    }
};

// The fields at the quadrature points from the staged nodal values, and
// nu_bar from the per-element array k_nu_points wrote.
template <int DIM, int LOC, typename R>
struct StagedNuQuad : NodalQuad<DIM, LOC, R, StagedNodal<DIM, R, true>> {
    static constexpr bool nu_reads_k = false;
    const R* __restrict__ nu_e;  // nu_bar at the element's points
    __device__ __forceinline__ R nu(int g, const R (&)[DIM], const R (&)[DIM][DIM], const R (&)[DIM][DIM],
                                    const AsmFlags&) const {
        return nu_e[g];
    }
};

// The element row with streamline-upwind / SUPG stabilisation: the
// Jacobian from the coordinate tables, velocity and the fields at every
// quadrature point from the staged nodal values, nu_bar from the
// per-element array (row_stab_q with the values the colour scheme's
// stabilised kernel computes, so the same bits).
template <int DIM, typename R, int FEAT>
struct StagedRowStab {
    const LTab<R>& tb;
    const AsmFlags& fl;
    const SlabLds<R>& L;
    int nc;
    int kstride;                      // diffusivity component stride in LDS (0: first only)
    const int32_t* __restrict__ inc;  // element of every slot
    const R* __restrict__ nub;        // nu_bar per element and point
    __device__ __forceinline__ void operator()(const Elem<DIM, DIM + 1, R>& e, const int (&lid)[DIM + 1], int i,
                                               int slot, R (&vals)[DIM + 1], R& r) const {
        constexpr int LOC = DIM + 1;
        const int ele = inc[slot];
        R JT[DIM][DIM], iJ[DIM][DIM], adet;
        geometry_jt<DIM, R>(e.X, tb, JT, iJ, adet);
        const StagedNodal<DIM, R, true> na{lid, L.sU, L.sK, L.sS, L.sB, nc, kstride};
        const StagedNuQuad<DIM, LOC, R> qv{{na, tb, iJ}, nub + (int64_t)ele * tb.ngi};
        row_stab_q<DIM, LOC, R, StagedNuQuad<DIM, LOC, R>, R[LOC], true, FEAT>(tb, fl, JT, iJ, adet, qv, e.T, i, vals,
                                                                               r);
    }
};

// Phase A: one element row per slot, from LDS (ROW computes it); its
// values to their places.
template <int DIM, typename R, class ROW>
__device__ __forceinline__ void staged_phase_a(const ROW& row, const StagedPlan& sp, const SlabMeta& m,
                                               const SlabLds<R>& L, const SlabWords& w, R* sc, R* srhs, int tid) {
    constexpr int LOC = DIM + 1;
    constexpr int NT = kStagedThreads;
    const int nc = sp.node_cap;
    // the slot words as plain register values (the empty asm keeps the
    // per-slot choice below a choice of values, not of addresses into a
    // scratch copy of the struct)
    uint32_t s0 = w.vs0, s1 = w.vs1, s2 = w.vs2, d0 = w.vd0, d1 = w.vd1, d2 = w.vd2;
    uint32_t q0 = w.vr0, q1 = w.vr1, q2 = w.vr2;
    __asm__("" : "+v"(s0), "+v"(s1), "+v"(s2), "+v"(d0), "+v"(d1), "+v"(d2), "+v"(q0), "+v"(q1), "+v"(q2));
#pragma unroll 1
    for (int u = 0; u < kStagedSlotsPerThread; u++) {
        const int sl = tid + u * NT;
        if (sl >= m.S) break;
        __asm__ __volatile__("" ::: "memory");
        const uint32_t vw = u == 0 ? s0 : u == 1 ? s1 : s2;
        const uint32_t dw = u == 0 ? d0 : u == 1 ? d1 : d2;
        const int l = (int)(u == 0 ? q0 : u == 1 ? q1 : q2);
        const int i = (int)(dw & 3u);
        int lid[LOC];
        Elem<DIM, LOC, R> e;
#pragma unroll
        for (int k = 0; k < LOC; k++) {
            lid[k] = (int)((vw >> (8 * k)) & 255u);
            e.T[k] = L.sT[lid[k]];
        }
        // vertices relative to the last one, differenced in FP64 (gather_elem)
#pragma unroll
        for (int a = 0; a < DIM; a++) {
            const double xo = L.sX[a * nc + lid[DIM]];
#pragma unroll
            for (int k = 0; k < DIM; k++) e.X[a][k] = (R)(L.sX[a * nc + lid[k]] - xo);
            e.X[a][DIM] = (R)0;
        }
        R vals[LOC], r;
        row(e, lid, i, m.s0 + sl, vals, r);
        const int ks = L.soff[l];
        R* cr = sc + LOC * ks;
        const int dg = L.sdg[l] + (sl - ks);
#pragma unroll
        for (int j = 0; j < LOC; j++) {
            const int q = j - (j > i ? 1 : 0);  // number among the off-diagonal vertices
            const int d = j == i ? dg : (int)((dw >> (2 + 9 * q)) & 511u);
            cr[d] = vals[j];
        }
        srhs[sl] = r;
    }
}

// Phase B: every entry sums its run of values (slot = colour order); every
// row its slots' rhs values.  The sums are FP64 whatever R is, as the
// colour scheme adds R values into the FP64 matrix.
template <int LOC, typename R>
__device__ __forceinline__ void staged_phase_b(const SlabMeta& m, const SlabLds<R>& L, const R* sc,
                                               const R* srhs, double* __restrict__ A,
                                               double* __restrict__ rhs, int tid) {
    constexpr int NT = kStagedThreads;
    const uint16_t* ecpl = reinterpret_cast<const uint16_t*>(L.scp) + (m.k0 - m.ce0);
    for (int t = tid; t < m.E; t += NT) {
        const int c0 = ecpl[t], c1 = t + 1 < m.E ? (int)ecpl[t + 1] : LOC * m.S;
        double acc = 0.0;
        for (int k = c0; k < c1; k++) acc = acc + (double)sc[k];
        A[m.k0 + t] = acc;
    }
    for (int l = tid; l < m.R; l += NT) {
        double acc = 0.0;
        for (int k = L.soff[l]; k < L.soff[l + 1]; k++) acc = acc + (double)srhs[k];
        rhs[m.r0 + l] = acc;
    }
}

// One workgroup per slab.  Three memory round trips stage the slab: (1)
// its record and the element tables, (2) what the record locates, (3) the
// nodes' coordinates and fields; every load of a step is issued before the
// first of them is used.
template <int DIM, int FEAT, typename R>
__global__ void __launch_bounds__(kStagedThreads, (FEAT & (F_ADV | F_ABS)) ? 4 : 5)
k_asm_staged(StagedPlan sp, const double* __restrict__ X4, AsmFields fd, AsmFlags fl, double* __restrict__ A,
             double* __restrict__ rhs) {
    constexpr int LOC = DIM + 1;
    constexpr int TW = (int)(sizeof(LTabLin<R>) / sizeof(uint32_t));
    static_assert(TW <= kStagedThreads, "one table word per thread");
    static_assert(sizeof(LTabLin<R>) % sizeof(uint32_t) == 0, "whole words");
    __shared__ LTabLin<R> tb;
    extern __shared__ double dyn[];
    R* sc = reinterpret_cast<R*>(dyn);                             // LOC * slot_cap: the contributions
    R* srhs = sc + (size_t)sp.slot_cap * LOC;                      // slot_cap
    const SlabLds<R> L = slab_lds<DIM, R>(reinterpret_cast<char*>(srhs + sp.slot_cap), sp);

    const int per = (sp.nslab + 7) >> 3;
    const int b = (int)(blockIdx.x & 7) * per + (int)(blockIdx.x >> 3);
    if (b >= sp.nslab) return;
    const int tid = threadIdx.x;
    const uint32_t tword = tid < TW ? sp.tabw[tid] : 0u;
    const SlabMeta m = slab_meta(sp.meta + 8 * (size_t)b);
    SlabWords w;
    load_slab_words(sp, m, tid, w);
    NodeVals<DIM> nv;
    load_node_vals<DIM, 1>(sp, X4, fd, w.vn, tid < m.NN, nv);
    if (tid < TW) reinterpret_cast<uint32_t*>(&tb)[tid] = tword;
    store_slab_words(w, m, L, tid);
    store_node_vals<DIM, 1, R>(sp, nv, m.NN, L, tid);
    __syncthreads();
    staged_phase_a<DIM, R>(StagedRowLinear<DIM, FEAT, R>{tb, fl, L, sp.node_cap}, sp, m, L, w, sc, srhs, tid);
    __syncthreads();
    staged_phase_b<LOC, R>(m, L, sc, srhs, A, rhs, tid);
}

// nu_bar at every quadrature point of every element, once per element (the
// staged rows of an element's vertices read it instead of each computing
// it): the values the colour scheme's stabilised kernel computes, from the
// same gathered coordinates, velocity and diffusivity.
template <int DIM, typename R, int SCHEME>
__global__ void __launch_bounds__(256)
k_nu_points(const uint32_t* __restrict__ ltab, int nele, const int32_t* __restrict__ conn,
            const int32_t* __restrict__ conn_x, const double* __restrict__ X4, AsmFields fd, AsmFlags fl,
            R* __restrict__ nub) {
    constexpr int LOC = DIM + 1;
    fl.nu_scheme = SCHEME;  // one scheme compiled in: the others' branches fold away
    __shared__ LTab<R> tb;
    stage_table_image<R, 256>(ltab, tb);
    __syncthreads();
    const int ele = (int)(blockIdx.x * blockDim.x + threadIdx.x);
    if (ele >= nele) return;
    Elem<DIM, LOC, R> e;
    gather_elem<DIM, LOC, R>(ele, conn, conn_x, X4, fd.T, e);
    R JT[DIM][DIM], iJ[DIM][DIM], adet;
    geometry_jt<DIM, R>(e.X, tb, JT, iJ, adet);
    const GlobalNodal<DIM, R> na{fd, e.nodes};
    const NodalQuad<DIM, LOC, R> nq{na, tb, iJ};
    for (int g = 0; g < tb.ngi; g++) {
        R uq[DIM], kt[DIM][DIM];
        nq.u(g, uq);
#pragma unroll
        for (int a = 0; a < DIM; a++)
#pragma unroll
            for (int b = 0; b < DIM; b++) kt[a][b] = (R)0;
        if (fl.diff) nq.k_tensor(g, kt);
        nub[(int64_t)ele * tb.ngi + g] = nq.nu(g, uq, JT, kt, fl);
    }
}

// The node-staged slab kernel with stabilisation: the same slabs, plan and
// summation order, the full element tables (the quadrature points' shape
// values and weights) in LDS beside the slab, nu_bar from k_nu_points, the
// diffusivity staged per node as its first component (isotropic) or the
// whole tensor.
template <int DIM, typename R, int FEAT>
__global__ void __launch_bounds__(kStagedThreads, (FEAT & (F_ABS | F_ANISO)) ? 2 : 3)
k_asm_staged_stab(StagedPlan sp, const uint32_t* __restrict__ ltab, const int32_t* __restrict__ inc,
                  const R* __restrict__ nub, const double* __restrict__ X4, AsmFields fd, AsmFlags fl,
                  double* __restrict__ A, double* __restrict__ rhs) {
    constexpr int LOC = DIM + 1;
    __shared__ LTab<R> tb;
    extern __shared__ double dyn[];
    R* sc = reinterpret_cast<R*>(dyn);
    R* srhs = sc + (size_t)sp.slot_cap * LOC;
    const SlabLds<R> L = slab_lds<DIM, R>(reinterpret_cast<char*>(srhs + sp.slot_cap), sp);

    const int per = (sp.nslab + 7) >> 3;
    const int b = (int)(blockIdx.x & 7) * per + (int)(blockIdx.x >> 3);
    if (b >= sp.nslab) return;
    const int tid = threadIdx.x;
    stage_table_image<R, kStagedThreads>(ltab, tb);
    const SlabMeta m = slab_meta(sp.meta + 8 * (size_t)b);
    SlabWords w;
    load_slab_words(sp, m, tid, w);
    NodeVals<DIM, DIM * DIM> nv;
    load_node_vals<DIM, DIM * DIM>(sp, X4, fd, w.vn, tid < m.NN, nv);
    store_slab_words(w, m, L, tid);
    store_node_vals<DIM, DIM * DIM, R>(sp, nv, m.NN, L, tid);
    __syncthreads();
    const int kstride = sp.kcomp > 1 ? sp.node_cap : 0;
    staged_phase_a<DIM, R>(StagedRowStab<DIM, R, FEAT>{tb, fl, L, sp.node_cap, kstride, inc, nub}, sp, m, L, w, sc,
                           srhs, tid);
    __syncthreads();
    staged_phase_b<LOC, R>(m, L, sc, srhs, A, rhs, tid);
}

static size_t blocked_lds(int loc, int rows_cap, int nnz_cap, int slot_cap) {
    return sizeof(double) * (size_t)slot_cap * (loc + 1) + sizeof(uint32_t) * (size_t)slot_cap +
           sizeof(int32_t) * ((size_t)nnz_cap + 2 * (size_t)(rows_cap + 1)) +
           sizeof(uint16_t) * ((size_t)slot_cap + nnz_cap);
}

static int build_block_plan_impl(flgpu_ctx* c);

// The plan once per mesh, sparsity and colouring; a failure is kept (and
// repeated) until one of them is bound again.
int build_block_plan(flgpu_ctx* c) {
    BlockPlan& P = c->plan;
    if (P.built) return FLGPU_OK;
    if (P.fail_rc) return set_error(P.fail_rc, "%s", P.fail_msg.c_str());
    const auto t0 = std::chrono::steady_clock::now();
    const int rc = build_block_plan_impl(c);
    P.build_seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    if (rc != FLGPU_OK) {
        P.fail_rc = rc;
        P.fail_msg = flgpu_last_error();
    }
    return rc;
}

static int build_block_plan_impl(flgpu_ctx* c) {
    if (c->plan.built) return FLGPU_OK;
    if (!c->linear) return set_error(FLGPU_EUNSUPPORTED, "block plan: linear field mesh only");
    const int N = c->rows, loc = c->loc;
    // LDS budget per slab: four to five workgroups per CU
    // Node staging needs the field nodes to be the coordinate nodes and each
    // slab's node list to hold at most 256 nodes (8-bit local numbers).
    bool stage = c->h_conn == c->h_conn_x && getenv("FLGPU_NO_STAGING") == nullptr;
    // the P1 element tables in the layout the kernel keeps in LDS
    LTabLin<double> lt;
    std::memset(&lt, 0, sizeof lt);
    {
        const ElemTables& t = c->tab;
        for (int k = 0; k < t.loc * kMaxDim && k < 4 * kMaxDim; k++) {
            const int nn = k % t.loc, d = k / t.loc;
            lt.dn[nn + t.loc * d] = t.dn[nn + t.loc * (0 + t.ngi * d)];
        }
        for (int k = 0; k < 4; k++) lt.what[k] = t.what[k];
        for (int k = 0; k < 16; k++) lt.mhat[k] = t.mhat[k];
        for (int k = 0; k < 64; k++) lt.that[k] = t.that[k];
        for (int k = 0; k < 4 * kMaxDim; k++) lt.xdn[k] = t.xdn[k];
        lt.loc = t.loc;
        lt.ngi = 1;
        lt.lx = t.loc_x;
    }
    // the staged kernel's geometry assumes the standard P1 tables
    if (!p1_standard_tables(lt, c->dim)) stage = false;
    int slot_cap = 384;
    if (const char* s = getenv("FLGPU_BLOCK_SLOTS")) slot_cap = std::max(64, std::min(4096, atoi(s)));
    // the staged kernel loads a slab's rows, entries and slots in one pass
    // of its workgroup: at most one row, four entries and two slots a thread
    if (stage) slot_cap = std::min(slot_cap, kStagedSlotsPerThread * kStagedThreads);
    const int rows_cap = stage ? kStagedThreads - 1 : 256;
    const int nnz_cap = stage ? kStagedColsPerThread * kStagedThreads - 4 : 1024;
    // node -> element incidence in colour order
    std::vector<int32_t> iptr((size_t)N + 1, 0);
    for (int64_t k = 0; k < (int64_t)c->nele * loc; k++) iptr[c->h_conn[k] + 1]++;
    for (int i = 0; i < N; i++) {
        const int d = iptr[i + 1];
        if (d > slot_cap)
            return set_error(FLGPU_EUNSUPPORTED, "block plan: node %d belongs to %d elements (a slab holds %d)",
                             i + 1, d, slot_cap);
        const int len = c->h_rowptr[i + 1] - c->h_rowptr[i];
        if (len > 255)
            return set_error(FLGPU_EUNSUPPORTED, "block plan: row %d has %d entries (max 255)", i + 1, len);
        if ((int64_t)iptr[i] + d > INT32_MAX) return set_error(FLGPU_EUNSUPPORTED, "block plan: too many slots");
        // the staged kernel places a row's values with 9-bit offsets
        if (loc * d > 512) stage = false;
        iptr[i + 1] += iptr[i];
    }
    std::vector<int32_t> inc(iptr[N]);
    {
        std::vector<int32_t> fill(iptr.begin(), iptr.end() - 1);
        for (int q = 0; q < c->nele; q++) {  // colour sets in order, ascending ids inside each
            const int e = c->h_col_elems[q];
            for (int i = 0; i < loc; i++) inc[fill[c->h_conn[(int64_t)e * loc + i]]++] = e;
        }
    }
    int node_cap_max = 192;
    if (const char* s = getenv("FLGPU_BLOCK_NODES")) node_cap_max = std::max(64, std::min(256, atoi(s)));
    // slabs of consecutive rows bounded by rows, entries, slots (and nodes)
    std::vector<int32_t> brow(1, 0), nptr(1, 0), nodes;
    std::vector<uint32_t> slotw;
    std::vector<int32_t> mark(stage ? N : 0, -1), fresh, stamp(stage ? N : 0, -1), posm(stage ? N : 0, 0);
    // per slot: destination word and slab-local row (bytes 4-byte padded);
    // per entry: slab-relative offset of its run of values (padded to whole
    // words past the end); per row: row-relative offset of the diagonal run
    std::vector<uint32_t> slotd;
    std::vector<uint8_t> srb(stage ? ((size_t)iptr[N] + 4) & ~(size_t)3 : 0, 0);
    std::vector<uint16_t> ecp(stage ? (((size_t)c->nnz + 2) & ~(size_t)1) + 2 : 0, 0);
    std::vector<uint16_t> rdg(stage ? (size_t)N : 0, 0), cur;
    int max_rows = 0, max_nnz = 0, max_slots = 0, max_nodes = 0;
    for (int r = 0; r < N;) {
        const int slab = (int)brow.size() - 1;
        int r1 = r, nn = 0;
        while (r1 < N && r1 - r < rows_cap && c->h_rowptr[r1 + 1] - c->h_rowptr[r] <= nnz_cap &&
               iptr[r1 + 1] - iptr[r] <= slot_cap) {
            if (stage) {  // nodes this row would add to the slab (the row's own node included)
                fresh.clear();
                if (mark[r1] != slab && mark[r1] != -2 - slab) {
                    mark[r1] = -2 - slab;
                    fresh.push_back(r1);
                }
                for (int q = iptr[r1]; q < iptr[r1 + 1]; q++)
                    for (int v = 0; v < loc; v++) {
                        const int g = c->h_conn[(int64_t)inc[q] * loc + v];
                        if (mark[g] != slab && mark[g] != -2 - slab) {
                            mark[g] = -2 - slab;
                            fresh.push_back(g);
                        }
                    }
                if (nn + (int)fresh.size() > node_cap_max) {
                    for (int g : fresh) mark[g] = -1;
                    if (r1 == r) return set_error(FLGPU_EUNSUPPORTED, "block plan: row %d touches > %d nodes", r + 1,
                                                    node_cap_max);
                    break;
                }
                for (int g : fresh) mark[g] = slab;
                nn += (int)fresh.size();
            }
            r1++;
        }
        max_rows = std::max(max_rows, r1 - r);
        max_nnz = std::max(max_nnz, c->h_rowptr[r1] - c->h_rowptr[r]);
        max_slots = std::max(max_slots, iptr[r1] - iptr[r]);
        if (stage) {
            // the slab's node list, ascending (its rows form one run in it),
            // and the slots' local vertex numbers
            const size_t first = nodes.size();
            for (int rr = r; rr < r1; rr++)
                if (mark[rr] == slab) {
                    mark[rr] = -3 - slab;
                    nodes.push_back(rr);
                }
            for (int q = iptr[r]; q < iptr[r1]; q++)
                for (int v = 0; v < loc; v++) {
                    const int g = c->h_conn[(int64_t)inc[q] * loc + v];
                    if (mark[g] == slab) {
                        mark[g] = -3 - slab;
                        nodes.push_back(g);
                    }
                }
            std::sort(nodes.begin() + first, nodes.end());
            for (size_t q = first; q < nodes.size(); q++) mark[nodes[q]] = (int)(q - first);  // local number
            for (int rr = r; rr < r1; rr++) {
                const int e0 = c->h_rowptr[rr], m = c->h_rowptr[rr + 1] - e0;
                for (int k = e0; k < e0 + m; k++) {
                    stamp[c->h_col[k]] = rr;
                    posm[c->h_col[k]] = k - e0;
                }
                // every vertex of the row's elements must be a column of the
                // row (csr_addto aborts otherwise, Sparse_Tools.F90:2645-2647)
                cur.assign((size_t)m + 1, 0);
                for (int q = iptr[rr]; q < iptr[rr + 1]; q++) {
                    srb[q] = (uint8_t)(rr - r);
                    for (int v = 0; v < loc; v++) {
                        const int g = c->h_conn[(int64_t)inc[q] * loc + v];
                        if (stamp[g] != rr)
                            return set_error(FLGPU_EPATTERN,
                                             "Attempting to set value in matrix outside sparsity pattern.");
                        cur[posm[g] + 1]++;
                    }
                }
                // runs of values in entry order; each run in slot order
                for (int p = 0; p < m; p++) cur[p + 1] += cur[p];
                const int rel = loc * (iptr[rr] - iptr[r]);
                for (int p = 0; p < m; p++) ecp[e0 + p] = (uint16_t)(rel + cur[p]);
                if (iptr[rr + 1] > iptr[rr]) rdg[rr] = cur[posm[rr]];
                for (int q = iptr[rr]; q < iptr[rr + 1]; q++) {
                    uint32_t dw = 0;
                    int nq = 0;
                    for (int v = 0; v < loc; v++) {
                        const int g = c->h_conn[(int64_t)inc[q] * loc + v];
                        const uint32_t d = cur[posm[g]]++;
                        if (g == rr) dw |= (uint32_t)v;
                        else dw |= d << (2 + 9 * nq++);
                    }
                    slotd.push_back(dw);
                }
            }
            for (int q = iptr[r]; q < iptr[r1]; q++) {
                uint32_t w = 0;
                for (int v = 0; v < loc; v++) w |= (uint32_t)mark[c->h_conn[(int64_t)inc[q] * loc + v]] << (8 * v);
                slotw.push_back(w);
            }
            for (size_t q = first; q < nodes.size(); q++) mark[nodes[q]] = -1;
            max_nodes = std::max(max_nodes, (int)(nodes.size() - first));
            nptr.push_back((int32_t)nodes.size());
        }
        brow.push_back(r1);
        r = r1;
    }
    BlockPlan& P = c->plan;
    P.staged = stage;
    if (stage) {
        const int nb = (int)brow.size() - 1;
        std::vector<int32_t> meta((size_t)nb * 8, 0);
        for (int b = 0; b < nb; b++) {
            int32_t* m = &meta[(size_t)b * 8];
            m[0] = brow[b];
            m[1] = brow[b + 1];
            m[2] = c->h_rowptr[brow[b]];
            m[3] = c->h_rowptr[brow[b + 1]];
            m[4] = iptr[brow[b]];
            m[5] = iptr[brow[b + 1]];
            m[6] = nptr[b];
            m[7] = nptr[b + 1];
        }
        FLGPU_TRY(P.meta.upload(meta.data(), meta.size(), c->stream));
        // the tables as the kernels keep them in LDS, FP64 and FP32
        std::vector<uint32_t> tw(sizeof lt / sizeof(uint32_t));
        std::memcpy(tw.data(), &lt, sizeof lt);
        FLGPU_TRY(P.tabw.upload(tw.data(), tw.size(), c->stream));
        LTabLin<float> lf;
        std::memset(&lf, 0, sizeof lf);
        for (int k = 0; k < 4 * kMaxDim; k++) lf.dn[k] = (float)lt.dn[k];
        for (int k = 0; k < 4; k++) lf.what[k] = (float)lt.what[k];
        for (int k = 0; k < 16; k++) lf.mhat[k] = (float)lt.mhat[k];
        for (int k = 0; k < 64; k++) lf.that[k] = (float)lt.that[k];
        for (int k = 0; k < 4 * kMaxDim; k++) lf.xdn[k] = (float)lt.xdn[k];
        lf.loc = lt.loc;
        lf.ngi = lt.ngi;
        lf.lx = lt.lx;
        std::vector<uint32_t> tw32(sizeof lf / sizeof(uint32_t));
        std::memcpy(tw32.data(), &lf, sizeof lf);
        FLGPU_TRY(P.tabw32.upload(tw32.data(), tw32.size(), c->stream));
        FLGPU_TRY(P.nptr.upload(nptr.data(), nptr.size(), c->stream));
        FLGPU_TRY(P.nodes.upload(nodes.data(), nodes.size(), c->stream));
        FLGPU_TRY(P.slotw.upload(slotw.data(), slotw.size(), c->stream));
        FLGPU_TRY(P.slotd.upload(slotd.data(), slotd.size(), c->stream));
        FLGPU_TRY(P.srb.upload(srb.data(), srb.size(), c->stream));
        FLGPU_TRY(P.ecp.upload(ecp.data(), ecp.size(), c->stream));
        FLGPU_TRY(P.rdg.upload(rdg.data(), rdg.size(), c->stream));
        P.node_cap = max_nodes;
        P.node_visits = (int64_t)nodes.size();
    }
    FLGPU_TRY(P.brow.upload(brow.data(), brow.size(), c->stream));
    FLGPU_TRY(P.iptr.upload(iptr.data(), iptr.size(), c->stream));
    FLGPU_TRY(P.inc.upload(inc.data(), inc.size(), c->stream));
    FLGPU_HIP(hipStreamSynchronize(c->stream));
    P.nblocks = (int)brow.size() - 1;
    P.rows_cap = max_rows;
    P.nnz_cap = max_nnz;
    P.slot_cap = max_slots;
    P.built = true;
    return FLGPU_OK;
}

template <int DIM, int FEAT>
static int launch_blocked_feat(flgpu_ctx* c, const AsmFields& fd, const AsmFlags& fl, int* d_err) {
    const BlockPlan& P = c->plan;
    constexpr int LOC = DIM + 1;
    const size_t lds = blocked_lds(LOC, P.rows_cap, P.nnz_cap, P.slot_cap);
    if (lds + sizeof(LTab<double>) + 64 > 160 * 1024)
        return set_error(FLGPU_EUNSUPPORTED, "blocked assembly: a slab needs %zu bytes of LDS", lds);
    FLGPU_HIP(hipFuncSetAttribute(reinterpret_cast<const void*>(&k_asm_blocked<DIM, FEAT>),
                                  hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
    const int grid = 8 * ((P.nblocks + 7) / 8);
    hipLaunchKernelGGL((k_asm_blocked<DIM, FEAT>), dim3(grid), dim3(256), lds, c->stream, c->d_tab.p, P.nblocks,
                       P.rows_cap, P.nnz_cap, P.slot_cap, P.brow.p, P.iptr.p, P.inc.p, c->conn.p, c->conn_x.p,
                       c->X4.p, fd, fl, c->rowptr.p, c->col.p, c->A.p, c->rhs.p, d_err);
    FLGPU_HIP(hipGetLastError());
    return FLGPU_OK;
}

template <int DIM, int FEAT, typename R>
static int launch_staged_feat(flgpu_ctx* c, const AsmFields& fd, const AsmFlags& fl, int* d_err) {
    const BlockPlan& P = c->plan;
    constexpr int LOC = DIM + 1;
    StagedPlan sp;
    sp.tabw = sizeof(R) == 8 ? P.tabw.p : P.tabw32.p;
    sp.meta = P.meta.p;
    sp.iptr = P.iptr.p;
    sp.nodes = P.nodes.p;
    sp.rdg = P.rdg.p;
    sp.ecp = P.ecp.p;
    sp.slotw = P.slotw.p;
    sp.slotd = P.slotd.p;
    sp.srb = P.srb.p;
    sp.nslab = P.nblocks;
    sp.rows_cap = P.rows_cap;
    sp.nnz_cap = P.nnz_cap;
    sp.slot_cap = P.slot_cap;
    sp.node_cap = P.node_cap;
    // the fields the compiled terms read, staged per node (constant ones too)
    sp.stage_u = (FEAT & F_ADV) && fl.adv;
    sp.stage_k = (FEAT & F_DIFF) && fl.diff;
    sp.stage_s = (FEAT & F_SRC) && fl.src;
    sp.stage_b = (FEAT & F_ABS) && fl.absorp;
    sp.kcomp = 1;
    const size_t head = sizeof(R) * (size_t)P.slot_cap * (LOC + 1), bb = slab_lds_bytes<DIM, R>(sp);
    const int per = (P.nblocks + 7) / 8;  // slabs per XCD
    const size_t lds = ((head + 15) & ~(size_t)15) + bb;
    if (lds + sizeof(LTabLin<R>) + 64 > 160 * 1024)
        return set_error(FLGPU_EUNSUPPORTED, "blocked assembly: a slab needs %zu bytes of LDS", lds);
    FLGPU_HIP(hipFuncSetAttribute(reinterpret_cast<const void*>(&k_asm_staged<DIM, FEAT, R>),
                                  hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
    hipLaunchKernelGGL((k_asm_staged<DIM, FEAT, R>), dim3(8 * per), dim3(kStagedThreads), lds, c->stream, sp,
                       c->X4.p, fd, fl, c->A.p, c->rhs.p);
    FLGPU_HIP(hipGetLastError());
    return FLGPU_OK;
}

// Stabilised assembly on the node-staged slabs (the plan's node lists and
// standard P1 tables); any diffusivity.  nu_bar once per element and
// quadrature point (k_nu_points), then the slabs.
template <int DIM, typename R>
static int launch_staged_stab(flgpu_ctx* c, const AsmFields& fd, const AsmFlags& fl) {
    const BlockPlan& P = c->plan;
    constexpr int LOC = DIM + 1;
    const uint32_t* img = nullptr;
    FLGPU_TRY(table_image<R>(c, &img));
    StagedPlan sp;
    sp.tabw = nullptr;
    sp.meta = P.meta.p;
    sp.iptr = P.iptr.p;
    sp.nodes = P.nodes.p;
    sp.rdg = P.rdg.p;
    sp.ecp = P.ecp.p;
    sp.slotw = P.slotw.p;
    sp.slotd = P.slotd.p;
    sp.srb = P.srb.p;
    sp.nslab = P.nblocks;
    sp.rows_cap = P.rows_cap;
    sp.nnz_cap = P.nnz_cap;
    sp.slot_cap = P.slot_cap;
    sp.node_cap = P.node_cap;
    sp.stage_u = 1;  // nu_bar needs the velocity whether or not advection is on
    sp.stage_k = fl.diff;
    sp.stage_s = fl.src;
    sp.stage_b = fl.absorp;
    // an isotropic row reads the first component only (nu_bar comes from
    // k_nu_points); an anisotropic one the whole tensor
    sp.kcomp = fl.iso ? 1 : DIM * DIM;
    const size_t head = sizeof(R) * (size_t)P.slot_cap * (LOC + 1), bb = slab_lds_bytes<DIM, R>(sp);
    const int per = (P.nblocks + 7) / 8;
    const size_t lds = ((head + 15) & ~(size_t)15) + bb;
    if (lds + sizeof(LTab<R>) + 64 > 160 * 1024)
        return set_error(FLGPU_EUNSUPPORTED, "stabilised blocked assembly: a slab needs %zu bytes of LDS", lds);
    FLGPU_TRY(c->plan.nub.reserve(((size_t)c->nele * c->tab.ngi * sizeof(R) + 7) / 8));
    R* nub = reinterpret_cast<R*>(c->plan.nub.p);
    auto nu_pass = [&](auto kern) {
        hipLaunchKernelGGL(kern, dim3(grid_for(c->nele, 256)), dim3(256), 0, c->stream, img, c->nele,
                           (const int32_t*)c->conn.p, (const int32_t*)c->conn_x.p, (const double*)c->X4.p, fd, fl,
                           nub);
    };
    switch (fl.nu_scheme) {
        case FLGPU_NU_BAR_OPTIMAL: nu_pass(&k_nu_points<DIM, R, FLGPU_NU_BAR_OPTIMAL>); break;
        case FLGPU_NU_BAR_DOUBLY_ASYMPTOTIC: nu_pass(&k_nu_points<DIM, R, FLGPU_NU_BAR_DOUBLY_ASYMPTOTIC>); break;
        case FLGPU_NU_BAR_CRITICAL_RULE: nu_pass(&k_nu_points<DIM, R, FLGPU_NU_BAR_CRITICAL_RULE>); break;
        default: nu_pass(&k_nu_points<DIM, R, FLGPU_NU_BAR_UNITY>); break;
    }
    FLGPU_HIP(hipGetLastError());
    // the common term set (no absorption, isotropic or no diffusivity)
    // compiled without the others, which frees registers for a third
    // workgroup per CU; the full set otherwise
    auto run = [&](auto kern) -> int {
        FLGPU_HIP(hipFuncSetAttribute(reinterpret_cast<const void*>(kern), hipFuncAttributeMaxDynamicSharedMemorySize,
                                      (int)lds));
        hipLaunchKernelGGL(kern, dim3(8 * per), dim3(kStagedThreads), lds, c->stream, sp, img, P.inc.p,
                           (const R*)nub, (const double*)c->X4.p, fd, fl, c->A.p, c->rhs.p);
        FLGPU_HIP(hipGetLastError());
        return FLGPU_OK;
    };
    if (!fl.absorp && (!fl.diff || fl.iso))
        return run(&k_asm_staged_stab<DIM, R, F_MASS | F_ADV | F_DIFF | F_SRC>);
    return run(&k_asm_staged_stab<DIM, R, F_ALL | F_ANISO>);
}

// The smallest compiled feature set that holds the terms switched on; the
// node-staged kernel when the plan has node lists and the diffusivity is
// isotropic (or constant).  FP32 element arithmetic has the node-staged
// kernel only.
template <int DIM, typename R>
static int launch_staged(flgpu_ctx* c, const AsmFields& fd, const AsmFlags& fl, int* d_err, int m) {
    auto fits = [m](int f) { return (m & ~f) == 0; };
    if (fits(F_DIFF | F_SRC)) return launch_staged_feat<DIM, F_DIFF | F_SRC, R>(c, fd, fl, d_err);
    if (fits(F_MASS | F_ADV | F_DIFF)) return launch_staged_feat<DIM, F_MASS | F_ADV | F_DIFF, R>(c, fd, fl, d_err);
    if (fits(F_MASS | F_ADV | F_DIFF | F_SRC))
        return launch_staged_feat<DIM, F_MASS | F_ADV | F_DIFF | F_SRC, R>(c, fd, fl, d_err);
    return launch_staged_feat<DIM, F_ALL, R>(c, fd, fl, d_err);
}

static bool stageable(const flgpu_ctx* c, const AsmFlags& fl) { return c->plan.staged && (!fl.diff || fl.iso); }

// The stabilised node-staged kernel forms the gradients once per row: the
// tables' derivatives must be the same at every quadrature point.
static bool gradients_constant(const ElemTables& t) {
    for (int d = 0; d < t.dim; d++)
        for (int g = 1; g < t.ngi; g++)
            for (int n = 0; n < t.loc; n++)
                if (t.dn[n + t.loc * (g + t.ngi * d)] != t.dn[n + t.loc * (0 + t.ngi * d)]) return false;
    return true;
}

template <int DIM>
static int launch_blocked(flgpu_ctx* c, const AsmFields& fd, const AsmFlags& fl, int* d_err, bool fp32) {
    const int m = (fl.have_mass ? F_MASS : 0) | (fl.adv ? F_ADV : 0) | (fl.absorp ? F_ABS : 0) |
                  (fl.diff ? F_DIFF : 0) | (fl.src ? F_SRC : 0);
    auto fits = [m](int f) { return (m & ~f) == 0; };
    if (fp32) return launch_staged<DIM, float>(c, fd, fl, d_err, m);
    if (stageable(c, fl)) return launch_staged<DIM, double>(c, fd, fl, d_err, m);
    if (fits(F_DIFF | F_SRC)) return launch_blocked_feat<DIM, F_DIFF | F_SRC>(c, fd, fl, d_err);
    if (fits(F_MASS | F_ADV | F_DIFF)) return launch_blocked_feat<DIM, F_MASS | F_ADV | F_DIFF>(c, fd, fl, d_err);
    if (fits(F_MASS | F_ADV | F_DIFF | F_SRC))
        return launch_blocked_feat<DIM, F_MASS | F_ADV | F_DIFF | F_SRC>(c, fd, fl, d_err);
    return launch_blocked_feat<DIM, F_ALL>(c, fd, fl, d_err);
}

static int assemble_launch(flgpu_ctx* c, const flgpu_advdif_params* p, int* d_err) {
    if (c->nele <= 0 || c->loc <= 0 || c->nnz <= 0 || c->ncol <= 0)
        return set_error(FLGPU_EINVAL, "flgpu assemble: bind mesh, field mesh, sparsity and colouring first");
    if (c->rows != c->nnodes)
        return set_error(FLGPU_EINVAL, "flgpu assemble: sparsity rows %d != field nodes %d", c->rows, c->nnodes);
    if (c->h_conn.size() != (size_t)c->nele * c->loc || c->h_conn_x.size() != (size_t)c->nele * c->loc_x ||
        c->h_col_elems.size() != (size_t)c->nele || c->h_col_ptr.size() != (size_t)c->ncol + 1)
        return set_error(FLGPU_EINVAL, "flgpu assemble: the field mesh or colouring was bound for another mesh "
                                       "(%zu connectivity entries, %zu coloured elements for %d elements)",
                         c->h_conn.size(), c->h_col_elems.size(), c->nele);
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
    fl.stab = p->stabilisation;
    fl.nu_scheme = p->nu_bar_scheme;
    fl.nu_scale = p->nu_bar_scale;
    if (fl.stab < 0 || fl.stab > FLGPU_STAB_SUPG)
        return set_error(FLGPU_EINVAL, "flgpu assemble: stabilisation %d", fl.stab);
    if (fl.stab) {
        if (fl.nu_scheme < FLGPU_NU_BAR_OPTIMAL || fl.nu_scheme > FLGPU_NU_BAR_UNITY)
            return set_error(FLGPU_EINVAL, "flgpu assemble: nu_bar scheme %d", fl.nu_scheme);
        if (!(fl.nu_scale >= 0.0))
            return set_error(FLGPU_EINVAL, "flgpu assemble: nu_bar scale %g < 0", fl.nu_scale);
        if (!c->fU.set) return set_error(FLGPU_EINVAL, "flgpu assemble: stabilisation needs the velocity field");
    }
    // the equation type (Advection_Diffusion_CG.F90:484-552)
    fl.eq = p->equation_type;
    if (fl.eq < FLGPU_EQ_ADVECTION_DIFFUSION || fl.eq > FLGPU_EQ_KEPSILON)
        return set_error(FLGPU_EINVAL, "flgpu assemble: equation type %d", fl.eq);
    fl.rho_theta = p->density_theta;
    fl.comp = fl.eq == FLGPU_EQ_INTERNAL_ENERGY && p->compressible != 0;
    fl.mph = fl.eq == FLGPU_EQ_INTERNAL_ENERGY && p->multiphase != 0;
    // mesh movement (:421-437: move_mesh needs the mass terms)
    fl.move = p->move_mesh != 0 && fl.have_mass;
    fl.dt = p->dt;
    if (fl.move) {
        if (c->motion_nx != c->nx || c->motion_dim != c->dim)
            return set_error(FLGPU_EINVAL, "flgpu assemble: mesh movement needs flgpu_mesh_motion_set for this mesh");
        if (!c->linear)
            return set_error(FLGPU_EUNSUPPORTED, "flgpu assemble: mesh movement takes linear T");
        // the reference stops here too (:447, :462, :500, :531)
        if (fl.stab || fl.eq)
            return set_error(FLGPU_EUNSUPPORTED, "flgpu assemble: mesh movement with stabilisation or the "
                                                 "internal-energy / k-epsilon equations");
        if (!(p->dt > 0.0)) return set_error(FLGPU_EINVAL, "flgpu assemble: mesh movement needs dt > 0");
    }
    if (fl.eq) {
        if (!c->fRho.set) return set_error(FLGPU_EINVAL, "flgpu assemble: density not set");
        if (fl.eq == FLGPU_EQ_INTERNAL_ENERGY && !c->fRhoOld.set)
            return set_error(FLGPU_EINVAL, "flgpu assemble: old density not set");
        if (fl.comp && (!c->fP.set || !c->fU.set))
            return set_error(FLGPU_EINVAL, "flgpu assemble: the p div(u) term needs pressure and velocity");
        if (fl.mph && !c->fV.set) return set_error(FLGPU_EINVAL, "flgpu assemble: phase volume fraction not set");
    }
    AsmFields fd;
    fd.T = c->fT.v.p;
    fd.U = c->fU.v.p; fd.u_c = c->fU.constant;
    fd.K = c->fK.v.p; fd.k_c = c->fK.constant;
    fd.S = c->fS.v.p; fd.s_c = c->fS.constant;
    fd.B = c->fA.v.p; fd.b_c = c->fA.constant;
    fd.Rho = c->fRho.v.p; fd.rho_c = c->fRho.constant;
    fd.RhoOld = c->fRhoOld.v.p; fd.rho_old_c = c->fRhoOld.constant;
    fd.P = c->fP.v.p; fd.p_c = c->fP.constant;
    fd.V = c->fV.v.p; fd.v_c = c->fV.constant;
    fd.Xo4 = c->Xo4.p;
    fd.Xn4 = c->Xn4.p;
    fd.UG4 = c->UG4.p;

    const int dim = c->dim, loc = c->loc;
    const bool fp32 = p->precision == FLGPU_FP32;
    int rc = FLGPU_OK;
    int scheme = p->scheme;
    // the blocked scheme: linear T; in FP32, and with stabilisation, with
    // the node-staged kernels only.  Chosen automatically, it gives way to
    // the colour scheme when the mesh does not fit its plan (a node in more
    // elements than a slab holds, a row longer than 255 entries).  The
    // colour scheme asked for by name needs no plan.
    // intervals (dim 1) take the colour scheme: a row has at most three
    // entries, too few for the slab staging to pay
    bool blocked_ok = c->linear && dim >= 2 && !fl.eq && !fl.move && scheme != FLGPU_ASM_COLOUR;
    if (blocked_ok && (fp32 || fl.stab || scheme == FLGPU_ASM_AUTO)) {
        const int rc = build_block_plan(c);
        if (rc == FLGPU_EUNSUPPORTED && scheme == FLGPU_ASM_AUTO)
            blocked_ok = false;
        else if (rc != FLGPU_OK)
            return rc;
        else if (fl.stab)
            blocked_ok = c->plan.staged && gradients_constant(c->tab);
        else if (fp32)
            blocked_ok = stageable(c, fl);
    }
    // quadratic T, unstabilised: the two-pass scheme is the blocked one
    const bool two_pass = !c->linear && dim >= 2 && !fl.stab && !fl.eq && scheme != FLGPU_ASM_COLOUR;
    // the equation types and mesh movement, unstabilised: two passes unless
    // the colour scheme is asked for by name (the same bits)
    const bool two_pass_eq = (fl.eq || fl.move) && !fl.stab && scheme != FLGPU_ASM_COLOUR;
    if (scheme == FLGPU_ASM_BLOCKED && !blocked_ok && !two_pass && !two_pass_eq)
        return set_error(FLGPU_EUNSUPPORTED, "flgpu assemble: the blocked scheme covers linear T (in FP32 and with "
                                             "stabilisation: with standard P1 tables and the field on the "
                                             "coordinate mesh; in FP32 unstabilised: isotropic diffusivity) and "
                                             "unstabilised quadratic T");
    if (scheme == FLGPU_ASM_AUTO) scheme = blocked_ok ? FLGPU_ASM_BLOCKED : FLGPU_ASM_COLOUR;
    if (scheme != FLGPU_ASM_COLOUR && scheme != FLGPU_ASM_BLOCKED)
        return set_error(FLGPU_EINVAL, "flgpu assemble: scheme %d", p->scheme);

    // a fresh assembly has no Dirichlet rows (the mask is all zero unless a
    // Dirichlet application or an upload set it)
    if (c->have_inactive) {
        FLGPU_HIP(hipMemsetAsync(c->inactive.p, 0, sizeof(int32_t) * c->rows, c->stream));
        c->have_inactive = false;
    }
    if (two_pass) {
        if (dim == 3 && loc == 10)
            return fp32 ? launch_gather<3, 10, float>(c, fd, fl, d_err) : launch_gather<3, 10, double>(c, fd, fl, d_err);
        if (dim == 2 && loc == 6)
            return fp32 ? launch_gather<2, 6, float>(c, fd, fl, d_err) : launch_gather<2, 6, double>(c, fd, fl, d_err);
        return set_error(FLGPU_EUNSUPPORTED, "flgpu assemble: dim %d loc %d", dim, loc);
    }
    if (two_pass_eq) {
#define FL_GATHER_EQ(D, L) \
    if (dim == D && loc == L) \
        return fp32 ? launch_gather_eq<D, L, float>(c, fd, fl, d_err) : launch_gather_eq<D, L, double>(c, fd, fl, d_err);
        FL_GATHER_EQ(1, 2)
        FL_GATHER_EQ(1, 3)
        FL_GATHER_EQ(2, 3)
        FL_GATHER_EQ(2, 6)
        FL_GATHER_EQ(3, 4)
        FL_GATHER_EQ(3, 10)
#undef FL_GATHER_EQ
        return set_error(FLGPU_EUNSUPPORTED, "flgpu assemble: dim %d loc %d", dim, loc);
    }
    if (scheme == FLGPU_ASM_BLOCKED) {
        FLGPU_TRY(build_block_plan(c));
        if (fl.stab) {
            if (dim == 3) return fp32 ? launch_staged_stab<3, float>(c, fd, fl) : launch_staged_stab<3, double>(c, fd, fl);
            return fp32 ? launch_staged_stab<2, float>(c, fd, fl) : launch_staged_stab<2, double>(c, fd, fl);
        }
        return dim == 3 ? launch_blocked<3>(c, fd, fl, d_err, fp32) : launch_blocked<2>(c, fd, fl, d_err, fp32);
    }
    FLGPU_HIP(hipMemsetAsync(c->A.p, 0, sizeof(double) * c->nnz, c->stream));
    FLGPU_HIP(hipMemsetAsync(c->rhs.p, 0, sizeof(double) * c->rows, c->stream));
    // linear T: a thread per element row; quadratic T: a thread per entry
#define FL_DISPATCH(D, L, LAUNCH)                                                            \
    if (dim == D && loc == L) {                                                              \
        rc = fp32 ? LAUNCH<D, L, float>(c, fd, fl, d_err) : LAUNCH<D, L, double>(c, fd, fl, d_err); \
    } else
    if (fl.stab) {
        FL_DISPATCH(1, 2, launch_colour_stab)
        FL_DISPATCH(1, 3, launch_colour_stab)
        FL_DISPATCH(2, 3, launch_colour_stab)
        FL_DISPATCH(2, 6, launch_colour_stab)
        FL_DISPATCH(3, 4, launch_colour_stab)
        FL_DISPATCH(3, 10, launch_colour_stab) {
            return set_error(FLGPU_EUNSUPPORTED, "flgpu assemble: dim %d loc %d", dim, loc);
        }
        return rc;
    }
    if (fl.eq || fl.move) {
        FL_DISPATCH(1, 2, launch_colour_eq)
        FL_DISPATCH(1, 3, launch_colour_eq)
        FL_DISPATCH(2, 3, launch_colour_eq)
        FL_DISPATCH(2, 6, launch_colour_eq)
        FL_DISPATCH(3, 4, launch_colour_eq)
        FL_DISPATCH(3, 10, launch_colour_eq) {
            return set_error(FLGPU_EUNSUPPORTED, "flgpu assemble: dim %d loc %d", dim, loc);
        }
        return rc;
    }
    FL_DISPATCH(1, 2, launch_colour)
    FL_DISPATCH(1, 3, launch_colour)  // a P2 interval row has three entries: a thread per row
    FL_DISPATCH(2, 3, launch_colour)
    FL_DISPATCH(2, 6, launch_colour_entry)
    FL_DISPATCH(3, 4, launch_colour)
    FL_DISPATCH(3, 10, launch_colour_entry) {
        return set_error(FLGPU_EUNSUPPORTED, "flgpu assemble: dim %d loc %d", dim, loc);
    }
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
    c->assembled = false;  // set again only when this assembly succeeds
    DevBuf<int> d_err;
    FLGPU_TRY(d_err.alloc(1));
    FLGPU_HIP(hipMemsetAsync(d_err.p, 0, sizeof(int), c->stream));
    FLGPU_TRY(assemble_launch(c, p, d_err.p));
    return assemble_check(c, d_err.p);
}

int assembly_plan_info(flgpu_ctx* c, int64_t* info, int n, double* build_seconds) {
    if (c->nele <= 0 || c->nnz <= 0 || c->ncol <= 0)
        return set_error(FLGPU_EINVAL, "flgpu_assembly_plan_info: bind mesh, sparsity and colouring first");
    FLGPU_TRY(build_block_plan(c));
    const BlockPlan& P = c->plan;
    StagedPlan sp{};
    sp.rows_cap = P.rows_cap;
    sp.nnz_cap = P.nnz_cap;
    sp.node_cap = P.node_cap;
    sp.stage_k = sp.stage_s = 1;
    sp.kcomp = 1;
    size_t lds = 0;
    if (P.staged) {
        const size_t head = (sizeof(double) * (size_t)P.slot_cap * (c->loc + 1) + 15) & ~(size_t)15;
        lds = head + (c->dim == 3 ? slab_lds_bytes<3, double>(sp) : slab_lds_bytes<2, double>(sp)) +
              sizeof(LTabLin<double>);
    }
    const int64_t v[10] = {P.staged ? 1 : 0, P.nblocks, P.rows_cap, P.nnz_cap, P.slot_cap, P.node_cap, P.node_visits,
                           (int64_t)c->nele * c->loc, (int64_t)lds, lds ? (int64_t)(160 * 1024 / lds) : 0};
    for (int q = 0; q < n && q < 10; q++) info[q] = v[q];
    if (build_seconds) *build_seconds = P.build_seconds;
    return FLGPU_OK;
}

int time_assemble(flgpu_ctx* c, const flgpu_advdif_params* p, int reps, double* ms) {
    if (reps < 1) return set_error(FLGPU_EINVAL, "flgpu_time_assemble: reps < 1");
    krylov_invalidate(c);
    c->assembled = false;
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
