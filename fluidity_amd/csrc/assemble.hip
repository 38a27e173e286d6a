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
#include <cmath>
#include <cstdlib>
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

// ---------------------------------------------------------------------
// Blocked scheme (linear T, FP64).  Rows are grouped into small blocks
// that are compact in space (Morton order of the node positions); one
// workgroup assembles one block's CSR rows:
//   phase A: every element touching the block computes the rows of its
//     element matrix that fall in the block and stores each row in an LDS
//     slot numbered (row, rank of the element among the row's elements in
//     colour order) -- the ranks are fixed by the colouring and computed
//     once per mesh;
//   phase B: one thread per CSR entry adds the entry's contributions slot
//     by slot, i.e. in colour order, ascending element id inside a colour
//     -- the order of the reference's colour loop (:574-598) and of the
//     colour scheme above, so the two schemes give the same bits.
// Each row is written to HBM once and the matrix needs no zeroing pass.
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

// Exclusive prefix sums of two per-row counts over the block's rows.
// Each thread owns a contiguous chunk of rows; returns the thread's
// starting offsets and leaves the totals in tot[0..1].
__device__ __forceinline__ void block_scan2(int a, int b, int* wsa, int* wsb, int& oa, int& ob, int* tot) {
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6, nw = blockDim.x >> 6;
    int ia = a, ib = b;
    for (int off = 1; off < 64; off <<= 1) {
        const int va = __shfl_up(ia, off, 64), vb = __shfl_up(ib, off, 64);
        if (lane >= off) {
            ia += va;
            ib += vb;
        }
    }
    if (lane == 63) {
        wsa[wave] = ia;
        wsb[wave] = ib;
    }
    __syncthreads();
    if (tid == 0) {
        int ra = 0, rb = 0;
        for (int w = 0; w < nw; w++) {
            const int ta = wsa[w], tb = wsb[w];
            wsa[w] = ra;
            wsb[w] = rb;
            ra += ta;
            rb += tb;
        }
        tot[0] = ra;
        tot[1] = rb;
    }
    __syncthreads();
    oa = wsa[wave] + ia - a;
    ob = wsb[wave] + ib - b;
}

template <int DIM>
__global__ void __launch_bounds__(256)
k_asm_blocked(const ElemTables* __restrict__ tab, int nnz_cap, int slot_cap, const int32_t* __restrict__ brow,
              const int32_t* __restrict__ brows, const int32_t* __restrict__ bvis, const int32_t* __restrict__ visits,
              const int32_t* __restrict__ node_block, const int32_t* __restrict__ node_lrow,
              const uint32_t* __restrict__ rank4, const int32_t* __restrict__ deg,
              const int32_t* __restrict__ conn, const int32_t* __restrict__ conn_x, const double* __restrict__ X4,
              AsmFields fd, AsmFlags fl, const int32_t* __restrict__ rowptr, const int32_t* __restrict__ col,
              double* __restrict__ A, double* __restrict__ rhs, int* __restrict__ err) {
    constexpr int LOC = DIM + 1;
    __shared__ LTab<double> tb;
    __shared__ int wsa[256 / 64], wsb[256 / 64], tot[2];
    extern __shared__ double dyn[];
    double* sval = dyn;                                            // slot_cap * LOC
    double* srhs = sval + (size_t)slot_cap * LOC;                  // slot_cap
    uint32_t* spos = reinterpret_cast<uint32_t*>(srhs + slot_cap); // slot_cap
    int32_t* lcol = reinterpret_cast<int32_t*>(spos + slot_cap);   // nnz_cap
    int32_t* loff = lcol + nnz_cap;                                // R + 1
    stage_tables<double>(tab, tb);

    const int b = blockIdx.x, tid = threadIdx.x, nt = blockDim.x;
    const int r0 = brow[b], R = brow[b + 1] - r0;
    int32_t* soff = loff + R + 1;                                  // R + 1
    const int per = (R + nt - 1) / nt;
    const int l0 = min(tid * per, R), l1 = min(l0 + per, R);
    int na = 0, ns = 0;
    for (int l = l0; l < l1; l++) {
        const int row = brows[r0 + l];
        na += rowptr[row + 1] - rowptr[row];
        ns += deg[row];
    }
    int oa, os;
    block_scan2(na, ns, wsa, wsb, oa, os, tot);
    for (int l = l0; l < l1; l++) {
        const int row = brows[r0 + l];
        const int s = rowptr[row], len = rowptr[row + 1] - s;
        loff[l] = oa;
        soff[l] = os;
        for (int k = 0; k < len; k++) lcol[oa + k] = col[s + k];
        oa += len;
        os += deg[row];
    }
    if (tid == 0) {
        loff[R] = tot[0];
        soff[R] = tot[1];
    }
    __syncthreads();

    // phase A: element rows into slots
    for (int v = bvis[b] + tid; v < bvis[b + 1]; v += nt) {
        const int ele = visits[v];
        Elem<DIM, LOC, double> e;
        gather_elem<DIM, LOC, double>(ele, conn, conn_x, X4, fd.T, e);
        GeomLin<DIM, LOC, double> gm;
        geom_linear<DIM, LOC, double>(e, tb, gm);
        const uint32_t rk = rank4[ele];
#pragma unroll
        for (int i = 0; i < LOC; i++) {
            const int node = e.nodes[i];
            if (node_block[node] != b) continue;
            const int l = node_lrow[node];
            const int slot = soff[l] + (int)((rk >> (8 * i)) & 255u);
            double vals[LOC], r;
            row_linear_g<DIM, LOC, double>(e, gm, tb, fl, fd, i, vals, r);
            const int lo = loff[l], hi = loff[l + 1];
            uint32_t packed = 0;
#pragma unroll
            for (int j = 0; j < LOC; j++) {
                const int q = lds_pos(lcol, lo, hi, e.nodes[j]);
                if (q < 0) atomicOr(err, 1);
                packed |= (uint32_t)(q < 0 ? 255 : q - lo) << (8 * j);
                sval[(size_t)slot * LOC + j] = vals[j];
            }
            spos[slot] = packed;
            srhs[slot] = r;
        }
    }
    __syncthreads();

    // phase B: one thread per entry, contributions in slot (= colour) order
    const int nent = loff[R];
    for (int t = tid; t < nent; t += nt) {
        int lo = 0, hi = R - 1;  // row of entry t: loff[l] <= t < loff[l+1]
        while (lo < hi) {
            const int mid = (lo + hi + 1) >> 1;
            if (loff[mid] <= t) lo = mid;
            else hi = mid - 1;
        }
        const int l = lo, p = t - loff[l];
        double acc = 0.0;
        for (int k = soff[l]; k < soff[l + 1]; k++) {
            const uint32_t pk = spos[k];
#pragma unroll
            for (int j = 0; j < LOC; j++) {
                if ((int)((pk >> (8 * j)) & 255u) == p) {
                    const double vv = sval[(size_t)k * LOC + j];
                    if (vv != 0.0) acc = acc + vv;  // csr_addto skips zeros
                }
            }
        }
        A[rowptr[brows[r0 + l]] + p] = acc;
    }
    for (int l = tid; l < R; l += nt) {
        double acc = 0.0;
        for (int k = soff[l]; k < soff[l + 1]; k++) acc = acc + srhs[k];
        rhs[brows[r0 + l]] = acc;
    }
}

static uint32_t spread_bits(uint32_t v, int dim) {
    uint32_t r = 0;
    for (int k = 0; k < 10; k++) r |= ((v >> k) & 1u) << (k * dim);
    return r;
}

int build_block_plan(flgpu_ctx* c) {
    if (c->plan.built) return FLGPU_OK;
    if (!c->linear) return set_error(FLGPU_EUNSUPPORTED, "block plan: linear field mesh only");
    const int N = c->rows, dim = c->dim, loc = c->loc, ncol = c->ncol;
    int rows_cap = 64;
    if (const char* s = getenv("FLGPU_BLOCK_ROWS")) rows_cap = std::max(16, std::min(1024, atoi(s)));
    // node positions of the field mesh (vertex i of T <-> vertex i of X)
    std::vector<float> pos((size_t)N * dim, 0.f);
    {
        std::vector<double> X4h((size_t)c->nx * 4);
        FLGPU_HIP(hipMemcpy(X4h.data(), c->X4.p, sizeof(double) * X4h.size(), hipMemcpyDeviceToHost));
        for (int64_t e = 0; e < c->nele; e++)
            for (int i = 0; i < loc; i++) {
                const int nd = c->h_conn[e * loc + i], xn = c->h_conn_x[e * c->loc_x + i];
                for (int d = 0; d < dim; d++) pos[(size_t)nd * dim + d] = (float)X4h[(size_t)xn * 4 + d];
            }
    }
    float lo[3] = {1e30f, 1e30f, 1e30f}, hi[3] = {-1e30f, -1e30f, -1e30f};
    for (int i = 0; i < N; i++)
        for (int d = 0; d < dim; d++) {
            lo[d] = std::min(lo[d], pos[(size_t)i * dim + d]);
            hi[d] = std::max(hi[d], pos[(size_t)i * dim + d]);
        }
    std::vector<uint64_t> key(N);
    for (int i = 0; i < N; i++) {
        uint32_t code = 0;
        for (int d = 0; d < dim; d++) {
            const float span = hi[d] > lo[d] ? hi[d] - lo[d] : 1.f;
            uint32_t q = (uint32_t)std::min(1023.f, std::max(0.f, (pos[(size_t)i * dim + d] - lo[d]) / span * 1023.f));
            code |= spread_bits(q, dim) << d;
        }
        key[i] = ((uint64_t)code << 32) | (uint32_t)i;
    }
    std::sort(key.begin(), key.end());
    // node -> element incidence in colour order (rank = position in it)
    std::vector<int32_t> rank(c->nele);
    for (int k = 0; k < ncol; k++)
        for (int q = c->h_col_ptr[k]; q < c->h_col_ptr[k + 1]; q++) rank[c->h_col_elems[q]] = q;
    std::vector<int64_t> iptr((size_t)N + 1, 0);
    for (int64_t k = 0; k < (int64_t)c->nele * loc; k++) iptr[c->h_conn[k] + 1]++;
    for (int i = 0; i < N; i++) iptr[i + 1] += iptr[i];
    std::vector<int32_t> deg(N);
    for (int i = 0; i < N; i++) {
        const int64_t d = iptr[i + 1] - iptr[i];
        if (d > 255) return set_error(FLGPU_EUNSUPPORTED, "block plan: node %d belongs to %lld elements (max 255)",
                                      i + 1, (long long)d);
        deg[i] = (int32_t)d;
    }
    std::vector<int32_t> inc(iptr[N]);
    {
        std::vector<int64_t> fill(iptr.begin(), iptr.end() - 1);
        for (int q = 0; q < c->nele; q++) {  // colour order: elements visited by rank
            const int e = c->h_col_elems[q];
            for (int i = 0; i < loc; i++) inc[fill[c->h_conn[(int64_t)e * loc + i]]++] = e;
        }
    }
    std::vector<uint32_t> rank4(c->nele, 0);
    for (int n = 0; n < N; n++)
        for (int64_t q = iptr[n]; q < iptr[n + 1]; q++) {
            const int e = inc[q];
            for (int i = 0; i < loc; i++)
                if (c->h_conn[(int64_t)e * loc + i] == n) rank4[e] |= (uint32_t)(q - iptr[n]) << (8 * i);
        }
    // blocks of consecutive Morton rows, bounded by rows, entries and slots
    std::vector<int32_t> brow(1, 0), brows(N), node_block(N), node_lrow(N);
    int cur = 0, cur_nnz = 0, cur_slots = 0, max_nnz = 0, max_slots = 0, max_rows = 0;
    const int nnz_cap = rows_cap * 24, slot_cap = rows_cap * 32;
    for (int k = 0; k < N; k++) {
        const int r = (int)(key[k] & 0xffffffffu);
        const int len = c->h_rowptr[r + 1] - c->h_rowptr[r];
        if (len > 255) return set_error(FLGPU_EUNSUPPORTED, "block plan: row %d has %d entries (max 255)", r + 1, len);
        if (cur > 0 && (cur == rows_cap || cur_nnz + len > nnz_cap || cur_slots + deg[r] > slot_cap)) {
            brow.push_back(k);
            max_nnz = std::max(max_nnz, cur_nnz);
            max_slots = std::max(max_slots, cur_slots);
            max_rows = std::max(max_rows, cur);
            cur = cur_nnz = cur_slots = 0;
        }
        brows[k] = r;
        node_block[r] = (int)brow.size() - 1;
        node_lrow[r] = cur;
        cur++;
        cur_nnz += len;
        cur_slots += deg[r];
    }
    brow.push_back(N);
    max_nnz = std::max(max_nnz, cur_nnz);
    max_slots = std::max(max_slots, cur_slots);
    max_rows = std::max(max_rows, cur);
    const int nb = (int)brow.size() - 1;
    // visits: elements touching each block, ascending ids
    std::vector<int32_t> stamp(c->nele, -1), bvis(1, 0), visits, tmp;
    visits.reserve((size_t)c->nele * 2);
    for (int b = 0; b < nb; b++) {
        tmp.clear();
        for (int k = brow[b]; k < brow[b + 1]; k++) {
            const int r = brows[k];
            for (int64_t q = iptr[r]; q < iptr[r + 1]; q++) {
                const int e = inc[q];
                if (stamp[e] != b) {
                    stamp[e] = b;
                    tmp.push_back(e);
                }
            }
        }
        std::sort(tmp.begin(), tmp.end());
        visits.insert(visits.end(), tmp.begin(), tmp.end());
        bvis.push_back((int32_t)visits.size());
    }
    BlockPlan& P = c->plan;
    FLGPU_TRY(P.brow.upload(brow.data(), brow.size(), c->stream));
    FLGPU_TRY(P.brows.upload(brows.data(), brows.size(), c->stream));
    FLGPU_TRY(P.node_block.upload(node_block.data(), node_block.size(), c->stream));
    FLGPU_TRY(P.node_lrow.upload(node_lrow.data(), node_lrow.size(), c->stream));
    FLGPU_TRY(P.seg.upload(bvis.data(), bvis.size(), c->stream));
    FLGPU_TRY(P.visits.upload(visits.data(), visits.size(), c->stream));
    FLGPU_TRY(P.rank4.upload(rank4.data(), rank4.size(), c->stream));
    FLGPU_TRY(P.deg.upload(deg.data(), deg.size(), c->stream));
    FLGPU_HIP(hipStreamSynchronize(c->stream));
    P.nblocks = nb;
    P.rows_cap = max_rows;
    P.nnz_cap = max_nnz;
    P.slot_cap = max_slots;
    P.nvisits = (int64_t)visits.size();
    P.built = true;
    return FLGPU_OK;
}

template <int DIM>
static int launch_blocked(flgpu_ctx* c, const AsmFields& fd, const AsmFlags& fl, int* d_err) {
    const BlockPlan& P = c->plan;
    constexpr int LOC = DIM + 1;
    const size_t lds = sizeof(double) * (size_t)P.slot_cap * (LOC + 1) + sizeof(uint32_t) * P.slot_cap +
                       sizeof(int32_t) * ((size_t)P.nnz_cap + 2 * (P.rows_cap + 1));
    if (lds + sizeof(LTab<double>) + 64 > 160 * 1024)
        return set_error(FLGPU_EUNSUPPORTED, "blocked assembly: a block needs %zu bytes of LDS", lds);
    FLGPU_HIP(hipFuncSetAttribute(reinterpret_cast<const void*>(&k_asm_blocked<DIM>),
                                  hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
    hipLaunchKernelGGL((k_asm_blocked<DIM>), dim3(P.nblocks), dim3(256), lds, c->stream, c->d_tab.p, P.nnz_cap,
                       P.slot_cap, P.brow.p, P.brows.p, P.seg.p, P.visits.p, P.node_block.p, P.node_lrow.p,
                       P.rank4.p, P.deg.p, c->conn.p, c->conn_x.p, c->X4.p, fd, fl, c->rowptr.p, c->col.p, c->A.p,
                       c->rhs.p, d_err);
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

    const int dim = c->dim, loc = c->loc;
    const bool fp32 = p->precision == FLGPU_FP32;
    int rc = FLGPU_OK;
    int scheme = p->scheme;
    const bool blocked_ok = c->linear && !fp32;
    if (scheme == FLGPU_ASM_BLOCKED && !blocked_ok)
        return set_error(FLGPU_EUNSUPPORTED, "flgpu assemble: the blocked scheme covers linear T in FP64");
    if (scheme == FLGPU_ASM_AUTO) scheme = blocked_ok ? FLGPU_ASM_BLOCKED : FLGPU_ASM_COLOUR;
    if (scheme != FLGPU_ASM_COLOUR && scheme != FLGPU_ASM_BLOCKED)
        return set_error(FLGPU_EINVAL, "flgpu assemble: scheme %d", p->scheme);

    FLGPU_HIP(hipMemsetAsync(c->inactive.p, 0, sizeof(int32_t) * c->rows, c->stream));
    c->have_inactive = false;
    if (scheme == FLGPU_ASM_BLOCKED) {
        FLGPU_TRY(build_block_plan(c));
        return dim == 3 ? launch_blocked<3>(c, fd, fl, d_err) : launch_blocked<2>(c, fd, fl, d_err);
    }
    FLGPU_HIP(hipMemsetAsync(c->A.p, 0, sizeof(double) * c->nnz, c->stream));
    FLGPU_HIP(hipMemsetAsync(c->rhs.p, 0, sizeof(double) * c->rows, c->stream));
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
    krylov_invalidate(c);
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
