// Device helpers shared by the Krylov and learned-preconditioner kernels:
// deterministic block reductions and CSR products.
#pragma once

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdint>
#include <cstdlib>
#include <vector>

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

constexpr int kVecBlock = 512;     // threads per block of the vector kernels (8 waves)
constexpr int kMaxVecGrid = 512;   // their fixed grid (2 blocks per CU)
constexpr int kRowBlock = 256;     // threads per block of the CSR row products
constexpr int kMaxRowGrid = 2048;  // their fixed grid
constexpr int kPartStride = 2048;  // partials[k * kPartStride + block], >= either grid

// Regions of the partial-sum buffer (in units of kPartStride).  A producer
// kernel writes one per-block value per component; the consumer sums them.
enum {
    PART_SPMV = 0,  // 2 components: x.y, y.y of the CSR product
    PART_VEC = 2,   // 4 components: r.z, r.r, z.z, z.w of the CG update
    PART_PC = 6,    // 3 components: r.z, z.z, z.w of the learned apply
    PART_MISC = 9,  // 2 components: b.b, diagonal min/max
    PART_VEC2 = 11, // 4 components: the single-reduction CG's update, alternating with PART_VEC
    PART_COMPONENTS = 15
};

// Sum over a block in a fixed order; result valid in thread 0.
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

// Per-block sums of NV values: thread 0 of each block stores them at
// partials[k * kPartStride + blockIdx.x].
template <int NV>
__device__ __forceinline__ void block_partials(double (&v)[NV], double* __restrict__ partials) {
    __shared__ double sh[kVecBlock / 64][NV];
    block_sum<NV>(v, sh);
    if (threadIdx.x == 0)
#pragma unroll
        for (int k = 0; k < NV; k++) partials[(size_t)k * kPartStride + blockIdx.x] = v[k];
}

// Total of the nparts per-block sums a previous kernel wrote.  Every block
// that calls it forms the same sum in the same order (thread t adds parts
// t, t + blockDim, ..., then the fixed block order), so all blocks agree
// bit for bit and no grid-wide atomics or last-block election are needed.
// Result valid in thread 0.
template <int NV>
__device__ __forceinline__ void sum_partials(const double* __restrict__ partials, int nparts,
                                             double (&out)[NV]) {
    __shared__ double sh[kVecBlock / 64][NV];
#pragma unroll
    for (int k = 0; k < NV; k++) out[k] = 0.0;
    for (int b = threadIdx.x; b < nparts; b += blockDim.x)
#pragma unroll
        for (int k = 0; k < NV; k++) out[k] += partials[(size_t)k * kPartStride + b];
    block_sum<NV>(out, sh);
}

// ---------------------------------------------------------------------
// Streamed CSR products
// ---------------------------------------------------------------------
constexpr int kStreamPerThread = 8;
constexpr int kStreamNnz = kRowBlock * kStreamPerThread;  // entries staged per row block

// Row blocks for k_csr_stream: runs of consecutive rows holding at most
// kStreamNnz entries and at most kRowBlock rows; a longer row is a block of
// its own.  rb has nblocks + 1 entries, rbk the matching first entries.
inline void stream_row_blocks(const int32_t* rowptr, int n, std::vector<int32_t>& rb,
                              std::vector<int32_t>& rbk) {
    rb.clear();
    rb.push_back(0);
    int r = 0;
    while (r < n) {
        const int k0 = rowptr[r];
        int r1 = r;
        while (r1 < n && r1 - r < kRowBlock && rowptr[r1 + 1] - k0 <= kStreamNnz) r1++;
        if (r1 == r) r1 = r + 1;
        rb.push_back(r1);
        r = r1;
    }
    rbk.resize(rb.size());
    for (size_t q = 0; q < rb.size(); q++) rbk[q] = rowptr[rb[q]];
}

enum { DOT_NONE = 0, DOT_XY = 1, DOT_PC = 2 };

// Row blocks handled by this workgroup: with a grid that is a multiple of 8
// each XCD (the hardware hands consecutive workgroups to the 8 XCDs in
// turn) takes a contiguous share of the row blocks, so the vector entries
// and nodes that neighbouring rows gather stay in one L2.
struct RowBlocks {
    int b, hi, step;
    __device__ __forceinline__ explicit RowBlocks(int nrb) {
        b = blockIdx.x;
        hi = nrb;
        step = gridDim.x;
        if ((gridDim.x & 7) == 0) {
            const int xcd = blockIdx.x & 7;
            b = (int)((int64_t)nrb * xcd / 8) + (int)(blockIdx.x >> 3);
            hi = (int)((int64_t)nrb * (xcd + 1) / 8);
            step = gridDim.x >> 3;
        }
    }
};

// A CSR operand of k_csr_stream.  Column indices are either int32 or, when
// every entry of a row block lies within int16 range of the block's first
// row, int16 offsets from that row (half the index bytes).  Optional
// scales: y_i = rs_i * sum_k val_k (cs_j x_j) -- a diagonally scaled matrix
// stored unscaled (low-precision values of order one).
template <typename AV, typename CT>
struct CsrOp {
    int nrb;
    const int32_t* rb;     // nrb+1 first rows of the row blocks
    const int32_t* rbk;    // nrb+1 first entries of the row blocks (rowptr[rb[b]])
    const int32_t* rowptr;
    const CT* col;
    const AV* val;
    const float* cs;
    const float* rs;
};

template <typename CT>
__device__ __forceinline__ int csr_col(const CT* __restrict__ col, int k, int r0) {
    if constexpr (sizeof(CT) == 2)
        return r0 + (int)col[k];
    else
        return (int)col[k];
}

// y = A x (or y = bsub - A x) over row blocks.  A block reads its entries
// once, coalesced, stages the products val*x[col] in LDS and one thread per
// row adds them in entry order: the sequential row sum of a plain CSR
// product, with the product rounded before the addition.  A fixed grid
// walks the row blocks (RowBlocks).
// Per-block dot products into partials: DOT_XY: x_i y_i, y_i^2;
// DOT_PC: r_i y_i, y_i^2, y_i w_i.
template <typename V, typename AV, typename CT, typename XV, typename YV, int DOTS>
__global__ void __launch_bounds__(kRowBlock)
k_csr_stream(CsrOp<AV, CT> m, const XV* __restrict__ x, YV* __restrict__ y, const double* __restrict__ bsub,
             const V* __restrict__ dr, const V* __restrict__ dw, double* __restrict__ partials,
             const int* __restrict__ done_flag) {
    if (done_flag && *done_flag) return;
    __shared__ V prod[kStreamNnz];
    double d0 = 0.0, d1 = 0.0, d2 = 0.0;
    const int32_t* __restrict__ rowptr = m.rowptr;
    const CT* __restrict__ col = m.col;
    const AV* __restrict__ val = m.val;
    const float* __restrict__ cs = m.cs;
    const RowBlocks it(m.nrb);
    // The bounds and column indices of the next row block are loaded while
    // the current block's values and x entries are gathered, so a row block
    // costs one round trip to memory instead of three.
    int r0 = 0, r1 = 0, k0 = 0, k1 = 0, c[kStreamPerThread];
    auto fetch = [&](int bb, int& a0, int& a1, int& e0, int& e1, int (&cc)[kStreamPerThread]) {
        a0 = m.rb[bb];
        a1 = m.rb[bb + 1];
        e0 = m.rbk[bb];
        e1 = m.rbk[bb + 1];
#pragma unroll
        for (int j = 0; j < kStreamPerThread; j++) {
            const int k = e0 + (int)threadIdx.x + j * kRowBlock;
            cc[j] = (k < e1 && e1 - e0 <= kStreamNnz) ? csr_col(col, k, a0) : 0;
        }
    };
    if (it.b < it.hi) fetch(it.b, r0, r1, k0, k1, c);
    for (int b = it.b; b < it.hi; b += it.step) {
        int n0 = 0, n1 = 0, e0 = 0, e1 = 0, nc[kStreamPerThread];
        if (b + it.step < it.hi) fetch(b + it.step, n0, n1, e0, e1, nc);
        const int row = r0 + threadIdx.x;
        V acc = (V)0;
        if (k1 - k0 <= kStreamNnz) {
            int kb = 0, ke = 0;
            if (row < r1) {
                kb = rowptr[row] - k0;
                ke = rowptr[row + 1] - k0;
            }
            V pv[kStreamPerThread];
#pragma unroll
            for (int j = 0; j < kStreamPerThread; j++) {
                const int k = k0 + (int)threadIdx.x + j * kRowBlock;
                pv[j] = (V)0;
                if (k < k1) {
                    V xv = (V)x[c[j]];
                    if (cs) xv = xv * (V)cs[c[j]];
                    pv[j] = (V)val[k] * xv;
                }
            }
#pragma unroll
            for (int j = 0; j < kStreamPerThread; j++) {
                const int k = k0 + (int)threadIdx.x + j * kRowBlock;
                if (k < k1) prod[k - k0] = pv[j];
            }
            __syncthreads();
            for (int k = kb; k < ke; k++) acc += prod[k];
            __syncthreads();
        } else {  // one row longer than the staging buffer: in order, chunk by chunk
            for (int c0 = k0; c0 < k1; c0 += kStreamNnz) {
                const int c1 = min(k1, c0 + kStreamNnz);
                for (int k = c0 + threadIdx.x; k < c1; k += kRowBlock) {
                    const int cj = csr_col(col, k, r0);
                    V xv = (V)x[cj];
                    if (cs) xv = xv * (V)cs[cj];
                    prod[k - c0] = (V)val[k] * xv;
                }
                __syncthreads();
                if (threadIdx.x == 0)
                    for (int k = 0; k < c1 - c0; k++) acc += prod[k];
                __syncthreads();
            }
        }
        if (row < r1) {
            if (m.rs) acc = acc * (V)m.rs[row];
            double yv = (double)acc;
            if (bsub) yv = bsub[row] - yv;
            const YV yo = (YV)yv;
            y[row] = yo;
            if (DOTS == DOT_XY) {
                d0 += (double)x[row] * yv;
                d1 += yv * yv;
            } else if (DOTS == DOT_PC) {
                const double z = (double)yo;
                d0 += (double)dr[row] * z;
                d1 += z * z;
                d2 += z * (double)dw[row];
            }
        }
        r0 = n0;
        r1 = n1;
        k0 = e0;
        k1 = e1;
#pragma unroll
        for (int j = 0; j < kStreamPerThread; j++) c[j] = nc[j];
    }
    if (DOTS == DOT_XY) {
        double v[2] = {d0, d1};
        block_partials<2>(v, partials);
    } else if (DOTS == DOT_PC) {
        double v[3] = {d0, d1, d2};
        block_partials<3>(v, partials);
    }
}

// ---------------------------------------------------------------------
// Sliced ELL products
// ---------------------------------------------------------------------
// A slice is 64 consecutive rows, one wave: entry j of the slice's rows
// sits at slots sptr[s] + 64 j + lane, so every load of values and column
// offsets is one contiguous line per wave, with no row bounds, no LDS and
// no barrier.  A slice is as wide as its longest row; the slots past a
// row's end are padding (value 0, column the row itself) and are skipped.
// Meshes numbered in space-filling order give slices of near-equal rows,
// so the padding is small; sell_plan reports it and the caller keeps the
// CSR stream product where it is not.
constexpr int kSlice = 64;

template <typename AV, typename CT>
struct SellOp {
    int nsl, n;
    const int32_t* sptr;    // nsl+1 first slots (multiples of 64)
    const int32_t* rowptr;  // row lengths from the CSR row pointers
    const CT* col;          // column, or int16 offset from the slice's first row
    const AV* val;
    const float* rs;        // optional row scale: y_i = rs_i (A x)_i
};

// y = A x (or y = bsub - A x) over slices; each lane adds its row's
// products in entry order, each product rounded before the addition (no
// contraction), the same sums as k_csr_stream.  Dot products as there.
// Slices go to waves in groups of kRowBlock/64 through RowBlocks, so each
// XCD takes a contiguous share.
template <typename V, typename AV, typename CT, typename XV, typename YV, int DOTS>
__global__ void __launch_bounds__(kRowBlock)
k_sell(SellOp<AV, CT> m, const XV* __restrict__ x, YV* __restrict__ y, const double* __restrict__ bsub,
       const V* __restrict__ dr, const V* __restrict__ dw, double* __restrict__ partials,
       const int* __restrict__ done_flag) {
#pragma clang fp contract(off)
    if (done_flag && *done_flag) return;
    constexpr int WPB = kRowBlock / kSlice;
    constexpr int U = 8;
    const int wave = threadIdx.x / kSlice, lane = threadIdx.x % kSlice;
    const CT* __restrict__ col = m.col;
    const AV* __restrict__ val = m.val;
    double d0 = 0.0, d1 = 0.0, d2 = 0.0;
    const RowBlocks it((m.nsl + WPB - 1) / WPB);
    for (int g = it.b; g < it.hi; g += it.step) {
        const int sl = g * WPB + wave;
        if (sl >= m.nsl) continue;
        const int r0 = sl * kSlice, row = r0 + lane;
        const int base = m.sptr[sl], w = (m.sptr[sl + 1] - base) / kSlice;
        const int len = row < m.n ? m.rowptr[row + 1] - m.rowptr[row] : 0;
        V acc = (V)0;
        for (int j0 = 0; j0 < w; j0 += U) {
            // Every load is unconditional (a step past the slice's width
            // re-reads its last slot, padding holds an in-range column), so
            // the U loads of a step are all in flight together; only the
            // additions are predicated on the row's length.
            int cj[U];
            AV av[U];
#pragma unroll
            for (int u = 0; u < U; u++) {
                const int slot = base + min(j0 + u, w - 1) * kSlice + lane;
                cj[u] = csr_col(col, slot, r0);
                av[u] = val[slot];
            }
            XV xv[U];
#pragma unroll
            for (int u = 0; u < U; u++) xv[u] = x[cj[u]];
#pragma unroll
            for (int u = 0; u < U; u++) {
                const V pv = (V)av[u] * (V)xv[u];
                if (j0 + u < len) acc += pv;
            }
        }
        if (row < m.n) {
            if (m.rs) acc = acc * (V)m.rs[row];
            double yv = (double)acc;
            if (bsub) yv = bsub[row] - yv;
            const YV yo = (YV)yv;
            y[row] = yo;
            if (DOTS == DOT_XY) {
                d0 += (double)x[row] * yv;
                d1 += yv * yv;
            } else if (DOTS == DOT_PC) {
                const double z = (double)yo;
                d0 += (double)dr[row] * z;
                d1 += z * z;
                d2 += z * (double)dw[row];
            }
        }
    }
    if (DOTS == DOT_XY) {
        double v[2] = {d0, d1};
        block_partials<2>(v, partials);
    } else if (DOTS == DOT_PC) {
        double v[3] = {d0, d1, d2};
        block_partials<3>(v, partials);
    }
}

// Host-side plan of a sliced ELL operand: slot offsets, column indices
// (int16 offsets from the slice's first row when they all fit) and, per
// slot, the entry of the source value array it takes (vmap[k] for CSR
// entry k, or k itself; -1 for padding).
struct SellHost {
    int nsl = 0;
    int64_t nslot = 0;
    bool use16 = false;
    std::vector<int32_t> sptr, src, c32;
    std::vector<int16_t> c16;
    double fill(int64_t nnz) const { return nnz > 0 ? (double)nslot / (double)nnz : 1.0; }
};

inline void sell_plan(const int32_t* rowptr, const int32_t* col, int n, const int32_t* vmap, SellHost& h) {
    h.nsl = (n + kSlice - 1) / kSlice;
    h.sptr.assign((size_t)h.nsl + 1, 0);
    for (int sl = 0; sl < h.nsl; sl++) {
        int w = 0;
        for (int r = sl * kSlice; r < std::min(n, (sl + 1) * kSlice); r++) w = std::max(w, rowptr[r + 1] - rowptr[r]);
        h.sptr[sl + 1] = h.sptr[sl] + w * kSlice;
    }
    h.nslot = h.sptr[h.nsl];
    h.src.assign((size_t)h.nslot, -1);
    h.c32.assign((size_t)h.nslot, 0);
    bool fits = true;
    for (int sl = 0; sl < h.nsl; sl++) {
        const int r0 = sl * kSlice, w = (h.sptr[sl + 1] - h.sptr[sl]) / kSlice;
        for (int lane = 0; lane < kSlice; lane++) {
            const int r = r0 + lane;
            const int len = r < n ? rowptr[r + 1] - rowptr[r] : 0;
            for (int j = 0; j < w; j++) {
                const size_t slot = (size_t)h.sptr[sl] + (size_t)j * kSlice + lane;
                if (j < len) {
                    const int k = rowptr[r] + j;
                    h.src[slot] = vmap ? vmap[k] : k;
                    h.c32[slot] = col[k];
                } else {
                    h.c32[slot] = r < n ? r : r0;
                }
                const int d = h.c32[slot] - r0;
                if (d < -32768 || d > 32767) fits = false;
            }
        }
    }
    h.use16 = fits;
    if (fits) {
        h.c16.resize((size_t)h.nslot);
        for (int sl = 0; sl < h.nsl; sl++)
            for (int64_t q = h.sptr[sl]; q < h.sptr[sl + 1]; q++) h.c16[q] = (int16_t)(h.c32[q] - sl * kSlice);
        std::vector<int32_t>().swap(h.c32);
    }
}

// out[t] = src[t] >= 0 ? in[src[t]] : 0 (a value array in slot order)
template <typename SV, typename DV>
__global__ void __launch_bounds__(256)
k_sell_gather(int64_t m, const int32_t* __restrict__ src, const SV* __restrict__ in, DV* __restrict__ out) {
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t < m) {
        const int32_t k = src[t];
        out[t] = k >= 0 ? (DV)in[k] : (DV)0;
    }
}

// Fixed grid for k_sell: the slice groups through stream_grid.
inline int sell_grid(int nsl);

// int16 column offsets from each row block's first row; false (and out
// untouched) when an offset does not fit.
inline bool compress_columns(const int32_t* rowptr, const int32_t* col, const std::vector<int32_t>& rb,
                             std::vector<int16_t>& out) {
    const int nrb = (int)rb.size() - 1;
    std::vector<int16_t> c16((size_t)rowptr[rb[nrb]]);
    for (int b = 0; b < nrb; b++) {
        const int r0 = rb[b];
        for (int k = rowptr[r0]; k < rowptr[rb[b + 1]]; k++) {
            const int d = col[k] - r0;
            if (d < -32768 || d > 32767) return false;
            c16[k] = (int16_t)d;
        }
    }
    out.swap(c16);
    return true;
}

// Fixed grid for k_csr_stream: a multiple of 8, at most kMaxRowGrid.
inline int stream_grid(int nrb) {
    static const int cap = [] {
        const char* e = getenv("FLGPU_STREAM_GRID");
        const int v = e ? atoi(e) : kMaxRowGrid;
        return v >= 8 && v <= kMaxRowGrid ? v : kMaxRowGrid;
    }();
    if (nrb < 8) return nrb > 0 ? nrb : 1;
    const int g = nrb < cap ? nrb : cap;
    return g & ~7;
}

inline int sell_grid(int nsl) { return stream_grid((nsl + kRowBlock / kSlice - 1) / (kRowBlock / kSlice)); }

}  // namespace flgpu
