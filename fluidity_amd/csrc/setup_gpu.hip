// Mesh set-up on the device, for meshes that change (every adapt rebuilds
// them in Fluidity):
//   flgpu_sparsity_build   make_sparsity of the bound field mesh
//                          (femtools/Sparsity_Patterns.F90:47-85, :299-428):
//                          row i = ascending union of the nodes of the
//                          elements containing i; the same arrays as the
//                          host routine, bit for bit;
//   flgpu_colouring_build  an element colouring (elements sharing a node get
//                          different colours; Colouring.F90:85-199 colours
//                          greedily in index order, which is sequential:
//                          here speculative first-fit rounds with hashed
//                          priorities, deterministic).
// Both then bind their result through the host entry points (which also
// keep the host copies the other set-up code reads).
#include <algorithm>
#include <climits>
#include <vector>

#include "flgpu_internal.h"

namespace flgpu {
namespace {

constexpr int kScanThreads = 256, kScanPer = 8, kScanTile = kScanThreads * kScanPer;

// Exclusive scan, tile by tile: each workgroup scans kScanTile values and
// leaves its total; the totals are scanned the same way; then added back.
__global__ void __launch_bounds__(kScanThreads)
k_scan_tiles(int64_t n, const int32_t* __restrict__ in, int64_t* __restrict__ out, int64_t* __restrict__ tsum) {
    __shared__ int64_t part[kScanThreads];
    const int64_t base = (int64_t)blockIdx.x * kScanTile + (int64_t)threadIdx.x * kScanPer;
    int64_t v[kScanPer], s = 0;
#pragma unroll
    for (int k = 0; k < kScanPer; k++) {
        v[k] = base + k < n ? in[base + k] : 0;
        s += v[k];
    }
    part[threadIdx.x] = s;
    __syncthreads();
    for (int off = 1; off < kScanThreads; off <<= 1) {  // inclusive Hillis-Steele over the thread sums
        const int64_t t = threadIdx.x >= off ? part[threadIdx.x - off] : 0;
        __syncthreads();
        part[threadIdx.x] += t;
        __syncthreads();
    }
    int64_t run = part[threadIdx.x] - s;
#pragma unroll
    for (int k = 0; k < kScanPer; k++) {
        if (base + k < n) out[base + k] = run;
        run += v[k];
    }
    if (threadIdx.x == kScanThreads - 1) tsum[blockIdx.x] = part[threadIdx.x];
}

__global__ void __launch_bounds__(kScanThreads)
k_scan_add(int64_t n, int64_t* __restrict__ out, const int64_t* __restrict__ toff) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) out[i] += toff[i / kScanTile];
}

}  // namespace

// out[0..n) = exclusive prefix sums of in[0..n), out[n] = total (int64).
int exclusive_scan(hipStream_t s, const int32_t* in, int64_t n, DevBuf<int64_t>& out) {
    FLGPU_TRY(out.alloc((size_t)n + 1));
    const int64_t tiles = (n + kScanTile - 1) / kScanTile;
    DevBuf<int64_t> tsum;
    FLGPU_TRY(tsum.alloc((size_t)tiles));
    hipLaunchKernelGGL(k_scan_tiles, dim3((unsigned)tiles), dim3(kScanThreads), 0, s, n, in, out.p, tsum.p);
    FLGPU_HIP(hipGetLastError());
    // the tile totals: few (n / 2048), scanned on the host
    std::vector<int64_t> h(tiles);
    FLGPU_HIP(hipMemcpyAsync(h.data(), tsum.p, sizeof(int64_t) * tiles, hipMemcpyDeviceToHost, s));
    FLGPU_HIP(hipStreamSynchronize(s));
    int64_t run = 0;
    for (int64_t t = 0; t < tiles; t++) {
        const int64_t v = h[t];
        h[t] = run;
        run += v;
    }
    FLGPU_HIP(hipMemcpyAsync(tsum.p, h.data(), sizeof(int64_t) * tiles, hipMemcpyHostToDevice, s));
    hipLaunchKernelGGL(k_scan_add, dim3(grid_for(n, kScanThreads)), dim3(kScanThreads), 0, s, n, out.p, tsum.p);
    FLGPU_HIP(hipGetLastError());
    FLGPU_HIP(hipMemcpyAsync(out.p + n, &run, sizeof(int64_t), hipMemcpyHostToDevice, s));
    FLGPU_HIP(hipStreamSynchronize(s));  // h and run go out of scope
    return FLGPU_OK;
}

namespace {

// node -> element incidence
__global__ void __launch_bounds__(256)
k_inc_count(int64_t n, const int32_t* __restrict__ conn, int32_t* __restrict__ cnt) {
    const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k < n) atomicAdd(&cnt[conn[k]], 1);
}

__global__ void __launch_bounds__(256)
k_inc_fill(int64_t n, int loc, const int32_t* __restrict__ conn, const int64_t* __restrict__ iptr,
           int32_t* __restrict__ cursor, int32_t* __restrict__ inc) {
    const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    const int node = conn[k];
    inc[iptr[node] + atomicAdd(&cursor[node], 1)] = (int32_t)(k / loc);
}

__global__ void __launch_bounds__(256)
k_max_count(int n, const int32_t* __restrict__ cnt, int* __restrict__ mx) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) atomicMax(mx, cnt[i]);
}

// One wavefront per row: the row's candidate columns (every node of every
// element containing it) bitonic-sorted in LDS, duplicates dropped.
// MODE 0 counts the unique columns; MODE 1 writes them (0-based) at the
// row's offset and the diagonal's position.  CAP (a power of two) bounds the
// candidates of a row; the host picks it from the largest node degree.
template <int MODE, int CAP>
__global__ void __launch_bounds__(256)
k_rows(int nrows, int loc, const int32_t* __restrict__ conn, const int64_t* __restrict__ iptr,
       const int32_t* __restrict__ inc, int32_t* __restrict__ rowcnt, const int64_t* __restrict__ rowptr,
       int32_t* __restrict__ col, int32_t* __restrict__ diag) {
    constexpr int W = 64, WAVES = 4;
    __shared__ int32_t buf[WAVES][CAP];
    const int wave = threadIdx.x / W, lane = threadIdx.x % W;
    const int row = blockIdx.x * WAVES + wave;
    const bool valid = row < nrows;
    const int64_t i0 = valid ? iptr[row] : 0;
    const int ncand = valid ? (int)(iptr[row + 1] - i0) * loc : 0;  // <= CAP by the host's choice
    for (int k = lane; k < CAP; k += W)
        buf[wave][k] = k < ncand ? conn[(int64_t)inc[i0 + k / loc] * loc + k % loc] : INT_MAX;
    __syncthreads();
    for (int size = 2; size <= CAP; size <<= 1)
        for (int stride = size >> 1; stride > 0; stride >>= 1) {
            for (int k = lane; k < CAP; k += W) {
                const int p = k ^ stride;
                if (p > k) {
                    const int a = buf[wave][k], b = buf[wave][p];
                    const bool up = (k & size) == 0;
                    if ((a > b) == up) {
                        buf[wave][k] = b;
                        buf[wave][p] = a;
                    }
                }
            }
            __syncthreads();
        }
    if (!valid) return;
    int64_t out = MODE == 1 ? rowptr[row] : 0;
    int total = 0;
    for (int base = 0; base < CAP; base += W) {
        const int k = base + lane;
        const int v = buf[wave][k];
        const bool u = v != INT_MAX && (k == 0 || buf[wave][k - 1] != v);
        const uint64_t m = __ballot(u);
        if (MODE == 1 && u) {
            const int at = __popcll(m & ((1ull << lane) - 1ull));
            col[out + at] = v;
            if (v == row) diag[row] = (int32_t)(out + at);
        }
        const int c = __popcll(m);
        out += c;
        total += c;
    }
    if (MODE == 0 && lane == 0) rowcnt[row] = total;
}

// Speculative first-fit rounds (Gebremedhin & Manne 2000), made
// deterministic by double buffering: (pick) every element without a final
// colour takes the smallest colour no neighbour (element sharing a node)
// holds as final; (resolve) it keeps it, as final, unless a neighbour picked
// the same colour and has the higher priority (hashed index, ties by index).
// Both read only the previous round's finals, so two neighbours never end
// with one colour, and the uncoloured element of highest priority always
// keeps its pick: every round makes progress.
__device__ __forceinline__ uint32_t prio_hash(uint32_t x) {
    x ^= x >> 16;
    x *= 0x7feb352du;
    x ^= x >> 15;
    x *= 0x846ca68bu;
    x ^= x >> 16;
    return x;
}

constexpr int kMaxColours = 256;

__global__ void __launch_bounds__(256)
k_colour_pick(int nele, int loc, const int32_t* __restrict__ conn, const int64_t* __restrict__ iptr,
              const int32_t* __restrict__ inc, const int32_t* __restrict__ fin, int32_t* __restrict__ pick,
              int* __restrict__ err) {
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= nele) return;
    if (fin[e] >= 0) {
        pick[e] = -1;
        return;
    }
    uint64_t used[kMaxColours / 64] = {};
    for (int i = 0; i < loc; i++) {
        const int node = conn[(int64_t)e * loc + i];
        for (int64_t q = iptr[node]; q < iptr[node + 1]; q++) {
            const int cn = fin[inc[q]];
            if (cn >= 0 && cn < kMaxColours) used[cn >> 6] |= 1ull << (cn & 63);
        }
    }
    int c = -1;
    for (int w = 0; w < kMaxColours / 64 && c < 0; w++)
        if (~used[w]) c = w * 64 + __ffsll((unsigned long long)~used[w]) - 1;
    if (c < 0) {
        atomicOr(err, 1);
        c = 0;
    }
    pick[e] = c;
}

__global__ void __launch_bounds__(256)
k_colour_resolve(int nele, int loc, const int32_t* __restrict__ conn, const int64_t* __restrict__ iptr,
                 const int32_t* __restrict__ inc, const int32_t* __restrict__ fin, const int32_t* __restrict__ pick,
                 int32_t* __restrict__ fout, int* __restrict__ left) {
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= nele) return;
    if (fin[e] >= 0) {
        fout[e] = fin[e];
        return;
    }
    const int pe = pick[e];
    const uint32_t he = prio_hash((uint32_t)e);
    bool keep = true;
    for (int i = 0; i < loc && keep; i++) {
        const int node = conn[(int64_t)e * loc + i];
        for (int64_t q = iptr[node]; q < iptr[node + 1]; q++) {
            const int n = inc[q];
            if (n == e || pick[n] != pe) continue;
            const uint32_t hn = prio_hash((uint32_t)n);
            if (hn > he || (hn == he && n > e)) {
                keep = false;
                break;
            }
        }
    }
    fout[e] = keep ? pe : -1;
    if (!keep) atomicAdd(left, 1);
}

// node -> element incidence of the bound field mesh (inc ordered
// arbitrarily within a node; only sets are read from it)
int build_incidence(flgpu_ctx* c, DevBuf<int64_t>& iptr, DevBuf<int32_t>& inc, int* max_deg) {
    const int64_t n = (int64_t)c->nele * c->loc;
    DevBuf<int32_t> cnt;
    FLGPU_TRY(cnt.alloc((size_t)c->nnodes));
    FLGPU_HIP(hipMemsetAsync(cnt.p, 0, sizeof(int32_t) * c->nnodes, c->stream));
    hipLaunchKernelGGL(k_inc_count, dim3(grid_for(n, 256)), dim3(256), 0, c->stream, n, c->conn.p, cnt.p);
    FLGPU_HIP(hipGetLastError());
    FLGPU_TRY(exclusive_scan(c->stream, cnt.p, c->nnodes, iptr));
    DevBuf<int> mx;
    FLGPU_TRY(mx.alloc(1));
    FLGPU_HIP(hipMemsetAsync(mx.p, 0, sizeof(int), c->stream));
    hipLaunchKernelGGL(k_max_count, dim3(grid_for(c->nnodes, 256)), dim3(256), 0, c->stream, c->nnodes, cnt.p,
                       mx.p);
    FLGPU_HIP(hipGetLastError());
    FLGPU_TRY(inc.alloc((size_t)n));
    FLGPU_HIP(hipMemsetAsync(cnt.p, 0, sizeof(int32_t) * c->nnodes, c->stream));
    hipLaunchKernelGGL(k_inc_fill, dim3(grid_for(n, 256)), dim3(256), 0, c->stream, n, c->loc, c->conn.p, iptr.p,
                       cnt.p, inc.p);
    FLGPU_HIP(hipGetLastError());
    FLGPU_HIP(hipMemcpyAsync(max_deg, mx.p, sizeof(int), hipMemcpyDeviceToHost, c->stream));
    FLGPU_HIP(hipStreamSynchronize(c->stream));
    return FLGPU_OK;
}

template <int MODE>
int launch_rows(flgpu_ctx* c, int cap, const DevBuf<int64_t>& iptr, const DevBuf<int32_t>& inc, int32_t* rowcnt,
                const int64_t* rowptr, int32_t* col, int32_t* diag) {
    const dim3 grid((unsigned)((c->nnodes + 3) / 4)), block(256);
#define ROWS_CAP(CAP)                                                                                          \
    if (cap == CAP) {                                                                                          \
        hipLaunchKernelGGL((k_rows<MODE, CAP>), grid, block, 0, c->stream, c->nnodes, c->loc, c->conn.p, iptr.p, \
                           inc.p, rowcnt, rowptr, col, diag);                                                  \
    } else
    ROWS_CAP(64)
    ROWS_CAP(128)
    ROWS_CAP(256)
    ROWS_CAP(512)
    ROWS_CAP(1024) {
        return set_error(FLGPU_EUNSUPPORTED, "flgpu_sparsity_build: row candidates %d", cap);
    }
#undef ROWS_CAP
    FLGPU_HIP(hipGetLastError());
    return FLGPU_OK;
}

__global__ void __launch_bounds__(256)
k_plus_one(int64_t n, const int64_t* __restrict__ a, int32_t* __restrict__ out) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) out[i] = (int32_t)(a[i] + 1);
}

__global__ void __launch_bounds__(256)
k_plus_one_i32(int64_t n, int32_t* __restrict__ a) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) a[i] = a[i] + 1;
}

}  // namespace
}  // namespace flgpu

using namespace flgpu;

extern "C" int flgpu_sparsity_build(flgpu_ctx* c, int64_t* nnz_out) {
    return guarded("flgpu_sparsity_build", [&]() -> int {
        if (!c) return set_error(FLGPU_EINVAL, "flgpu_sparsity_build: null ctx");
        if (c->nnodes <= 0 || c->nele <= 0) return set_error(FLGPU_EINVAL, "flgpu_sparsity_build: bind the meshes first");
        FLGPU_HIP(hipSetDevice(c->device));
        DevBuf<int64_t> iptr;
        DevBuf<int32_t> inc;
        int max_deg = 0;
        FLGPU_TRY(build_incidence(c, iptr, inc, &max_deg));
        int cap = 64;
        while (cap < max_deg * c->loc) cap <<= 1;
        if (cap > 1024)
            return set_error(FLGPU_EUNSUPPORTED, "flgpu_sparsity_build: a node in %d elements (at most %d)", max_deg,
                             1024 / c->loc);
        DevBuf<int32_t> cnt;
        FLGPU_TRY(cnt.alloc((size_t)c->nnodes));
        FLGPU_TRY(launch_rows<0>(c, cap, iptr, inc, cnt.p, nullptr, nullptr, nullptr));
        DevBuf<int64_t> rp;
        FLGPU_TRY(exclusive_scan(c->stream, cnt.p, c->nnodes, rp));
        int64_t nnz = 0;
        FLGPU_HIP(hipMemcpyAsync(&nnz, rp.p + c->nnodes, sizeof(int64_t), hipMemcpyDeviceToHost, c->stream));
        FLGPU_HIP(hipStreamSynchronize(c->stream));
        if (nnz > INT32_MAX - 1) return set_error(FLGPU_EUNSUPPORTED, "flgpu_sparsity_build: %lld entries", (long long)nnz);
        DevBuf<int32_t> col, diag;
        FLGPU_TRY(col.alloc((size_t)nnz));
        FLGPU_TRY(diag.alloc((size_t)c->nnodes));
        FLGPU_HIP(hipMemsetAsync(diag.p, 0xff, sizeof(int32_t) * c->nnodes, c->stream));  // -1: no diagonal
        FLGPU_TRY(launch_rows<1>(c, cap, iptr, inc, nullptr, rp.p, col.p, diag.p));
        // 1-based csr_sparsity arrays, then the host binding
        DevBuf<int32_t> findrm;
        FLGPU_TRY(findrm.alloc((size_t)c->nnodes + 1));
        hipLaunchKernelGGL(k_plus_one, dim3(grid_for(c->nnodes + 1, 256)), dim3(256), 0, c->stream,
                           (int64_t)c->nnodes + 1, rp.p, findrm.p);
        hipLaunchKernelGGL(k_plus_one_i32, dim3(grid_for(nnz, 256)), dim3(256), 0, c->stream, nnz, col.p);
        hipLaunchKernelGGL(k_plus_one_i32, dim3(grid_for(c->nnodes, 256)), dim3(256), 0, c->stream,
                           (int64_t)c->nnodes, diag.p);
        FLGPU_HIP(hipGetLastError());
        std::vector<int32_t> hf((size_t)c->nnodes + 1), hc((size_t)nnz), hd((size_t)c->nnodes);
        FLGPU_HIP(hipMemcpyAsync(hf.data(), findrm.p, sizeof(int32_t) * hf.size(), hipMemcpyDeviceToHost, c->stream));
        FLGPU_HIP(hipMemcpyAsync(hc.data(), col.p, sizeof(int32_t) * hc.size(), hipMemcpyDeviceToHost, c->stream));
        FLGPU_HIP(hipMemcpyAsync(hd.data(), diag.p, sizeof(int32_t) * hd.size(), hipMemcpyDeviceToHost, c->stream));
        FLGPU_HIP(hipStreamSynchronize(c->stream));
        FLGPU_TRY(flgpu_sparsity_bind(c, c->nnodes, c->nnodes, hf.data(), hc.data(), hd.data()));
        if (nnz_out) *nnz_out = nnz;
        return FLGPU_OK;
    });
}

extern "C" int flgpu_sparsity_get(flgpu_ctx* c, int32_t* findrm, int32_t* colm, int32_t* centrm) {
    return guarded("flgpu_sparsity_get", [&]() -> int {
        if (!c) return set_error(FLGPU_EINVAL, "flgpu_sparsity_get: null ctx");
        if (c->nnz <= 0) return set_error(FLGPU_EINVAL, "flgpu_sparsity_get: no sparsity bound");
        if (findrm)
            for (int i = 0; i <= c->rows; i++) findrm[i] = c->h_rowptr[i] + 1;
        if (colm)
            for (int64_t k = 0; k < c->nnz; k++) colm[k] = c->h_col[k] + 1;
        if (centrm) {
            std::vector<int32_t> d(c->rows);
            FLGPU_HIP(hipSetDevice(c->device));
            FLGPU_HIP(hipMemcpy(d.data(), c->diag.p, sizeof(int32_t) * c->rows, hipMemcpyDeviceToHost));
            for (int i = 0; i < c->rows; i++) centrm[i] = d[i] + 1;
        }
        return FLGPU_OK;
    });
}

extern "C" int flgpu_colouring_build(flgpu_ctx* c, int32_t* ncol_out) {
    return guarded("flgpu_colouring_build", [&]() -> int {
        if (!c) return set_error(FLGPU_EINVAL, "flgpu_colouring_build: null ctx");
        if (c->nnodes <= 0 || c->nele <= 0) return set_error(FLGPU_EINVAL, "flgpu_colouring_build: bind the meshes first");
        FLGPU_HIP(hipSetDevice(c->device));
        DevBuf<int64_t> iptr;
        DevBuf<int32_t> inc;
        int max_deg = 0;
        FLGPU_TRY(build_incidence(c, iptr, inc, &max_deg));
        DevBuf<int32_t> ca, cb, pick;
        DevBuf<int> cnt;  // [0] uncoloured after the round, [1] error
        FLGPU_TRY(ca.alloc((size_t)c->nele));
        FLGPU_TRY(cb.alloc((size_t)c->nele));
        FLGPU_TRY(pick.alloc((size_t)c->nele));
        FLGPU_TRY(cnt.alloc(2));
        FLGPU_HIP(hipMemsetAsync(ca.p, 0xff, sizeof(int32_t) * c->nele, c->stream));
        FLGPU_HIP(hipMemsetAsync(cnt.p, 0, sizeof(int) * 2, c->stream));
        int32_t* cin = ca.p;
        int32_t* cout = cb.p;
        // every round finalises at least the uncoloured element of highest
        // priority, so nele rounds always suffice
        int left = c->nele, rounds = 0;
        const dim3 grid(grid_for(c->nele, 256)), block(256);
        while (left > 0) {
            if (rounds++ > c->nele) return set_error(FLGPU_EHIP, "flgpu_colouring_build: no progress");
            FLGPU_HIP(hipMemsetAsync(cnt.p, 0, sizeof(int), c->stream));
            hipLaunchKernelGGL(k_colour_pick, grid, block, 0, c->stream, c->nele, c->loc, c->conn.p, iptr.p, inc.p, cin,
                               pick.p, cnt.p + 1);
            hipLaunchKernelGGL(k_colour_resolve, grid, block, 0, c->stream, c->nele, c->loc, c->conn.p, iptr.p, inc.p,
                               cin, pick.p, cout, cnt.p);
            FLGPU_HIP(hipGetLastError());
            int h[2];
            FLGPU_HIP(hipMemcpyAsync(h, cnt.p, sizeof(int) * 2, hipMemcpyDeviceToHost, c->stream));
            FLGPU_HIP(hipStreamSynchronize(c->stream));
            if (h[1]) return set_error(FLGPU_EUNSUPPORTED, "flgpu_colouring_build: more than %d colours", kMaxColours);
            left = h[0];
            std::swap(cin, cout);
        }
        std::vector<int32_t> colour(c->nele);
        FLGPU_HIP(hipMemcpyAsync(colour.data(), cin, sizeof(int32_t) * c->nele, hipMemcpyDeviceToHost, c->stream));
        FLGPU_HIP(hipStreamSynchronize(c->stream));
        // colour sets, ascending element ids inside each (counting sort)
        int ncol = 0;
        for (int e = 0; e < c->nele; e++) ncol = std::max(ncol, colour[e] + 1);
        std::vector<int32_t> ptr(ncol + 1, 0), elems(c->nele);
        for (int e = 0; e < c->nele; e++) ptr[colour[e] + 1]++;
        for (int k = 0; k < ncol; k++) ptr[k + 1] += ptr[k];
        {
            std::vector<int32_t> fill(ptr.begin(), ptr.end() - 1);
            for (int e = 0; e < c->nele; e++) elems[fill[colour[e]]++] = e;
        }
        // valid by construction (no two finals of one colour are neighbours):
        // bound without the host check of flgpu_colouring_bind
        FLGPU_TRY(colouring_adopt(c, ptr, elems));
        if (ncol_out) *ncol_out = ncol;
        return FLGPU_OK;
    });
}

extern "C" int flgpu_colouring_get(flgpu_ctx* c, int32_t* col_ptr, int32_t* elems) {
    return guarded("flgpu_colouring_get", [&]() -> int {
        if (!c) return set_error(FLGPU_EINVAL, "flgpu_colouring_get: null ctx");
        if (c->ncol <= 0) return set_error(FLGPU_EINVAL, "flgpu_colouring_get: no colouring bound");
        if (col_ptr)
            for (int k = 0; k <= c->ncol; k++) col_ptr[k] = c->h_col_ptr[k] + 1;
        if (elems)
            for (int e = 0; e < c->nele; e++) elems[e] = c->h_col_elems[e] + 1;
        return FLGPU_OK;
    });
}
