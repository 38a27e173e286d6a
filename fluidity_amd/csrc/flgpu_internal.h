// Internal declarations of the gfx950 CG advection-diffusion library.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>
#include <string>
#include <vector>

#include "../../include/flgpu.h"

namespace flgpu {

constexpr int kMaxLoc = 10;   // P2 tetrahedron
constexpr int kMaxGi = 11;    // degree-4 tetrahedron rule
constexpr int kMaxDim = 3;

int set_error(int code, const char* fmt, ...);

#define FLGPU_HIP(call)                                                     \
    do {                                                                    \
        hipError_t e_ = (call);                                             \
        if (e_ != hipSuccess)                                               \
            return ::flgpu::set_error(FLGPU_EHIP, "%s failed: %s (%s:%d)",  \
                                      #call, hipGetErrorString(e_),         \
                                      __FILE__, __LINE__);                  \
    } while (0)

#define FLGPU_TRY(expr)                 \
    do {                                \
        int rc_ = (expr);               \
        if (rc_ != FLGPU_OK) return rc_; \
    } while (0)

// Runs an entry point's body and turns a C++ exception thrown by its host
// code (a std::vector that cannot be allocated, above all) into an error
// status, so that no exception crosses the C ABI.
template <typename F>
int guarded(const char* name, F&& body) noexcept {
    try {
        return body();
    } catch (const std::bad_alloc&) {
        return set_error(FLGPU_ENOMEM, "%s: host memory allocation failed", name);
    } catch (const std::exception& e) {
        return set_error(FLGPU_EINVAL, "%s: %s", name, e.what());
    } catch (...) {
        return set_error(FLGPU_EINVAL, "%s: unexpected exception", name);
    }
}

// Device buffer: owning (alloc / upload), or a view of memory owned
// elsewhere (borrow), which release() forgets without freeing.
template <typename T>
struct DevBuf {
    T* p = nullptr;
    size_t n = 0;
    bool owned = true;
    DevBuf() = default;
    DevBuf(const DevBuf&) = delete;
    DevBuf& operator=(const DevBuf&) = delete;
    ~DevBuf() { release(); }
    void release() {
        if (p && owned) (void)hipFree(p);
        p = nullptr;
        n = 0;
        owned = true;
    }
    void borrow(T* ptr, size_t count) {
        release();
        p = ptr;
        n = count;
        owned = false;
    }
    int alloc(size_t count) {
        if (count == n && p && owned) return FLGPU_OK;
        release();
        if (count == 0) return FLGPU_OK;
        hipError_t e = hipMalloc(&p, count * sizeof(T));
        if (e != hipSuccess) {
            p = nullptr;
            return set_error(FLGPU_ENOMEM, "hipMalloc of %zu bytes failed: %s",
                             count * sizeof(T), hipGetErrorString(e));
        }
        n = count;
        return FLGPU_OK;
    }
    // grow only: keeps the buffer when it is already large enough
    int reserve(size_t count) {
        if (p && owned && count <= n) return FLGPU_OK;
        return alloc(count);
    }
    int upload(const T* host, size_t count, hipStream_t s) {
        FLGPU_TRY(alloc(count));
        if (count)
            FLGPU_HIP(hipMemcpyAsync(p, host, count * sizeof(T),
                                     hipMemcpyHostToDevice, s));
        return FLGPU_OK;
    }
};

// Element tables, device-side copy (staged into LDS by the kernels).
struct ElemTables {
    int32_t dim, loc, ngi, loc_x;
    double n[kMaxLoc * kMaxGi];             // n[i + loc*g]
    double dn[kMaxLoc * kMaxGi * kMaxDim];  // dn[i + loc*(g + ngi*d)]
    double w[kMaxGi];
    double xdn[4 * kMaxDim];                // coordinate dn at g=1: [i + loc_x*d]
    // Contractions used when the physical gradients are constant over the
    // element (linear T on linear simplices):
    double what[kMaxLoc];                   // sum_g w_g n(i,g)
    double mhat[kMaxLoc * kMaxLoc];         // sum_g w_g n(i,g) n(j,g)
    double that[4 * 4 * 4];                 // sum_g w_g n_i n_j n_k (P1 only)
};

// Fields on the device.  A constant field keeps one value (ncomp numbers)
// and `constant` is set.
struct DevField {
    DevBuf<double> v;
    int32_t constant = 0;
    int32_t set = 0;
};

// Row-block plan of the blocked P1 assembly: rows grouped in Morton order
// of their positions; per block the elements touching it; per element the
// rank of the element in each of its rows' colour-ordered element lists.
struct BlockPlan {
    DevBuf<int32_t> brow;        // nblocks+1 row boundaries: block b owns rows [brow[b], brow[b+1])
    DevBuf<int32_t> iptr;        // nodes+1: start of every node's incidence list in inc
    DevBuf<int32_t> inc;         // elements of every node, in colour order (ascending id in a colour)
    int32_t nblocks = 0, rows_cap = 0, nnz_cap = 0, slot_cap = 0;
    bool built = false;
    int fail_rc = 0;             // a plan that cannot be built for this mesh: the status and message
    std::string fail_msg;
    // node-staged slabs (field mesh = coordinate mesh): every slab's nodes
    // (its rows and their neighbours) are loaded into LDS once; slots name
    // their element's vertices by slab-local node number
    bool staged = false;
    DevBuf<int32_t> nptr;        // nblocks+1 offsets into nodes
    DevBuf<int32_t> nodes;       // global node ids of each slab, ascending
    DevBuf<int32_t> meta;        // per slab: r0 r1 k0 k1 s0 s1 n0 n1, one 32-byte record
    DevBuf<uint32_t> tabw, tabw32;  // the P1 element tables (LTabLin<double>, LTabLin<float>) as words
    DevBuf<uint32_t> slotw;      // per slot: 4 x 8-bit slab-local vertex numbers
    DevBuf<uint32_t> slotd;      // per slot: row-relative places of its off-diagonal values (9 bits
                                 // each) and the number of the row's vertex in the element (2 bits)
    DevBuf<uint8_t> srb;         // per slot: slab-local row
    DevBuf<uint16_t> ecp;        // per CSR entry: slab-relative offset of its run of values
    DevBuf<uint16_t> rdg;        // per row: row-relative offset of its diagonal run
    int32_t node_cap = 0;
    int64_t node_visits = 0;     // staged nodes summed over the slabs
    double build_seconds = 0.0;  // host time of the last plan build
    // per element, its LOC^2 entries' CSR positions (-1: not in the
    // pattern), for the quadratic-T colour kernel; the element tables as
    // the kernel keeps them in LDS (FP64 and FP32 images)
    DevBuf<int32_t> epos;
    bool epos_built = false;
    DevBuf<uint32_t> ltab64, ltab32;
    std::vector<uint32_t> h_ltab64, h_ltab32;  // what the images hold
    // quadratic T in two passes (element values, then per-entry sums): for
    // every CSR entry (gptr, glist) and rhs row (rptr, rlist) the element
    // values it receives, as indices e*LOC^2 + entry / e*LOC + row into
    // gval / grv (e the element id), in colour order, ascending element id
    // inside a colour; gbad lists the element entries the pattern lacks
    bool gather_built = false;
    DevBuf<int32_t> gptr, glist, rptr, rlist, gbad;
    int32_t ngbad = 0;
    DevBuf<double> gval, grv;
    // stabilised node-staged assembly: nu_bar per element and quadrature
    // point (FP64 or FP32 values, as 8-byte words)
    DevBuf<double> nub;
};

// Surface elements of T for the face loop (surface.hip).
constexpr int kMaxFaceLoc = 6;   // P2 triangle
constexpr int kMaxFaceGi = 16;

struct FaceTables {
    int32_t floc, fngi, floc_x, pad;
    double fn[kMaxFaceLoc * kMaxFaceGi];      // fn[i + floc*g]: face_shape(t)%n
    double fw[kMaxFaceGi];
    double fxdn[kMaxDim * kMaxFaceGi * 2];    // fxdn[k + floc_x*(g + fngi*d)]: face_shape(positions)%dn
};

struct SurfaceState {
    bool bound = false;
    int32_t nfaces = 0, floc = 0, fngi = 0;
    DevBuf<int32_t> fnodes, fnodes_x, fele;   // 0-based: T nodes, coordinate nodes, element of each face
    std::vector<int32_t> h_fnodes;
    FaceTables tab{};
    DevBuf<FaceTables> d_tab;
    // addition lists for the bound sparsity: per CSR entry / rhs row the
    // face values it receives, in face order
    int64_t lists_gen = -1;
    int32_t nent = 0, nrow = 0;
    DevBuf<int32_t> tpos, tptr, tsrc, rrow, rptr, rsrc;
    DevBuf<double> fm, fr;                    // per face: floc^2 matrix values, floc rhs values
    DevBuf<int32_t> bct;
    DevBuf<double> bcv, bcv2;
};

struct KrylovWork;   // krylov.hip
struct LearnedState; // learned.hip
struct PressureState; // pressure.hip

}  // namespace flgpu

struct flgpu_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    hipDeviceProp_t prop{};

    // coordinate mesh
    int64_t key_id = -1, key_event = -1;
    int32_t dim = 0, nele = 0, loc_x = 0, nx = 0;
    flgpu::DevBuf<int32_t> conn_x;   // loc_x*nele, 0-based, element-major
    flgpu::DevBuf<double> X4;        // 4 doubles per node (x,y,z,0)
    // mesh movement: old and new positions and the grid velocity, laid out
    // as X4, for a mesh of motion_nx nodes in motion_dim dimensions
    flgpu::DevBuf<double> Xo4, Xn4, UG4;
    int motion_nx = 0, motion_dim = 0;
    std::vector<int32_t> h_conn_x;

    // field mesh
    int32_t loc = 0, ngi = 0, nnodes = 0;
    flgpu::DevBuf<int32_t> conn;     // loc*nele, 0-based, element-major
    std::vector<int32_t> h_conn;
    flgpu::ElemTables tab{};
    flgpu::DevBuf<flgpu::ElemTables> d_tab;
    bool linear = false;             // T is degree 1

    // sparsity (0-based)
    int32_t rows = 0, cols = 0;
    int64_t nnz = 0;
    int64_t sparsity_gen = 0;        // bumped by every flgpu_sparsity_bind
    bool all_diag = false;           // every row holds its diagonal entry
    flgpu::DevBuf<int32_t> rowptr, col, diag;
    std::vector<int32_t> h_rowptr, h_col;

    // colouring (0-based)
    int32_t ncol = 0;
    std::vector<int32_t> h_col_ptr, h_col_elems;
    flgpu::DevBuf<int32_t> col_elems;

    // fields
    flgpu::DevField fT, fU, fK, fS, fA;
    flgpu::DevField fRho, fRhoOld, fP, fV;  // density, old density, pressure, phase volume fraction

    // assembled system
    flgpu::DevBuf<double> A, rhs;
    flgpu::DevBuf<int32_t> inactive;
    bool have_inactive = false;
    bool assembled = false;

    flgpu::BlockPlan plan;
    flgpu::SurfaceState surf;
    flgpu::KrylovWork* kw = nullptr;
    flgpu::LearnedState* ls = nullptr;
    flgpu::PressureState* ps = nullptr;
    int64_t plan_gen = 0;            // bumped whenever the meshes, sparsity or colouring change
};

namespace flgpu {
// assemble.hip
int assemble(flgpu_ctx* c, const flgpu_advdif_params* p);
int build_block_plan(flgpu_ctx* c);
int time_assemble(flgpu_ctx* c, const flgpu_advdif_params* p, int reps, double* ms);
int assembly_plan_info(flgpu_ctx* c, int64_t* info, int n, double* build_seconds);
// krylov.hip
int solve(flgpu_ctx* c, const flgpu_solver_params* sp, double* x_host,
          flgpu_solve_info* info);
// the allocation-free solve: its work lives in a fixed device array set
// up once per device by noalloc_setup (flgpu_init)
int solve_noalloc(flgpu_ctx* c, const flgpu_solver_params* sp, double* x_host, flgpu_solve_info* info);
int noalloc_setup(int device);
void krylov_free(flgpu_ctx* c);
void krylov_invalidate(flgpu_ctx* c);  // the assembled system changed
int solver_prepare(flgpu_ctx* c, int precision);
int time_spmv(flgpu_ctx* c, int precision, int reps, double* ms);
int time_spmv_cold(flgpu_ctx* c, int precision, int reps, int flush_mib, double* ms);
int csr_mult(flgpu_ctx* c, const double* x, double* y);
int solution_get(flgpu_ctx* c, double* x);
int pc_apply_host(flgpu_ctx* c, const flgpu_solver_params* sp, const double* r, double* z);
// learned.hip
void learned_free(flgpu_ctx* c);
// pressure.hip
void pressure_free(flgpu_ctx* c);
// shared helpers
unsigned grid_for(int64_t n, int block);
// setup_gpu.hip: out[0..n) = exclusive prefix sums of in[0..n), out[n] = the total
int exclusive_scan(hipStream_t s, const int32_t* in, int64_t n, DevBuf<int64_t>& out);
// flgpu_context.hip: bind a colouring already known to be valid (0-based)
int colouring_adopt(flgpu_ctx* c, std::vector<int32_t>& col_ptr, std::vector<int32_t>& elems);
}  // namespace flgpu
