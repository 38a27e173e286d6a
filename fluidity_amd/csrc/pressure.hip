// The pressure path of the momentum solve on gfx950: the divergence block
// ct_m of construct_momentum_cg with the lumped velocity mass it assembles
// alongside (assemble/Momentum_CG.F90:1377-1403, :1468-1469, :1492-1600), and
// C_P^T M_l^-1 C of assemble_masslumped_cmc (assemble/Assemble_CMC.F90:
// 119-135 -> femtools/Sparse_Matrices_Fields.F90:590-671).
//
// The velocity mesh is the bound field mesh (its colouring, quadrature and
// shape tables); the pressure mesh is bound here with its own connectivity
// and tables on the same quadrature, and ct_m's sparsity (rows: pressure
// nodes, columns: velocity nodes).
//
// ct_m and the lumped mass in two passes, as the quadratic-T assembly: one
// thread per element entry (p node, u node) forms its dim values (threads of
// the first p node also the lumped mass of their u node) and stores them by
// element id; then one thread per ct entry (per velocity node) adds its
// values in colour order, ascending element id inside a colour -- the order
// of the reference's colour loop, with zeros added (which leaves the sums'
// bits as block_csr_blocks_addto's skip of zeros does).
//
// Arithmetic: the oracle's operation order, no contraction
// (-ffp-contract=off for this file), so results are the oracle's bit for bit.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <type_traits>
#include <vector>

#include "flgpu_internal.h"

namespace flgpu {

struct PTables {
    int32_t p_loc, ngi;
    double n[kMaxLoc * kMaxGi];             // n[i + p_loc*g]
    double dn[kMaxLoc * kMaxGi * kMaxDim];  // dn[i + p_loc*(g + ngi*d)]
};

struct PressureState {
    bool bound = false;
    int32_t p_loc = 0, p_nodes = 0;
    int64_t ct_nnz = 0;
    DevBuf<int32_t> p_conn;                  // p_loc * nele, 0-based
    std::vector<int32_t> h_p_conn;
    PTables tab{};
    DevBuf<PTables> d_tab;
    DevBuf<int32_t> ct_rowptr, ct_col;       // 0-based
    std::vector<int32_t> h_ct_rowptr;
    // per ct entry / velocity node, its element values in colour order
    int64_t plan_key = -1;                   // plan_gen the lists were built for
    DevBuf<int32_t> cptr, clist, cbad, mptr, mlist;
    int32_t ncbad = 0;
    DevBuf<double> cval, mval;               // element values: dim per entry, one per u row
    DevBuf<double> ct, ml;                   // ct_m (dim blocks of ct_nnz) and the lumped mass
    bool assembled = false;
    DevBuf<int32_t> cmc_rowptr, cmc_col;     // the last cmc pattern
    int32_t cmc_rows = 0;
    int64_t cmc_nnz = -1;                    // -1: no cmc assembled yet
    DevBuf<double> cmc, inv;
    DevBuf<double> rho;                      // the last density (one value when constant)
    DevBuf<int> err;
    int last_ibp = 0, last_rho_const = 1;
};

namespace {

// Position of column cj in row r of a sorted 0-based CSR; -1 when absent.
__device__ __forceinline__ int64_t csr_find(const int32_t* __restrict__ rowptr, const int32_t* __restrict__ col,
                                            int r, int cj) {
    int lo = rowptr[r], hi = rowptr[r + 1] - 1;
    while (lo <= hi) {
        const int mid = (lo + hi) >> 1;
        const int v = col[mid];
        if (v == cj) return mid;
        if (v < cj) lo = mid + 1;
        else hi = mid - 1;
    }
    return -1;
}

// Lists: item t of the colour-order numbering (q = t / per) goes to the
// list of its key; KIND 0: ct entries (per = p_loc*u_loc, key = the ct
// position of (p node, u node)), KIND 1: lumped mass (per = u_loc, key = the
// u node).  A ct item outside the pattern has key -1.
template <int KIND>
__device__ __forceinline__ int64_t item_key(int64_t t, int per, int uloc, int ploc, const int32_t* __restrict__ elems,
                                            const int32_t* __restrict__ conn, const int32_t* __restrict__ pconn,
                                            const int32_t* __restrict__ rowptr, const int32_t* __restrict__ col) {
    const int64_t q = t / per;
    const int r = (int)(t - q * per);
    const int64_t ele = elems[q];
    if (KIND == 1) return conn[ele * uloc + r];
    const int pi = r / uloc, uj = r - pi * uloc;
    return csr_find(rowptr, col, pconn[ele * ploc + pi], conn[ele * uloc + uj]);
}

template <int KIND>
__global__ void __launch_bounds__(256)
k_list_count(int64_t n, int per, int uloc, int ploc, const int32_t* __restrict__ elems,
             const int32_t* __restrict__ conn, const int32_t* __restrict__ pconn, const int32_t* __restrict__ rowptr,
             const int32_t* __restrict__ col, int32_t* __restrict__ cnt, int32_t* __restrict__ nbad) {
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n) return;
    const int64_t key = item_key<KIND>(t, per, uloc, ploc, elems, conn, pconn, rowptr, col);
    if (key >= 0)
        atomicAdd(&cnt[key], 1);
    else
        atomicAdd(nbad, 1);
}

template <int KIND>
__global__ void __launch_bounds__(256)
k_list_fill(int64_t n, int per, int uloc, int ploc, const int32_t* __restrict__ elems,
            const int32_t* __restrict__ conn, const int32_t* __restrict__ pconn, const int32_t* __restrict__ rowptr,
            const int32_t* __restrict__ col, const int32_t* __restrict__ ptr, int32_t* __restrict__ cur,
            int32_t* __restrict__ list, int32_t* __restrict__ bad, int32_t* __restrict__ nbad) {
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n) return;
    const int64_t key = item_key<KIND>(t, per, uloc, ploc, elems, conn, pconn, rowptr, col);
    if (key >= 0)
        list[ptr[key] + atomicAdd(&cur[key], 1)] = (int32_t)t;
    else
        bad[atomicAdd(nbad, 1)] = (int32_t)t;
}

// Each list ascending (one value per element holding the entry's nodes).
__global__ void __launch_bounds__(256)
k_list_sort(int n, const int32_t* __restrict__ ptr, int32_t* __restrict__ list) {
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

// Colour-order items rewritten as element-id items (elems[q]*per + r).
__global__ void __launch_bounds__(256)
k_list_remap(int64_t n, int per, const int32_t* __restrict__ elems, int32_t* __restrict__ list) {
    const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    const int t = list[k], q = t / per;
    list[k] = elems[q] * per + (t - q * per);
}

__global__ void __launch_bounds__(256)
k_to32(int64_t n, const int64_t* __restrict__ in, int32_t* __restrict__ out) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i <= n) out[i] = (int32_t)in[i];
}

// transform_to_physical's geometry of a linear element
// (femtools/Transform_elements.F90:227-306): J^T from the coordinate dn,
// the inverse by cofactors over the determinant.
template <int DIM>
__device__ __forceinline__ void elem_geometry(int ele, const int32_t* __restrict__ conn_x,
                                              const double* __restrict__ X4, const ElemTables* __restrict__ ut,
                                              double (&iJ)[DIM][DIM], double& detJ) {
    constexpr int LX = DIM + 1;
    double Xv[DIM][LX], JT[DIM][DIM];
#pragma unroll
    for (int n = 0; n < LX; n++) {
        const int xn = conn_x[(int64_t)ele * LX + n];
#pragma unroll
        for (int a = 0; a < DIM; a++) Xv[a][n] = X4[(int64_t)xn * 4 + a];
    }
#pragma unroll
    for (int a = 0; a < DIM; a++)
#pragma unroll
        for (int b = 0; b < DIM; b++) {
            double s = 0.0;
#pragma unroll
            for (int n = 0; n < LX; n++) s = s + Xv[a][n] * ut->xdn[n + LX * b];
            JT[a][b] = s;
        }
    if constexpr (DIM == 3) {
#pragma unroll
        for (int i = 0; i < 3; i++)
#pragma unroll
            for (int k = 0; k < 3; k++)
                iJ[i][k] = JT[(i + 1) % 3][(k + 1) % 3] * JT[(i + 2) % 3][(k + 2) % 3] -
                           JT[(i + 2) % 3][(k + 1) % 3] * JT[(i + 1) % 3][(k + 2) % 3];
    } else {
        iJ[0][0] = JT[1][1];
        iJ[1][0] = -JT[0][1];
        iJ[0][1] = -JT[1][0];
        iJ[1][1] = JT[0][0];
    }
    detJ = 0.0;
#pragma unroll
    for (int a = 0; a < DIM; a++) detJ = detJ + JT[a][0] * iJ[a][0];
#pragma unroll
    for (int i = 0; i < DIM; i++)
#pragma unroll
        for (int k = 0; k < DIM; k++) iJ[i][k] = iJ[i][k] / detJ;
}

// First pass: thread per (element, u node j), for every p node i the
// element entry grad_p_u_mat(d, i, j) = sum_g detwei_g (n_p(i,g)
// du_t(j,g,d)) (shape_dshape, femtools/FETools.F90:332-362), or by parts
// -sum_g detwei_g (dp_t(i,g,d) n_u(j,g)) (dshape_shape, :364-389), then
// the lumped mass of u node j, sum_k sum_g (n(j,g) n(k,g)) (rho_g
// detwei_g) (add_mass_element_cg, Momentum_CG.F90:1519-1532): the
// element's geometry once for its p_loc entries, every thread the same
// work.
template <int DIM>
__global__ void __launch_bounds__(256)
k_ct_elem(int nele, int ploc, int uloc, int ngi, const int32_t* __restrict__ conn, const int32_t* __restrict__ conn_x,
          const double* __restrict__ X4, const ElemTables* __restrict__ ut, const PTables* __restrict__ pt,
          const double* __restrict__ rho, int rho_const, int ibp, double* __restrict__ cval,
          double* __restrict__ mval) {
    const int per = ploc * uloc;
    const int64_t tj = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (tj >= (int64_t)nele * uloc) return;
    const int ele = (int)(tj / uloc), j = (int)(tj - (int64_t)ele * uloc);
    double iJ[DIM][DIM], detJ;
    elem_geometry<DIM>(ele, conn_x, X4, ut, iJ, detJ);
    const bool u_nonlinear = uloc != DIM + 1, p_nonlinear = ploc != DIM + 1;
    for (int i = 0; i < ploc; i++) {
        const int64_t t = (int64_t)ele * per + (int64_t)i * uloc + j;
        double s[DIM];
#pragma unroll
        for (int d = 0; d < DIM; d++) s[d] = 0.0;
        for (int g = 0; g < ngi; g++) {
            const double detwei = fabs(detJ) * ut->w[g];
            if (ibp) {
                const int gp = p_nonlinear ? g : 0;
                const double nj = ut->n[j + uloc * g];
#pragma unroll
                for (int d = 0; d < DIM; d++) {
                    double dp = 0.0;
#pragma unroll
                    for (int k = 0; k < DIM; k++) dp = dp + pt->dn[i + ploc * (gp + ngi * k)] * iJ[d][k];
                    s[d] = s[d] + detwei * (dp * nj);
                }
            } else {
                const int gu = u_nonlinear ? g : 0;
                const double ni = pt->n[i + ploc * g];
#pragma unroll
                for (int d = 0; d < DIM; d++) {
                    double du = 0.0;
#pragma unroll
                    for (int k = 0; k < DIM; k++) du = du + ut->dn[j + uloc * (gu + ngi * k)] * iJ[d][k];
                    s[d] = s[d] + detwei * (ni * du);
                }
            }
        }
#pragma unroll
        for (int d = 0; d < DIM; d++) cval[t * DIM + d] = ibp ? -s[d] : s[d];
    }
    {
        double lump = 0.0;
        for (int k = 0; k < uloc; k++) {
            double mjk = 0.0;
            for (int g = 0; g < ngi; g++) {
                double rg = 0.0;
                for (int q = 0; q < uloc; q++) {
                    const double v = rho_const ? rho[0] : rho[conn[(int64_t)ele * uloc + q]];
                    rg = rg + v * ut->n[q + uloc * g];
                }
                const double cdw = rg * (fabs(detJ) * ut->w[g]);
                mjk = mjk + (ut->n[j + uloc * g] * ut->n[k + uloc * g]) * cdw;
            }
            lump = lump + mjk;
        }
        mval[(int64_t)ele * uloc + j] = lump;
    }
}

// Second pass: thread per ct entry (dim sums), then per velocity node.
template <int DIM>
__global__ void __launch_bounds__(256)
k_ct_sum(int64_t nnz, int nodes, const int32_t* __restrict__ cptr, const int32_t* __restrict__ clist,
         const double* __restrict__ cval, const int32_t* __restrict__ mptr, const int32_t* __restrict__ mlist,
         const double* __restrict__ mval, double* __restrict__ ct, double* __restrict__ ml) {
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    // U list items a step: their index loads, then their value gathers, all
    // in flight before the additions in list order (a step past the end
    // re-reads the last item and adds nothing)
    constexpr int U = 4;
    if (t < nnz) {
        double s[DIM];
#pragma unroll
        for (int d = 0; d < DIM; d++) s[d] = 0.0;
        const int k0 = cptr[t], k1 = cptr[t + 1];
        for (int kb = k0; kb < k1; kb += U) {
            int64_t it[U];
            double v[U][DIM];
#pragma unroll
            for (int u = 0; u < U; u++) it[u] = clist[min(kb + u, k1 - 1)];
#pragma unroll
            for (int u = 0; u < U; u++)
#pragma unroll
                for (int d = 0; d < DIM; d++) v[u][d] = cval[it[u] * DIM + d];
#pragma unroll
            for (int u = 0; u < U; u++)
                if (kb + u < k1)
#pragma unroll
                    for (int d = 0; d < DIM; d++) s[d] = s[d] + v[u][d];
        }
#pragma unroll
        for (int d = 0; d < DIM; d++) ct[d * nnz + t] = s[d];
    } else if (t - nnz < nodes) {
        const int n = (int)(t - nnz);
        double s = 0.0;
        const int k0 = mptr[n], k1 = mptr[n + 1];
        for (int kb = k0; kb < k1; kb += U) {
            double v[U];
#pragma unroll
            for (int u = 0; u < U; u++) v[u] = mval[mlist[min(kb + u, k1 - 1)]];
#pragma unroll
            for (int u = 0; u < U; u++)
                if (kb + u < k1) s = s + v[u];
        }
        ml[n] = s;
    }
}

// An element entry outside the pattern is an error when one of its values
// is not zero (block_csr_blocks_addto looks up only such entries).
template <int DIM>
__global__ void __launch_bounds__(256)
k_ct_bad(int nbad, const int32_t* __restrict__ bad, const double* __restrict__ cval, int* __restrict__ err) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= nbad) return;
    bool any = false;
#pragma unroll
    for (int d = 0; d < DIM; d++) any |= cval[(int64_t)bad[t] * DIM + d] != 0.0;
    if (any) atomicOr(err, 1);
}

// inverse lumped mass, the same for every component (invert of the
// vector field, femtools/Fields_Manipulation.F90)
template <int DIM>
__global__ void __launch_bounds__(256) k_invert_ml(int n, const double* __restrict__ ml, double* __restrict__ inv) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const double v = 1.0 / ml[i];
#pragma unroll
    for (int d = 0; d < DIM; d++) inv[(int64_t)i * DIM + d] = v;
}

// C_P^T M_l^-1 C a wave per cmc row i: row i of ct (columns, the dim
// values and the inverse lumped mass of each column) staged in the wave's
// LDS slice once, then each lane takes an entry (i, j) and walks row j of
// ct in ascending column order, its column loads issued eight at a time, finding
// each column in the staged row by a
// pointer that only moves forward: for each common velocity node k and
// each component d (inner) ct_d(i,k) ct_d(j,k) inv(d,k) is added, the
// order and the products of k_cmc, so the same bits.  A row i longer than
// the slice (kCmcRow entries) takes k_cmc's walk from global memory.
constexpr int kCmcRow = 96;

template <int DIM>
__global__ void __launch_bounds__(256)
k_cmc_rows(int rows, const int32_t* __restrict__ cmc_rowptr, const int32_t* __restrict__ cmc_col,
           const int32_t* __restrict__ rowptr, const int32_t* __restrict__ col, const double* __restrict__ ct,
           int64_t ct_nnz, const double* __restrict__ inv, double* __restrict__ out) {
    __shared__ int32_t s_col[4][kCmcRow];
    __shared__ double s_val[4][DIM][kCmcRow], s_inv[4][DIM][kCmcRow];
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int i = blockIdx.x * 4 + wave;
    if (i >= rows) return;
    const int a0 = rowptr[i], a1 = rowptr[i + 1], la = a1 - a0;
    const bool staged = la <= kCmcRow;
    if (staged)
        for (int q = lane; q < la; q += 64) {
            const int k = col[a0 + q];
            s_col[wave][q] = k;
#pragma unroll
            for (int d = 0; d < DIM; d++) {
                s_val[wave][d][q] = ct[d * ct_nnz + a0 + q];
                s_inv[wave][d][q] = inv[(int64_t)k * DIM + d];
            }
        }
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    constexpr int U = 8;
    for (int e = cmc_rowptr[i] + lane; e < cmc_rowptr[i + 1]; e += 64) {
        const int j = cmc_col[e];
        const int b0 = rowptr[j], b1 = rowptr[j + 1];
        double entry0 = 0.0;
        if (staged) {
            int p = 0;
            for (int q0 = b0; q0 < b1; q0 += U) {
                int cq[U];
#pragma unroll
                for (int u = 0; u < U; u++) cq[u] = col[min(q0 + u, b1 - 1)];
#pragma unroll
                for (int u = 0; u < U; u++) {
                    if (q0 + u >= b1) break;
                    const int k = cq[u];
                    while (p < la && s_col[wave][p] < k) p++;
                    if (p < la && s_col[wave][p] == k) {
#pragma unroll
                        for (int d = 0; d < DIM; d++)
                            entry0 = entry0 + s_val[wave][d][p] * ct[d * ct_nnz + q0 + u] * s_inv[wave][d][p];
                        p++;
                    }
                }
            }
        } else {
            int k1 = a0, k2 = b0;
            while (k1 < a1 && k2 < b1) {
                const int c1 = col[k1], c2 = col[k2];
                if (c1 < c2) {
                    k1++;
                } else if (c1 == c2) {
#pragma unroll
                    for (int d = 0; d < DIM; d++)
                        entry0 = entry0 + ct[d * ct_nnz + k1] * ct[d * ct_nnz + k2] * inv[(int64_t)c1 * DIM + d];
                    k1++;
                    k2++;
                } else {
                    k2++;
                }
            }
        }
        out[e] = entry0;
    }
}

// One set of lists (KIND 0 or 1), sorted in colour order and remapped to
// element-id items.  keys: number of lists.
template <int KIND>
int build_lists(flgpu_ctx* c, PressureState& S, int64_t keys, int per, DevBuf<int32_t>& ptr, DevBuf<int32_t>& list,
                DevBuf<int32_t>* bad, int32_t* nbad_out) {
    const hipStream_t s = c->stream;
    const int64_t n = (int64_t)c->nele * per;
    DevBuf<int32_t> cnt, nbad;
    FLGPU_TRY(cnt.alloc((size_t)keys));
    FLGPU_TRY(nbad.alloc(1));
    FLGPU_HIP(hipMemsetAsync(cnt.p, 0, sizeof(int32_t) * keys, s));
    FLGPU_HIP(hipMemsetAsync(nbad.p, 0, sizeof(int32_t), s));
    hipLaunchKernelGGL(k_list_count<KIND>, dim3(grid_for(n, 256)), dim3(256), 0, s, n, per, c->loc, S.p_loc,
                       c->col_elems.p, c->conn.p, S.p_conn.p, S.ct_rowptr.p, S.ct_col.p, cnt.p, nbad.p);
    FLGPU_HIP(hipGetLastError());
    DevBuf<int64_t> p64;
    FLGPU_TRY(exclusive_scan(s, cnt.p, keys, p64));
    int64_t total = 0;
    int32_t h_nbad = 0;
    FLGPU_HIP(hipMemcpyAsync(&total, p64.p + keys, sizeof(int64_t), hipMemcpyDeviceToHost, s));
    FLGPU_HIP(hipMemcpyAsync(&h_nbad, nbad.p, sizeof(int32_t), hipMemcpyDeviceToHost, s));
    FLGPU_HIP(hipStreamSynchronize(s));
    FLGPU_TRY(ptr.alloc((size_t)keys + 1));
    hipLaunchKernelGGL(k_to32, dim3(grid_for(keys + 1, 256)), dim3(256), 0, s, keys, p64.p, ptr.p);
    FLGPU_TRY(list.alloc((size_t)std::max<int64_t>(total, 1)));
    DevBuf<int32_t> dummy;
    DevBuf<int32_t>& b = bad ? *bad : dummy;
    FLGPU_TRY(b.alloc((size_t)std::max(h_nbad, 1)));
    FLGPU_HIP(hipMemsetAsync(cnt.p, 0, sizeof(int32_t) * keys, s));
    FLGPU_HIP(hipMemsetAsync(nbad.p, 0, sizeof(int32_t), s));
    hipLaunchKernelGGL(k_list_fill<KIND>, dim3(grid_for(n, 256)), dim3(256), 0, s, n, per, c->loc, S.p_loc,
                       c->col_elems.p, c->conn.p, S.p_conn.p, S.ct_rowptr.p, S.ct_col.p, ptr.p, cnt.p, list.p, b.p,
                       nbad.p);
    hipLaunchKernelGGL(k_list_sort, dim3(grid_for(keys, 256)), dim3(256), 0, s, (int)keys, ptr.p, list.p);
    if (total > 0)
        hipLaunchKernelGGL(k_list_remap, dim3(grid_for(total, 256)), dim3(256), 0, s, total, per, c->col_elems.p,
                           list.p);
    if (h_nbad > 0)
        hipLaunchKernelGGL(k_list_remap, dim3(grid_for(h_nbad, 256)), dim3(256), 0, s, (int64_t)h_nbad, per,
                           c->col_elems.p, b.p);
    FLGPU_HIP(hipGetLastError());
    FLGPU_HIP(hipStreamSynchronize(s));  // the counters are freed on return
    if (nbad_out) *nbad_out = h_nbad;
    return FLGPU_OK;
}

int build_plan(flgpu_ctx* c, PressureState& S) {
    if (S.plan_key == c->plan_gen) return FLGPU_OK;
    const int64_t per = (int64_t)S.p_loc * c->loc;
    if ((int64_t)c->nele * per >= INT32_MAX || S.ct_nnz >= INT32_MAX)
        return set_error(FLGPU_EUNSUPPORTED, "flgpu ct assembly: more than 2^31 element entries");
    FLGPU_TRY(build_lists<0>(c, S, S.ct_nnz, (int)per, S.cptr, S.clist, &S.cbad, &S.ncbad));
    FLGPU_TRY(build_lists<1>(c, S, c->nnodes, c->loc, S.mptr, S.mlist, nullptr, nullptr));
    FLGPU_TRY(S.cval.alloc((size_t)c->nele * per * c->dim));
    FLGPU_TRY(S.mval.alloc((size_t)c->nele * c->loc));
    FLGPU_TRY(S.ct.alloc((size_t)S.ct_nnz * c->dim));
    FLGPU_TRY(S.ml.alloc((size_t)c->nnodes));
    S.plan_key = c->plan_gen;
    return FLGPU_OK;
}

// The ct_m and lumped-mass kernels on the device-resident inputs (density
// in S.rho, the error flag in S.err).
static int ct_launch(flgpu_ctx* c, PressureState& S, int ibp, int density_const) {
    const hipStream_t s = c->stream;
    auto launch = [&](auto dimtag) {
        constexpr int D = decltype(dimtag)::value;
        hipLaunchKernelGGL(k_ct_elem<D>, dim3(grid_for((int64_t)c->nele * c->loc, 256)), dim3(256), 0, s, c->nele,
                           S.p_loc, c->loc, c->ngi, c->conn.p, c->conn_x.p, c->X4.p, c->d_tab.p, S.d_tab.p, S.rho.p,
                           density_const, ibp, S.cval.p, S.mval.p);
        hipLaunchKernelGGL(k_ct_sum<D>, dim3(grid_for(S.ct_nnz + c->nnodes, 256)), dim3(256), 0, s, S.ct_nnz,
                           c->nnodes, S.cptr.p, S.clist.p, S.cval.p, S.mptr.p, S.mlist.p, S.mval.p, S.ct.p, S.ml.p);
        if (S.ncbad > 0)
            hipLaunchKernelGGL(k_ct_bad<D>, dim3(grid_for(S.ncbad, 256)), dim3(256), 0, s, S.ncbad, S.cbad.p,
                               S.cval.p, S.err.p);
    };
    if (c->dim == 3)
        launch(std::integral_constant<int, 3>());
    else
        launch(std::integral_constant<int, 2>());
    FLGPU_HIP(hipGetLastError());
    return FLGPU_OK;
}

// C^T M_l^-1 C on the bound cmc pattern (own_inverse: 1 / the lumped mass).
static int cmc_launch(flgpu_ctx* c, PressureState& S, bool own_inverse) {
    const hipStream_t s = c->stream;
    auto launch = [&](auto dimtag) {
        constexpr int D = decltype(dimtag)::value;
        if (own_inverse)
            hipLaunchKernelGGL(k_invert_ml<D>, dim3(grid_for(c->nnodes, 256)), dim3(256), 0, s, c->nnodes, S.ml.p,
                               S.inv.p);
        if (S.cmc_nnz > 0)
            hipLaunchKernelGGL(k_cmc_rows<D>, dim3((S.cmc_rows + 3) / 4), dim3(256), 0, s, S.cmc_rows,
                               S.cmc_rowptr.p, S.cmc_col.p, S.ct_rowptr.p, S.ct_col.p, S.ct.p, S.ct_nnz, S.inv.p,
                               S.cmc.p);
    };
    if (c->dim == 3)
        launch(std::integral_constant<int, 3>());
    else
        launch(std::integral_constant<int, 2>());
    FLGPU_HIP(hipGetLastError());
    return FLGPU_OK;
}

int ct_assemble(flgpu_ctx* c, PressureState& S, int ibp, int density_const, const double* density, double* ct_out,
                double* ml_out) {
    const int dim = c->dim;
    FLGPU_TRY(build_plan(c, S));
    S.assembled = false;
    const hipStream_t s = c->stream;
    FLGPU_TRY(S.rho.upload(density, density_const ? 1 : (size_t)c->nnodes, s));
    FLGPU_TRY(S.err.alloc(1));
    FLGPU_HIP(hipMemsetAsync(S.err.p, 0, sizeof(int), s));
    if (!c->d_tab.p) FLGPU_TRY(c->d_tab.upload(&c->tab, 1, s));
    FLGPU_TRY(ct_launch(c, S, ibp, density_const));
    S.last_ibp = ibp;
    S.last_rho_const = density_const;
    int h_err = 0;
    FLGPU_HIP(hipMemcpyAsync(&h_err, S.err.p, sizeof(int), hipMemcpyDeviceToHost, s));
    if (ct_out)
        FLGPU_HIP(hipMemcpyAsync(ct_out, S.ct.p, sizeof(double) * S.ct_nnz * dim, hipMemcpyDeviceToHost, s));
    if (ml_out) FLGPU_HIP(hipMemcpyAsync(ml_out, S.ml.p, sizeof(double) * c->nnodes, hipMemcpyDeviceToHost, s));
    FLGPU_HIP(hipStreamSynchronize(s));
    if (h_err) return set_error(FLGPU_EPATTERN, "Attempting to set value in matrix outside sparsity pattern.");
    S.assembled = true;
    return FLGPU_OK;
}

int cmc_assemble(flgpu_ctx* c, PressureState& S, const double* inverse_masslump, int rows, const int32_t* findrm,
                 const int32_t* colm, double* out) {
    const int dim = c->dim;
    const hipStream_t s = c->stream;
    if (rows != S.p_nodes)
        return set_error(FLGPU_EINVAL, "flgpu_cmc_assemble: %d rows for %d pressure nodes", rows, S.p_nodes);
    if (findrm[0] != 1) return set_error(FLGPU_EINVAL, "flgpu_cmc_assemble: findrm(1) = %d, not 1", findrm[0]);
    std::vector<int32_t> rp(rows + 1);
    for (int i = 0; i <= rows; i++) {
        rp[i] = findrm[i] - 1;
        if (i > 0 && rp[i] < rp[i - 1]) return set_error(FLGPU_EINVAL, "flgpu_cmc_assemble: findrm decreases at %d", i);
    }
    const int64_t nnz = rp[rows];
    std::vector<int32_t> cl((size_t)nnz);
    for (int64_t k = 0; k < nnz; k++) {
        const int v = colm[k] - 1;
        if (v < 0 || v >= S.p_nodes)
            return set_error(FLGPU_EINVAL, "flgpu_cmc_assemble: colm(%lld) = %d out of 1..%d", (long long)k + 1,
                             colm[k], S.p_nodes);
        cl[k] = v;
    }
    FLGPU_TRY(S.cmc_rowptr.upload(rp.data(), rp.size(), s));
    FLGPU_TRY(S.cmc_col.upload(cl.data(), cl.size(), s));
    FLGPU_TRY(S.cmc.alloc((size_t)std::max<int64_t>(nnz, 1)));
    FLGPU_TRY(S.inv.alloc((size_t)c->nnodes * dim));
    if (inverse_masslump)
        FLGPU_HIP(hipMemcpyAsync(S.inv.p, inverse_masslump, sizeof(double) * c->nnodes * dim, hipMemcpyHostToDevice,
                                 s));
    S.cmc_rows = rows;
    S.cmc_nnz = nnz;
    FLGPU_TRY(cmc_launch(c, S, !inverse_masslump));
    if (out && nnz > 0) FLGPU_HIP(hipMemcpyAsync(out, S.cmc.p, sizeof(double) * nnz, hipMemcpyDeviceToHost, s));
    FLGPU_HIP(hipStreamSynchronize(s));  // rp, cl and the host arrays
    return FLGPU_OK;
}

}  // namespace

// Device milliseconds per call of the ct_m + lumped-mass kernels and of
// the C^T M_l^-1 C kernel (with the device's own inverse lumped mass), on
// the inputs of the last flgpu_ct_assemble / flgpu_cmc_assemble (host
// copies and checks excluded).
int time_pressure(flgpu_ctx* c, int reps, double* ct_ms, double* cmc_ms) {
    PressureState* S = c->ps;
    if (!S || !S->assembled || S->cmc_nnz < 0)
        return set_error(FLGPU_EINVAL, "flgpu_time_pressure: assemble ct_m and C^T M^-1 C once first");
    hipEvent_t e0, e1, e2;
    FLGPU_HIP(hipEventCreate(&e0));
    FLGPU_HIP(hipEventCreate(&e1));
    FLGPU_HIP(hipEventCreate(&e2));
    int rc = FLGPU_OK;
    if (hipEventRecord(e0, c->stream) != hipSuccess) rc = set_error(FLGPU_EHIP, "flgpu_time_pressure: event");
    for (int r = 0; r < reps && rc == FLGPU_OK; r++) rc = ct_launch(c, *S, S->last_ibp, S->last_rho_const);
    if (rc == FLGPU_OK && hipEventRecord(e1, c->stream) != hipSuccess)
        rc = set_error(FLGPU_EHIP, "flgpu_time_pressure: event");
    for (int r = 0; r < reps && rc == FLGPU_OK; r++) rc = cmc_launch(c, *S, true);
    if (rc == FLGPU_OK && hipEventRecord(e2, c->stream) != hipSuccess)
        rc = set_error(FLGPU_EHIP, "flgpu_time_pressure: event");
    if (rc == FLGPU_OK && hipEventSynchronize(e2) != hipSuccess)
        rc = set_error(FLGPU_EHIP, "flgpu_time_pressure: synchronize");
    float a = 0.f, b = 0.f;
    if (rc == FLGPU_OK && (hipEventElapsedTime(&a, e0, e1) != hipSuccess ||
                           hipEventElapsedTime(&b, e1, e2) != hipSuccess))
        rc = set_error(FLGPU_EHIP, "flgpu_time_pressure: elapsed time");
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    (void)hipEventDestroy(e2);
    if (rc == FLGPU_OK) {
        *ct_ms = a / reps;
        *cmc_ms = b / reps;
    }
    return rc;
}

void pressure_free(flgpu_ctx* c) {
    delete c->ps;
    c->ps = nullptr;
}

}  // namespace flgpu

using namespace flgpu;

extern "C" int flgpu_pressure_bind(flgpu_ctx* c, int32_t p_loc, int32_t p_nodes, const int32_t* p_ndglno,
                                   const double* p_n, const double* p_dn, int32_t ct_rows, const int32_t* ct_findrm,
                                   const int32_t* ct_colm) {
    return guarded("flgpu_pressure_bind", [&]() -> int {
        if (!c || !p_ndglno || !p_n || !p_dn || !ct_findrm || !ct_colm)
            return set_error(FLGPU_EINVAL, "flgpu_pressure_bind: null argument");
        if (c->nele <= 0 || c->loc <= 0 || c->ncol <= 0)
            return set_error(FLGPU_EINVAL, "flgpu_pressure_bind: bind the meshes and the colouring first");
        const int dim = c->dim, ngi = c->ngi;
        if (dim < 2) return set_error(FLGPU_EUNSUPPORTED, "flgpu_pressure_bind: 2-D and 3-D meshes only");
        if (p_loc < 1 || p_loc > kMaxLoc)
            return set_error(FLGPU_EUNSUPPORTED, "flgpu_pressure_bind: p_loc %d", p_loc);
        if (p_nodes < 1 || ct_rows != p_nodes)
            return set_error(FLGPU_EINVAL, "flgpu_pressure_bind: %d ct rows for %d pressure nodes", ct_rows, p_nodes);
        if (ct_findrm[0] != 1) return set_error(FLGPU_EINVAL, "flgpu_pressure_bind: ct findrm(1) = %d", ct_findrm[0]);
        FLGPU_HIP(hipSetDevice(c->device));
        std::vector<int32_t> pc((size_t)c->nele * p_loc);
        for (size_t k = 0; k < pc.size(); k++) {
            const int v = p_ndglno[k];
            if (v < 1 || v > p_nodes)
                return set_error(FLGPU_EINVAL, "flgpu_pressure_bind: p_ndglno[%zu]=%d out of 1..%d", k, v, p_nodes);
            pc[k] = v - 1;
        }
        std::vector<int32_t> rp(p_nodes + 1);
        for (int i = 0; i <= p_nodes; i++) {
            rp[i] = ct_findrm[i] - 1;
            if (i > 0 && rp[i] < rp[i - 1])
                return set_error(FLGPU_EINVAL, "flgpu_pressure_bind: ct findrm decreases at %d", i);
        }
        const int64_t nnz = rp[p_nodes];
        std::vector<int32_t> cl((size_t)nnz);
        for (int i = 0; i < p_nodes; i++)
            for (int k = rp[i]; k < rp[i + 1]; k++) {
                const int v = ct_colm[k] - 1;
                if (v < 0 || v >= c->nnodes)
                    return set_error(FLGPU_EINVAL, "flgpu_pressure_bind: ct colm(%d) = %d out of 1..%d", k + 1,
                                     ct_colm[k], c->nnodes);
                if (k > rp[i] && v <= cl[k - 1])
                    return set_error(FLGPU_EINVAL, "flgpu_pressure_bind: ct row %d not sorted", i + 1);
                cl[k] = v;
            }
        PTables t;
        std::memset(&t, 0, sizeof t);
        t.p_loc = p_loc;
        t.ngi = ngi;
        for (int k = 0; k < p_loc * ngi; k++) t.n[k] = p_n[k];
        for (int k = 0; k < p_loc * ngi * dim; k++) t.dn[k] = p_dn[k];
        if (!c->ps) c->ps = new PressureState();
        PressureState& S = *c->ps;
        FLGPU_TRY(S.p_conn.upload(pc.data(), pc.size(), c->stream));
        FLGPU_TRY(S.ct_rowptr.upload(rp.data(), rp.size(), c->stream));
        FLGPU_TRY(S.ct_col.upload(cl.data(), cl.size(), c->stream));
        FLGPU_TRY(S.d_tab.upload(&t, 1, c->stream));
        FLGPU_HIP(hipStreamSynchronize(c->stream));
        S.h_p_conn.swap(pc);
        S.h_ct_rowptr.swap(rp);
        S.tab = t;
        S.p_loc = p_loc;
        S.p_nodes = p_nodes;
        S.ct_nnz = nnz;
        S.plan_key = -1;
        S.assembled = false;
        S.bound = true;
        return FLGPU_OK;
    });
}

extern "C" int flgpu_time_pressure(flgpu_ctx* c, int32_t reps, double* ct_ms, double* cmc_ms) {
    return flgpu::guarded("flgpu_time_pressure", [&]() -> int {
        if (!c || !ct_ms || !cmc_ms || reps < 1) return flgpu::set_error(FLGPU_EINVAL, "flgpu_time_pressure: bad argument");
        FLGPU_HIP(hipSetDevice(c->device));
        return flgpu::time_pressure(c, reps, ct_ms, cmc_ms);
    });
}

extern "C" int flgpu_ct_assemble(flgpu_ctx* c, int32_t integrate_by_parts, int32_t density_constant,
                                 const double* density, double* ct_val, double* masslump) {
    return guarded("flgpu_ct_assemble", [&]() -> int {
        if (!c || !density) return set_error(FLGPU_EINVAL, "flgpu_ct_assemble: null argument");
        if (!c->ps || !c->ps->bound) return set_error(FLGPU_EINVAL, "flgpu_ct_assemble: bind the pressure mesh first");
        PressureState& S = *c->ps;
        if (S.h_p_conn.size() != (size_t)c->nele * S.p_loc || c->h_col_elems.size() != (size_t)c->nele)
            return set_error(FLGPU_EINVAL, "flgpu_ct_assemble: the pressure mesh was bound for another mesh");
        if (c->ngi != S.tab.ngi)
            return set_error(FLGPU_EINVAL, "flgpu_ct_assemble: pressure tables on %d points, velocity on %d",
                             S.tab.ngi, c->ngi);
        FLGPU_HIP(hipSetDevice(c->device));
        return ct_assemble(c, S, integrate_by_parts != 0, density_constant != 0, density, ct_val, masslump);
    });
}

extern "C" int flgpu_cmc_assemble(flgpu_ctx* c, const double* inverse_masslump, int32_t cmc_rows,
                                  const int32_t* cmc_findrm, const int32_t* cmc_colm, double* cmc_val) {
    return guarded("flgpu_cmc_assemble", [&]() -> int {
        if (!c || !cmc_findrm || !cmc_colm) return set_error(FLGPU_EINVAL, "flgpu_cmc_assemble: null argument");
        if (!c->ps || !c->ps->assembled)
            return set_error(FLGPU_EINVAL, "flgpu_cmc_assemble: assemble ct_m first (flgpu_ct_assemble)");
        FLGPU_HIP(hipSetDevice(c->device));
        return cmc_assemble(c, *c->ps, inverse_masslump, cmc_rows, cmc_findrm, cmc_colm, cmc_val);
    });
}
