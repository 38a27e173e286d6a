// Learned preconditioner on gfx950: a message-passing network on the matrix
// graph predicts a lower-triangular factor G on the IC(0) pattern (the
// lower triangle of A, diagonal included); the preconditioner is
// M^-1 = G^T G, applied as two sparse products that share A's column
// indices (no triangular solves).
//
// Scaling: with D = diag(A), d_i = D_ii^-1/2 and s_ij = d_i A_ij d_j, the
// network predicts Gs (Jacobi = identity) and G = Gs D^-1/2:
//   Gs_ii = exp(psi_i),  Gs_ij = s_ij * phi_ij  (j < i)
//   phi_ij = alpha_i + beta_j + gamma1 s_ij + gamma2 |s_ij| + b_o
// Network (hidden H = 32, 3 layers, residual):
//   h0  = W_in f_i + b_in                     f_i: 4 node features
//   r_ij = relu(P_i + Q_j + c1 s_ij + c2 |s_ij| + b_m),  [P|Q] = h [Wa|Wb]^T
//   m_i = sum_{j != i} r_ij
//   h'  = h + relu(Wu [h; m] + b_u)
//   alpha = h.w_a, beta = h.w_b, psi = h.w_d + b_d
// The node products run on BF16 MFMA (32x32x16, FP32 accumulation); the
// weight gradients are MFMA reductions over the nodes.  Training loss:
//   L = 1/(N P) sum_p || z_p - G^T G A z_p ||^2 over P Rademacher probes.
// Every reduction uses a fixed partition and order: runs repeat bit for bit.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <string>
#include <vector>

#include "flgpu_internal.h"
#include "krylov_common.h"
#include "learned.h"

namespace flgpu {

constexpr int H = 32;        // hidden width
constexpr int NFEAT = 4;     // node features
constexpr int NLAYER = 3;
constexpr int MAXPROBE = 16;

// ---- parameter layout -----------------------------------------------------
struct Lay {
    static constexpr int W_IN = 0;                       // [H][NFEAT]
    static constexpr int B_IN = W_IN + H * NFEAT;        // [H]
    static constexpr int L0 = B_IN + H;
    // per layer (relative)
    static constexpr int WA = 0;                         // [H][H]
    static constexpr int WB = WA + H * H;                // [H][H]
    static constexpr int C1 = WB + H * H;                // [H]
    static constexpr int C2 = C1 + H;                    // [H]
    static constexpr int BM = C2 + H;                    // [H]
    static constexpr int WU = BM + H;                    // [H][2H]
    static constexpr int BU = WU + H * 2 * H;            // [H]
    static constexpr int LSIZE = BU + H;
    static constexpr int OUT = L0 + NLAYER * LSIZE;
    static constexpr int WOA = OUT;                      // [H]
    static constexpr int WOB = WOA + H;                  // [H]
    static constexpr int WD = WOB + H;                   // [H]
    static constexpr int G1 = WD + H;                    // scalar
    static constexpr int G2 = G1 + 1;
    static constexpr int BO = G2 + 1;
    static constexpr int BD = BO + 1;
    static constexpr int TOTAL = BD + 1;
    static constexpr int layer(int l) { return L0 + l * LSIZE; }
};

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x16 __attribute__((ext_vector_type(16)));

struct LearnedState {
    DevBuf<float> theta, grad, am, av;
    int64_t step = 0;
    bool have_params = false;
    DevBuf<__bf16> wbf;   // bf16 copies: per layer [Wa;Wb] (2H x H), Wu (H x 2H),
                          // Wu^T (2H x H), [Wa|Wb]^T stacked for backward (H x 2H)
    // graph
    int64_t graph_key = -1;
    int n = 0;
    int64_t nnz = 0;
    DevBuf<int32_t> tmap;
    DevBuf<int32_t> fsai_rows;  // FSAI: the rows longer than the first kernel takes, then their count
    DevBuf<float> dsq, sval, feat;
    // activations (training keeps every layer)
    DevBuf<float> h, PQ, M, U, ab;  // ab: alpha,beta,psi per node (3N)
    DevBuf<float> gv;               // preconditioner values in A's pattern
    DevBuf<double> gv64;
    // solver copies of the factor: the lower part (j <= i, with the
    // diagonal) and the upper part (j >= i) as two compact CSR matrices,
    // so that each product streams only its own half
    int64_t lu_key = -1;
    DevBuf<int32_t> lptr, lcol, lsrc, uptr, ucol, usrc, rbl, rbu, rbkl, rbku;
    DevBuf<int16_t> lcol16, ucol16;  // int16 column offsets (when they fit)
    bool l16 = false, u16 = false;
    int nrbl = 0, nrbu = 0;
    int64_t nnzl = 0, nnzu = 0;
    // the halves as sliced ELL operands (when both pad little): the value
    // arrays gl/gu then hold slots, gathered from the factor through lss/uss
    bool sell = false, ls16 = false, us16 = false;
    int nsll = 0, nslu = 0;
    int64_t nslotl = 0, nslotu = 0;
    DevBuf<int32_t> lsp, usp, lss, uss, lsc32, usc32;
    DevBuf<int16_t> lsc16, usc16;
    // factor of the scaled matrix, Gs: FP16 for the mixed solve, FP64 for
    // the FP64 solve; the scaling D is applied by the products
    DevBuf<_Float16> gl16, gu16;
    DevBuf<double> gl64, gu64;
    // training
    DevBuf<float> z, u, v, w, dv, dw, dphi, dpsi, dab, dh, dHM, dU, dPQ, dM;
    DevBuf<float> wpart;            // weight-gradient partial tiles
    DevBuf<double> lpart;           // loss partials
    DevBuf<double> lossbuf;
};

static inline int params_count() { return Lay::TOTAL; }

// ---------------------------------------------------------------------------
// helpers
// ---------------------------------------------------------------------------

__device__ __forceinline__ uint64_t smix(uint64_t x) {
    x += 0x9E3779B97F4A7C15ull;
    x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
    x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
    return x ^ (x >> 31);
}

__global__ void k_to_bf16(int n, const float* __restrict__ src, __bf16* __restrict__ dst) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) dst[i] = (__bf16)src[i];
}

// Graph preparation from the reduced system: d_i = a_ii^(-1/2).
template <typename V>
__global__ void __launch_bounds__(256)
k_graph_prep(int n, const int32_t* __restrict__ rowptr, const int32_t* __restrict__ col,
             const int32_t* __restrict__ diag, const V* __restrict__ val, float* __restrict__ dsq,
             float* __restrict__ sval, float* __restrict__ feat, int pass) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int d = diag[i];
    const double a = d >= 0 ? (double)val[d] : 1.0;
    dsq[i] = (float)(a > 0 ? 1.0 / sqrt(a) : 1.0);
}

// The scaled values s_ij = d_i a_ij d_j and the node features (row sum of
// s, off-diagonal count / 16, largest off-diagonal |s|, no-neighbour flag),
// half a wave per row: entries read and written coalesced, the row sums
// formed in entry order.
template <typename V>
__global__ void __launch_bounds__(256)
k_graph_prep_rows(int n, const int32_t* __restrict__ rowptr, const int32_t* __restrict__ col,
                  const V* __restrict__ val, const float* __restrict__ dsq, float* __restrict__ sval,
                  float* __restrict__ feat) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    const int i = t >> 5, c = t & 31;
    if (i >= n) return;
    const float di = dsq[i];
    float rowsum = 0.f, maxoff = 0.f;
    int nb = 0;
    const int kb = rowptr[i], ke = rowptr[i + 1];
    for (int k0 = kb; k0 < ke; k0 += 32) {
        const int len = min(32, ke - k0);
        int j = i;
        float s = 0.f;
        if (c < len) {
            j = col[k0 + c];
            s = (float)((double)val[k0 + c] * (double)di * (double)dsq[j]);
            sval[k0 + c] = s;
        }
        for (int q = 0; q < len; q++) {
            const int jq = __shfl(j, q, 32);
            const float sq = __shfl(s, q, 32);
            rowsum += sq;
            if (jq != i && sq != 0.f) {
                nb++;
                maxoff = fmaxf(maxoff, fabsf(sq));
            }
        }
    }
    if (c == 0) {
        feat[(size_t)i * NFEAT + 0] = rowsum;
        feat[(size_t)i * NFEAT + 1] = nb / 16.f;
        feat[(size_t)i * NFEAT + 2] = maxoff;
        feat[(size_t)i * NFEAT + 3] = nb == 0 ? 1.f : 0.f;
    }
}

// tmap[k]: position of the transposed entry (col[k], row) in row col[k]
__global__ void __launch_bounds__(256)
k_tmap(int n, const int32_t* __restrict__ rowptr, const int32_t* __restrict__ col, int32_t* __restrict__ tmap) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    for (int k = rowptr[i]; k < rowptr[i + 1]; k++) {
        const int j = col[k];
        int lo = rowptr[j], hi = rowptr[j + 1] - 1, pos = -1;
        while (lo <= hi) {
            const int mid = (lo + hi) >> 1;
            const int v = col[mid];
            if (v == i) {
                pos = mid;
                break;
            }
            if (v < i) lo = mid + 1;
            else hi = mid - 1;
        }
        tmap[k] = pos;
    }
}

// h0 = W_in f + b_in
__global__ void __launch_bounds__(256)
k_input(int n, const float* __restrict__ feat, const float* __restrict__ th, float* __restrict__ h0) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n * H) return;
    const int i = t / H, c = t % H;
    float s = th[Lay::B_IN + c];
    for (int f = 0; f < NFEAT; f++) s = fmaf(th[Lay::W_IN + c * NFEAT + f], feat[(size_t)i * NFEAT + f], s);
    h0[t] = s;
}

// ---------------------------------------------------------------------------
// MFMA node product: Y[n][o] = act(sum_k X[n][k] W[o][k] + b[o]) (+ R[n][o])
// X = [X1 (d1 cols) | X2 (d2 cols)], K = d1 + d2 (multiple of 16),
// O = DOUT (multiple of 32).  One wave computes a 32-node x 32-output tile.
// Operand layout of v_mfma_f32_32x32x16_bf16: lane l holds A[l%32][8*(l/32)+t]
// and B[8*(l/32)+t][l%32]; result r of lane l is D[8*(r/4)+4*(l/32)+r%4][l%32].
// ---------------------------------------------------------------------------
template <int ACT>  // 0 none, 1 relu-residual (Y = R + relu(.)), stores pre-activation in Ypre
__global__ void __launch_bounds__(256)
k_node_gemm(int n, int d1, int d2, const float* __restrict__ X1, const float* __restrict__ X2, int dout,
            const __bf16* __restrict__ W, const float* __restrict__ bias, const float* __restrict__ R,
            float* __restrict__ Y, float* __restrict__ Ypre) {
    const int wave = (blockIdx.x * blockDim.x + threadIdx.x) >> 6;
    const int lane = threadIdx.x & 63;
    const int otiles = dout / 32;
    const int ntile = wave / otiles, otile = wave % otiles;
    const int n0 = ntile * 32, o0 = otile * 32;
    if (n0 >= n) return;
    const int K = d1 + d2;
    const int row = lane & 31, kg = lane >> 5;
    f32x16 acc = {};
    const int nr = n0 + row;
    for (int k0 = 0; k0 < K; k0 += 16) {
        // 8 consecutive k of one node row (d1, d2 are multiples of 16, so a
        // run never straddles X1 and X2): two 16-byte loads; 8 weights: one
        const int k = k0 + 8 * kg;
        float4 p0 = make_float4(0.f, 0.f, 0.f, 0.f), p1 = p0;
        if (nr < n) {
            const float* xr = k < d1 ? X1 + (size_t)nr * d1 + k : X2 + (size_t)nr * d2 + (k - d1);
            p0 = *reinterpret_cast<const float4*>(xr);
            p1 = *reinterpret_cast<const float4*>(xr + 4);
        }
        const bf16x8 b = *reinterpret_cast<const bf16x8*>(W + (size_t)(o0 + row) * K + k);
        bf16x8 a;
        a[0] = (__bf16)p0.x;
        a[1] = (__bf16)p0.y;
        a[2] = (__bf16)p0.z;
        a[3] = (__bf16)p0.w;
        a[4] = (__bf16)p1.x;
        a[5] = (__bf16)p1.y;
        a[6] = (__bf16)p1.z;
        a[7] = (__bf16)p1.w;
        acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, b, acc, 0, 0, 0);
    }
#pragma unroll
    for (int r = 0; r < 16; r++) {
        const int i = 8 * (r / 4) + 4 * kg + (r % 4);
        const int nn = n0 + i, o = o0 + row;
        if (nn >= n) continue;
        float v = acc[r] + (bias ? bias[o] : 0.f);
        if (ACT == 1) {
            Ypre[(size_t)nn * dout + o] = v;
            v = R[(size_t)nn * dout + o] + fmaxf(v, 0.f);
        }
        Y[(size_t)nn * dout + o] = v;
    }
}

// MFMA weight gradient: part[blk][o][k] = sum_{n in chunk} dY[n][o] X[n][k]
// with X = [X1 | X2] (K cols), O = dout.  One block per node chunk; every
// wave of the block owns a (o,k) tile and walks the chunk.
__global__ void __launch_bounds__(256)
k_wgrad(int n, int chunk, int d1, int d2, const float* __restrict__ X1, const float* __restrict__ X2,
        int dout, const float* __restrict__ dY, float* __restrict__ part) {
    const int K = d1 + d2;
    const int otiles = dout / 32, ktiles = K / 32;
    const int ntile = otiles * ktiles;
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63, nw = blockDim.x >> 6;
    const int nb0 = blockIdx.x * chunk, nb1 = min(nb0 + chunk, n);
    const int row = lane & 31, kg = lane >> 5;
    for (int tile = wave; tile < ntile; tile += nw) {
        const int o0 = (tile / ktiles) * 32, k0 = (tile % ktiles) * 32;
        f32x16 acc = {};
        for (int s = nb0; s < nb1; s += 16) {
            bf16x8 a, b;
#pragma unroll
            for (int t = 0; t < 8; t++) {
                const int nn = s + 8 * kg + t;
                float av = 0.f, bv = 0.f;
                if (nn < nb1) {
                    av = dY[(size_t)nn * dout + o0 + row];
                    const int k = k0 + row;
                    bv = k < d1 ? X1[(size_t)nn * d1 + k] : X2[(size_t)nn * d2 + (k - d1)];
                }
                a[t] = (__bf16)av;
                b[t] = (__bf16)bv;
            }
            acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, b, acc, 0, 0, 0);
        }
        float* out = part + (size_t)blockIdx.x * dout * K;
#pragma unroll
        for (int r = 0; r < 16; r++) {
            const int i = 8 * (r / 4) + 4 * kg + (r % 4);
            out[(size_t)(o0 + i) * K + k0 + row] = acc[r];
        }
    }
}

// g[e] += sum_blk part[blk][e]  (fixed order)
__global__ void k_wgrad_reduce(int nblk, int nel, const float* __restrict__ part, float* __restrict__ g) {
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= nel) return;
    float s = 0.f;
    for (int b = 0; b < nblk; b++) s += part[(size_t)b * nel + e];
    g[e] += s;
}

// column sums: g[o] += sum_n dY[n][o]  (fixed chunks, then fixed order)
__global__ void k_colsum(int n, int dout, int chunk, const float* __restrict__ dY, float* __restrict__ part) {
    const int o = threadIdx.x;
    if (o >= dout) return;
    const int n0 = blockIdx.x * chunk, n1 = min(n0 + chunk, n);
    float s = 0.f;
    for (int i = n0; i < n1; i++) s += dY[(size_t)i * dout + o];
    part[(size_t)blockIdx.x * dout + o] = s;
}

// ---------------------------------------------------------------------------
// message passing: M_i = sum_{j != i} relu(P_i + Q_j + c1 s + c2 |s| + bm)
// 32 lanes per row (one per channel)
// ---------------------------------------------------------------------------
// Half a wave per row (lane = channel).  The row's columns and values are
// read once, coalesced, and handed to the lanes with shuffles, so the
// gathers of Q_j for the row's edges are independent and overlap; the sum
// keeps the row's entry order.  Rows are dealt to the XCDs in contiguous
// runs so that neighbouring rows' Q_j meet in one L2 (the grid must be a
// multiple of 8 for the mapping to cover every row).
__global__ void __launch_bounds__(256)
k_edge_agg(int n, const int32_t* __restrict__ rowptr, const int32_t* __restrict__ col,
           const float* __restrict__ sval, const float* __restrict__ PQ, const float* __restrict__ lp,
           float* __restrict__ M) {
    const int per = (int)((gridDim.x + 7) >> 3);
    const int blk = (int)(blockIdx.x & 7) * per + (int)(blockIdx.x >> 3);
    const int i = (blk * 256 + (int)threadIdx.x) >> 5, c = threadIdx.x & 31;
    if (i >= n) return;
    const float pi = PQ[(size_t)i * 2 * H + c];
    const float c1 = lp[Lay::C1 + c], c2 = lp[Lay::C2 + c], bm = lp[Lay::BM + c];
    const int kb = rowptr[i], ke = rowptr[i + 1];
    float m = 0.f;
    constexpr int U = 16;
    for (int k0 = kb; k0 < ke; k0 += 32) {
        const int len = min(32, ke - k0);
        const int myj = c < len ? col[k0 + c] : i;
        const float mys = c < len ? sval[k0 + c] : 0.f;
        // U gathers of Q_j in flight per step (a P1 tet row in one step):
        // the loads are unconditional
        // (a step past the row's end re-reads its last edge) and only the
        // additions are predicated, in entry order
        for (int t0 = 0; t0 < len; t0 += U) {
            int jj[U];
            float sv[U], q[U];
#pragma unroll
            for (int u = 0; u < U; u++) {
                const int t = min(t0 + u, len - 1);
                jj[u] = __shfl(myj, t, 32);
                sv[u] = __shfl(mys, t, 32);
            }
#pragma unroll
            for (int u = 0; u < U; u++) q[u] = PQ[(size_t)jj[u] * 2 * H + H + c];
#pragma unroll
            for (int u = 0; u < U; u++) {
                const float e = fmaxf(pi + q[u] + c1 * sv[u] + c2 * fabsf(sv[u]) + bm, 0.f);
                if (t0 + u < len && jj[u] != i) m += e;
            }
        }
    }
    M[(size_t)i * H + c] = m;
}

// alpha, beta, psi per node
__global__ void __launch_bounds__(256)
k_out_node(int n, const float* __restrict__ h, const float* __restrict__ th, float* __restrict__ ab) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    float a = 0.f, b = 0.f, p = th[Lay::BD];
    for (int c = 0; c < H; c++) {
        const float hv = h[(size_t)i * H + c];
        a = fmaf(th[Lay::WOA + c], hv, a);
        b = fmaf(th[Lay::WOB + c], hv, b);
        p = fmaf(th[Lay::WD + c], hv, p);
    }
    ab[3 * (size_t)i + 0] = a;
    ab[3 * (size_t)i + 1] = b;
    ab[3 * (size_t)i + 2] = p;
}

__device__ __forceinline__ float gs_lower(float s, float ai, float bj, const float* th) {
    return s * (ai + bj + th[Lay::G1] * s + th[Lay::G2] * fabsf(s) + th[Lay::BO]);
}

// Preconditioner values in A's pattern: lower slots hold Gs_ij d_j, upper
// slots hold d_i Gs_ji, the diagonal Gs_ii d_i; with scaled != 0 the
// factor of the scaled matrix alone (Gs_ij, Gs_ji, Gs_ii).
template <typename GV>
__global__ void __launch_bounds__(256)
k_build_g(int n, const int32_t* __restrict__ rowptr, const int32_t* __restrict__ col,
          const int32_t* __restrict__ tmap, const float* __restrict__ sval, const float* __restrict__ dsq,
          const float* __restrict__ ab, const float* __restrict__ th, GV* __restrict__ gv, int scaled, int upper) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    const int i = t >> 5, lane = t & 31;  // half a wave per row: coalesced entries
    if (i >= n) return;
    const float di = scaled ? 1.f : dsq[i];
    for (int k = rowptr[i] + lane; k < rowptr[i + 1]; k += 32) {
        const int j = col[k];
        float g;
        if (j < i) {
            g = gs_lower(sval[k], ab[3 * (size_t)i], ab[3 * (size_t)j + 1], th) * (scaled ? 1.f : dsq[j]);
        } else if (j == i) {
            g = expf(ab[3 * (size_t)i + 2]) * di;
        } else {
            if (!upper) continue;
            const int kt = tmap[k];
            const float s = kt >= 0 ? sval[kt] : 0.f;
            g = di * gs_lower(s, ab[3 * (size_t)j], ab[3 * (size_t)i + 1], th);
        }
        gv[k] = (GV)g;
    }
}

// ---------------------------------------------------------------------------
// The exact factorised sparse approximate inverse (FSAI) on the same
// pattern: the ceiling of the network's factor family, and a preconditioner
// of its own.  Row i of Gs, on the lower pattern P_i of row i of
// S = D A D (columns j <= i, the diagonal last), minimises the Kaporin
// condition of Gs S Gs^T: with S[P_i, P_i] = L L^T (Cholesky),
// Gs[i, P_i] = L^-T e_last, so that (Gs S Gs^T)_ii = 1 and
// (Gs S)_ij = 0 for the other j in P_i (Kolotilina & Yeremin 1993).
// ---------------------------------------------------------------------------
constexpr int kFsaiMax = 48;   // largest lower pattern solved; longer rows keep the Jacobi row
constexpr int kFsaiShort = 8;  // rows up to this long are solved in registers (P1 tetrahedra: all)

// Rows whose lower pattern has at most K entries, the local matrix in
// registers: every index below is a compile-time constant after
// unrolling.  The pattern is padded to K with the identity (P[a] = i for
// a >= k): the padded Cholesky factor is L extended by an identity block
// and the padded solve gives g = 0 there, so no step needs the row's own
// length.  Longer rows are listed for k_fsai_lower.
template <typename GV, int K>
__global__ void __launch_bounds__(256)
k_fsai_short(int n, const int32_t* __restrict__ rowptr, const int32_t* __restrict__ col,
             const int32_t* __restrict__ diag, const double* __restrict__ val, const float* __restrict__ dsq,
             GV* __restrict__ gv, int32_t* __restrict__ long_rows, int32_t* __restrict__ nlong);

template <typename GV>
__device__ inline void fsai_jacobi_row(int k0, int kd, GV* gv) {
    for (int q = k0; q < kd; q++) gv[q] = (GV)0;
    gv[kd] = (GV)1;  // S_ii = 1
}

// The listed long rows (long_rows[0 .. n-1], at most KMAX lower entries;
// longer ones keep the Jacobi row), one thread per row, NT threads per
// block: the local matrix S[P_i, P_i] in the block's LDS, entry-major so
// that a block's threads touch consecutive words, formed in FP64 from the
// reduced FP64 values and D.  Writes Gs at the row's lower positions of
// A's pattern, the diagonal included.
template <typename GV, int KMAX, int NT>
__global__ void __launch_bounds__(NT)
k_fsai_lower(int n, const int32_t* __restrict__ rowptr, const int32_t* __restrict__ col,
             const int32_t* __restrict__ diag, const double* __restrict__ val, const float* __restrict__ dsq,
             GV* __restrict__ gv, const int32_t* __restrict__ long_rows) {
    __shared__ double sl[KMAX * (KMAX + 1) / 2][NT];
    __shared__ double sg[KMAX][NT];
    const int idx = blockIdx.x * NT + threadIdx.x;
    if (idx >= n) return;
    const int i = long_rows[idx];
    const int t = threadIdx.x;
    const int k0 = rowptr[i], kd = diag[i];
    const int k = kd - k0 + 1;  // lower pattern size, the diagonal last (columns sorted)
    if (k > KMAX) {
        fsai_jacobi_row(k0, kd, gv);
        return;
    }
#define B(a, b) sl[(a) * ((a) + 1) / 2 + (b)][t]
    // B(a, b) = S[P_a, P_b] for b <= a: row P_a of A merged with P
    for (int a = 0; a < k; a++) {
        const int pa = col[k0 + a];
        const double da = (double)dsq[pa];
        int q = rowptr[pa];
        const int qe = rowptr[pa + 1];
        for (int b = 0; b <= a; b++) {
            const int pb = col[k0 + b];
            while (q < qe && col[q] < pb) q++;
            B(a, b) = (q < qe && col[q] == pb) ? val[q] * da * (double)dsq[pb] : 0.0;
        }
    }
    // Cholesky B = L L^T in place
    for (int a = 0; a < k; a++) {
        double d = B(a, a);
        for (int c = 0; c < a; c++) d -= B(a, c) * B(a, c);
        if (!(d > 0.0)) {  // not positive definite
            fsai_jacobi_row(k0, kd, gv);
            return;
        }
        const double l = sqrt(d);
        B(a, a) = l;
        for (int r = a + 1; r < k; r++) {
            double v = B(r, a);
            for (int c = 0; c < a; c++) v -= B(r, c) * B(a, c);
            B(r, a) = v / l;
        }
    }
    // L^T g = e_last, backwards
    for (int a = k - 1; a >= 0; a--) {
        double v = a == k - 1 ? 1.0 : 0.0;
        for (int r = a + 1; r < k; r++) v -= B(r, a) * sg[r][t];
        sg[a][t] = v / B(a, a);
    }
#undef B
    for (int a = 0; a < k; a++) gv[k0 + a] = (GV)sg[a][t];
}

template <typename GV, int K>
__global__ void __launch_bounds__(256)
k_fsai_short(int n, const int32_t* __restrict__ rowptr, const int32_t* __restrict__ col,
             const int32_t* __restrict__ diag, const double* __restrict__ val, const float* __restrict__ dsq,
             GV* __restrict__ gv, int32_t* __restrict__ long_rows, int32_t* __restrict__ nlong) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int k0 = rowptr[i], kd = diag[i];
    const int k = kd - k0 + 1;
    if (k > K) {
        long_rows[atomicAdd(nlong, 1)] = i;
        return;
    }
    int P[K];
    double dp[K];
#pragma unroll
    for (int a = 0; a < K; a++) {
        P[a] = a < k ? col[k0 + a] : i;
        dp[a] = (double)dsq[P[a]];
    }
    // B(a, b) = S[P_a, P_b] (b <= a): entry P_b of row P_a lies in that
    // row's lower part; a lower part of at most K entries is loaded whole
    // (its loads all in flight) and matched against P by comparisons, a
    // longer one searched entry by entry
    double B[K * (K + 1) / 2];
#pragma unroll
    for (int a = 0; a < K; a++) {
        const int q0 = rowptr[P[a]], ql = diag[P[a]] + 1;
        if (a < k && ql - q0 <= K) {
            int lc[K];
            double lv[K];
#pragma unroll
            for (int q = 0; q < K; q++) {
                lc[q] = q0 + q < ql ? col[q0 + q] : -1;
                lv[q] = q0 + q < ql ? val[q0 + q] : 0.0;
            }
#pragma unroll
            for (int b = 0; b <= a; b++) {
                double v = 0.0;
#pragma unroll
                for (int q = 0; q < K; q++) v = lc[q] == P[b] ? lv[q] : v;
                B[a * (a + 1) / 2 + b] = v * dp[a] * dp[b];
            }
        } else {
            int q = q0;
#pragma unroll
            for (int b = 0; b <= a; b++) {
                double v = (a == b) ? 1.0 : 0.0;  // the identity padding
                if (a < k) {
                    while (q < ql && col[q] < P[b]) q++;
                    v = (q < ql && col[q] == P[b]) ? val[q] * dp[a] * dp[b] : 0.0;
                }
                B[a * (a + 1) / 2 + b] = v;
            }
        }
    }
    bool pd = true;
#pragma unroll
    for (int a = 0; a < K; a++) {
        double d = B[a * (a + 1) / 2 + a];
#pragma unroll
        for (int c = 0; c < a; c++) d -= B[a * (a + 1) / 2 + c] * B[a * (a + 1) / 2 + c];
        pd = pd && d > 0.0;
        const double l = sqrt(d > 0.0 ? d : 1.0);
        B[a * (a + 1) / 2 + a] = l;
#pragma unroll
        for (int r = a + 1; r < K; r++) {
            double v = B[r * (r + 1) / 2 + a];
#pragma unroll
            for (int c = 0; c < a; c++) v -= B[r * (r + 1) / 2 + c] * B[a * (a + 1) / 2 + c];
            B[r * (r + 1) / 2 + a] = v / l;
        }
    }
    if (!pd) {  // not positive definite: the Jacobi row
        fsai_jacobi_row(k0, kd, gv);
        return;
    }
    double g[K];
#pragma unroll
    for (int a = K - 1; a >= 0; a--) {
        double v = a == k - 1 ? 1.0 : 0.0;
#pragma unroll
        for (int r = a + 1; r < K; r++) v -= B[r * (r + 1) / 2 + a] * g[r];
        g[a] = v / B[a * (a + 1) / 2 + a];
    }
#pragma unroll
    for (int a = 0; a < K; a++)
        if (a < k) gv[k0 + a] = (GV)g[a];
}

// gathers the factor values of one half from A's pattern
template <typename GV, typename HV>
__global__ void k_gather_half(int64_t m, const int32_t* __restrict__ src, const GV* __restrict__ gv,
                              HV* __restrict__ out) {
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t < m) {
        const int32_t k = src[t];
        out[t] = k >= 0 ? (HV)gv[k] : (HV)0;
    }
}

// ---------------------------------------------------------------------------
// training: batched products over P probes (probe-minor layout [N][P])
// ---------------------------------------------------------------------------
__global__ void k_probes(int n, int P, uint64_t seed, float* __restrict__ z) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n * P) return;
    z[t] = (smix(seed ^ (uint64_t)t) >> 63) ? 1.f : -1.f;
}

// y = A x over P probes; part selects the CSR range: 0 full row, 1 lower
// (k <= diag), 2 upper (k >= diag).  Values: vals (double A) or gv (float).
template <typename VT>
__global__ void __launch_bounds__(256)
k_batch_spmv(int n, int P, const int32_t* __restrict__ rowptr, const int32_t* __restrict__ col,
             const int32_t* __restrict__ diag, const VT* __restrict__ val, const float* __restrict__ x,
             float* __restrict__ y, int part) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    const int i = t / P, p = t % P;
    if (i >= n) return;
    const int b = part == 2 ? diag[i] : rowptr[i];
    const int e = part == 1 ? diag[i] + 1 : rowptr[i + 1];
    float acc = 0.f;
    for (int k = b; k < e; k++) acc = fmaf((float)val[k], x[(size_t)col[k] * P + p], acc);
    y[t] = acc;
}

// res = z - w; loss partials; dw = -2 res / (N P)
__global__ void __launch_bounds__(kVecBlock)
k_loss(int np, float scale, const float* __restrict__ z, const float* __restrict__ w, float* __restrict__ dw,
       double* __restrict__ partials) {
    double s = 0.0;
    for (int t = blockIdx.x * blockDim.x + threadIdx.x; t < np; t += gridDim.x * blockDim.x) {
        const float r = z[t] - w[t];
        s += (double)r * (double)r;
        dw[t] = -2.f * r * scale;
    }
    double v[1] = {s};
    block_partials<1>(v, partials);
}

// One block: loss = scale * sum of the k_loss partials.
__global__ void __launch_bounds__(kVecBlock)
k_loss_fin(int nparts, float scale, const double* __restrict__ partials, double* lossout) {
    double v[1];
    sum_partials<1>(partials, nparts, v);
    if (threadIdx.x == 0) lossout[0] = v[0] * (double)scale;
}

// dGs per stored lower/diagonal entry -> dphi (lower), dpsi (diag)
__global__ void __launch_bounds__(256)
k_dgs(int n, int P, const int32_t* __restrict__ rowptr, const int32_t* __restrict__ col,
      const int32_t* __restrict__ diag, const float* __restrict__ sval, const float* __restrict__ dsq,
      const float* __restrict__ gv, const float* __restrict__ u, const float* __restrict__ v,
      const float* __restrict__ dv, const float* __restrict__ dw, float* __restrict__ dphi,
      float* __restrict__ dpsi) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    for (int k = rowptr[i]; k <= diag[i]; k++) {
        const int j = col[k];
        float s = 0.f;
        for (int p = 0; p < P; p++)
            s += dv[(size_t)i * P + p] * u[(size_t)j * P + p] + dw[(size_t)j * P + p] * v[(size_t)i * P + p];
        const float dg = s * dsq[j];
        if (j < i) {
            dphi[k] = dg * sval[k];
        } else {
            dphi[k] = 0.f;
            dpsi[i] = dg * (gv[k] / dsq[i]);  // dGs_ii * Gs_ii
        }
    }
    for (int k = diag[i] + 1; k < rowptr[i + 1]; k++) dphi[k] = 0.f;
}

// per node: dalpha (row sum of dphi), dbeta (column sum via tmap), and the
// gradient w.r.t. the last hidden layer; partials for gamma1, gamma2, b_o.
__global__ void __launch_bounds__(kVecBlock)
k_out_back(int n, const int32_t* __restrict__ rowptr, const int32_t* __restrict__ col,
           const int32_t* __restrict__ diag, const int32_t* __restrict__ tmap, const float* __restrict__ sval,
           const float* __restrict__ dphi, const float* __restrict__ dpsi, const float* __restrict__ th,
           float* __restrict__ dab, float* __restrict__ dh, double* __restrict__ partials) {
    double g1 = 0.0, g2 = 0.0, gb = 0.0;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        float da = 0.f, db = 0.f;
        for (int k = rowptr[i]; k < diag[i]; k++) {
            const float d = dphi[k], s = sval[k];
            da += d;
            g1 += (double)d * s;
            g2 += (double)d * fabsf(s);
            gb += d;
        }
        for (int k = diag[i] + 1; k < rowptr[i + 1]; k++) {
            const int kt = tmap[k];
            if (kt >= 0) db += dphi[kt];
        }
        const float dp = dpsi[i];
        dab[3 * (size_t)i] = da;
        dab[3 * (size_t)i + 1] = db;
        dab[3 * (size_t)i + 2] = dp;
        for (int c = 0; c < H; c++)
            dh[(size_t)i * H + c] = da * th[Lay::WOA + c] + db * th[Lay::WOB + c] + dp * th[Lay::WD + c];
    }
    double v[3] = {g1, g2, gb};
    block_partials<3>(v, partials);
}

// One block: gradients of gamma1, gamma2, b_o from the k_out_back partials.
__global__ void __launch_bounds__(kVecBlock)
k_out_back_fin(int nparts, const double* __restrict__ partials, float* __restrict__ grad) {
    double v[3];
    sum_partials<3>(partials, nparts, v);
    if (threadIdx.x == 0) {
        grad[Lay::G1] += (float)v[0];
        grad[Lay::G2] += (float)v[1];
        grad[Lay::BO] += (float)v[2];
    }
}

// dU = dh_out * [Upre > 0]
__global__ void k_relu_back(int nt, const float* __restrict__ dho, const float* __restrict__ Upre,
                            float* __restrict__ dU) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t < nt) dU[t] = Upre[t] > 0.f ? dho[t] : 0.f;
}

// edge backward: dP_i = sum_j dr_ij, dQ_j = sum_i dr_ij (gathered over the
// neighbours of j), with dr_ij = dM_i [r_ij > 0]; partials of dc1, dc2, dbm
__global__ void __launch_bounds__(256)
k_edge_back(int n, const int32_t* __restrict__ rowptr, const int32_t* __restrict__ col,
            const int32_t* __restrict__ tmap, const float* __restrict__ sval, const float* __restrict__ PQ,
            const float* __restrict__ lp, const float* __restrict__ dM, float* __restrict__ dPQ,
            float* __restrict__ part3, int chunk_rows) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    const int i = t >> 5, c = t & 31;
    const float c1 = lp[Lay::C1 + c], c2 = lp[Lay::C2 + c], bm = lp[Lay::BM + c];
    float dp = 0.f, dq = 0.f, g1 = 0.f, g2 = 0.f, gb = 0.f;
    if (i < n) {
        const float pi = PQ[(size_t)i * 2 * H + c];
        const float qi = PQ[(size_t)i * 2 * H + H + c];
        const float dmi = dM[(size_t)i * H + c];
        for (int k = rowptr[i]; k < rowptr[i + 1]; k++) {
            const int j = col[k];
            if (j == i) continue;
            const float s = sval[k];
            // edge (i <- j): contributes to dP_i
            const float qj = PQ[(size_t)j * 2 * H + H + c];
            if (pi + qj + c1 * s + c2 * fabsf(s) + bm > 0.f) {
                dp += dmi;
                g1 += dmi * s;
                g2 += dmi * fabsf(s);
                gb += dmi;
            }
            // edge (j <- i): Q_i enters row j's message with s_ji
            const int kt = tmap[k];
            const float sj = kt >= 0 ? sval[kt] : 0.f;
            const float pj = PQ[(size_t)j * 2 * H + c];
            if (kt >= 0 && pj + qi + c1 * sj + c2 * fabsf(sj) + bm > 0.f) dq += dM[(size_t)j * H + c];
        }
        dPQ[(size_t)i * 2 * H + c] = dp;
        dPQ[(size_t)i * 2 * H + H + c] = dq;
    }
    // per-block partials of (c1, c2, bm) gradients: fixed-order block sums
    __shared__ float sh[3][256];
    sh[0][threadIdx.x] = g1;
    sh[1][threadIdx.x] = g2;
    sh[2][threadIdx.x] = gb;
    __syncthreads();
    if (threadIdx.x < 32) {
        float a = 0.f, b = 0.f, d = 0.f;
        for (int q = threadIdx.x; q < (int)blockDim.x; q += 32) {
            a += sh[0][q];
            b += sh[1][q];
            d += sh[2][q];
        }
        part3[((size_t)blockIdx.x * 3 + 0) * H + threadIdx.x] = a;
        part3[((size_t)blockIdx.x * 3 + 1) * H + threadIdx.x] = b;
        part3[((size_t)blockIdx.x * 3 + 2) * H + threadIdx.x] = d;
    }
    (void)chunk_rows;
}

__global__ void k_part3_reduce(int nblk, const float* __restrict__ part3, float* __restrict__ g) {
    const int e = threadIdx.x;  // 0..3H-1
    if (e >= 3 * H) return;
    float s = 0.f;
    for (int b = 0; b < nblk; b++) s += part3[(size_t)b * 3 * H + e];
    g[e] += s;  // layout C1, C2, BM are consecutive
}

// dh_in += dA (N x H) + dB (N x H) elementwise helper: y += x
__global__ void k_axpy1(int nt, const float* __restrict__ x, float* __restrict__ y) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t < nt) y[t] += x[t];
}

__global__ void k_adam(int np, float lr, float b1, float b2, float eps, float bc1, float bc2,
                       float* __restrict__ th, const float* __restrict__ g, float* __restrict__ m,
                       float* __restrict__ v) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= np) return;
    const float gi = g[i];
    const float mi = b1 * m[i] + (1.f - b1) * gi;
    const float vi = b2 * v[i] + (1.f - b2) * gi * gi;
    m[i] = mi;
    v[i] = vi;
    th[i] -= lr * (mi / bc1) / (sqrtf(vi / bc2) + eps);
}


// output-layer weight gradients, per node chunk: part[chunk][q][c] for
// q = 0 (w_a), 1 (w_b), 2 (w_d); part[chunk][3][0] = sum dpsi (b_d)
__global__ void k_outgrad(int n, int chunk, const float* __restrict__ dab, const float* __restrict__ h,
                          float* __restrict__ part) {
    const int t = threadIdx.x;  // 0 .. 4H-1
    const int q = t / H, cc = t % H;
    const int n0 = blockIdx.x * chunk, n1 = min(n0 + chunk, n);
    float s = 0.f;
    if (q < 3) {
        for (int i = n0; i < n1; i++) s += dab[3 * (size_t)i + q] * h[(size_t)i * H + cc];
    } else if (cc == 0) {
        for (int i = n0; i < n1; i++) s += dab[3 * (size_t)i + 2];
    }
    part[(size_t)blockIdx.x * 4 * H + t] = s;
}

__global__ void k_outgrad_reduce(int nchunk, const float* __restrict__ part, float* __restrict__ g) {
    const int t = threadIdx.x;
    float s = 0.f;
    for (int b = 0; b < nchunk; b++) s += part[(size_t)b * 4 * H + t];
    const int q = t / H, cc = t % H;
    if (q == 0) g[Lay::WOA + cc] += s;
    else if (q == 1) g[Lay::WOB + cc] += s;
    else if (q == 2) g[Lay::WD + cc] += s;
    else if (cc == 0) g[Lay::BD] += s;
}

// dh += dHM[:, :H]; dM = dHM[:, H:]
__global__ void k_split_dm(int n, const float* __restrict__ dHM, float* __restrict__ dh, float* __restrict__ dM) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n * H) return;
    const int i = t / H, cc = t % H;
    dh[t] += dHM[(size_t)i * 2 * H + cc];
    dM[t] = dHM[(size_t)i * 2 * H + H + cc];
}

// input layer: part[chunk][c][f] = sum dh0[n][c] feat[n][f], f = NFEAT -> bias
__global__ void k_ingrad(int n, int chunk, const float* __restrict__ dh, const float* __restrict__ feat,
                         float* __restrict__ part) {
    const int t = threadIdx.x;  // 0 .. H*(NFEAT+1)-1
    const int cc = t / (NFEAT + 1), f = t % (NFEAT + 1);
    const int n0 = blockIdx.x * chunk, n1 = min(n0 + chunk, n);
    float s = 0.f;
    for (int i = n0; i < n1; i++) s += dh[(size_t)i * H + cc] * (f < NFEAT ? feat[(size_t)i * NFEAT + f] : 1.f);
    part[(size_t)blockIdx.x * H * (NFEAT + 1) + t] = s;
}

__global__ void k_ingrad_reduce(int nchunk, const float* __restrict__ part, float* __restrict__ g) {
    const int t = threadIdx.x;
    float s = 0.f;
    for (int b = 0; b < nchunk; b++) s += part[(size_t)b * H * (NFEAT + 1) + t];
    const int cc = t / (NFEAT + 1), f = t % (NFEAT + 1);
    if (f < NFEAT) g[Lay::W_IN + cc * NFEAT + f] += s;
    else g[Lay::B_IN + cc] += s;
}

// ---------------------------------------------------------------------------
// host
// ---------------------------------------------------------------------------

static LearnedState* state(flgpu_ctx* c) {
    if (!c->ls) c->ls = new LearnedState();
    return c->ls;
}

// bf16 weight copies: per layer l, block of 2H*H ([Wa;Wb] rows = outputs),
// then Wu (H x 2H), then Wu^T (2H x H), then [Wa|Wb]^T as (H x 2H) for the
// backward products dh += dP Wa + dQ Wb.
static constexpr int WBF_LAYER = 2 * H * H + H * 2 * H + 2 * H * H + H * 2 * H;

static int refresh_bf16(flgpu_ctx* c, LearnedState* s) {
    std::vector<float> th(Lay::TOTAL);
    FLGPU_HIP(hipMemcpyAsync(th.data(), s->theta.p, sizeof(float) * Lay::TOTAL, hipMemcpyDeviceToHost, c->stream));
    FLGPU_HIP(hipStreamSynchronize(c->stream));
    std::vector<float> w((size_t)NLAYER * WBF_LAYER);
    for (int l = 0; l < NLAYER; l++) {
        const float* lp = th.data() + Lay::layer(l);
        float* o = w.data() + (size_t)l * WBF_LAYER;
        // [Wa;Wb]: output o<H -> Wa[o][k], o>=H -> Wb[o-H][k]
        for (int oo = 0; oo < 2 * H; oo++)
            for (int k = 0; k < H; k++) o[oo * H + k] = oo < H ? lp[Lay::WA + oo * H + k] : lp[Lay::WB + (oo - H) * H + k];
        float* wu = o + 2 * H * H;
        for (int i = 0; i < H * 2 * H; i++) wu[i] = lp[Lay::WU + i];
        float* wut = wu + H * 2 * H;  // (2H x H): wut[k][o] = Wu[o][k]
        for (int oo = 0; oo < H; oo++)
            for (int k = 0; k < 2 * H; k++) wut[k * H + oo] = lp[Lay::WU + oo * 2 * H + k];
        float* wabt = wut + 2 * H * H;  // (H x 2H): wabt[c][q] = (q<H ? Wa[q][c] : Wb[q-H][c])
        for (int cc = 0; cc < H; cc++)
            for (int q = 0; q < 2 * H; q++)
                wabt[cc * 2 * H + q] = q < H ? lp[Lay::WA + q * H + cc] : lp[Lay::WB + (q - H) * H + cc];
    }
    DevBuf<float> tmp;
    FLGPU_TRY(tmp.upload(w.data(), w.size(), c->stream));
    FLGPU_TRY(s->wbf.alloc(w.size()));
    hipLaunchKernelGGL(k_to_bf16, dim3(grid_for((int64_t)w.size(), 256)), dim3(256), 0, c->stream, (int)w.size(),
                       tmp.p, s->wbf.p);
    FLGPU_HIP(hipGetLastError());
    FLGPU_HIP(hipStreamSynchronize(c->stream));
    return FLGPU_OK;
}

static int init_params_host(std::vector<float>& th, uint64_t seed) {
    th.assign(Lay::TOTAL, 0.f);
    auto rnd = [&](uint64_t i) {
        uint64_t x = seed ^ (i * 0x9E3779B97F4A7C15ull);
        x += 0x9E3779B97F4A7C15ull;
        x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
        x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
        x ^= x >> 31;
        return (double)(x >> 11) * (1.0 / 9007199254740992.0) * 2.0 - 1.0;  // [-1,1)
    };
    uint64_t q = 1;
    for (int i = 0; i < H * NFEAT; i++) th[Lay::W_IN + i] = (float)(rnd(q++) * std::sqrt(3.0 / NFEAT));
    for (int l = 0; l < NLAYER; l++) {
        float* lp = th.data() + Lay::layer(l);
        for (int i = 0; i < H * H; i++) lp[Lay::WA + i] = (float)(rnd(q++) * std::sqrt(3.0 / H));
        for (int i = 0; i < H * H; i++) lp[Lay::WB + i] = (float)(rnd(q++) * std::sqrt(3.0 / H));
        for (int i = 0; i < H; i++) lp[Lay::C1 + i] = (float)(rnd(q++));
        for (int i = 0; i < H; i++) lp[Lay::C2 + i] = (float)(rnd(q++));
        for (int i = 0; i < H * 2 * H; i++) lp[Lay::WU + i] = (float)(rnd(q++) * std::sqrt(3.0 / (2 * H)) * 0.5);
    }
    for (int i = 0; i < H; i++) th[Lay::WOA + i] = (float)(rnd(q++) * 0.01);
    for (int i = 0; i < H; i++) th[Lay::WOB + i] = (float)(rnd(q++) * 0.01);
    for (int i = 0; i < H; i++) th[Lay::WD + i] = (float)(rnd(q++) * 0.01);
    return FLGPU_OK;
}

static int ensure_params(flgpu_ctx* c, LearnedState* s) {
    if (s->have_params) return FLGPU_OK;
    std::vector<float> th;
    init_params_host(th, 0x5eed);
    FLGPU_TRY(s->theta.upload(th.data(), th.size(), c->stream));
    FLGPU_TRY(s->am.alloc(Lay::TOTAL));
    FLGPU_TRY(s->av.alloc(Lay::TOTAL));
    FLGPU_TRY(s->grad.alloc(Lay::TOTAL));
    FLGPU_HIP(hipMemsetAsync(s->am.p, 0, sizeof(float) * Lay::TOTAL, c->stream));
    FLGPU_HIP(hipMemsetAsync(s->av.p, 0, sizeof(float) * Lay::TOTAL, c->stream));
    s->step = 0;
    s->have_params = true;
    return refresh_bf16(c, s);
}

// The graph of the reduced system: d, s, features, tmap.  val64 is the
// reduced FP64 matrix built by the Krylov setup.
static int graph_prepare(flgpu_ctx* c, LearnedState* s, const double* val64, bool features = true) {
    const int n = c->rows;
    const unsigned g = grid_for(n, 256);
    if (s->graph_key != c->sparsity_gen || s->n != n || s->nnz != c->nnz) {
        FLGPU_TRY(s->tmap.alloc(c->nnz));
        hipLaunchKernelGGL(k_tmap, dim3(g), dim3(256), 0, c->stream, n, c->rowptr.p, c->col.p, s->tmap.p);
        FLGPU_HIP(hipGetLastError());
        s->graph_key = c->sparsity_gen;
        s->n = n;
        s->nnz = c->nnz;
    }
    FLGPU_TRY(s->dsq.alloc(n));
    FLGPU_TRY(s->sval.alloc(c->nnz));
    FLGPU_TRY(s->feat.alloc((size_t)n * NFEAT));
    hipLaunchKernelGGL((k_graph_prep<double>), dim3(g), dim3(256), 0, c->stream, n, c->rowptr.p, c->col.p, c->diag.p,
                       val64, s->dsq.p, s->sval.p, s->feat.p, 0);
    if (features)  // the scaled values and node features (the network's inputs)
        hipLaunchKernelGGL((k_graph_prep_rows<double>), dim3(grid_for((int64_t)n * 32, 256)), dim3(256), 0,
                           c->stream, n, c->rowptr.p, c->col.p, val64, s->dsq.p, s->sval.p, s->feat.p);
    FLGPU_HIP(hipGetLastError());
    return FLGPU_OK;
}

static unsigned gemm_grid(int n, int dout) {
    const int64_t waves = (int64_t)((n + 31) / 32) * (dout / 32);
    return (unsigned)std::max<int64_t>(1, (waves + 3) / 4);
}

// Forward pass; keep_all: store every layer (training).  Leaves alpha,beta,
// psi in s->ab and the last h.
static int forward(flgpu_ctx* c, LearnedState* s, bool keep_all) {
    const int n = c->rows;
    const int nl = keep_all ? NLAYER : 1;
    FLGPU_TRY(s->h.alloc((size_t)(keep_all ? NLAYER + 1 : 2) * n * H));
    FLGPU_TRY(s->PQ.alloc((size_t)nl * n * 2 * H));
    FLGPU_TRY(s->M.alloc((size_t)nl * n * H));
    FLGPU_TRY(s->U.alloc((size_t)nl * n * H));
    FLGPU_TRY(s->ab.alloc((size_t)3 * n));
    const float* th = s->theta.p;
    hipLaunchKernelGGL(k_input, dim3(grid_for((int64_t)n * H, 256)), dim3(256), 0, c->stream, n, s->feat.p, th,
                       s->h.p);
    for (int l = 0; l < NLAYER; l++) {
        const int li = keep_all ? l : 0;
        float* hin = s->h.p + (size_t)(keep_all ? l : (l & 1)) * n * H;
        float* hout = s->h.p + (size_t)(keep_all ? l + 1 : ((l + 1) & 1)) * n * H;
        float* PQ = s->PQ.p + (size_t)li * n * 2 * H;
        float* M = s->M.p + (size_t)li * n * H;
        float* U = s->U.p + (size_t)li * n * H;
        const __bf16* wl = s->wbf.p + (size_t)l * WBF_LAYER;
        const float* lp = th + Lay::layer(l);
        hipLaunchKernelGGL((k_node_gemm<0>), dim3(gemm_grid(n, 2 * H)), dim3(256), 0, c->stream, n, H, 0, hin,
                           (const float*)nullptr, 2 * H, wl, (const float*)nullptr, (const float*)nullptr, PQ,
                           (float*)nullptr);
        hipLaunchKernelGGL(k_edge_agg, dim3(8 * ((grid_for((int64_t)n * 32, 256) + 7) / 8)), dim3(256), 0, c->stream, n,
                           c->rowptr.p,
                           c->col.p, s->sval.p, PQ, lp, M);
        hipLaunchKernelGGL((k_node_gemm<1>), dim3(gemm_grid(n, H)), dim3(256), 0, c->stream, n, H, H, hin,
                           (const float*)M, H, wl + 2 * H * H, lp + Lay::BU, (const float*)hin, hout, U);
    }
    const float* hlast = s->h.p + (size_t)(keep_all ? NLAYER : (NLAYER & 1)) * n * H;
    hipLaunchKernelGGL(k_out_node, dim3(grid_for(n, 256)), dim3(256), 0, c->stream, n, hlast, th, s->ab.p);
    FLGPU_HIP(hipGetLastError());
    return FLGPU_OK;
}

// upper = 0: the lower pattern and the diagonal only (what the solver's
// halves read: the upper half gathers the transposed lower values)
template <typename GV>
static int build_g(flgpu_ctx* c, LearnedState* s, GV* out, int scaled = 0, int upper = 1) {
    const int n = c->rows;
    hipLaunchKernelGGL((k_build_g<GV>), dim3(grid_for((int64_t)n * 32, 256)), dim3(256), 0, c->stream, n, c->rowptr.p, c->col.p,
                       s->tmap.p, s->sval.p, s->dsq.p, s->ab.p, s->theta.p, out, scaled, upper);
    FLGPU_HIP(hipGetLastError());
    return FLGPU_OK;
}

// val64 of the reduced system (from the Krylov work area)
const double* reduced_values(flgpu_ctx* c);
int reduced_prepare(flgpu_ctx* c);

template <typename GV, typename HV>
static int split_halves(flgpu_ctx* c, LearnedState* s, const GV* gv, DevBuf<HV>& gl, DevBuf<HV>& gu);

template <typename GV>
static int build_fsai(flgpu_ctx* c, LearnedState* s, GV* out) {
    const int n = c->rows;
    const double* val = reduced_values(c);
    FLGPU_TRY(s->fsai_rows.alloc((size_t)n + 1));
    int32_t* cnt = s->fsai_rows.p + n;
    FLGPU_HIP(hipMemsetAsync(cnt, 0, sizeof(int32_t), c->stream));
    hipLaunchKernelGGL((k_fsai_short<GV, kFsaiShort>), dim3(grid_for(n, 256)), dim3(256), 0, c->stream, n,
                       c->rowptr.p, c->col.p, c->diag.p, val, s->dsq.p, out, s->fsai_rows.p, cnt);
    FLGPU_HIP(hipGetLastError());
    int32_t nlong = 0;
    FLGPU_HIP(hipMemcpyAsync(&nlong, cnt, sizeof(int32_t), hipMemcpyDeviceToHost, c->stream));
    FLGPU_HIP(hipStreamSynchronize(c->stream));
    if (nlong > 0)
        hipLaunchKernelGGL((k_fsai_lower<GV, kFsaiMax, 8>), dim3(grid_for(nlong, 8)), dim3(8), 0, c->stream, nlong,
                           c->rowptr.p, c->col.p, c->diag.p, val, s->dsq.p, out, (const int32_t*)s->fsai_rows.p);
    FLGPU_HIP(hipGetLastError());
    return FLGPU_OK;
}

int learned_prepare(flgpu_ctx* c, bool mixed, bool fsai) {
    if (!c->all_diag) return set_error(FLGPU_EUNSUPPORTED, "learned preconditioner: a row has no diagonal entry");
    LearnedState* s = state(c);
    if (!fsai) FLGPU_TRY(ensure_params(c, s));
    FLGPU_TRY(graph_prepare(c, s, reduced_values(c), !fsai));
    if (!fsai) FLGPU_TRY(forward(c, s, false));
    if (mixed) {
        FLGPU_TRY(s->gv.alloc(c->nnz));
        FLGPU_TRY(fsai ? build_fsai<float>(c, s, s->gv.p) : build_g<float>(c, s, s->gv.p, 1, 0));
        FLGPU_TRY((split_halves<float, _Float16>(c, s, s->gv.p, s->gl16, s->gu16)));
    } else {
        FLGPU_TRY(s->gv64.alloc(c->nnz));
        FLGPU_TRY(fsai ? build_fsai<double>(c, s, s->gv64.p) : build_g<double>(c, s, s->gv64.p, 1, 0));
        FLGPU_TRY((split_halves<double, double>(c, s, s->gv64.p, s->gl64, s->gu64)));
    }
    FLGPU_HIP(hipStreamSynchronize(c->stream));
    return FLGPU_OK;
}

// Compact lower/upper CSR of A's pattern (positions into A for the values)
// and their stream row blocks; rebuilt when the sparsity changes.
// The factor halves take the sliced ELL product when each pads at most this much.
constexpr double kSellMaxFillHalf = 1.2;

static int lu_prepare(flgpu_ctx* c, LearnedState* s) {
    if (s->lu_key == c->sparsity_gen) return FLGPU_OK;
    const int n = c->rows;
    const int32_t* rp = c->h_rowptr.data();
    const int32_t* cl = c->h_col.data();
    std::vector<int32_t> lp(n + 1), up(n + 1), lc, uc, ls, us;
    lc.reserve((size_t)c->nnz / 2 + n);
    uc.reserve((size_t)c->nnz / 2 + n);
    ls.reserve((size_t)c->nnz / 2 + n);
    us.reserve((size_t)c->nnz / 2 + n);
    for (int i = 0; i < n; i++) {
        lp[i] = (int32_t)lc.size();
        up[i] = (int32_t)uc.size();
        for (int k = rp[i]; k < rp[i + 1]; k++) {
            if (cl[k] <= i) {
                lc.push_back(cl[k]);
                ls.push_back(k);
            }
            if (cl[k] >= i) {
                // the upper half takes each value from its transposed
                // lower position (j, i) of row j, so the factor is only
                // ever built on the lower pattern (-1: not in the pattern)
                int src = k;
                if (cl[k] > i) {
                    const int32_t* b0 = cl + rp[cl[k]];
                    const int32_t* b1 = cl + rp[cl[k] + 1];
                    const int32_t* hit = std::lower_bound(b0, b1, i);
                    src = (hit != b1 && *hit == i) ? (int)(hit - cl) : -1;
                }
                uc.push_back(cl[k]);
                us.push_back(src);
            }
        }
    }
    lp[n] = (int32_t)lc.size();
    up[n] = (int32_t)uc.size();
    std::vector<int32_t> rbl, rbu, rbkl, rbku;
    stream_row_blocks(lp.data(), n, rbl, rbkl);
    stream_row_blocks(up.data(), n, rbu, rbku);
    FLGPU_TRY(s->rbkl.upload(rbkl.data(), rbkl.size(), c->stream));
    FLGPU_TRY(s->rbku.upload(rbku.data(), rbku.size(), c->stream));
    FLGPU_TRY(s->lptr.upload(lp.data(), lp.size(), c->stream));
    FLGPU_TRY(s->lcol.upload(lc.data(), lc.size(), c->stream));
    FLGPU_TRY(s->lsrc.upload(ls.data(), ls.size(), c->stream));
    FLGPU_TRY(s->uptr.upload(up.data(), up.size(), c->stream));
    FLGPU_TRY(s->ucol.upload(uc.data(), uc.size(), c->stream));
    FLGPU_TRY(s->usrc.upload(us.data(), us.size(), c->stream));
    FLGPU_TRY(s->rbl.upload(rbl.data(), rbl.size(), c->stream));
    FLGPU_TRY(s->rbu.upload(rbu.data(), rbu.size(), c->stream));
    std::vector<int16_t> l16, u16;
    s->l16 = compress_columns(lp.data(), lc.data(), rbl, l16);
    s->u16 = compress_columns(up.data(), uc.data(), rbu, u16);
    if (s->l16) FLGPU_TRY(s->lcol16.upload(l16.data(), l16.size(), c->stream));
    if (s->u16) FLGPU_TRY(s->ucol16.upload(u16.data(), u16.size(), c->stream));
    // sliced ELL plans of both halves, the slots mapped to the factor's entries
    SellHost hl, hu;
    const char* e = getenv("FLGPU_SPMV");
    s->sell = false;
    if (!(e && std::string(e) == "csr")) {
        sell_plan(lp.data(), lc.data(), n, ls.data(), hl);
        sell_plan(up.data(), uc.data(), n, us.data(), hu);
        s->sell = hl.fill((int64_t)lc.size()) <= kSellMaxFillHalf && hu.fill((int64_t)uc.size()) <= kSellMaxFillHalf;
    }
    if (s->sell) {
        FLGPU_TRY(s->lsp.upload(hl.sptr.data(), hl.sptr.size(), c->stream));
        FLGPU_TRY(s->usp.upload(hu.sptr.data(), hu.sptr.size(), c->stream));
        FLGPU_TRY(s->lss.upload(hl.src.data(), hl.src.size(), c->stream));
        FLGPU_TRY(s->uss.upload(hu.src.data(), hu.src.size(), c->stream));
        if (hl.use16) FLGPU_TRY(s->lsc16.upload(hl.c16.data(), hl.c16.size(), c->stream));
        else FLGPU_TRY(s->lsc32.upload(hl.c32.data(), hl.c32.size(), c->stream));
        if (hu.use16) FLGPU_TRY(s->usc16.upload(hu.c16.data(), hu.c16.size(), c->stream));
        else FLGPU_TRY(s->usc32.upload(hu.c32.data(), hu.c32.size(), c->stream));
        s->ls16 = hl.use16;
        s->us16 = hu.use16;
        s->nsll = hl.nsl;
        s->nslu = hu.nsl;
        s->nslotl = hl.nslot;
        s->nslotu = hu.nslot;
    }
    FLGPU_HIP(hipStreamSynchronize(c->stream));
    s->nnzl = (int64_t)lc.size();
    s->nnzu = (int64_t)uc.size();
    s->nrbl = (int)rbl.size() - 1;
    s->nrbu = (int)rbu.size() - 1;
    s->lu_key = c->sparsity_gen;
    return FLGPU_OK;
}

template <typename GV, typename HV>
static int split_halves(flgpu_ctx* c, LearnedState* s, const GV* gv, DevBuf<HV>& gl, DevBuf<HV>& gu) {
    FLGPU_TRY(lu_prepare(c, s));
    if (s->sell) {
        FLGPU_TRY(gl.alloc(s->nslotl));
        FLGPU_TRY(gu.alloc(s->nslotu));
        hipLaunchKernelGGL((k_sell_gather<GV, HV>), dim3(grid_for(s->nslotl, 256)), dim3(256), 0, c->stream,
                           s->nslotl, s->lss.p, gv, gl.p);
        hipLaunchKernelGGL((k_sell_gather<GV, HV>), dim3(grid_for(s->nslotu, 256)), dim3(256), 0, c->stream,
                           s->nslotu, s->uss.p, gv, gu.p);
        FLGPU_HIP(hipGetLastError());
        return FLGPU_OK;
    }
    FLGPU_TRY(gl.alloc(s->nnzl));
    FLGPU_TRY(gu.alloc(s->nnzu));
    hipLaunchKernelGGL((k_gather_half<GV, HV>), dim3(grid_for(s->nnzl, 256)), dim3(256), 0, c->stream, s->nnzl,
                       s->lsrc.p, gv, gl.p);
    hipLaunchKernelGGL((k_gather_half<GV, HV>), dim3(grid_for(s->nnzu, 256)), dim3(256), 0, c->stream, s->nnzu,
                       s->usrc.p, gv, gu.p);
    FLGPU_HIP(hipGetLastError());
    return FLGPU_OK;
}

template <typename HV, typename CT>
static CsrOp<HV, CT> half_op(int nrb, const int32_t* rb, const int32_t* rbk, const int32_t* ptr, const CT* col,
                             const HV* val, const float* cs, const float* rs) {
    return CsrOp<HV, CT>{nrb, rb, rbk, ptr, col, val, cs, rs};
}

// y = Gs rt (rt = D r) over the lower half, then z = D (Gs^T y) over the
// upper half with the sums r.z, z.z, z.w
template <typename V, typename HV, typename LC, typename UC>
static void apply_launch(flgpu_ctx* c, LearnedState* s, const HV* gl, const HV* gu, const LC* lcol,
                         const UC* ucol, const V* r, const V* rt, V* z, V* y, const V* w, double* partials,
                         const int* done) {
    const int gl_grid = stream_grid(s->nrbl), gu_grid = stream_grid(s->nrbu);
    const auto lo = half_op<HV, LC>(s->nrbl, s->rbl.p, s->rbkl.p, s->lptr.p, lcol, gl, nullptr, nullptr);
    const auto up = half_op<HV, UC>(s->nrbu, s->rbu.p, s->rbku.p, s->uptr.p, ucol, gu, nullptr, s->dsq.p);
    hipLaunchKernelGGL((k_csr_stream<V, HV, LC, V, V, DOT_NONE>), dim3(gl_grid), dim3(kRowBlock), 0, c->stream, lo,
                       rt, y, (const double*)nullptr, (const V*)nullptr, (const V*)nullptr, partials, done);
    hipLaunchKernelGGL((k_csr_stream<V, HV, UC, V, V, DOT_PC>), dim3(gu_grid), dim3(kRowBlock), 0, c->stream, up,
                       (const V*)y, z, (const double*)nullptr, r, w, partials, done);
}

// The same two products over the sliced ELL halves.
template <typename V, typename HV, typename LC, typename UC>
static void apply_launch_sell(flgpu_ctx* c, LearnedState* s, const HV* gl, const HV* gu, const LC* lcol,
                              const UC* ucol, const V* r, const V* rt, V* z, V* y, const V* w, double* partials,
                              const int* done) {
    const int n = c->rows;
    const SellOp<HV, LC> lo{s->nsll, n, s->lsp.p, s->lptr.p, lcol, gl, nullptr};
    const SellOp<HV, UC> up{s->nslu, n, s->usp.p, s->uptr.p, ucol, gu, s->dsq.p};
    hipLaunchKernelGGL((k_sell<V, HV, LC, V, V, DOT_NONE>), dim3(sell_grid(s->nsll)), dim3(kRowBlock), 0, c->stream,
                       lo, rt, y, (const double*)nullptr, (const V*)nullptr, (const V*)nullptr, partials, done);
    hipLaunchKernelGGL((k_sell<V, HV, UC, V, V, DOT_PC>), dim3(sell_grid(s->nslu)), dim3(kRowBlock), 0, c->stream,
                       up, (const V*)y, z, (const double*)nullptr, r, w, partials, done);
}

template <typename V, typename HV>
static int apply_impl(flgpu_ctx* c, LearnedState* s, const HV* gl, const HV* gu, const V* r, const V* rt, V* z,
                      V* y, const V* w, double* partials, int* nparts, const int* done) {
    if (s->sell) {
        *nparts = sell_grid(s->nslu);
        if (s->ls16 && s->us16)
            apply_launch_sell<V, HV, int16_t, int16_t>(c, s, gl, gu, s->lsc16.p, s->usc16.p, r, rt, z, y, w,
                                                       partials, done);
        else if (s->ls16)
            apply_launch_sell<V, HV, int16_t, int32_t>(c, s, gl, gu, s->lsc16.p, s->usc32.p, r, rt, z, y, w,
                                                       partials, done);
        else if (s->us16)
            apply_launch_sell<V, HV, int32_t, int16_t>(c, s, gl, gu, s->lsc32.p, s->usc16.p, r, rt, z, y, w,
                                                       partials, done);
        else
            apply_launch_sell<V, HV, int32_t, int32_t>(c, s, gl, gu, s->lsc32.p, s->usc32.p, r, rt, z, y, w,
                                                       partials, done);
        FLGPU_HIP(hipGetLastError());
        return FLGPU_OK;
    }
    *nparts = stream_grid(s->nrbu);
    if (s->l16 && s->u16)
        apply_launch<V, HV, int16_t, int16_t>(c, s, gl, gu, s->lcol16.p, s->ucol16.p, r, rt, z, y, w, partials,
                                              done);
    else
        apply_launch<V, HV, int32_t, int32_t>(c, s, gl, gu, s->lcol.p, s->ucol.p, r, rt, z, y, w, partials, done);
    FLGPU_HIP(hipGetLastError());
    return FLGPU_OK;
}

template <typename V>
int learned_precond_apply(flgpu_ctx* c, const V* r, const V* rt, V* z, V* y, const V* w, double* partials,
                          int* nparts, const int* done_flag) {
    LearnedState* s = state(c);
    if (std::is_same<V, float>::value)
        return apply_impl<V, _Float16>(c, s, s->gl16.p, s->gu16.p, r, rt, z, y, w, partials, nparts, done_flag);
    return apply_impl<V, double>(c, s, s->gl64.p, s->gu64.p, r, rt, z, y, w, partials, nparts, done_flag);
}
template int learned_precond_apply<double>(flgpu_ctx*, const double*, const double*, double*, double*,
                                           const double*, double*, int*, const int*);
template int learned_precond_apply<float>(flgpu_ctx*, const float*, const float*, float*, float*, const float*,
                                          double*, int*, const int*);

const float* learned_scale(flgpu_ctx* c) { return c->ls ? c->ls->dsq.p : nullptr; }

void learned_free(flgpu_ctx* c) {
    delete c->ls;
    c->ls = nullptr;
}

// Loss and gradient on the device-resident reduced system (grad in s->grad).
static int loss_grad(flgpu_ctx* c, int P, uint64_t seed, double* loss_out) {
    if (!c->all_diag) return set_error(FLGPU_EUNSUPPORTED, "learned preconditioner: a row has no diagonal entry");
    LearnedState* s = state(c);
    const int n = c->rows;
    FLGPU_TRY(ensure_params(c, s));
    FLGPU_TRY(reduced_prepare(c));
    const double* val64 = reduced_values(c);
    FLGPU_TRY(graph_prepare(c, s, val64));
    FLGPU_TRY(forward(c, s, true));
    FLGPU_TRY(s->gv.alloc(c->nnz));
    FLGPU_TRY(build_g<float>(c, s, s->gv.p));
    const size_t np = (size_t)n * P;
    for (DevBuf<float>* b : {&s->z, &s->u, &s->v, &s->w, &s->dv, &s->dw}) FLGPU_TRY(b->alloc(np));
    FLGPU_TRY(s->dphi.alloc(c->nnz));
    FLGPU_TRY(s->dpsi.alloc(n));
    FLGPU_TRY(s->dab.alloc((size_t)3 * n));
    FLGPU_TRY(s->dh.alloc((size_t)n * H));
    FLGPU_TRY(s->dU.alloc((size_t)n * H));
    FLGPU_TRY(s->dHM.alloc((size_t)n * 2 * H));
    FLGPU_TRY(s->dPQ.alloc((size_t)n * 2 * H));
    FLGPU_TRY(s->lpart.alloc((size_t)kPartStride * 4));
    FLGPU_TRY(s->lossbuf.alloc(1));
    FLGPU_HIP(hipMemsetAsync(s->grad.p, 0, sizeof(float) * Lay::TOTAL, c->stream));
    const unsigned gnp = grid_for((int64_t)np, 256);
    hipLaunchKernelGGL(k_probes, dim3(gnp), dim3(256), 0, c->stream, n, P, seed, s->z.p);
    hipLaunchKernelGGL((k_batch_spmv<double>), dim3(gnp), dim3(256), 0, c->stream, n, P, c->rowptr.p, c->col.p,
                       c->diag.p, val64, (const float*)s->z.p, s->u.p, 0);
    hipLaunchKernelGGL((k_batch_spmv<float>), dim3(gnp), dim3(256), 0, c->stream, n, P, c->rowptr.p, c->col.p,
                       c->diag.p, (const float*)s->gv.p, (const float*)s->u.p, s->v.p, 1);
    hipLaunchKernelGGL((k_batch_spmv<float>), dim3(gnp), dim3(256), 0, c->stream, n, P, c->rowptr.p, c->col.p,
                       c->diag.p, (const float*)s->gv.p, (const float*)s->v.p, s->w.p, 2);
    const int lgrid = (int)std::min<int64_t>(kMaxVecGrid, grid_for((int64_t)np, kVecBlock));
    hipLaunchKernelGGL(k_loss, dim3(lgrid), dim3(kVecBlock), 0, c->stream, (int)np, 1.0f / (float)np, s->z.p, s->w.p,
                       s->dw.p, s->lpart.p);
    hipLaunchKernelGGL(k_loss_fin, dim3(1), dim3(kVecBlock), 0, c->stream, lgrid, 1.0f / (float)np, s->lpart.p,
                       s->lossbuf.p);
    // dv = G dw
    hipLaunchKernelGGL((k_batch_spmv<float>), dim3(gnp), dim3(256), 0, c->stream, n, P, c->rowptr.p, c->col.p,
                       c->diag.p, (const float*)s->gv.p, (const float*)s->dw.p, s->dv.p, 1);
    hipLaunchKernelGGL(k_dgs, dim3(grid_for(n, 256)), dim3(256), 0, c->stream, n, P, c->rowptr.p, c->col.p, c->diag.p,
                       s->sval.p, s->dsq.p, s->gv.p, s->u.p, s->v.p, s->dv.p, s->dw.p, s->dphi.p, s->dpsi.p);
    const int ogrid = (int)std::min<int64_t>(kMaxVecGrid, grid_for(n, kVecBlock));
    hipLaunchKernelGGL(k_out_back, dim3(ogrid), dim3(kVecBlock), 0, c->stream, n, c->rowptr.p, c->col.p, c->diag.p,
                       s->tmap.p, s->sval.p, s->dphi.p, s->dpsi.p, s->theta.p, s->dab.p, s->dh.p, s->lpart.p);
    hipLaunchKernelGGL(k_out_back_fin, dim3(1), dim3(kVecBlock), 0, c->stream, ogrid, s->lpart.p, s->grad.p);
    // output-layer weights: grad[WOA/WOB/WD][c] = sum_n dab[n][q] h_L[n][c], b_d = sum dpsi
    const int chunk = 4096;
    const int nchunk = (n + chunk - 1) / chunk;
    FLGPU_TRY(s->wpart.alloc((size_t)nchunk * 2 * H * 2 * H));
    const float* hL = s->h.p + (size_t)NLAYER * n * H;
    hipLaunchKernelGGL(k_outgrad, dim3(nchunk), dim3(4 * H), 0, c->stream, n, chunk, s->dab.p, hL, s->wpart.p);
    hipLaunchKernelGGL(k_outgrad_reduce, dim3(1), dim3(4 * H), 0, c->stream, nchunk, s->wpart.p, s->grad.p);
    // hidden layers, last to first
    FLGPU_TRY(s->dM.alloc((size_t)n * H));  // scratch for dh_in assembly
    const int eblk = (int)grid_for((int64_t)n * 32, 256);
    DevBuf<float> part3;
    FLGPU_TRY(part3.alloc((size_t)eblk * 3 * H));
    for (int l = NLAYER - 1; l >= 0; l--) {
        const float* hin = s->h.p + (size_t)l * n * H;
        const float* PQ = s->PQ.p + (size_t)l * n * 2 * H;
        const float* M = s->M.p + (size_t)l * n * H;
        const float* U = s->U.p + (size_t)l * n * H;
        const __bf16* wl = s->wbf.p + (size_t)l * WBF_LAYER;
        const float* lp = s->theta.p + Lay::layer(l);
        float* gl = s->grad.p + Lay::layer(l);
        const int nt = n * H;
        hipLaunchKernelGGL(k_relu_back, dim3(grid_for(nt, 256)), dim3(256), 0, c->stream, nt, s->dh.p, U, s->dU.p);
        // Wu, b_u
        hipLaunchKernelGGL(k_wgrad, dim3(nchunk), dim3(256), 0, c->stream, n, chunk, H, H, hin, M, H,
                           (const float*)s->dU.p, s->wpart.p);
        hipLaunchKernelGGL(k_wgrad_reduce, dim3(grid_for(H * 2 * H, 256)), dim3(256), 0, c->stream, nchunk,
                           H * 2 * H, (const float*)s->wpart.p, gl + Lay::WU);
        hipLaunchKernelGGL(k_colsum, dim3(nchunk), dim3(64), 0, c->stream, n, H, chunk, (const float*)s->dU.p,
                           s->wpart.p);
        hipLaunchKernelGGL(k_wgrad_reduce, dim3(1), dim3(H), 0, c->stream, nchunk, H, (const float*)s->wpart.p,
                           gl + Lay::BU);
        // [dh_direct | dM] = dU Wu
        hipLaunchKernelGGL((k_node_gemm<0>), dim3(gemm_grid(n, 2 * H)), dim3(256), 0, c->stream, n, H, 0,
                           (const float*)s->dU.p, (const float*)nullptr, 2 * H, wl + 2 * H * H + H * 2 * H,
                           (const float*)nullptr, (const float*)nullptr, s->dHM.p, (float*)nullptr);
        hipLaunchKernelGGL(k_split_dm, dim3(grid_for(nt, 256)), dim3(256), 0, c->stream, n, (const float*)s->dHM.p,
                           s->dh.p, s->dM.p);
        // edge backward -> dP, dQ and (c1, c2, bm) gradients
        hipLaunchKernelGGL(k_edge_back, dim3(eblk), dim3(256), 0, c->stream, n, c->rowptr.p, c->col.p, s->tmap.p,
                           s->sval.p, PQ, lp, (const float*)s->dM.p, s->dPQ.p, part3.p, 0);
        hipLaunchKernelGGL(k_part3_reduce, dim3(1), dim3(3 * H), 0, c->stream, eblk, (const float*)part3.p,
                           gl + Lay::C1);
        // [Wa; Wb]
        hipLaunchKernelGGL(k_wgrad, dim3(nchunk), dim3(256), 0, c->stream, n, chunk, H, 0, hin,
                           (const float*)nullptr, 2 * H, (const float*)s->dPQ.p, s->wpart.p);
        hipLaunchKernelGGL(k_wgrad_reduce, dim3(grid_for(2 * H * H, 256)), dim3(256), 0, c->stream, nchunk,
                           2 * H * H, (const float*)s->wpart.p, gl + Lay::WA);
        // dh_in += dP Wa + dQ Wb
        hipLaunchKernelGGL((k_node_gemm<0>), dim3(gemm_grid(n, H)), dim3(256), 0, c->stream, n, 2 * H, 0,
                           (const float*)s->dPQ.p, (const float*)nullptr, H, wl + 2 * H * H + H * 2 * H + 2 * H * H,
                           (const float*)nullptr, (const float*)nullptr, s->dU.p, (float*)nullptr);
        hipLaunchKernelGGL(k_axpy1, dim3(grid_for(nt, 256)), dim3(256), 0, c->stream, nt, (const float*)s->dU.p,
                           s->dh.p);
    }
    // input layer
    hipLaunchKernelGGL(k_ingrad, dim3(nchunk), dim3(H * (NFEAT + 1)), 0, c->stream, n, chunk, (const float*)s->dh.p,
                       (const float*)s->feat.p, s->wpart.p);
    hipLaunchKernelGGL(k_ingrad_reduce, dim3(1), dim3(H * (NFEAT + 1)), 0, c->stream, nchunk,
                       (const float*)s->wpart.p, s->grad.p);
    FLGPU_HIP(hipGetLastError());
    if (loss_out) {
        FLGPU_HIP(hipMemcpyAsync(loss_out, s->lossbuf.p, sizeof(double), hipMemcpyDeviceToHost, c->stream));
    }
    FLGPU_HIP(hipStreamSynchronize(c->stream));
    return FLGPU_OK;
}

static int adam_step(flgpu_ctx* c, LearnedState* s, float lr) {
    s->step += 1;
    const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
    const float bc1 = 1.f - std::pow(b1, (float)s->step), bc2 = 1.f - std::pow(b2, (float)s->step);
    hipLaunchKernelGGL(k_adam, dim3(grid_for(Lay::TOTAL, 256)), dim3(256), 0, c->stream, Lay::TOTAL, lr, b1, b2, eps,
                       bc1, bc2, s->theta.p, (const float*)s->grad.p, s->am.p, s->av.p);
    FLGPU_HIP(hipGetLastError());
    return refresh_bf16(c, s);
}

}  // namespace flgpu

using namespace flgpu;

extern "C" {

int flgpu_learned_param_count(void) { return params_count(); }

int flgpu_learned_set_params(flgpu_ctx* c, const float* params) {
    return guarded("flgpu_learned_set_params", [&]() -> int {
        if (!c || !params) return set_error(FLGPU_EINVAL, "flgpu_learned_set_params: null argument");
        FLGPU_HIP(hipSetDevice(c->device));
        LearnedState* s = state(c);
        FLGPU_TRY(ensure_params(c, s));
        FLGPU_HIP(hipMemcpyAsync(s->theta.p, params, sizeof(float) * Lay::TOTAL, hipMemcpyHostToDevice, c->stream));
        FLGPU_HIP(hipMemsetAsync(s->am.p, 0, sizeof(float) * Lay::TOTAL, c->stream));
        FLGPU_HIP(hipMemsetAsync(s->av.p, 0, sizeof(float) * Lay::TOTAL, c->stream));
        s->step = 0;
        return refresh_bf16(c, s);
    });
}

int flgpu_learned_get_params(flgpu_ctx* c, float* params) {
    return guarded("flgpu_learned_get_params", [&]() -> int {
        if (!c || !params) return set_error(FLGPU_EINVAL, "flgpu_learned_get_params: null argument");
        FLGPU_HIP(hipSetDevice(c->device));
        LearnedState* s = state(c);
        FLGPU_TRY(ensure_params(c, s));
        FLGPU_HIP(hipMemcpyAsync(params, s->theta.p, sizeof(float) * Lay::TOTAL, hipMemcpyDeviceToHost, c->stream));
        FLGPU_HIP(hipStreamSynchronize(c->stream));
        return FLGPU_OK;
    });
}

int flgpu_learned_init_params(flgpu_ctx* c, uint64_t seed) {
    return guarded("flgpu_learned_init_params", [&]() -> int {
        if (!c) return set_error(FLGPU_EINVAL, "flgpu_learned_init_params: null ctx");
        std::vector<float> th;
        init_params_host(th, seed);
        return flgpu_learned_set_params(c, th.data());
    });
}

int flgpu_learned_loss_grad(flgpu_ctx* c, int32_t nprobe, uint64_t seed, float* loss, float* grad) {
    return guarded("flgpu_learned_loss_grad", [&]() -> int {
        if (!c || !loss) return set_error(FLGPU_EINVAL, "flgpu_learned_loss_grad: null argument");
        if (nprobe < 1 || nprobe > MAXPROBE) return set_error(FLGPU_EINVAL, "flgpu_learned_loss_grad: nprobe %d", nprobe);
        if (!c->assembled) return set_error(FLGPU_EINVAL, "flgpu_learned_loss_grad: no system on the device");
        FLGPU_HIP(hipSetDevice(c->device));
        double l = 0.0;
        FLGPU_TRY(loss_grad(c, nprobe, seed, &l));
        *loss = (float)l;
        if (grad) {
            FLGPU_HIP(hipMemcpyAsync(grad, c->ls->grad.p, sizeof(float) * Lay::TOTAL, hipMemcpyDeviceToHost, c->stream));
            FLGPU_HIP(hipStreamSynchronize(c->stream));
        }
        return FLGPU_OK;
    });
}

int flgpu_learned_train_step(flgpu_ctx* c, int32_t nprobe, uint64_t seed, float lr, float* loss) {
    return guarded("flgpu_learned_train_step", [&]() -> int {
        if (!c || !loss) return set_error(FLGPU_EINVAL, "flgpu_learned_train_step: null argument");
        if (nprobe < 1 || nprobe > MAXPROBE) return set_error(FLGPU_EINVAL, "flgpu_learned_train_step: nprobe %d", nprobe);
        if (!c->assembled) return set_error(FLGPU_EINVAL, "flgpu_learned_train_step: no system on the device");
        FLGPU_HIP(hipSetDevice(c->device));
        double l = 0.0;
        FLGPU_TRY(loss_grad(c, nprobe, seed, &l));
        *loss = (float)l;
        return adam_step(c, c->ls, lr);
    });
}

int flgpu_learned_factor(flgpu_ctx* c, float* G_val, int64_t* nnz) {
    return guarded("flgpu_learned_factor", [&]() -> int {
        if (!c) return set_error(FLGPU_EINVAL, "flgpu_learned_factor: null ctx");
        if (!c->assembled) return set_error(FLGPU_EINVAL, "flgpu_learned_factor: no system on the device");
        FLGPU_HIP(hipSetDevice(c->device));
        FLGPU_TRY(reduced_prepare(c));
        FLGPU_TRY(learned_prepare(c, true));
        FLGPU_TRY(build_g<float>(c, c->ls, c->ls->gv.p, 0));
        if (nnz) *nnz = c->nnz;
        if (G_val) {
            FLGPU_HIP(hipMemcpyAsync(G_val, c->ls->gv.p, sizeof(float) * c->nnz, hipMemcpyDeviceToHost, c->stream));
            FLGPU_HIP(hipStreamSynchronize(c->stream));
        }
        return FLGPU_OK;
    });
}

}  // extern "C"
