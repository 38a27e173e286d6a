// Context lifetime, mesh / sparsity / colouring binding, field mirrors,
// strong Dirichlet and the host-side setup helpers of the C ABI.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstring>
#include <vector>

#include "flgpu_internal.h"

namespace flgpu {

static thread_local std::string g_err;

int set_error(int code, const char* fmt, ...) {
    char buf[1024];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    g_err = buf;
    return code;
}

unsigned grid_for(int64_t n, int block) {
    int64_t g = (n + block - 1) / block;
    return (unsigned)(g < 1 ? 1 : g);
}

__global__ void k_dirichlet(int nbc, const int32_t* __restrict__ nodes,
                            const double* __restrict__ vals,
                            const double* __restrict__ T, int t_const,
                            double dt, int32_t* __restrict__ inactive,
                            double* __restrict__ rhs) {
    int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= nbc) return;
    int nd = nodes[j];
    inactive[nd] = 1;
    double t = t_const ? T[0] : T[nd];
    rhs[nd] = dt > 0.0 ? (vals[j] - t) / dt : vals[j];
}

// rhs += f (nodal, or one value for a constant field)
__global__ void k_rhs_add(int n, const double* __restrict__ f, int f_const, double* __restrict__ rhs) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    rhs[i] = rhs[i] + (f_const ? f[0] : f[i]);
}

// A new coordinate mesh topology: the field mesh, its fields, the sparsity,
// the colouring, the surface and the assembled system named the previous
// mesh's elements and nodes; all are unbound (the calls that need them
// then fail with a message until they are bound again).
static void unbind_from_mesh(flgpu_ctx* c) {
    c->loc = 0;
    c->ngi = 0;
    c->nnodes = 0;
    c->linear = false;
    c->h_conn.clear();
    c->conn.release();
    for (DevField* f : {&c->fT, &c->fU, &c->fK, &c->fS, &c->fA, &c->fRho, &c->fRhoOld, &c->fP, &c->fV}) f->set = 0;
    c->motion_nx = 0;
    c->rows = 0;
    c->cols = 0;
    c->nnz = 0;
    c->sparsity_gen++;
    c->h_rowptr.clear();
    c->h_col.clear();
    c->ncol = 0;
    c->h_col_ptr.clear();
    c->h_col_elems.clear();
    c->col_elems.release();
    c->surf.bound = false;
    c->have_inactive = false;
    c->assembled = false;
    krylov_invalidate(c);
}

}  // namespace flgpu

using namespace flgpu;

extern "C" {

const char* flgpu_last_error(void) { return g_err.c_str(); }

int flgpu_init(flgpu_ctx** out, int32_t device) {
    return guarded("flgpu_init", [&]() -> int {
        if (!out) return set_error(FLGPU_EINVAL, "flgpu_init: null ctx pointer");
        *out = nullptr;
        int ndev = 0;
        FLGPU_HIP(hipGetDeviceCount(&ndev));
        if (device < 0 || device >= ndev)
            return set_error(FLGPU_EINVAL, "flgpu_init: device %d of %d", device, ndev);
        FLGPU_HIP(hipSetDevice(device));
        auto* c = new flgpu_ctx();
        c->device = device;
        hipError_t e = hipGetDeviceProperties(&c->prop, device);
        if (e == hipSuccess) e = hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking);
        if (e != hipSuccess) {
            delete c;
            return set_error(FLGPU_EHIP, "flgpu_init: %s", hipGetErrorString(e));
        }
        const int rc = noalloc_setup(device);  // the device address of flgpu_solve_noalloc's array
        if (rc != FLGPU_OK) {
            (void)hipStreamDestroy(c->stream);
            delete c;
            return rc;
        }
        *out = c;
        return FLGPU_OK;
    });
}

int flgpu_finalize(flgpu_ctx* c) {
    return guarded("flgpu_finalize", [&]() -> int {
        if (!c) return FLGPU_OK;
        (void)hipSetDevice(c->device);
        (void)hipStreamSynchronize(c->stream);
        krylov_free(c);
        learned_free(c);
        pressure_free(c);
        (void)hipStreamDestroy(c->stream);
        delete c;
        return FLGPU_OK;
    });
}

int flgpu_mesh_bind(flgpu_ctx* c, int64_t key_id, int64_t key_event, int32_t dim,
                    int32_t loc_x, int32_t nele, int32_t nx, const int32_t* ndglno_x,
                    const double* X) {
    return guarded("flgpu_mesh_bind", [&]() -> int {
        if (!c || !ndglno_x || !X) return set_error(FLGPU_EINVAL, "flgpu_mesh_bind: null argument");
        if (dim < 1 || dim > 3 || loc_x != dim + 1 || nele <= 0 || nx <= 0)
            return set_error(FLGPU_EUNSUPPORTED,
                             "flgpu_mesh_bind: linear simplices of dim 1-3 only (dim=%d loc_x=%d)",
                             dim, loc_x);
        FLGPU_HIP(hipSetDevice(c->device));
        if (key_id == c->key_id && key_event == c->key_event && c->nele == nele && c->nx == nx)
            return FLGPU_OK;  // cache valid (is_cache_valid)
        std::vector<int32_t> cx((size_t)nele * loc_x);
        for (size_t k = 0; k < cx.size(); k++) {
            int v = ndglno_x[k];
            if (v < 1 || v > nx)
                return set_error(FLGPU_EINVAL, "flgpu_mesh_bind: ndglno_x[%zu]=%d out of 1..%d", k, v, nx);
            cx[k] = v - 1;
        }
        std::vector<double> x4((size_t)nx * 4, 0.0);
        for (int64_t i = 0; i < nx; i++)
            for (int d = 0; d < dim; d++) x4[i * 4 + d] = X[i * dim + d];
        // Moved coordinates on the same connectivity (a new event count, or
        // a new key for the same topology) keep the field mesh, sparsity,
        // colouring and surface; any other mesh unbinds everything derived
        // from the previous one, so a later call cannot index the new
        // arrays with the old element or node numbering.
        const bool same_topology = dim == c->dim && loc_x == c->loc_x && nele == c->nele && nx == c->nx &&
                                   cx == c->h_conn_x;
        c->assembled = false;
        c->key_id = -1;  // the device mirror is being replaced
        c->plan.built = false;
        c->plan.epos_built = false;
        c->plan_gen++;
        c->plan.fail_rc = 0;
        if (!same_topology) unbind_from_mesh(c);
        FLGPU_TRY(c->conn_x.upload(cx.data(), cx.size(), c->stream));
        FLGPU_TRY(c->X4.upload(x4.data(), x4.size(), c->stream));
        FLGPU_HIP(hipStreamSynchronize(c->stream));
        c->h_conn_x.swap(cx);
        c->key_id = key_id;
        c->key_event = key_event;
        c->dim = dim;
        c->loc_x = loc_x;
        c->nele = nele;
        c->nx = nx;
        c->plan.built = false;
        c->plan.epos_built = false;
        c->plan_gen++;
        c->plan.fail_rc = 0;
        return FLGPU_OK;
    });
}

int flgpu_field_mesh_bind(flgpu_ctx* c, int32_t loc, int32_t ngi, int32_t nnodes,
                          const int32_t* ndglno, const double* n, const double* dn,
                          const double* w, const double* x_dn) {
    return guarded("flgpu_field_mesh_bind", [&]() -> int {
        if (!c || !ndglno || !n || !dn || !w || !x_dn)
            return set_error(FLGPU_EINVAL, "flgpu_field_mesh_bind: null argument");
        if (c->nele <= 0) return set_error(FLGPU_EINVAL, "flgpu_field_mesh_bind: bind the coordinate mesh first");
        const int dim = c->dim;
        const int lin = dim + 1, quad = (dim + 1) * (dim + 2) / 2;
        if (!(loc == lin || loc == quad) || ngi < 1 || ngi > kMaxGi || nnodes <= 0)
            return set_error(FLGPU_EUNSUPPORTED,
                             "flgpu_field_mesh_bind: P1/P2 simplices with ngi<=%d only (loc=%d ngi=%d)",
                             kMaxGi, loc, ngi);
        FLGPU_HIP(hipSetDevice(c->device));
        std::vector<int32_t> cn((size_t)c->nele * loc);
        for (size_t k = 0; k < cn.size(); k++) {
            int v = ndglno[k];
            if (v < 1 || v > nnodes)
                return set_error(FLGPU_EINVAL, "flgpu_field_mesh_bind: ndglno[%zu]=%d out of 1..%d", k, v, nnodes);
            cn[k] = v - 1;
        }
        ElemTables t;
        std::memset(&t, 0, sizeof t);
        t.dim = dim;
        t.loc = loc;
        t.ngi = ngi;
        t.loc_x = c->loc_x;
        for (int i = 0; i < loc * ngi; i++) t.n[i] = n[i];
        for (int i = 0; i < loc * ngi * dim; i++) t.dn[i] = dn[i];
        for (int g = 0; g < ngi; g++) t.w[g] = w[g];
        for (int i = 0; i < c->loc_x; i++)
            for (int d = 0; d < dim; d++) t.xdn[i + c->loc_x * d] = x_dn[i + c->loc_x * (0 + ngi * d)];
        for (int i = 0; i < loc; i++) {
            double s = 0.0;
            for (int g = 0; g < ngi; g++) s += w[g] * n[i + loc * g];
            t.what[i] = s;
            for (int j = 0; j < loc; j++) {
                double m = 0.0;
                for (int g = 0; g < ngi; g++) m += w[g] * n[i + loc * g] * n[j + loc * g];
                t.mhat[i + loc * j] = m;
            }
        }
        if (loc == lin)
            for (int i = 0; i < loc; i++)
                for (int j = 0; j < loc; j++)
                    for (int k = 0; k < loc; k++) {
                        double m = 0.0;
                        for (int g = 0; g < ngi; g++)
                            m += w[g] * n[i + loc * g] * n[j + loc * g] * n[k + loc * g];
                        t.that[i + 4 * (j + 4 * k)] = m;
                    }
        // A field mesh of other nodes leaves the fields and the sparsity of
        // the previous one unbound; other connectivity, the colouring too.
        if (nnodes != c->nnodes) {
            for (DevField* f : {&c->fT, &c->fU, &c->fK, &c->fS, &c->fA, &c->fRho, &c->fRhoOld, &c->fP, &c->fV}) f->set = 0;
            c->rows = c->cols = 0;
            c->nnz = 0;
            c->sparsity_gen++;
        }
        if (cn != c->h_conn) {
            c->ncol = 0;
            c->h_col_ptr.clear();
            c->h_col_elems.clear();
        }
        c->assembled = false;
        c->have_inactive = false;
        krylov_invalidate(c);
        FLGPU_TRY(c->conn.upload(cn.data(), cn.size(), c->stream));
        FLGPU_TRY(c->d_tab.upload(&t, 1, c->stream));
        FLGPU_HIP(hipStreamSynchronize(c->stream));
        c->h_conn.swap(cn);
        c->tab = t;
        c->loc = loc;
        c->ngi = ngi;
        c->nnodes = nnodes;
        c->linear = (loc == lin);
        c->plan.built = false;
        c->plan.epos_built = false;
        c->plan_gen++;
        c->plan.fail_rc = 0;
        c->surf.bound = false;  // the surface names nodes of the previous field mesh
        return FLGPU_OK;
    });
}

int flgpu_sparsity_bind(flgpu_ctx* c, int32_t rows, int32_t cols, const int32_t* findrm,
                        const int32_t* colm, const int32_t* centrm) {
    return guarded("flgpu_sparsity_bind", [&]() -> int {
        if (!c || !findrm || !colm || !centrm)
            return set_error(FLGPU_EINVAL, "flgpu_sparsity_bind: null argument");
        if (rows <= 0 || cols <= 0 || findrm[0] != 1)
            return set_error(FLGPU_EINVAL, "flgpu_sparsity_bind: bad shape or findrm(1) != 1");
        FLGPU_HIP(hipSetDevice(c->device));
        int64_t nnz = (int64_t)findrm[rows] - 1;
        std::vector<int32_t> rp(rows + 1), cl(nnz), dg(rows);
        for (int i = 0; i <= rows; i++) rp[i] = findrm[i] - 1;
        for (int i = 0; i < rows; i++) {
            if (rp[i + 1] < rp[i]) return set_error(FLGPU_EINVAL, "flgpu_sparsity_bind: findrm decreasing at %d", i + 1);
            for (int64_t k = rp[i]; k < rp[i + 1]; k++) {
                int v = colm[k];
                if (v < 1 || v > cols) return set_error(FLGPU_EINVAL, "flgpu_sparsity_bind: colm(%lld)=%d", (long long)k + 1, v);
                if (k > rp[i] && v <= colm[k - 1])
                    return set_error(FLGPU_EINVAL, "flgpu_sparsity_bind: row %d not sorted/unique", i + 1);
                cl[k] = v - 1;
            }
            dg[i] = centrm[i] - 1;  // -1 when the diagonal is absent (centrm = 0)
        }
        FLGPU_TRY(c->rowptr.upload(rp.data(), rp.size(), c->stream));
        FLGPU_TRY(c->col.upload(cl.data(), cl.size(), c->stream));
        FLGPU_TRY(c->diag.upload(dg.data(), dg.size(), c->stream));
        FLGPU_TRY(c->A.alloc(nnz));
        FLGPU_TRY(c->rhs.alloc(rows));
        FLGPU_TRY(c->inactive.alloc(rows));
        FLGPU_HIP(hipMemsetAsync(c->inactive.p, 0, sizeof(int32_t) * rows, c->stream));
        FLGPU_HIP(hipStreamSynchronize(c->stream));
        c->all_diag = true;
        for (int i = 0; i < rows; i++) c->all_diag = c->all_diag && dg[i] >= 0;
        c->h_rowptr.swap(rp);
        c->h_col.swap(cl);
        c->rows = rows;
        c->cols = cols;
        c->nnz = nnz;
        c->have_inactive = false;
        c->assembled = false;
        c->plan.built = false;
        c->plan.epos_built = false;
        c->plan_gen++;
        c->plan.fail_rc = 0;
        c->sparsity_gen++;
        krylov_invalidate(c);
        return FLGPU_OK;
    });
}

int flgpu_colouring_bind(flgpu_ctx* c, int32_t ncol, const int32_t* col_ptr, const int32_t* elems) {
    return guarded("flgpu_colouring_bind", [&]() -> int {
        if (!c || !col_ptr || !elems || ncol <= 0)
            return set_error(FLGPU_EINVAL, "flgpu_colouring_bind: bad argument");
        if (c->nele <= 0 || c->loc <= 0)
            return set_error(FLGPU_EINVAL, "flgpu_colouring_bind: bind the meshes first");
        if (col_ptr[0] != 1 || col_ptr[ncol] - 1 != c->nele)
            return set_error(FLGPU_ECOLOUR, "flgpu_colouring_bind: sets cover %d of %d elements",
                             col_ptr[ncol] - 1, c->nele);
        std::vector<int32_t> cp(ncol + 1), el(c->nele);
        std::vector<int32_t> seen(c->nele, 0), stamp(c->nnodes, -1);
        for (int k = 0; k <= ncol; k++) cp[k] = col_ptr[k] - 1;
        for (int k = 0; k < ncol; k++) {
            if (cp[k + 1] < cp[k]) return set_error(FLGPU_ECOLOUR, "flgpu_colouring_bind: col_ptr decreasing");
            for (int q = cp[k]; q < cp[k + 1]; q++) {
                int e = elems[q] - 1;
                if (e < 0 || e >= c->nele || seen[e])
                    return set_error(FLGPU_ECOLOUR, "flgpu_colouring_bind: element %d repeated or out of range", e + 1);
                if (q > cp[k] && elems[q] <= elems[q - 1])
                    return set_error(FLGPU_ECOLOUR, "flgpu_colouring_bind: colour %d not ascending", k + 1);
                seen[e] = 1;
                el[q] = e;
                for (int i = 0; i < c->loc; i++) {
                    int nd = c->h_conn[(size_t)e * c->loc + i];
                    if (stamp[nd] == k)
                        return set_error(FLGPU_ECOLOUR,
                                         "flgpu_colouring_bind: colour %d has two elements on node %d",
                                         k + 1, nd + 1);
                    stamp[nd] = k;
                }
            }
        }
        return colouring_adopt(c, cp, el);
    });
}

}  // extern "C"

namespace flgpu {
// Take a checked colouring (0-based set offsets and element ids).
int colouring_adopt(flgpu_ctx* c, std::vector<int32_t>& cp, std::vector<int32_t>& el) {
    FLGPU_HIP(hipSetDevice(c->device));
    FLGPU_TRY(c->col_elems.upload(el.data(), el.size(), c->stream));
    FLGPU_HIP(hipStreamSynchronize(c->stream));
    c->ncol = (int32_t)cp.size() - 1;
    c->h_col_ptr.swap(cp);
    c->h_col_elems.swap(el);
    c->plan.built = false;
    c->plan.epos_built = false;
    c->plan_gen++;
    c->plan.fail_rc = 0;
    return FLGPU_OK;
}
}  // namespace flgpu

extern "C" {

int flgpu_make_sparsity(int32_t nnodes, int32_t nele, int32_t loc, const int32_t* ndglno,
                        int32_t* findrm, int32_t* colm, int32_t* centrm, int64_t* nnz_out) {
    return guarded("flgpu_make_sparsity", [&]() -> int {
        if (!ndglno || !findrm || !centrm || !nnz_out || nnodes <= 0 || nele <= 0 || loc <= 0)
            return set_error(FLGPU_EINVAL, "flgpu_make_sparsity: bad argument");
        // node -> element incidence
        std::vector<int64_t> ptr((size_t)nnodes + 1, 0);
        for (int64_t k = 0; k < (int64_t)nele * loc; k++) {
            int v = ndglno[k];
            if (v < 1 || v > nnodes) return set_error(FLGPU_EINVAL, "flgpu_make_sparsity: ndglno out of range");
            ptr[v]++;
        }
        for (int i = 0; i < nnodes; i++) ptr[i + 1] += ptr[i];
        std::vector<int32_t> inc(ptr[nnodes]);
        {
            std::vector<int64_t> fill(ptr.begin(), ptr.end() - 1);
            for (int e = 0; e < nele; e++)
                for (int i = 0; i < loc; i++) inc[fill[ndglno[(int64_t)e * loc + i] - 1]++] = e;
        }
        std::vector<int32_t> mark(nnodes, -1), buf;
        int64_t count = 0;
        for (int r = 0; r < nnodes; r++) {
            buf.clear();
            for (int64_t q = ptr[r]; q < ptr[r + 1]; q++) {
                const int32_t* en = ndglno + (int64_t)inc[q] * loc;
                for (int i = 0; i < loc; i++) {
                    int nd = en[i] - 1;
                    if (mark[nd] != r) {
                        mark[nd] = r;
                        buf.push_back(nd);
                    }
                }
            }
            std::sort(buf.begin(), buf.end());
            if (count + (int64_t)buf.size() > INT32_MAX - 1)
                return set_error(FLGPU_EUNSUPPORTED, "flgpu_make_sparsity: nnz exceeds int32");
            findrm[r] = (int32_t)(count + 1);
            centrm[r] = 0;
            for (size_t k = 0; k < buf.size(); k++) {
                if (colm) colm[count + k] = buf[k] + 1;
                if (buf[k] == r) centrm[r] = (int32_t)(count + k + 1);
            }
            count += buf.size();
        }
        findrm[nnodes] = (int32_t)(count + 1);
        *nnz_out = count;
        return FLGPU_OK;
    });
}

int flgpu_colour_mesh(int32_t nnodes, int32_t nele, int32_t loc, const int32_t* ndglno,
                      int32_t* colour, int32_t* ncol_out) {
    return guarded("flgpu_colour_mesh", [&]() -> int {
        if (!ndglno || !colour || !ncol_out || nnodes <= 0 || nele <= 0 || loc <= 0)
            return set_error(FLGPU_EINVAL, "flgpu_colour_mesh: bad argument");
        std::vector<int64_t> ptr((size_t)nnodes + 1, 0);
        for (int64_t k = 0; k < (int64_t)nele * loc; k++) {
            int v = ndglno[k];
            if (v < 1 || v > nnodes) return set_error(FLGPU_EINVAL, "flgpu_colour_mesh: ndglno out of range");
            ptr[v]++;
        }
        for (int i = 0; i < nnodes; i++) ptr[i + 1] += ptr[i];
        std::vector<int32_t> inc(ptr[nnodes]);
        {
            std::vector<int64_t> fill(ptr.begin(), ptr.end() - 1);
            for (int e = 0; e < nele; e++)
                for (int i = 0; i < loc; i++) inc[fill[ndglno[(int64_t)e * loc + i] - 1]++] = e;
        }
        // Greedy first fit over elements in index order; only lower-numbered
        // neighbours are coloured yet (colour_sparsity, Colouring.F90:159-199).
        std::vector<int32_t> used(64, -1);
        int ncol = 0;
        for (int e = 0; e < nele; e++) {
            for (int i = 0; i < loc; i++) {
                int nd = ndglno[(int64_t)e * loc + i] - 1;
                for (int64_t q = ptr[nd]; q < ptr[nd + 1]; q++) {
                    int f = inc[q];
                    if (f < e) used[colour[f]] = e;
                }
            }
            int cc = 1;
            while (cc <= ncol && used[cc] == e) cc++;
            colour[e] = cc;
            if (cc > ncol) {
                ncol = cc;
                if ((size_t)ncol + 1 >= used.size()) used.resize(used.size() * 2, -1);
            }
        }
        *ncol_out = ncol;
        return FLGPU_OK;
    });
}

int flgpu_mesh_motion_set(flgpu_ctx* c, const double* X_old, const double* X_new, const double* grid_velocity) {
    return guarded("flgpu_mesh_motion_set", [&]() -> int {
        if (!c || !X_old || !X_new || !grid_velocity)
            return set_error(FLGPU_EINVAL, "flgpu_mesh_motion_set: null argument");
        if (c->nx <= 0) return set_error(FLGPU_EINVAL, "flgpu_mesh_motion_set: bind the coordinate mesh first");
        FLGPU_HIP(hipSetDevice(c->device));
        const int dim = c->dim;
        std::vector<double> pad((size_t)c->nx * 4, 0.0);
        const double* src[3] = {X_old, X_new, grid_velocity};
        DevBuf<double>* dst[3] = {&c->Xo4, &c->Xn4, &c->UG4};
        for (int w = 0; w < 3; w++) {
            for (int64_t i = 0; i < c->nx; i++)
                for (int d = 0; d < dim; d++) pad[i * 4 + d] = src[w][i * dim + d];
            FLGPU_TRY(dst[w]->upload(pad.data(), pad.size(), c->stream));
            FLGPU_HIP(hipStreamSynchronize(c->stream));  // pad is reused
        }
        c->motion_nx = c->nx;
        c->motion_dim = dim;
        return FLGPU_OK;
    });
}

int flgpu_field_set(flgpu_ctx* c, int32_t which, int32_t constant, const double* values) {
    return guarded("flgpu_field_set", [&]() -> int {
        if (!c || !values) return set_error(FLGPU_EINVAL, "flgpu_field_set: null argument");
        if (c->nnodes <= 0) return set_error(FLGPU_EINVAL, "flgpu_field_set: bind the field mesh first");
        DevField* f;
        int ncomp;
        switch (which) {
            case FLGPU_FIELD_T: f = &c->fT; ncomp = 1; break;
            case FLGPU_FIELD_VELOCITY: f = &c->fU; ncomp = c->dim; break;
            case FLGPU_FIELD_DIFFUSIVITY: f = &c->fK; ncomp = c->dim * c->dim; break;
            case FLGPU_FIELD_SOURCE: f = &c->fS; ncomp = 1; break;
            case FLGPU_FIELD_ABSORPTION: f = &c->fA; ncomp = 1; break;
            case FLGPU_FIELD_DENSITY: f = &c->fRho; ncomp = 1; break;
            case FLGPU_FIELD_OLD_DENSITY: f = &c->fRhoOld; ncomp = 1; break;
            case FLGPU_FIELD_PRESSURE: f = &c->fP; ncomp = 1; break;
            case FLGPU_FIELD_VOLUME_FRACTION: f = &c->fV; ncomp = 1; break;
            default: return set_error(FLGPU_EINVAL, "flgpu_field_set: unknown field %d", which);
        }
        if (which == FLGPU_FIELD_T && constant)
            return set_error(FLGPU_EINVAL, "flgpu_field_set: the prognostic field cannot be constant");
        FLGPU_HIP(hipSetDevice(c->device));
        size_t count = constant ? (size_t)ncomp : (size_t)ncomp * c->nnodes;
        FLGPU_TRY(f->v.upload(values, count, c->stream));
        FLGPU_HIP(hipStreamSynchronize(c->stream));
        f->constant = constant ? 1 : 0;
        f->set = 1;
        return FLGPU_OK;
    });
}

int flgpu_advdif_assemble_resident(flgpu_ctx* c, const flgpu_advdif_params* p) {
    return guarded("flgpu_advdif_assemble_resident", [&]() -> int {
        if (!c || !p) return set_error(FLGPU_EINVAL, "flgpu_advdif_assemble_resident: null argument");
        FLGPU_HIP(hipSetDevice(c->device));
        krylov_invalidate(c);
        return assemble(c, p);
    });
}

int flgpu_advdif_assemble(flgpu_ctx* c, const flgpu_advdif_params* p, const double* T,
                          const double* U, const double* kappa, const double* src,
                          const double* absorp, double* A_val, double* rhs) {
    return guarded("flgpu_advdif_assemble", [&]() -> int {
        if (!c || !p) return set_error(FLGPU_EINVAL, "flgpu_advdif_assemble: null argument");
        if (T) FLGPU_TRY(flgpu_field_set(c, FLGPU_FIELD_T, 0, T));
        if (U) FLGPU_TRY(flgpu_field_set(c, FLGPU_FIELD_VELOCITY, 0, U));
        if (kappa) FLGPU_TRY(flgpu_field_set(c, FLGPU_FIELD_DIFFUSIVITY, 0, kappa));
        if (src) FLGPU_TRY(flgpu_field_set(c, FLGPU_FIELD_SOURCE, 0, src));
        if (absorp) FLGPU_TRY(flgpu_field_set(c, FLGPU_FIELD_ABSORPTION, 0, absorp));
        FLGPU_TRY(flgpu_advdif_assemble_resident(c, p));
        return flgpu_matrix_get(c, A_val, rhs, nullptr);
    });
}

int flgpu_rhs_add_field(flgpu_ctx* c, int32_t which) {
    return guarded("flgpu_rhs_add_field", [&]() -> int {
        if (!c) return set_error(FLGPU_EINVAL, "flgpu_rhs_add_field: null ctx");
        if (!c->assembled) return set_error(FLGPU_EINVAL, "flgpu_rhs_add_field: assemble first");
        DevField* f;
        switch (which) {
            case FLGPU_FIELD_SOURCE: f = &c->fS; break;
            case FLGPU_FIELD_ABSORPTION: f = &c->fA; break;
            case FLGPU_FIELD_T: f = &c->fT; break;
            default: return set_error(FLGPU_EINVAL, "flgpu_rhs_add_field: field %d is not scalar", which);
        }
        if (!f->set) return set_error(FLGPU_EINVAL, "flgpu_rhs_add_field: field %d not set", which);
        if (c->nnodes != c->rows)
            return set_error(FLGPU_EINVAL, "flgpu_rhs_add_field: field nodes %d != rows %d", c->nnodes, c->rows);
        FLGPU_HIP(hipSetDevice(c->device));
        krylov_invalidate(c);
        hipLaunchKernelGGL(k_rhs_add, dim3(grid_for(c->rows, 256)), dim3(256), 0, c->stream, c->rows, f->v.p,
                           f->constant, c->rhs.p);
        FLGPU_HIP(hipGetLastError());
        FLGPU_HIP(hipStreamSynchronize(c->stream));
        return FLGPU_OK;
    });
}

int flgpu_dirichlet_apply(flgpu_ctx* c, int32_t nbc, const int32_t* nodes, const double* values,
                          double dt) {
    return guarded("flgpu_dirichlet_apply", [&]() -> int {
        if (!c || (nbc > 0 && (!nodes || !values)))
            return set_error(FLGPU_EINVAL, "flgpu_dirichlet_apply: null argument");
        if (!c->assembled) return set_error(FLGPU_EINVAL, "flgpu_dirichlet_apply: assemble first");
        if (!c->fT.set) return set_error(FLGPU_EINVAL, "flgpu_dirichlet_apply: field T not set");
        if (nbc == 0) return FLGPU_OK;
        FLGPU_HIP(hipSetDevice(c->device));
        std::vector<int32_t> nd(nbc);
        for (int j = 0; j < nbc; j++) {
            if (nodes[j] < 1 || nodes[j] > c->rows)
                return set_error(FLGPU_EINVAL, "flgpu_dirichlet_apply: node %d out of range", nodes[j]);
            nd[j] = nodes[j] - 1;
        }
        DevBuf<int32_t> dn;
        DevBuf<double> dv;
        FLGPU_TRY(dn.upload(nd.data(), nbc, c->stream));
        FLGPU_TRY(dv.upload(values, nbc, c->stream));
        hipLaunchKernelGGL(k_dirichlet, dim3(grid_for(nbc, 256)), dim3(256), 0, c->stream, nbc, dn.p,
                           dv.p, c->fT.v.p, c->fT.constant, dt, c->inactive.p, c->rhs.p);
        FLGPU_HIP(hipGetLastError());
        FLGPU_HIP(hipStreamSynchronize(c->stream));
        c->have_inactive = true;
        krylov_invalidate(c);
        return FLGPU_OK;
    });
}

int flgpu_matrix_get(flgpu_ctx* c, double* A_val, double* rhs, int32_t* inactive) {
    return guarded("flgpu_matrix_get", [&]() -> int {
        if (!c) return set_error(FLGPU_EINVAL, "flgpu_matrix_get: null ctx");
        if ((A_val || rhs) && !c->assembled) return set_error(FLGPU_EINVAL, "flgpu_matrix_get: nothing assembled");
        FLGPU_HIP(hipSetDevice(c->device));
        if (A_val)
            FLGPU_HIP(hipMemcpyAsync(A_val, c->A.p, sizeof(double) * c->nnz, hipMemcpyDeviceToHost, c->stream));
        if (rhs)
            FLGPU_HIP(hipMemcpyAsync(rhs, c->rhs.p, sizeof(double) * c->rows, hipMemcpyDeviceToHost, c->stream));
        if (inactive)
            FLGPU_HIP(hipMemcpyAsync(inactive, c->inactive.p, sizeof(int32_t) * c->rows,
                                     hipMemcpyDeviceToHost, c->stream));
        FLGPU_HIP(hipStreamSynchronize(c->stream));
        return FLGPU_OK;
    });
}

int flgpu_matrix_set(flgpu_ctx* c, const double* A_val, const double* rhs, const int32_t* inactive) {
    return guarded("flgpu_matrix_set", [&]() -> int {
        if (!c) return set_error(FLGPU_EINVAL, "flgpu_matrix_set: null ctx");
        if (c->nnz <= 0) return set_error(FLGPU_EINVAL, "flgpu_matrix_set: bind a sparsity first");
        FLGPU_HIP(hipSetDevice(c->device));
        krylov_invalidate(c);
        if (A_val) {
            FLGPU_HIP(hipMemcpyAsync(c->A.p, A_val, sizeof(double) * c->nnz, hipMemcpyHostToDevice, c->stream));
            c->assembled = true;
            if (!inactive) {
                FLGPU_HIP(hipMemsetAsync(c->inactive.p, 0, sizeof(int32_t) * c->rows, c->stream));
                c->have_inactive = false;
            }
        }
        if (rhs)
            FLGPU_HIP(hipMemcpyAsync(c->rhs.p, rhs, sizeof(double) * c->rows, hipMemcpyHostToDevice, c->stream));
        if (inactive) {
            FLGPU_HIP(hipMemcpyAsync(c->inactive.p, inactive, sizeof(int32_t) * c->rows,
                                     hipMemcpyHostToDevice, c->stream));
            c->have_inactive = true;
        }
        FLGPU_HIP(hipStreamSynchronize(c->stream));
        return FLGPU_OK;
    });
}

int flgpu_solve(flgpu_ctx* c, const flgpu_solver_params* sp, double* x, flgpu_solve_info* info) {
    return guarded("flgpu_solve", [&]() -> int {
        if (!c || !sp) return set_error(FLGPU_EINVAL, "flgpu_solve: null argument");
        if (!c->assembled) return set_error(FLGPU_EINVAL, "flgpu_solve: no system on the device");
        FLGPU_HIP(hipSetDevice(c->device));
        return solve(c, sp, x, info);
    });
}

int flgpu_solve_noalloc(flgpu_ctx* c, const flgpu_solver_params* sp, double* x, flgpu_solve_info* info) {
    return guarded("flgpu_solve_noalloc", [&]() -> int {
        if (!c || !sp) return set_error(FLGPU_EINVAL, "flgpu_solve_noalloc: null argument");
        if (!c->assembled) return set_error(FLGPU_EINVAL, "flgpu_solve_noalloc: no system on the device");
        FLGPU_HIP(hipSetDevice(c->device));
        return solve_noalloc(c, sp, x, info);
    });
}

int flgpu_petsc_solve(flgpu_ctx* c, const flgpu_solver_params* sp, const double* A_val,
                      const int32_t* inactive, const double* rhs, double* x, int32_t* its,
                      int32_t* reason, double* true_relres) {
    return guarded("flgpu_petsc_solve", [&]() -> int {
        if (!c || !sp || !x) return set_error(FLGPU_EINVAL, "flgpu_petsc_solve: null argument");
        FLGPU_TRY(flgpu_matrix_set(c, A_val, rhs, inactive));
        flgpu_solve_info info{};
        FLGPU_TRY(flgpu_solve(c, sp, x, &info));
        if (its) *its = info.its;
        if (reason) *reason = info.reason;
        if (true_relres) *true_relres = info.true_relres;
        return FLGPU_OK;
    });
}

int flgpu_pc_apply(flgpu_ctx* c, const flgpu_solver_params* sp, const double* r, double* z) {
    return guarded("flgpu_pc_apply", [&]() -> int {
        if (!c || !sp || !r || !z) return set_error(FLGPU_EINVAL, "flgpu_pc_apply: null argument");
        if (!c->assembled) return set_error(FLGPU_EINVAL, "flgpu_pc_apply: no system on the device");
        FLGPU_HIP(hipSetDevice(c->device));
        return pc_apply_host(c, sp, r, z);
    });
}

int flgpu_solution_get(flgpu_ctx* c, double* x) {
    return guarded("flgpu_solution_get", [&]() -> int {
        if (!c || !x) return set_error(FLGPU_EINVAL, "flgpu_solution_get: null argument");
        FLGPU_HIP(hipSetDevice(c->device));
        return solution_get(c, x);
    });
}

int flgpu_csr_mult(flgpu_ctx* c, const double* x, double* y) {
    return guarded("flgpu_csr_mult", [&]() -> int {
        if (!c || !x || !y) return set_error(FLGPU_EINVAL, "flgpu_csr_mult: null argument");
        if (!c->assembled) return set_error(FLGPU_EINVAL, "flgpu_csr_mult: no matrix");
        FLGPU_HIP(hipSetDevice(c->device));
        return csr_mult(c, x, y);
    });
}

int flgpu_time_assemble(flgpu_ctx* c, const flgpu_advdif_params* p, int32_t reps, double* ms) {
    return guarded("flgpu_time_assemble", [&]() -> int {
        if (!c || !p || !ms) return set_error(FLGPU_EINVAL, "flgpu_time_assemble: null argument");
        FLGPU_HIP(hipSetDevice(c->device));
        return time_assemble(c, p, reps, ms);
    });
}

int flgpu_assembly_plan_info(flgpu_ctx* c, int64_t* info, int32_t n, double* build_seconds) {
    return guarded("flgpu_assembly_plan_info", [&]() -> int {
        if (!c || (!info && n > 0)) return set_error(FLGPU_EINVAL, "flgpu_assembly_plan_info: null argument");
        FLGPU_HIP(hipSetDevice(c->device));
        return assembly_plan_info(c, info, n, build_seconds);
    });
}

int flgpu_time_spmv(flgpu_ctx* c, int32_t precision, int32_t reps, double* ms) {
    return guarded("flgpu_time_spmv", [&]() -> int {
        if (!c || !ms || reps < 1) return set_error(FLGPU_EINVAL, "flgpu_time_spmv: bad argument");
        if (!c->assembled) return set_error(FLGPU_EINVAL, "flgpu_time_spmv: no system on the device");
        FLGPU_HIP(hipSetDevice(c->device));
        return time_spmv(c, precision, reps, ms);
    });
}

int flgpu_time_spmv_cold(flgpu_ctx* c, int32_t precision, int32_t reps, int32_t flush_mib, double* ms) {
    return guarded("flgpu_time_spmv_cold", [&]() -> int {
        if (!c || !ms || reps < 1 || flush_mib < 1 || flush_mib > 8192)
            return set_error(FLGPU_EINVAL, "flgpu_time_spmv_cold: bad argument");
        if (!c->assembled) return set_error(FLGPU_EINVAL, "flgpu_time_spmv_cold: no system on the device");
        FLGPU_HIP(hipSetDevice(c->device));
        return time_spmv_cold(c, precision, reps, flush_mib, ms);
    });
}

int flgpu_solver_prepare(flgpu_ctx* c, int32_t precision) {
    return guarded("flgpu_solver_prepare", [&]() -> int {
        if (!c) return set_error(FLGPU_EINVAL, "flgpu_solver_prepare: null ctx");
        if (!c->assembled) return set_error(FLGPU_EINVAL, "flgpu_solver_prepare: no system on the device");
        FLGPU_HIP(hipSetDevice(c->device));
        return solver_prepare(c, precision);
    });
}

int flgpu_sizes(flgpu_ctx* c, int32_t* rows, int64_t* nnz, int32_t* ncol) {
    return guarded("flgpu_sizes", [&]() -> int {
        if (!c) return set_error(FLGPU_EINVAL, "flgpu_sizes: null ctx");
        if (rows) *rows = c->rows;
        if (nnz) *nnz = c->nnz;
        if (ncol) *ncol = c->ncol;
        return FLGPU_OK;
    });
}

}  // extern "C"
