// Face loop of the CG advection-diffusion assembly on gfx950
// (assemble/Advection_Diffusion_CG.F90:606-638 -> :1228-1379): Neumann and
// Robin terms of the diffusivity, the boundary term of advection integrated
// by parts (with weak Dirichlet values on inflow faces).
//
// Two launches per call:
//   k_face_local   one thread per face computes the face's matrix and rhs
//                  (FP64, in the order the Fortran source writes them) into
//                  a per-face buffer;
//   k_face_add     one thread per CSR entry (then per rhs row) touched by
//                  faces adds that entry's face values in face order, the
//                  order of the reference's sequential face loop, skipping
//                  zero matrix values as csr_addto does
//                  (femtools/Sparse_Tools.F90:2640).
// The lists of k_face_add are built on the host once per surface and
// sparsity.  The surface is a few percent of the elements' nodes, so both
// kernels are small next to the element assembly.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <numeric>

#include "flgpu_internal.h"

namespace flgpu {

namespace {

constexpr int kBcNeumann = FLGPU_BC_NEUMANN, kBcWeak = FLGPU_BC_WEAKDIRICHLET,
              kBcInternal = FLGPU_BC_INTERNAL, kBcRobin = FLGPU_BC_ROBIN;

struct FaceFlags {
    int adv_face;        // advection integrated by parts
    int diff;            // diffusivity
    int u_c;             // velocity constant
    int mat_terms;       // |dt_theta| > epsilon(0.0)
    double dt_theta, theta;
    // the internal energy's density-weighted face advection (:1318-1327:
    // rho_theta density + (1 - rho_theta) old density, times the volume
    // fraction when multiphase) and, with mesh movement, the velocity
    // relative to the grid (:1315-1317)
    int ie, mph, move;
    double rho_theta;
};

// The fields those read: scalars on T's mesh (constant: one value) and the
// grid velocity by coordinate node, 4 doubles a node.
struct FaceFields {
    const double *rho, *rho_old, *nv, *ug4;
    int rho_c, rho_old_c, nv_c;
};

// One face, FP64.  Arithmetic order follows assemble_advection_diffusion_
// face_cg and its callees term by term (:1228-1379, Transform_elements.F90:
// 1353-1555, FETools.F90:38-54 and :206-226); the file is compiled without
// contraction of multiply-adds.
template <int DIM, int FL>
__global__ void __launch_bounds__(128)
k_face_local(const FaceTables* __restrict__ tb, int nfaces, const int32_t* __restrict__ fnodes,
             const int32_t* __restrict__ fnodes_x, const int32_t* __restrict__ fele,
             const int32_t* __restrict__ conn_x, const double* __restrict__ X4, const double* __restrict__ T,
             const double* __restrict__ U, const int32_t* __restrict__ bct, const double* __restrict__ bv,
             const double* __restrict__ bv2, FaceFlags ff, FaceFields fx, double* __restrict__ fm,
             double* __restrict__ fr, int* __restrict__ err) {
    constexpr int FLX = DIM;   // vertices of a linear face
    constexpr int LX = DIM + 1;
    const int f = blockIdx.x * blockDim.x + threadIdx.x;
    if (f >= nfaces) return;
    const int fg = tb->fngi;
    const int bc = bct[f];
    double mat[FL][FL], rhs[FL], Tf[FL];
#pragma unroll
    for (int i = 0; i < FL; i++) {
        rhs[i] = 0.0;
#pragma unroll
        for (int j = 0; j < FL; j++) mat[i][j] = 0.0;
    }
    int nd[FL];
#pragma unroll
    for (int i = 0; i < FL; i++) nd[i] = fnodes[(int64_t)f * FL + i];
#pragma unroll
    for (int i = 0; i < FL; i++) Tf[i] = T[nd[i]];

    if (bc != kBcInternal) {
        const bool neu_rob = ff.diff && (bc == kBcNeumann || bc == kBcRobin);
        double detJ = 0.0, nrm[DIM];
#pragma unroll
        for (int a = 0; a < DIM; a++) nrm[a] = 0.0;
        if (ff.adv_face || neu_rob) {
            // transform_facet_to_physical: one Jacobian per linear face
            double Xf[DIM][FLX], J[DIM][DIM > 1 ? DIM - 1 : 1];
#pragma unroll
            for (int k = 0; k < FLX; k++) {
                const int v = fnodes_x[(int64_t)f * FLX + k];
#pragma unroll
                for (int a = 0; a < DIM; a++) Xf[a][k] = X4[(int64_t)v * 4 + a];
            }
#pragma unroll
            for (int a = 0; a < DIM; a++)
#pragma unroll
                for (int d = 0; d < DIM - 1; d++) {
                    double t = 0.0;
#pragma unroll
                    for (int k = 0; k < FLX; k++) t = t + Xf[a][k] * tb->fxdn[k + FLX * (0 + fg * d)];
                    J[a][d] = t;
                }
            if constexpr (DIM == 1) {
                detJ = 1.0;  // a point facet: detwei = the weight (1)
            } else if constexpr (DIM == 2) {
                detJ = sqrt(J[0][0] * J[0][0] + J[1][0] * J[1][0]);
            } else {
#pragma unroll
                for (int i = 1; i <= 3; i++) {  // cyc3 (Transform_elements.F90:79)
                    const int a2 = (i + 1) % 3, a1 = i % 3;
                    const double t = J[a2][0] * J[a1][1] - J[a2][1] * J[a1][0];
                    detJ = detJ + t * t;
                }
                detJ = sqrt(detJ);
            }
            if (ff.adv_face) {
                // outward vector: face centroid minus element centroid
                const int ele = fele[f];
                double o[DIM], dp = 0.0, len = 0.0;
#pragma unroll
                for (int a = 0; a < DIM; a++) {
                    double sf = 0.0, se = 0.0;
#pragma unroll
                    for (int k = 0; k < FLX; k++) sf = sf + Xf[a][k];
#pragma unroll
                    for (int k = 0; k < LX; k++) se = se + X4[(int64_t)conn_x[(int64_t)ele * LX + k] * 4 + a];
                    o[a] = sf / FLX - se / LX;
                }
                // facet_normal; cross_product (femtools/Vector_Tools.F90:76-85)
                if constexpr (DIM == 1) {
                    nrm[0] = 1.0;  // its sign from the outward vector below
                } else if constexpr (DIM == 2) {
                    nrm[0] = -J[1][0];
                    nrm[1] = J[0][0];
                } else {
                    nrm[0] = J[1][0] * J[2][1] - J[2][0] * J[1][1];
                    nrm[1] = J[2][0] * J[0][1] - J[0][0] * J[2][1];
                    nrm[2] = J[0][0] * J[1][1] - J[1][0] * J[0][1];
                }
#pragma unroll
                for (int a = 0; a < DIM; a++) dp = dp + nrm[a] * o[a];
#pragma unroll
                for (int a = 0; a < DIM; a++) nrm[a] = nrm[a] * dp;
#pragma unroll
                for (int a = 0; a < DIM; a++) len = len + nrm[a] * nrm[a];
                len = sqrt(len);
#pragma unroll
                for (int a = 0; a < DIM; a++) nrm[a] = nrm[a] / len;
            }
        }

        // add_advection_face_cg (:1292-1347)
        if (ff.adv_face) {
            double Uf[DIM][FL], adv[FL][FL];
#pragma unroll
            for (int d = 0; d < DIM; d++)
#pragma unroll
                for (int k = 0; k < FL; k++) Uf[d][k] = ff.u_c ? U[d] : U[(int64_t)nd[k] * DIM + d];
#pragma unroll
            for (int i = 0; i < FL; i++)
#pragma unroll
                for (int j = 0; j < FL; j++) adv[i][j] = 0.0;
#pragma unroll 1
            for (int g = 0; g < fg; g++) {
                double n_g[FL];
#pragma unroll
                for (int k = 0; k < FL; k++) n_g[k] = tb->fn[k + FL * g];
                double un = 0.0;
#pragma unroll
                for (int d = 0; d < DIM; d++) {
                    double uq = 0.0;
#pragma unroll
                    for (int k = 0; k < FL; k++) uq = uq + Uf[d][k] * n_g[k];
                    if (ff.move) {  // linear T: the face's vertices are its nodes
                        double ugq = 0.0;
#pragma unroll
                        for (int k = 0; k < FLX && k < FL; k++)
                            ugq = ugq + fx.ug4[(int64_t)fnodes_x[(int64_t)f * FLX + k] * 4 + d] * n_g[k];
                        uq = uq - ugq;
                    }
                    un = un + uq * nrm[d];
                }
                double cq = (detJ * tb->fw[g]) * un;
                if (ff.ie) {
                    double r = 0.0, ro = 0.0, nv = 0.0;
#pragma unroll
                    for (int k = 0; k < FL; k++) {
                        r = r + (fx.rho_c ? fx.rho[0] : fx.rho[nd[k]]) * n_g[k];
                        ro = ro + (fx.rho_old_c ? fx.rho_old[0] : fx.rho_old[nd[k]]) * n_g[k];
                        if (ff.mph) nv = nv + (fx.nv_c ? fx.nv[0] : fx.nv[nd[k]]) * n_g[k];
                    }
                    cq = cq * (ff.rho_theta * r + (1.0 - ff.rho_theta) * ro);
                    if (ff.mph) cq = cq * nv;
                }
#pragma unroll
                for (int i = 0; i < FL; i++)
#pragma unroll
                    for (int j = 0; j < FL; j++) adv[i][j] = adv[i][j] + (n_g[i] * n_g[j]) * cq;
            }
            if (ff.mat_terms) {
                if (bc == kBcWeak) {
                    double gv[FL];
#pragma unroll
                    for (int k = 0; k < FL; k++) gv[k] = bv[(int64_t)f * FL + k];
#pragma unroll
                    for (int i = 0; i < FL; i++) {
                        double t = 0.0;
#pragma unroll
                        for (int j = 0; j < FL; j++) t = t + adv[i][j] * (gv[j] - Tf[j]);
                        rhs[i] = rhs[i] - ff.theta * t;
                    }
                } else {
#pragma unroll
                    for (int i = 0; i < FL; i++)
#pragma unroll
                        for (int j = 0; j < FL; j++) mat[i][j] = mat[i][j] + ff.dt_theta * adv[i][j];
                }
            }
#pragma unroll
            for (int i = 0; i < FL; i++) {
                double t = 0.0;
#pragma unroll
                for (int j = 0; j < FL; j++) t = t + adv[i][j] * Tf[j];
                rhs[i] = rhs[i] - t;
            }
        }

        // add_diffusivity_face_cg (:1349-1379)
        if (neu_rob) {
            double gv[FL], ri[FL];
#pragma unroll
            for (int k = 0; k < FL; k++) gv[k] = bv[(int64_t)f * FL + k];
#pragma unroll
            for (int i = 0; i < FL; i++) ri[i] = 0.0;
#pragma unroll 1
            for (int g = 0; g < fg; g++) {
                double n_g[FL];
#pragma unroll
                for (int k = 0; k < FL; k++) n_g[k] = tb->fn[k + FL * g];
                double t = 0.0;
#pragma unroll
                for (int k = 0; k < FL; k++) t = t + gv[k] * n_g[k];
                const double wq = (detJ * tb->fw[g]) * t;
#pragma unroll
                for (int i = 0; i < FL; i++) ri[i] = ri[i] + n_g[i] * wq;
            }
#pragma unroll
            for (int i = 0; i < FL; i++) rhs[i] = rhs[i] + ri[i];
        }
        if (ff.diff && bc == kBcRobin) {
            double hv[FL], rob[FL][FL];
#pragma unroll
            for (int k = 0; k < FL; k++) hv[k] = bv2[(int64_t)f * FL + k];
#pragma unroll
            for (int i = 0; i < FL; i++)
#pragma unroll
                for (int j = 0; j < FL; j++) rob[i][j] = 0.0;
#pragma unroll 1
            for (int g = 0; g < fg; g++) {
                double n_g[FL];
#pragma unroll
                for (int k = 0; k < FL; k++) n_g[k] = tb->fn[k + FL * g];
                double t = 0.0;
#pragma unroll
                for (int k = 0; k < FL; k++) t = t + hv[k] * n_g[k];
                const double wq = (detJ * tb->fw[g]) * t;
#pragma unroll
                for (int i = 0; i < FL; i++)
#pragma unroll
                    for (int j = 0; j < FL; j++) rob[i][j] = rob[i][j] + (n_g[i] * n_g[j]) * wq;
            }
            if (ff.mat_terms)
#pragma unroll
                for (int i = 0; i < FL; i++)
#pragma unroll
                    for (int j = 0; j < FL; j++) mat[i][j] = mat[i][j] + ff.dt_theta * rob[i][j];
#pragma unroll
            for (int i = 0; i < FL; i++) {
                double t = 0.0;
#pragma unroll
                for (int j = 0; j < FL; j++) t = t + rob[i][j] * Tf[j];
                rhs[i] = rhs[i] - t;
            }
        } else if (ff.diff && bc == kBcWeak) {
            atomicOr(err, 2);  // weak Dirichlet with diffusivity (:1374-1377)
        }
    }
#pragma unroll
    for (int i = 0; i < FL; i++) {
#pragma unroll
        for (int j = 0; j < FL; j++) fm[(int64_t)f * FL * FL + i * FL + j] = mat[i][j];
        fr[(int64_t)f * FL + i] = rhs[i];
    }
}

// Thread per touched CSR entry, then per touched rhs row: the face values
// in face order.  Matrix zeros are skipped (csr_addto); an entry outside
// the pattern (pos < 0) is an error only for a nonzero value.
__global__ void __launch_bounds__(256)
k_face_add(int nent, const int32_t* __restrict__ tpos, const int32_t* __restrict__ tptr,
           const int32_t* __restrict__ tsrc, const double* __restrict__ fm, int nrow,
           const int32_t* __restrict__ rrow, const int32_t* __restrict__ rptr, const int32_t* __restrict__ rsrc,
           const double* __restrict__ fr, double* __restrict__ A, double* __restrict__ rhs,
           int* __restrict__ err) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t < nent) {
        const int pos = tpos[t];
        double s = pos >= 0 ? A[pos] : 0.0;
        bool bad = false;
        for (int k = tptr[t]; k < tptr[t + 1]; k++) {
            const double v = fm[tsrc[k]];
            if (v == 0.0) continue;
            if (pos < 0) bad = true;
            else s = s + v;
        }
        if (bad) atomicOr(err, 1);
        if (pos >= 0) A[pos] = s;
    } else if (t < nent + nrow) {
        const int q = t - nent;
        const int row = rrow[q];
        double s = rhs[row];
        for (int k = rptr[q]; k < rptr[q + 1]; k++) s = s + fr[rsrc[k]];
        rhs[row] = s;
    }
}

int64_t find_pos(const flgpu_ctx* c, int r, int cj) {
    int lo = c->h_rowptr[r], hi = c->h_rowptr[r + 1] - 1;
    while (lo <= hi) {
        const int mid = (lo + hi) >> 1;
        const int v = c->h_col[mid];
        if (v == cj) return mid;
        if (v < cj) lo = mid + 1;
        else hi = mid - 1;
    }
    return -1;
}

// The addition lists for the bound surface and sparsity (host, once).
int build_lists(flgpu_ctx* c) {
    SurfaceState& S = c->surf;
    if (S.lists_gen == c->sparsity_gen) return FLGPU_OK;
    const int FL = S.floc, nf = S.nfaces;
    std::vector<std::pair<int64_t, int32_t>> ent;  // (pos, source), sources ascend with the face
    std::vector<std::pair<int32_t, int32_t>> rws;  // (row, source)
    ent.reserve((size_t)nf * FL * FL);
    rws.reserve((size_t)nf * FL);
    for (int k = 0; k < nf * FL; k++)
        if (S.h_fnodes[k] >= c->rows)
            return set_error(FLGPU_EINVAL, "flgpu_advdif_surface: face node %d beyond the %d sparsity rows",
                             S.h_fnodes[k] + 1, c->rows);
    for (int f = 0; f < nf; f++)
        for (int i = 0; i < FL; i++) {
            const int r = S.h_fnodes[(size_t)f * FL + i];
            for (int j = 0; j < FL; j++)
                ent.emplace_back(find_pos(c, r, S.h_fnodes[(size_t)f * FL + j]), f * FL * FL + i * FL + j);
            rws.emplace_back(r, f * FL + i);
        }
    std::stable_sort(ent.begin(), ent.end(), [](const auto& a, const auto& b) { return a.first < b.first; });
    std::stable_sort(rws.begin(), rws.end(), [](const auto& a, const auto& b) { return a.first < b.first; });
    std::vector<int32_t> tpos, tptr{0}, tsrc, rrow, rptr{0}, rsrc;
    tsrc.reserve(ent.size());
    rsrc.reserve(rws.size());
    // entries outside the pattern (pos -1) stay one list per pair
    for (size_t k = 0; k < ent.size(); k++) {
        if (k == 0 || ent[k].first != ent[k - 1].first || ent[k].first < 0) {
            if (k) tptr.push_back((int32_t)tsrc.size());
            tpos.push_back((int32_t)ent[k].first);
        }
        tsrc.push_back(ent[k].second);
    }
    if (!ent.empty()) tptr.push_back((int32_t)tsrc.size());
    for (size_t k = 0; k < rws.size(); k++) {
        if (k == 0 || rws[k].first != rws[k - 1].first) {
            if (k) rptr.push_back((int32_t)rsrc.size());
            rrow.push_back(rws[k].first);
        }
        rsrc.push_back(rws[k].second);
    }
    if (!rws.empty()) rptr.push_back((int32_t)rsrc.size());
    FLGPU_TRY(S.tpos.upload(tpos.data(), tpos.size(), c->stream));
    FLGPU_TRY(S.tptr.upload(tptr.data(), tptr.size(), c->stream));
    FLGPU_TRY(S.tsrc.upload(tsrc.data(), tsrc.size(), c->stream));
    FLGPU_TRY(S.rrow.upload(rrow.data(), rrow.size(), c->stream));
    FLGPU_TRY(S.rptr.upload(rptr.data(), rptr.size(), c->stream));
    FLGPU_TRY(S.rsrc.upload(rsrc.data(), rsrc.size(), c->stream));
    FLGPU_HIP(hipStreamSynchronize(c->stream));  // the host vectors go out of scope
    S.nent = (int32_t)tpos.size();
    S.nrow = (int32_t)rrow.size();
    S.lists_gen = c->sparsity_gen;
    return FLGPU_OK;
}

template <int DIM, int FL>
void launch_face_local(flgpu_ctx* c, const FaceFlags& ff, int* d_err) {
    SurfaceState& S = c->surf;
    const FaceFields fx{c->fRho.v.p, c->fRhoOld.v.p, c->fV.v.p, c->UG4.p, c->fRho.constant, c->fRhoOld.constant,
                        c->fV.constant};
    hipLaunchKernelGGL((k_face_local<DIM, FL>), dim3(grid_for(S.nfaces, 128)), dim3(128), 0, c->stream,
                       S.d_tab.p, S.nfaces, S.fnodes.p, S.fnodes_x.p, S.fele.p, c->conn_x.p, c->X4.p, c->fT.v.p,
                       c->fU.v.p, S.bct.p, S.bcv.p, S.bcv2.p, ff, fx, S.fm.p, S.fr.p, d_err);
}

}  // namespace

int surface_assemble(flgpu_ctx* c, const flgpu_advdif_params* p, const int32_t* bc_type, const double* bc_value,
                     const double* bc_value2) {
    SurfaceState& S = c->surf;
    if (!S.bound) return set_error(FLGPU_EINVAL, "flgpu_advdif_surface: bind the surface first");
    if (!c->assembled) return set_error(FLGPU_EINVAL, "flgpu_advdif_surface: assemble the elements first");
    if (!bc_type) return set_error(FLGPU_EINVAL, "flgpu_advdif_surface: null bc_type");
    const bool adv_face = p->have_advection && p->integrate_by_parts;
    const bool diff = p->have_diffusivity != 0;
    // the face loop runs only with advection by parts or diffusivity (:608-611)
    if ((!adv_face && !diff) || S.nfaces == 0) return FLGPU_OK;
    if (!c->fT.set) return set_error(FLGPU_EINVAL, "flgpu_advdif_surface: field T not set");
    if (adv_face && !c->fU.set) return set_error(FLGPU_EINVAL, "flgpu_advdif_surface: velocity not set");
    bool need_v = false, need_v2 = false;
    for (int f = 0; f < S.nfaces; f++) {
        const int b = bc_type[f];
        if (b < 0 || b > kBcRobin)
            return set_error(FLGPU_EINVAL, "flgpu_advdif_surface: bc_type(%d)=%d", f + 1, b);
        if (b == kBcWeak && diff)
            return set_error(FLGPU_EUNSUPPORTED,
                             "flgpu_advdif_surface: weak Dirichlet boundary conditions with diffusivity are not "
                             "supported by CG advection-diffusion (face %d)", f + 1);
        need_v |= (b == kBcNeumann || b == kBcRobin) && diff;
        need_v |= b == kBcWeak && adv_face;
        need_v2 |= b == kBcRobin && diff;
    }
    if ((need_v && !bc_value) || (need_v2 && !bc_value2))
        return set_error(FLGPU_EINVAL, "flgpu_advdif_surface: boundary values required by the bc types are NULL");
    FLGPU_TRY(build_lists(c));
    const size_t nv = (size_t)S.nfaces * S.floc;
    std::vector<double> zeros;
    if (!bc_value || !bc_value2) zeros.assign(nv, 0.0);
    FLGPU_TRY(S.bct.upload(bc_type, S.nfaces, c->stream));
    FLGPU_TRY(S.bcv.upload(bc_value ? bc_value : zeros.data(), nv, c->stream));
    FLGPU_TRY(S.bcv2.upload(bc_value2 ? bc_value2 : zeros.data(), nv, c->stream));
    FLGPU_TRY(S.fm.reserve(nv * S.floc));
    FLGPU_TRY(S.fr.reserve(nv));
    DevBuf<int> d_err;
    FLGPU_TRY(d_err.alloc(1));
    FLGPU_HIP(hipMemsetAsync(d_err.p, 0, sizeof(int), c->stream));
    const double dt_theta = p->dt * p->theta;
    const bool ie = adv_face && p->equation_type == FLGPU_EQ_INTERNAL_ENERGY;
    const bool mph = ie && p->multiphase != 0;
    const bool move = adv_face && p->move_mesh != 0 && p->have_mass != 0;
    if (ie && (!c->fRho.set || !c->fRhoOld.set || (mph && !c->fV.set)))
        return set_error(FLGPU_EINVAL, "flgpu_advdif_surface: internal-energy faces need density, old density%s",
                         mph ? " and volume fraction" : "");
    if (move && (c->motion_nx != c->nx || c->motion_dim != c->dim))
        return set_error(FLGPU_EINVAL, "flgpu_advdif_surface: mesh movement needs flgpu_mesh_motion_set");
    if (move && !c->linear) return set_error(FLGPU_EUNSUPPORTED, "flgpu_advdif_surface: mesh movement takes linear T");
    const FaceFlags ff{adv_face ? 1 : 0, diff ? 1 : 0, c->fU.constant, std::fabs(dt_theta) > 2.220446049250313e-16,
                       dt_theta, p->theta, ie ? 1 : 0, mph ? 1 : 0, move ? 1 : 0, p->density_theta};
    const int key = c->dim * 10 + S.floc;
    switch (key) {
        case 11: launch_face_local<1, 1>(c, ff, d_err.p); break;
        case 22: launch_face_local<2, 2>(c, ff, d_err.p); break;
        case 23: launch_face_local<2, 3>(c, ff, d_err.p); break;
        case 33: launch_face_local<3, 3>(c, ff, d_err.p); break;
        case 36: launch_face_local<3, 6>(c, ff, d_err.p); break;
        default: return set_error(FLGPU_EUNSUPPORTED, "flgpu_advdif_surface: dim %d, face loc %d", c->dim, S.floc);
    }
    FLGPU_HIP(hipGetLastError());
    const int n = S.nent + S.nrow;
    if (n > 0) {
        hipLaunchKernelGGL(k_face_add, dim3(grid_for(n, 256)), dim3(256), 0, c->stream, S.nent, S.tpos.p, S.tptr.p,
                           S.tsrc.p, S.fm.p, S.nrow, S.rrow.p, S.rptr.p, S.rsrc.p, S.fr.p, c->A.p, c->rhs.p,
                           d_err.p);
        FLGPU_HIP(hipGetLastError());
    }
    int h_err = 0;
    FLGPU_HIP(hipMemcpyAsync(&h_err, d_err.p, sizeof(int), hipMemcpyDeviceToHost, c->stream));
    FLGPU_HIP(hipStreamSynchronize(c->stream));
    if (h_err & 1)
        return set_error(FLGPU_EPATTERN, "flgpu_advdif_surface: face entry outside the sparsity pattern");
    if (h_err & 2)
        return set_error(FLGPU_EUNSUPPORTED, "flgpu_advdif_surface: weak Dirichlet with diffusivity");
    return FLGPU_OK;
}

}  // namespace flgpu

using namespace flgpu;

extern "C" int flgpu_surface_bind(flgpu_ctx* c, int32_t nfaces, int32_t floc, int32_t fngi,
                                  const int32_t* face_ndglno, const int32_t* face_ndglno_x,
                                  const int32_t* face_ele, const double* fn, const double* fw,
                                  const double* fx_dn) {
    return guarded("flgpu_surface_bind", [&]() -> int {
        if (!c || !face_ndglno || !face_ndglno_x || !face_ele || !fn || !fw || (!fx_dn && c->dim > 1))
            return set_error(FLGPU_EINVAL, "flgpu_surface_bind: null argument");
        if (c->nnodes <= 0 || c->nele <= 0)
            return set_error(FLGPU_EINVAL, "flgpu_surface_bind: bind the coordinate and field meshes first");
        const int dim = c->dim;
        const int flin = dim, fquad = dim * (dim + 1) / 2;  // an interval's face is one node at any degree
        if (nfaces < 0 || !(floc == flin || floc == fquad) || floc > kMaxFaceLoc || fngi < 1 || fngi > kMaxFaceGi)
            return set_error(FLGPU_EUNSUPPORTED, "flgpu_surface_bind: face loc %d / ngi %d (dim %d)", floc, fngi, dim);
        if (dim > 1 && (floc == flin) != c->linear)
            return set_error(FLGPU_EINVAL, "flgpu_surface_bind: face degree differs from the field mesh's");
        FLGPU_HIP(hipSetDevice(c->device));
        std::vector<int32_t> fnd((size_t)nfaces * floc), fx((size_t)nfaces * dim), fe(nfaces);
        for (size_t k = 0; k < fnd.size(); k++) {
            const int v = face_ndglno[k];
            if (v < 1 || v > c->nnodes)
                return set_error(FLGPU_EINVAL, "flgpu_surface_bind: face_ndglno[%zu]=%d out of 1..%d", k, v, c->nnodes);
            fnd[k] = v - 1;
        }
        for (size_t k = 0; k < fx.size(); k++) {
            const int v = face_ndglno_x[k];
            if (v < 1 || v > c->nx)
                return set_error(FLGPU_EINVAL, "flgpu_surface_bind: face_ndglno_x[%zu]=%d out of 1..%d", k, v, c->nx);
            fx[k] = v - 1;
        }
        for (int f = 0; f < nfaces; f++) {
            const int e = face_ele[f];
            if (e < 1 || e > c->nele)
                return set_error(FLGPU_EINVAL, "flgpu_surface_bind: face_ele(%d)=%d out of 1..%d", f + 1, e, c->nele);
            fe[f] = e - 1;
        }
        FaceTables t;
        std::memset(&t, 0, sizeof t);
        t.floc = floc;
        t.fngi = fngi;
        t.floc_x = dim;
        for (int k = 0; k < floc * fngi; k++) t.fn[k] = fn[k];
        for (int g = 0; g < fngi; g++) t.fw[g] = fw[g];
        for (int k = 0; k < dim * fngi * (dim - 1); k++) t.fxdn[k] = fx_dn[k];
        SurfaceState& S = c->surf;
        FLGPU_TRY(S.fnodes.upload(fnd.data(), fnd.size(), c->stream));
        FLGPU_TRY(S.fnodes_x.upload(fx.data(), fx.size(), c->stream));
        FLGPU_TRY(S.fele.upload(fe.data(), fe.size(), c->stream));
        FLGPU_TRY(S.d_tab.upload(&t, 1, c->stream));
        FLGPU_HIP(hipStreamSynchronize(c->stream));
        S.h_fnodes.swap(fnd);
        S.tab = t;
        S.nfaces = nfaces;
        S.floc = floc;
        S.fngi = fngi;
        S.lists_gen = -1;
        S.bound = true;
        return FLGPU_OK;
    });
}

extern "C" int flgpu_advdif_surface(flgpu_ctx* c, const flgpu_advdif_params* p, const int32_t* bc_type,
                                    const double* bc_value, const double* bc_value2) {
    return guarded("flgpu_advdif_surface", [&]() -> int {
        if (!c || !p) return set_error(FLGPU_EINVAL, "flgpu_advdif_surface: null argument");
        // the internal-energy face advection is weighted by the density
        // (add_advection_face_cg, :1318-1327), which this face loop does
        // not stage; k-epsilon faces take the advection-diffusion form there
        FLGPU_HIP(hipSetDevice(c->device));
        krylov_invalidate(c);
        return surface_assemble(c, p, bc_type, bc_value, bc_value2);
    });
}
