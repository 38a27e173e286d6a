// Element arithmetic of the CG advection-diffusion assembly: one row of
// the element matrix and rhs per call.  Shared by the gfx950 kernels
// (assemble.hip) and by a host build used to unit-test the same code on
// the CPU (tests compile it with g++); the device path never runs it on
// the host.
#pragma once

#include <cmath>
#include <cstdint>

#ifdef __HIPCC__
#include <hip/hip_runtime.h>
#define FLGPU_HD __host__ __device__
#define FLGPU_HD_INLINE __host__ __device__ __forceinline__
#else
#define FLGPU_HD
#define FLGPU_HD_INLINE inline
#endif

namespace flgpu {

#ifndef FLGPU_ELEM_LIMITS
#define FLGPU_ELEM_LIMITS
constexpr int kElemMaxLoc = 10;   // P2 tetrahedron
constexpr int kElemMaxGi = 11;    // degree-4 tetrahedron rule
constexpr int kElemMaxDim = 3;
#endif

using std::fabs;

FLGPU_HD_INLINE double FMA(double a, double b, double c) { return fma(a, b, c); }
FLGPU_HD_INLINE float FMA(float a, float b, float c) { return fmaf(a, b, c); }

struct AsmFlags {
    double dt_theta, beta;
    int have_mass, lump, adv, ibp, diff, iso, src, absorp;
    int need_div, dtt_on;
    int stab;        // 0 none, 1 streamline upwind, 2 SUPG
    int nu_scheme;   // nu_bar: 1 optimal, 2 doubly asymptotic, 3 critical rule, 4 unity
    double nu_scale;
    // equation type (Advection_Diffusion_CG.F90:484-552): 0 advection-
    // diffusion, 1 internal energy, 2 k-epsilon; mph: internal energy with
    // the phase volume fraction; comp: its p div(u) term; rho_theta: the
    // internal energy's density theta
    int eq = 0, mph = 0, comp = 0;
    double rho_theta = 1.0;
    // mesh movement (:421-437, with mass): the mass at the new positions,
    // the (N^{n+1} - N^n) T / dt rhs term, the velocity relative to the
    // grid and the grid velocity's divergence term; dt is local_dt
    int move = 0;
    double dt = 0.0;
};

// An element's mesh-movement data: |detJ| of its old and new positions and
// the grid velocity at its vertices (linear T: the grid velocity's element
// is T's).
template <int DIM, typename R>
struct MoveData {
    R adet_old, adet_new;
    R ug[DIM + 1][DIM];
};

// Field pointers for the kernels (constant fields point at one value).
struct AsmFields {
    const double* T;
    const double* U; int u_c;
    const double* K; int k_c;
    const double* S; int s_c;
    const double* B; int b_c;
    // density, old density, pressure, phase volume fraction (scalars)
    const double* Rho = nullptr; int rho_c = 0;
    const double* RhoOld = nullptr; int rho_old_c = 0;
    const double* P = nullptr; int p_c = 0;
    const double* V = nullptr; int v_c = 0;
    // mesh movement: old and new positions and the grid velocity, 4 doubles
    // per coordinate node (as the positions)
    const double* Xo4 = nullptr;
    const double* Xn4 = nullptr;
    const double* UG4 = nullptr;
};

// Compile-time feature sets: a kernel instantiated for a subset of terms
// drops the code (and registers) of the others.  A term is computed when
// its bit is set here and its flag is set at run time; the arithmetic of a
// term does not depend on the set, so every instantiation gives the same
// bits for the same flags.
enum : int { F_MASS = 1, F_ADV = 2, F_ABS = 4, F_DIFF = 8, F_SRC = 16, F_ALL = 31, F_ANISO = 32 };

// Tables in LDS, converted to the arithmetic type R.
template <typename R>
struct LTab {
    R n[kElemMaxLoc * kElemMaxGi];
    R dn[kElemMaxLoc * kElemMaxGi * kElemMaxDim];
    R w[kElemMaxGi];
    R what[kElemMaxLoc];
    R mhat[kElemMaxLoc * kElemMaxLoc];
    R that[64];
    R xdn[4 * kElemMaxDim];
    int loc, ngi, lx;
};

// The tables a linear element needs (contracted mass tables, coordinate
// derivatives and the shape derivatives at the first quadrature point, with
// ngi = 1 so that dn indexes like LTab's), a fraction of LTab's size.
template <typename R>
struct LTabLin {
    R dn[4 * kElemMaxDim];
    R what[4];
    R mhat[16];
    R that[64];
    R xdn[4 * kElemMaxDim];
    int loc, ngi, lx;
};

// Gathered element data.
template <int DIM, int LOC, typename R>
struct Elem {
    int nodes[LOC];
    R X[DIM][DIM + 1];
    R T[LOC];
};

// Cofactor inverse and determinant of a Jacobian (transposed, JT[a][b] =
// dx_a/dxi_b), as construct_cache forms them
// (femtools/Transform_elements.F90:263-304).
template <int DIM, typename R>
FLGPU_HD_INLINE void invert_jacobian(const R (&JT)[DIM][DIM], R (&iJ)[DIM][DIM], R& adet) {
    R C[DIM][DIM];
    if constexpr (DIM == 3) {
#pragma unroll
        for (int i = 0; i < 3; i++)
#pragma unroll
            for (int k = 0; k < 3; k++) {
                const int i1 = (i + 1) % 3, i2 = (i + 2) % 3, k1 = (k + 1) % 3, k2 = (k + 2) % 3;
                C[i][k] = FMA(JT[i1][k1], JT[i2][k2], -(JT[i2][k1] * JT[i1][k2]));
            }
    } else if constexpr (DIM == 1) {
        C[0][0] = (R)1;  // an interval: J = dx/dxi, inverse 1/J
    } else {
        C[0][0] = JT[1][1];
        C[1][0] = -JT[0][1];
        C[0][1] = -JT[1][0];
        C[1][1] = JT[0][0];
    }
    R det = JT[0][0] * C[0][0];
#pragma unroll
    for (int a = 1; a < DIM; a++) det = FMA(JT[a][0], C[a][0], det);
    const R rdet = (R)1 / det;
#pragma unroll
    for (int i = 0; i < DIM; i++)
#pragma unroll
        for (int k = 0; k < DIM; k++) iJ[i][k] = C[i][k] * rdet;
    adet = fabs(det);
}

// Jacobian of the linear coordinate element from its derivative table,
// then its inverse and determinant.
template <int DIM, typename R, class TB = LTab<R>>
FLGPU_HD_INLINE void geometry(const R (&X)[DIM][DIM + 1], const TB& tb, R (&iJ)[DIM][DIM], R& adet) {
    R JT[DIM][DIM];
#pragma unroll
    for (int a = 0; a < DIM; a++)
#pragma unroll
        for (int b = 0; b < DIM; b++) {
            R s = (R)0;
#pragma unroll
            for (int nn = 0; nn < DIM + 1; nn++) s = FMA(X[a][nn], tb.xdn[nn + (DIM + 1) * b], s);
            JT[a][b] = s;
        }
    invert_jacobian<DIM, R>(JT, iJ, adet);
}

template <typename R>
FLGPU_HD_INLINE R fld(const double* __restrict__ f, int c, int node, int comp, int ncomp) {
    return (R)(c ? f[comp] : f[(int64_t)node * ncomp + comp]);
}

// Nodal field values of an element's vertices, read from the global field
// arrays (constant fields hold one value).  A kernel that stages its nodes'
// data elsewhere passes its own accessor with the same four methods.
template <int DIM, typename R>
struct GlobalNodal {
    static constexpr bool eq_fields = true;  // has the equation types' fields
    const AsmFields& fd;
    const int* nodes;
    FLGPU_HD_INLINE R U(int k, int d) const { return fld<R>(fd.U, fd.u_c, nodes[k], d, DIM); }
    FLGPU_HD_INLINE R K(int k, int comp) const { return fld<R>(fd.K, fd.k_c, nodes[k], comp, DIM * DIM); }
    FLGPU_HD_INLINE R S(int k) const { return fld<R>(fd.S, fd.s_c, nodes[k], 0, 1); }
    FLGPU_HD_INLINE R B(int k) const { return fld<R>(fd.B, fd.b_c, nodes[k], 0, 1); }
    FLGPU_HD_INLINE R Rho(int k) const { return fld<R>(fd.Rho, fd.rho_c, nodes[k], 0, 1); }
    FLGPU_HD_INLINE R RhoOld(int k) const { return fld<R>(fd.RhoOld, fd.rho_old_c, nodes[k], 0, 1); }
    FLGPU_HD_INLINE R P(int k) const { return fld<R>(fd.P, fd.p_c, nodes[k], 0, 1); }
    FLGPU_HD_INLINE R V(int k) const { return fld<R>(fd.V, fd.v_c, nodes[k], 0, 1); }
};

// Geometry of a linear element: physical gradients G[n][d] = sum_k
// dn(n,k) invJ(d,k) (femtools/Transform_elements.F90:777-780) and |detJ|.
template <int DIM, int LOC, typename R>
struct GeomLin {
    R G[LOC][DIM];
    R adet;
};

template <int DIM, int LOC, typename R, class TB = LTab<R>>
FLGPU_HD_INLINE void geom_linear(const Elem<DIM, LOC, R>& e, const TB& tb, GeomLin<DIM, LOC, R>& gm) {
    R iJ[DIM][DIM];
    geometry<DIM, R, TB>(e.X, tb, iJ, gm.adet);
    for (int nn = 0; nn < LOC; nn++)
        for (int d = 0; d < DIM; d++) {
            R s = (R)0;
            for (int k = 0; k < DIM; k++) s = FMA(tb.dn[nn + LOC * (0 + tb.ngi * k)], iJ[d][k], s);
            gm.G[nn][d] = s;
        }
}

// geom_linear for the standard P1 simplex tables (dn(n,k) = [n == k] for
// n < DIM, dn(DIM,k) = -1, the same for the coordinate element), with the
// multiplications by 0 and +-1 left out.  Each remaining sum runs in the
// table order, so the results equal geom_linear's bit for bit (zeros
// aside, whose sign may differ).
template <int DIM, int LOC, typename R>
FLGPU_HD_INLINE void geom_linear_p1(const Elem<DIM, LOC, R>& e, GeomLin<DIM, LOC, R>& gm) {
    static_assert(LOC == DIM + 1, "linear simplex");
    R JT[DIM][DIM], iJ[DIM][DIM];
#pragma unroll
    for (int a = 0; a < DIM; a++)
#pragma unroll
        for (int b = 0; b < DIM; b++) JT[a][b] = e.X[a][b] - e.X[a][DIM];
    invert_jacobian<DIM, R>(JT, iJ, gm.adet);
#pragma unroll
    for (int d = 0; d < DIM; d++) {
        R s = -iJ[d][0];
#pragma unroll
        for (int k = 0; k < DIM; k++) {
            gm.G[k][d] = iJ[d][k];
            if (k > 0) s = s - iJ[d][k];
        }
        gm.G[DIM][d] = s;
    }
}

// True when the tables are the standard P1 simplex ones geom_linear_p1
// assumes: dn laid out dn[n + loc*k] (ngi = 1), xdn[n + (dim+1)*k].
template <class TB>
FLGPU_HD_INLINE bool p1_standard_tables(const TB& tb, int dim) {
    const int loc = dim + 1;
    for (int n = 0; n < loc; n++)
        for (int k = 0; k < dim; k++) {
            const double want = n == dim ? -1.0 : (n == k ? 1.0 : 0.0);
            if ((double)tb.dn[n + loc * k] != want || (double)tb.xdn[n + loc * k] != want) return false;
        }
    return true;
}

// Row i of the element matrix (vals[j], j = 0..LOC-1) and of the element
// rhs, for linear T (gradients constant over the element).
template <int DIM, int LOC, typename R, int FEAT = F_ALL, class NA = GlobalNodal<DIM, R>, class TB = LTab<R>>
FLGPU_HD_INLINE void row_linear_g(const Elem<DIM, LOC, R>& e, const GeomLin<DIM, LOC, R>& gm, const TB& tb,
                                  const AsmFlags& fl_in, const NA& na, int i, R (&vals)[LOC], R& rhs_i) {
    AsmFlags fl = fl_in;
    fl.have_mass = (FEAT & F_MASS) ? fl.have_mass : 0;
    fl.adv = (FEAT & F_ADV) ? fl.adv : 0;
    fl.absorp = (FEAT & F_ABS) ? fl.absorp : 0;
    fl.diff = (FEAT & F_DIFF) ? fl.diff : 0;
    fl.src = (FEAT & F_SRC) ? fl.src : 0;
    const R(&G)[LOC][DIM] = gm.G;
    const R adet = gm.adet;
    const R dtt = (R)fl.dt_theta;
    // gradient of N_i picked by selects: i may be a run-time value, and an
    // indexed register array would be moved to scratch memory
    R Gi[DIM];
#pragma unroll
    for (int d = 0; d < DIM; d++) {
        Gi[d] = G[0][d];
#pragma unroll
        for (int k = 1; k < LOC; k++)
            if (i == k) Gi[d] = G[k][d];
    }
    R M[LOC], Aa[LOC], Bb[LOC], Kk[LOC];
#pragma unroll
    for (int j = 0; j < LOC; j++) {
        M[j] = (R)0;
        Aa[j] = (R)0;
        Bb[j] = (R)0;
        Kk[j] = (R)0;
    }
    if (fl.have_mass) {
#pragma unroll
        for (int j = 0; j < LOC; j++) M[j] = adet * tb.mhat[i + LOC * j];
    }
    if (fl.adv) {
        R u[LOC][DIM];
#pragma unroll
        for (int k = 0; k < LOC; k++)
#pragma unroll
            for (int d = 0; d < DIM; d++) u[k][d] = na.U(k, d);
        R div = (R)0;
        if (fl.need_div) {
#pragma unroll
            for (int d = 0; d < DIM; d++)
#pragma unroll
                for (int k = 0; k < LOC; k++) div = FMA(u[k][d], G[k][d], div);
        }
        if (!fl.ibp) {
            // A(i,j) = |J| sum_k mhat(i,k) u_k . G_j  (shape_vector_dot_dshape)
            R ut[DIM];
#pragma unroll
            for (int d = 0; d < DIM; d++) {
                R s = (R)0;
#pragma unroll
                for (int k = 0; k < LOC; k++) s = FMA(tb.mhat[i + LOC * k], u[k][d], s);
                ut[d] = s;
            }
#pragma unroll
            for (int j = 0; j < LOC; j++) {
                R s = ut[0] * G[j][0];
#pragma unroll
                for (int d = 1; d < DIM; d++) s = FMA(ut[d], G[j][d], s);
                R a = adet * s;
                if (fl.need_div) a = FMA((R)fl.beta * div, adet * tb.mhat[i + LOC * j], a);
                Aa[j] = a;
            }
        } else {
            // A(i,j) = -|J| G_i . sum_k mhat(k,j) u_k  (dshape_dot_vector_shape)
#pragma unroll
            for (int j = 0; j < LOC; j++) {
                R s = (R)0;
#pragma unroll
                for (int d = 0; d < DIM; d++) {
                    R t = (R)0;
#pragma unroll
                    for (int k = 0; k < LOC; k++) t = FMA(tb.mhat[k + LOC * j], u[k][d], t);
                    s = FMA(Gi[d], t, s);
                }
                R a = -(adet * s);
                if (fl.need_div) a = FMA(-((R)1 - (R)fl.beta) * div, adet * tb.mhat[i + LOC * j], a);
                Aa[j] = a;
            }
        }
    }
    if (fl.absorp) {
        R sg[LOC];
#pragma unroll
        for (int k = 0; k < LOC; k++) sg[k] = na.B(k);
#pragma unroll
        for (int j = 0; j < LOC; j++) {
            R s = (R)0;
#pragma unroll
            for (int k = 0; k < LOC; k++) s = FMA(tb.that[i + 4 * (j + 4 * k)], sg[k], s);
            Bb[j] = adet * s;
        }
    }
    if (fl.diff) {
        if (fl.iso) {
            R kw = (R)0;
#pragma unroll
            for (int k = 0; k < LOC; k++)
                kw = FMA(tb.what[k], na.K(k, 0), kw);
            const R scale = adet * kw;
#pragma unroll
            for (int j = 0; j < LOC; j++) {
                R s = Gi[0] * G[j][0];
#pragma unroll
                for (int d = 1; d < DIM; d++) s = FMA(Gi[d], G[j][d], s);
                Kk[j] = s * scale;
            }
        } else {
            R kb[DIM][DIM];
#pragma unroll
            for (int a = 0; a < DIM; a++)
#pragma unroll
                for (int b = 0; b < DIM; b++) {
                    R s = (R)0;
#pragma unroll
                    for (int k = 0; k < LOC; k++)
                        s = FMA(tb.what[k], na.K(k, a + DIM * b), s);
                    kb[a][b] = s;
                }
            R v[DIM];
#pragma unroll
            for (int b = 0; b < DIM; b++) {
                R s = (R)0;
#pragma unroll
                for (int a = 0; a < DIM; a++) s = FMA(Gi[a], kb[a][b], s);
                v[b] = s;
            }
#pragma unroll
            for (int j = 0; j < LOC; j++) {
                R s = (R)0;
#pragma unroll
                for (int b = 0; b < DIM; b++) s = FMA(v[b], G[j][b], s);
                Kk[j] = adet * s;
            }
        }
    }
    // A term compiled in but switched off holds zeros, and adding its
    // products leaves a sum unchanged (up to the sign of a zero): no
    // run-time test in the sums below.
    R r = (R)0;
    if constexpr ((FEAT & F_ADV) != 0)
#pragma unroll
        for (int j = 0; j < LOC; j++) r = FMA(-Aa[j], e.T[j], r);
    if constexpr ((FEAT & F_ABS) != 0)
#pragma unroll
        for (int j = 0; j < LOC; j++) r = FMA(-Bb[j], e.T[j], r);
    if constexpr ((FEAT & F_DIFF) != 0)
#pragma unroll
        for (int j = 0; j < LOC; j++) r = FMA(-Kk[j], e.T[j], r);
    if (fl.src) {
        R s = (R)0;
#pragma unroll
        for (int k = 0; k < LOC; k++) s = FMA(tb.mhat[i + LOC * k], na.S(k), s);
        r = FMA(adet, s, r);
    }
    rhs_i = r;
    R lump = (R)0;
    if (fl.have_mass && fl.lump) {
#pragma unroll
        for (int j = 0; j < LOC; j++) lump += M[j];
    }
    const R dtte = fl.dtt_on ? dtt : (R)0;
#pragma unroll
    for (int j = 0; j < LOC; j++) {
        R m = (fl.have_mass && !fl.lump) ? M[j] : (R)0;
        if (fl.have_mass && fl.lump && j == i) m = lump;
        if constexpr ((FEAT & F_ADV) != 0) m = FMA(dtte, Aa[j], m);
        if constexpr ((FEAT & F_ABS) != 0) m = FMA(dtte, Bb[j], m);
        if constexpr ((FEAT & F_DIFF) != 0) m = FMA(dtte, Kk[j], m);
        vals[j] = m;
    }
}

template <int DIM, int LOC, typename R>
FLGPU_HD_INLINE void row_linear(const Elem<DIM, LOC, R>& e, const LTab<R>& tb, const AsmFlags& fl,
                                const AsmFields& fd, int i, R (&vals)[LOC], R& rhs_i) {
    GeomLin<DIM, LOC, R> gm;
    geom_linear<DIM, LOC, R>(e, tb, gm);
    row_linear_g<DIM, LOC, R>(e, gm, tb, fl, GlobalNodal<DIM, R>{fd, e.nodes}, i, vals, rhs_i);
}

// Nodal fields at quadrature point g, interpolated from the element's
// nodes as ele_val_at_quad does (femtools/Fields_Base.F90:2256-2310), and
// the divergence of the velocity at g.  The nodal values come from NA
// (GlobalNodal, or a kernel's staged copy).
template <int DIM, typename R>
FLGPU_HD_INLINE R nu_bar_scaled(const R (&u)[DIM], const R (&JT)[DIM][DIM], const R (&K)[DIM][DIM], bool have_k,
                                int scheme, R scale);

template <int DIM, int LOC, typename R, class NA = GlobalNodal<DIM, R>>
struct NodalQuad {
    // nu() reads the diffusivity tensor (a source with nu_bar precomputed
    // says false, and an isotropic row then skips the tensor)
    static constexpr bool nu_reads_k = true;
    static constexpr bool eq_fields = NA::eq_fields;
    const NA& na;
    const LTab<R>& tb;
    const R (&iJ)[DIM][DIM];
    FLGPU_HD_INLINE void u(int g, R (&uq)[DIM]) const {
#pragma unroll
        for (int d = 0; d < DIM; d++) {
            R s = (R)0;
#pragma unroll
            for (int k = 0; k < LOC; k++) s = FMA(na.U(k, d), tb.n[k + LOC * g], s);
            uq[d] = s;
        }
    }
    FLGPU_HD_INLINE R div(int g) const {
        const int ngi = tb.ngi;
        R div = (R)0;
#pragma unroll
        for (int k = 0; k < LOC; k++)
#pragma unroll
            for (int d = 0; d < DIM; d++) {
                R s = (R)0;
#pragma unroll
                for (int q = 0; q < DIM; q++) s = FMA(tb.dn[k + LOC * (g + ngi * q)], iJ[d][q], s);
                div = FMA(na.U(k, d), s, div);
            }
        return div;
    }
    FLGPU_HD_INLINE R k_iso(int g) const {
        R kq = (R)0;
#pragma unroll
        for (int k = 0; k < LOC; k++) kq = FMA(na.K(k, 0), tb.n[k + LOC * g], kq);
        return kq;
    }
    FLGPU_HD_INLINE void k_tensor(int g, R (&kt)[DIM][DIM]) const {
#pragma unroll
        for (int a = 0; a < DIM; a++)
#pragma unroll
            for (int b = 0; b < DIM; b++) {
                R s = (R)0;
#pragma unroll
                for (int k = 0; k < LOC; k++)
                    s = FMA(na.K(k, a + DIM * b), tb.n[k + LOC * g], s);
                kt[a][b] = s;
            }
    }
    FLGPU_HD_INLINE R absorp(int g) const {
        R sq = (R)0;
#pragma unroll
        for (int k = 0; k < LOC; k++) sq = FMA(na.B(k), tb.n[k + LOC * g], sq);
        return sq;
    }
    FLGPU_HD_INLINE R src(int g) const {
        R fq = (R)0;
#pragma unroll
        for (int k = 0; k < LOC; k++) fq = FMA(na.S(k), tb.n[k + LOC * g], fq);
        return fq;
    }
    // a scalar field of the internal-energy / k-epsilon equations at g
    // (ele_val_at_quad), and its gradient there (ele_grad_at_quad_scalar,
    // Fields_Base.F90:2494-2510, with the field on T's mesh): the nodal
    // values by F (a GlobalNodal method)
    template <class F>
    FLGPU_HD_INLINE R at(int g, F f) const {
        R q = (R)0;
#pragma unroll
        for (int k = 0; k < LOC; k++) q = FMA(f(k), tb.n[k + LOC * g], q);
        return q;
    }
    template <class F>
    FLGPU_HD_INLINE void grad(int g, F f, R (&gq)[DIM]) const {
        const int ngi = tb.ngi;
#pragma unroll
        for (int d = 0; d < DIM; d++) gq[d] = (R)0;
#pragma unroll
        for (int k = 0; k < LOC; k++) {
            const R v = f(k);
#pragma unroll
            for (int d = 0; d < DIM; d++) {
                R s = (R)0;
#pragma unroll
                for (int q = 0; q < DIM; q++) s = FMA(tb.dn[k + LOC * (g + ngi * q)], iJ[d][q], s);
                gq[d] = FMA(v, s, gq[d]);
            }
        }
    }
    FLGPU_HD_INLINE R rho(int g) const { return at(g, [&](int k) { return na.Rho(k); }); }
    FLGPU_HD_INLINE R rho_old(int g) const { return at(g, [&](int k) { return na.RhoOld(k); }); }
    FLGPU_HD_INLINE R pres(int g) const { return at(g, [&](int k) { return na.P(k); }); }
    FLGPU_HD_INLINE R nv(int g) const { return at(g, [&](int k) { return na.V(k); }); }
    // the gradient of the advected density: theta-weighted between the
    // density and the old density (internal energy), the density (k-epsilon)
    FLGPU_HD_INLINE void grad_rho(int g, int eq, R th, R (&gq)[DIM]) const {
        grad(g, [&](int k) { return na.Rho(k); }, gq);
        if (eq == 1) {
            R go[DIM];
            grad(g, [&](int k) { return na.RhoOld(k); }, go);
#pragma unroll
            for (int d = 0; d < DIM; d++) gq[d] = th * gq[d] + ((R)1 - th) * go[d];
        }
    }
    FLGPU_HD_INLINE void grad_nv(int g, R (&gq)[DIM]) const { grad(g, [&](int k) { return na.V(k); }, gq); }
    // the stabilisation's nu_bar at g, from u and the diffusivity there
    FLGPU_HD_INLINE R nu(int g, const R (&uq)[DIM], const R (&JT)[DIM][DIM], const R (&kt)[DIM][DIM],
                         const AsmFlags& fl) const {
        return nu_bar_scaled<DIM, R>(uq, JT, kt, fl.diff != 0, fl.nu_scheme, (R)fl.nu_scale);
    }
};

// Row i for quadratic T: per-quadrature-point loops
// (Transform_elements.F90:771-775; FETools.F90 contractions), with the
// fields at the quadrature points from QV (NodalQuad, or a kernel's staged
// copy of the same values).  The mass row and the sum of the advection,
// absorption and diffusion rows are kept (the matrix row is
// M + dt*theta*(A + B + K) and the rhs -(A + B + K) T): two accumulators per
// column instead of four, which is what keeps a P2 row in registers.
template <int DIM, int LOC, typename R, class QV, class TV>
FLGPU_HD_INLINE void row_general_q(const LTab<R>& tb, const AsmFlags& fl, const R (&iJ)[DIM][DIM], R adet,
                                   const QV& qv, const TV& T, int i, R (&vals)[LOC], R& rhs_i,
                                   const MoveData<DIM, R>* mv = nullptr) {
    const int ngi = tb.ngi;
    const R dtt = (R)fl.dt_theta;
    R M[LOC], Sm[LOC], Mv[LOC];
#pragma unroll
    for (int j = 0; j < LOC; j++) {
        M[j] = (R)0;
        Sm[j] = (R)0;
        Mv[j] = (R)0;
    }
    R src_i = (R)0, pdv_i = (R)0;
#pragma unroll 1
    for (int g = 0; g < ngi; g++) {
        const R dw = adet * tb.w[g];
        const R ni = tb.n[i + LOC * g];
        // gradient of N_i at g
        R Gi[DIM];
#pragma unroll
        for (int d = 0; d < DIM; d++) {
            R s = (R)0;
#pragma unroll
            for (int k = 0; k < DIM; k++) s = FMA(tb.dn[i + LOC * (g + ngi * k)], iJ[d][k], s);
            Gi[d] = s;
        }
        R uq[DIM] = {};
        R div = (R)0, div_ug = (R)0;
        if (fl.adv) {
            qv.u(g, uq);
            if (fl.need_div) div = qv.div(g);
        }
        if (mv && fl.adv) {
            // the velocity relative to the grid (:983-985) and, not by
            // parts, div(u_grid) (:1099-1102), from the vertex values
#pragma unroll
            for (int k = 0; k < LOC; k++) {
                const R nk = tb.n[k + LOC * g];
#pragma unroll
                for (int d = 0; d < DIM; d++) uq[d] = uq[d] - mv->ug[k][d] * nk;
            }
            if (!fl.ibp) {
#pragma unroll
                for (int k = 0; k < LOC; k++)
#pragma unroll
                    for (int d = 0; d < DIM; d++) {
                        R s = (R)0;
#pragma unroll
                        for (int q = 0; q < DIM; q++) s = FMA(tb.dn[k + LOC * (g + ngi * q)], iJ[d][q], s);
                        div_ug = FMA(mv->ug[k][d], s, div_ug);
                    }
            }
        }
        R kq = (R)0, sq = (R)0;
        R kt[DIM][DIM];
        if (fl.diff) {
            if (fl.iso)
                kq = qv.k_iso(g);
            else
                qv.k_tensor(g, kt);
        }
        if (fl.absorp) sq = qv.absorp(g);
        if (fl.src) src_i = FMA(ni, dw * qv.src(g), src_i);
        // The internal-energy and k-epsilon equations (:884-903, :983-1097,
        // :1181-1190, :1204-1226): the mass weighted by the old density
        // (internal energy) or the density (k-epsilon), the advection by the
        // (theta-weighted) density, times the phase volume fraction when
        // multiphase; the div term's coefficient div(density u) split as
        // density div(u) + u.grad(density) (and + density u.grad(nvfrac));
        // the diffusion times nvfrac (multiphase); the rhs' -p div(u).
        R dwm = dw, ca = (R)1, cdiv = div, dk = dw, dwv = (R)0;
        if (mv) {
            dwm = mv->adet_new * tb.w[g];
            dwv = dwm - mv->adet_old * tb.w[g];
        }
        if (fl.eq) {
            const R th = (R)fl.rho_theta;
            const R rq = qv.rho(g);
            R rb = rq;
            R nvq = (R)1;
            if (fl.mph) nvq = qv.nv(g);
            if (fl.eq == 1) {
                const R ro = qv.rho_old(g);
                rb = th * rq + ((R)1 - th) * ro;
                dwm = fl.mph ? dw * ro * nvq : dw * ro;
            } else {
                dwm = dw * rq;
            }
            ca = fl.mph ? rb * nvq : rb;
            if (fl.adv && fl.need_div) {
                R gr[DIM];
                qv.grad_rho(g, fl.eq, th, gr);
                R udgr = (R)0;
#pragma unroll
                for (int d = 0; d < DIM; d++) udgr = FMA(gr[d], uq[d], udgr);
                if (fl.mph) {
                    R gn[DIM];
                    qv.grad_nv(g, gn);
                    R udgn = (R)0;
#pragma unroll
                    for (int d = 0; d < DIM; d++) udgn = FMA(gn[d], uq[d], udgn);
                    cdiv = div * rb * nvq + nvq * udgr + rb * udgn;
                } else {
                    cdiv = div * rb + udgr;
                }
            }
            if (fl.mph) dk = dw * nvq;
            if (fl.comp) {
                R pd = qv.div(g) * qv.pres(g);
                if (fl.mph) pd = pd * nvq;
                pdv_i = FMA(ni, dw * pd, pdv_i);
            }
        }
        R ug_i = (R)0;
        if (fl.adv && fl.ibp) {
#pragma unroll
            for (int d = 0; d < DIM; d++) ug_i = FMA(Gi[d], uq[d], ug_i);
        }
        R vK[DIM];
        if (fl.diff && !fl.iso) {
#pragma unroll
            for (int b = 0; b < DIM; b++) {
                R s = (R)0;
#pragma unroll
                for (int a = 0; a < DIM; a++) s = FMA(Gi[a], kt[a][b], s);
                vK[b] = s;
            }
        }
#pragma unroll
        for (int j = 0; j < LOC; j++) {
#ifdef __HIP_DEVICE_COMPILE__
            // keep each column's table reads in its own step (hoisting all
            // of them ahead of the arithmetic needs more registers than a
            // P2 row can have)
            if constexpr (LOC > DIM + 1) __asm__ __volatile__("" ::: "memory");
#endif
            const R nj = tb.n[j + LOC * g];
            const R nn = ni * nj;
            if (fl.have_mass) M[j] = FMA(nn, dwm, M[j]);
            if (mv) Mv[j] = FMA(nn, dwv, Mv[j]);
            if (fl.adv || fl.diff) {
                R Gj[DIM];
#pragma unroll
                for (int d = 0; d < DIM; d++) {
                    R s = (R)0;
#pragma unroll
                    for (int k = 0; k < DIM; k++) s = FMA(tb.dn[j + LOC * (g + ngi * k)], iJ[d][k], s);
                    Gj[d] = s;
                }
                if (fl.adv) {
                    R a;
                    if (!fl.ibp) {
                        R t = (R)0;
#pragma unroll
                        for (int d = 0; d < DIM; d++) t = FMA(Gj[d], uq[d], t);
                        a = ni * t;
                        if (fl.eq) a = a * ca;
                        if (fl.need_div) a = FMA((R)fl.beta * nn, cdiv, a);
                        if (mv) a = FMA(-nn, div_ug, a);
                    } else {
                        a = -(ug_i * nj);
                        if (fl.eq) a = a * ca;
                        if (fl.need_div) a = FMA(-((R)1 - (R)fl.beta) * nn, cdiv, a);
                    }
                    Sm[j] = FMA(a, dw, Sm[j]);
                }
                if (fl.diff) {
                    R t = (R)0;
                    if (fl.iso) {
#pragma unroll
                        for (int d = 0; d < DIM; d++) t = FMA(Gi[d], Gj[d], t);
                        Sm[j] = FMA(t, dk * kq, Sm[j]);
                    } else {
#pragma unroll
                        for (int d = 0; d < DIM; d++) t = FMA(vK[d], Gj[d], t);
                        Sm[j] = FMA(t, dw, Sm[j]);
                    }
                }
            }
            if (fl.absorp) Sm[j] = FMA(nn, dw * sq, Sm[j]);
        }
    }
    R r = (R)0;
#pragma unroll
    for (int j = 0; j < LOC; j++) r = FMA(-Sm[j], T[j], r);
    if (fl.src) r = r + src_i;
    if (fl.comp) r = r - pdv_i;
    if (mv) {  // - (N^{n+1} - N^n) T / dt (:922-938)
        R s = (R)0;
        if (fl.lump) {
#pragma unroll
            for (int j = 0; j < LOC; j++) s += Mv[j];
            s = s * T[i];
        } else {
#pragma unroll
            for (int j = 0; j < LOC; j++) s = FMA(Mv[j], T[j], s);
        }
        r = r - s / (R)fl.dt;
    }
    rhs_i = r;
    R lump = (R)0;
    if (fl.have_mass && fl.lump) {
#pragma unroll
        for (int j = 0; j < LOC; j++) lump += M[j];
    }
    const R dtte = fl.dtt_on ? dtt : (R)0;
#pragma unroll
    for (int j = 0; j < LOC; j++) {
        R m = (fl.have_mass && !fl.lump) ? M[j] : (R)0;
        if (fl.have_mass && fl.lump && j == i) m = lump;
        vals[j] = FMA(dtte, Sm[j], m);
    }
}

// Component d of the physical gradient of shape function i at point g:
// sum_k dn(i,g,k) invJ(d,k), accumulated in k order (the dshape of
// transform_to_physical, Transform_elements.F90:769-781).
template <int DIM, int LOC, typename R>
FLGPU_HD_INLINE R phys_grad(const LTab<R>& tb, const R (&iJ)[DIM][DIM], int i, int g, int d) {
    R s = (R)0;
#pragma unroll
    for (int k = 0; k < DIM; k++) s = FMA(tb.dn[i + LOC * (g + tb.ngi * k)], iJ[d][k], s);
    return s;
}

// Entries (i, j0) .. (i, j0+JB-1) of the quadratic-T element sums of
// row_general_q: the mass term, the summed advection/absorption/diffusion
// term, and row i's source integral, each entry accumulated over the
// quadrature points in the same order with the same operations as there
// (so the same bits); the row's point values are formed once for its JB
// entries.  gt: the physical gradients [(g*LOC + i)*DIM + d] when staged.
// FEAT: the terms compiled in (without F_ANISO the diffusivity is taken as
// isotropic whatever fl says; the caller picks a set that holds the flags).
template <int DIM, int LOC, typename R, class QV, int FEAT = F_ALL | F_ANISO, int JB = 1>
FLGPU_HD_INLINE void entry_general_q(const LTab<R>& tb, const AsmFlags& fl_in, const R (&iJ)[DIM][DIM], R adet,
                                     const QV& qv, int i, int j0, R (&m_ij)[JB], R (&s_ij)[JB], R& src_i,
                                     const R* gt = nullptr) {
    AsmFlags fl = fl_in;
    fl.iso = (FEAT & F_ANISO) ? fl.iso : 1;
    const int ngi = tb.ngi;
    R M[JB], Sm[JB], src = (R)0;
#pragma unroll
    for (int b = 0; b < JB; b++) M[b] = Sm[b] = (R)0;
#pragma unroll 1
    for (int g = 0; g < ngi; g++) {
        const R dw = adet * tb.w[g];
        const R ni = tb.n[i + LOC * g];
        R Gi[DIM];
        if (gt) {
#pragma unroll
            for (int d = 0; d < DIM; d++) Gi[d] = gt[(g * LOC + i) * DIM + d];
        } else {
#pragma unroll
            for (int d = 0; d < DIM; d++) Gi[d] = phys_grad<DIM, LOC, R>(tb, iJ, i, g, d);
        }
        R uq[DIM] = {};
        R div = (R)0;
        if (fl.adv) {
            qv.u(g, uq);
            if (fl.need_div) div = qv.div(g);
        }
        R kq = (R)0, sq = (R)0;
        R kt[DIM][DIM];
        if (fl.diff) {
            if (fl.iso)
                kq = qv.k_iso(g);
            else
                qv.k_tensor(g, kt);
        }
        if (fl.absorp) sq = qv.absorp(g);
        if (fl.src) src = FMA(ni, dw * qv.src(g), src);
        R ug_i = (R)0;
        if (fl.adv && fl.ibp) {
#pragma unroll
            for (int d = 0; d < DIM; d++) ug_i = FMA(Gi[d], uq[d], ug_i);
        }
        R vK[DIM];
        if (fl.diff && !fl.iso) {
#pragma unroll
            for (int b = 0; b < DIM; b++) {
                R s = (R)0;
#pragma unroll
                for (int a = 0; a < DIM; a++) s = FMA(Gi[a], kt[a][b], s);
                vK[b] = s;
            }
        }
        const R dwk = dw * kq;
#pragma unroll
        for (int b = 0; b < JB; b++) {
            const int j = j0 + b;
            const R nj = tb.n[j + LOC * g];
            const R nn = ni * nj;
            if (fl.have_mass) M[b] = FMA(nn, dw, M[b]);
            if (fl.adv || fl.diff) {
                R Gj[DIM];
                if (gt) {
#pragma unroll
                    for (int d = 0; d < DIM; d++) Gj[d] = gt[(g * LOC + j) * DIM + d];
                } else {
#pragma unroll
                    for (int d = 0; d < DIM; d++) Gj[d] = phys_grad<DIM, LOC, R>(tb, iJ, j, g, d);
                }
                if (fl.adv) {
                    R a;
                    if (!fl.ibp) {
                        R t = (R)0;
#pragma unroll
                        for (int d = 0; d < DIM; d++) t = FMA(Gj[d], uq[d], t);
                        a = ni * t;
                        if (fl.need_div) a = FMA((R)fl.beta * nn, div, a);
                    } else {
                        a = -(ug_i * nj);
                        if (fl.need_div) a = FMA(-((R)1 - (R)fl.beta) * nn, div, a);
                    }
                    Sm[b] = FMA(a, dw, Sm[b]);
                }
                if (fl.diff) {
                    R t = (R)0;
                    if (fl.iso) {
#pragma unroll
                        for (int d = 0; d < DIM; d++) t = FMA(Gi[d], Gj[d], t);
                        Sm[b] = FMA(t, dwk, Sm[b]);
                    } else {
#pragma unroll
                        for (int d = 0; d < DIM; d++) t = FMA(vK[d], Gj[d], t);
                        Sm[b] = FMA(t, dw, Sm[b]);
                    }
                }
            }
            if (fl.absorp) Sm[b] = FMA(nn, dw * sq, Sm[b]);
        }
    }
#pragma unroll
    for (int b = 0; b < JB; b++) {
        m_ij[b] = M[b];
        s_ij[b] = Sm[b];
    }
    src_i = src;
}

// ---------------------------------------------------------------------
// Upwind stabilisation (assemble/Upwind_Stabilisation.F90).

// nu_bar_scaled_q (:243-319) at one quadrature point.  transform_to_physical
// hands the stabilisation j = transpose(J_T) (Transform_elements.F90:
// 783-787), so (u j)_a = sum_c u_c JT[a][c]; K is the diffusivity tensor at
// the point, used when have_k (else xi = sign(Pe), the unity scheme).  The
// inverse of K by cofactors over the determinant (Vector_Tools.F90:231-283).
template <int DIM, typename R>
FLGPU_HD_INLINE R nu_bar_scaled(const R (&u)[DIM], const R (&JT)[DIM][DIM], const R (&K)[DIM][DIM], bool have_k,
                                int scheme, R scale) {
    const int sch = have_k ? scheme : 4;
    R norm_u = (R)0;
#pragma unroll
    for (int d = 0; d < DIM; d++) norm_u = FMA(u[d], u[d], norm_u);
    if (norm_u < (R)1.0e-10) return (R)0;
    R uJ[DIM];
#pragma unroll
    for (int a = 0; a < DIM; a++) {
        R t = (R)0;
#pragma unroll
        for (int c = 0; c < DIM; c++) t = FMA(u[c], JT[a][c], t);
        uJ[a] = t;
    }
    R acc = (R)0;
    if (sch == 4) {
#pragma unroll
        for (int a = 0; a < DIM; a++) acc += fabs(uJ[a]);
    } else {
        R C[DIM][DIM], det;
        if constexpr (DIM == 3) {
#pragma unroll
            for (int i = 0; i < 3; i++)
#pragma unroll
                for (int k = 0; k < 3; k++) {
                    const int i1 = (i + 1) % 3, i2 = (i + 2) % 3, k1 = (k + 1) % 3, k2 = (k + 2) % 3;
                    // C[i][k]: inverse(K) times det, entry (i, k)
                    C[i][k] = K[k1][i1] * K[k2][i2] - K[k2][i1] * K[k1][i2];
                }
            det = K[0][0] * C[0][0] + K[1][0] * C[0][1] + K[2][0] * C[0][2];
        } else if constexpr (DIM == 1) {
            C[0][0] = (R)1;
            det = K[0][0];
        } else {
            C[0][0] = K[1][1];
            C[1][1] = K[0][0];
            C[0][1] = -K[0][1];
            C[1][0] = -K[1][0];
            det = K[0][0] * K[1][1] - K[0][1] * K[1][0];
        }
#pragma unroll
        for (int a = 0; a < DIM; a++) {
            // Pe_a = 0.5 (u j K^-1)_a = 0.5 sum_b (u j)_b Kinv[b][a]
            R t = (R)0;
#pragma unroll
            for (int b = 0; b < DIM; b++) t = FMA(uJ[b], C[b][a], t);
            const R pe = (R)0.5 * (t / det);
            R xi;
            if (sch == 1) {  // xi_optimal (:128-160)
                if (fabs(pe) < (R)1.0e-10) xi = (R)0;
                else if (pe > (R)11.859499013855018) xi = (R)1 - (R)1 / pe;
                else if (pe < -(R)11.859499013855018) xi = -(R)1 - (R)1 / pe;
                else xi = (R)1 / tanh(pe) - (R)1 / pe;
            } else if (sch == 2) {  // doubly asymptotic (:162-194)
                if (fabs(pe) <= (R)3) xi = pe / (R)3;
                else xi = pe > (R)0 ? (R)1 : -(R)1;
            } else {  // critical rule (:196-228)
                if (fabs(pe) <= (R)1) xi = (R)0;
                else xi = (pe > (R)0 ? (R)1 : -(R)1) - (R)1 / pe;
            }
            acc = FMA(xi, uJ[a], acc);
        }
    }
    return (acc / norm_u) * scale;
}

// Row i with upwind stabilisation, any degree: the per-quadrature-point
// loop of row_general_q with (Advection_Diffusion_CG.F90:812-825) the SUPG
// test function N_i + nu_bar u.grad N_i in the mass, advection (its
// non-by-parts product and the divergence term), absorption and source
// terms, the diffusion and the by-parts u.grad N_i term keeping N_i; and
// (:1110-1121, Upwind_Stabilisation.F90:52-84) for streamline upwind the
// term nu_bar (u.grad N_i)(u.grad N_j) added to the advection matrix.
// CG: the tables' shape derivatives are the same at every quadrature point
// (linear T, checked on the host), so the physical gradients and the
// velocity divergence are formed once, not per point: the same values.
// FEAT: the terms compiled in (F_ANISO: anisotropic diffusivity); the caller
// picks a set that holds the flags switched on.
template <int DIM, int LOC, typename R, class QV, class TV, bool CG = false, int FEAT = F_ALL | F_ANISO>
FLGPU_HD_INLINE void row_stab_q(const LTab<R>& tb, const AsmFlags& fl_in, const R (&JT)[DIM][DIM],
                                const R (&iJ)[DIM][DIM], R adet, const QV& qv, const TV& T, int i, R (&vals)[LOC],
                                R& rhs_i) {
    AsmFlags fl = fl_in;
    fl.have_mass = (FEAT & F_MASS) ? fl.have_mass : 0;
    fl.adv = (FEAT & F_ADV) ? fl.adv : 0;
    fl.absorp = (FEAT & F_ABS) ? fl.absorp : 0;
    fl.diff = (FEAT & F_DIFF) ? fl.diff : 0;
    fl.src = (FEAT & F_SRC) ? fl.src : 0;
    fl.iso = (FEAT & F_ANISO) ? fl.iso : 1;
    const int ngi = tb.ngi;
    const bool supg = fl.stab == 2, su = fl.stab == 1 && fl.adv;
    R Gc[CG ? LOC : 1][DIM];
    R div_c = (R)0;
    if constexpr (CG) {
#pragma unroll
        for (int j = 0; j < LOC; j++)
#pragma unroll
            for (int d = 0; d < DIM; d++) {
                R s = (R)0;
#pragma unroll
                for (int k = 0; k < DIM; k++) s = FMA(tb.dn[j + LOC * (0 + ngi * k)], iJ[d][k], s);
                Gc[j][d] = s;
            }
        if (fl.adv && fl.need_div) div_c = qv.div(0);
    }
    R M[LOC], Sm[LOC];
#pragma unroll
    for (int j = 0; j < LOC; j++) {
        M[j] = (R)0;
        Sm[j] = (R)0;
    }
    R src_i = (R)0, pdv_i = (R)0;
#pragma unroll 1
    for (int g = 0; g < ngi; g++) {
        const R dw = adet * tb.w[g];
        const R ni = tb.n[i + LOC * g];
        R Gi[DIM];
        if constexpr (CG) {
            // select, not index: i is a run-time value
#pragma unroll
            for (int d = 0; d < DIM; d++) {
                R s = Gc[0][d];
#pragma unroll
                for (int j = 1; j < LOC; j++) s = i == j ? Gc[j][d] : s;
                Gi[d] = s;
            }
        } else {
#pragma unroll
            for (int d = 0; d < DIM; d++) {
                R s = (R)0;
#pragma unroll
                for (int k = 0; k < DIM; k++) s = FMA(tb.dn[i + LOC * (g + ngi * k)], iJ[d][k], s);
                Gi[d] = s;
            }
        }
        R uq[DIM];
        qv.u(g, uq);
        const R div = (fl.adv && fl.need_div) ? (CG ? div_c : qv.div(g)) : (R)0;
        R kt[DIM][DIM];
#pragma unroll
        for (int a = 0; a < DIM; a++)
#pragma unroll
            for (int b = 0; b < DIM; b++) kt[a][b] = (R)0;
        R kq = (R)0;
        if (fl.diff) {
            if (!fl.iso || QV::nu_reads_k) qv.k_tensor(g, kt);
            if (fl.iso) kq = qv.k_iso(g);
        }
        const R nu = qv.nu(g, uq, JT, kt, fl);
        R ug_i = (R)0;
#pragma unroll
        for (int d = 0; d < DIM; d++) ug_i = FMA(Gi[d], uq[d], ug_i);
        const R tn_i = supg ? FMA(nu, ug_i, ni) : ni;
        const R sq = fl.absorp ? qv.absorp(g) : (R)0;
        // the internal-energy / k-epsilon weights, as in row_general_q, with
        // the SUPG test function where the reference takes it (mass,
        // advection, the p div(u) term); the streamline-upwind term is not
        // density-weighted (:1105-1116)
        R dwm = dw, ca = (R)1, cdiv = div, dk = dw;
        if constexpr (QV::eq_fields) {
            if (fl.eq) {
                const R th = (R)fl.rho_theta;
                const R rq = qv.rho(g);
                R rb = rq, nvq = (R)1;
                if (fl.mph) nvq = qv.nv(g);
                if (fl.eq == 1) {
                    const R ro = qv.rho_old(g);
                    rb = th * rq + ((R)1 - th) * ro;
                    dwm = fl.mph ? dw * ro * nvq : dw * ro;
                } else {
                    dwm = dw * rq;
                }
                ca = fl.mph ? rb * nvq : rb;
                if (fl.adv && fl.need_div) {
                    R gr[DIM];
                    qv.grad_rho(g, fl.eq, th, gr);
                    R udgr = (R)0;
#pragma unroll
                    for (int d = 0; d < DIM; d++) udgr = FMA(gr[d], uq[d], udgr);
                    if (fl.mph) {
                        R gn[DIM];
                        qv.grad_nv(g, gn);
                        R udgn = (R)0;
#pragma unroll
                        for (int d = 0; d < DIM; d++) udgn = FMA(gn[d], uq[d], udgn);
                        cdiv = div * rb * nvq + nvq * udgr + rb * udgn;
                    } else {
                        cdiv = div * rb + udgr;
                    }
                }
                if (fl.mph) dk = dw * nvq;
                if (fl.comp) {
                    R pd = qv.div(g) * qv.pres(g);
                    if (fl.mph) pd = pd * nvq;
                    pdv_i = FMA(tn_i, dw * pd, pdv_i);
                }
            }
        }
        if (fl.src) src_i = FMA(tn_i, dw * qv.src(g), src_i);
        R vK[DIM];
        if (fl.diff && !fl.iso) {
#pragma unroll
            for (int b = 0; b < DIM; b++) {
                R s = (R)0;
#pragma unroll
                for (int a = 0; a < DIM; a++) s = FMA(Gi[a], kt[a][b], s);
                vK[b] = s;
            }
        }
#pragma unroll
        for (int j = 0; j < LOC; j++) {
#ifdef __HIP_DEVICE_COMPILE__
            if constexpr (LOC > DIM + 1) __asm__ __volatile__("" ::: "memory");
#endif
            const R nj = tb.n[j + LOC * g];
            const R tnn = tn_i * nj;
            if (fl.have_mass) M[j] = FMA(tnn, dwm, M[j]);
            R Gj[DIM];
#pragma unroll
            for (int d = 0; d < DIM; d++) {
                if constexpr (CG) {
                    Gj[d] = Gc[j][d];
                } else {
                    R s = (R)0;
#pragma unroll
                    for (int k = 0; k < DIM; k++) s = FMA(tb.dn[j + LOC * (g + ngi * k)], iJ[d][k], s);
                    Gj[d] = s;
                }
            }
            if (fl.adv) {
                R ugj = (R)0;
#pragma unroll
                for (int d = 0; d < DIM; d++) ugj = FMA(Gj[d], uq[d], ugj);
                R a;
                if (!fl.ibp) {
                    a = tn_i * ugj;
                    if (fl.eq) a = a * ca;
                    if (fl.need_div) a = FMA((R)fl.beta * tnn, cdiv, a);
                } else {
                    a = -(ug_i * nj);
                    if (fl.eq) a = a * ca;
                    if (fl.need_div) a = FMA(-((R)1 - (R)fl.beta) * tnn, cdiv, a);
                }
                if (su) a = FMA(ug_i * nu, ugj, a);
                Sm[j] = FMA(a, dw, Sm[j]);
            }
            if (fl.diff) {
                R t = (R)0;
                if (fl.iso) {
#pragma unroll
                    for (int d = 0; d < DIM; d++) t = FMA(Gi[d], Gj[d], t);
                    Sm[j] = FMA(t, dk * kq, Sm[j]);
                } else {
#pragma unroll
                    for (int d = 0; d < DIM; d++) t = FMA(vK[d], Gj[d], t);
                    Sm[j] = FMA(t, dw, Sm[j]);
                }
            }
            if (fl.absorp) Sm[j] = FMA(tnn, dw * sq, Sm[j]);
        }
    }
    R r = (R)0;
#pragma unroll
    for (int j = 0; j < LOC; j++) r = FMA(-Sm[j], T[j], r);
    if (fl.src) r = r + src_i;
    if (fl.comp) r = r - pdv_i;
    rhs_i = r;
    R lump = (R)0;
    if (fl.have_mass && fl.lump) {
#pragma unroll
        for (int j = 0; j < LOC; j++) lump += M[j];
    }
    const R dtte = fl.dtt_on ? (R)fl.dt_theta : (R)0;
#pragma unroll
    for (int j = 0; j < LOC; j++) {
        R m = (fl.have_mass && !fl.lump) ? M[j] : (R)0;
        if (fl.have_mass && fl.lump && j == i) m = lump;
        vals[j] = FMA(dtte, Sm[j], m);
    }
}

// Jacobian (JT[a][b] = dx_a/dxi_b), its inverse and |det| of the linear
// coordinate element.
template <int DIM, typename R, class TB = LTab<R>>
FLGPU_HD_INLINE void geometry_jt(const R (&X)[DIM][DIM + 1], const TB& tb, R (&JT)[DIM][DIM], R (&iJ)[DIM][DIM],
                                 R& adet) {
#pragma unroll
    for (int a = 0; a < DIM; a++)
#pragma unroll
        for (int b = 0; b < DIM; b++) {
            R s = (R)0;
#pragma unroll
            for (int nn = 0; nn < DIM + 1; nn++) s = FMA(X[a][nn], tb.xdn[nn + (DIM + 1) * b], s);
            JT[a][b] = s;
        }
    invert_jacobian<DIM, R>(JT, iJ, adet);
}

template <int DIM, int LOC, typename R>
FLGPU_HD_INLINE void row_stab(const Elem<DIM, LOC, R>& e, const LTab<R>& tb, const AsmFlags& fl,
                              const AsmFields& fd, int i, R (&vals)[LOC], R& rhs_i) {
    R JT[DIM][DIM], iJ[DIM][DIM], adet;
    geometry_jt<DIM, R>(e.X, tb, JT, iJ, adet);
    const GlobalNodal<DIM, R> na{fd, e.nodes};
    row_stab_q<DIM, LOC, R>(tb, fl, JT, iJ, adet, NodalQuad<DIM, LOC, R>{na, tb, iJ}, e.T, i, vals, rhs_i);
}

template <int DIM, int LOC, typename R, class NA>
FLGPU_HD_INLINE void row_general_na(const Elem<DIM, LOC, R>& e, const LTab<R>& tb, const AsmFlags& fl, const NA& na,
                                    int i, R (&vals)[LOC], R& rhs_i, const MoveData<DIM, R>* mv = nullptr) {
    R iJ[DIM][DIM], adet;
    geometry<DIM, R>(e.X, tb, iJ, adet);
    row_general_q<DIM, LOC, R>(tb, fl, iJ, adet, NodalQuad<DIM, LOC, R, NA>{na, tb, iJ}, e.T, i, vals, rhs_i, mv);
}

template <int DIM, int LOC, typename R>
FLGPU_HD void row_general(const Elem<DIM, LOC, R>& e, const LTab<R>& tb, const AsmFlags& fl,
                            const AsmFields& fd, int i, R (&vals)[LOC], R& rhs_i,
                            const MoveData<DIM, R>* mv = nullptr) {
    const GlobalNodal<DIM, R> na{fd, e.nodes};
    row_general_na<DIM, LOC, R>(e, tb, fl, na, i, vals, rhs_i, mv);
}

// An element's nodal field values staged in shared memory by its row
// threads (node k's values at base[k * kSharedNodalFields + ...]: velocity,
// the diffusivity tensor, source, absorption, density, old density,
// pressure, volume fraction), read with GlobalNodal's methods.
template <int DIM>
constexpr int shared_nodal_fields() {
    return DIM + DIM * DIM + 6;
}

template <int DIM, typename R>
struct SharedNodal {
    static constexpr bool eq_fields = true;
    static constexpr int NF = shared_nodal_fields<DIM>();
    const R* base;
    FLGPU_HD_INLINE R U(int k, int d) const { return base[k * NF + d]; }
    FLGPU_HD_INLINE R K(int k, int comp) const { return base[k * NF + DIM + comp]; }
    FLGPU_HD_INLINE R S(int k) const { return base[k * NF + DIM + DIM * DIM]; }
    FLGPU_HD_INLINE R B(int k) const { return base[k * NF + DIM + DIM * DIM + 1]; }
    FLGPU_HD_INLINE R Rho(int k) const { return base[k * NF + DIM + DIM * DIM + 2]; }
    FLGPU_HD_INLINE R RhoOld(int k) const { return base[k * NF + DIM + DIM * DIM + 3]; }
    FLGPU_HD_INLINE R P(int k) const { return base[k * NF + DIM + DIM * DIM + 4]; }
    FLGPU_HD_INLINE R V(int k) const { return base[k * NF + DIM + DIM * DIM + 5]; }
};

// Node `node`'s values into dst (the SharedNodal layout); only the fields
// the flags use are read, the others are zero.
template <int DIM, typename R>
FLGPU_HD_INLINE void stage_shared_nodal(const AsmFields& fd, const AsmFlags& fl, int node, R* dst) {
    const GlobalNodal<DIM, R> g{fd, &node};
    const bool u = fl.adv || fl.comp;
#pragma unroll
    for (int d = 0; d < DIM; d++) dst[d] = u ? g.U(0, d) : (R)0;
#pragma unroll
    for (int c = 0; c < DIM * DIM; c++) dst[DIM + c] = (fl.diff && (!fl.iso || c == 0)) ? g.K(0, c) : (R)0;
    R* t = dst + DIM + DIM * DIM;
    t[0] = fl.src ? g.S(0) : (R)0;
    t[1] = fl.absorp ? g.B(0) : (R)0;
    t[2] = fl.eq ? g.Rho(0) : (R)0;
    t[3] = fl.eq == 1 ? g.RhoOld(0) : (R)0;
    t[4] = fl.comp ? g.P(0) : (R)0;
    t[5] = fl.mph ? g.V(0) : (R)0;
}

template <int DIM, int LOC, typename R>
FLGPU_HD_INLINE void row_contrib(const Elem<DIM, LOC, R>& e, const LTab<R>& tb,
                                            const AsmFlags& fl, const AsmFields& fd, int i,
                                            R (&vals)[LOC], R& rhs_i) {
    if constexpr (LOC == DIM + 1)
        row_linear<DIM, LOC, R>(e, tb, fl, fd, i, vals, rhs_i);
    else
        row_general<DIM, LOC, R>(e, tb, fl, fd, i, vals, rhs_i);
}

}  // namespace flgpu
