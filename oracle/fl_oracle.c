/*
 * fl_oracle.c -- TEST INFRASTRUCTURE ONLY (the oracle).
 *
 * A plain-C restatement of Fluidity's CG advection-diffusion hot path, used
 * as the checker for the GPU library and as the timed CPU baseline.  It is
 * compiled with -ffp-contract=off so every expression is evaluated in the
 * order the Fortran source writes it (Fluidity's own release build adds
 * -ffast-math, so its bits are only defined up to reassociation).
 *
 * What pins each part (tests/test_oracle_*.py):
 *   orc_quadrature      monomial exactness p!/(p+d)! (femtools/tests/
 *                       test_quadrature.F90), sum w = 1/2 (tri) and 1/6
 *                       (tet), literal constants of Quadrature.F90.
 *   orc_shape_tables    partition of unity, sum of gradients = 0, nodal
 *                       interpolation of P2, exact integrals
 *                       (femtools/tests/test_shape_functions.F90).
 *   element arithmetic  closed forms: P1 stiffness |T| G G^T, consistent
 *                       mass |T|/((d+1)(d+2))(1+delta), zero row sums of K
 *                       and of the non-IBP advection matrix, P1/P2 patch
 *                       tests, sum detwei = |T| (test_jacobian.F90).
 *   orc_make_sparsity   brute-force dense set union on small meshes.
 *   orc_colour_elements validity, partition, and an independent greedy.
 *   orc_assemble        dense element-by-element sum on small meshes.
 *   orc_petsc_solve     dense reduced-system solve (numpy) on small meshes;
 *                       scipy spsolve on the 2-D case; true residual.
 * Iteration counts of PETSc KSPSolve are "parity unpinned": PETSc is not
 * vendored by the reference and only "converged" (reason > 0) is tested
 * there.
 */
#include "fl_oracle.h"

#include <float.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define MAXLOC 10
#define MAXGI 16
#define MAXDIM 3

static double wall_now(void)
{
#ifdef _OPENMP
    return omp_get_wtime();
#else
    return 0.0;
#endif
}

/* ------------------------------------------------------------------ */
/* Quadrature (femtools/Quadrature.F90)                                 */
/* ------------------------------------------------------------------ */

/* Expand one generator: permutation columns give, per point, the index
 * (1-based) into coords of each barycentric coordinate
 * (expand_quadrature_template, Quadrature.F90:572-605). */
static int expand(int nv, int npts, const int *perm, const double *coords,
                  double weight, double *l, double *w, int dk)
{
    for (int k = 0; k < npts; k++) {
        for (int j = 0; j < nv; j++) {
            int pj = perm[k * nv + j];
            l[(dk + k) * nv + j] = pj ? coords[pj - 1] : 0.0;
        }
        w[dk + k] = weight;
    }
    return dk + npts;
}

int orc_quadrature(int dim, int degree, double *l, double *w)
{
    /* tet_permutations(7) (centroid), (8) Quadrature.F90 tet perm table,
     * (9); tri_permutations(4), (5). */
    static const int tet7[4] = {1, 1, 1, 1};
    static const int tet8[16] = {1, 1, 1, 2, 1, 1, 2, 1, 1, 2, 1, 1, 2, 1, 1, 1};
    static const int tet9[24] = {1, 1, 2, 2, 1, 2, 1, 2, 1, 2, 2, 1,
                                 2, 1, 1, 2, 2, 1, 2, 1, 2, 2, 1, 1};
    static const int tri4[3] = {1, 1, 1};
    static const int tri5[9] = {1, 1, 2, 1, 2, 1, 2, 1, 1};
    double c[2];
    int dk = 0;
    if (dim == 3) {
        /* template chosen as the lowest degree >= requested
         * (Quadrature.F90:304-321) */
        if (degree <= 1) { /* Str71 1 point, :663-672 */
            c[0] = 0.25;
            dk = expand(4, 1, tet7, c, 0.166666666666666666666666666666666, l, w, dk);
        } else if (degree == 2) { /* str71 4 point, :676-688 */
            c[0] = 0.138196601125010515179541316563436;
            c[1] = 1.0 - 3.0 * c[0];
            dk = expand(4, 4, tet8, c, 0.041666666666666666666666666666666, l, w, dk);
        } else if (degree == 3) { /* str71 5 point, :692-707 */
            c[0] = 0.25;
            dk = expand(4, 1, tet7, c, -0.133333333333333333333333333333333, l, w, dk);
            c[0] = 0.166666666666666666666666666666666;
            c[1] = 1.0 - 3.0 * c[0];
            dk = expand(4, 4, tet8, c, 0.075, l, w, dk);
        } else if (degree == 4) { /* kea86 11 point, :711-735 */
            c[0] = 0.25;
            dk = expand(4, 1, tet7, c, -0.013155555555555555555555555555555, l, w, dk);
            c[0] = 0.0714285714285714285714285714285714;
            c[1] = 1.0 - 3.0 * c[0];
            dk = expand(4, 4, tet8, c, 7.62222222222222222222222222222222E-3, l, w, dk);
            c[0] = 0.399403576166799204996102147461640;
            c[1] = 0.5 - c[0];
            dk = expand(4, 6, tet9, c, 0.0248888888888888888888888888888888, l, w, dk);
        } else {
            return -1;
        }
        return dk;
    }
    if (dim == 2) {
        if (degree <= 1) { /* :921-931 */
            c[0] = 0.333333333333333333333333333333333;
            dk = expand(3, 1, tri4, c, 0.5, l, w, dk);
        } else if (degree == 2) { /* :935-949 */
            c[0] = 0.166666666666666666666666666666666;
            c[1] = 1.0 - 2.0 * c[0];
            dk = expand(3, 3, tri5, c, 0.166666666666666666666666666666666, l, w, dk);
        } else if (degree == 3) { /* str71 4 point, :953-970 */
            c[0] = 0.333333333333333333333333333333333;
            dk = expand(3, 1, tri4, c, -0.28125, l, w, dk);
            c[0] = 0.2;
            c[1] = 1.0 - 2.0 * c[0];
            dk = expand(3, 3, tri5, c, 0.260416666666666666666666666666666, l, w, dk);
        } else if (degree == 4) { /* 6 point, :974-993 */
            c[0] = 0.0915762135097707434595714634022015;
            c[1] = 1.0 - 2.0 * c[0];
            dk = expand(3, 3, tri5, c, 0.0549758718276609338191631624501052, l, w, dk);
            c[0] = 0.445948490915964886318329253883051;
            c[1] = 1.0 - 2.0 * c[0];
            dk = expand(3, 3, tri5, c, 0.111690794839005732847503504216561, l, w, dk);
        } else {
            return -1;
        }
        return dk;
    }
    if (dim == 1) {
        /* interval rules (:1140-1230); interval_permutations(1) = (1,1),
         * (2) = columns (1,2), (2,1) (:1893-1905) */
        static const int int1[2] = {1, 1};
        static const int int2[4] = {1, 2, 2, 1};
        if (degree <= 1) {
            c[0] = 0.5;
            dk = expand(2, 1, int1, c, 1.0, l, w, dk);
        } else if (degree == 2) {
            c[0] = 0.788675134594813;
            c[1] = 1.0 - c[0];
            dk = expand(2, 2, int2, c, 0.5, l, w, dk);
        } else if (degree == 3) {
            c[0] = 0.887298334620742;
            c[1] = 1.0 - c[0];
            dk = expand(2, 2, int2, c, 0.277777777777777, l, w, dk);
            c[0] = 0.5;
            dk = expand(2, 1, int1, c, 0.444444444444444, l, w, dk);
        } else if (degree == 4) {
            c[0] = 0.9305681557970262;
            c[1] = 1.0 - c[0];
            dk = expand(2, 2, int2, c, 0.173927422568727, l, w, dk);
            c[0] = 0.6699905217924281;
            c[1] = 1.0 - c[0];
            dk = expand(2, 2, int2, c, 0.326072577431273, l, w, dk);
        } else {
            return -1;
        }
        return dk;
    }
    return -1;
}

/* ------------------------------------------------------------------ */
/* Shape tables (Shape_Functions.F90:128-366, Elements.F90:615-717)     */
/* ------------------------------------------------------------------ */

/* number2count of the simplex Lagrange numbering: counts per barycentric
 * coordinate of each local node (Element_Numbering; the P2 order is
 * v1,e12,v2,e13,e23,v3,e14,e24,e34,v4 as in Elements.F90:527-537). */
static int simplex_counts(int dim, int degree, int counts[MAXLOC][4])
{
    int nv = dim + 1;
    memset(counts, 0, sizeof(int) * MAXLOC * 4);
    if (degree == 1) {
        for (int i = 0; i < nv; i++) counts[i][i] = 1;
        return nv;
    }
    if (degree == 2) {
        int k = 0;
        /* ordering: for vertex b = 1..nv: edges (a,b) a<b, then vertex b */
        for (int b = 0; b < nv; b++) {
            for (int a = 0; a < b; a++) {
                counts[k][a] = 1;
                counts[k][b] = 1;
                k++;
            }
            counts[k][b] = 2;
            k++;
        }
        return k;
    }
    return -1;
}

/* lagrange_polynomial(c, c, dx) = prod_{i<c} (x - i dx) / prod_{i<c} (c dx - i dx)
 * and its derivative (Shape_Functions.F90:331-366). */
static double lag(int c, double dx, double x)
{
    double num = 1.0, den = 1.0;
    for (int i = 0; i < c; i++) {
        num *= (x - i * dx);
        den *= (c * dx - i * dx);
    }
    return num / den;
}

static double dlag(int c, double dx, double x)
{
    double den = 1.0, s = 0.0;
    for (int i = 0; i < c; i++) den *= (c * dx - i * dx);
    for (int k = 0; k < c; k++) {
        double t = 1.0;
        for (int i = 0; i < c; i++)
            if (i != k) t *= (x - i * dx);
        s += t;
    }
    return s / den;
}

int orc_shape_tables(int dim, int degree, int ngi, const double *l, double *n,
                     double *dn)
{
    int counts[MAXLOC][4];
    int loc = simplex_counts(dim, degree, counts);
    int nv = dim + 1;
    double dx = 1.0 / degree;
    if (loc < 0) return -1;
    for (int i = 0; i < loc; i++) {
        for (int g = 0; g < ngi; g++) {
            const double *lg = l + g * nv;
            double v = 1.0;
            for (int j = 0; j < nv; j++) v *= lag(counts[i][j], dx, lg[j]);
            n[i + loc * g] = v;
            /* eval_dshape_simplex: dependent coordinate is the last one,
             * dl4dl = -1 (diffl4, Elements.F90:693-717) */
            for (int d = 0; d < dim; d++) {
                double e = dlag(counts[i][d], dx, lg[d]) *
                               lag(counts[i][dim], dx, lg[dim]) +
                           (-1.0) * dlag(counts[i][dim], dx, lg[dim]) *
                               lag(counts[i][d], dx, lg[d]);
                for (int j = 0; j < dim; j++) {
                    if (j == d) continue;
                    e *= lag(counts[i][j], dx, lg[j]);
                }
                dn[i + loc * (g + ngi * d)] = e;
            }
        }
    }
    return loc;
}

/* ------------------------------------------------------------------ */
/* Sparsity (Sparsity_Patterns.F90:47-85, :299-428)                    */
/* ------------------------------------------------------------------ */

static int cmp_int(const void *a, const void *b)
{
    int x = *(const int *)a, y = *(const int *)b;
    return (x > y) - (x < y);
}

/* node -> element adjacency (0-based), CSR */
static void node_elements(int nnodes, int nele, int loc, const int32_t *ndglno,
                          int **ptr_out, int **adj_out)
{
    int *ptr = calloc((size_t)nnodes + 1, sizeof(int));
    int *adj = malloc(sizeof(int) * (size_t)nele * loc);
    for (int64_t k = 0; k < (int64_t)nele * loc; k++) ptr[ndglno[k]]++;
    for (int i = 0; i < nnodes; i++) ptr[i + 1] += ptr[i];
    int *fill = malloc(sizeof(int) * (size_t)nnodes);
    memcpy(fill, ptr, sizeof(int) * (size_t)nnodes);
    for (int e = 0; e < nele; e++)
        for (int i = 0; i < loc; i++) {
            int nd = ndglno[(int64_t)e * loc + i] - 1;
            adj[fill[nd]++] = e;
        }
    free(fill);
    *ptr_out = ptr;
    *adj_out = adj;
}

int64_t orc_make_sparsity(int nnodes, int nele, int loc, const int32_t *ndglno,
                          int32_t *findrm, int32_t *colm, int32_t *centrm,
                          int64_t cap)
{
    int *ptr, *adj;
    node_elements(nnodes, nele, loc, ndglno, &ptr, &adj);
    int *buf = malloc(sizeof(int) * (size_t)(loc * 64 + 64));
    int bufcap = loc * 64 + 64;
    int64_t count = 1;
    for (int r = 0; r < nnodes; r++) {
        int nb = 0;
        int need = (ptr[r + 1] - ptr[r]) * loc;
        if (need > bufcap) {
            bufcap = need;
            buf = realloc(buf, sizeof(int) * (size_t)bufcap);
        }
        for (int k = ptr[r]; k < ptr[r + 1]; k++) {
            int e = adj[k];
            for (int j = 0; j < loc; j++) buf[nb++] = ndglno[(int64_t)e * loc + j];
        }
        /* insert_ascending of every node of every element containing r */
        qsort(buf, nb, sizeof(int), cmp_int);
        int u = 0;
        for (int k = 0; k < nb; k++)
            if (u == 0 || buf[k] != buf[u - 1]) buf[u++] = buf[k];
        findrm[r] = (int32_t)count;
        if (count - 1 + u > cap) {
            free(buf); free(ptr); free(adj);
            return -1;
        }
        centrm[r] = 0;
        for (int k = 0; k < u; k++) {
            colm[count - 1 + k] = buf[k];
            if (buf[k] == r + 1) centrm[r] = (int32_t)(count + k);
        }
        count += u;
    }
    findrm[nnodes] = (int32_t)count;
    free(buf); free(ptr); free(adj);
    return count - 1;
}

/* ------------------------------------------------------------------ */
/* Colouring (Colouring.F90:85-262)                                     */
/* ------------------------------------------------------------------ */

int orc_colour_elements(int nnodes, int nele, int loc, const int32_t *ndglno,
                        int32_t *colour)
{
    int *ptr, *adj;
    node_elements(nnodes, nele, loc, ndglno, &ptr, &adj);
    int ncol = 0;
    int cap = 64;
    int *seen = calloc((size_t)cap + 2, sizeof(int)); /* stamp per colour */
    for (int e = 0; e < nele; e++) {
        /* neighbours = elements sharing a node (second-order sparsity of
         * the P0 mesh against the topology); only lower ids count */
        for (int i = 0; i < loc; i++) {
            int nd = ndglno[(int64_t)e * loc + i] - 1;
            for (int k = ptr[nd]; k < ptr[nd + 1]; k++) {
                int f = adj[k];
                if (f < e) seen[colour[f]] = e + 1;
            }
        }
        int c = 1;
        while (c <= ncol && seen[c] == e + 1) c++;
        colour[e] = c;
        if (c > ncol) {
            ncol = c;
            if (ncol + 1 >= cap) {
                int nc = cap * 2;
                seen = realloc(seen, sizeof(int) * ((size_t)nc + 2));
                memset(seen + cap, 0, sizeof(int) * ((size_t)nc + 2 - cap));
                cap = nc;
            }
        }
    }
    free(seen); free(ptr); free(adj);
    return ncol;
}

/* ------------------------------------------------------------------ */
/* Element assembly (Advection_Diffusion_CG.F90:702-1202)               */
/* ------------------------------------------------------------------ */

/* ------------------------------------------------------------------ */
/* Upwind stabilisation (assemble/Upwind_Stabilisation.F90)             */
/* ------------------------------------------------------------------ */

enum { STAB_NONE = 0, STAB_SU = 1, STAB_SUPG = 2 };
enum { NU_OPTIMAL = 1, NU_DOUBLY_ASYMPTOTIC = 2, NU_CRITICAL = 3, NU_UNITY = 4 };
static const double upw_tolerance = 1.0e-10, tanh_tolerance = 11.859499013855018; /* :21-22 */

/* inverse (femtools/Vector_Tools.F90:231-283): cofactors over det */
static void inverse_small(int dim, double A[MAXDIM][MAXDIM], double B[MAXDIM][MAXDIM])
{
    if (dim == 3) {
        double det = A[0][0] * (A[1][1] * A[2][2] - A[2][1] * A[1][2]) -
                     A[1][0] * (A[0][1] * A[2][2] - A[2][1] * A[0][2]) +
                     A[2][0] * (A[0][1] * A[1][2] - A[1][1] * A[0][2]);
        B[0][0] = A[1][1] * A[2][2] - A[2][1] * A[1][2];
        B[0][1] = A[2][1] * A[0][2] - A[0][1] * A[2][2];
        B[0][2] = A[0][1] * A[1][2] - A[1][1] * A[0][2];
        B[1][0] = A[1][2] * A[2][0] - A[2][2] * A[1][0];
        B[1][1] = A[2][2] * A[0][0] - A[0][2] * A[2][0];
        B[1][2] = A[0][2] * A[1][0] - A[1][2] * A[0][0];
        B[2][0] = A[1][0] * A[2][1] - A[2][0] * A[1][1];
        B[2][1] = A[2][0] * A[0][1] - A[0][0] * A[2][1];
        B[2][2] = A[0][0] * A[1][1] - A[1][0] * A[0][1];
        for (int i = 0; i < 3; i++)
            for (int k = 0; k < 3; k++) B[i][k] = B[i][k] / det;
    } else if (dim == 2) {
        double det = A[0][0] * A[1][1] - A[0][1] * A[1][0];
        B[0][0] = A[1][1] / det;
        B[1][1] = A[0][0] / det;
        B[0][1] = -A[0][1] / det;
        B[1][0] = -A[1][0] / det;
    } else {
        B[0][0] = 1.0 / A[0][0];
    }
}

/* nu_bar_scaled_q (:243-319) at one quadrature point: u (dim), J with
 * J[c][a] = dx_a/dxi_c (transform_to_physical's j, the transpose of the
 * cached J_T, Transform_elements.F90:783-787), K the diffusivity there (used
 * when have_k; else xi = sign(pe), NU_BAR_UNITY). */
double orc_nu_bar_scaled(int dim, const double *u, const double *Jflat, const double *Kflat, int have_k,
                         int scheme, double scale)
{
    double J[MAXDIM][MAXDIM], K[MAXDIM][MAXDIM], Ki[MAXDIM][MAXDIM], JK[MAXDIM][MAXDIM];
    double uJ[MAXDIM], pe[MAXDIM], xi[MAXDIM];
    for (int c = 0; c < dim; c++)
        for (int a = 0; a < dim; a++) {
            J[c][a] = Jflat[c * dim + a];
            K[c][a] = have_k ? Kflat[c * dim + a] : 0.0;
        }
    int sch = have_k ? scheme : NU_UNITY;
    double norm_u = 0.0, nu;
    for (int d = 0; d < dim; d++) norm_u = norm_u + u[d] * u[d];
    for (int a = 0; a < dim; a++) { /* matmul(u, j_mat) */
        double t = 0.0;
        for (int c = 0; c < dim; c++) t = t + u[c] * J[c][a];
        uJ[a] = t;
    }
    if (norm_u < upw_tolerance) {
        nu = 0.0;
    } else if (sch == NU_UNITY) {
        double t = 0.0;
        for (int a = 0; a < dim; a++) t = t + fabs(uJ[a]);
        nu = t / norm_u;
    } else {
        /* pe = 0.5 * matmul(u, matmul(j_mat, inverse(diff_q))) (:148, :182, :215) */
        inverse_small(dim, K, Ki);
        for (int c = 0; c < dim; c++)
            for (int a = 0; a < dim; a++) {
                double t = 0.0;
                for (int b = 0; b < dim; b++) t = t + J[c][b] * Ki[b][a];
                JK[c][a] = t;
            }
        for (int a = 0; a < dim; a++) {
            double t = 0.0;
            for (int c = 0; c < dim; c++) t = t + u[c] * JK[c][a];
            pe[a] = 0.5 * t;
        }
        for (int a = 0; a < dim; a++) {
            double p = pe[a];
            if (sch == NU_OPTIMAL) { /* xi_optimal :128-160 */
                if (fabs(p) < upw_tolerance) xi[a] = 0.0;
                else if (p > tanh_tolerance) xi[a] = 1.0 - (1.0 / p);
                else if (p < -tanh_tolerance) xi[a] = -1.0 - (1.0 / p);
                else xi[a] = (1.0 / tanh(p)) - (1.0 / p);
            } else if (sch == NU_DOUBLY_ASYMPTOTIC) { /* :162-194 */
                if (fabs(p) <= 3.0) xi[a] = p / 3.0;
                else if (p > 0.0) xi[a] = 1.0;
                else xi[a] = -1.0;
            } else { /* critical rule :196-228 */
                if (fabs(p) <= 1.0) xi[a] = 0.0;
                else if (p > 0.0) xi[a] = 1.0 - 1.0 / p;
                else xi[a] = -1.0 - 1.0 / p;
            }
        }
        double t = 0.0;
        for (int a = 0; a < dim; a++) t = t + xi[a] * uJ[a];
        nu = t / norm_u;
    }
    return nu * scale;
}

#define N_(i, g) (m->n[(i) + loc * (g)])
#define DN_(i, g, d) (m->dn[(i) + loc * ((g) + ngi * (d))])
#define XDN_(i, g, d) (m->x_dn[(i) + lx * ((g) + ngi * (d))])

/* transform_to_physical's cached geometry of a linear element
 * (Transform_elements.F90:227-306 construct_cache, :766-802 retrieval):
 * J^T from the coordinate element's dn, its inverse by cofactors over the
 * determinant. */
static void elem_geometry_of(const orc_mesh *m, const double *X, int ele, double JT[MAXDIM][MAXDIM],
                             double iJ[MAXDIM][MAXDIM], double *detJ_out)
{
    const int dim = m->dim, ngi = m->ngi, lx = m->loc_x;
    const int32_t *exn = m->ndglno_x + (int64_t)ele * lx;
    double Xv[MAXDIM][MAXLOC];
    for (int n = 0; n < lx; n++)
        for (int a = 0; a < dim; a++)
            Xv[a][n] = X[a + dim * (int64_t)(exn[n] - 1)];
    for (int a = 0; a < dim; a++)
        for (int b = 0; b < dim; b++) {
            double s = 0.0;
            for (int n = 0; n < lx; n++) s = s + Xv[a][n] * XDN_(n, 0, b);
            JT[a][b] = s;
        }
    if (dim == 3) {
        for (int i = 0; i < 3; i++)
            for (int k = 0; k < 3; k++)
                iJ[i][k] = JT[(i + 1) % 3][(k + 1) % 3] * JT[(i + 2) % 3][(k + 2) % 3] -
                           JT[(i + 2) % 3][(k + 1) % 3] * JT[(i + 1) % 3][(k + 2) % 3];
    } else if (dim == 2) {
        iJ[0][0] = JT[1][1];
        iJ[1][0] = -JT[0][1];
        iJ[0][1] = -JT[1][0];
        iJ[1][1] = JT[0][0];
    } else {
        iJ[0][0] = 1.0;
    }
    double detJ = 0.0;
    for (int a = 0; a < dim; a++) detJ = detJ + JT[a][0] * iJ[a][0];
    for (int i = 0; i < dim; i++)
        for (int k = 0; k < dim; k++) iJ[i][k] = iJ[i][k] / detJ;
    *detJ_out = detJ;
}

static void elem_geometry(const orc_mesh *m, int ele, double JT[MAXDIM][MAXDIM], double iJ[MAXDIM][MAXDIM],
                          double *detJ_out)
{
    elem_geometry_of(m, m->X, ele, JT, iJ, detJ_out);
}

enum { EQ_ADVECTION_DIFFUSION = 0, EQ_INTERNAL_ENERGY = 1, EQ_KEPSILON = 2 };

/* ele_val_at_quad of a scalar field on T's mesh (constant: one value) */
static void scalar_at_quad(const orc_mesh *m, const double *f, int fc, const int32_t *en, double *q)
{
    for (int g = 0; g < m->ngi; g++) {
        double s = 0.0;
        for (int i = 0; i < m->loc; i++) {
            double v = fc ? f[0] : f[en[i] - 1];
            s = s + v * m->n[i + m->loc * g];
        }
        q[g] = s;
    }
}

/* ele_grad_at_quad_scalar (Fields_Base.F90:2494-2510) with the field's
 * shape = T's (drho_t = dt_t, Advection_Diffusion_CG.F90:794-809):
 * quad_grad(d,:) = matmul(ele_val, dn(:,:,d)) */
static void scalar_grad_at_quad(const orc_mesh *m, const double *f, int fc, const int32_t *en,
                                double dsh[MAXLOC][MAXGI][MAXDIM], double gq[MAXDIM][MAXGI])
{
    for (int d = 0; d < m->dim; d++)
        for (int g = 0; g < m->ngi; g++) {
            double s = 0.0;
            for (int i = 0; i < m->loc; i++) {
                double v = fc ? f[0] : f[en[i] - 1];
                s = s + v * dsh[i][g][d];
            }
            gq[d][g] = s;
        }
}

static int element_cg(const orc_params *p, const orc_mesh *m, int ele,
                      double *mat, double *rhs)
{
    const int dim = m->dim, loc = m->loc, ngi = m->ngi, lx = m->loc_x;
    const int32_t *en = m->ndglno + (int64_t)ele * loc;
    const int32_t *exn = m->ndglno_x + (int64_t)ele * lx;
    const double eps = DBL_EPSILON; /* epsilon(0.0) with real*8 */
    const double dt_theta = p->dt * p->theta;
    double Xv[MAXDIM][MAXLOC], JT[MAXDIM][MAXDIM], iJ[MAXDIM][MAXDIM];
    double detJ;
    double dsh[MAXLOC][MAXGI][MAXDIM];
    double detwei[MAXGI];
    double Te[MAXLOC];

    for (int i = 0; i < loc * loc; i++) mat[i] = 0.0;
    for (int i = 0; i < loc; i++) rhs[i] = 0.0;

    /* --- transform_to_physical, cache path --- */
    (void)Xv;
    elem_geometry(m, ele, JT, iJ, &detJ);

    /* dshape = dn . invJ^T; once for linear T (m_nonlinear false),
     * per quadrature point otherwise (Transform_elements.F90:769-781) */
    const int m_nonlinear = (loc != dim + 1);
    for (int g = 0; g < ngi; g++) {
        int gs = m_nonlinear ? g : 0;
        for (int n = 0; n < loc; n++)
            for (int d = 0; d < dim; d++) {
                double s = 0.0;
                for (int k = 0; k < dim; k++) s = s + DN_(n, gs, k) * iJ[d][k];
                dsh[n][g][d] = s;
            }
    }
    for (int g = 0; g < ngi; g++) detwei[g] = fabs(detJ) * m->w[g];

    for (int i = 0; i < loc; i++) Te[i] = m->T[en[i] - 1];

    /* --- mesh movement (:421-437, :788-791): detwei of the old and new
     * positions, the grid velocity (on the coordinate mesh) and its
     * divergence at the points; linear T here, so the grid velocity's shape
     * and derivatives are T's.  The reference refuses it with stabilisation
     * and with the internal-energy and k-epsilon equations (:447, :462,
     * :500, :531). --- */
    const int mv = p->move_mesh && p->have_mass;
    double detwei_old[MAXGI], detwei_new[MAXGI], ugq[MAXDIM][MAXGI], divug[MAXGI];
    if (mv) {
        if (m_nonlinear || p->stabilisation != STAB_NONE || p->equation_type != 0) return -3;
        double JTo[MAXDIM][MAXDIM], iJo[MAXDIM][MAXDIM], dto, dtn;
        elem_geometry_of(m, m->X_old, ele, JTo, iJo, &dto);
        elem_geometry_of(m, m->X_new, ele, JTo, iJo, &dtn);
        for (int g = 0; g < ngi; g++) {
            detwei_old[g] = fabs(dto) * m->w[g];
            detwei_new[g] = fabs(dtn) * m->w[g];
            divug[g] = 0.0;
        }
        for (int d = 0; d < dim; d++)
            for (int g = 0; g < ngi; g++) {
                double s = 0.0;
                for (int i = 0; i < loc; i++) s = s + m->ug[d + dim * (int64_t)(exn[i] - 1)] * N_(i, g);
                ugq[d][g] = s;
            }
        for (int d = 0; d < dim; d++)
            for (int g = 0; g < ngi; g++) {
                double s = 0.0;
                for (int i = 0; i < loc; i++) s = s + m->ug[d + dim * (int64_t)(exn[i] - 1)] * dsh[i][g][d];
                divug[g] = divug[g] + s;
            }
    }

    /* --- the equation type's fields (:484-552, :983-1004) --- */
    const int eqt = p->equation_type;
    const int ie = eqt == EQ_INTERNAL_ENERGY, ke = eqt == EQ_KEPSILON;
    const int mph = ie && p->multiphase;
    double rhoq[MAXGI], rhooq[MAXGI], nvq[MAXGI];
    if (ie || ke) scalar_at_quad(m, m->rho, m->rho_const, en, rhoq);
    if (ie) scalar_at_quad(m, m->rho_old, m->rho_old_const, en, rhooq);
    if (mph) scalar_at_quad(m, m->nvfrac, m->nvfrac_const, en, nvq);

    /* --- test function (:812-825): N, or the SUPG shape
     * N + nu_bar u.grad N (supg_test_function, Upwind_Stabilisation.F90:
     * 420-452); with it, for streamline upwind, u.grad N and nu_bar --- */
    double tn[MAXLOC][MAXGI], udn[MAXLOC][MAXGI], nuq[MAXGI];
    for (int i = 0; i < loc; i++)
        for (int g = 0; g < ngi; g++) tn[i][g] = N_(i, g);
    if (p->stabilisation != STAB_NONE) {
        double Jf[MAXDIM * MAXDIM];
        for (int c = 0; c < dim; c++)
            for (int a = 0; a < dim; a++) Jf[c * dim + a] = JT[a][c]; /* j = transpose(J_T) */
        for (int g = 0; g < ngi; g++) {
            double u[MAXDIM], Kf[MAXDIM * MAXDIM];
            for (int d = 0; d < dim; d++) {
                double s = 0.0;
                for (int i = 0; i < loc; i++) {
                    double ui = m->u_const ? m->U[d] : m->U[d + dim * (int64_t)(en[i] - 1)];
                    s = s + ui * N_(i, g);
                }
                u[d] = s;
            }
            if (p->have_diffusivity)
                for (int a = 0; a < dim; a++)
                    for (int b = 0; b < dim; b++) {
                        double s = 0.0;
                        for (int i = 0; i < loc; i++) {
                            double v = m->kappa_const
                                           ? m->kappa[a + dim * b]
                                           : m->kappa[a + dim * b + dim * dim * (int64_t)(en[i] - 1)];
                            s = s + v * N_(i, g);
                        }
                        Kf[a * dim + b] = s;
                    }
            nuq[g] = orc_nu_bar_scaled(dim, u, Jf, Kf, p->have_diffusivity, p->nu_bar_scheme, p->nu_bar_scale);
            for (int j = 0; j < loc; j++) { /* u_nl_dn(j,i) = dot_product(u, dshape(j,i,:)) */
                double t = 0.0;
                for (int d = 0; d < dim; d++) t = t + u[d] * dsh[j][g][d];
                udn[j][g] = t;
            }
            if (p->stabilisation == STAB_SUPG)
                for (int i = 0; i < loc; i++) tn[i][g] = N_(i, g) + nuq[g] * udn[i][g];
        }
    }
#define TN_(i, g) (tn[i][g])

    /* --- mass (add_mass_element_cg, :867-941) --- */
    if (p->have_mass) {
        /* internal energy: detwei*olddensity[*nvfrac]; k-epsilon:
         * detwei*density (:884-903) */
        double wm[MAXGI], M[MAXLOC][MAXLOC];
        for (int g = 0; g < ngi; g++) {
            if (mph)
                wm[g] = (detwei[g] * rhooq[g]) * nvq[g];
            else if (ie)
                wm[g] = detwei[g] * rhooq[g];
            else if (ke)
                wm[g] = detwei[g] * rhoq[g];
            else
                wm[g] = mv ? detwei_new[g] : detwei[g];
        }
        for (int i = 0; i < loc; i++)
            for (int j = 0; j < loc; j++) {
                double s = 0.0;
                for (int g = 0; g < ngi; g++) s = s + (TN_(i, g) * N_(j, g)) * wm[g];
                M[i][j] = s;
            }
        if (p->lump_mass) {
            for (int i = 0; i < loc; i++) {
                double s = 0.0;
                for (int j = 0; j < loc; j++) s = s + M[i][j];
                mat[i + loc * i] = mat[i + loc * i] + s;
            }
        } else {
            for (int i = 0; i < loc; i++)
                for (int j = 0; j < loc; j++)
                    mat[i + loc * j] = mat[i + loc * j] + M[i][j];
        }
        if (mv) { /* (N^{n+1} - N^n) T^n / dt on the rhs (:922-938) */
            for (int i = 0; i < loc; i++)
                for (int j = 0; j < loc; j++) {
                    double s = 0.0;
                    for (int g = 0; g < ngi; g++) s = s + (TN_(i, g) * N_(j, g)) * (detwei_new[g] - detwei_old[g]);
                    M[i][j] = s;
                }
            for (int i = 0; i < loc; i++) {
                double s = 0.0;
                if (p->lump_mass) {
                    for (int j = 0; j < loc; j++) s = s + M[i][j];
                    s = s * Te[i];
                } else {
                    for (int j = 0; j < loc; j++) s = s + M[i][j] * Te[j];
                }
                rhs[i] = rhs[i] - s / p->dt;
            }
        }
    }

    /* --- advection (add_advection_element_cg, :943-1127) --- */
    if (p->have_advection) {
        double uq[MAXDIM][MAXGI], divq[MAXGI], adv[MAXLOC][MAXLOC];
        for (int d = 0; d < dim; d++)
            for (int g = 0; g < ngi; g++) {
                double s = 0.0;
                for (int i = 0; i < loc; i++) {
                    double ui = m->u_const ? m->U[d] : m->U[d + dim * (int64_t)(en[i] - 1)];
                    s = s + ui * N_(i, g);
                }
                uq[d][g] = s;
            }
        if (mv) /* the velocity relative to the grid (:983-985) */
            for (int d = 0; d < dim; d++)
                for (int g = 0; g < ngi; g++) uq[d][g] = uq[d][g] - ugq[d][g];
        int need_div = p->integrate_by_parts ? (fabs(1.0 - p->beta) > eps)
                                             : (fabs(p->beta) > eps);
        /* density at the points and u.grad(density) (:986-1004): internal
         * energy theta-weighted between density and old density, k-epsilon
         * density; wa = detwei*density[*nvfrac] weights the main term */
        double rbq[MAXGI], udgr[MAXGI], udgnv[MAXGI], wa[MAXGI];
        if (ie || ke) {
            double gr[MAXDIM][MAXGI], gro[MAXDIM][MAXGI];
            scalar_grad_at_quad(m, m->rho, m->rho_const, en, dsh, gr);
            if (ie) scalar_grad_at_quad(m, m->rho_old, m->rho_old_const, en, dsh, gro);
            for (int g = 0; g < ngi; g++) {
                rbq[g] = ie ? p->density_theta * rhoq[g] + (1. - p->density_theta) * rhooq[g] : rhoq[g];
                double s = 0.0;
                for (int d = 0; d < dim; d++) {
                    double gd = ie ? p->density_theta * gr[d][g] + (1. - p->density_theta) * gro[d][g] : gr[d][g];
                    s = s + gd * uq[d][g];
                }
                udgr[g] = s;
            }
            if (mph) {
                double gnv[MAXDIM][MAXGI];
                scalar_grad_at_quad(m, m->nvfrac, m->nvfrac_const, en, dsh, gnv);
                for (int g = 0; g < ngi; g++) {
                    double s = 0.0;
                    for (int d = 0; d < dim; d++) s = s + gnv[d][g] * uq[d][g];
                    udgnv[g] = s;
                }
            }
        }
        for (int g = 0; g < ngi; g++)
            wa[g] = mph ? (detwei[g] * rbq[g]) * nvq[g] : (ie || ke) ? detwei[g] * rbq[g] : detwei[g];
        if (need_div) { /* ele_div_at_quad, Fields_Base.F90:2557-2575 */
            for (int g = 0; g < ngi; g++) divq[g] = 0.0;
            for (int d = 0; d < dim; d++)
                for (int g = 0; g < ngi; g++) {
                    double s = 0.0;
                    for (int i = 0; i < loc; i++) {
                        double ui = m->u_const ? m->U[d] : m->U[d + dim * (int64_t)(en[i] - 1)];
                        s = s + ui * dsh[i][g][d];
                    }
                    divq[g] = divq[g] + s;
                }
        }
        for (int i = 0; i < loc; i++)
            for (int j = 0; j < loc; j++) {
                double s = 0.0;
                if (!p->integrate_by_parts) {
                    /* shape_vector_dot_dshape (FETools.F90:749-772) */
                    for (int g = 0; g < ngi; g++) {
                        double t = 0.0;
                        for (int d = 0; d < dim; d++) t = t + dsh[j][g][d] * uq[d][g];
                        s = s + (TN_(i, g) * t) * wa[g];
                    }
                } else {
                    /* -dshape_dot_vector_shape (FETools.F90:700-721) */
                    for (int g = 0; g < ngi; g++) {
                        double t = 0.0;
                        for (int d = 0; d < dim; d++) t = t + dsh[i][g][d] * uq[d][g];
                        s = s + (t * N_(j, g)) * wa[g];
                    }
                    s = -s;
                }
                adv[i][j] = s;
            }
        if (need_div) {
            double coef = p->integrate_by_parts ? (1.0 - p->beta) : p->beta;
            /* the div(density u) split (:1024-1033, :1066-1081): one
             * weight, or for multiphase three shape_shape terms summed */
            double wd[3][MAXGI];
            int nd = 1;
            for (int g = 0; g < ngi; g++) {
                if (mph) {
                    wd[0][g] = ((detwei[g] * divq[g]) * rbq[g]) * nvq[g];
                    wd[1][g] = (detwei[g] * nvq[g]) * udgr[g];
                    wd[2][g] = (detwei[g] * rbq[g]) * udgnv[g];
                } else if (ie || ke) {
                    wd[0][g] = (divq[g] * rbq[g] + udgr[g]) * detwei[g];
                } else {
                    wd[0][g] = divq[g] * detwei[g];
                }
            }
            if (mph) nd = 3;
            for (int i = 0; i < loc; i++)
                for (int j = 0; j < loc; j++) {
                    double s = 0.0;
                    for (int q = 0; q < nd; q++) {
                        double sq = 0.0;
                        for (int g = 0; g < ngi; g++) sq = sq + (TN_(i, g) * N_(j, g)) * wd[q][g];
                        s = q == 0 ? sq : s + sq;
                    }
                    if (p->integrate_by_parts)
                        adv[i][j] = adv[i][j] - coef * s;
                    else
                        adv[i][j] = adv[i][j] + coef * s;
                }
        }
        if (mv && !p->integrate_by_parts) /* - N_A div(u_grid) N_B (:1099-1102) */
            for (int i = 0; i < loc; i++)
                for (int j = 0; j < loc; j++) {
                    double s = 0.0;
                    for (int g = 0; g < ngi; g++) s = s + (TN_(i, g) * N_(j, g)) * (divug[g] * detwei[g]);
                    adv[i][j] = adv[i][j] - s;
                }
        if (p->stabilisation == STAB_SU) {
            /* element_upwind_stabilisation (Upwind_Stabilisation.F90:52-84) */
            for (int i = 0; i < loc; i++)
                for (int j = 0; j < loc; j++) {
                    double s = 0.0;
                    for (int g = 0; g < ngi; g++) s = s + ((udn[i][g] * detwei[g]) * nuq[g]) * udn[j][g];
                    adv[i][j] = adv[i][j] + s;
                }
        }
        if (fabs(dt_theta) > eps)
            for (int i = 0; i < loc; i++)
                for (int j = 0; j < loc; j++)
                    mat[i + loc * j] = mat[i + loc * j] + dt_theta * adv[i][j];
        for (int i = 0; i < loc; i++) {
            double s = 0.0;
            for (int j = 0; j < loc; j++) s = s + adv[i][j] * Te[j];
            rhs[i] = rhs[i] - s;
        }
    }

    /* --- absorption (add_absorption_element_cg, :1143-1162) --- */
    if (p->have_absorption) {
        double sq[MAXGI], ab[MAXLOC][MAXLOC];
        for (int g = 0; g < ngi; g++) {
            double s = 0.0;
            for (int i = 0; i < loc; i++) {
                double v = m->absorp_const ? m->absorp[0] : m->absorp[en[i] - 1];
                s = s + v * N_(i, g);
            }
            sq[g] = detwei[g] * s;
        }
        for (int i = 0; i < loc; i++)
            for (int j = 0; j < loc; j++) {
                double s = 0.0;
                for (int g = 0; g < ngi; g++) s = s + (TN_(i, g) * N_(j, g)) * sq[g];
                ab[i][j] = s;
            }
        if (fabs(dt_theta) > eps)
            for (int i = 0; i < loc; i++)
                for (int j = 0; j < loc; j++)
                    mat[i + loc * j] = mat[i + loc * j] + dt_theta * ab[i][j];
        for (int i = 0; i < loc; i++) {
            double s = 0.0;
            for (int j = 0; j < loc; j++) s = s + ab[i][j] * Te[j];
            rhs[i] = rhs[i] - s;
        }
    }

    /* --- diffusivity (add_diffusivity_element_cg, :1164-1202) --- */
    if (p->have_diffusivity) {
        double kq[MAXDIM][MAXDIM][MAXGI], df[MAXLOC][MAXLOC];
        for (int a = 0; a < dim; a++)
            for (int b = 0; b < dim; b++)
                for (int g = 0; g < ngi; g++) {
                    double s = 0.0;
                    for (int i = 0; i < loc; i++) {
                        double v = m->kappa_const
                                       ? m->kappa[a + dim * b]
                                       : m->kappa[a + dim * b + dim * dim * (int64_t)(en[i] - 1)];
                        s = s + v * N_(i, g);
                    }
                    kq[a][b][g] = s;
                }
        for (int i = 0; i < loc; i++)
            for (int j = 0; j < loc; j++) df[i][j] = 0.0;
        if (p->isotropic_diffusivity) {
            /* dshape_dot_dshape (FETools.F90:391-453): gi outermost */
            /* multiphase internal energy: times nvfrac (:1181-1190) */
            double wk[MAXGI];
            for (int g = 0; g < ngi; g++) wk[g] = mph ? (detwei[g] * kq[0][0][g]) * nvq[g] : detwei[g] * kq[0][0][g];
            for (int g = 0; g < ngi; g++)
                for (int i = 0; i < loc; i++)
                    for (int j = 0; j < loc; j++) {
                        double t;
                        if (dim == 3)
                            t = ((dsh[i][g][0] * dsh[j][g][0]) + (dsh[i][g][1] * dsh[j][g][1])) +
                                (dsh[i][g][2] * dsh[j][g][2]);
                        else if (dim == 2)
                            t = (dsh[i][g][0] * dsh[j][g][0]) + (dsh[i][g][1] * dsh[j][g][1]);
                        else
                            t = dsh[i][g][0] * dsh[j][g][0];
                        df[i][j] = df[i][j] + t * wk[g];
                    }
        } else {
            /* dshape_tensor_dshape (FETools.F90:668-698) */
            for (int g = 0; g < ngi; g++)
                for (int i = 0; i < loc; i++)
                    for (int j = 0; j < loc; j++) {
                        double v[MAXDIM], t = 0.0;
                        for (int b = 0; b < dim; b++) {
                            double s = 0.0;
                            for (int a = 0; a < dim; a++) s = s + dsh[i][g][a] * kq[a][b][g];
                            v[b] = s;
                        }
                        for (int b = 0; b < dim; b++) t = t + v[b] * dsh[j][g][b];
                        df[i][j] = df[i][j] + t * detwei[g];
                    }
        }
        if (fabs(dt_theta) > eps)
            for (int i = 0; i < loc; i++)
                for (int j = 0; j < loc; j++)
                    mat[i + loc * j] = mat[i + loc * j] + dt_theta * df[i][j];
        for (int i = 0; i < loc; i++) {
            double s = 0.0;
            for (int j = 0; j < loc; j++) s = s + df[i][j] * Te[j];
            rhs[i] = rhs[i] - s;
        }
    }

    /* --- source (add_source_element_cg, :1129-1141; shape_rhs
     * FETools.F90:38-54) --- */
    if (p->have_source) {
        double fw[MAXGI];
        for (int g = 0; g < ngi; g++) {
            double s = 0.0;
            for (int i = 0; i < loc; i++) {
                double v = m->src_const ? m->src[0] : m->src[en[i] - 1];
                s = s + v * N_(i, g);
            }
            fw[g] = detwei[g] * s;
        }
        for (int i = 0; i < loc; i++) {
            double s = 0.0;
            for (int g = 0; g < ngi; g++) s = s + TN_(i, g) * fw[g];
            rhs[i] = rhs[i] + s;
        }
    }

    /* --- pressure (add_pressurediv_element_cg, :1204-1226): compressible
     * internal energy, rhs -= N_i div(u) p detwei [nvfrac] --- */
    if (ie && p->compressible) {
        double dv[MAXGI], pq[MAXGI], wp[MAXGI];
        for (int g = 0; g < ngi; g++) dv[g] = 0.0;
        for (int d = 0; d < dim; d++)
            for (int g = 0; g < ngi; g++) {
                double s = 0.0;
                for (int i = 0; i < loc; i++) {
                    double ui = m->u_const ? m->U[d] : m->U[d + dim * (int64_t)(en[i] - 1)];
                    s = s + ui * dsh[i][g][d];
                }
                dv[g] = dv[g] + s;
            }
        scalar_at_quad(m, m->pres, m->pres_const, en, pq);
        for (int g = 0; g < ngi; g++) wp[g] = mph ? ((dv[g] * pq[g]) * detwei[g]) * nvq[g] : (dv[g] * pq[g]) * detwei[g];
        for (int i = 0; i < loc; i++) {
            double s = 0.0;
            for (int g = 0; g < ngi; g++) s = s + TN_(i, g) * wp[g];
            rhs[i] = rhs[i] - s;
        }
    }
#undef TN_
    return 0;
}

int orc_element(const orc_params *p, const orc_mesh *m, int ele, double *mat,
                double *rhs)
{
    return element_cg(p, m, ele, mat, rhs);
}

/* csr_sparsity_pos bisection (Sparse_Tools.F90:2411-2517); returns the
 * 0-based position or -1 */
static int64_t csr_pos(const int32_t *findrm, const int32_t *colm, int i1, int j1)
{
    int64_t base = findrm[i1 - 1] - 1; /* 0-based start of row */
    int64_t len = findrm[i1] - findrm[i1 - 1];
    if (len <= 0) return -1;
    int64_t lo = 0, hi = len - 1;
    int lo_j = colm[base + lo], hi_j = colm[base + hi];
    if (hi_j < j1) return -1;
    if (hi_j == j1) return base + hi;
    if (lo_j > j1) return -1;
    if (lo_j == j1) return base + lo;
    while (hi - lo > 1) {
        int64_t mid = (hi + lo) / 2;
        int mj = colm[base + mid];
        if (mj == j1) return base + mid;
        if (mj > j1) hi = mid;
        else lo = mid;
    }
    return -1;
}

int orc_assemble(const orc_params *p, const orc_mesh *m, double *A,
                 double *rhs, int nthreads, double *seconds)
{
    const int loc = m->loc;
    int64_t nnz = m->findrm[m->nrows] - 1;
    int err = 0;
    /* mesh movement: linear T, no stabilisation, advection-diffusion only */
    if (p->move_mesh && p->have_mass &&
        (loc != m->dim + 1 || p->stabilisation != STAB_NONE || p->equation_type != 0))
        return -3;
    /* zero(matrix); zero(rhs) (:560-561) */
    for (int64_t k = 0; k < nnz; k++) A[k] = 0.0;
    for (int i = 0; i < m->nrows; i++) rhs[i] = 0.0;
#ifdef _OPENMP
    if (nthreads <= 0) nthreads = omp_get_max_threads();
#endif
    double t0 = wall_now();
    /* colour loop (:574-598): colours in order, elements ascending */
#pragma omp parallel num_threads(nthreads > 0 ? nthreads : 1)
    {
        double mat[MAXLOC * MAXLOC], r[MAXLOC];
        for (int c = 0; c < m->ncolours; c++) {
#pragma omp for schedule(static)
            for (int k = m->col_ptr[c]; k < m->col_ptr[c + 1]; k++) {
                int ele = m->col_elems[k] - 1;
                const int32_t *en = m->ndglno + (int64_t)ele * loc;
                element_cg(p, m, ele, mat, r);
                /* addto(matrix, nodes, nodes, matrix_addto) -> csr_vaddto
                 * (:2680-2703): rows outer, columns inner, zeros skipped */
                for (int i = 0; i < loc; i++)
                    for (int j = 0; j < loc; j++) {
                        double v = mat[i + loc * j];
                        if (v == 0.0) continue;
                        int64_t pos = csr_pos(m->findrm, m->colm, en[i], en[j]);
                        if (pos < 0) {
#pragma omp atomic write
                            err = 1;
                            continue;
                        }
                        A[pos] = A[pos] + v;
                    }
                /* addto(rhs, nodes, rhs_addto) (Fields_Manipulation.F90:
                 * 255-273) */
                for (int i = 0; i < loc; i++) rhs[en[i] - 1] = rhs[en[i] - 1] + r[i];
            }
        }
    }
    if (seconds) *seconds = wall_now() - t0;
    return err ? -1 : 0;
}

/* ------------------------------------------------------------------ */
/* The pressure path (Momentum_CG.F90, Assemble_CMC.F90)               */
/* ------------------------------------------------------------------ */

#define PN_(i, g) (pp->p_n[(i) + pl * (g)])
#define PDN_(i, g, d) (pp->p_dn[(i) + pl * ((g) + ngi * (d))])

/* construct_momentum_element_cg's divergence block (:1377-1403) and
 * add_mass_element_cg's lumped mass (:1492-1600) for one element. */
int orc_ct_element(const orc_mesh *m, const orc_pressure *pp, int ele, double *mat, double *mass_lump)
{
    const int dim = m->dim, loc = m->loc, ngi = m->ngi, pl = pp->p_loc;
    const int32_t *en = m->ndglno + (int64_t)ele * loc;
    double JT[MAXDIM][MAXDIM], iJ[MAXDIM][MAXDIM], detJ;
    double du[MAXLOC][MAXGI][MAXDIM], dp[MAXLOC][MAXGI][MAXDIM], detwei[MAXGI];
    elem_geometry(m, ele, JT, iJ, &detJ);
    /* transform_to_physical(X, ele, u_shape, dshape=du_t, detwei=detwei)
     * and, by parts, (X, ele, p_shape, dshape=dp_t): dshape = dn . invJ^T,
     * once for a linear element, per point otherwise */
    const int u_nonlinear = (loc != dim + 1), p_nonlinear = (pl != dim + 1);
    for (int g = 0; g < ngi; g++) {
        const int gu = u_nonlinear ? g : 0, gp = p_nonlinear ? g : 0;
        for (int d = 0; d < dim; d++) {
            for (int n = 0; n < loc; n++) {
                double s = 0.0;
                for (int k = 0; k < dim; k++) s = s + DN_(n, gu, k) * iJ[d][k];
                du[n][g][d] = s;
            }
            for (int n = 0; n < pl; n++) {
                double s = 0.0;
                for (int k = 0; k < dim; k++) s = s + PDN_(n, gp, k) * iJ[d][k];
                dp[n][g][d] = s;
            }
        }
        detwei[g] = fabs(detJ) * m->w[g];
    }
    /* grad_p_u_mat(dim, p_loc, u_loc): shape_dshape(p_shape, du_t, detwei)
     * (FETools.F90:332-362: sum_g detwei_g (n_p(i,g) du_t(j,g,d))), or by
     * parts -dshape_shape(dp_t, u_shape, detwei) (:364-389) */
    for (int i = 0; i < pl; i++)
        for (int j = 0; j < loc; j++)
            for (int d = 0; d < dim; d++) {
                double s = 0.0;
                if (pp->integrate_by_parts) {
                    for (int g = 0; g < ngi; g++) s = s + detwei[g] * (dp[i][g][d] * N_(j, g));
                    s = -s;
                } else {
                    for (int g = 0; g < ngi; g++) s = s + detwei[g] * (PN_(i, g) * du[j][g][d]);
                }
                mat[d + dim * (i + pl * j)] = s;
            }
    /* mass_mat = shape_shape(u_shape, u_shape, density_gi*detwei)
     * (FETools.F90:206-226: dot_product(n(i,:)*n(j,:), coefficient));
     * mass_lump = sum(mass_mat, 2) */
    double cdw[MAXGI];
    for (int g = 0; g < ngi; g++) {
        double rho = 0.0;
        for (int i = 0; i < loc; i++) {
            const double v = pp->density_const ? pp->density[0] : pp->density[en[i] - 1];
            rho = rho + v * N_(i, g);
        }
        cdw[g] = rho * detwei[g];
    }
    for (int i = 0; i < loc; i++) {
        double lump = 0.0;
        for (int j = 0; j < loc; j++) {
            double mij = 0.0;
            for (int g = 0; g < ngi; g++) mij = mij + (N_(i, g) * N_(j, g)) * cdw[g];
            lump = lump + mij;
        }
        mass_lump[i] = lump;
    }
    return 0;
}

int orc_ct_assemble(const orc_mesh *m, const orc_pressure *pp, double *ct_val, double *masslump)
{
    const int dim = m->dim, loc = m->loc, pl = pp->p_loc;
    const int64_t nnz = pp->ct_findrm[pp->p_nodes] - 1;
    int err = 0;
    for (int64_t k = 0; k < (int64_t)dim * nnz; k++) ct_val[k] = 0.0;
    for (int i = 0; i < m->nnodes; i++) masslump[i] = 0.0;
    double mat[MAXDIM * MAXLOC * MAXLOC], ml[MAXLOC];
    int64_t pos[MAXLOC][MAXLOC];
    /* colour loop (Momentum_CG.F90:722-741): colours in order, elements
     * ascending (sequential here: elements of one colour share no node) */
    for (int c = 0; c < m->ncolours; c++)
        for (int k = m->col_ptr[c]; k < m->col_ptr[c + 1]; k++) {
            const int ele = m->col_elems[k] - 1;
            const int32_t *en = m->ndglno + (int64_t)ele * loc;
            const int32_t *pn = pp->p_ndglno + (int64_t)ele * pl;
            orc_ct_element(m, pp, ele, mat, ml);
            /* addto(ct_m, p_ele, u_ele, spread(grad_p_u_mat, 1, 1)) ->
             * block_csr_blocks_addto (Sparse_Tools.F90:2764-2812): a
             * position is looked up unless all blocks' values are zero, and
             * a zero value is skipped */
            for (int i = 0; i < pl; i++)
                for (int j = 0; j < loc; j++) {
                    int any = 0;
                    for (int d = 0; d < dim; d++) any |= mat[d + dim * (i + pl * j)] != 0.0;
                    pos[i][j] = any ? csr_pos(pp->ct_findrm, pp->ct_colm, pn[i], en[j]) : -2;
                    if (any && pos[i][j] < 0) err = 1;
                }
            for (int d = 0; d < dim; d++)
                for (int i = 0; i < pl; i++)
                    for (int j = 0; j < loc; j++) {
                        const double v = mat[d + dim * (i + pl * j)];
                        if (v == 0.0 || pos[i][j] < 0) continue;
                        ct_val[d * nnz + pos[i][j]] = ct_val[d * nnz + pos[i][j]] + v;
                    }
            /* addto(masslump, dim, u_ele, mass_lump) (:1560-1564) */
            for (int i = 0; i < loc; i++) masslump[en[i] - 1] = masslump[en[i] - 1] + ml[i];
        }
    return err ? -1 : 0;
}

void orc_cmc(int rows, const int32_t *ctp_findrm, const int32_t *ctp_colm, const double *ctp_val, int64_t ctp_nnz,
             const int32_t *ct_findrm, const int32_t *ct_colm, const double *ct_val, int64_t ct_nnz, int dim,
             const double *inv_ml, const int32_t *cmc_findrm, const int32_t *cmc_colm, double *cmc_val)
{
    /* mult_div_vector_div_T (Sparse_Matrices_Fields.F90:590-671):
     * M_ij = sum_k A_ik D_k B_jk, rows of A and B walked together */
    for (int i = 0; i < rows; i++)
        for (int e = cmc_findrm[i] - 1; e < cmc_findrm[i + 1] - 1; e++) {
            const int j = cmc_colm[e] - 1;
            int64_t k1 = ctp_findrm[i] - 1, k2 = ct_findrm[j] - 1;
            const int64_t e1 = ctp_findrm[i + 1] - 1, e2 = ct_findrm[j + 1] - 1;
            double entry0 = 0.0;
            while (k1 < e1 && k2 < e2) {
                if (ctp_colm[k1] < ct_colm[k2]) {
                    k1++;
                } else if (ctp_colm[k1] == ct_colm[k2]) {
                    const int64_t node = ctp_colm[k1] - 1;
                    for (int d = 0; d < dim; d++)
                        entry0 = entry0 + ctp_val[d * ctp_nnz + k1] * ct_val[d * ct_nnz + k2] * inv_ml[d + dim * node];
                    k1++;
                    k2++;
                } else {
                    k2++;
                }
            }
            cmc_val[e] = entry0;
        }
}

/* ------------------------------------------------------------------ */
/* Surface integrals (Advection_Diffusion_CG.F90:606-638, :1228-1379)   */
/* ------------------------------------------------------------------ */

enum { BC_NEUMANN = 1, BC_WEAKDIRICHLET = 2, BC_INTERNAL = 3, BC_ROBIN = 4 }; /* :74-75 */

#define FN_(i, g) (s->fn[(i) + fl * (g)])
#define FXDN_(i, g, d) (s->fx_dn[(i) + flx * ((g) + fg * (d))])

/* assemble_advection_diffusion_face_cg (:1228-1290) for one face:
 * mat(floc,floc) column-major, rhs(floc).  Returns 0, or -2 for a weak
 * Dirichlet face with diffusivity (the FLExit of :1374-1377). */
static int face_cg(const orc_params *p, const orc_mesh *m, const orc_surface *s, int face,
                   double *mat, double *rhs)
{
    const int dim = m->dim, fl = s->floc, fg = s->fngi, flx = s->floc_x, lx = m->loc_x;
    const int bc = s->bc_type[face];
    const int32_t *fnod = s->face_nodes + (int64_t)face * fl;
    const int32_t *fxn = s->face_nodes_x + (int64_t)face * flx;
    const double *bv = s->bc_value ? s->bc_value + (int64_t)face * fl : NULL;
    const double *bv2 = s->bc_value2 ? s->bc_value2 + (int64_t)face * fl : NULL;
    const double eps = DBL_EPSILON;
    const double dt_theta = p->dt * p->theta;
    const int adv_face = p->have_advection && p->integrate_by_parts;
    double detwei[MAXGI], nrm[MAXDIM] = {0.0, 0.0, 0.0}, Tf[MAXLOC];

    for (int i = 0; i < fl * fl; i++) mat[i] = 0.0;
    for (int i = 0; i < fl; i++) rhs[i] = 0.0;
    for (int g = 0; g < fg; g++) detwei[g] = 0.0;

    /* Step 1: transform_facet_to_physical (Transform_elements.F90:1353-1525
     * with the normal, :1557-1640 without), linear simplex coordinates:
     * one Jacobian per face */
    if (adv_face || (p->have_diffusivity && (bc == BC_NEUMANN || bc == BC_ROBIN))) {
        double Xf[MAXDIM][MAXLOC], J[MAXDIM][2], detJ = 0.0;
        for (int k = 0; k < flx; k++)
            for (int a = 0; a < dim; a++) Xf[a][k] = m->X[a + dim * (int64_t)(fxn[k] - 1)];
        for (int a = 0; a < dim; a++)
            for (int d = 0; d < dim - 1; d++) {
                double t = 0.0;
                for (int k = 0; k < flx; k++) t = t + Xf[a][k] * FXDN_(k, 0, d);
                J[a][d] = t;
            }
        if (dim == 2) {
            detJ = sqrt(J[0][0] * J[0][0] + J[1][0] * J[1][0]);
        } else if (dim == 3) {
            for (int i = 1; i <= 3; i++) { /* cyc3(i) = mod(i-1,3)+1 */
                int a2 = (i + 1) % 3, a1 = i % 3;
                double t = J[a2][0] * J[a1][1] - J[a2][1] * J[a1][0];
                detJ = detJ + t * t;
            }
            detJ = sqrt(detJ);
        } else {
            detJ = 1.0;
        }
        for (int g = 0; g < fg; g++) detwei[g] = detJ * s->fw[g];
        if (adv_face) {
            /* outward vector: face centroid minus element centroid (:1436-1437) */
            const int32_t *exn = m->ndglno_x + (int64_t)(s->face_ele[face] - 1) * lx;
            double o[MAXDIM], dp = 0.0, len = 0.0;
            for (int a = 0; a < dim; a++) {
                double sf = 0.0, se = 0.0;
                for (int k = 0; k < flx; k++) sf = sf + Xf[a][k];
                for (int k = 0; k < lx; k++) se = se + m->X[a + dim * (int64_t)(exn[k] - 1)];
                o[a] = sf / flx - se / lx;
            }
            /* facet_normal (:1527-1555); cross_product (Vector_Tools.F90:76-85) */
            if (dim == 2) {
                nrm[0] = -J[1][0];
                nrm[1] = J[0][0];
            } else if (dim == 3) {
                nrm[0] = J[1][0] * J[2][1] - J[2][0] * J[1][1];
                nrm[1] = J[2][0] * J[0][1] - J[0][0] * J[2][1];
                nrm[2] = J[0][0] * J[1][1] - J[1][0] * J[0][1];
            } else {
                nrm[0] = 1.0;
            }
            for (int a = 0; a < dim; a++) dp = dp + nrm[a] * o[a];
            for (int a = 0; a < dim; a++) nrm[a] = nrm[a] * dp;
            for (int a = 0; a < dim; a++) len = len + nrm[a] * nrm[a];
            len = sqrt(len);
            for (int a = 0; a < dim; a++) nrm[a] = nrm[a] / len;
        }
    }
    for (int i = 0; i < fl; i++) Tf[i] = m->T[fnod[i] - 1];

    /* Step 2a: add_advection_face_cg (:1292-1347) */
    if (adv_face) {
        double cq[MAXGI], adv[MAXLOC][MAXLOC];
        /* internal energy: times the theta-weighted density [and nvfrac]
         * (:1318-1327); mesh movement: the velocity relative to the grid
         * (:1315-1317; linear T, the grid velocity's face is T's) */
        const int ie = p->equation_type == EQ_INTERNAL_ENERGY;
        const int mv = p->move_mesh && p->have_mass;
        if (mv && fl != flx) return -3;
        for (int g = 0; g < fg; g++) {
            double un = 0.0;
            for (int d = 0; d < dim; d++) {
                double uq = 0.0; /* face_val_at_quad (Fields_Base.F90:2392-2400) */
                for (int k = 0; k < fl; k++) {
                    double uk = m->u_const ? m->U[d] : m->U[d + dim * (int64_t)(fnod[k] - 1)];
                    uq = uq + uk * FN_(k, g);
                }
                if (mv) {
                    double ugq = 0.0;
                    for (int k = 0; k < flx; k++) ugq = ugq + m->ug[d + dim * (int64_t)(fxn[k] - 1)] * FN_(k, g);
                    uq = uq - ugq;
                }
                un = un + uq * nrm[d];
            }
            cq[g] = detwei[g] * un;
            if (ie) {
                double r = 0.0, ro = 0.0, nv = 0.0;
                for (int k = 0; k < fl; k++) {
                    r = r + (m->rho_const ? m->rho[0] : m->rho[fnod[k] - 1]) * FN_(k, g);
                    ro = ro + (m->rho_old_const ? m->rho_old[0] : m->rho_old[fnod[k] - 1]) * FN_(k, g);
                    if (p->multiphase) nv = nv + (m->nvfrac_const ? m->nvfrac[0] : m->nvfrac[fnod[k] - 1]) * FN_(k, g);
                }
                cq[g] = cq[g] * (p->density_theta * r + (1.0 - p->density_theta) * ro);
                if (p->multiphase) cq[g] = cq[g] * nv;
            }
        }
        for (int i = 0; i < fl; i++)
            for (int j = 0; j < fl; j++) {
                double t = 0.0;
                for (int g = 0; g < fg; g++) t = t + (FN_(i, g) * FN_(j, g)) * cq[g];
                adv[i][j] = t;
            }
        if (fabs(dt_theta) > eps) {
            if (bc == BC_WEAKDIRICHLET) {
                for (int i = 0; i < fl; i++) {
                    double t = 0.0;
                    for (int j = 0; j < fl; j++) t = t + adv[i][j] * (bv[j] - Tf[j]);
                    rhs[i] = rhs[i] - p->theta * t;
                }
            } else {
                for (int i = 0; i < fl; i++)
                    for (int j = 0; j < fl; j++)
                        mat[i + fl * j] = mat[i + fl * j] + dt_theta * adv[i][j];
            }
        }
        for (int i = 0; i < fl; i++) {
            double t = 0.0;
            for (int j = 0; j < fl; j++) t = t + adv[i][j] * Tf[j];
            rhs[i] = rhs[i] - t;
        }
    }

    /* Step 2b: add_diffusivity_face_cg (:1349-1379) */
    if (p->have_diffusivity) {
        if (bc == BC_NEUMANN || bc == BC_ROBIN) {
            double wq[MAXGI];
            for (int g = 0; g < fg; g++) {
                double t = 0.0;
                for (int k = 0; k < fl; k++) t = t + bv[k] * FN_(k, g);
                wq[g] = detwei[g] * t;
            }
            for (int i = 0; i < fl; i++) { /* shape_rhs (FETools.F90:38-54) */
                double t = 0.0;
                for (int g = 0; g < fg; g++) t = t + FN_(i, g) * wq[g];
                rhs[i] = rhs[i] + t;
            }
        }
        if (bc == BC_ROBIN) {
            double wq[MAXGI], rob[MAXLOC][MAXLOC];
            for (int g = 0; g < fg; g++) {
                double t = 0.0;
                for (int k = 0; k < fl; k++) t = t + bv2[k] * FN_(k, g);
                wq[g] = detwei[g] * t;
            }
            for (int i = 0; i < fl; i++)
                for (int j = 0; j < fl; j++) {
                    double t = 0.0;
                    for (int g = 0; g < fg; g++) t = t + (FN_(i, g) * FN_(j, g)) * wq[g];
                    rob[i][j] = t;
                }
            if (fabs(dt_theta) > eps)
                for (int i = 0; i < fl; i++)
                    for (int j = 0; j < fl; j++)
                        mat[i + fl * j] = mat[i + fl * j] + dt_theta * rob[i][j];
            for (int i = 0; i < fl; i++) {
                double t = 0.0;
                for (int j = 0; j < fl; j++) t = t + rob[i][j] * Tf[j];
                rhs[i] = rhs[i] - t;
            }
        } else if (bc == BC_WEAKDIRICHLET) {
            return -2;
        }
    }
    return 0;
}

int orc_face(const orc_params *p, const orc_mesh *m, const orc_surface *s, int face,
             double *mat, double *rhs)
{
    return face_cg(p, m, s, face, mat, rhs);
}

int orc_surface_assemble(const orc_params *p, const orc_mesh *m, const orc_surface *s,
                         double *A, double *rhs)
{
    const int fl = s->floc;
    double mat[MAXLOC * MAXLOC], r[MAXLOC];
    /* the face loop runs only with advection integrated by parts or with
     * diffusivity (:608-611); internal faces are skipped (:627) */
    if (!((p->integrate_by_parts && p->have_advection) || p->have_diffusivity)) return 0;
    for (int f = 0; f < s->nfaces; f++) {
        if (s->bc_type[f] == BC_INTERNAL) continue;
        int rc = face_cg(p, m, s, f, mat, r);
        if (rc) return rc;
        const int32_t *fnod = s->face_nodes + (int64_t)f * fl;
        for (int i = 0; i < fl; i++)
            for (int j = 0; j < fl; j++) {
                double v = mat[i + fl * j];
                if (v == 0.0) continue; /* csr_addto skips zeros */
                int64_t pos = csr_pos(m->findrm, m->colm, fnod[i], fnod[j]);
                if (pos < 0) return -1;
                A[pos] = A[pos] + v;
            }
        for (int i = 0; i < fl; i++) rhs[fnod[i] - 1] = rhs[fnod[i] - 1] + r[i];
    }
    return 0;
}

/* ------------------------------------------------------------------ */
/* Dirichlet, SpMV                                                      */
/* ------------------------------------------------------------------ */

void orc_dirichlet(int nbc, const int32_t *nodes, const double *values,
                   const double *T, double dt, int32_t *inactive, double *rhs)
{
    for (int j = 0; j < nbc; j++) {
        int nd = nodes[j] - 1;
        inactive[nd] = 1;
        if (dt > 0.0)
            rhs[nd] = (values[j] - T[nd]) / dt;
        else
            rhs[nd] = values[j];
    }
}

void orc_csr_mult(int nrows, const int32_t *findrm, const int32_t *colm,
                  const double *val, const double *x, double *y)
{
#pragma omp parallel for schedule(static)
    for (int i = 0; i < nrows; i++) {
        double s = 0.0;
        for (int64_t k = findrm[i] - 1; k < findrm[i + 1] - 1; k++)
            s = s + val[k] * x[colm[k] - 1];
        y[i] = s;
    }
}

double orc_ghost_pivot(int n, const int32_t *centrm, const double *A)
{
    double mindiag = DBL_MAX, maxdiag = -DBL_MAX;
    for (int i = 0; i < n; i++) {
        double d = centrm[i] > 0 ? fabs(A[centrm[i] - 1]) : 0.0;
        if (d < mindiag) mindiag = d;
        if (d > maxdiag) maxdiag = d;
    }
    return sqrt((maxdiag + mindiag) * maxdiag / 2.0);
}

/* ------------------------------------------------------------------ */
/* Krylov (PETSc KSPCG / KSPGMRES semantics on the reduced system)      */
/* ------------------------------------------------------------------ */

typedef struct {
    int n;
    const int32_t *findrm, *colm;
    double *val; /* reduced values (inactive columns zeroed, ghost pivot) */
} rsys;

static void rmult(const rsys *s, const double *x, double *y)
{
#pragma omp parallel for schedule(static)
    for (int i = 0; i < s->n; i++) {
        double t = 0.0;
        for (int64_t k = s->findrm[i] - 1; k < s->findrm[i + 1] - 1; k++)
            t = t + s->val[k] * x[s->colm[k] - 1];
        y[i] = t;
    }
}

static double dot(int n, const double *a, const double *b)
{
    double s = 0.0;
#pragma omp parallel for reduction(+ : s) schedule(static)
    for (int i = 0; i < n; i++) s += a[i] * b[i];
    return s;
}

enum {
    KSP_CONVERGED_RTOL = 2,
    KSP_CONVERGED_ATOL = 3,
    KSP_DIVERGED_ITS = -3,
    KSP_DIVERGED_DTOL = -4,
    KSP_DIVERGED_INDEFINITE_PC = -8,
    KSP_DIVERGED_NANORINF = -9,
    KSP_DIVERGED_INDEFINITE_MAT = -10
};

/* ---- preconditioners of the reduced system -------------------------------
 * Jacobi (PCJACOBI) and "sor" (femtools/Solvers.F90:1844-1854 selects
 * PCSOR: one local symmetric Gauss-Seidel sweep, omega = 1).  PETSc sweeps
 * in the natural row order; this library (and so this checker) sweeps
 * colour by colour over a first-fit colouring of the matrix graph in row
 * order, ascending rows inside a colour: SSOR of the symmetrically permuted
 * matrix.  Each row is summed in entry order with fused multiply-adds. */

/* First-fit colouring of the rows of a 1-based CSR pattern in row order;
 * rows listed colour by colour (0-based), ptr[ncol+1].  Returns ncol. */
int orc_colour_rows(int n, const int32_t *findrm, const int32_t *colm, int32_t *ptr_out,
                    int32_t *rows_out)
{
    int32_t *colour = malloc(sizeof(int32_t) * (size_t)n);
    int32_t *mark = malloc(sizeof(int32_t) * (size_t)(n + 1));
    int ncol = 0;
    for (int q = 0; q <= n; q++) mark[q] = -1;
    for (int i = 0; i < n; i++) {
        for (int64_t k = findrm[i] - 1; k < findrm[i + 1] - 1; k++) {
            int j = colm[k] - 1;
            if (j >= 0 && j < i) mark[colour[j]] = i;
        }
        int q = 0;
        while (mark[q] == i) q++;
        colour[i] = q;
        if (q + 1 > ncol) ncol = q + 1;
    }
    if (ptr_out) {
        for (int q = 0; q <= ncol; q++) ptr_out[q] = 0;
        for (int i = 0; i < n; i++) ptr_out[colour[i] + 1]++;
        for (int q = 0; q < ncol; q++) ptr_out[q + 1] += ptr_out[q];
        int32_t *pos = malloc(sizeof(int32_t) * (size_t)(ncol + 1));
        for (int q = 0; q < ncol; q++) pos[q] = ptr_out[q];
        for (int i = 0; i < n; i++) rows_out[pos[colour[i]]++] = i;
        free(pos);
    }
    free(colour);
    free(mark);
    return ncol;
}

typedef struct {
    int kind;              /* 0 none, 1 jacobi, 3 sor */
    const double *dinv;
    int ncol;
    const int32_t *ptr, *rows;
} pcond;

static void sor_row(const rsys *s, int i, const double *r, double *z)
{
    double acc = r[i], aii = 0.0;
    for (int64_t k = s->findrm[i] - 1; k < s->findrm[i + 1] - 1; k++) {
        int j = s->colm[k] - 1;
        if (j == i) aii = s->val[k];
        else acc = fma(-s->val[k], z[j], acc);
    }
    z[i] = acc / aii;
}

static void ssor(const rsys *s, const pcond *pc, const double *r, double *z)
{
    for (int i = 0; i < s->n; i++) z[i] = 0.0;
    for (int q = 0; q < pc->ncol; q++)
        for (int t = pc->ptr[q]; t < pc->ptr[q + 1]; t++) sor_row(s, pc->rows[t], r, z);
    for (int q = pc->ncol - 2; q >= 0; q--)
        for (int t = pc->ptr[q]; t < pc->ptr[q + 1]; t++) sor_row(s, pc->rows[t], r, z);
}

static void pc_apply(const rsys *s, const pcond *pc, const double *r, double *z)
{
    if (pc->kind == 3) {
        ssor(s, pc, r, z);
        return;
    }
    for (int i = 0; i < s->n; i++) z[i] = pc->kind == 1 ? pc->dinv[i] * r[i] : r[i];
}

/* z = M^-1 r of the multicolour SSOR on a 1-based CSR matrix (test hook). */
void orc_ssor_apply(int n, const int32_t *findrm, const int32_t *colm, const double *val,
                    const double *r, double *z)
{
    int32_t *ptr = malloc(sizeof(int32_t) * (size_t)(n + 1)), *rows = malloc(sizeof(int32_t) * (size_t)n);
    pcond pc = {3, NULL, 0, ptr, rows};
    pc.ncol = orc_colour_rows(n, findrm, colm, ptr, rows);
    rsys s = {n, findrm, colm, (double *)val};
    ssor(&s, &pc, r, z);
    free(ptr);
    free(rows);
}

static int conv_test(double rnorm, double rnorm0, double rtol, double atol)
{
    if (isnan(rnorm) || isinf(rnorm)) return KSP_DIVERGED_NANORINF;
    if (rnorm <= atol) return KSP_CONVERGED_ATOL;
    if (rnorm <= rtol * rnorm0) return KSP_CONVERGED_RTOL;
    if (rnorm >= 1e5 * rnorm0 && rnorm0 > 0) return KSP_DIVERGED_DTOL;
    return 0;
}

static int ksp_cg(const rsys *s, const pcond *pc, const double *b, double *x,
                  double rtol, double atol, int maxits, int *its)
{
    int n = s->n;
    double *r = malloc(sizeof(double) * n), *z = malloc(sizeof(double) * n);
    double *p = malloc(sizeof(double) * n), *w = malloc(sizeof(double) * n);
    int reason = 0;
    rmult(s, x, w);
    for (int i = 0; i < n; i++) r[i] = b[i] - w[i];
    pc_apply(s, pc, r, z);
    double dp = sqrt(dot(n, z, z)); /* KSP_NORM_PRECONDITIONED */
    double rnorm0 = dp;
    double beta = dot(n, z, r), betaold = 1.0;
    *its = 0;
    reason = conv_test(dp, rnorm0, rtol, atol);
    if (dp == 0.0) reason = KSP_CONVERGED_ATOL;
    for (int it = 0; it < maxits && !reason; it++) {
        if (it == 0) {
            for (int i = 0; i < n; i++) p[i] = z[i];
        } else {
            double bb = beta / betaold;
            for (int i = 0; i < n; i++) p[i] = z[i] + bb * p[i];
        }
        rmult(s, p, w);
        double dpi = dot(n, p, w);
        if (dpi <= 0.0) {
            reason = (dpi == 0.0) ? KSP_DIVERGED_INDEFINITE_MAT : KSP_DIVERGED_INDEFINITE_MAT;
            break;
        }
        double a = beta / dpi;
#pragma omp parallel for schedule(static)
        for (int i = 0; i < n; i++) {
            x[i] = x[i] + a * p[i];
            r[i] = r[i] - a * w[i];
        }
        pc_apply(s, pc, r, z);
        dp = sqrt(dot(n, z, z));
        *its = it + 1;
        betaold = beta;
        beta = dot(n, z, r);
        reason = conv_test(dp, rnorm0, rtol, atol);
        if (!reason && beta < 0.0) reason = KSP_DIVERGED_INDEFINITE_PC;
    }
    if (!reason) reason = KSP_DIVERGED_ITS;
    free(r); free(z); free(p); free(w);
    return reason;
}

/* restarted GMRES, left preconditioning, modified Gram-Schmidt */
static int ksp_gmres(const rsys *s, const pcond *pc, const double *b,
                     double *x, double rtol, double atol, int maxits,
                     int restart, int *its)
{
    int n = s->n, m = restart > 0 ? restart : 30;
    double *V = malloc(sizeof(double) * (size_t)n * (m + 1));
    double *H = calloc((size_t)(m + 1) * m, sizeof(double));
    double *cs = malloc(sizeof(double) * m), *sn = malloc(sizeof(double) * m);
    double *g = malloc(sizeof(double) * (m + 1)), *y = malloc(sizeof(double) * m);
    double *w = malloc(sizeof(double) * n), *t = malloc(sizeof(double) * n);
    int reason = 0, total = 0;
    double rnorm0 = -1.0;
    *its = 0;
    while (!reason) {
        rmult(s, x, w);
        for (int i = 0; i < n; i++) w[i] = b[i] - w[i];
        pc_apply(s, pc, w, V);
        double beta = sqrt(dot(n, V, V));
        if (rnorm0 < 0) {
            rnorm0 = beta;
            reason = conv_test(beta, rnorm0, rtol, atol);
            if (beta == 0.0) reason = KSP_CONVERGED_ATOL;
            if (reason) break;
        }
        for (int i = 0; i < n; i++) V[i] /= beta;
        memset(g, 0, sizeof(double) * (m + 1));
        g[0] = beta;
        int k;
        for (k = 0; k < m && !reason; k++) {
            double *vk = V + (size_t)n * k, *vn = V + (size_t)n * (k + 1);
            rmult(s, vk, t);
            pc_apply(s, pc, t, vn);
            for (int j = 0; j <= k; j++) {
                double *vj = V + (size_t)n * j;
                double h = dot(n, vn, vj);
                H[j + (m + 1) * k] = h;
                for (int i = 0; i < n; i++) vn[i] -= h * vj[i];
            }
            double hn = sqrt(dot(n, vn, vn));
            H[k + 1 + (m + 1) * k] = hn;
            if (hn != 0.0)
                for (int i = 0; i < n; i++) vn[i] /= hn;
            for (int j = 0; j < k; j++) {
                double a = H[j + (m + 1) * k], bb = H[j + 1 + (m + 1) * k];
                H[j + (m + 1) * k] = cs[j] * a + sn[j] * bb;
                H[j + 1 + (m + 1) * k] = -sn[j] * a + cs[j] * bb;
            }
            double a = H[k + (m + 1) * k], bb = H[k + 1 + (m + 1) * k];
            double den = sqrt(a * a + bb * bb);
            cs[k] = den == 0.0 ? 1.0 : a / den;
            sn[k] = den == 0.0 ? 0.0 : bb / den;
            H[k + (m + 1) * k] = cs[k] * a + sn[k] * bb;
            H[k + 1 + (m + 1) * k] = 0.0;
            g[k + 1] = -sn[k] * g[k];
            g[k] = cs[k] * g[k];
            total++;
            *its = total;
            reason = conv_test(fabs(g[k + 1]), rnorm0, rtol, atol);
            if (!reason && total >= maxits) reason = KSP_DIVERGED_ITS;
        }
        /* solve H y = g and update x */
        for (int i = k - 1; i >= 0; i--) {
            double acc = g[i];
            for (int j = i + 1; j < k; j++) acc -= H[i + (m + 1) * j] * y[j];
            y[i] = acc / H[i + (m + 1) * i];
        }
        for (int j = 0; j < k; j++) {
            double *vj = V + (size_t)n * j;
            for (int i = 0; i < n; i++) x[i] += y[j] * vj[i];
        }
    }
    free(V); free(H); free(cs); free(sn); free(g); free(y); free(w); free(t);
    return reason;
}

int orc_petsc_solve(const orc_solver *sv, int n, const int32_t *findrm,
                    const int32_t *colm, const int32_t *centrm,
                    const double *A, const int32_t *inactive,
                    const double *rhs, double *x, int32_t *its_out,
                    int32_t *reason_out, double *true_relres, double *seconds)
{
#ifdef _OPENMP
    if (sv->nthreads > 0) omp_set_num_threads(sv->nthreads);
#endif
    int64_t nnz = findrm[n] - 1;
    rsys s = {n, findrm, colm, malloc(sizeof(double) * (size_t)nnz)};
    double *b = malloc(sizeof(double) * n), *ghost = malloc(sizeof(double) * n);
    double *tmp = malloc(sizeof(double) * n), *dinv = NULL;
    double *y = malloc(sizeof(double) * n);
    int any_inactive = 0;
    if (inactive)
        for (int i = 0; i < n; i++) any_inactive |= inactive[i];
    /* lifting (Solvers.F90:1072-1099): tmp = A * ghost_rhs */
    for (int i = 0; i < n; i++) ghost[i] = (any_inactive && inactive[i]) ? rhs[i] : 0.0;
    orc_csr_mult(n, findrm, colm, A, ghost, tmp);
    double pivot = orc_ghost_pivot(n, centrm, A);
    for (int i = 0; i < n; i++) {
        int ghost_row = any_inactive && inactive[i];
        b[i] = ghost_row ? 0.0 : rhs[i] - tmp[i];
        y[i] = ghost_row ? 0.0 : x[i];
        for (int64_t k = findrm[i] - 1; k < findrm[i + 1] - 1; k++) {
            int c = colm[k] - 1;
            double v = A[k];
            if (ghost_row) v = (c == i) ? pivot : 0.0;
            else if (any_inactive && inactive[c]) v = 0.0;
            s.val[k] = v;
        }
    }
    pcond pc = {sv->pc, NULL, 0, NULL, NULL};
    int32_t *mc_ptr = NULL, *mc_rows = NULL;
    if (sv->pc == 1) {
        dinv = malloc(sizeof(double) * n);
        for (int i = 0; i < n; i++) {
            double d = centrm[i] > 0 ? s.val[centrm[i] - 1] : 0.0;
            dinv[i] = d != 0.0 ? 1.0 / d : 1.0;
        }
        pc.dinv = dinv;
    } else if (sv->pc == 3) {
        mc_ptr = malloc(sizeof(int32_t) * (size_t)(n + 1));
        mc_rows = malloc(sizeof(int32_t) * (size_t)n);
        pc.ncol = orc_colour_rows(n, findrm, colm, mc_ptr, mc_rows);
        pc.ptr = mc_ptr;
        pc.rows = mc_rows;
    }
    double t0 = wall_now();
    int its = 0, reason;
    if (sv->ksp == 1)
        reason = ksp_gmres(&s, &pc, b, y, sv->rtol, sv->atol, sv->max_its, sv->restart, &its);
    else
        reason = ksp_cg(&s, &pc, b, y, sv->rtol, sv->atol, sv->max_its, &its);
    free(mc_ptr);
    free(mc_rows);
    if (seconds) *seconds = wall_now() - t0;
    /* true residual of the reduced system */
    rmult(&s, y, tmp);
    double rr = 0.0, bb = 0.0;
    for (int i = 0; i < n; i++) {
        double d = b[i] - tmp[i];
        rr += d * d;
        bb += b[i] * b[i];
    }
    *true_relres = bb > 0 ? sqrt(rr / bb) : sqrt(rr);
    /* petsc2field with ghost copy-back (Petsc_Tools.F90:724-731) */
    for (int i = 0; i < n; i++) x[i] = (any_inactive && inactive[i]) ? rhs[i] : y[i];
    *its_out = its;
    *reason_out = reason;
    free(s.val); free(b); free(ghost); free(tmp); free(dinv); free(y);
    return 0;
}
