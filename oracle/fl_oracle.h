/*
 * fl_oracle.h -- TEST INFRASTRUCTURE ONLY.
 *
 * CPU restatement ("oracle") of Fluidity's continuous-Galerkin
 * advection-diffusion assembly, CSR insertion, strong-Dirichlet elimination
 * and the Krylov hand-off.  Only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py may load it, and only as the checker or the
 * timed CPU baseline.  The product library (fluidity_amd) never links it.
 *
 * Conventions follow Fluidity's Fortran data layout: 1-based int32 node
 * numbers, element-major ndglno, column-major dim x nodes vector values,
 * column-major n(loc,ngi) and dn(loc,ngi,dim) shape tables.
 */
#ifndef FL_ORACLE_H
#define FL_ORACLE_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Quadrature.F90:164-240 (template selection), :572-605 (expansion),
 * :649-735 (tet rules), :909-993 (tri rules), :1140-1230 (interval
 * rules).  l is ngi x (dim+1), row per point; returns ngi or -1. */
int orc_quadrature(int dim, int degree, double *l, double *w);

/* Shape_Functions.F90:128-329 + Elements.F90:615-663: Lagrange simplex
 * tables.  n(loc,ngi), dn(loc,ngi,dim) column-major.  Returns loc or -1. */
int orc_shape_tables(int dim, int degree, int ngi, const double *l,
                     double *n, double *dn);

/* Sparsity_Patterns.F90:47-85, :299-428.  All arrays 1-based.
 * findrm[nnodes+1], centrm[nnodes]; colm capacity cap.  Returns nnz or -1. */
int64_t orc_make_sparsity(int nnodes, int nele, int loc, const int32_t *ndglno,
                          int32_t *findrm, int32_t *colm, int32_t *centrm,
                          int64_t cap);

/* Colouring.F90:85-199: greedy first-fit colouring of the element graph
 * (elements adjacent when they share a node), elements visited in index
 * order.  colour[e] is 1-based.  Returns the number of colours. */
int orc_colour_elements(int nnodes, int nele, int loc, const int32_t *ndglno,
                        int32_t *colour);

typedef struct {
    double dt, theta, beta;
    int32_t have_mass, lump_mass, have_advection, integrate_by_parts;
    int32_t have_diffusivity, isotropic_diffusivity, have_source,
        have_absorption;
    /* stabilisation (Advection_Diffusion_CG.F90:71, :440-465): 0 none,
     * 1 streamline upwind, 2 SUPG; nu_bar scheme (Upwind_Stabilisation.F90:
     * 18-19): 1 optimal, 2 doubly asymptotic, 3 critical rule, 4 unity */
    int32_t stabilisation, nu_bar_scheme;
    double nu_bar_scale;
    /* equation type (Advection_Diffusion_CG.F90:484-552): 0 advection-
     * diffusion, 1 internal energy, 2 k-epsilon; density_theta (internal
     * energy); compressible: the p div(u) term (internal energy);
     * multiphase: the PhaseVolumeFraction factor (internal energy) */
    double density_theta;
    int32_t equation_type, compressible, multiphase;
    /* mesh movement (:421-437, move_mesh with mass): the old and new
     * positions and the grid velocity of orc_mesh */
    int32_t move_mesh;
} orc_params;

typedef struct {
    int32_t dim, nele;
    /* coordinate mesh (linear simplices) */
    int32_t loc_x, nx;
    const int32_t *ndglno_x;   /* loc_x*nele, 1-based */
    const double *X;           /* dim x nx */
    const double *x_dn;        /* loc_x x ngi x dim */
    /* field mesh of T (and of every other field) */
    int32_t loc, ngi, nnodes;
    const int32_t *ndglno;     /* loc*nele, 1-based */
    const double *n, *dn, *w;  /* n(loc,ngi) dn(loc,ngi,dim) w(ngi) */
    /* fields; *_const != 0 means FIELD_TYPE_CONSTANT (one value) */
    const double *T;
    const double *U;      int32_t u_const;      /* dim x N */
    const double *kappa;  int32_t kappa_const;  /* dim x dim x N */
    const double *src;    int32_t src_const;
    const double *absorp; int32_t absorp_const;
    /* density, old density, pressure and phase volume fraction (scalars on
     * the field mesh), read by the internal-energy and k-epsilon equations */
    const double *rho;    int32_t rho_const;
    const double *rho_old; int32_t rho_old_const;
    const double *pres;   int32_t pres_const;
    const double *nvfrac; int32_t nvfrac_const;
    /* mesh movement: OldCoordinate and IteratedCoordinate (dim x nx, on the
     * coordinate mesh) and GridVelocity (dim x nx, on the coordinate mesh) */
    const double *X_old, *X_new, *ug;
    /* colour sets: ncolours+1 offsets into elems (0-based offsets,
     * 1-based element ids, ascending inside each colour) */
    int32_t ncolours;
    const int32_t *col_ptr, *col_elems;
    /* sparsity, 1-based */
    int32_t nrows;
    const int32_t *findrm, *colm;
} orc_mesh;

/* Advection_Diffusion_CG.F90:558-600 (zero + colour loop), :702-1202
 * (element), Sparse_Tools.F90:2411-2703 (bisection addto skipping zeros).
 * Returns 0, or -1 if an entry falls outside the pattern.  *seconds gets
 * the wall time of the colour loop (profiler key advection_diffusion_loop).
 * nthreads <= 0 uses the OpenMP default. */
int orc_assemble(const orc_params *p, const orc_mesh *m, double *A,
                 double *rhs, int nthreads, double *seconds);

/* nu_bar_scaled_q of assemble/Upwind_Stabilisation.F90:243-319 at one
 * quadrature point: u(dim); J(dim,dim) row-major with J[c][a] = dx_a/dxi_c
 * (transform_to_physical's j); K(dim,dim) row-major, used when have_k. */
double orc_nu_bar_scaled(int dim, const double *u, const double *J, const double *K, int have_k,
                         int scheme, double scale);

/* Same element routine, for one element: mat(loc,loc) column-major and
 * rhs(loc).  Used by the unit tests of the element arithmetic. */
int orc_element(const orc_params *p, const orc_mesh *m, int ele,
                double *mat, double *rhs);

/* Surface elements of T with their boundary conditions, as
 * assemble_advection_diffusion_cg passes them to the face loop
 * (Advection_Diffusion_CG.F90:606-638): face_global_nodes of T and of the
 * positions, face_ele, face_shape(t) tables on the surface quadrature, the
 * bc type from get_entire_boundary_condition (0, 1 neumann,
 * 2 weakdirichlet, 3 internal, 4 robin; :74-75) and the bc fields'
 * values at each face's nodes (ele_val(t_bc, face), ele_val(t_bc_2, face)). */
typedef struct {
    int32_t nfaces, floc, fngi, floc_x;
    const int32_t *face_nodes;    /* floc*nfaces, 1-based T nodes */
    const int32_t *face_nodes_x;  /* floc_x*nfaces, 1-based coordinate nodes */
    const int32_t *face_ele;      /* nfaces, 1-based */
    const double *fn, *fw;        /* fn(floc,fngi), fw(fngi) */
    const double *fx_dn;          /* fx_dn(floc_x,fngi,dim-1) */
    const int32_t *bc_type;       /* nfaces */
    const double *bc_value;       /* floc*nfaces */
    const double *bc_value2;      /* floc*nfaces (robin) */
} orc_surface;

/* The face loop (:606-638 -> :1228-1379): faces in index order, each
 * face's matrix and rhs added with the zero-skipping csr addto.  Adds to A
 * and rhs (after orc_assemble, before orc_dirichlet).  Returns 0, -1 for an
 * entry outside the pattern, -2 for weak Dirichlet with diffusivity. */
int orc_surface_assemble(const orc_params *p, const orc_mesh *m,
                         const orc_surface *s, double *A, double *rhs);

/* One face: mat(floc,floc) column-major, rhs(floc). */
int orc_face(const orc_params *p, const orc_mesh *m, const orc_surface *s,
             int face, double *mat, double *rhs);

/* Boundary_Conditions.F90:1983-2025.  nodes 1-based; inactive[N] set to 1;
 * rhs(node) = (g - T(node))/dt when dt > 0, else g. */
void orc_dirichlet(int nbc, const int32_t *nodes, const double *values,
                   const double *T, double dt, int32_t *inactive,
                   double *rhs);

/* Sparse_Tools.F90:3822-3844 */
void orc_csr_mult(int nrows, const int32_t *findrm, const int32_t *colm,
                  const double *val, const double *x, double *y);

typedef struct {
    int32_t ksp;      /* 0 = cg, 1 = gmres */
    int32_t pc;       /* 0 = none, 1 = jacobi, 3 = sor (multicolour SSOR) */
    double rtol, atol;
    int32_t max_its, restart;
    int32_t nthreads;
} orc_solver;

/* Solvers.F90:111-187 / :1052-1126 (lifting), Petsc_Tools.F90:891-1074
 * (ghost pivot), :671-749 (ghost copy-back), KSPCG/KSPGMRES semantics.
 * x is in/out (initial guess).  Returns 0.  reason follows PETSc's
 * KSPConvergedReason numbering. */
int orc_petsc_solve(const orc_solver *s, int n, const int32_t *findrm,
                    const int32_t *colm, const int32_t *centrm,
                    const double *A, const int32_t *inactive,
                    const double *rhs, double *x, int32_t *its,
                    int32_t *reason, double *true_relres, double *seconds);

/* First-fit colouring of the rows of a 1-based CSR pattern in row order;
 * ptr_out[ncol+1] and rows_out[n] (0-based rows, colour by colour,
 * ascending inside a colour) when ptr_out is not NULL.  Returns ncol. */
int orc_colour_rows(int n, const int32_t *findrm, const int32_t *colm, int32_t *ptr_out,
                    int32_t *rows_out);

/* z = M^-1 r of the multicolour SSOR (forward sweep over the colours from
 * z = 0, backward over all but the last) on a 1-based CSR matrix. */
void orc_ssor_apply(int n, const int32_t *findrm, const int32_t *colm, const double *val,
                    const double *r, double *z);

/* ---- the pressure path ------------------------------------------------
 * The divergence block of construct_momentum_cg (assemble/Momentum_CG.F90:
 * 1377-1403, addto :1468-1469) and the lumped velocity mass it assembles
 * alongside (add_mass_element_cg, :1492-1600, the sum of shape_shape(u, u,
 * density*detwei) over its columns), then assemble_masslumped_cmc
 * (assemble/Assemble_CMC.F90:119-135 -> mult_div_vector_div_T,
 * femtools/Sparse_Matrices_Fields.F90:590-671).  The velocity mesh is the
 * field mesh of orc_mesh (its colouring, quadrature and shape tables); the
 * pressure mesh has its own connectivity and tables on the same quadrature. */
typedef struct {
    int32_t p_loc, p_nodes;
    const int32_t *p_ndglno;      /* p_loc*nele, 1-based */
    const double *p_n, *p_dn;     /* p_n(p_loc,ngi), p_dn(p_loc,ngi,dim) */
    /* ct_m's sparsity: rows = pressure nodes, columns = velocity nodes */
    const int32_t *ct_findrm, *ct_colm;
    const double *density;        /* on the velocity mesh (one value if const) */
    int32_t density_const;
    int32_t integrate_by_parts;   /* integrate_continuity_by_parts */
} orc_pressure;

/* ct_val: dim blocks of nnz(ct) values, block (1,d) at d*nnz; masslump:
 * one value per velocity node (every component holds the same).  Colour
 * loop, block addto skipping zeros.  Returns 0, or -1 for a nonzero value
 * outside the pattern. */
int orc_ct_assemble(const orc_mesh *m, const orc_pressure *pp, double *ct_val, double *masslump);

/* One element's grad_p_u_mat(dim, p_loc, u_loc) (column-major) and
 * mass_lump(u_loc). */
int orc_ct_element(const orc_mesh *m, const orc_pressure *pp, int ele, double *mat, double *mass_lump);

/* cmc = ctp diag(inv_ml) ct^T on the given pattern: entry (i, j) walks rows
 * i of ctp and j of ct together (ascending columns), adding
 * ctp_d(i,k) ct_d(j,k) inv_ml(d,k) for each common k, d inner.  inv_ml is
 * dim x velocity nodes (vector_field%val).  All CSR arrays 1-based. */
void orc_cmc(int rows, const int32_t *ctp_findrm, const int32_t *ctp_colm, const double *ctp_val, int64_t ctp_nnz,
             const int32_t *ct_findrm, const int32_t *ct_colm, const double *ct_val, int64_t ct_nnz, int dim,
             const double *inv_ml, const int32_t *cmc_findrm, const int32_t *cmc_colm, double *cmc_val);

/* Ghost pivot of Petsc_Tools.F90:959-984 over all rows' |diagonal|. */
double orc_ghost_pivot(int n, const int32_t *centrm, const double *A);

#ifdef __cplusplus
}
#endif
#endif
