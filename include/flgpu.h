/*
 * flgpu.h -- C ABI of the MI355X (gfx950) continuous-Galerkin
 * advection-diffusion library.
 *
 * The library replaces two cut points of Fluidity's CG scalar-field solve:
 *   (i)  the element loop of assemble_advection_diffusion_cg
 *        (assemble/Advection_Diffusion_CG.F90:565-598: transform cache,
 *        colouring, colour x element loop, csr addto), and
 *   (ii) the petsc_solve call of solve_advection_diffusion_cg
 *        (assemble/Advection_Diffusion_CG.F90:1389 ->
 *        assemble/Petsc_Solve_State.F90:64-117 -> femtools/Solvers.F90:111-187).
 * Fortran reaches it through ISO_C_BINDING (fluidity_amd/fortran/flgpu.F90);
 * every argument is a plain pointer or scalar.
 *
 * Conventions (the Fortran ones, so arrays can be passed straight from
 * mesh_type / field / csr_sparsity components):
 *   - integers are int32 (Fortran default integer), reals are double;
 *   - node and element numbers are 1-based; ndglno is element-major
 *     (ndglno(loc*(ele-1)+i));
 *   - vector fields are dim x nodes column-major (vector_field%val),
 *     tensors dim x dim x nodes (tensor_field%val);
 *   - shape tables are element_type%n(loc,ngi), %dn(loc,ngi,dim),
 *     quadrature%weight(ngi), all column-major.
 * Every entry point returns 0 on success and a nonzero FLGPU_E* code on
 * failure; flgpu_last_error() then holds a message (the Fortran shim turns
 * it into FLAbort, as csr_addto does for an out-of-pattern insertion,
 * femtools/Sparse_Tools.F90:2645-2647).
 *
 * Threading: one context per process and GPU, not re-entrant; call it from
 * the serial region, as Fluidity does for construct_cache
 * (femtools/Transform_elements.F90:238).
 */
#ifndef FLGPU_H
#define FLGPU_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum {
    FLGPU_OK = 0,
    FLGPU_EINVAL = 1,     /* bad argument or call order */
    FLGPU_EHIP = 2,       /* HIP runtime error */
    FLGPU_EPATTERN = 3,   /* element entry outside the CSR sparsity */
    FLGPU_ECOLOUR = 4,    /* colouring invalid for the mesh */
    FLGPU_ENOMEM = 5,
    FLGPU_EUNSUPPORTED = 6
};

typedef struct flgpu_ctx flgpu_ctx;

/* ---- lifetime ------------------------------------------------------ */

/* New context on HIP device `device` (one HIP stream).  No counterpart in
 * Fluidity: it replaces the implicit module state (transform cache,
 * colouring sets, cached ksp) that the Fortran modules keep. */
int flgpu_init(flgpu_ctx **ctx, int32_t device);
int flgpu_finalize(flgpu_ctx *ctx);
/* Message of the last failure on this thread. */
const char *flgpu_last_error(void);

/* ---- mesh binding (setup; cached on the device) ----------------------- */

/* Coordinate field: replaces the element geometry cache
 * (femtools/Transform_elements.F90:81-100 arrays, :131-168 validity keyed
 * by X%refcount%id and eventcount(EVENT_MESH_MOVEMENT), :227-306
 * construct_cache).  A call with the same (key_id, key_event) is a no-op.
 * Linear simplices only (loc_x == dim+1), dim 1, 2 or 3 (intervals
 * assemble on the colour scheme and have point faces; the pressure path
 * takes dim 2 and 3).
 * ndglno_x: loc_x*nele; X: dim x nx. */
int flgpu_mesh_bind(flgpu_ctx *ctx, int64_t key_id, int64_t key_event,
                    int32_t dim, int32_t loc_x, int32_t nele, int32_t nx,
                    const int32_t *ndglno_x, const double *X);

/* Field mesh and element tables of T (ele_shape(t,ele)): element_type%n,
 * %dn, quadrature%weight (femtools/Elements.F90:38-65) and the coordinate
 * element's dn at the same quadrature (x_dn(loc_x,ngi,dim)).  Lagrange
 * simplices of degree 1 or 2.  Velocity, diffusivity, source and
 * absorption live on this mesh (or are constant). */
int flgpu_field_mesh_bind(flgpu_ctx *ctx, int32_t loc, int32_t ngi,
                          int32_t nnodes, const int32_t *ndglno,
                          const double *n, const double *dn, const double *w,
                          const double *x_dn);

/* First-order sparsity of the field mesh: csr_sparsity%findrm(rows+1),
 * %colm(nnz), %centrm(rows), 1-based, sorted rows
 * (femtools/Sparse_Tools.F90:48-71; built by make_sparsity,
 * femtools/Sparsity_Patterns.F90:47-85). */
int flgpu_sparsity_bind(flgpu_ctx *ctx, int32_t rows, int32_t cols,
                        const int32_t *findrm, const int32_t *colm,
                        const int32_t *centrm);

/* COLOURING_CG1 element colour sets (femtools/Colouring.F90:85-262),
 * flattened: col_ptr(ncol+1) 1-based offsets into elems; each set lists
 * its 1-based element ids ascending (the Judy1 fetch order,
 * femtools/Data_structures_C.c:42-47).  Validated on bind: no two
 * elements of one colour may share a node (verify_colour_sparsity,
 * femtools/Colouring.F90:202-221). */
int flgpu_colouring_bind(flgpu_ctx *ctx, int32_t ncol, const int32_t *col_ptr,
                         const int32_t *elems);

/* ---- host-side setup helpers (the Fortran caller has its own) -------- */

/* make_sparsity (femtools/Sparsity_Patterns.F90:47-85, :299-428) for a
 * continuous mesh: row i = ascending union of the nodes of the elements
 * that contain i.  Call with colm == NULL to get nnz in *nnz; then again
 * with colm of that size.  findrm(nnodes+1), centrm(nnodes), 1-based. */
int flgpu_make_sparsity(int32_t nnodes, int32_t nele, int32_t loc,
                        const int32_t *ndglno, int32_t *findrm, int32_t *colm,
                        int32_t *centrm, int64_t *nnz);

/* get_mesh_colouring(..., COLOURING_CG1, ...) (femtools/Colouring.F90:
 * 85-199): greedy first-fit over elements in index order, neighbours =
 * elements sharing a node.  colour(nele) 1-based; *ncol = colours used. */
int flgpu_colour_mesh(int32_t nnodes, int32_t nele, int32_t loc,
                      const int32_t *ndglno, int32_t *colour, int32_t *ncol);

/* ---- set-up on the device (meshes that change) ------------------------ */
/* make_sparsity of the bound field mesh computed on the GPU (the arrays of
 * flgpu_make_sparsity, bit for bit) and bound as by flgpu_sparsity_bind;
 * *nnz (if non-NULL) gets the entry count.  No counterpart call in Fluidity:
 * it replaces the CPU make_sparsity (femtools/Sparsity_Patterns.F90:47-85,
 * :299-428) that every adapt repeats. */
int flgpu_sparsity_build(flgpu_ctx *ctx, int64_t *nnz);
/* Copy the bound sparsity out: findrm(rows+1), colm(nnz), centrm(rows),
 * 1-based (any may be NULL). */
int flgpu_sparsity_get(flgpu_ctx *ctx, int32_t *findrm, int32_t *colm,
                       int32_t *centrm);
/* An element colouring computed on the GPU and bound as by
 * flgpu_colouring_bind (elements sharing a node differ; speculative
 * first-fit rounds with hashed priorities, deterministic).  It is not the greedy
 * index-order colouring of get_mesh_colouring (femtools/Colouring.F90:
 * 85-199), so matrix values equal the reference's up to the order of the
 * colour sums; bind the reference colouring where that order matters. */
int flgpu_colouring_build(flgpu_ctx *ctx, int32_t *ncol);
/* Copy the bound colour sets out: col_ptr(ncol+1), elems(nele), 1-based. */
int flgpu_colouring_get(flgpu_ctx *ctx, int32_t *col_ptr, int32_t *elems);
/* ---- fields ------------------------------------------------------- */

enum {
    FLGPU_FIELD_T = 0,          /* scalar, the prognostic field (old value) */
    FLGPU_FIELD_VELOCITY = 1,   /* vector dim x N */
    FLGPU_FIELD_DIFFUSIVITY = 2,/* tensor dim x dim x N */
    FLGPU_FIELD_SOURCE = 3,     /* scalar */
    FLGPU_FIELD_ABSORPTION = 4, /* scalar */
    /* scalars of the internal-energy and k-epsilon equations
     * (assemble/Advection_Diffusion_CG.F90:484-552), on the field mesh */
    FLGPU_FIELD_DENSITY = 5,
    FLGPU_FIELD_OLD_DENSITY = 6,
    FLGPU_FIELD_PRESSURE = 7,
    FLGPU_FIELD_VOLUME_FRACTION = 8 /* PhaseVolumeFraction (nvfrac) */
};

/* Mesh movement inputs (Advection_Diffusion_CG.F90:421-437):
 * OldCoordinate and IteratedCoordinate positions and GridVelocity, each
 * dim x nx on the coordinate mesh (vector_field%val).  Used by assemblies
 * with flgpu_advdif_params.move_mesh; kept until the coordinate mesh
 * changes topology. */
int flgpu_mesh_motion_set(flgpu_ctx *ctx, const double *X_old, const double *X_new,
                          const double *grid_velocity);

/* Upload a field to its device mirror.  constant != 0: FIELD_TYPE_CONSTANT,
 * values holds one value (ncomp numbers); it is interpolated like a nodal
 * field, as ele_val_at_quad does (femtools/Fields_Base.F90:2256-2310). */
int flgpu_field_set(flgpu_ctx *ctx, int32_t which, int32_t constant,
                    const double *values);

/* ---- assembly --------------------------------------------------------- */

typedef struct {
    double dt, theta, beta;          /* dt_theta = dt*theta; beta =
                                        conservative_advection */
    int32_t have_mass, lump_mass;
    int32_t have_advection, integrate_by_parts;
    int32_t have_diffusivity, isotropic_diffusivity;
    int32_t have_source, have_absorption;
    int32_t precision;               /* FLGPU_FP64 or FLGPU_FP32 element
                                        arithmetic (values stored FP64) */
    int32_t scheme;                  /* FLGPU_ASM_* */
    /* spatial_discretisation/continuous_galerkin/stabilisation
     * (assemble/Advection_Diffusion_CG.F90:440-465): FLGPU_STAB_*, the
     * nu_bar scheme (FLGPU_NU_BAR_*, assemble/Upwind_Stabilisation.F90:
     * 18-19) and nu_scale.  Stabilised assembly needs the velocity field;
     * FLGPU_ASM_BLOCKED takes it for linear T on the coordinate mesh (the
     * same bits as FLGPU_ASM_COLOUR), the colour scheme for any degree. */
    int32_t stabilisation, nu_bar_scheme;
    double nu_bar_scale;
    /* the equation type (Advection_Diffusion_CG.F90:484-552, FLGPU_EQ_*):
     * internal energy weights the mass by the old density and the
     * advection by density_theta*density + (1-density_theta)*old density,
     * with compressible != 0 adding -p div(u) to the rhs and multiphase != 0
     * the PhaseVolumeFraction factor (mass, advection, isotropic
     * diffusion, pressure term); k-epsilon weights mass and advection by
     * the density.  Unstabilised they assemble in two passes (FLGPU_ASM_AUTO
     * or FLGPU_ASM_BLOCKED: rows in one launch, then per-entry colour-order
     * sums) or on FLGPU_ASM_COLOUR, the same bits; stabilised on the colour
     * scheme. */
    double density_theta;
    int32_t equation_type, compressible, multiphase;
    /* mesh movement (/mesh_adaptivity/mesh_movement with mass terms,
     * :421-437): the mass at the new positions, -(N^{n+1} - N^n) T / dt on
     * the rhs, the velocity relative to the grid and, not by parts,
     * -N_A div(u_grid) N_B (flgpu_mesh_motion_set gives the positions and
     * grid velocity).  Linear T (schemes as for the equation types), advection-diffusion
     * without stabilisation (the reference stops on the other cases). */
    int32_t move_mesh;
} flgpu_advdif_params;
enum { FLGPU_EQ_ADVECTION_DIFFUSION = 0, FLGPU_EQ_INTERNAL_ENERGY = 1, FLGPU_EQ_KEPSILON = 2 };
enum { FLGPU_STAB_NONE = 0, FLGPU_STAB_STREAMLINE_UPWIND = 1, FLGPU_STAB_SUPG = 2 };
enum {
    FLGPU_NU_BAR_OPTIMAL = 1,
    FLGPU_NU_BAR_DOUBLY_ASYMPTOTIC = 2,
    FLGPU_NU_BAR_CRITICAL_RULE = 3,
    FLGPU_NU_BAR_UNITY = 4
};

enum { FLGPU_FP64 = 0, FLGPU_FP32 = 1, FLGPU_MIXED = 1 };
enum {
    FLGPU_ASM_AUTO = 0,     /* fastest available */
    FLGPU_ASM_COLOUR = 1,   /* one launch per colour, element per thread,
                               read-modify-write of the CSR in HBM */
    FLGPU_ASM_BLOCKED = 2   /* linear T: row blocks in LDS, colours in
                               order inside each block; quadratic T
                               (unstabilised): element values in one pass,
                               then one thread per CSR entry adds them in
                               colour order (same bits as FLGPU_ASM_COLOUR) */
};

/* assemble_advection_diffusion_cg's element loop
 * (assemble/Advection_Diffusion_CG.F90:558-600 -> :702-1202): zero the
 * matrix and rhs, then add every element's contribution colour by colour.
 * Inputs are the device mirrors set by flgpu_field_set; the result stays
 * on the device (flgpu_matrix_get copies it out). */
int flgpu_advdif_assemble_resident(flgpu_ctx *ctx, const flgpu_advdif_params *p);

/* One-shot form for the Fortran drop-in: upload the non-NULL fields
 * (nodal), assemble, and copy matrix%val (nnz) and rhs%val (rows) back
 * when A_val / rhs are non-NULL. */
int flgpu_advdif_assemble(flgpu_ctx *ctx, const flgpu_advdif_params *p,
                          const double *T, const double *U,
                          const double *kappa, const double *src,
                          const double *absorp, double *A_val, double *rhs);

/* rhs += field on the device-resident rhs (scalar_field addto,
 * femtools/Fields_Manipulation.F90:255-285): the source's
 * "add_directly_to_rhs" option (assemble/Advection_Diffusion_CG.F90:321,
 * :604) -- assemble with have_source = 0, then add FLGPU_FIELD_SOURCE here,
 * before the face loop and Dirichlet.  A constant field adds its value to
 * every row. */
int flgpu_rhs_add_field(flgpu_ctx *ctx, int32_t which);
/* ---- surface integrals ------------------------------------------------- */
/* Boundary-condition types of the face loop, the numbering of
 * assemble/Advection_Diffusion_CG.F90:74-75 (0 = none of them, e.g. a strong
 * Dirichlet face). */
enum {
    FLGPU_BC_NONE = 0,
    FLGPU_BC_NEUMANN = 1,
    FLGPU_BC_WEAKDIRICHLET = 2,
    FLGPU_BC_INTERNAL = 3,
    FLGPU_BC_ROBIN = 4
};
/* Surface elements of T for the face loop
 * (assemble/Advection_Diffusion_CG.F90:606-638): per face
 * face_global_nodes(t, face) (face_ndglno, floc*nfaces),
 * face_global_nodes(positions, face) (face_ndglno_x, dim*nfaces) and
 * face_ele (nfaces), all 1-based (femtools/Fields_Base.F90:1867-1932);
 * the face element of T on the surface quadrature, face_shape(t)%n
 * (fn(floc,fngi)) and %quadrature%weight (fw(fngi)), and the coordinate
 * face element's face_shape(positions)%dn (fx_dn(dim,fngi,dim-1)).  Faces of
 * the field mesh's degree (P1 or P2 segments / triangles); the coordinate
 * mesh is linear.  Rebinding the field mesh unbinds the surface. */
int flgpu_surface_bind(flgpu_ctx *ctx, int32_t nfaces, int32_t floc,
                       int32_t fngi, const int32_t *face_ndglno,
                       const int32_t *face_ndglno_x, const int32_t *face_ele,
                       const double *fn, const double *fw,
                       const double *fx_dn);
/* The face loop of assemble_advection_diffusion_cg (:606-638 ->
 * assemble_advection_diffusion_face_cg :1228-1379): for every face not of
 * type FLGPU_BC_INTERNAL, the boundary term of advection integrated by parts
 * (with the weak Dirichlet value on FLGPU_BC_WEAKDIRICHLET faces) and the
 * Neumann / Robin terms of the diffusivity, added to the device-resident
 * matrix and rhs in face order with csr_addto's skip of zeros.  Call after
 * the element assembly and before flgpu_dirichlet_apply, with the same
 * params.  bc_type(nfaces) as get_entire_boundary_condition returns it;
 * bc_value / bc_value2 (floc*nfaces) are the bc fields at each face's
 * nodes, ele_val(t_bc, face) and ele_val(t_bc_2, face) (the Robin
 * coefficient); either may be NULL when no face needs it.  Weak Dirichlet
 * with diffusivity is FLGPU_EUNSUPPORTED, as it is an FLExit in Fluidity
 * (:1374-1377).  The internal-energy equation weights the face advection by
 * its theta-weighted density (and volume fraction), and mesh movement takes
 * the velocity relative to the grid there (:1313-1327).  FP64 arithmetic
 * whatever p->precision says. */
int flgpu_advdif_surface(flgpu_ctx *ctx, const flgpu_advdif_params *p,
                         const int32_t *bc_type, const double *bc_value,
                         const double *bc_value2);
/* Strong Dirichlet (femtools/Boundary_Conditions.F90:1983-2025): mark the
 * rows inactive (set_inactive, femtools/Sparse_Tools.F90:374-376) and set
 * rhs(node) = (value - T(node))/dt (dt > 0) or value (dt <= 0) on the
 * device-resident rhs.  nodes are 1-based. */
int flgpu_dirichlet_apply(flgpu_ctx *ctx, int32_t nbc, const int32_t *nodes,
                          const double *values, double dt);

/* Copy the device-resident matrix values / rhs / inactive mask to host
 * arrays (any may be NULL). */
int flgpu_matrix_get(flgpu_ctx *ctx, double *A_val, double *rhs,
                     int32_t *inactive);

/* Replace the device-resident matrix values, rhs and inactive mask (any
 * may be NULL; inactive NULL clears the mask only when A_val is given). */
int flgpu_matrix_set(flgpu_ctx *ctx, const double *A_val, const double *rhs,
                     const int32_t *inactive);

/* ---- solve ------------------------------------------------------------ */

/* FLGPU_KSP_PIPECG: "pipecg", the single-reduction CG recurrence of
 * Chronopoulos & Gear (1989) that PETSc's KSPPIPECG is built on (u = M r,
 * w = A u, gamma = r.u and delta = u.w in one reduction phase, s = A p by
 * recurrence): one vector kernel an iteration instead of two; PETSc's
 * overlap of the reduction with the next product is not needed on one
 * in-order stream.  The same preconditioners as cg, and the same FP64
 * reliable updates in mixed precision (each recomputes s = A p and keeps
 * the direction, which restarts after a residual drop above 1e3). */
enum { FLGPU_KSP_CG = 0, FLGPU_KSP_GMRES = 1, FLGPU_KSP_FCG = 2, FLGPU_KSP_PIPECG = 3 };
/* FLGPU_PC_SOR: "sor" (and "eisenstat", SSOR by Eisenstat's trick) of
 * femtools/Solvers.F90:1844-1854 -- one symmetric Gauss-Seidel sweep
 * (omega = 1) taken colour by colour over a first-fit colouring of the
 * matrix graph, instead of PETSc's sequential natural-order sweep. */
/* FLGPU_PC_FSAI: the factorised sparse approximate inverse M^-1 = G^T G
 * with G lower triangular on A's lower pattern, each row the exact
 * minimiser of the Kaporin condition number (Kolotilina & Yeremin): the
 * best factor of the family the learned preconditioner predicts, built
 * per solve from one small Cholesky solve per row; no counterpart in
 * Fluidity's option tree (a preconditioner PETSc's PCHPDDM/PCSPAI family
 * would supply).  A is taken as symmetric, as CG takes it: each row's
 * local matrix is read from A's lower triangle.  Rows with more than 48
 * lower entries keep the Jacobi row. */
enum { FLGPU_PC_NONE = 0, FLGPU_PC_JACOBI = 1, FLGPU_PC_LEARNED = 2, FLGPU_PC_SOR = 3, FLGPU_PC_FSAI = 4 };

typedef struct {
    int32_t ksp, pc;
    double rtol, atol;          /* KSPSetTolerances (femtools/Solvers.F90:1646) */
    int32_t max_its, restart;
    int32_t precision;          /* FLGPU_FP64, or FLGPU_MIXED: FP32 inner
                                   iterations with FP64 residual, update and
                                   reductions */
    int32_t start_from_zero;    /* ignore the initial x (start_from_zero,
                                   femtools/Solvers.F90:1648-1655) */
    int32_t check_every;        /* host polls convergence every k its (0=16) */
    double stall_factor;        /* learned pc: fall back to Jacobi when the
                                   residual has not dropped by this factor
                                   over stall_window its (0 = 0.9) */
    int32_t stall_window;       /* (0 = 50) */
    int32_t norm_type;          /* FLGPU_NORM_*: the norm of the convergence
                                   test (PETSc KSPSetNormType) */
} flgpu_solver_params;

enum {
    FLGPU_NORM_PRECONDITIONED = 0,   /* ||M^-1 r||, PETSc's default for CG */
    FLGPU_NORM_UNPRECONDITIONED = 1  /* ||r||: the true-residual test */
};

typedef struct {
    int32_t its;          /* iterations taken (iterations_taken) */
    int32_t reason;       /* PETSc KSPConvergedReason numbering: >0 converged,
                             -3 max its, -9 NaN/Inf, ... */
    double true_relres;   /* ||b - A x||_2 / ||b||_2 of the reduced system,
                             recomputed in FP64 after the solve */
    double precond_relres;/* last convergence-test norm / first one */
    int32_t fallback;     /* 1 if the learned preconditioner stalled and the
                             solve finished with Jacobi */
    int32_t refinements;  /* FP64 residual replacements (mixed precision) */
    double seconds;       /* device time of the Krylov loop */
} flgpu_solve_info;

/* petsc_solve on the device-resident system (femtools/Solvers.F90:111-187):
 * the inactive rows become ghost rows with the diagonal ghost_pivot
 * (femtools/Petsc_Tools.F90:959-984, :1045-1057), the inactive columns are
 * lifted into the rhs (femtools/Solvers.F90:1072-1099), the Krylov method
 * runs on the reduced system, and the ghost entries of x are set to the
 * rhs values (femtools/Petsc_Tools.F90:724-731).  x: in = initial guess
 * (unless start_from_zero), out = solution; NULL keeps x on the device. */
int flgpu_solve(flgpu_ctx *ctx, const flgpu_solver_params *sp, double *x,
                flgpu_solve_info *info);

/* flgpu_solve without allocation.  No counterpart in the reference: it is
 * for callers that cannot allocate device memory (code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc).  The
 * Krylov work (reduced matrix values, seven vectors, partial sums, solver
 * scalars, row blocks: about 8*nnz + 56*rows bytes plus 180 KiB) lives in
 * one fixed device array of 64 MiB per GPU, set up by flgpu_init and held
 * until the process ends; a larger system is rejected with
 * FLGPU_EUNSUPPORTED.  FP64 CG or FCG with pc none or jacobi; x (host,
 * rows values) is required: in = initial guess (unless start_from_zero),
 * out = solution (the solution is not kept for flgpu_solution_get).  All
 * contexts on a GPU share that GPU's array, which keeps the last call's
 * contents: calls that overlap on one GPU (several threads or contexts)
 * race. */
int flgpu_solve_noalloc(flgpu_ctx *ctx, const flgpu_solver_params *sp,
                        double *x, flgpu_solve_info *info);

/* Drop-in form of (ii): takes host arrays like petsc_solve(x, matrix, rhs).
 * A_val / rhs / inactive replace the device-resident system when non-NULL. */
int flgpu_petsc_solve(flgpu_ctx *ctx, const flgpu_solver_params *sp,
                      const double *A_val, const int32_t *inactive,
                      const double *rhs, double *x, int32_t *its,
                      int32_t *reason, double *true_relres);

/* PCApply of the preconditioner flgpu_solve builds for the reduced system
 * (setup_pc_from_options, femtools/Solvers.F90:1844-1854): z = M^-1 r for
 * pc none, jacobi, sor, learned or fsai (the factor built and applied in
 * FP64, as flgpu_solve does with precision FLGPU_FP64), host arrays of rows
 * doubles. */
int flgpu_pc_apply(flgpu_ctx *ctx, const flgpu_solver_params *sp,
                   const double *r, double *z);

/* Copy the last device-resident solution (rows) to the host. */
int flgpu_solution_get(flgpu_ctx *ctx, double *x);

/* y = A x with the device-resident assembled matrix (csr_mult,
 * femtools/Sparse_Tools.F90:3822-3844); host arrays. */
int flgpu_csr_mult(flgpu_ctx *ctx, const double *x, double *y);

/* ---- matrix dumps -------------------------------------------------------
 * The PETSc binary file DumpMatrixEquation writes after a failed solve
 * (femtools/Petsc_Tools.F90:1487-1504, called from ConvergenceCheck,
 * femtools/Solvers.F90:1369-1479): MatView(A), VecView(b), VecView(x0),
 * big-endian, 32-bit indices.  Host-side; the arrays are 1-based CSR as
 * flgpu_sparsity_bind / flgpu_matrix_set take them, so a dump replays
 * through flgpu_solve (the GPU form of tools/petsc_readnsolve.F90). */
/* Read a dump.  With findrm == NULL only *n and *nnz are returned (size
 * query); then findrm(n+1), colm(nnz), val(nnz), b(n) and, if non-NULL,
 * x0(n) are filled; *has_x0 (if non-NULL) says whether the file held x0.
 * Square matrices with ascending columns per row only. */
int flgpu_matrixdump_read(const char *filename, int32_t *n, int64_t *nnz,
                          int32_t *findrm, int32_t *colm, double *val,
                          double *b, double *x0, int32_t *has_x0);
/* Write a dump of a CSR system (x0 may be NULL: then only A and b). */
int flgpu_matrixdump_write(const char *filename, int32_t n,
                           const int32_t *findrm, const int32_t *colm,
                           const double *val, const double *b,
                           const double *x0);
/* ---- the pressure path ---------------------------------------------------
 * construct_momentum_cg's divergence block ct_m and lumped velocity mass
 * (assemble/Momentum_CG.F90:1377-1403, :1468-1469, :1492-1600) and
 * assemble_masslumped_cmc (assemble/Assemble_CMC.F90:119-135).  The
 * velocity mesh is the bound field mesh (flgpu_field_mesh_bind: u's
 * connectivity and element tables) with its colouring. */

/* The pressure mesh: ndglno of p (p_loc per element, 1-based),
 * ele_shape(p)%n(p_loc,ngi) and %dn(p_loc,ngi,dim) on the velocity's
 * quadrature, and ct_m's sparsity (get_csr_sparsity_firstorder(state,
 * p%mesh, u%mesh): ct_rows = pressure nodes, columns = velocity nodes,
 * 1-based, sorted rows). */
int flgpu_pressure_bind(flgpu_ctx *ctx, int32_t p_loc, int32_t p_nodes,
                        const int32_t *p_ndglno, const double *p_n,
                        const double *p_dn, int32_t ct_rows,
                        const int32_t *ct_findrm, const int32_t *ct_colm);

/* ct_m = block (1,d) of grad_p_u_mat, shape_dshape(p_shape, du_t, detwei)
 * or, with integrate_by_parts (integrate_continuity_by_parts),
 * -dshape_shape(dp_t, u_shape, detwei), added colour by colour as
 * addto(ct_m, p_ele, u_ele, ...) adds it; and the lumped mass
 * sum(shape_shape(u_shape, u_shape, density*detwei), 2) added per velocity
 * node (the non-inverted inverse_masslump, one component).  density: one
 * value (density_constant) or one per velocity node.  ct_val (dim blocks of
 * nnz(ct), block d at d*nnz) and masslump (velocity nodes) are copied out
 * when not NULL; both stay on the device for flgpu_cmc_assemble. */
int flgpu_ct_assemble(flgpu_ctx *ctx, int32_t integrate_by_parts,
                      int32_t density_constant, const double *density,
                      double *ct_val, double *masslump);

/* cmc = ct_m diag(inverse_masslump) ct_m^T on the given pattern
 * (mult_div_vector_div_T, femtools/Sparse_Matrices_Fields.F90:590-671):
 * inverse_masslump is dim x velocity nodes (vector_field%val, Dirichlet
 * rows already zeroed by the caller), or NULL for 1/masslump in every
 * component; cmc_findrm(cmc_rows+1) / cmc_colm 1-based, sorted; cmc_val
 * (nnz) copied out when not NULL. */
int flgpu_cmc_assemble(flgpu_ctx *ctx, const double *inverse_masslump,
                       int32_t cmc_rows, const int32_t *cmc_findrm,
                       const int32_t *cmc_colm, double *cmc_val);
/* Device milliseconds per call of the ct_m + lumped-mass kernels and of
 * the C^T M_l^-1 C kernel (the inverse lumped mass from the device's own
 * lumped mass), reps calls each on the inputs of the last
 * flgpu_ct_assemble / flgpu_cmc_assemble, host copies and checks excluded
 * (no counterpart in Fluidity; measurement).  The device copies of ct_m
 * and C^T M_l^-1 C are recomputed in place (cmc with 1/masslump). */
int flgpu_time_pressure(flgpu_ctx *ctx, int32_t reps, double *ct_ms,
                        double *cmc_ms);

/* ---- learned preconditioner --------------------------------------------
 * A message-passing network on the matrix graph (3 layers, hidden 32,
 * BF16 MFMA with FP32 accumulation) predicts a lower-triangular factor G on
 * the IC(0) pattern (lower triangle of A); the preconditioner is
 * M^-1 = G^T G.  No counterpart in Fluidity (new component). */

/* Number of float parameters of the network. */
int flgpu_learned_param_count(void);
/* Set / get the network parameters (float32, flgpu_learned_param_count). */
int flgpu_learned_set_params(flgpu_ctx *ctx, const float *params);
int flgpu_learned_get_params(flgpu_ctx *ctx, float *params);
/* Deterministic initialisation from a seed. */
int flgpu_learned_init_params(flgpu_ctx *ctx, uint64_t seed);
/* One Adam step on the device-resident system (after assembly + Dirichlet):
 * loss = E_z ||z - G^T G A z||^2 / ||z||^2 over nprobe Rademacher probes.
 * *loss gets the loss before the step. */
int flgpu_learned_train_step(flgpu_ctx *ctx, int32_t nprobe, uint64_t seed,
                             float lr, float *loss);
/* Loss and the parameter gradient without updating (grad may be NULL). */
int flgpu_learned_loss_grad(flgpu_ctx *ctx, int32_t nprobe, uint64_t seed,
                            float *loss, float *grad);
/* Build G for the device-resident system (done by flgpu_solve when
 * pc == FLGPU_PC_LEARNED).  *nnz = nnz of A; with G_val != NULL the factor
 * is copied out in A's pattern, with D = diag(A)^(-1/2) and Gs the factor
 * of the scaled matrix D A D: entry (i,j) of the lower triangle holds
 * Gs_ij D_jj, entry (i,j) of the upper triangle holds D_ii Gs_ji and the
 * diagonal holds Gs_ii D_ii (so G = Gs D is the lower part with the
 * diagonal, and G^T the upper part with the diagonal). */
int flgpu_learned_factor(flgpu_ctx *ctx, float *G_val, int64_t *nnz);

/* ---- measurement --------------------------------------------------------
 * Device-time measurement on the library's own stream with HIP events (no
 * counterpart in Fluidity, which times with its profiler keys
 * advection_diffusion_loop / solve, femtools/Profiler.cpp:33-174). */

/* reps calls of the resident assembly; *ms = mean device milliseconds per
 * call (the element loop only: zeroing included, host checks excluded). */
int flgpu_time_assemble(flgpu_ctx *ctx, const flgpu_advdif_params *p,
                        int32_t reps, double *ms);
/* The blocked scheme's plan for the bound mesh (built here if it is not
 * yet): info[0..n-1] of staged (1: node-staged slabs), slabs, the largest
 * slab's rows, entries, slots and staged nodes, staged nodes summed over
 * the slabs, element rows (slots) in all, the LDS bytes of a slab for the
 * FP64 kernel with diffusivity and source staged, and the workgroups per
 * CU that LDS allows; *build_seconds (may be NULL) = host seconds of the
 * last plan build. */
int flgpu_assembly_plan_info(flgpu_ctx *ctx, int64_t *info, int32_t n,
                             double *build_seconds);
/* reps solver-matrix SpMVs (FLGPU_FP64 or FLGPU_MIXED values) after a solve
 * or flgpu_solver_prepare; *ms = mean device milliseconds per product. */
int flgpu_time_spmv(flgpu_ctx *ctx, int32_t precision, int32_t reps,
                    double *ms);
/* As flgpu_time_spmv with cold caches: before each product a kernel reads
 * and writes flush_mib MiB (1..8192) of scratch memory, evicting the L2
 * caches and the Infinity Cache; each product is timed alone; *ms = their
 * mean. */
int flgpu_time_spmv_cold(flgpu_ctx *ctx, int32_t precision, int32_t reps,
                         int32_t flush_mib, double *ms);
/* Build the reduced system of flgpu_solve without iterating (for timing). */
int flgpu_solver_prepare(flgpu_ctx *ctx, int32_t precision);
/* Number of CSR rows / nonzeros / colours bound (any pointer may be NULL). */
int flgpu_sizes(flgpu_ctx *ctx, int32_t *rows, int64_t *nnz, int32_t *ncol);

#ifdef __cplusplus
}
#endif
#endif /* FLGPU_H */
