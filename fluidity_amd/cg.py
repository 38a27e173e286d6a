"""Host-side mirror of Fluidity's CG scalar-field solve over the gfx950
library (the caller side of include/flgpu.h).

The call sequence is the one of solve_field_equation_cg
(assemble/Advection_Diffusion_CG.F90:127-174):

    problem = CGProblem(mesh, element)            # state, sparsity, colouring
    problem.set_fields(T=..., velocity=..., ...)  # scalar/vector/tensor fields
    assemble_advection_diffusion_cg(problem, opts, dt)      # :228-659
    apply_dirichlet_conditions(problem, nodes, values, dt)  # :643
    delta_t, info = petsc_solve(problem, solver_options)    # :1389
    T += dt * delta_t                                       # :1396-1407

Option names follow the .flml schema (schemas/solvers.rnc,
temporal_discretisation/theta, conservative_advection, ...).  Every call
goes to libflgpu.so; there is no CPU path.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass

import numpy as np

from . import lib as L
from .elements import Element, make_element, make_face_element
from .meshgen import Mesh, Surface


@dataclass
class CGOptions:
    """spatial_discretisation / temporal_discretisation options read by
    assemble_advection_diffusion_cg (:282-556)."""
    theta: float = 0.5
    beta: float = 0.0                    # conservative_advection
    exclude_mass_terms: bool = False
    lump_mass_matrix: bool = False
    exclude_advection_terms: bool = False
    integrate_advection_by_parts: bool = False
    have_diffusivity: bool = True
    isotropic_diffusivity: bool = True
    have_source: bool = False
    have_absorption: bool = False
    precision: str = "fp64"              # element arithmetic: fp64 | fp32
    scheme: str = "auto"                 # auto | colour | blocked
    add_source_directly_to_rhs: bool = False  # source/diagnostic/add_directly_to_rhs
    stabilisation: str = "none"          # none | streamline_upwind | supg
    nu_bar_scheme: str = "optimal"       # optimal | doubly_asymptotic | critical_rule | unity
    nu_scale: float = 0.5
    # the field's equation (prognostic/equation::AdvectionDiffusion,
    # InternalEnergy or KEpsilon, :484-552): the internal energy's density
    # theta (the density field's temporal theta when compressible),
    # compressible (equation_of_state/compressible: the p div(u) term) and
    # multiphase (more than one prognostic velocity: PhaseVolumeFraction)
    equation_type: str = "advection_diffusion"  # advection_diffusion | internal_energy | k_epsilon
    density_theta: float = 1.0
    compressible: bool = False
    multiphase: bool = False
    # /mesh_adaptivity/mesh_movement (with the mass terms): positions and
    # grid velocity from CGProblem.set_mesh_motion
    move_mesh: bool = False

    def params(self, dt: float) -> L.AdvdifParams:
        return L.AdvdifParams(
            dt=dt, theta=self.theta, beta=self.beta,
            have_mass=int(not self.exclude_mass_terms), lump_mass=int(self.lump_mass_matrix),
            have_advection=int(not self.exclude_advection_terms),
            integrate_by_parts=int(self.integrate_advection_by_parts),
            have_diffusivity=int(self.have_diffusivity),
            isotropic_diffusivity=int(self.isotropic_diffusivity),
            have_source=int(self.have_source and not self.add_source_directly_to_rhs),
            have_absorption=int(self.have_absorption),
            precision={"fp64": L.FLGPU_FP64, "fp32": L.FLGPU_FP32}[self.precision],
            scheme={"auto": L.FLGPU_ASM_AUTO, "colour": L.FLGPU_ASM_COLOUR,
                    "blocked": L.FLGPU_ASM_BLOCKED}[self.scheme],
            stabilisation=L.FLGPU_STAB[self.stabilisation], nu_bar_scheme=L.FLGPU_NU_BAR[self.nu_bar_scheme],
            nu_bar_scale=self.nu_scale, density_theta=self.density_theta,
            equation_type=L.FLGPU_EQ[self.equation_type], compressible=int(self.compressible),
            multiphase=int(self.multiphase), move_mesh=int(self.move_mesh))


# preconditioner names of the solver options (setup_pc_from_options,
# femtools/Solvers.F90:1844-1854): "eisenstat" is SSOR applied with
# Eisenstat's trick, the same preconditioner as "sor" (one symmetric sweep).
PRECONDITIONERS = {"none": L.FLGPU_PC_NONE, "jacobi": L.FLGPU_PC_JACOBI, "sor": L.FLGPU_PC_SOR,
                   "eisenstat": L.FLGPU_PC_SOR, "learned": L.FLGPU_PC_LEARNED,
                   "fsai": L.FLGPU_PC_FSAI}


@dataclass
class SolverOptions:
    """solver options (schemas/solvers.rnc:107-612)."""
    iterative_method: str = "cg"         # cg | gmres | fcg | pipecg
    preconditioner: str = "jacobi"       # jacobi | none | sor | eisenstat | learned | fsai
    relative_error: float = 1e-8
    absolute_error: float = 0.0
    max_iterations: int = 10000
    restart: int = 30
    precision: str = "fp64"              # fp64 | mixed
    start_from_zero: bool = True
    norm: str = "preconditioned"         # preconditioned | unpreconditioned
    check_every: int = 16
    stall_factor: float = 0.0
    stall_window: int = 0

    def params(self) -> L.SolverParams:
        return L.SolverParams(
            ksp={"cg": L.FLGPU_KSP_CG, "gmres": L.FLGPU_KSP_GMRES, "fcg": L.FLGPU_KSP_FCG,
                 "pipecg": L.FLGPU_KSP_PIPECG}[self.iterative_method],
            pc=PRECONDITIONERS[self.preconditioner],
            rtol=self.relative_error, atol=self.absolute_error, max_its=self.max_iterations,
            restart=self.restart, precision={"fp64": L.FLGPU_FP64, "mixed": L.FLGPU_MIXED}[self.precision],
            start_from_zero=int(self.start_from_zero), check_every=self.check_every,
            stall_factor=self.stall_factor, stall_window=self.stall_window,
            norm_type={"preconditioned": L.FLGPU_NORM_PRECONDITIONED,
                       "unpreconditioned": L.FLGPU_NORM_UNPRECONDITIONED}[self.norm])


@dataclass
class Sparsity:
    """csr_sparsity (femtools/Sparse_Tools.F90:48-71), 1-based."""
    findrm: np.ndarray
    colm: np.ndarray
    centrm: np.ndarray

    @property
    def nnz(self) -> int:
        return int(self.colm.size)


def make_sparsity(mesh: Mesh) -> Sparsity:
    """make_sparsity(mesh, mesh) through the library's host helper."""
    lib = L.load()
    n = mesh.nnodes
    findrm = np.zeros(n + 1, np.int32)
    centrm = np.zeros(n, np.int32)
    nnz = C.c_int64(0)
    L.check(lib.flgpu_make_sparsity(n, mesh.nele, mesh.loc, L.ptr(mesh.ndglno), L.ptr(findrm), None,
                                    L.ptr(centrm), C.byref(nnz)), "flgpu_make_sparsity")
    colm = np.zeros(nnz.value, np.int32)
    L.check(lib.flgpu_make_sparsity(n, mesh.nele, mesh.loc, L.ptr(mesh.ndglno), L.ptr(findrm), L.ptr(colm),
                                    L.ptr(centrm), C.byref(nnz)), "flgpu_make_sparsity")
    return Sparsity(findrm, colm, centrm)


def get_mesh_colouring(mesh: Mesh) -> tuple[np.ndarray, np.ndarray, np.ndarray]:
    """COLOURING_CG1 colour per element (1-based) and the flattened sets
    (col_ptr 1-based offsets, elems ascending inside each colour)."""
    lib = L.load()
    colour = np.zeros(mesh.nele, np.int32)
    ncol = C.c_int32(0)
    L.check(lib.flgpu_colour_mesh(mesh.nnodes, mesh.nele, mesh.loc, L.ptr(mesh.ndglno), L.ptr(colour),
                                  C.byref(ncol)), "flgpu_colour_mesh")
    order = np.argsort(colour, kind="stable").astype(np.int32)  # ascending ids inside a colour
    counts = np.bincount(colour, minlength=ncol.value + 1)[1:]
    col_ptr = np.concatenate([[1], 1 + np.cumsum(counts)]).astype(np.int32)
    return colour, col_ptr, (order + 1).astype(np.int32)


class CGProblem:
    """One mesh + element bound to a device context (the state_type of the
    path: positions, the field mesh, its sparsity and colouring)."""

    _mesh_key = 0  # X%refcount%id stand-in: every bound mesh gets a new key

    def __init__(self, mesh: Mesh, element: Element | None = None, quad_degree: int = 2,
                 device: int = 0, sparsity: Sparsity | None = None,
                 colouring: tuple | None = None, setup: str = "host"):
        self.lib = L.load()
        self.ctx = C.c_void_p()
        self.sig_int = False  # set when a solve failed and the run should dump and finish
        L.check(self.lib.flgpu_init(C.byref(self.ctx), device), "flgpu_init")
        self.bind(mesh, element, quad_degree, sparsity, colouring, setup)

    def bind(self, mesh: Mesh, element: Element | None = None, quad_degree: int = 2,
             sparsity: Sparsity | None = None, colouring: tuple | None = None, setup: str = "host") -> None:
        """(Re)bind a mesh to this context; the learned network state of the
        context is kept (used to train over many meshes).  setup="device"
        builds the sparsity and an element colouring on the GPU
        (flgpu_sparsity_build, flgpu_colouring_build) instead of taking the
        host ones."""
        self.mesh = mesh
        self.element = element or make_element(mesh.dim, mesh.degree, quad_degree)
        xel = make_element(mesh.dim, 1, self.element.quad_degree)
        if setup == "device":
            self._bind_meshes(mesh, xel)
            nnz = C.c_int64(0)
            L.check(self.lib.flgpu_sparsity_build(self.ctx, C.byref(nnz)), "flgpu_sparsity_build")
            f = np.zeros(mesh.nnodes + 1, np.int32)
            cm = np.zeros(nnz.value, np.int32)
            cen = np.zeros(mesh.nnodes, np.int32)
            L.check(self.lib.flgpu_sparsity_get(self.ctx, L.ptr(f), L.ptr(cm), L.ptr(cen)), "flgpu_sparsity_get")
            self.sparsity = Sparsity(f, cm, cen)
            ncol = C.c_int32(0)
            L.check(self.lib.flgpu_colouring_build(self.ctx, C.byref(ncol)), "flgpu_colouring_build")
            cp = np.zeros(ncol.value + 1, np.int32)
            el = np.zeros(mesh.nele, np.int32)
            L.check(self.lib.flgpu_colouring_get(self.ctx, L.ptr(cp), L.ptr(el)), "flgpu_colouring_get")
            colour = np.zeros(mesh.nele, np.int32)
            for k in range(ncol.value):
                colour[el[cp[k] - 1:cp[k + 1] - 1] - 1] = k + 1
            self.colouring = (colour, cp, el)
            self.ncolours = ncol.value
            self.T = None
            return
        self.sparsity = sparsity or make_sparsity(mesh)
        self.colouring = colouring or get_mesh_colouring(mesh)
        self._bind_meshes(mesh, xel)
        s = self.sparsity
        L.check(self.lib.flgpu_sparsity_bind(self.ctx, mesh.nnodes, mesh.nnodes, L.ptr(s.findrm),
                                             L.ptr(s.colm), L.ptr(s.centrm)), "flgpu_sparsity_bind")
        _, col_ptr, elems = self.colouring
        self.ncolours = col_ptr.size - 1
        L.check(self.lib.flgpu_colouring_bind(self.ctx, self.ncolours, L.ptr(col_ptr), L.ptr(elems)),
                "flgpu_colouring_bind")
        self.T = None

    def _bind_meshes(self, mesh: Mesh, xel: Element) -> None:
        e = self.element
        X = np.ascontiguousarray(mesh.X, dtype=np.float64)
        CGProblem._mesh_key += 1
        L.check(self.lib.flgpu_mesh_bind(self.ctx, CGProblem._mesh_key, 0, mesh.dim, mesh.dim + 1, mesh.nele, mesh.nx,
                                         L.ptr(mesh.ndglno_x), L.ptr(X)), "flgpu_mesh_bind")
        L.check(self.lib.flgpu_field_mesh_bind(self.ctx, e.loc, e.ngi, mesh.nnodes, L.ptr(mesh.ndglno),
                                               L.ptr(e.fortran(e.n)), L.ptr(e.fortran(e.dn)),
                                               L.ptr(np.ascontiguousarray(e.w)), L.ptr(xel.fortran(xel.dn))),
                "flgpu_field_mesh_bind")

    def bind_surface(self, surface: Surface) -> None:
        """Surface elements for the face loop: face_global_nodes of T and of
        the positions, face_ele, and the face elements on the surface rule
        of the volume rule's degree."""
        mesh = self.mesh
        fe = make_face_element(mesh.dim, mesh.degree, self.element.quad_degree)
        fxe = make_face_element(mesh.dim, 1, self.element.quad_degree)
        L.check(self.lib.flgpu_surface_bind(self.ctx, surface.nfaces, surface.floc, fe.ngi,
                                            L.ptr(surface.face_nodes), L.ptr(surface.face_nodes_x),
                                            L.ptr(surface.face_ele), L.ptr(fe.fortran(fe.n)),
                                            L.ptr(np.ascontiguousarray(fe.w)), L.ptr(fxe.fortran(fxe.dn))),
                "flgpu_surface_bind")
        self.surface = surface

    # -- fields --------------------------------------------------------
    def set_mesh_motion(self, old_positions, new_positions, grid_velocity) -> None:
        """OldCoordinate, IteratedCoordinate and GridVelocity, (nx, dim)
        arrays on the coordinate mesh, for assemblies with move_mesh."""
        arrs = [np.ascontiguousarray(a, dtype=np.float64).reshape(-1) for a in
                (old_positions, new_positions, grid_velocity)]
        L.check(self.lib.flgpu_mesh_motion_set(self.ctx, *[L.ptr(a) for a in arrs]), "flgpu_mesh_motion_set")

    def set_field(self, which: int, values, ncomp: int) -> None:
        """Nodal array (nnodes,) / (nnodes, ncomp) or a constant."""
        if np.isscalar(values) or np.asarray(values).size == ncomp:
            v = np.ascontiguousarray(np.broadcast_to(np.asarray(values, np.float64), (ncomp,)))
            const = 1
        else:
            v = np.ascontiguousarray(values, dtype=np.float64).reshape(-1)
            const = 0
        L.check(self.lib.flgpu_field_set(self.ctx, which, const, L.ptr(v)), "flgpu_field_set")

    def set_fields(self, T=None, velocity=None, diffusivity=None, source=None, absorption=None,
                   isotropic: bool = True, density=None, old_density=None, pressure=None,
                   volume_fraction=None) -> None:
        d = self.mesh.dim
        if T is not None:
            self.T = np.ascontiguousarray(T, dtype=np.float64)
            self.set_field(L.FIELD_T, self.T, 1)
        if velocity is not None:
            self.set_field(L.FIELD_VELOCITY, velocity, d)
        if diffusivity is not None:
            k = diffusivity
            if isotropic and not np.isscalar(k) and np.asarray(k).ndim == 1 and np.asarray(k).size == self.mesh.nnodes:
                full = np.zeros((self.mesh.nnodes, d, d))
                for a in range(d):
                    full[:, a, a] = k
                k = full
            elif np.isscalar(k):
                k = (np.eye(d) * k).ravel()
            self.set_field(L.FIELD_DIFFUSIVITY, k, d * d)
        if source is not None:
            self.set_field(L.FIELD_SOURCE, source, 1)
        if absorption is not None:
            self.set_field(L.FIELD_ABSORPTION, absorption, 1)
        for which, v in ((L.FIELD_DENSITY, density), (L.FIELD_OLD_DENSITY, old_density),
                         (L.FIELD_PRESSURE, pressure), (L.FIELD_VOLUME_FRACTION, volume_fraction)):
            if v is not None:
                self.set_field(which, v, 1)

    # -- system --------------------------------------------------------
    def matrix(self):
        """(A values, rhs, inactive mask) copied from the device."""
        A = np.zeros(self.sparsity.nnz)
        rhs = np.zeros(self.mesh.nnodes)
        inact = np.zeros(self.mesh.nnodes, np.int32)
        L.check(self.lib.flgpu_matrix_get(self.ctx, L.ptr(A), L.ptr(rhs), L.ptr(inact)), "flgpu_matrix_get")
        return A, rhs, inact

    def close(self) -> None:
        if self.ctx:
            self.lib.flgpu_finalize(self.ctx)
            self.ctx = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


@dataclass
class SurfaceBCs:
    """What get_entire_boundary_condition hands the face loop
    (assemble/Advection_Diffusion_CG.F90:613-622): per surface element its
    type (BC_NONE, BC_NEUMANN, BC_WEAKDIRICHLET, BC_INTERNAL, BC_ROBIN) and
    the bc fields at its nodes (floc per face; value2 is the Robin
    coefficient)."""
    bc_type: np.ndarray
    value: np.ndarray | None = None
    value2: np.ndarray | None = None


def assemble_advection_diffusion_cg(problem: CGProblem, opts: CGOptions, dt: float,
                                    bcs: SurfaceBCs | None = None) -> None:
    """assemble_advection_diffusion_cg on the device up to the strong
    Dirichlet conditions: the element loop (zero matrix and rhs, colour
    loop, insertion) and, with bcs, the face loop (:606-638) over the bound
    surface."""
    p = opts.params(dt)
    L.check(problem.lib.flgpu_advdif_assemble_resident(problem.ctx, C.byref(p)),
            "flgpu_advdif_assemble_resident")
    if opts.have_source and opts.add_source_directly_to_rhs:  # :604, before the face loop
        L.check(problem.lib.flgpu_rhs_add_field(problem.ctx, L.FIELD_SOURCE), "flgpu_rhs_add_field")
    if bcs is not None:
        t = np.ascontiguousarray(bcs.bc_type, dtype=np.int32)
        v1 = None if bcs.value is None else np.ascontiguousarray(bcs.value, dtype=np.float64)
        v2 = None if bcs.value2 is None else np.ascontiguousarray(bcs.value2, dtype=np.float64)
        L.check(problem.lib.flgpu_advdif_surface(problem.ctx, C.byref(p), L.ptr(t), L.ptr(v1), L.ptr(v2)),
                "flgpu_advdif_surface")


def apply_dirichlet_conditions(problem: CGProblem, nodes: np.ndarray, values, dt: float | None) -> None:
    """apply_dirichlet_conditions_scalar (femtools/Boundary_Conditions.F90:
    1983-2025) in rate-of-change form when dt is given."""
    nodes = np.ascontiguousarray(nodes, dtype=np.int32)
    vals = np.ascontiguousarray(np.broadcast_to(np.asarray(values, np.float64), nodes.shape))
    L.check(problem.lib.flgpu_dirichlet_apply(problem.ctx, nodes.size, L.ptr(nodes), L.ptr(vals),
                                              -1.0 if dt is None else float(dt)), "flgpu_dirichlet_apply")


def petsc_solve(problem: CGProblem, solver: SolverOptions, x0: np.ndarray | None = None):
    """petsc_solve(x, matrix, rhs) on the device-resident system; returns
    (x, SolveInfo).  info.reason > 0 means converged (KSPConvergedReason)."""
    x = np.zeros(problem.mesh.nnodes) if x0 is None else np.ascontiguousarray(x0, dtype=np.float64).copy()
    info = L.SolveInfo()
    sp = solver.params()
    L.check(problem.lib.flgpu_solve(problem.ctx, C.byref(sp), L.ptr(x), C.byref(info)), "flgpu_solve")
    return x, info


def petsc_system(problem: CGProblem):
    """The system petsc_solve hands to PETSc, from the device-resident one
    (femtools/Petsc_Tools.F90:891-1074, femtools/Solvers.F90:1072-1099):
    inactive (ghost) rows keep only the diagonal ghost_pivot
    sqrt((maxdiag + mindiag) maxdiag / 2) over every row's |diagonal|, their
    columns are dropped from the other rows and lifted into the rhs, and the
    ghost rows' rhs stays 0.  Returns 1-based (findrm, colm, val, b)."""
    A, rhs, inact = problem.matrix()
    sp = problem.sparsity
    n = sp.findrm.size - 1
    row = np.repeat(np.arange(n), np.diff(sp.findrm))
    col = sp.colm.astype(np.int64) - 1
    ghost = inact.astype(bool)
    dg = np.abs(np.where(sp.centrm > 0, A[np.maximum(sp.centrm, 1) - 1], 0.0))
    pivot = np.sqrt((dg.max() + dg.min()) * dg.max() / 2.0)
    g = np.where(ghost, rhs, 0.0)
    lift = np.zeros(n)
    np.add.at(lift, row, A * g[col])
    b = np.where(ghost, 0.0, rhs - lift)
    keep = np.where(ghost[row], col == row, ~ghost[col])
    val = np.where(ghost[row] & (col == row), pivot, A)[keep]
    lens = np.bincount(row[keep], minlength=n)
    findrm = np.concatenate([[1], 1 + np.cumsum(lens)]).astype(np.int32)
    return findrm, (col[keep] + 1).astype(np.int32), np.ascontiguousarray(val), np.ascontiguousarray(b)


KSP_REASONS = {-1: "Undefined", -2: "KSP_DIVERGED_NULL", -3: "KSP_DIVERGED_ITS", -4: "KSP_DIVERGED_DTOL",
               -5: "KSP_DIVERGED_BREAKDOWN", -6: "KSP_DIVERGED_BREAKDOWN_BICG", -7: "KSP_DIVERGED_NONSYMMETRIC",
               -8: "KSP_DIVERGED_INDEFINITE_PC", -9: "KSP_DIVERGED_NAN", -10: "KSP_DIVERGED_INDEFINITE_MAT"}
_matrixdumped = False


def convergence_check(problem: CGProblem, info, name: str = "", x0: np.ndarray | None = None,
                      filename: str = "matrixdump", ignore_all_solver_failures: bool = False,
                      no_matrixdump: bool = False) -> bool:
    """ConvergenceCheck (femtools/Solvers.F90:1369-1479) after a solve: a
    failure (reason <= 0) is reported and, the first time in the process,
    the system is written as a matrix dump (DumpMatrixEquation) with x0 (zero
    when None, start_from_zero).  Returns sig_int: whether the run should
    dump and finish (every failure unless ignore_all_solver_failures)."""
    global _matrixdumped
    import sys
    if info.reason > 0:
        return False
    print(f"WARNING: Failed to converge.\nPETSc did not converge for matrix solve of: {name}\n"
          f"Reason for non-convergence: {KSP_REASONS.get(info.reason, info.reason)}\n"
          f"Number of iterations: {info.its}", file=sys.stderr, flush=True)
    if no_matrixdump:
        _matrixdumped = True
    if not _matrixdumped:
        from .readnsolve import matrixdump_write
        findrm, colm, val, b = petsc_system(problem)
        matrixdump_write(filename, findrm, colm, val, b, np.zeros(b.size) if x0 is None else x0)
        _matrixdumped = True
    return not ignore_all_solver_failures


def solve_field_equation_cg(problem: CGProblem, opts: CGOptions, solver: SolverOptions, dt: float,
                            bc_nodes: np.ndarray | None = None, bc_values=0.0, bcs: SurfaceBCs | None = None,
                            name: str = "Tracer", matrixdump: str = "matrixdump",
                            ignore_all_solver_failures: bool = False):
    """One solve_field_equation_cg step: assemble (elements, then faces
    with bcs), strong Dirichlet, solve for the rate of change, check the
    solve as petsc_solve does (ConvergenceCheck, Solvers.F90:1369-1479: a
    failure is reported, the first one writes the matrix dump and sets
    problem.sig_int, the "dump and finish" signal of Signal_Vars) and update
    T (T += dt*dT, which the reference applies whatever the outcome,
    Advection_Diffusion_CG.F90:1389-1407).  Returns (T_new,
    iterations_taken, info)."""
    assemble_advection_diffusion_cg(problem, opts, dt, bcs)
    if bc_nodes is not None and bc_nodes.size:
        apply_dirichlet_conditions(problem, bc_nodes, bc_values, dt)
    dT, info = petsc_solve(problem, solver)
    if convergence_check(problem, info, name, filename=matrixdump,
                         ignore_all_solver_failures=ignore_all_solver_failures):
        problem.sig_int = True
    T_new = problem.T + dt * dT
    return T_new, info.its, info


# ---------------------------------------------------------------------------
# The pressure path of the momentum solve (assemble/Momentum_CG.F90,
# assemble/Assemble_CMC.F90) on the velocity mesh bound to a CGProblem
# ---------------------------------------------------------------------------

@dataclass
class PressureMesh:
    """The pressure mesh of a momentum solve: connectivity (1-based,
    p_loc per element), node count, element tables on the velocity's
    quadrature, and ct_m's sparsity (rows: pressure nodes, columns: velocity
    nodes, 1-based)."""
    ndglno: np.ndarray
    loc: int
    nnodes: int
    element: Element
    ct_findrm: np.ndarray
    ct_colm: np.ndarray


PLAN_INFO_KEYS = ("staged", "slabs", "max_rows", "max_entries", "max_slots", "max_nodes", "node_visits",
                  "element_rows", "lds_bytes_fp64", "workgroups_per_cu")


def assembly_plan_info(problem) -> dict:
    """The blocked scheme's plan for the bound mesh (flgpu_assembly_plan_info):
    slabs, their largest extents, staged nodes in all (node_visits /
    nodes = how often a node is staged), the LDS of a slab and the host
    seconds of the plan build."""
    info = np.zeros(len(PLAN_INFO_KEYS), np.int64)
    secs = C.c_double(0)
    L.check(problem.lib.flgpu_assembly_plan_info(problem.ctx, L.ptr(info), len(PLAN_INFO_KEYS), C.byref(secs)),
            "flgpu_assembly_plan_info")
    out = {k: int(v) for k, v in zip(PLAN_INFO_KEYS, info)}
    out["build_seconds"] = round(secs.value, 4)
    return out


def vertex_pressure_mesh(problem: CGProblem) -> PressureMesh:
    """The P1 pressure mesh on the vertices of the problem's mesh (P1-P1,
    or Taylor-Hood P2-P1 for a quadratic velocity), with ct_m's sparsity:
    row i holds every velocity node of every element containing pressure
    node i (make_sparsity(p_mesh, u_mesh), femtools/Sparsity_Patterns.F90:
    47-85)."""
    mesh = problem.mesh
    pl = mesh.dim + 1
    pe = np.asarray(mesh.ndglno_x, np.int64).reshape(-1, pl) - 1
    ue = np.asarray(mesh.ndglno, np.int64).reshape(-1, mesh.loc) - 1
    rows = np.repeat(pe, mesh.loc, axis=1).ravel()
    cols = np.tile(ue, (1, pl)).ravel()
    key = np.unique(rows * mesh.nnodes + cols)
    r, c = key // mesh.nnodes, key % mesh.nnodes
    findrm = np.concatenate([[0], np.cumsum(np.bincount(r, minlength=mesh.nx))]).astype(np.int32) + 1
    return PressureMesh(np.asarray(mesh.ndglno_x, np.int32), pl, mesh.nx,
                        make_element(mesh.dim, 1, problem.element.quad_degree), findrm, (c + 1).astype(np.int32))


def bind_pressure_mesh(problem: CGProblem, pm: PressureMesh) -> None:
    el = pm.element
    L.check(problem.lib.flgpu_pressure_bind(problem.ctx, pm.loc, pm.nnodes, L.ptr(pm.ndglno),
                                            L.ptr(el.fortran(el.n)), L.ptr(el.fortran(el.dn)), pm.nnodes,
                                            L.ptr(pm.ct_findrm), L.ptr(pm.ct_colm)), "flgpu_pressure_bind")


def construct_momentum_ct(problem: CGProblem, pm: PressureMesh, integrate_continuity_by_parts: bool = False,
                          density=1.0):
    """ct_m and the lumped velocity mass of construct_momentum_cg
    (Momentum_CG.F90:1377-1403, :1492-1600) for a bound pressure mesh:
    returns (ct_val (dim, nnz), masslump (velocity nodes))."""
    d = problem.mesh.dim
    rho = np.atleast_1d(np.ascontiguousarray(density, dtype=np.float64))
    ct = np.zeros((d, pm.ct_colm.size))
    ml = np.zeros(problem.mesh.nnodes)
    L.check(problem.lib.flgpu_ct_assemble(problem.ctx, int(integrate_continuity_by_parts), int(rho.size == 1),
                                          L.ptr(rho), L.ptr(ct), L.ptr(ml)), "flgpu_ct_assemble")
    return ct, ml


def cmc_sparsity(pm: PressureMesh, u_nodes: int):
    """The pattern of ct_m ct_m^T (get_csr_sparsity_secondorder): row i holds
    every pressure node whose ct_m row shares a velocity node with row i."""
    import scipy.sparse as sp
    c = sp.csr_matrix((np.ones(pm.ct_colm.size), pm.ct_colm - 1, pm.ct_findrm - 1), shape=(pm.nnodes, u_nodes))
    m = (c @ c.T).tocsr()
    m.sort_indices()
    return (m.indptr + 1).astype(np.int32), (m.indices + 1).astype(np.int32)


def assemble_masslumped_cmc(problem: CGProblem, cmc_findrm: np.ndarray, cmc_colm: np.ndarray,
                            inverse_masslump: np.ndarray | None = None) -> np.ndarray:
    """C_P^T M_l^-1 C (Assemble_CMC.F90:119-135) from the last
    construct_momentum_ct, on the given pattern; inverse_masslump (dim,
    velocity nodes) or None for 1/masslump in every component."""
    im = None if inverse_masslump is None else np.ascontiguousarray(
        np.asarray(inverse_masslump, np.float64).reshape(problem.mesh.dim, -1).T)
    out = np.zeros(cmc_colm.size)
    L.check(problem.lib.flgpu_cmc_assemble(problem.ctx, None if im is None else L.ptr(im), cmc_findrm.size - 1,
                                           L.ptr(cmc_findrm), L.ptr(cmc_colm), L.ptr(out)), "flgpu_cmc_assemble")
    return out


def problem_from_config(cfg, device: int = 0, **kw) -> CGProblem:
    """Bind a meshgen.Config and upload its fields."""
    pr = CGProblem(cfg.mesh, quad_degree=cfg.quad_degree, device=device, **kw)
    pr.set_fields(T=cfg.T, velocity=cfg.U, diffusivity=cfg.kappa, source=cfg.source,
                  absorption=cfg.absorption)
    return pr


def options_from_config(cfg, precision: str = "fp64", scheme: str = "auto") -> CGOptions:
    return CGOptions(theta=cfg.theta, beta=cfg.beta, exclude_mass_terms=not cfg.have_mass,
                     exclude_advection_terms=not cfg.have_advection, have_diffusivity=cfg.kappa is not None,
                     have_source=cfg.source is not None, have_absorption=cfg.absorption is not None,
                     precision=precision, scheme=scheme)
