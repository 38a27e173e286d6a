!! ISO_C_BINDING interface of the gfx950 CG advection-diffusion library
!! (include/flgpu.h).  Fortran callers pass the arrays of Fluidity's own
!! types straight through: mesh%ndglno, positions%val, element_type%n/%dn,
!! quadrature%weight, csr_sparsity%findrm/%colm/%centrm, field%val.
!!
!! Cut points it replaces:
!!   assemble/Advection_Diffusion_CG.F90:565-598  (element loop)
!!     -> flgpu_advdif_assemble
!!   assemble/Advection_Diffusion_CG.F90:606-638  (face loop)
!!     -> flgpu_surface_bind + flgpu_advdif_surface
!!   assemble/Advection_Diffusion_CG.F90:1389      (petsc_solve)
!!     -> flgpu_petsc_solve
!! Errors: a nonzero return is turned into FLAbort by flgpu_check, as
!! csr_addto does for an insertion outside the pattern
!! (femtools/Sparse_Tools.F90:2645-2647).
module flgpu
  use iso_c_binding
  implicit none
  private

  integer(c_int32_t), parameter, public :: FLGPU_OK = 0
  integer(c_int32_t), parameter, public :: FLGPU_FP64 = 0, FLGPU_MIXED = 1
  integer(c_int32_t), parameter, public :: FLGPU_ASM_AUTO = 0, FLGPU_ASM_COLOUR = 1, FLGPU_ASM_BLOCKED = 2
  integer(c_int32_t), parameter, public :: FLGPU_KSP_CG = 0, FLGPU_KSP_GMRES = 1, FLGPU_KSP_FCG = 2, &
                                            FLGPU_KSP_PIPECG = 3
  integer(c_int32_t), parameter, public :: FLGPU_PC_NONE = 0, FLGPU_PC_JACOBI = 1, FLGPU_PC_LEARNED = 2, &
                                            FLGPU_PC_SOR = 3, FLGPU_PC_FSAI = 4
  integer(c_int32_t), parameter, public :: FLGPU_NORM_PRECONDITIONED = 0, FLGPU_NORM_UNPRECONDITIONED = 1
  integer(c_int32_t), parameter, public :: FLGPU_FIELD_T = 0, FLGPU_FIELD_VELOCITY = 1, &
       FLGPU_FIELD_DIFFUSIVITY = 2, FLGPU_FIELD_SOURCE = 3, FLGPU_FIELD_ABSORPTION = 4, &
       FLGPU_FIELD_DENSITY = 5, FLGPU_FIELD_OLD_DENSITY = 6, FLGPU_FIELD_PRESSURE = 7, &
       FLGPU_FIELD_VOLUME_FRACTION = 8, &
       FLGPU_EQ_ADVECTION_DIFFUSION = 0, FLGPU_EQ_INTERNAL_ENERGY = 1, FLGPU_EQ_KEPSILON = 2
  !! stabilisation (Advection_Diffusion_CG.F90:71) and nu_bar schemes (Upwind_Stabilisation.F90:18-19)
  integer(c_int32_t), parameter, public :: FLGPU_STAB_NONE = 0, FLGPU_STAB_STREAMLINE_UPWIND = 1, &
       FLGPU_STAB_SUPG = 2
  integer(c_int32_t), parameter, public :: FLGPU_NU_BAR_OPTIMAL = 1, FLGPU_NU_BAR_DOUBLY_ASYMPTOTIC = 2, &
       FLGPU_NU_BAR_CRITICAL_RULE = 3, FLGPU_NU_BAR_UNITY = 4
  !! boundary-condition types of the face loop (Advection_Diffusion_CG.F90:74-75)
  integer(c_int32_t), parameter, public :: FLGPU_BC_NONE = 0, FLGPU_BC_NEUMANN = 1, &
       FLGPU_BC_WEAKDIRICHLET = 2, FLGPU_BC_INTERNAL = 3, FLGPU_BC_ROBIN = 4

  !! flgpu_advdif_params: the options assemble_advection_diffusion_cg reads
  !! (Advection_Diffusion_CG.F90:282-556); logicals as 0/1 integers.
  type, bind(c), public :: flgpu_advdif_params
     real(c_double) :: dt, theta, beta
     integer(c_int32_t) :: have_mass, lump_mass
     integer(c_int32_t) :: have_advection, integrate_by_parts
     integer(c_int32_t) :: have_diffusivity, isotropic_diffusivity
     integer(c_int32_t) :: have_source, have_absorption
     integer(c_int32_t) :: precision, scheme
     integer(c_int32_t) :: stabilisation, nu_bar_scheme
     real(c_double) :: nu_bar_scale
     !! equation type (FLGPU_EQ_*), the internal energy's density theta,
     !! its p div(u) term and PhaseVolumeFraction factor (0/1)
     real(c_double) :: density_theta
     integer(c_int32_t) :: equation_type, compressible, multiphase
     !! mesh movement (0/1; positions and grid velocity by flgpu_mesh_motion_set)
     integer(c_int32_t) :: move_mesh
  end type flgpu_advdif_params

  !! flgpu_solver_params: the solver options of schemas/solvers.rnc
  type, bind(c), public :: flgpu_solver_params
     integer(c_int32_t) :: ksp, pc
     real(c_double) :: rtol, atol
     integer(c_int32_t) :: max_its, restart
     integer(c_int32_t) :: precision, start_from_zero
     integer(c_int32_t) :: check_every
     real(c_double) :: stall_factor
     integer(c_int32_t) :: stall_window, norm_type
  end type flgpu_solver_params

  public :: flgpu_init, flgpu_finalize, flgpu_mesh_bind, flgpu_field_mesh_bind, &
       flgpu_sparsity_bind, flgpu_colouring_bind, flgpu_make_sparsity, flgpu_colour_mesh, &
       flgpu_sparsity_build, flgpu_sparsity_get, flgpu_colouring_build, flgpu_colouring_get, &
       flgpu_field_set, flgpu_mesh_motion_set, flgpu_advdif_assemble, flgpu_rhs_add_field, flgpu_surface_bind, flgpu_advdif_surface, &
       flgpu_dirichlet_apply, flgpu_matrix_get, &
       flgpu_petsc_solve, flgpu_check, flgpu_pc_from_name

  interface
     integer(c_int) function flgpu_init(ctx, device) bind(c, name="flgpu_init")
       import :: c_int, c_ptr, c_int32_t
       type(c_ptr), intent(out) :: ctx
       integer(c_int32_t), value :: device
     end function flgpu_init

     integer(c_int) function flgpu_finalize(ctx) bind(c, name="flgpu_finalize")
       import :: c_int, c_ptr
       type(c_ptr), value :: ctx
     end function flgpu_finalize

     function flgpu_last_error_c() bind(c, name="flgpu_last_error")
       import :: c_ptr
       type(c_ptr) :: flgpu_last_error_c
     end function flgpu_last_error_c

     integer(c_int) function flgpu_mesh_bind(ctx, key_id, key_event, dim, loc_x, nele, nx, ndglno_x, X) &
          bind(c, name="flgpu_mesh_bind")
       import :: c_int, c_ptr, c_int32_t, c_int64_t, c_double
       type(c_ptr), value :: ctx
       integer(c_int64_t), value :: key_id, key_event
       integer(c_int32_t), value :: dim, loc_x, nele, nx
       integer(c_int32_t), intent(in) :: ndglno_x(*)
       real(c_double), intent(in) :: X(*)
     end function flgpu_mesh_bind

     integer(c_int) function flgpu_field_mesh_bind(ctx, loc, ngi, nnodes, ndglno, n, dn, w, x_dn) &
          bind(c, name="flgpu_field_mesh_bind")
       import :: c_int, c_ptr, c_int32_t, c_double
       type(c_ptr), value :: ctx
       integer(c_int32_t), value :: loc, ngi, nnodes
       integer(c_int32_t), intent(in) :: ndglno(*)
       real(c_double), intent(in) :: n(*), dn(*), w(*), x_dn(*)
     end function flgpu_field_mesh_bind

     integer(c_int) function flgpu_sparsity_bind(ctx, rows, cols, findrm, colm, centrm) &
          bind(c, name="flgpu_sparsity_bind")
       import :: c_int, c_ptr, c_int32_t
       type(c_ptr), value :: ctx
       integer(c_int32_t), value :: rows, cols
       integer(c_int32_t), intent(in) :: findrm(*), colm(*), centrm(*)
     end function flgpu_sparsity_bind

     integer(c_int) function flgpu_colouring_bind(ctx, ncol, col_ptr, elems) bind(c, name="flgpu_colouring_bind")
       import :: c_int, c_ptr, c_int32_t
       type(c_ptr), value :: ctx
       integer(c_int32_t), value :: ncol
       integer(c_int32_t), intent(in) :: col_ptr(*), elems(*)
     end function flgpu_colouring_bind

     integer(c_int) function flgpu_make_sparsity(nnodes, nele, loc, ndglno, findrm, colm, centrm, nnz) &
          bind(c, name="flgpu_make_sparsity")
       import :: c_int, c_int32_t, c_int64_t, c_ptr
       integer(c_int32_t), value :: nnodes, nele, loc
       integer(c_int32_t), intent(in) :: ndglno(*)
       integer(c_int32_t), intent(out) :: findrm(*), centrm(*)
       type(c_ptr), value :: colm
       integer(c_int64_t), intent(out) :: nnz
     end function flgpu_make_sparsity

     !! make_sparsity and an element colouring of the bound field mesh, on the GPU
     integer(c_int) function flgpu_sparsity_build(ctx, nnz) bind(c, name="flgpu_sparsity_build")
       import :: c_int, c_ptr, c_int64_t
       type(c_ptr), value :: ctx
       integer(c_int64_t), intent(out) :: nnz
     end function flgpu_sparsity_build

     integer(c_int) function flgpu_sparsity_get(ctx, findrm, colm, centrm) bind(c, name="flgpu_sparsity_get")
       import :: c_int, c_ptr, c_int32_t
       type(c_ptr), value :: ctx
       integer(c_int32_t), intent(out) :: findrm(*), colm(*), centrm(*)
     end function flgpu_sparsity_get

     integer(c_int) function flgpu_colouring_build(ctx, ncol) bind(c, name="flgpu_colouring_build")
       import :: c_int, c_ptr, c_int32_t
       type(c_ptr), value :: ctx
       integer(c_int32_t), intent(out) :: ncol
     end function flgpu_colouring_build

     integer(c_int) function flgpu_colouring_get(ctx, col_ptr, elems) bind(c, name="flgpu_colouring_get")
       import :: c_int, c_ptr, c_int32_t
       type(c_ptr), value :: ctx
       integer(c_int32_t), intent(out) :: col_ptr(*), elems(*)
     end function flgpu_colouring_get

     integer(c_int) function flgpu_colour_mesh(nnodes, nele, loc, ndglno, colour, ncol) &
          bind(c, name="flgpu_colour_mesh")
       import :: c_int, c_int32_t
       integer(c_int32_t), value :: nnodes, nele, loc
       integer(c_int32_t), intent(in) :: ndglno(*)
       integer(c_int32_t), intent(out) :: colour(*)
       integer(c_int32_t), intent(out) :: ncol
     end function flgpu_colour_mesh

     integer(c_int) function flgpu_field_set(ctx, which, constant, values) bind(c, name="flgpu_field_set")
       import :: c_int, c_ptr, c_int32_t, c_double
       type(c_ptr), value :: ctx
       integer(c_int32_t), value :: which, constant
       real(c_double), intent(in) :: values(*)
     end function flgpu_field_set

     !! OldCoordinate, IteratedCoordinate, GridVelocity (dim x nx each)
     integer(c_int) function flgpu_mesh_motion_set(ctx, X_old, X_new, grid_velocity) &
          bind(c, name="flgpu_mesh_motion_set")
       import :: c_int, c_ptr, c_double
       type(c_ptr), value :: ctx
       real(c_double), intent(in) :: X_old(*), X_new(*), grid_velocity(*)
     end function flgpu_mesh_motion_set

     integer(c_int) function flgpu_advdif_assemble(ctx, p, T, U, kappa, src, absorp, A_val, rhs) &
          bind(c, name="flgpu_advdif_assemble")
       import :: c_int, c_ptr, flgpu_advdif_params
       type(c_ptr), value :: ctx
       type(flgpu_advdif_params), intent(in) :: p
       type(c_ptr), value :: T, U, kappa, src, absorp, A_val, rhs
     end function flgpu_advdif_assemble

     !! surface elements of T for the face loop (:606-638): face_global_nodes
     !! of t and of positions, face_ele, face_shape(t)%n, its weights and
     !! face_shape(positions)%dn
     integer(c_int) function flgpu_surface_bind(ctx, nfaces, floc, fngi, face_ndglno, face_ndglno_x, face_ele, &
          fn, fw, fx_dn) bind(c, name="flgpu_surface_bind")
       import :: c_int, c_ptr, c_int32_t, c_double
       type(c_ptr), value :: ctx
       integer(c_int32_t), value :: nfaces, floc, fngi
       integer(c_int32_t), intent(in) :: face_ndglno(*), face_ndglno_x(*), face_ele(*)
       real(c_double), intent(in) :: fn(*), fw(*), fx_dn(*)
     end function flgpu_surface_bind

     !! the face loop itself, after flgpu_advdif_assemble: t_bc_types from
     !! get_entire_boundary_condition, ele_val(t_bc, face) and
     !! ele_val(t_bc_2, face) flattened per face (c_null_ptr when unused)
     integer(c_int) function flgpu_advdif_surface(ctx, p, bc_type, bc_value, bc_value2) &
          bind(c, name="flgpu_advdif_surface")
       import :: c_int, c_ptr, c_int32_t, flgpu_advdif_params
       type(c_ptr), value :: ctx
       type(flgpu_advdif_params), intent(in) :: p
       integer(c_int32_t), intent(in) :: bc_type(*)
       type(c_ptr), value :: bc_value, bc_value2
     end function flgpu_advdif_surface

     !! rhs += field (the source's add_directly_to_rhs, :604)
     integer(c_int) function flgpu_rhs_add_field(ctx, which) bind(c, name="flgpu_rhs_add_field")
       import :: c_int, c_ptr, c_int32_t
       type(c_ptr), value :: ctx
       integer(c_int32_t), value :: which
     end function flgpu_rhs_add_field

     integer(c_int) function flgpu_dirichlet_apply(ctx, nbc, nodes, values, dt) bind(c, name="flgpu_dirichlet_apply")
       import :: c_int, c_ptr, c_int32_t, c_double
       type(c_ptr), value :: ctx
       integer(c_int32_t), value :: nbc
       integer(c_int32_t), intent(in) :: nodes(*)
       real(c_double), intent(in) :: values(*)
       real(c_double), value :: dt
     end function flgpu_dirichlet_apply

     integer(c_int) function flgpu_matrix_get(ctx, A_val, rhs, inactive) bind(c, name="flgpu_matrix_get")
       import :: c_int, c_ptr
       type(c_ptr), value :: ctx, A_val, rhs, inactive
     end function flgpu_matrix_get

     integer(c_int) function flgpu_petsc_solve(ctx, sp, A_val, inactive, rhs, x, its, reason, true_relres) &
          bind(c, name="flgpu_petsc_solve")
       import :: c_int, c_ptr, c_int32_t, c_double, flgpu_solver_params
       type(c_ptr), value :: ctx
       type(flgpu_solver_params), intent(in) :: sp
       type(c_ptr), value :: A_val, inactive, rhs
       real(c_double), intent(inout) :: x(*)
       integer(c_int32_t), intent(out) :: its, reason
       real(c_double), intent(out) :: true_relres
     end function flgpu_petsc_solve

     !! the pressure path (Momentum_CG.F90 / Assemble_CMC.F90): the pressure
     !! mesh with ct_m's sparsity, then ct_m and the lumped velocity mass,
     !! then assemble_masslumped_cmc's C_P^T M_l^-1 C
     integer(c_int) function flgpu_pressure_bind(ctx, p_loc, p_nodes, p_ndglno, p_n, p_dn, ct_rows, &
          ct_findrm, ct_colm) bind(c, name="flgpu_pressure_bind")
       import :: c_int, c_ptr, c_int32_t, c_double
       type(c_ptr), value :: ctx
       integer(c_int32_t), value :: p_loc, p_nodes, ct_rows
       integer(c_int32_t), intent(in) :: p_ndglno(*), ct_findrm(*), ct_colm(*)
       real(c_double), intent(in) :: p_n(*), p_dn(*)
     end function flgpu_pressure_bind

     integer(c_int) function flgpu_ct_assemble(ctx, integrate_by_parts, density_constant, density, ct_val, &
          masslump) bind(c, name="flgpu_ct_assemble")
       import :: c_int, c_ptr, c_int32_t, c_double
       type(c_ptr), value :: ctx
       integer(c_int32_t), value :: integrate_by_parts, density_constant
       real(c_double), intent(in) :: density(*)
       type(c_ptr), value :: ct_val, masslump
     end function flgpu_ct_assemble

     integer(c_int) function flgpu_cmc_assemble(ctx, inverse_masslump, cmc_rows, cmc_findrm, cmc_colm, cmc_val) &
          bind(c, name="flgpu_cmc_assemble")
       import :: c_int, c_ptr, c_int32_t
       type(c_ptr), value :: ctx, inverse_masslump
       integer(c_int32_t), value :: cmc_rows
       integer(c_int32_t), intent(in) :: cmc_findrm(*), cmc_colm(*)
       type(c_ptr), value :: cmc_val
     end function flgpu_cmc_assemble

     !! flgpu_solve for callers that cannot allocate device memory (its work
     !! in a fixed 64 MiB device array per GPU)
     integer(c_int) function flgpu_solve_noalloc(ctx, sp, x, info) bind(c, name="flgpu_solve_noalloc")
       import :: c_int, c_ptr, c_double, flgpu_solver_params
       type(c_ptr), value :: ctx
       type(flgpu_solver_params), intent(in) :: sp
       real(c_double), intent(inout) :: x(*)
       type(c_ptr), value :: info
     end function flgpu_solve_noalloc
  end interface

contains

  !! FLAbort-style check: print the library's message and stop.
  subroutine flgpu_check(rc, what)
    integer(c_int), intent(in) :: rc
    character(len=*), intent(in) :: what
    type(c_ptr) :: msg
    character(kind=c_char), pointer :: chars(:)
    integer :: n
    interface
       integer(c_size_t) function c_strlen(s) bind(c, name="strlen")
         import :: c_ptr, c_size_t
         type(c_ptr), value :: s
       end function c_strlen
    end interface
    if (rc == FLGPU_OK) return
    msg = flgpu_last_error_c()
    n = int(c_strlen(msg))
    call c_f_pointer(msg, chars, [n])
    write(0, '(a,a,i0,a,*(a))') trim(what), " failed (", rc, "): ", chars(1:n)
    error stop 1
  end subroutine flgpu_check

  !! The pc field of flgpu_solver_params for a preconditioner name of the
  !! solver options, as setup_pc_from_options maps it onto PETSc
  !! (femtools/Solvers.F90:1844-1854); -1 for a name the library lacks.
  !! "eisenstat" is SSOR applied with Eisenstat's trick: the same
  !! preconditioner as "sor".
  integer(c_int32_t) function flgpu_pc_from_name(name) result(pc)
    character(len=*), intent(in) :: name
    select case (trim(name))
    case ("none")
       pc = FLGPU_PC_NONE
    case ("jacobi")
       pc = FLGPU_PC_JACOBI
    case ("sor", "eisenstat")
       pc = FLGPU_PC_SOR
    case ("learned")
       pc = FLGPU_PC_LEARNED
    case ("fsai")
       pc = FLGPU_PC_FSAI
    case default
       pc = -1
    end select
  end function flgpu_pc_from_name

end module flgpu
