!! Example Fortran caller of the gfx950 library: one solve_field_equation_cg
!! step (assemble/Advection_Diffusion_CG.F90:127-174) for the rotating
!! Gaussian bump on the unit square, written the way Fluidity hands its
!! arrays over (1-based ndglno, positions%val(dim,nodes), element_type
!! tables, csr_sparsity, colour sets).
!!
!! usage: flgpu_driver [n [faces]]   (n x n cells, two P1 triangles per cell)
!! With "faces": sparsity and colouring built on the GPU
!! (flgpu_sparsity_build, flgpu_colouring_build) and the face loop of
!! assemble_advection_diffusion_cg (:606-638) on the x sides -- Neumann flux
!! 1 + y/2 on x = 0, Robin g = 0.5, h = 2 + y on x = 1 -- with strong
!! Dirichlet 0 on the y sides only.
!! prints: nodes nnz colours its reason relres sum(T) max(T)
program flgpu_driver
  use iso_c_binding
  use flgpu
  implicit none
  integer(c_int32_t) :: n, m, nn, ne, i, j, c, e, k, g, ncol, its, reason, nbc
  integer(c_int32_t), allocatable, target :: ndglno(:), findrm(:), colm(:), centrm(:), colour(:)
  integer(c_int32_t), allocatable :: col_ptr(:), elems(:), fill(:), bnodes(:)
  real(c_double), allocatable, target :: X(:,:), T(:), U(:,:), kap(:,:,:), dT(:), bvals(:)
  real(c_double) :: shn(3,3), shdn(3,3,2), w(3), l(3,3), relres, h, x0, y0
  integer(c_int64_t) :: nnz
  type(c_ptr) :: ctx
  type(flgpu_advdif_params) :: p
  type(flgpu_solver_params) :: sp
  character(len=32) :: arg
  logical :: faces
  integer(c_int32_t), allocatable :: fnodes(:), fele(:), ftype(:)
  real(c_double), allocatable, target :: fv1(:), fv2(:)
  real(c_double) :: fn(2,2), fw(2), fxdn(2,2), gq(2)
  integer(c_int32_t) :: nf

  n = 16
  if (command_argument_count() >= 1) then
     call get_command_argument(1, arg)
     read(arg, *) n
  end if
  faces = .false.
  if (command_argument_count() >= 2) then
     call get_command_argument(2, arg)
     faces = trim(arg) == "faces"
  end if
  m = n + 1
  nn = m * m
  ne = 2 * n * n
  h = 1.0d0 / n

  ! --- mesh: node (i,j) -> i + m*j + 1; cell diagonal (i,j)-(i+1,j+1)
  allocate(X(2, nn), ndglno(3 * ne))
  do j = 0, n
     do i = 0, n
        X(1, i + m * j + 1) = i * h
        X(2, i + m * j + 1) = j * h
     end do
  end do
  e = 0
  do j = 0, n - 1
     do i = 0, n - 1
        k = i + m * j + 1
        ndglno(3 * e + 1:3 * e + 3) = [k, k + 1, k + m + 1]
        e = e + 1
        ndglno(3 * e + 1:3 * e + 3) = [k, k + m + 1, k + m]
        e = e + 1
     end do
  end do

  ! --- element_type of P1 triangles with the 3-point degree-2 rule
  !     (femtools/Quadrature.F90:935-949; n = barycentric coordinates)
  l(:, 1) = [1.0d0/6, 1.0d0/6, 1.0d0 - 2.0d0/6]
  l(:, 2) = [1.0d0/6, 1.0d0 - 2.0d0/6, 1.0d0/6]
  l(:, 3) = [1.0d0 - 2.0d0/6, 1.0d0/6, 1.0d0/6]
  w = 1.0d0 / 6
  do g = 1, 3
     shn(:, g) = l(:, g)
     shdn(1, g, :) = [1.0d0, 0.0d0]
     shdn(2, g, :) = [0.0d0, 1.0d0]
     shdn(3, g, :) = [-1.0d0, -1.0d0]
  end do

  ! --- fields: bump, solid-body rotation, kappa = 1e-3
  allocate(T(nn), U(2, nn), kap(2, 2, nn), dT(nn))
  do k = 1, nn
     x0 = X(1, k) - 0.3d0
     y0 = X(2, k) - 0.5d0
     T(k) = exp(-(x0 * x0 + y0 * y0) / (2 * 0.05d0**2))
     U(:, k) = [-(X(2, k) - 0.5d0), X(1, k) - 0.5d0]
     kap(:, :, k) = reshape([1.0d-3, 0.0d0, 0.0d0, 1.0d-3], [2, 2])
  end do

  ! --- sparsity (make_sparsity) and COLOURING_CG1 sets
  allocate(findrm(nn + 1), centrm(nn))
  call flgpu_check(flgpu_make_sparsity(nn, ne, 3, ndglno, findrm, c_null_ptr, centrm, nnz), "make_sparsity")
  allocate(colm(nnz))
  call flgpu_check(flgpu_make_sparsity(nn, ne, 3, ndglno, findrm, c_loc(colm), centrm, nnz), "make_sparsity")
  allocate(colour(ne))
  call flgpu_check(flgpu_colour_mesh(nn, ne, 3, ndglno, colour, ncol), "colour_mesh")
  allocate(col_ptr(ncol + 1), elems(ne), fill(ncol))
  col_ptr = 0
  do e = 1, ne
     col_ptr(colour(e) + 1) = col_ptr(colour(e) + 1) + 1
  end do
  col_ptr(1) = 1
  do c = 1, ncol
     col_ptr(c + 1) = col_ptr(c + 1) + col_ptr(c)
  end do
  fill = col_ptr(1:ncol)
  do e = 1, ne                              ! ascending inside each colour
     elems(fill(colour(e))) = e
     fill(colour(e)) = fill(colour(e)) + 1
  end do

  ! --- bind (once per mesh) and one solve_field_equation_cg step
  call flgpu_check(flgpu_init(ctx, 0_c_int32_t), "init")
  call flgpu_check(flgpu_mesh_bind(ctx, 1_c_int64_t, 0_c_int64_t, 2, 3, ne, nn, ndglno, X), "mesh_bind")
  call flgpu_check(flgpu_field_mesh_bind(ctx, 3, 3, nn, ndglno, shn, shdn, w, shdn), "field_mesh_bind")
  if (faces) then
     ! make_sparsity and a colouring on the GPU instead of the host ones
     call flgpu_check(flgpu_sparsity_build(ctx, nnz), "sparsity_build")
     deallocate(colm)
     allocate(colm(nnz))
     call flgpu_check(flgpu_sparsity_get(ctx, findrm, colm, centrm), "sparsity_get")
     call flgpu_check(flgpu_colouring_build(ctx, ncol), "colouring_build")
  else
     call flgpu_check(flgpu_sparsity_bind(ctx, nn, nn, findrm, colm, centrm), "sparsity_bind")
     call flgpu_check(flgpu_colouring_bind(ctx, ncol, col_ptr, elems), "colouring_bind")
  end if

  p = flgpu_advdif_params(dt=0.01d0, theta=0.5d0, beta=0.0d0, have_mass=1, lump_mass=0, &
       have_advection=1, integrate_by_parts=0, have_diffusivity=1, isotropic_diffusivity=1, &
       have_source=0, have_absorption=0, precision=FLGPU_FP64, scheme=FLGPU_ASM_AUTO, &
       stabilisation=FLGPU_STAB_NONE, nu_bar_scheme=FLGPU_NU_BAR_OPTIMAL, nu_bar_scale=0.5d0, &
       density_theta=1.0d0, equation_type=FLGPU_EQ_ADVECTION_DIFFUSION, compressible=0, multiphase=0, &
       move_mesh=0)
  call flgpu_check(flgpu_advdif_assemble(ctx, p, c_loc(T), c_loc(U), c_loc(kap), c_null_ptr, c_null_ptr, &
       c_null_ptr, c_null_ptr), "advdif_assemble")

  if (faces) then
     ! surface elements on x = 0 and x = 1: face_global_nodes in the order of
     ! the element's boundary numbering (the edge off the element's vertex
     ! 2 for x = 0, off vertex 3 for x = 1 ... here simply bottom to top),
     ! the 2-point Gauss rule of degree 2 (Quadrature.F90:1166-1180)
     nf = 2 * n
     allocate(fnodes(2 * nf), fele(nf), ftype(nf), fv1(2 * nf), fv2(2 * nf))
     do j = 0, n - 1
        ! x = 0: element 2 of cell (0, j) = (k, k+m+1, k+m); its edge off vertex 2 is (k+m, k)
        k = m * j + 1
        fnodes(2 * j + 1:2 * j + 2) = [k + m, k]
        fele(j + 1) = 2 * (n * j) + 2
        ftype(j + 1) = FLGPU_BC_NEUMANN
        fv1(2 * j + 1:2 * j + 2) = [1.0d0 + 0.5d0 * X(2, k + m), 1.0d0 + 0.5d0 * X(2, k)]
        fv2(2 * j + 1:2 * j + 2) = 0.0d0
        ! x = 1: element 1 of cell (n-1, j) = (k', k'+1, k'+m+1); edge off vertex 1 is (k'+1, k'+m+1)
        k = n - 1 + m * j + 1
        fnodes(2 * (n + j) + 1:2 * (n + j) + 2) = [k + 1, k + m + 1]
        fele(n + j + 1) = 2 * (n - 1 + n * j) + 1
        ftype(n + j + 1) = FLGPU_BC_ROBIN
        fv1(2 * (n + j) + 1:2 * (n + j) + 2) = 0.5d0
        fv2(2 * (n + j) + 1:2 * (n + j) + 2) = [2.0d0 + X(2, k + 1), 2.0d0 + X(2, k + m + 1)]
     end do
     gq = [0.788675134594813d0, 1.0d0 - 0.788675134594813d0]
     fn(1, :) = gq
     fn(2, :) = 1.0d0 - gq
     fw = 0.5d0
     fxdn(1, :) = 1.0d0
     fxdn(2, :) = -1.0d0
     call flgpu_check(flgpu_surface_bind(ctx, nf, 2, 2, fnodes, fnodes, fele, fn, fw, fxdn), "surface_bind")
     call flgpu_check(flgpu_advdif_surface(ctx, p, ftype, c_loc(fv1), c_loc(fv2)), "advdif_surface")
     nbc = 2 * (n + 1)
  else
     nbc = 4 * n
  end if
  ! strong Dirichlet 0 on the boundary (the y sides with faces), rate-of-change form
  allocate(bnodes(nbc), bvals(nbc))
  k = 0
  do j = 0, n
     do i = 0, n
        if (j == 0 .or. j == n .or. (.not. faces .and. (i == 0 .or. i == n))) then
           k = k + 1
           bnodes(k) = i + m * j + 1
        end if
     end do
  end do
  bvals = 0.0d0
  call flgpu_check(flgpu_dirichlet_apply(ctx, nbc, bnodes, bvals, p%dt), "dirichlet_apply")

  sp = flgpu_solver_params(ksp=FLGPU_KSP_GMRES, pc=FLGPU_PC_JACOBI, rtol=1.0d-10, atol=0.0d0, &
       max_its=10000, restart=30, precision=FLGPU_FP64, start_from_zero=1, check_every=16, &
       stall_factor=0.0d0, stall_window=0, norm_type=FLGPU_NORM_PRECONDITIONED)
  dT = 0.0d0
  call flgpu_check(flgpu_petsc_solve(ctx, sp, c_null_ptr, c_null_ptr, c_null_ptr, dT, its, reason, relres), &
       "petsc_solve")
  T = T + p%dt * dT                          ! apply_advection_diffusion_cg_change

  write(*, '(i0,1x,i0,1x,i0,1x,i0,1x,i0,1x,es24.16,1x,es24.16,1x,es24.16)') &
       nn, nnz, ncol, its, reason, relres, sum(T), maxval(T)
  call flgpu_check(flgpu_finalize(ctx), "finalize")
  ! ConvergenceCheck (Solvers.F90:1369-1479): a failed solve is reported
  ! and the run ends ("dump and finish") with a nonzero status
  if (reason <= 0) then
     write(0, '(a,i0,a,i0)') 'WARNING: Failed to converge. PETSc reason ', reason, ', iterations ', its
     stop 3
  end if
end program flgpu_driver
