"""Benchmark of the CG advection-diffusion hot path on one MI355X.

Workload (default): the graded pressure-Poisson configuration, 110^3 hexes x
6 Kuhn tetrahedra = 7,986,000 P1 tets, 1,367,631 nodes, z graded 100:1
toward z = 0.5, log-normal diffusivity (contrast 1e3), f = sin sin sin,
Dirichlet 0, synthetic inputs from a splitmix64 seed.

A step = one assembly of the whole mesh (matrix + rhs) with every input
resident in HBM.  K timed steps give `value` (Melem/s).  After the steps:
the solves to a true relative residual of 1e-8 (FP64 and mixed Jacobi-PCG,
learned FCG in FP64 and mixed), the SpMV with warm and with flushed caches,
the full-size parity leg against the oracle (matrix, rhs, pattern, blocked
== colour bit for bit, solutions against the oracle's FP64 solution), the
rows beyond the element loop, the 2-D and 3-D P1 configurations, the P2
mixed-precision configuration, the held-out learned-preconditioner system,
and the oracle timed on the host cores (median of 5 runs).

Prints ONE JSON line.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

PEAK_HBM_GBS = 8000.0


def log(msg: str) -> None:
    print(f"[bench {time.strftime('%H:%M:%S')}] {msg}", file=sys.stderr, flush=True)


def asm_bytes(cfg, nnz: int) -> int:
    """Compulsory assembly traffic: conn 4*loc*E + coords 8*dim*Nx + nodal
    inputs (T, kappa if nodal, f) 8*N each + A 8*nnz + rhs 8*N."""
    m = cfg.mesh
    nodal = 1  # T
    if cfg.kappa is not None and not np.isscalar(cfg.kappa):
        nodal += 1
    if cfg.source is not None:
        nodal += 1
    if cfg.U is not None:
        nodal += m.dim
    return 4 * m.loc * m.nele + 8 * m.dim * m.nx + 8 * nodal * m.nnodes + 8 * nnz + 8 * m.nnodes


def spmv_bytes(n: int, nnz: int, value_bytes: int = 8) -> int:
    return (value_bytes + 4) * nnz + 4 * (n + 1) + (value_bytes * 2) * n


def measured_traffic(kernel_prefixes) -> int | None:
    """HBM bytes per dispatch of the first listed kernel whose name starts
    with one of kernel_prefixes, from the committed rocprofv3 counter summary
    (profiles/hbm_traffic_c5.txt: FETCH_SIZE and WRITE_SIZE per dispatch,
    scaled by a 1 GiB copy); None when the summary does not list it."""
    path = os.path.join(ROOT, "profiles", "hbm_traffic_c5.txt")
    if not os.path.exists(path):
        return None
    for line in open(path):
        f = line.split(None, 4)
        if len(f) == 5 and f[0].isdigit() and f[4].startswith(tuple(kernel_prefixes)):
            return int(f[3])
    return None


def c3_mixed_report(log_each) -> dict:
    """The P2 advection-diffusion configuration (40^3 x 6 P2 tets, 11-point
    rule): FP32 element arithmetic against FP64 (A and b, norm-wise relative
    error), and the mixed-precision GMRES solution against the FP64 one, both
    to a 1e-8 true relative residual.  The FP64 GPU assembly is itself equal
    to the oracle to 1e-12 (tests/test_gpu_parity.py)."""
    from fluidity_amd import cg, meshgen as M

    cfg = M.config("c3", seed=1)
    pr = cg.problem_from_config(cfg)
    out = {"elements": cfg.mesh.nele, "nodes": cfg.mesh.nnodes, "nnz": pr.sparsity.nnz}
    mats = {}
    for prec in ("fp64", "fp32"):
        t0 = time.perf_counter()
        cg.assemble_advection_diffusion_cg(pr, cg.options_from_config(cfg, precision=prec), cfg.dt)
        mats[prec] = pr.matrix()
        out[f"assembly_{prec}_seconds"] = round(time.perf_counter() - t0, 4)

    def rel(a, b):
        return float(np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300))

    out["A_rel_err_fp32_vs_fp64"] = rel(mats["fp32"][0], mats["fp64"][0])
    out["b_rel_err_fp32_vs_fp64"] = rel(mats["fp32"][1], mats["fp64"][1])
    # device-resident assembly time (HIP events over 5 launches after a warm-up)
    from fluidity_amd import lib as L
    import ctypes as C
    lib = L.load()
    for prec in ("fp64", "fp32"):
        p = cg.options_from_config(cfg, precision=prec).params(cfg.dt)
        L.check(lib.flgpu_advdif_assemble_resident(pr.ctx, C.byref(p)), "assemble")
        ms = C.c_double(0)
        L.check(lib.flgpu_time_assemble(pr.ctx, C.byref(p), 5, C.byref(ms)), "time_assemble")
        out[f"assembly_{prec}_ms_resident"] = round(ms.value, 4)
        out[f"assembly_{prec}_Melem_per_s"] = round(cfg.mesh.nele / ms.value / 1e3, 2)
    xs = {}
    for label, prec in (("fp64", "fp64"), ("mixed", "mixed")):
        so = cg.SolverOptions("gmres", "jacobi", relative_error=1e-8, restart=30, max_iterations=20000,
                              precision=prec)
        cg.assemble_advection_diffusion_cg(pr, cg.options_from_config(cfg, precision="fp64"), cfg.dt)
        cg.apply_dirichlet_conditions(pr, cfg.mesh.boundary, 0.0, cfg.dt)
        t0 = time.perf_counter()
        x, info = cg.petsc_solve(pr, so)
        xs[label] = x
        out[f"gmres_{label}"] = {"seconds": round(time.perf_counter() - t0, 4), "its": info.its,
                                 "reason": info.reason, "true_relres": info.true_relres}
    out["x_rel_err_mixed_vs_fp64"] = rel(xs["mixed"], xs["fp64"])
    pr.close()
    log_each(f"c3 mixed: A err {out['A_rel_err_fp32_vs_fp64']:.2e}, b err {out['b_rel_err_fp32_vs_fp64']:.2e}, "
             f"x err {out['x_rel_err_mixed_vs_fp64']:.2e}; assembly fp64 {out['assembly_fp64_ms_resident']:.3f} ms, "
             f"fp32 {out['assembly_fp32_ms_resident']:.3f} ms")
    return out


def c4_heldout_report(weights, log_each) -> dict:
    """The learned preconditioner on a held-out system: the 64^3 x 6 cube
    (1,572,864 P1 tets) with a log-normal diffusivity of contrast 1e4 from a
    seed the training set does not use, against FP64 Jacobi-PCG, to a 1e-8
    true relative residual: iterations and seconds."""
    from fluidity_amd import cg, learned, meshgen as M

    cfg = M.config("c4", seed=9001, scale=64)
    pr = cg.problem_from_config(cfg)
    cg.assemble_advection_diffusion_cg(pr, cg.options_from_config(cfg), cfg.dt)
    cg.apply_dirichlet_conditions(pr, cfg.mesh.boundary, 0.0, cfg.dt)
    learned.set_params(pr, weights)
    out = {"elements": cfg.mesh.nele, "nodes": cfg.mesh.nnodes, "seed": 9001}
    for label, so in (
        ("fp64_jacobi_pcg", cg.SolverOptions("cg", "jacobi", relative_error=1e-8, norm="unpreconditioned",
                                             max_iterations=50000, check_every=32)),
        ("mixed_jacobi_pcg", cg.SolverOptions("cg", "jacobi", relative_error=1e-8, norm="unpreconditioned",
                                              max_iterations=50000, precision="mixed", check_every=32)),
        ("learned_fp64_fcg", cg.SolverOptions("fcg", "learned", relative_error=1e-8, norm="unpreconditioned",
                                              max_iterations=50000, check_every=32)),
        ("learned_mixed_fcg", cg.SolverOptions("fcg", "learned", relative_error=1e-8, norm="unpreconditioned",
                                               max_iterations=50000, precision="mixed", check_every=32)),
        ("fsai_fp64_pcg", cg.SolverOptions("cg", "fsai", relative_error=1e-8, norm="unpreconditioned",
                                           max_iterations=50000, check_every=32)),
        ("fsai_mixed_pcg", cg.SolverOptions("cg", "fsai", relative_error=1e-8, norm="unpreconditioned",
                                            max_iterations=50000, precision="mixed", check_every=32)),
    ):
        cg.petsc_solve(pr, so)
        t0 = time.perf_counter()
        _, info = cg.petsc_solve(pr, so)
        out[label] = {"seconds": round(time.perf_counter() - t0, 4), "its": info.its, "reason": info.reason,
                      "true_relres": info.true_relres, "fallback": info.fallback}
    pr.close()
    j, lm = out["fp64_jacobi_pcg"], out["learned_mixed_fcg"]
    out["iteration_ratio_jacobi_over_learned"] = round(j["its"] / max(out["learned_fp64_fcg"]["its"], 1), 3)
    out["speedup_learned_mixed_vs_fp64_jacobi"] = round(j["seconds"] / lm["seconds"], 3)
    out["speedup_learned_mixed_vs_mixed_jacobi"] = round(out["mixed_jacobi_pcg"]["seconds"] / lm["seconds"], 3)
    out["iteration_ratio_learned_over_fsai"] = round(out["learned_fp64_fcg"]["its"] /
                                                     max(out["fsai_fp64_pcg"]["its"], 1), 3)
    log_each(f"c4 held-out: jacobi {j['its']} its {j['seconds']:.4f}s, learned {out['learned_fp64_fcg']['its']} its, "
             f"learned+mixed {lm['seconds']:.4f}s ({out['speedup_learned_mixed_vs_fp64_jacobi']}x), "
             f"fsai {out['fsai_fp64_pcg']['its']} its, fsai+mixed {out['fsai_mixed_pcg']['seconds']:.4f}s")
    return out


def next_rows_report(cfg, pr, log_each) -> dict:
    """Timings of the rows beyond the element assembly and the solve, on the
    same mesh: the face loop (Neumann on x = 0, Robin on x = 1), the
    SUPG-stabilised assembly (a constant velocity field is set for it; the
    node-staged and colour schemes compared bit for bit), and the sparsity
    and colouring built on the GPU against the host routines, with the
    host-side block-plan build.  Leaves the context with the GPU-built
    sparsity and colouring bound."""
    import ctypes as C
    from fluidity_amd import cg, lib as L, meshgen as M
    lib = L.load()
    rep = {}
    t0 = time.perf_counter()
    surf = M.surface_of(cfg.mesh)
    bt, b1, b2, _ = M.unit_box_bcs(cfg.mesh, surf, "neumann_robin")
    pr.bind_surface(surf)
    log_each(f"surface: {surf.nfaces} faces ({time.perf_counter() - t0:.1f}s host extraction)")
    opts = cg.options_from_config(cfg)
    p = opts.params(cfg.dt)
    bcs = cg.SurfaceBCs(bt, b1, b2)
    cg.assemble_advection_diffusion_cg(pr, opts, cfg.dt, bcs)  # builds the face lists once
    reps = 10
    t0 = time.perf_counter()
    for _ in range(reps):
        L.check(lib.flgpu_advdif_surface(pr.ctx, C.byref(p), L.ptr(np.ascontiguousarray(bt, np.int32)),
                                         L.ptr(b1), L.ptr(b2)), "surface")
    face_ms = (time.perf_counter() - t0) / reps * 1e3
    rep["face_loop"] = {"faces": surf.nfaces, "ms_per_call": round(face_ms, 4),
                        "note": "wall time of flgpu_advdif_surface incl. the upload of the bc values"}
    log_each(f"face loop: {face_ms:.3f} ms per call")
    # SUPG: constant velocity on the same mesh, the stabilised row kernel
    pr.set_fields(velocity=np.array([1.0, 0.5, 0.25]))
    ms = C.c_double(0)
    for stab in ("supg", "streamline_upwind"):
        mats = {}
        for scheme in ("blocked", "colour"):  # node-staged slabs, then the colour scheme
            so = cg.options_from_config(cfg, scheme=scheme)
            so.exclude_advection_terms = False
            so.stabilisation = stab
            ps = so.params(cfg.dt)
            L.check(lib.flgpu_advdif_assemble_resident(pr.ctx, C.byref(ps)), "assemble " + stab)
            mats[scheme] = pr.matrix()[:2]
            L.check(lib.flgpu_time_assemble(pr.ctx, C.byref(ps), 5, C.byref(ms)), "time " + stab)
            key = f"assembly_{stab}" + ("" if scheme == "blocked" else "_colour_scheme")
            rep[key] = {"ms": round(ms.value, 4), "Melem/s": round(cfg.mesh.nele / ms.value / 1e3, 2)}
            log_each(f"assembly with {stab} ({scheme} scheme): {ms.value:.3f} ms "
                     f"({cfg.mesh.nele / ms.value / 1e3:.1f} Melem/s)")
        same = bool(np.array_equal(mats["blocked"][0], mats["colour"][0]) and
                    np.array_equal(mats["blocked"][1], mats["colour"][1]))
        rep[f"assembly_{stab}"]["blocked_equals_colour_bitwise"] = same
        log_each(f"{stab}: node-staged == colour scheme bit for bit: {same}")
    # the element routine's other branches on the same mesh (colour scheme,
    # every row by the general per-point row): the compressible multiphase
    # internal-energy equation, and mesh movement
    x = cfg.mesh.node_pos
    pr.set_fields(density=1.0 + 0.5 * x[:, 0], old_density=0.8 + 0.4 * x[:, 1], pressure=2.0 + x[:, 2],
                  volume_fraction=0.3 + 0.4 * x[:, 0] * x[:, 1])
    X = cfg.mesh.X
    pr.set_mesh_motion(X, X + 1e-3 * np.sin(np.pi * X), 1e-3 * np.sin(np.pi * X) / cfg.dt)
    for label, kw in (("assembly_internal_energy", dict(equation_type="internal_energy", compressible=True,
                                                        multiphase=True, density_theta=0.5, beta=0.5)),
                      ("assembly_moving_mesh", dict(move_mesh=True, beta=0.5))):
        so = cg.options_from_config(cfg)
        so.exclude_advection_terms = False
        so.exclude_mass_terms = False
        for k, v in kw.items():
            setattr(so, k, v)
        ps = so.params(cfg.dt)
        L.check(lib.flgpu_advdif_assemble_resident(pr.ctx, C.byref(ps)), "assemble " + label)
        L.check(lib.flgpu_time_assemble(pr.ctx, C.byref(ps), 5, C.byref(ms)), "time " + label)
        rep[label] = {"ms": round(ms.value, 4), "Melem/s": round(cfg.mesh.nele / ms.value / 1e3, 2)}
        log_each(f"{label}: {ms.value:.3f} ms ({cfg.mesh.nele / ms.value / 1e3:.1f} Melem/s)")
    # the host-side block plan (slabs, slots, places), rebuilt after every
    # adapt: a fresh bind of the same mesh, then two assemblies; the first
    # builds the plan
    pr2 = cg.problem_from_config(cfg)
    pa = cg.options_from_config(cfg).params(cfg.dt)
    t0 = time.perf_counter()
    L.check(lib.flgpu_advdif_assemble_resident(pr2.ctx, C.byref(pa)), "assemble")
    t1 = time.perf_counter()
    L.check(lib.flgpu_advdif_assemble_resident(pr2.ctx, C.byref(pa)), "assemble")
    t2 = time.perf_counter()
    pr2.close()
    rep["block_plan_build_s"] = round((t1 - t0) - (t2 - t1), 4)
    log_each(f"block plan build (host, once per mesh): {rep['block_plan_build_s']:.3f}s")
    # set-up: host routines, then the GPU ones (bound afterwards)
    t0 = time.perf_counter()
    cg.make_sparsity(cfg.mesh)
    t1 = time.perf_counter()
    cg.get_mesh_colouring(cfg.mesh)
    t2 = time.perf_counter()
    nnz = C.c_int64(0)
    ncol = C.c_int32(0)
    L.check(lib.flgpu_sparsity_build(pr.ctx, C.byref(nnz)), "sparsity_build")  # warm
    t3 = time.perf_counter()
    L.check(lib.flgpu_sparsity_build(pr.ctx, C.byref(nnz)), "sparsity_build")
    t4 = time.perf_counter()
    L.check(lib.flgpu_colouring_build(pr.ctx, C.byref(ncol)), "colouring_build")
    t5 = time.perf_counter()
    rep["setup"] = {"sparsity_gpu_s": round(t4 - t3, 4), "sparsity_host_s": round(t1 - t0, 4),
                    "colouring_gpu_s": round(t5 - t4, 4), "colouring_host_s": round(t2 - t1, 4),
                    "colours_gpu": ncol.value, "nnz": nnz.value}
    log_each(f"set-up: sparsity {t4 - t3:.3f}s GPU / {t1 - t0:.3f}s host, colouring {t5 - t4:.3f}s GPU "
             f"({ncol.value} colours) / {t2 - t1:.3f}s host")
    return rep


def rel(a, b) -> float:
    return float(np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300))


def oracle_mod():
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as O
    return O


def threads() -> int:
    return int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))


def spmv_report(lib, L, pr, nnodes, nnz, log_each) -> dict:
    """The solver-matrix SpMV, FP64 and FP32 values: warm (50 products back
    to back) and cold (each product after a 1 GiB flush of the caches)."""
    import ctypes as C
    ms = C.c_double(0)
    out = {}
    for prec, vb in (("fp64", 8), ("mixed", 4)):
        code = L.FLGPU_FP64 if prec == "fp64" else L.FLGPU_MIXED
        L.check(lib.flgpu_solver_prepare(pr.ctx, code), "prep")
        L.check(lib.flgpu_time_spmv(pr.ctx, code, 50, C.byref(ms)), "time_spmv")
        warm = ms.value
        L.check(lib.flgpu_time_spmv_cold(pr.ctx, code, 10, 1024, C.byref(ms)), "time_spmv_cold")
        cold = ms.value
        b = spmv_bytes(nnodes, nnz, vb)
        gw, gc = b / (warm * 1e-3) / 1e9, b / (cold * 1e-3) / 1e9
        out[f"spmv_{prec}"] = {"ms": round(warm, 4), "GB/s": round(gw, 1), "frac": round(gw / PEAK_HBM_GBS, 4),
                               "cold_ms": round(cold, 4), "cold_GB/s": round(gc, 1),
                               "cold_frac": round(gc / PEAK_HBM_GBS, 4), "algorithmic_bytes": b}
        log_each(f"spmv {prec}: warm {warm:.4f} ms ({gw:.0f} GB/s), cold {cold:.4f} ms ({gc:.0f} GB/s)")
    return out


def p1_config_report(name, log_each) -> dict:
    """A P1 configuration at its stated size (C1: 64^2 x 2 triangles; C2:
    64^3 x 6 tets): assembly Melem/s against its HBM roofline, the matrix
    and rhs against the oracle, SpMV warm and cold, and the configuration's
    solve (C1: GMRES(30) and Jacobi-PCG to 1e-10; C2: Jacobi-PCG to 1e-8)
    against the HBM floor of its iterations."""
    import ctypes as C
    from fluidity_amd import cg, lib as L, meshgen as M
    O = oracle_mod()
    lib = L.load()
    cfg = M.config(name, seed=1)
    pr = cg.problem_from_config(cfg)
    nnz = pr.sparsity.nnz
    p = cg.options_from_config(cfg).params(cfg.dt)
    ms = C.c_double(0)
    for _ in range(3):
        L.check(lib.flgpu_advdif_assemble_resident(pr.ctx, C.byref(p)), "assemble")
    L.check(lib.flgpu_time_assemble(pr.ctx, C.byref(p), 20, C.byref(ms)), "time_assemble")
    B = asm_bytes(cfg, nnz)
    gbs = B / (ms.value * 1e-3) / 1e9
    out = {"elements": cfg.mesh.nele, "nodes": cfg.mesh.nnodes, "nnz": nnz,
           "assembly": {"ms": round(ms.value, 4), "Melem/s": round(cfg.mesh.nele / ms.value / 1e3, 2),
                        "GB/s": round(gbs, 1), "frac": round(gbs / PEAK_HBM_GBS, 4), "algorithmic_bytes": B}}
    ref = O.reference_system(cfg, nthreads=threads())
    cg.assemble_advection_diffusion_cg(pr, cg.options_from_config(cfg), cfg.dt)
    cg.apply_dirichlet_conditions(pr, cfg.mesh.boundary, 0.0, cfg.dt)
    A, rhs, inact = pr.matrix()
    f, c, _ = ref["sparsity"]
    out["parity"] = {"pattern_bit_exact": bool(np.array_equal(pr.sparsity.findrm, f) and
                                               np.array_equal(pr.sparsity.colm, c)),
                     "A_rel_err_vs_oracle": rel(A, ref["A"]), "rhs_rel_err_vs_oracle": rel(rhs, ref["rhs"]),
                     "inactive_equal": bool(np.array_equal(inact, ref["inactive"]))}
    # C1 is non-symmetric (advection): GMRES(30); C2: Jacobi-PCG
    solves = ([("fp64_gmres", cg.SolverOptions("gmres", "jacobi", relative_error=1e-10, restart=30))] if name == "c1"
              else [("fp64_jacobi_pcg", cg.SolverOptions("cg", "jacobi", relative_error=cfg.rtol,
                                                         norm="unpreconditioned", max_iterations=50000))])
    b_it = spmv_bytes(cfg.mesh.nnodes, nnz, 8) + 128 * cfg.mesh.nnodes
    for label, so in solves:
        cg.petsc_solve(pr, so)
        t0 = time.perf_counter()
        x, info = cg.petsc_solve(pr, so)
        tot = time.perf_counter() - t0
        d = {"seconds": round(tot, 5), "krylov_seconds": round(info.seconds, 5), "its": info.its,
             "reason": info.reason, "true_relres": info.true_relres}
        if label == "fp64_jacobi_pcg":  # Jacobi-PCG iterations against their HBM floor
            floor = info.its * b_it / (PEAK_HBM_GBS * 1e9)
            d["hbm_floor_seconds"] = round(floor, 6)
            d["roofline_frac"] = round(floor / max(info.seconds, 1e-12), 4)
        xo, oits, _, _, _ = O.petsc_solve("gmres" if name == "c1" else "cg", "jacobi", 1e-12, ref["sparsity"], ref["A"],
                                          ref["inactive"], ref["rhs"], max_its=50000, nthreads=threads())
        d["x_rel_err_vs_oracle_1e-12"] = rel(x, xo)
        out[label] = d
    out.update(spmv_report(lib, L, pr, cfg.mesh.nnodes, nnz, lambda m: None))
    pr.close()
    a = out["assembly"]
    sv = out[solves[0][0]]
    log_each(f"{name}: assembly {a['ms']:.4f} ms ({a['Melem/s']:.0f} Melem/s, {a['frac']:.3f} of HBM), "
             f"A err {out['parity']['A_rel_err_vs_oracle']:.1e}, {solves[0][0]} {sv['its']} its "
             f"{sv['seconds']:.4f}s, x err {sv['x_rel_err_vs_oracle_1e-12']:.1e}")
    return out


def pressure_report(log_each) -> dict:
    """The pressure path on the 3-D P1 configuration's mesh (64^3 x 6 tets,
    P1 velocity and pressure): ct_m + lumped mass and C^T M_l^-1 C, device
    time per call of their kernels (flgpu_time_pressure) and the wall time
    of the whole API calls (host checks of the cmc pattern and copies
    included), ct_m's compulsory traffic against HBM, and both results
    against the oracle bit for bit."""
    from fluidity_amd import cg, meshgen as M
    O = oracle_mod()
    cfg = M.config("c2", seed=1)
    mesh = cfg.mesh
    pr = cg.CGProblem(mesh, quad_degree=2)
    pm = cg.vertex_pressure_mesh(pr)
    cg.bind_pressure_mesh(pr, pm)
    lib = pr.lib
    import ctypes as C
    from fluidity_amd import lib as L
    one = np.ones(1)
    L.check(lib.flgpu_ct_assemble(pr.ctx, 0, 1, L.ptr(one), None, None), "ct_assemble")  # builds the lists
    reps = 10
    t0 = time.perf_counter()
    for _ in range(reps):
        L.check(lib.flgpu_ct_assemble(pr.ctx, 0, 1, L.ptr(one), None, None), "ct_assemble")
    ct_call_ms = (time.perf_counter() - t0) / reps * 1e3
    cf, cc = cg.cmc_sparsity(pm, mesh.nnodes)
    cg.assemble_masslumped_cmc(pr, cf, cc, None)
    t0 = time.perf_counter()
    for _ in range(reps):
        L.check(lib.flgpu_cmc_assemble(pr.ctx, None, cf.size - 1, L.ptr(cf), L.ptr(cc), None), "cmc_assemble")
    cmc_call_ms = (time.perf_counter() - t0) / reps * 1e3
    ct_d, cmc_d = C.c_double(0), C.c_double(0)
    L.check(lib.flgpu_time_pressure(pr.ctx, reps, C.byref(ct_d), C.byref(cmc_d)), "time_pressure")
    ct_ms, cmc_ms = ct_d.value, cmc_d.value
    nnz_ct = pm.ct_colm.size
    # compulsory bytes: both connectivities, coordinates, ct_m and the lumped mass
    b_ct = 4 * (mesh.loc + pm.loc) * mesh.nele + 8 * mesh.dim * mesh.nx + 8 * mesh.dim * nnz_ct + 8 * mesh.nnodes
    ct_g, ml_g = cg.construct_momentum_ct(pr, pm)
    cmc_g = cg.assemble_masslumped_cmc(pr, cf, cc, None)
    asm, pd, cts, pn, p_nodes = O.pressure_setup(mesh, 2, colouring=(pr.colouring[1] - 1, pr.colouring[2]))
    ct_o, ml_o = O.ct_assemble(asm, pd)
    cmc_o = O.cmc(cts, ct_o, np.tile(1.0 / ml_o, (mesh.dim, 1)), (cf, cc))
    pr.close()
    out = {"elements": mesh.nele, "pressure_nodes": pm.nnodes, "ct_nnz": int(nnz_ct), "cmc_nnz": int(cc.size),
           "ct_and_masslump_ms": round(ct_ms, 4), "ct_GB/s": round(b_ct / (ct_ms * 1e-3) / 1e9, 1),
           "ct_frac": round(b_ct / (ct_ms * 1e-3) / 1e9 / PEAK_HBM_GBS, 4), "ct_algorithmic_bytes": int(b_ct),
           "cmc_ms": round(cmc_ms, 4), "ct_call_ms": round(ct_call_ms, 4), "cmc_call_ms": round(cmc_call_ms, 4),
           "ct_bitwise_equal_oracle": bool(np.array_equal(ct_g, ct_o) and np.array_equal(ml_g, ml_o)),
           "cmc_bitwise_equal_oracle": bool(np.array_equal(cmc_g, cmc_o))}
    log_each(f"pressure path (c2 mesh): device ct_m + lumped mass {ct_ms:.3f} ms, cmc {cmc_ms:.3f} ms "
             f"({cc.size} entries; whole calls {ct_call_ms:.3f} / {cmc_call_ms:.3f} ms), oracle bitwise: ct "
             f"{out['ct_bitwise_equal_oracle']}, "
             f"cmc {out['cmc_bitwise_equal_oracle']}")
    return out


def cpu_baseline(cfg, ref, log_each) -> dict:
    """The oracle (C restatement, OpenMP over the colour loop like the
    reference) on this host's cores: the assembly of the whole mesh, one
    warm-up then the median of 5 runs; and the time to solution of FP64
    Jacobi-PCG to a 1e-8 relative residual (the reference's KSPCG default
    preconditioned norm), the median of 5 runs."""
    O = oracle_mod()
    nt = threads()
    O.assemble(ref["params"], ref["asm"], nt)
    asm = [O.assemble(ref["params"], ref["asm"], nt)[2] for _ in range(5)]
    sol = []
    its = 0
    for _ in range(5):
        t0 = time.perf_counter()
        _, its, reason, relres, _ = O.petsc_solve("cg", "jacobi", 1e-8, ref["sparsity"], ref["A"], ref["inactive"],
                                                  ref["rhs"], max_its=50000, nthreads=nt)
        sol.append(time.perf_counter() - t0)
    a, s = float(np.median(asm)), float(np.median(sol))
    log_each(f"cpu baseline ({nt} threads): assembly median {a:.3f}s, PCG to 1e-8 median {s:.3f}s ({its} its)")
    return {"value": round(cfg.mesh.nele / a / 1e6, 3), "unit": "Melem/s", "cores": nt, "kind": "port",
            "sample": f"oracle assembly of the whole mesh (colour loop, {nt} OpenMP threads), median of 5 runs after "
                      f"1 warm-up; oracle Jacobi-PCG to a 1e-8 relative residual, median of 5 runs",
            "assembly_seconds_median": round(a, 4), "assembly_seconds_runs": [round(t, 4) for t in asm],
            "pcg_time_to_solution_s_median": round(s, 4), "pcg_its": its,
            "pcg_ms_per_iteration": round(s / max(its, 1) * 1e3, 3)}


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--config", default="c5")
    ap.add_argument("--scale", type=int, default=None)
    ap.add_argument("--no-cpu", action="store_true", help="skip the oracle: no parity leg, no CPU baseline")
    ap.add_argument("--no-solve", action="store_true")
    ap.add_argument("--scheme", default="auto")
    ap.add_argument("--no-c3", action="store_true", help="skip the C3 and C4 configurations")
    ap.add_argument("--no-next", action="store_true")
    ap.add_argument("--no-p1", action="store_true", help="skip the C1 and C2 configurations")
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU; --gpus must be 1 (sharding the path "
                                   "over GPUs, by element partition with a node halo, is not implemented)"}))
        return 2

    from fluidity_amd import cg, lib as L, meshgen as M
    import ctypes as C

    t0 = time.time()
    cfg = M.config(args.config, seed=1, scale=args.scale)
    log(f"config {cfg.name}: {cfg.mesh.nele} elements, {cfg.mesh.nnodes} nodes ({time.time() - t0:.1f}s)")
    pr = cg.problem_from_config(cfg)
    nnz = pr.sparsity.nnz
    log(f"bound: nnz {nnz}, colours {pr.ncolours} ({time.time() - t0:.1f}s)")
    opts = cg.options_from_config(cfg, scheme=args.scheme)
    p = opts.params(cfg.dt)
    lib = L.load()
    ms = C.c_double(0)
    for _ in range(max(args.warmup, 0)):
        L.check(lib.flgpu_advdif_assemble_resident(pr.ctx, C.byref(p)), "assemble")
    L.check(lib.flgpu_time_assemble(pr.ctx, C.byref(p), args.steps, C.byref(ms)), "time_assemble")
    ms_step = ms.value
    t_asm = ms_step * 1e-3
    melems = cfg.mesh.nele / t_asm / 1e6
    B = asm_bytes(cfg, nnz)
    achieved = B / t_asm / 1e9
    log(f"assembly: {ms.value:.3f} ms/step, {melems:.1f} Melem/s, {achieved:.1f} GB/s algorithmic")
    plan = cg.assembly_plan_info(pr)
    plan["node_stagings_per_node"] = round(plan["node_visits"] / cfg.mesh.nnodes, 3)
    log(f"assembly plan: {plan['slabs']} slabs, {plan['node_stagings_per_node']} stagings per node, "
        f"{plan['lds_bytes_fp64']} B LDS ({plan['workgroups_per_cu']} per CU), built in {plan['build_seconds']} s")
    L.check(lib.flgpu_advdif_assemble_resident(pr.ctx, C.byref(p)), "assemble")
    A_blocked, rhs_blocked, _ = pr.matrix()
    schemes = {}
    if args.scheme == "auto":
        p_col = cg.options_from_config(cfg, scheme="colour").params(cfg.dt)
        L.check(lib.flgpu_advdif_assemble_resident(pr.ctx, C.byref(p_col)), "assemble colour")
        A_col, rhs_col, _ = pr.matrix()
        L.check(lib.flgpu_time_assemble(pr.ctx, C.byref(p_col), max(2, args.steps // 4), C.byref(ms)), "time")
        schemes["colour"] = {"ms": round(ms.value, 4), "Melem/s": round(cfg.mesh.nele / ms.value / 1e3, 2),
                             "equals_blocked_bitwise": bool(np.array_equal(A_col, A_blocked) and
                                                            np.array_equal(rhs_col, rhs_blocked))}
        del A_col, rhs_col
        log(f"assembly colour scheme: {ms.value:.3f} ms/step, bitwise equal to the blocked scheme: "
            f"{schemes['colour']['equals_blocked_bitwise']}")
        # FP32 element arithmetic (mixed precision), values widened into the FP64 matrix
        p32 = cg.options_from_config(cfg, precision="fp32").params(cfg.dt)
        for _ in range(3):
            L.check(lib.flgpu_advdif_assemble_resident(pr.ctx, C.byref(p32)), "assemble fp32")
        A32, rhs32, _ = pr.matrix()
        L.check(lib.flgpu_time_assemble(pr.ctx, C.byref(p32), args.steps, C.byref(ms)), "time")
        schemes["blocked_fp32_elements"] = {"ms": round(ms.value, 4),
                                            "Melem/s": round(cfg.mesh.nele / ms.value / 1e3, 2),
                                            "A32": A32, "rhs32": rhs32}
        log(f"assembly with FP32 element arithmetic: {ms.value:.3f} ms/step")
        L.check(lib.flgpu_advdif_assemble_resident(pr.ctx, C.byref(p)), "assemble")

    out = {
        "metric": "Melem/s assembly; solve time to 1e-8 res: learned+mixed vs FP64 PCG",
        "value": round(melems, 3),
        "unit": "Melem/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_step, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (splitmix64-seeded mesh jitter and log-normal diffusivity)",
        "config": {"workload": f"{cfg.name}: graded pressure Poisson, {cfg.mesh.nele} P1 tets, "
                               f"{cfg.mesh.nnodes} nodes, nnz {nnz}", "elements": cfg.mesh.nele,
                   "nodes": cfg.mesh.nnodes, "nnz": nnz, "colours": pr.ncolours,
                   "parallelism": "single GPU"},
        "roofline": {"bound": "hbm", "achieved": round(achieved, 2), "peak": PEAK_HBM_GBS, "unit": "GB/s",
                     "frac": round(achieved / PEAK_HBM_GBS, 4),
                     "traffic": measured_traffic(("void flgpu::k_asm_staged", "void flgpu::k_asm_blocked")),
                     "kernel": "k_asm_staged / k_asm_blocked (one launch per step)", "algorithmic_bytes": B},
    }
    A32 = schemes.get("blocked_fp32_elements", {}).pop("A32", None)
    rhs32 = schemes.get("blocked_fp32_elements", {}).pop("rhs32", None)
    out["assembly_plan"] = plan
    if schemes:
        out["assembly_other_schemes"] = schemes

    xs = {}
    if not args.no_solve:
        cg.apply_dirichlet_conditions(pr, cfg.mesh.boundary, 0.0, cfg.dt)
        solves = {}
        from fluidity_amd import learned
        learned.set_params(pr, learned.load())
        for label, so in (
            ("fp64_jacobi_pcg", cg.SolverOptions("cg", "jacobi", relative_error=1e-8, norm="unpreconditioned",
                                                 max_iterations=50000, check_every=32)),
            ("mixed_jacobi_pcg", cg.SolverOptions("cg", "jacobi", relative_error=1e-8, norm="unpreconditioned",
                                                  max_iterations=50000, precision="mixed", check_every=32)),
            ("learned_fp64_fcg", cg.SolverOptions("fcg", "learned", relative_error=1e-8, norm="unpreconditioned",
                                                  max_iterations=50000, check_every=32)),
            ("learned_mixed_fcg", cg.SolverOptions("fcg", "learned", relative_error=1e-8, norm="unpreconditioned",
                                                   max_iterations=50000, precision="mixed", check_every=32)),
            ("fsai_fp64_pcg", cg.SolverOptions("cg", "fsai", relative_error=1e-8, norm="unpreconditioned",
                                               max_iterations=50000, check_every=32)),
            ("fsai_mixed_pcg", cg.SolverOptions("cg", "fsai", relative_error=1e-8, norm="unpreconditioned",
                                                max_iterations=50000, precision="mixed", check_every=32)),
            ("mixed_jacobi_pipecg", cg.SolverOptions("pipecg", "jacobi", relative_error=1e-8, norm="unpreconditioned",
                                                     max_iterations=50000, precision="mixed", check_every=32)),
            ("learned_mixed_pipecg", cg.SolverOptions("pipecg", "learned", relative_error=1e-8,
                                                      norm="unpreconditioned", max_iterations=50000,
                                                      precision="mixed", check_every=32)),
            ("fsai_mixed_pipecg", cg.SolverOptions("pipecg", "fsai", relative_error=1e-8, norm="unpreconditioned",
                                                   max_iterations=50000, precision="mixed", check_every=32)),
        ):
            cg.petsc_solve(pr, so)  # warm-up (allocations, first-touch)
            t0s = time.perf_counter()
            x, info = cg.petsc_solve(pr, so)
            tot = time.perf_counter() - t0s
            xs[label] = x
            solves[label] = {"seconds": round(tot, 4), "krylov_seconds": round(info.seconds, 4), "its": info.its,
                             "reason": info.reason, "true_relres": info.true_relres,
                             "refinements": info.refinements, "fallback": info.fallback}
            log(f"solve {label}: {tot:.4f}s total ({info.seconds:.4f}s Krylov), {info.its} its, "
                f"reason {info.reason}, relres {info.true_relres:.2e}, fallback {info.fallback}")
        ref = xs["fp64_jacobi_pcg"]
        for label, x in xs.items():
            solves[label]["x_rel_diff_vs_fp64_jacobi"] = float(np.max(np.abs(x - ref)) / np.max(np.abs(ref)))
        # time-to-solution against its HBM roofline: every Jacobi-PCG iteration
        # moves at least B_spmv + 128 N bytes (the SpMV plus the unfused
        # vector updates, FP64), so its / (8 TB/s / B_it) is the floor
        b_it = spmv_bytes(cfg.mesh.nnodes, nnz, 8) + 128 * cfg.mesh.nnodes
        j = solves["fp64_jacobi_pcg"]
        floor_s = j["its"] * b_it / (PEAK_HBM_GBS * 1e9)
        j["hbm_floor_seconds"] = round(floor_s, 5)
        j["roofline_frac"] = round(floor_s / j["krylov_seconds"], 4)
        log(f"FP64 Jacobi-PCG: HBM floor {floor_s:.4f}s for {j['its']} its of {b_it / 1e6:.0f} MB, "
            f"{j['roofline_frac']:.2f} of roofline")
        solves["speedup_learned_mixed_vs_fp64_jacobi"] = round(
            solves["fp64_jacobi_pcg"]["seconds"] / solves["learned_mixed_fcg"]["seconds"], 3)
        solves["speedup_learned_mixed_vs_mixed_jacobi"] = round(
            solves["mixed_jacobi_pcg"]["seconds"] / solves["learned_mixed_fcg"]["seconds"], 3)
        solves["speedup_fsai_mixed_vs_fp64_jacobi"] = round(
            solves["fp64_jacobi_pcg"]["seconds"] / solves["fsai_mixed_pcg"]["seconds"], 3)
        for lab in ("learned_mixed_pipecg", "fsai_mixed_pipecg", "mixed_jacobi_pipecg"):
            solves[f"speedup_{lab}_vs_fp64_jacobi"] = round(
                solves["fp64_jacobi_pcg"]["seconds"] / solves[lab]["seconds"], 3)
        # the exact FSAI on the same pattern is the best factor of the family
        # the network predicts: its iteration count is the network's ceiling
        solves["iteration_ratio_learned_over_fsai"] = round(
            solves["learned_fp64_fcg"]["its"] / max(solves["fsai_fp64_pcg"]["its"], 1), 3)
        log(f"speedup learned+mixed vs FP64 Jacobi-PCG: {solves['speedup_learned_mixed_vs_fp64_jacobi']}x, "
            f"vs mixed Jacobi-PCG: {solves['speedup_learned_mixed_vs_mixed_jacobi']}x; "
            f"fsai+mixed {solves['speedup_fsai_mixed_vs_fp64_jacobi']}x; learned/fsai its "
            f"{solves['iteration_ratio_learned_over_fsai']}; single-reduction CG: learned+mixed "
            f"{solves['speedup_learned_mixed_pipecg_vs_fp64_jacobi']}x, fsai+mixed "
            f"{solves['speedup_fsai_mixed_pipecg_vs_fp64_jacobi']}x, mixed Jacobi "
            f"{solves['speedup_mixed_jacobi_pipecg_vs_fp64_jacobi']}x")
        solves.update(spmv_report(lib, L, pr, cfg.mesh.nnodes, nnz, log))
        out["solve"] = solves

    ref_sys = None
    if not args.no_cpu:
        # full-size parity against the oracle (after the timed region)
        O = oracle_mod()
        t0 = time.time()
        ref_sys = O.reference_system(cfg, nthreads=threads())
        f, c, _ = ref_sys["sparsity"]
        cg.assemble_advection_diffusion_cg(pr, opts, cfg.dt)
        cg.apply_dirichlet_conditions(pr, cfg.mesh.boundary, 0.0, cfg.dt)
        A, rhs, inact = pr.matrix()
        par = {"pattern_bit_exact": bool(np.array_equal(pr.sparsity.findrm, f) and
                                         np.array_equal(pr.sparsity.colm, c)),
               "inactive_equal": bool(np.array_equal(inact, ref_sys["inactive"])),
               "A_rel_err_vs_oracle": rel(A, ref_sys["A"]), "rhs_rel_err_vs_oracle": rel(rhs, ref_sys["rhs"]),
               "A_rtol": 1e-12}
        # the matrix of the timed launches (Dirichlet changes only the rhs
        # and the mask) against the oracle's
        par["timed_assembly_A_rel_err_vs_oracle"] = rel(A_blocked, ref_sys["A"])
        if "colour" in schemes:
            par["blocked_equals_colour_bitwise"] = schemes["colour"]["equals_blocked_bitwise"]
        if A32 is not None:
            par["fp32_elements_A_rel_err_vs_oracle"] = rel(A32, ref_sys["A"])
            par["fp32_elements_rhs_rel_err_vs_fp64"] = rel(rhs32, rhs_blocked)
            par["fp32_rtol"] = 1e-5
        del A, rhs, A_blocked, rhs_blocked, A32, rhs32
        if xs:
            xo, oits, oreason, orel, _ = O.petsc_solve("cg", "jacobi", 1e-12, ref_sys["sparsity"], ref_sys["A"],
                                                       ref_sys["inactive"], ref_sys["rhs"], max_its=50000,
                                                       nthreads=threads())
            par["oracle_fp64_pcg"] = {"rtol": 1e-12, "its": oits, "reason": oreason, "true_relres": orel}
            par["x_bound"] = 1e-8
            for label, x in xs.items():
                par[f"x_rel_err_{label}_vs_oracle"] = rel(x, xo)
        out["parity_full"] = par
        log(f"parity (full size, {time.time() - t0:.1f}s): pattern {par['pattern_bit_exact']}, A err "
            f"{par['A_rel_err_vs_oracle']:.2e}, rhs err {par['rhs_rel_err_vs_oracle']:.2e}"
            + (f", learned+mixed x err {par['x_rel_err_learned_mixed_fcg_vs_oracle']:.2e}" if xs else ""))

    if not args.no_next:
        out["next_rows"] = next_rows_report(cfg, pr, log)
    pr.close()
    if not args.no_p1:
        out["c1"] = p1_config_report("c1", log)
        out["c2"] = p1_config_report("c2", log)
        if not args.no_cpu:
            out["pressure_path"] = pressure_report(log)
    if not args.no_c3:
        out["c3_mixed_precision"] = c3_mixed_report(log)
        if not args.no_solve:
            from fluidity_amd import learned
            out["c4_heldout"] = c4_heldout_report(learned.load(), log)
    if ref_sys is not None:
        out["cpu_baseline"] = cpu_baseline(cfg, ref_sys, log)
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
