"""Standalone launches of the hot kernels, for profiling under rocprofv3.

    python -m fluidity_amd.profile_kernels [--config c5] [--reps 3] [--copy-mib 1024]
                                           [--solves fsai:mixed,learned:mixed,jacobi:mixed:pipecg,...]

runs, on the device-resident system of the configuration: `reps` blocked
assemblies, `reps` colour-scheme assemblies, the solver-matrix SpMV (FP64 and
FP32 values) `reps` times each, then each listed solve (preconditioner:
precision[:ksp], by default CG, FCG with the learned preconditioner, to a
1e-8 true relative residual) once, `pressure` ct_m and C^T M_l^-1 C assemblies on the 3-D P1
configuration's mesh, and finally one plain device-to-device copy of `copy-mib`
MiB through the HIP runtime (its blit kernel reads and writes exactly that
many bytes), which calibrates the FETCH_SIZE / WRITE_SIZE counters.
"""
from __future__ import annotations

import argparse
import ctypes as C
import sys


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--config", default="c5")
    ap.add_argument("--reps", type=int, default=3)
    ap.add_argument("--copy-mib", type=int, default=1024)
    ap.add_argument("--solves", default="")
    ap.add_argument("--pressure", type=int, default=0, help="then this many ct_m and C^T M^-1 C assemblies "
                                                             "on the 3-D P1 configuration's mesh")
    args = ap.parse_args()
    from . import cg, lib as L, meshgen as M

    cfg = M.config(args.config, seed=1)
    pr = cg.problem_from_config(cfg)
    lib = L.load()
    for scheme in ("blocked", "colour"):
        p = cg.options_from_config(cfg, scheme=scheme).params(cfg.dt)
        for _ in range(args.reps):
            L.check(lib.flgpu_advdif_assemble_resident(pr.ctx, C.byref(p)), "assemble")
        print(f"assembled {args.reps}x ({scheme})", flush=True)
    cg.apply_dirichlet_conditions(pr, cfg.mesh.boundary, 0.0, cfg.dt)
    ms = C.c_double(0)
    for prec in (L.FLGPU_FP64, L.FLGPU_MIXED):
        L.check(lib.flgpu_solver_prepare(pr.ctx, prec), "solver_prepare")
        L.check(lib.flgpu_time_spmv(pr.ctx, prec, args.reps, C.byref(ms)), "time_spmv")
    print(f"spmv {args.reps}x per precision", flush=True)
    if args.solves:
        from . import learned
        learned.set_params(pr, learned.load())
    for item in filter(None, args.solves.split(",")):
        pc, prec, *ksp = item.split(":")
        ksp = ksp[0] if ksp else ("fcg" if pc == "learned" else "cg")
        so = cg.SolverOptions(ksp, pc, relative_error=1e-8, norm="unpreconditioned",
                              max_iterations=50000, precision=prec, check_every=32)
        _, info = cg.petsc_solve(pr, so)
        print(f"solve {pc} {prec} {ksp}: {info.its} its, reason {info.reason}", flush=True)
    pr.close()
    if args.pressure > 0:
        import numpy as np
        cfg2 = M.config("c2", seed=1)
        pq = cg.CGProblem(cfg2.mesh, quad_degree=2)
        pm = cg.vertex_pressure_mesh(pq)
        cg.bind_pressure_mesh(pq, pm)
        cf, cc = cg.cmc_sparsity(pm, cfg2.mesh.nnodes)
        one = np.ones(1)
        for _ in range(args.pressure):
            L.check(lib.flgpu_ct_assemble(pq.ctx, 0, 1, L.ptr(one), None, None), "ct_assemble")
            L.check(lib.flgpu_cmc_assemble(pq.ctx, None, cf.size - 1, L.ptr(cf), L.ptr(cc), None), "cmc_assemble")
        print(f"pressure path {args.pressure}x", flush=True)
        pq.close()
    if args.copy_mib > 0:
        _device_copy(args.copy_mib << 20)
        print(f"copied {args.copy_mib} MiB", flush=True)
    return 0


def _device_copy(nbytes: int) -> None:
    """One device-to-device hipMemcpy of nbytes (the runtime's blit copy
    kernel), between two fresh allocations."""
    import ctypes.util
    import os

    path = ctypes.util.find_library("amdhip64") or os.path.join(os.environ.get("ROCM_PATH", "/opt/rocm"), "lib",
                                                                   "libamdhip64.so")
    hip = C.CDLL(path)
    a, b = C.c_void_p(), C.c_void_p()
    for ptr in (a, b):
        if hip.hipMalloc(C.byref(ptr), C.c_size_t(nbytes)) != 0:
            raise RuntimeError("hipMalloc failed")
    hip.hipMemset(a, 1, C.c_size_t(nbytes))
    hip.hipDeviceSynchronize()
    if hip.hipMemcpy(b, a, C.c_size_t(nbytes), 3) != 0:  # hipMemcpyDeviceToDevice
        raise RuntimeError("hipMemcpy failed")
    hip.hipDeviceSynchronize()
    hip.hipFree(a)
    hip.hipFree(b)


if __name__ == "__main__":
    sys.exit(main())
