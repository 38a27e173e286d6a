"""Train the learned preconditioner on the GPU (the C4 training set) and
write the weights file.

    python -m fluidity_amd.train_learned [--systems 512] [--steps-per-system 4]
        [--lr 1e-3] [--out fluidity_amd/weights/learned_ic0_h32_l3.bin]

Every system is assembled by the library's own assembly on a seeded P1 tet
mesh of 20k-200k nodes with a log-normal diffusivity (correlation length
0.1, contrast up to 1e4), then the network takes Adam steps on the probed
loss ||I - G^T G A|| (8 Rademacher probes), all on the device.
"""
from __future__ import annotations

import argparse
import json
import sys
import time

import numpy as np

from . import cg, learned


def main(argv=None) -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--systems", type=int, default=512)
    ap.add_argument("--meshes", type=int, default=32)
    ap.add_argument("--steps-per-system", type=int, default=4)
    ap.add_argument("--lr", type=float, default=1e-3)
    ap.add_argument("--nprobe", type=int, default=8)
    ap.add_argument("--seed", type=int, default=2024)
    ap.add_argument("--n-min", type=int, default=27)
    ap.add_argument("--n-max", type=int, default=57)
    ap.add_argument("--out", default=learned.DEFAULT_WEIGHTS)
    ap.add_argument("--init", default=None, help="start from this weights file")
    args = ap.parse_args(argv)

    def log(msg):
        print(f"[train {time.strftime('%H:%M:%S')}] {msg}", flush=True)

    systems = learned.training_systems(args.systems, args.seed, args.n_min, args.n_max, args.meshes)
    params = learned.load(args.init) if args.init else None
    # a held-out system: the loss with fixed probes every 64 steps, and the
    # iterations of learned FCG against Jacobi-PCG at the start and the end
    vcfg = learned.validation_system()
    vpr = cg.problem_from_config(vcfg)
    cg.assemble_advection_diffusion_cg(vpr, cg.options_from_config(vcfg), vcfg.dt)
    cg.apply_dirichlet_conditions(vpr, vcfg.mesh.boundary, 0.0, vcfg.dt)
    _, jinfo = cg.petsc_solve(vpr, cg.SolverOptions("cg", "jacobi", relative_error=1e-8, norm="unpreconditioned"))
    history = []

    def validate(step, th):
        learned.set_params(vpr, th)
        vloss = learned.loss(vpr, nprobe=8, seed=4242)
        _, info = cg.petsc_solve(vpr, cg.SolverOptions("fcg", "learned", relative_error=1e-8,
                                                       norm="unpreconditioned"))
        history.append({"step": int(step), "val_loss": vloss, "val_its": int(info.its),
                        "fallback": int(info.fallback)})
        log(f"validation step {step}: loss {vloss:.5f}, learned FCG {info.its} its "
            f"(Jacobi-PCG {jinfo.its})")

    t0 = time.time()
    pr, losses = learned.train(lambda cfg: cg.CGProblem(cfg.mesh, quad_degree=cfg.quad_degree), systems,
                               args.steps_per_system, args.lr, args.nprobe, args.seed, log, params=params,
                               validate=validate)
    th = learned.get_params(pr)
    learned.save(args.out, th)
    summary = {"steps": int(losses.size), "loss_first64": float(np.mean(losses[:64])),
               "loss_last64": float(np.mean(losses[-64:])), "seconds": round(time.time() - t0, 1),
               "validation": {"nodes": vcfg.mesh.nnodes, "jacobi_its": int(jinfo.its), "history": history},
               "out": args.out}
    log(json.dumps(summary))
    pr.close()
    vpr.close()
    return 0


if __name__ == "__main__":
    sys.exit(main())
