"""Host side of the learned preconditioner: parameter files and the
training driver (the kernels live in csrc/learned.hip).

Weights file: b"FLGPUNN1", then int32 version, hidden, layers, count, then
count float32 parameters (little endian).
"""
from __future__ import annotations

import ctypes as C
import os
import struct
import time

import numpy as np

from . import lib as L

MAGIC = b"FLGPUNN1"
VERSION, HIDDEN, LAYERS = 1, 32, 3
HERE = os.path.dirname(os.path.abspath(__file__))
DEFAULT_WEIGHTS = os.path.join(HERE, "weights", "learned_ic0_h32_l3.bin")


def param_count() -> int:
    return int(L.load().flgpu_learned_param_count())


def get_params(problem) -> np.ndarray:
    p = np.zeros(param_count(), np.float32)
    L.check(problem.lib.flgpu_learned_get_params(problem.ctx, L.ptr(p)), "flgpu_learned_get_params")
    return p


def set_params(problem, params: np.ndarray) -> None:
    p = np.ascontiguousarray(params, dtype=np.float32)
    if p.size != param_count():
        raise ValueError(f"expected {param_count()} parameters, got {p.size}")
    L.check(problem.lib.flgpu_learned_set_params(problem.ctx, L.ptr(p)), "flgpu_learned_set_params")


def init_params(problem, seed: int) -> None:
    L.check(problem.lib.flgpu_learned_init_params(problem.ctx, C.c_uint64(seed)), "flgpu_learned_init_params")


def save(path: str, params: np.ndarray) -> None:
    p = np.ascontiguousarray(params, dtype="<f4")
    os.makedirs(os.path.dirname(os.path.abspath(path)), exist_ok=True)
    with open(path, "wb") as f:
        f.write(MAGIC)
        f.write(struct.pack("<iiii", VERSION, HIDDEN, LAYERS, p.size))
        f.write(p.tobytes())


def load(path: str = DEFAULT_WEIGHTS) -> np.ndarray:
    with open(path, "rb") as f:
        if f.read(8) != MAGIC:
            raise ValueError(f"{path}: not a learned-preconditioner weights file")
        version, hidden, layers, count = struct.unpack("<iiii", f.read(16))
        if (version, hidden, layers) != (VERSION, HIDDEN, LAYERS):
            raise ValueError(f"{path}: version/shape {version, hidden, layers} not supported")
        p = np.frombuffer(f.read(4 * count), dtype="<f4").astype(np.float32)
    if p.size != count:
        raise ValueError(f"{path}: truncated")
    return p


def loss_grad(problem, nprobe: int = 8, seed: int = 0):
    loss = C.c_float(0)
    g = np.zeros(param_count(), np.float32)
    L.check(problem.lib.flgpu_learned_loss_grad(problem.ctx, nprobe, C.c_uint64(seed), C.byref(loss), L.ptr(g)),
            "flgpu_learned_loss_grad")
    return float(loss.value), g


def loss(problem, nprobe: int = 8, seed: int = 0) -> float:
    """The probed loss alone (no gradient), e.g. on a fixed validation system
    with fixed probes."""
    val = C.c_float(0)
    L.check(problem.lib.flgpu_learned_loss_grad(problem.ctx, nprobe, C.c_uint64(seed), C.byref(val), None),
            "flgpu_learned_loss_grad")
    return float(val.value)


def train_step(problem, lr: float = 1e-3, nprobe: int = 8, seed: int = 0) -> float:
    loss = C.c_float(0)
    L.check(problem.lib.flgpu_learned_train_step(problem.ctx, nprobe, C.c_uint64(seed), C.c_float(lr),
                                                 C.byref(loss)), "flgpu_learned_train_step")
    return float(loss.value)


def factor(problem) -> np.ndarray:
    """Preconditioner values in A's pattern (see include/flgpu.h)."""
    nnz = C.c_int64(0)
    L.check(problem.lib.flgpu_learned_factor(problem.ctx, None, C.byref(nnz)), "flgpu_learned_factor")
    g = np.zeros(nnz.value, np.float32)
    L.check(problem.lib.flgpu_learned_factor(problem.ctx, L.ptr(g), C.byref(nnz)), "flgpu_learned_factor")
    return g


def training_systems(count: int = 512, seed: int = 2024, n_min: int = 27, n_max: int = 57,
                     meshes: int = 32, contrast_max: float = 1e4):
    """The training set: `count` systems on `meshes` seeded P1 tet meshes of
    n^3 x 6 elements with n in [n_min, n_max] ((n+1)^3 = 22k..195k nodes),
    each with a log-normal diffusivity field (correlation length 0.1,
    contrast drawn log-uniformly up to contrast_max).  Yields (mesh index,
    Config) in a fixed order."""
    from . import meshgen as M

    rng = np.random.default_rng(seed)
    sizes = rng.integers(n_min, n_max + 1, size=meshes)
    per_mesh = max(1, count // meshes)
    k = 0
    for mi, n in enumerate(sizes):
        grading = None if mi % 2 == 0 else float(rng.choice([10.0, 100.0]))
        mesh = M.cube_tets(int(n), jitter=0.15, seed=seed + mi, grading=grading)
        src = M.sin_source(mesh.node_pos)
        for j in range(per_mesh):
            if k >= count:
                return
            contrast = float(10 ** rng.uniform(1, np.log10(contrast_max)))
            kap = M.grf_lognormal(mesh.node_pos, seed * 1000 + k, 0.1, contrast)
            yield mi, M.Config(f"train{k}", mesh, 2, np.zeros(mesh.nnodes), None, kap, src, None,
                               1.0, 1.0, 0.0, False, False, "cg", 1e-8)
            k += 1


def validation_system(seed: int = 777, n: int = 40, contrast: float = 1e3):
    """A held-out system for tracking training: an n^3 x 6 jittered cube
    mesh and a log-normal diffusivity field from seeds the training set does
    not use."""
    from . import meshgen as M

    mesh = M.cube_tets(n, jitter=0.15, seed=seed)
    kap = M.grf_lognormal(mesh.node_pos, seed * 7919 + 1, 0.1, contrast)
    return M.Config("validation", mesh, 2, np.zeros(mesh.nnodes), None, kap, M.sin_source(mesh.node_pos), None,
                    1.0, 1.0, 0.0, False, False, "cg", 1e-8)


def train(problem_factory, systems, steps_per_system: int = 4, lr: float = 1e-3, nprobe: int = 8,
          seed: int = 0, log=None, init_seed: int | None = 1, params: np.ndarray | None = None,
          validate=None, every: int = 64):
    """Adam on the probed loss over a stream of systems.  problem_factory()
    returns a CGProblem whose context keeps the network across rebinds.
    validate(step, params), if given, runs every `every` steps and at the
    end."""
    from . import cg

    pr = None
    losses = []
    step = 0
    last_mesh = None
    t0 = time.time()
    for mi, cfg in systems:
        if pr is None:
            pr = problem_factory(cfg)
            if params is not None:
                set_params(pr, params)
            elif init_seed is not None:
                init_params(pr, init_seed)
        elif mi != last_mesh:
            pr.bind(cfg.mesh, quad_degree=cfg.quad_degree)
        last_mesh = mi
        pr.set_fields(T=cfg.T, diffusivity=cfg.kappa, source=cfg.source)
        opts = cg.options_from_config(cfg)
        cg.assemble_advection_diffusion_cg(pr, opts, cfg.dt)
        cg.apply_dirichlet_conditions(pr, cfg.mesh.boundary, 0.0, cfg.dt)
        for _ in range(steps_per_system):
            losses.append(train_step(pr, lr=lr, nprobe=nprobe, seed=seed * 100003 + step))
            step += 1
        if log and step % every < steps_per_system:
            log(f"step {step}: loss {np.mean(losses[-64:]):.4f} ({time.time() - t0:.1f}s)")
        if validate and step % every < steps_per_system:
            validate(step, get_params(pr))
    if validate and pr is not None:
        validate(step, get_params(pr))
    return pr, np.array(losses)
