"""CPU unit tests of the element arithmetic the gfx950 kernels run
(fluidity_amd/csrc/element_kernels.h, compiled here with g++) against the
oracle's element routine (orc_element, the Fortran source order of
assemble_advection_diffusion_element_cg)."""
import ctypes as C
import os
import subprocess

import numpy as np
import pytest

import oracle as O
from fluidity_amd import meshgen as M
from fluidity_amd.elements import make_element

HERE = os.path.dirname(os.path.abspath(__file__))
BUILD = os.path.join(os.path.dirname(HERE), "fluidity_amd", "_build")
SO = os.path.join(BUILD, "libelemhost.so")


@pytest.fixture(scope="module")
def elemlib():
    os.makedirs(BUILD, exist_ok=True)
    subprocess.run(["g++", "-O2", "-std=c++17", "-fPIC", "-shared", "-ffp-contract=off", "-o", SO,
                    os.path.join(HERE, "elem_host.cpp")], check=True)
    lib = C.CDLL(SO)
    lib.elem_host.restype = C.c_int
    return lib


def contracted(el):
    """what, mhat, that as flgpu_field_mesh_bind computes them."""
    loc, ngi = el.n.shape
    what = np.zeros(loc)
    mhat = np.zeros(loc * loc)
    that = np.zeros(64)
    for i in range(loc):
        s = 0.0
        for g in range(ngi):
            s += el.w[g] * el.n[i, g]
        what[i] = s
        for j in range(loc):
            m = 0.0
            for g in range(ngi):
                m += el.w[g] * el.n[i, g] * el.n[j, g]
            mhat[i + loc * j] = m
    if loc == el.dim + 1:
        for i in range(loc):
            for j in range(loc):
                for k in range(loc):
                    m = 0.0
                    for g in range(ngi):
                        m += el.w[g] * el.n[i, g] * el.n[j, g] * el.n[k, g]
                    that[i + 4 * (j + 4 * k)] = m
    return what, mhat, that


def _p(a):
    if a is None:
        return None
    p = C.c_void_p(a.ctypes.data)
    p._keep = a
    return p


def host_element(lib, cfg, el, xel, ele, variant, fields, fp32=False):
    m = cfg.mesh
    d = m.dim
    what, mhat, that = contracted(el)
    xn = m.ndglno_x[ele * (d + 1):(ele + 1) * (d + 1)] - 1
    X = np.ascontiguousarray(m.X[xn].T.ravel(order="F"))  # X[a + dim*n]
    nodes = np.ascontiguousarray(m.ndglno[ele * el.loc:(ele + 1) * el.loc] - 1, dtype=np.int32)
    xdn = np.ascontiguousarray(xel.dn[:, 0, :].ravel(order="F"))
    eps = np.finfo(float).eps
    beta = variant.get("beta", cfg.beta)
    ibp = variant.get("integrate_advection_by_parts", False)
    adv = cfg.have_advection
    need_div = adv and ((abs(1 - beta) > eps) if ibp else (abs(beta) > eps))
    iparams = np.array([cfg.have_mass, variant.get("lump_mass_matrix", False), adv, ibp,
                        fields["kappa"] is not None, True, fields["src"] is not None,
                        fields["absorp"] is not None, need_div, abs(cfg.dt * cfg.theta) > eps,
                        O.STABILISATION[variant.get("stabilisation", "none")],
                        O.NU_BAR[variant.get("nu_bar_scheme", "optimal")]], np.int32)
    fparams = np.array([cfg.dt * cfg.theta, beta, variant.get("nu_scale", 0.5)])
    mat = np.zeros(el.loc * el.loc)
    rhs = np.zeros(el.loc)

    def f(key, ncomp):
        v = fields[key]
        if v is None:
            return None, 0
        a = np.ascontiguousarray(v, dtype=np.float64).ravel()
        return a, int(a.size == ncomp)

    U, uc = f("U", d)
    K, kc = f("kappa", d * d)
    S, sc = f("src", 1)
    B, bc = f("absorp", 1)
    T = np.ascontiguousarray(fields["T"], dtype=np.float64)
    nF, dnF, wC = el.fortran(el.n), el.fortran(el.dn), np.ascontiguousarray(el.w)  # keep alive
    rc = lib.elem_host(d, el.loc, int(fp32), _p(nF), _p(dnF),
                       _p(wC), _p(xdn), _p(what), _p(mhat), _p(that), el.ngi,
                       _p(X), _p(nodes), _p(T), _p(U), uc, _p(K), kc, _p(S), sc, _p(B), bc, _p(fparams),
                       _p(iparams), _p(mat), _p(rhs))
    assert rc == 0
    return mat.reshape((el.loc, el.loc), order="F"), rhs, (U, K, S, B)


def oracle_setup(cfg, variant):
    m = cfg.mesh
    el = make_element(m.dim, m.degree, cfg.quad_degree)
    xel = make_element(m.dim, 1, cfg.quad_degree)
    d = m.dim
    kap = cfg.kappa
    if kap is not None:
        if np.isscalar(kap):
            kap = (np.eye(d) * kap).ravel()
        else:
            full = np.zeros((m.nnodes, d, d))
            for a in range(d):
                full[:, a, a] = kap
            kap = full
    fields = {"T": cfg.T, "U": cfg.U, "kappa": kap, "src": cfg.source, "absorp": cfg.absorption}
    sp = O.make_sparsity(m)
    col, ncol = O.colour_elements(m)
    asm = O.Assembly(m, el, xel, fields, sp, O.colour_sets(col, ncol))
    p = O.params(cfg.dt, cfg.theta, variant.get("beta", cfg.beta), have_mass=cfg.have_mass,
                 lump=variant.get("lump_mass_matrix", False), adv=cfg.have_advection,
                 ibp=variant.get("integrate_advection_by_parts", False), diff=cfg.kappa is not None,
                 src=cfg.source is not None, absorp=cfg.absorption is not None,
                 stab=variant.get("stabilisation", "none"), nu_bar=variant.get("nu_bar_scheme", "optimal"),
                 nu_scale=variant.get("nu_scale", 0.5))
    return el, xel, fields, asm, p


CASES = [("c1", 6, {}), ("c2", 3, {}), ("c3", 2, {}), ("c4", 3, {}), ("c5", 3, {}),
         ("c1", 6, {"lump_mass_matrix": True}), ("c1", 6, {"integrate_advection_by_parts": True}),
         ("c1", 6, {"beta": 1.0}), ("c1", 6, {"integrate_advection_by_parts": True, "beta": 0.3}),
         ("c3", 2, {"beta": 0.7}), ("c3", 2, {"integrate_advection_by_parts": True}), ("c1p2", 4, {}),
         ("c1p2", 4, {"lump_mass_matrix": True}),
         # upwind stabilisation (row_stab: the general quadrature path)
         ("c1", 6, {"stabilisation": "supg"}), ("c1", 6, {"stabilisation": "streamline_upwind"}),
         ("c1", 6, {"stabilisation": "supg", "integrate_advection_by_parts": True, "beta": 0.3}),
         ("c1", 6, {"stabilisation": "supg", "nu_bar_scheme": "critical_rule", "lump_mass_matrix": True}),
         ("c1", 6, {"stabilisation": "streamline_upwind", "nu_bar_scheme": "unity", "nu_scale": 1.0}),
         ("c1p2", 4, {"stabilisation": "supg", "beta": 0.4}),
         ("c3", 2, {"stabilisation": "supg", "nu_bar_scheme": "doubly_asymptotic"}),
         ("c3", 2, {"stabilisation": "streamline_upwind"})]


@pytest.mark.parametrize("name,scale,variant", CASES)
def test_element_rows_match_oracle(elemlib, name, scale, variant):
    cfg = M.config(name, seed=9, scale=scale)
    if name == "c1":
        cfg.source = np.cos(3 * cfg.mesh.node_pos[:, 0])
        cfg.absorption = 0.5 + cfg.mesh.node_pos[:, 1]
    el, xel, fields, asm, p = oracle_setup(cfg, variant)
    rng = np.random.default_rng(0)
    for ele in rng.choice(cfg.mesh.nele, size=min(12, cfg.mesh.nele), replace=False):
        mref, rref = O.element(p, asm, int(ele))
        mh, rh, _ = host_element(elemlib, cfg, el, xel, int(ele), variant, fields)
        scale_m = np.max(np.abs(mref))
        assert np.max(np.abs(mh - mref)) <= 1e-13 * scale_m, (ele, np.max(np.abs(mh - mref)) / scale_m)
        scale_r = max(np.max(np.abs(rref)), 1e-300)
        assert np.max(np.abs(rh - rref)) <= 1e-12 * scale_r, (ele, np.max(np.abs(rh - rref)) / scale_r)
        mf, rf, _ = host_element(elemlib, cfg, el, xel, int(ele), variant, fields, fp32=True)
        assert np.max(np.abs(mf - mref)) <= 2e-6 * scale_m
