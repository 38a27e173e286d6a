"""flgpu_assembly_plan_info: the blocked scheme's slab plan, as the bench
reports it (slabs, extents, stagings per node, LDS, build time)."""
import pytest

from fluidity_amd import cg, meshgen as M

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("name,scale", [("c2", 10), ("c1", 24)])
def test_plan_info_is_consistent(require_gpu, name, scale):
    cfg = M.config(name, seed=3, scale=scale)
    print(f"\n[case] plan info {name} scale {scale}", flush=True)
    pr = cg.problem_from_config(cfg)
    info = cg.assembly_plan_info(pr)
    m = cfg.mesh
    assert info["staged"] == 1
    assert info["element_rows"] == m.nele * (m.dim + 1)
    assert 1 <= info["slabs"] <= m.nnodes
    # every slab stages its rows and their neighbours: each node at least once
    assert m.nnodes <= info["node_visits"] <= info["slabs"] * info["max_nodes"]
    assert 0 < info["max_nodes"] <= 256 and 0 < info["max_slots"] <= 768
    assert info["lds_bytes_fp64"] > 8 * info["max_slots"] * (m.dim + 2)
    assert info["workgroups_per_cu"] == 160 * 1024 // info["lds_bytes_fp64"]
    assert info["build_seconds"] >= 0.0
    # the plan is built once: a second call reports the same plan
    assert cg.assembly_plan_info(pr) == info
    pr.close()
