"""Device binned-SAH builder against the host builder and through traversal.

The device builder (pt_bvh_build with PT_BUILD_DEVICE) must give the same tree as the
host builder (pt_bvh_build_host), which the CPU tests pin against the oracle's
restatement of BinnedSAHBVH::Build (Core/Accel/BinnedSAHBVH.cpp:15-225): the same node
table in the reference's numbering (boxes, children, ranges, leaf flags), the same leaf
order and depth. Boxes compare as floats (min/max merges may differ in the sign of a
zero coordinate). Traversal through both trees must then be bit-identical, work
counters included.
"""
import numpy as np
import pytest

import petrichor_amd as P
from gpu_util import bits_equal, ctx, log, to_dev
from petrichor_amd import scenes
from test_gpu_parity import rays_into

pytestmark = pytest.mark.gpu


def clustered(n=3000, seed=4):
    """Random triangles with groups of identical copies: 20 copies (a zero-width leaf sorted
    on the device), 100 copies (sorted on the host) and 7 copies (a small leaf)."""
    sc = scenes.random_triangles(n, seed=seed, mu=-3.0)
    t = sc["tri_pos"].copy()
    t[10:30] = t[10]
    t[200:300] = t[200]
    t[500:507] = t[500]
    sc["tri_pos"] = t
    return sc


def with_positions(t):
    sc = scenes.random_triangles(len(t), seed=1)
    sc["tri_pos"] = np.ascontiguousarray(t, np.float32)
    return sc


CASES = {
    "cornell": (scenes.cornell_box, 16, 4),
    "random50k": (lambda: scenes.random_triangles(50000, seed=21, mu=-4.0), 16, 4),
    "random50k_32bins": (lambda: scenes.random_triangles(50000, seed=22, mu=-4.0), 32, 4),
    "random50k_64bins": (lambda: scenes.random_triangles(50000, seed=25, mu=-4.0), 64, 4),
    "random_leaf1": (lambda: scenes.random_triangles(4000, seed=23, mu=-4.0), 16, 1),
    "random_leaf8": (lambda: scenes.random_triangles(4000, seed=24, mu=-4.0), 12, 8),
    "spheres": (lambda: scenes.spheres_scene(n_spheres=8, seg=16, bands=9, terrain_n=32, env_w=32, env_h=16), 16, 4),
    "terrain_grid": (lambda: with_positions(scenes.terrain(64)), 16, 4),
    "clustered": (clustered, 16, 4),
    "identical40": (lambda: with_positions(np.repeat(scenes.random_triangles(1, seed=2)["tri_pos"], 40, 0)), 16, 4),
    "n1": (lambda: scenes.random_triangles(1, seed=5), 16, 4),
    "n4": (lambda: scenes.random_triangles(4, seed=6), 16, 4),
    "n5": (lambda: scenes.random_triangles(5, seed=7), 16, 4),
    "n17": (lambda: scenes.random_triangles(17, seed=8), 16, 4),
    "n300k": (lambda: scenes.random_triangles(300000, seed=9, mu=-5.0), 16, 4),
}


@pytest.mark.parametrize("name", list(CASES))
def test_device_builder_equals_host_builder(gpu, name):
    log("device build", name)
    make, bins, leaf_max = CASES[name]
    sc = make()
    c = ctx()
    b = c.scene(sc).build_bvh(bins, leaf_max, builder=P.BUILD_DEVICE)
    info = b.info()
    e = b.export_nodes()
    order = b.export_order()
    h = P.build_bvh_host(sc["tri_pos"], bins, leaf_max)
    log("  nodes", info["num_nodes"], "depth", info["max_depth"], "device build ms %.2f" % info["build_ms"])
    assert info["num_nodes"] == len(h["leaf"])
    assert info["max_depth"] == h["max_depth"]
    assert np.array_equal(e["boxes"], h["boxes"])
    for k in ("child", "range", "leaf"):
        assert np.array_equal(e[k], h[k]), k
    assert np.array_equal(order, h["prim_ids"])


@pytest.mark.parametrize("name", ["clustered", "spheres", "random50k", "identical40"])
def test_traversal_same_through_both_builders(gpu, name):
    log("trace through both builders", name)
    make, bins, leaf_max = CASES[name]
    sc = make()
    c = ctx()
    s = c.scene(sc)
    bd = s.build_bvh(bins, leaf_max, builder=P.BUILD_DEVICE)
    bh = s.build_bvh(bins, leaf_max, builder=P.BUILD_HOST)
    r = np.concatenate([rays_into(sc, 8000, 3), scenes.surface_rays(sc["tri_pos"], 2000, seed=4)])
    td, pd, sd = bd.trace_stats(to_dev(r))
    th, ph, sh = bh.trace_stats(to_dev(r))
    td, pd, th, ph = td.cpu().numpy(), pd.cpu().numpy(), th.cpu().numpy(), ph.cpu().numpy()
    log("  hits", int((ph >= 0).sum()), "work", sd, sh)
    assert (ph >= 0).sum() > 100
    assert np.array_equal(pd, ph)
    assert bits_equal(td, th)
    assert np.array_equal(sd, sh)
    assert np.array_equal(bd.trace_any(to_dev(r)).cpu().numpy(), bh.trace_any(to_dev(r)).cpu().numpy())


def test_device_builder_rejects_bad_parameters(gpu):
    c = ctx()
    s = c.scene(scenes.random_triangles(100, seed=1))
    with pytest.raises(P.PtError):
        s.build_bvh(65, 4, builder=P.BUILD_DEVICE)
    with pytest.raises(P.PtError):
        s.build_bvh(1, 4, builder=P.BUILD_DEVICE)
    with pytest.raises(P.PtError):
        s.build_bvh(16, 0, builder=P.BUILD_DEVICE)
    with pytest.raises(P.PtError):
        s.build_bvh(16, 4, builder=7)
    # the host builder still takes up to 256 bins
    assert s.build_bvh(128, 4, builder=P.BUILD_HOST).info()["num_nodes"] >= 1
