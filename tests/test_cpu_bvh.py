"""BVH build and closest-hit checks on the CPU (no GPU).

Pins the oracle's binned-SAH build (Core/Accel/BinnedSAHBVH.cpp:15-270) by invariants
and by an independent numpy evaluation of the root split, pins its traversal
(BinnedSAHBVH.cpp:272-374) against the layout-independent BruteForce scan
(Core/Accel/BruteForce.cpp:21-50), and checks the product's host builder
(pt_bvh_build_host) against the oracle's node table.
"""
import numpy as np
import pytest

import oracle_lib as O
import petrichor_amd as P
from petrichor_amd import scenes

f32 = np.float32


def rays_at(tri_pos, n, seed):
    rng = np.random.default_rng(seed)
    lo, hi = tri_pos.reshape(-1, 3).min(0), tri_pos.reshape(-1, 3).max(0)
    o = rng.uniform(lo - 0.5, hi + 0.5, (n, 3))
    tgt = rng.uniform(lo, hi, (n, 3))
    d = tgt - o
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    r = np.zeros((n, 8), np.float32)
    r[:, 0:3], r[:, 4:7] = o, d
    r[:, 3], r[:, 7] = 0.0, np.finfo(np.float32).max
    return r


def bare_scene(tri_pos):
    """A scene of the given triangles with one Lambert material (geometry tests only)."""
    sc = dict(scenes.random_triangles(1, seed=1))
    sc["tri_pos"] = np.ascontiguousarray(tri_pos, np.float32).reshape(-1, 9)
    sc["tri_mat"] = np.zeros(sc["tri_pos"].shape[0], np.int32)
    return sc


def prim_boxes(tri_pos):
    t = tri_pos.reshape(-1, 3, 3)
    return t.min(1), t.max(1)


@pytest.mark.parametrize("name", ["random", "cornell", "spheres"])
def test_oracle_bvh_invariants(name):
    sc = {"random": lambda: scenes.random_triangles(5000, seed=11),
          "cornell": scenes.cornell_box,
          "spheres": lambda: scenes.spheres_scene(n_spheres=4, seg=12, bands=7, terrain_n=16, env_w=32, env_h=16)}[name]()
    o = O.OracleScene(sc).build_bvh(16)
    e = o.bvh_export()
    n = len(sc["tri_mat"])
    ids = e["prim_ids"]
    assert np.array_equal(np.sort(ids), np.arange(n))  # a permutation
    lo, hi = prim_boxes(sc["tri_pos"])
    nn = len(e["leaf"])
    assert nn <= 2 * n - 1
    for i in range(nn):
        b, en = e["range"][i]
        assert 0 <= b < en <= n
        sub = ids[b:en]
        # node box = merge of its primitives' boxes (AABB::Merge is exact min/max)
        assert np.array_equal(e["boxes"][i, :3], lo[sub].min(0)) and np.array_equal(e["boxes"][i, 3:], hi[sub].max(0))
        if e["leaf"][i]:
            assert e["child"][i].tolist() == [0, 0] or e["child"][i].tolist() == [-1, -1]
        else:
            l, r = e["child"][i]
            assert e["range"][l][0] == b and e["range"][l][1] == e["range"][r][0] and e["range"][r][1] == en
            assert e["range"][l][1] > b and e["range"][r][0] < en  # both children non-empty


def test_root_split_matches_numpy_sah():
    """Recompute the root's 16-bin SAH split independently in numpy float32."""
    sc = scenes.random_triangles(3000, seed=5)
    t = sc["tri_pos"].reshape(-1, 3, 3)
    cen = ((t[:, 0] + t[:, 1]) + t[:, 2]) / f32(3.0)
    lo, hi = t.min(1), t.max(1)
    clo, chi = cen.min(0), cen.max(0)
    w = chi - clo
    axis = 0
    for a in (1, 2):
        if w[a] > w[axis]:
            axis = a
    ln = f32(chi[axis] - clo[axis])
    l = (cen[:, axis] - clo[axis]).astype(np.float32)
    pos = (f32(16.0) * l / ln).astype(np.float32)
    bins = np.nextafter(pos, f32(0.0)).astype(np.int32)

    def area(mask):
        if not mask.any():
            return f32(0.0)
        d = (hi[mask].max(0) - lo[mask].min(0)).astype(np.float32)
        return f32(2.0) * f32(f32(f32(d[0] * d[1]) + f32(d[1] * d[2])) + f32(d[2] * d[0]))

    costs = []
    for k in range(16):
        left = bins < k
        nl, nr = left.sum(), (~left).sum()
        costs.append(f32(f32(f32(nl) * area(left)) + f32(f32(nr) * area(~left))))
    best = int(np.argmin(costs))  # first minimum, as the strict '<' scan
    e = O.OracleScene(sc).build_bvh(16).bvh_export()
    l_child = e["child"][0][0]
    assert best > 0
    assert e["range"][l_child][1] == int((bins < best).sum())


@pytest.mark.parametrize("seed", [1, 2])
def test_bvh_closest_equals_brute_force(seed):
    sc = scenes.random_triangles(2000, seed=seed, mu=-2.5)
    o = O.OracleScene(sc).build_bvh(16)
    r = rays_at(sc["tri_pos"], 3000, seed + 100)
    t1, p1, st = o.trace_closest(r)
    t2, p2 = o.trace_brute(r)
    assert (p1 >= 0).sum() > 500
    assert np.array_equal(p1, p2)
    assert np.array_equal(t1.view(np.uint32), t2.view(np.uint32))
    assert st[0] > len(r) and st[1] > 0


def test_distance_window_is_respected():
    sc = scenes.random_triangles(2000, seed=3, mu=-2.5)
    o = O.OracleScene(sc).build_bvh(16)
    r = rays_at(sc["tri_pos"], 2000, 7)
    t_full, p_full, _ = o.trace_closest(r)
    r2 = r.copy()
    hit = p_full >= 0
    r2[hit, 7] = np.nextafter(t_full[hit], np.float32(0))  # dist_max just below the hit
    t2, p2, _ = o.trace_closest(r2)
    assert np.all((p2[hit] < 0) | (t2[hit] < t_full[hit]))
    tb, pb = o.trace_brute(r2)
    assert np.array_equal(pb, p2)


@pytest.mark.parametrize("name", ["random", "cornell", "spheres"])
def test_product_host_builder_matches_oracle(name):
    sc = {"random": lambda: scenes.random_triangles(20000, seed=7),
          "cornell": scenes.cornell_box,
          "spheres": lambda: scenes.spheres_scene(n_spheres=6, seg=16, bands=9, terrain_n=24, env_w=32, env_h=16)}[name]()
    e = O.OracleScene(sc).build_bvh(16).bvh_export()
    h = P.build_bvh_host(sc["tri_pos"], 16)
    for k in ("boxes", "child", "range", "leaf"):
        assert np.array_equal(e[k], h[k]), k
    assert e["max_depth"] == h["max_depth"]
    t = sc["tri_pos"].reshape(-1, 3, 3)
    cen = ((t[:, 0] + t[:, 1]) + t[:, 2]) / f32(3.0)
    for i in np.nonzero(e["leaf"])[0]:
        b, en = e["range"][i]
        a1, a2 = e["prim_ids"][b:en], h["prim_ids"][b:en]
        assert set(a1) == set(a2)
        if not np.array_equal(a1, a2):  # only where the sort keys tie
            assert any(np.array_equal(cen[a1, ax], cen[a2, ax]) for ax in range(3))
    if name == "random":
        assert np.array_equal(e["prim_ids"], h["prim_ids"])


def test_host_builder_32_bins_and_edge_cases():
    sc = scenes.random_triangles(5000, seed=9)
    e = O.OracleScene(sc).build_bvh(32).bvh_export()
    h = P.build_bvh_host(sc["tri_pos"], 32)
    for k in ("boxes", "child", "range", "leaf", "prim_ids"):
        assert np.array_equal(e[k], h[k]), k
    # 12 bins: bins * l / len rounds up to 12 at some nodes (the last bin takes it)
    sc12 = scenes.random_triangles(4000, seed=24, mu=-4.0)
    e12 = O.OracleScene(sc12).build_bvh(12).bvh_export()
    h12 = P.build_bvh_host(sc12["tri_pos"], 12)
    for k in ("boxes", "child", "range", "leaf", "prim_ids"):
        assert np.array_equal(e12[k], h12[k]), k
    # a single triangle, four triangles (one leaf), identical triangles (zero-width bins)
    one = scenes.random_triangles(1, seed=1)["tri_pos"]
    assert len(P.build_bvh_host(one)["leaf"]) == 1
    same = np.repeat(one, 40, axis=0)
    hs = P.build_bvh_host(same)
    assert len(hs["leaf"]) == 1 and hs["leaf"][0] == 1
    empty = P.build_bvh_host(np.zeros((0, 9), np.float32))
    assert len(empty["leaf"]) == 1


def test_deep_tree_host_builder_and_oracle():
    """A 117-level tree (two bins over the halving chain): the product's host builder gives
    the oracle's node table, and the oracle's walk (unbounded stack) finds the brute-force
    hits on the chain's rays, down to its deepest leaf."""
    sc = scenes.deep_chain(120)
    e = O.OracleScene(sc).build_bvh(2).bvh_export()
    assert e["max_depth"] == 117
    h = P.build_bvh_host(sc["tri_pos"], 2)
    for k in ("boxes", "child", "range", "leaf", "prim_ids"):
        assert np.array_equal(e[k], h[k]), k
    assert h["max_depth"] == 117
    o = O.OracleScene(sc).build_bvh(2)
    r = scenes.deep_chain_rays(sc["tri_pos"])
    t1, p1, _ = o.trace_closest(r)
    t2, p2 = o.trace_brute(r)
    assert np.array_equal(p1, p2)
    assert np.array_equal(t1.view(np.uint32), t2.view(np.uint32))
    assert np.array_equal(p1[0:240:2], np.arange(120))


@pytest.mark.parametrize("name", ["cornell", "spheres", "grid_ties", "identical"])
def test_reference_order_host_builder_equals_oracle_exactly(name):
    """PT_BUILD_HOST_REFERENCE_ORDER sorts every split node's range as
    BinnedSAHBVH::Build does (std::sort, Core/Accel/BinnedSAHBVH.cpp:104-107), so its leaf
    order equals the oracle's (which follows the same sort sequence) also among equal
    centroid keys, where the default builders order by triangle index. The scenes are
    chosen so that the default order differs: the Cornell box, a terrain grid, a shuffled
    grid of triangles whose centroids tie on every axis in rows, and 40 identical
    triangles (one leaf of equal keys, permuted by introsort's partitions)."""
    if name == "cornell":
        sc = scenes.cornell_box()
    elif name == "spheres":
        sc = scenes.spheres_scene(n_spheres=6, seg=16, bands=9, terrain_n=48, env_w=32, env_h=16)
    elif name == "grid_ties":
        g = np.arange(40, dtype=np.float32)
        x, y = np.meshgrid(g, g)
        x, y = x.ravel(), y.ravel()
        t = np.zeros((x.size, 3, 3), np.float32)
        t[:, 0, :2] = np.stack([x, y], 1)
        t[:, 1, :2] = np.stack([x + 1, y], 1)
        t[:, 2, :2] = np.stack([x, y + 1], 1)
        t[:, :, 2] = (x % 3)[:, None]
        t = t[np.random.default_rng(5).permutation(len(t))]  # ties not in index order
        sc = bare_scene(t.reshape(-1, 9))
    else:
        sc = bare_scene(np.repeat(scenes.random_triangles(1, seed=1)["tri_pos"], 40, axis=0))
    tri = np.ascontiguousarray(sc["tri_pos"], np.float32).reshape(-1, 9)
    e = O.OracleScene(sc).build_bvh(16).bvh_export()
    r = P.build_bvh_host(tri, 16, builder=P.BUILD_HOST_REFERENCE_ORDER)
    for k in ("boxes", "range", "leaf", "prim_ids"):
        assert np.array_equal(e[k], r[k]), k
    inner = e["leaf"] == 0  # a leaf's children are never read (the reference's root leaf keeps {0, 0})
    assert np.array_equal(e["child"][inner], r["child"][inner])
    assert e["max_depth"] == r["max_depth"]
    d = P.build_bvh_host(tri, 16)
    for k in ("boxes", "child", "range", "leaf"):
        assert np.array_equal(d[k], r[k]), k
    assert not np.array_equal(d["prim_ids"], r["prim_ids"]), "scene has no tied keys"
