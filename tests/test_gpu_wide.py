"""Wide BVHs (widths 4 and 8): nearest hits and renders against the oracle.

The wide traversal visits boxes nearest first and culls by the nearest hit so far,
not in the reference's order (BinnedSAHBVH::Intersect, Core/Accel/BinnedSAHBVH.cpp:
272-374); leaves, triangle tests and the strictly-nearer rule are the reference's. The
nearest hit distance is therefore the reference's for every ray that does not graze a
box within rounding of its hit, and the triangle is the reference's unless two
triangles are hit at exactly the same distance. The tests require identical distances
and any-hit flags for all rays, and identical triangles wherever the oracle's nearest
distance is unique.
"""
import numpy as np
import pytest

import oracle_lib as O
import petrichor_amd as P
from gpu_util import bits_equal, ctx, log, rel_err, to_dev
from petrichor_amd import scenes
from test_gpu_build import clustered
from test_gpu_parity import rays_into

pytestmark = pytest.mark.gpu

SCENES = {
    "cornell": scenes.cornell_box,
    "random": lambda: scenes.random_triangles(50000, seed=21, mu=-4.0),
    "spheres": lambda: scenes.spheres_scene(n_spheres=8, seg=16, bands=9, terrain_n=32, env_w=64, env_h=32),
    "clustered": clustered,
}


@pytest.mark.parametrize("width", [4, 8])
@pytest.mark.parametrize("name", list(SCENES))
@pytest.mark.parametrize("builder", [P.BUILD_DEVICE, P.BUILD_HOST])
def test_wide_nearest_hits_match_oracle(gpu, name, builder, width):
    log("%d-wide trace" % width, name, builder)
    sc = SCENES[name]()
    c = ctx()
    s = c.scene(sc)
    b2 = s.build_bvh(16, builder=builder)
    b8 = s.build_bvh(16, width=width, builder=builder)
    o = O.OracleScene(sc).use_tree(b2.export_nodes() | {"prim_ids": b2.export_order()})
    r = np.concatenate([rays_into(sc, 20000, 5), scenes.surface_rays(sc["tri_pos"], 5000, seed=6)])
    r[20000:, 3] = 1e-5  # surface rays leave their own triangle
    t8, p8, st8 = b8.trace_stats(to_dev(r))
    t8, p8 = t8.cpu().numpy(), p8.cpu().numpy()
    to, po, so = o.trace_closest(r)
    hits = po >= 0
    log("  hits", int(hits.sum()), "of", len(r), "wide work", st8, "reference-order work", so)
    assert hits.sum() > 1000
    assert bits_equal(t8, to)
    same = p8 == po
    if not same.all():
        # another triangle only where two are hit at exactly the same distance (the
        # distances themselves are equal on every ray, checked above)
        bad = np.nonzero(~same)[0]
        log("  triangles differ on", len(bad), "rays (equal distances)")
        assert len(bad) <= max(2, len(r) // 10000)
        assert bits_equal(t8[bad], to[bad])
    hit8 = b8.trace_any(to_dev(r)).cpu().numpy()
    assert np.array_equal(hit8, hits.astype(np.int32))


@pytest.mark.parametrize("width", [4, 8])
@pytest.mark.parametrize("integrator", [0, 1])
@pytest.mark.parametrize("name", ["cornell", "spheres"])
def test_wide_render_matches_oracle(gpu, name, integrator, width):
    log("%d-wide render" % width, name, integrator)
    sc = SCENES[name]()
    c = ctx()
    s = c.scene(sc)
    b2 = s.build_bvh(16)
    b8 = s.build_bvh(16, width=width)
    w, h, spp, depth = (96, 96, 8, 8) if name == "cornell" else (128, 72, 8, 12)
    img8, _ = b8.render(w, h, spp, depth, integrator, 3)
    img2, _ = b2.render(w, h, spp, depth, integrator, 3)
    o = O.OracleScene(sc).use_tree(b2.export_nodes() | {"prim_ids": b2.export_order()})
    ref, _ = o.render(w, h, spp, depth, integrator, 3, 0)
    err = rel_err(img8, ref.reshape(img8.shape))
    exact = np.mean(img8 == img2)
    log("  max rel err %.3g, identical to the binary image: %.4f" % (err.max(), exact))
    assert err.max() <= 1e-4


@pytest.mark.parametrize("width", [4, 8])
def test_wide_batch_equals_single(gpu, width):
    log("%d-wide batch" % width)
    c = ctx()
    cb = c.scene(scenes.cornell_box()).build_bvh(16, width=width)
    sp = c.scene(SCENES["spheres"]()).build_bvh(16, width=width)
    jobs = [P.Job(cb, 64, 64, 4, 8, 0, 5), P.Job(sp, 80, 60, 6, 12, 0, 6), P.Job(cb, 48, 40, 3, 8, 1, 7)]
    eng = P.Engine(c, 1 << 16, max_jobs=8)
    eng.render_batch(jobs)
    for j in jobs:
        single, _ = j.bvh.render(j.width, j.height, j.spp, j.max_bounces, j.integrator, j.seed)
        assert bits_equal(j.result, single)
    # jobs of both widths in one batch run one width at a time, with the same images
    b2 = c.scene(scenes.cornell_box()).build_bvh(16)
    mixed = [P.Job(b2, 32, 32, 2, 8, 0, 1), P.Job(cb, 32, 32, 2, 8, 0, 1), P.Job(b2, 40, 24, 3, 8, 1, 2)]
    eng.render_batch(mixed)
    for j in mixed:
        single, _ = j.bvh.render(j.width, j.height, j.spp, j.max_bounces, j.integrator, j.seed)
        assert bits_equal(j.result, single)
