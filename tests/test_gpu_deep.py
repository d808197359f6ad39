"""Trees deeper than the traversal stack (62 levels): the 2-bin tree of scenes.deep_chain
is 117 levels deep on both builders. Its walk goes by parent links instead of a stack
and must visit the same nodes in the same order as the oracle's restatement of
BinnedSAHBVH::Intersect (Core/Accel/BinnedSAHBVH.cpp:272-374), whose stack is
unbounded: closest hit, any hit and the work counters bit-identical, renders, AOVs and
batched renders (a deep tree beside a shallow one in one multi-job launch) within the
renders' 1e-4 bar; the same tree collapsed to 8-wide nodes as well."""
import numpy as np
import pytest

import oracle_lib as O
import petrichor_amd as P
from gpu_util import bits_equal, ctx, log, rel_err, to_dev
from petrichor_amd import scenes

pytestmark = pytest.mark.gpu

DEPTH = 117


@pytest.mark.parametrize("builder", [P.BUILD_DEVICE, P.BUILD_HOST])
def test_deep_tree_trace_matches_oracle(gpu, builder):
    sc = scenes.deep_chain(120)
    b = ctx().scene(sc).build_bvh(2, 4, builder=builder)
    info = b.info()
    log("deep chain, builder %d:" % builder, info)
    assert info["max_depth"] == DEPTH
    o = O.OracleScene(sc).use_tree(b.export_nodes() | {"prim_ids": b.export_order()})
    r = scenes.deep_chain_rays(sc["tri_pos"])
    tg, pg, st = b.trace_stats(to_dev(r))
    tg, pg = tg.cpu().numpy(), pg.cpu().numpy()
    to, po, so = o.trace_closest(r)
    log("  hits", int((po >= 0).sum()), "of", len(r), "gpu work", st, "oracle work", so)
    assert np.array_equal(pg, po)
    assert bits_equal(tg, to)
    assert np.array_equal(st, so)
    # every centroid ray finds its own triangle, down to the deepest leaf
    assert np.array_equal(pg[0:240:2], np.arange(120))
    hit = b.trace_any(to_dev(r)).cpu().numpy()
    assert np.array_equal(hit, (po >= 0).astype(np.int32))
    r2 = r.copy()
    h = po >= 0
    r2[h, 7] = np.nextafter(to[h], np.float32(0))
    t2, p2 = b.trace_closest(to_dev(r2))
    _, p2o, _ = o.trace_closest(r2)
    assert np.array_equal(p2.cpu().numpy(), p2o)


@pytest.mark.parametrize("integrator", [0, 1])
def test_deep_tree_render_matches_oracle(gpu, integrator):
    sc = scenes.deep_chain(120)
    b = ctx().scene(sc).build_bvh(2, 4)
    assert b.info()["max_depth"] == DEPTH
    o = O.OracleScene(sc).use_tree(b.export_nodes() | {"prim_ids": b.export_order()})
    img, st = b.render(48, 32, 4, 6, integrator, 3)
    ref, q = o.render(48, 32, 4, 6, integrator, 3, 0)
    err = rel_err(img, ref)
    log("deep render, integrator %d: queries %d / %d, max rel err %g" % (integrator, st["queries"], q, err.max()))
    assert np.abs(ref).mean() > 0.01
    assert st["queries"] == q
    assert err.max() <= 1e-4
    a, _ = b.render_aov(48, 32, 2, 6, P.AOV_DENOISING_NORMAL, 3)
    aref, _ = o.render(48, 32, 2, 6, 2 + P.AOV_DENOISING_NORMAL, 3, 0)
    assert rel_err(a, aref).max() <= 1e-4


def test_deep_tree_in_a_batch_with_a_shallow_one(gpu):
    c = ctx()
    deep = c.scene(scenes.deep_chain(120)).build_bvh(2, 4)
    cb = c.scene(scenes.cornell_box()).build_bvh(16)
    specs = [(deep, 40, 24, 4, 6, 0, 7), (cb, 32, 32, 4, 8, 0, 8), (deep, 24, 24, 3, 6, 1, 9)]
    jobs = [P.Job(b, w, h, spp, d, integ, seed) for (b, w, h, spp, d, integ, seed) in specs]
    eng = P.Engine(c, max_paths=1 << 14, max_jobs=4)
    eng.render_batch(jobs, iterations_per_sync=4)
    for j, (b, w, h, spp, d, integ, seed) in zip(jobs, specs):
        plain, _ = b.render(w, h, spp, d, integ, seed)
        assert bits_equal(j.result, plain), (w, h, integ)
    eng.close()


def test_deep_tree_8_wide_matches_oracle(gpu):
    # collapsed to 8-wide nodes the chain is 17 levels deep: its stack of up to 7 pending
    # children per level fits, and no ray here grazes a box within rounding of its hit
    sc = scenes.deep_chain(120)
    b = ctx().scene(sc).build_bvh(2, 4, width=8)
    o = O.OracleScene(sc).use_tree(b.export_nodes() | {"prim_ids": b.export_order()})
    r = scenes.deep_chain_rays(sc["tri_pos"])
    tg, pg = b.trace_closest(to_dev(r))
    to, po, _ = o.trace_closest(r)
    assert np.array_equal(pg.cpu().numpy(), po)
    assert bits_equal(tg.cpu().numpy(), to)
    img, _ = b.render(48, 32, 4, 6, 0, 3)
    ref, _ = o.render(48, 32, 4, 6, 0, 3, 0)
    assert rel_err(img, ref).max() <= 1e-4
