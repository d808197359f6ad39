"""Batched rendering: jobs of different scenes and settings in one engine give, per job,
the same image as the plain per-job render (same kernels, same seeds) and as the oracle."""
import numpy as np
import pytest

import oracle_lib as O
import petrichor_amd as P
from gpu_util import bits_equal, ctx, log, rel_err
from petrichor_amd import scenes

pytestmark = pytest.mark.gpu


def test_batch_matches_plain_and_oracle(gpu):
    c = ctx()
    cb = c.scene(scenes.cornell_box()).build_bvh(16)
    sp_sc = scenes.spheres_scene(n_spheres=8, seg=16, bands=9, terrain_n=32, env_w=256, env_h=128)
    sp = c.scene(sp_sc).build_bvh(16)
    specs = [(cb, 48, 40, 6, 8, 0, 11), (sp, 40, 32, 5, 12, 0, 12), (cb, 32, 32, 3, 8, 1, 13),
             (sp, 24, 24, 7, 10, 1, 14), (cb, 64, 16, 2, 8, 0, 15)]
    jobs = [P.Job(b, w, h, spp, d, integ, seed) for (b, w, h, spp, d, integ, seed) in specs]
    eng = P.Engine(c, max_paths=4096, max_jobs=8)  # smaller than the total: admission waits for slots
    st = eng.render_batch(jobs, iterations_per_sync=4)
    log("batch", st)
    total_q = 0
    for j, (b, w, h, spp, d, integ, seed) in zip(jobs, specs):
        plain, pst = b.render(w, h, spp, d, integ, seed)
        total_q += pst["queries"]
        assert bits_equal(j.result, plain), (w, h, spp, integ, seed)
        o = O.OracleScene(b.scene.src).use_tree(b.export_nodes() | {"prim_ids": b.export_order()})
        ref, q = o.render(w, h, spp, d, integ, seed, 0)
        assert rel_err(j.result, ref).max() <= 1e-4
    assert st["queries"] == total_q


def test_timed_render_matches_graph_render(gpu):
    c = ctx()
    b = c.scene(scenes.cornell_box()).build_bvh(16)
    a, sa = b.render(40, 30, 4, 8, 0, 3)
    t, stt = b.render(40, 30, 4, 8, 0, 3, time_kernels=True)
    assert bits_equal(a, t)
    assert sa["queries"] == stt["queries"]
    k = stt["kernels"]
    assert k["extension"]["items"] + k["shadow_any"]["items"] + k["shadow_closest"]["items"] == stt["queries"]


def test_autotune_picks_a_width_whose_batch_equals_plain(gpu):
    """The autotuner keeps to binary BVHs unless told otherwise; asked to, it times the
    calibration batch over BVHs of widths 2, 4 and 8 and returns one width's jobs and launch
    shapes; the batch rendered with them equals the per-job plain renders of the same BVHs."""
    from petrichor_amd import autotune
    c = ctx()
    scs = [c.scene(scenes.cornell_box()),
           c.scene(scenes.spheres_scene(n_spheres=8, seg=16, bands=9, terrain_n=32, env_w=256, env_h=128))]
    specs = [(0, 48, 40, 6, 8, 11), (1, 40, 32, 5, 12, 12), (1, 24, 24, 7, 10, 14)]
    by_width = {}
    for w in (2, 4, 8):
        bv = [s.build_bvh(16, width=w) for s in scs]
        by_width[w] = [P.Job(bv[i], W, H, spp, d, 0, seed) for (i, W, H, spp, d, seed) in specs]
    exact = autotune.tune(c, None, pool_sizes=[4096], passes=0, by_width=by_width)
    assert exact["width"] == 2  # the wide walks are not the reference's: candidates only on request
    best = autotune.tune(c, None, pool_sizes=[4096], passes=0, by_width=by_width, exact_only=False)
    log("autotune", {k: v for k, v in best.items() if k != "jobs"})
    assert best["width"] in (2, 4, 8) and best["jobs"] is by_width[best["width"]]
    eng = P.Engine(c, best["max_paths"], max_jobs=4)
    eng.set_tuning(**best["tuning"])
    jobs = [P.Job(j.bvh, j.width, j.height, j.spp, j.max_bounces, j.integrator, j.seed) for j in best["jobs"]]
    eng.render_batch(jobs)
    for j in jobs:
        plain, _ = j.bvh.render(j.width, j.height, j.spp, j.max_bounces, j.integrator, j.seed, plain=True)
        assert j.bvh.width == best["width"]
        assert bits_equal(j.result, plain)


def test_large_image_in_two_halves_equals_one_piece(gpu):
    """With two_streams, an image of at least 2^19 pixels renders in two halves on two
    engines (streams); it equals the same render in one piece, with plain launches and
    with the graph path, and the oracle."""
    c = ctx()
    b = c.scene(scenes.cornell_box()).build_bvh(16)
    w, h = 1024, 576
    halves, st = b.render(w, h, 2, 8, 0, 21, two_streams=True)
    one, _ = b.render(w, h, 2, 8, 0, 21)
    assert bits_equal(halves, one)
    whole, st1 = b.render(w, h, 2, 8, 0, 21, plain=True)
    assert bits_equal(halves, whole)
    assert st["queries"] == st1["queries"]
    o = O.OracleScene(b.scene.src).use_tree(b.export_nodes() | {"prim_ids": b.export_order()})
    rows = [0, 287, 288, 575]  # both halves' first and last rows
    pix = np.array([r * w + x for r in rows for x in range(w)], np.int32)
    ref, _ = o.render(w, h, 2, 8, 0, 21, 0, pixels=pix)
    assert rel_err(halves.reshape(-1, 3)[pix], ref).max() <= 1e-4


@pytest.mark.parametrize("solo,admission", [(12, P.ADMIT_FIT), (0, P.ADMIT_TAILS), (14, P.ADMIT_TAILS)])
def test_batch_scheduling_options_keep_images(gpu, solo, admission):
    """pt_render_batch's scheduling options: jobs of at least 2^solo samples rendered alone
    on the context's single-job engine beside the shared engine, and admission only once the
    running jobs have handed out their samples. Every image equals its plain per-job render
    (both integrators, two scenes)."""
    log("batch scheduling: solo %d, admission %d" % (solo, admission))
    c = ctx()
    cb = c.scene(scenes.cornell_box()).build_bvh(16)
    sp = c.scene(scenes.spheres_scene(n_spheres=8, seg=16, bands=9, terrain_n=32, env_w=64, env_h=32)).build_bvh(16)
    jobs = [P.Job(cb, 64, 64, 4, 8, 0, 5), P.Job(sp, 80, 60, 6, 12, 0, 6), P.Job(cb, 48, 40, 3, 8, 1, 7),
            P.Job(sp, 32, 32, 2, 12, 0, 9), P.Job(sp, 72, 40, 9, 10, 1, 3)]
    eng = P.Engine(c, 1 << 14, max_jobs=4)
    eng.set_tuning(solo_log2_samples=solo, admission=admission)
    assert eng.get_tuning()["solo_log2_samples"] == solo and eng.get_tuning()["admission"] == admission
    st = eng.render_batch(jobs)
    q = 0
    for j in jobs:
        single, s1 = j.bvh.render(j.width, j.height, j.spp, j.max_bounces, j.integrator, j.seed, plain=True)
        assert bits_equal(j.result, single)
        q += s1["queries"]
    assert st["queries"] == q
    eng.render_batch(jobs)  # again on the same engines
    for j in jobs:
        single, _ = j.bvh.render(j.width, j.height, j.spp, j.max_bounces, j.integrator, j.seed)
        assert bits_equal(j.result, single)
