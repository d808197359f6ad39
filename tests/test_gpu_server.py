"""Render server: continuous batching with resident scenes, an HBM budget and
progressive results.

A request's image must not depend on what runs beside it: every path draws from its
own per-pixel sampler streams, so the image of a request served alone, served in a mix
of other requests, and rendered by the plain pt_render entry are bit-identical (the
1e-4 per-pixel bar holds a fortiori). The plain image is itself pinned against the
oracle in test_gpu_parity.py.
"""
import time

import numpy as np
import pytest

import petrichor_amd as P
from gpu_util import bits_equal, ctx, log
from petrichor_amd import scenes

pytestmark = pytest.mark.gpu

GiB = 1 << 30
WAIT_MS = 120000


def small_spheres(seed=2):
    return scenes.spheres_scene(seed=seed, n_spheres=6, seg=16, bands=9, terrain_n=24, env_w=128, env_h=64)


def plain(sc, w, h, spp, depth, integrator, seed):
    img, _ = ctx().scene(sc).build_bvh(16).render(w, h, spp, depth, integrator, seed)
    return img


def done(srv, rid):
    info = srv.wait(rid, WAIT_MS)
    assert info["state"] == P.REQ_DONE, info
    return info


@pytest.mark.parametrize("solo", [0, 15])
def test_request_alone_equals_mix_and_plain(gpu, solo):
    """A request's image alone, in a mix of requests and through pt_render are the same;
    with solo = 15 the requests of at least 2^15 samples run one at a time on the server's
    single-job engine beside the shared pool."""
    log("server: alone vs mix, solo %d" % solo)
    srv = P.Server(8 * GiB, max_paths=1 << 20, max_jobs=16, iterations_per_sync=4, solo_log2_samples=solo,
                   solo_paths=1 << 20)
    cb, sp = scenes.cornell_box(), small_spheres()
    a_id, b_id = srv.add_scene(cb), srv.add_scene(sp)
    req_a = (128, 96, 8, 8, P.INTEGRATOR_MIS, 3)
    ra = srv.submit(a_id, *req_a)
    done(srv, ra)
    alone = srv.fetch(ra)
    mix = [srv.submit(b_id, 160, 90, 16, 12, P.INTEGRATOR_MIS, 5),
           srv.submit(a_id, 64, 64, 4, 8, P.INTEGRATOR_SIMPLE, 9),
           srv.submit(a_id, *req_a),
           srv.submit(b_id, 200, 100, 8, 12, P.INTEGRATOR_MIS, 7)]
    for r in mix:
        done(srv, r)
    imgs = [srv.fetch(r) for r in mix]
    assert bits_equal(imgs[2], alone)
    assert bits_equal(alone, plain(cb, *req_a))
    assert bits_equal(imgs[0], plain(sp, 160, 90, 16, 12, P.INTEGRATOR_MIS, 5))
    assert bits_equal(imgs[1], plain(cb, 64, 64, 4, 8, P.INTEGRATOR_SIMPLE, 9))
    st = srv.stats()
    log("  stats", st)
    assert st["completed"] == 5 and st["queries"] > 0 and st["running"] == 0 and st["queued"] == 0
    srv.close()


@pytest.mark.parametrize("solo", [0, 20])
def test_requests_join_a_running_one_and_previews_progress(gpu, solo):
    log("server: continuous admission and progressive results, solo %d" % solo)
    srv = P.Server(8 * GiB, max_paths=1 << 20, max_jobs=16, iterations_per_sync=2, solo_log2_samples=solo,
                   solo_paths=1 << 19)
    sp = small_spheres(4)
    sid = srv.add_scene(sp)
    long_req = (512, 384, 2048, 12, P.INTEGRATOR_MIS, 11)  # 400M samples on half the pool
    rl = srv.submit(sid, *long_req)
    t0 = time.time()
    while srv.poll(rl)["state"] == P.REQ_QUEUED and time.time() - t0 < 60:
        time.sleep(0.001)
    assert srv.poll(rl)["state"] == P.REQ_RUNNING
    rs = srv.submit(sid, 32, 32, 2, 4, P.INTEGRATOR_MIS, 12)  # arrives while the long one runs
    done(srv, rs)
    info_l = srv.poll(rl)
    log("  short done, long state", info_l)
    assert info_l["state"] == P.REQ_RUNNING, "the short request must finish while the long one is on the device"
    prev, progress = srv.preview(rl)
    log("  preview progress %.3f" % progress)
    assert 0.0 < progress < 1.0
    assert np.isfinite(prev).all() and prev.min() >= 0 and prev.max() > 0
    done(srv, rl)
    final = srv.fetch(rl)
    assert bits_equal(final, plain(sp, *long_req))
    srv.close()


def test_budget_admission_and_eviction(gpu):
    log("server: budget and eviction")
    # three scenes of equal size (the same geometry uploaded three times)
    sc = [scenes.random_triangles(20000, seed=1, mu=-4.0)] * 3
    probe = P.Server(4 * GiB, max_paths=65536, max_jobs=8)
    pool = probe.stats()["bytes_in_use"]
    s0 = probe.add_scene(sc[0])
    actual = probe.stats()["bytes_in_use"] - pool
    probe.close()
    estimate = 20000 * 320 + 64 + (1 << 20)  # the reservation made while a scene is built
    budget = pool + 2 * actual + estimate - 1
    srv = P.Server(budget, max_paths=65536, max_jobs=8)
    ids = [srv.add_scene(sc[0]), srv.add_scene(sc[1])]
    r = srv.submit(ids[1], 64, 64, 2, 4)  # keeps scene 1 busy until it is done
    done(srv, r)
    srv.fetch(r)
    ids.append(srv.add_scene(sc[2]))  # evicts scene 0, the least recently used idle scene
    st = srv.stats()
    log("  stats", st)
    assert st["evicted_scenes"] == 1 and not srv.resident(ids[0]) and srv.resident(ids[1]) and srv.resident(ids[2])
    assert st["bytes_in_use"] <= budget
    with pytest.raises(P.PtError):
        srv.submit(ids[0], 16, 16, 1, 2)  # evicted
    with pytest.raises(P.PtError):
        srv.submit(ids[1], 512, 512, 64, 2)  # its image and samples (12 B each) exceed the budget
    with pytest.raises(P.PtError):
        srv.add_scene(scenes.random_triangles(400000, seed=9))  # larger than the whole budget
    r = srv.submit(ids[2], 48, 48, 2, 4, P.INTEGRATOR_MIS, 2)
    done(srv, r)
    assert bits_equal(srv.fetch(r), plain(sc[2], 48, 48, 2, 4, P.INTEGRATOR_MIS, 2))
    srv.remove_scene(ids[2])
    assert not srv.resident(ids[2])
    srv.close()


def test_scene_textures_are_charged_to_the_budget(gpu):
    """A scene's material-map textures (12 bytes per texel) and texture coordinates count
    against the server's budget like its triangles do."""
    log("server: texture bytes in the budget")
    base = small_spheres(3)
    tex = dict(base)
    n, m = len(base["tri_mat"]), len(base["mat_type"])
    tex["tri_uv"] = np.zeros((n, 6), np.float32)
    tex["textures"] = [np.full((1024, 2048, 3), 0.5, np.float32)]  # 24 MiB
    maps = -np.ones((m, 2), np.int32)
    maps[:, 0] = [0 if t == P.PT_MAT_LAMBERT else -1 for t in base["mat_type"]]
    tex["mat_maps"] = maps
    srv = P.Server(8 * GiB, max_paths=1 << 16, max_jobs=4)
    b0 = srv.stats()["bytes_in_use"]
    srv.add_scene(base)
    b1 = srv.stats()["bytes_in_use"]
    srv.add_scene(tex)
    b2 = srv.stats()["bytes_in_use"]
    srv.close()
    assert (b2 - b1) - (b1 - b0) >= 12 * 1024 * 2048 + 24 * n
