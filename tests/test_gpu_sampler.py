"""Renders with the micro-jittered 2D sampler (MicroJitteredSampler over Halton sequences,
Core/Sampler/MicroJitteredSampler.cpp:10-48; pt_render_desc.sampler_length) against the
oracle's restatement, on both integrators and several sequence lengths (1: every draw
is point 0 plus the offsets; 7 and 64: wraps within a path), in one piece, in a batch
beside random-sampler jobs, and through the C++ API."""
import numpy as np
import pytest

import cpp_api
import oracle_lib as O
import petrichor_amd as P
from gpu_util import bits_equal, ctx, log, rel_err
from petrichor_amd import scenes

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("length", [1, 7, 64])
@pytest.mark.parametrize("integrator", [0, 1])
def test_microjittered_render_matches_oracle(gpu, integrator, length):
    sc = scenes.spheres_scene(n_spheres=8, seg=16, bands=9, terrain_n=32, env_w=64, env_h=32)
    b = ctx().scene(sc).build_bvh(16)
    o = O.OracleScene(sc).use_tree(b.export_nodes() | {"prim_ids": b.export_order()})
    img, st = b.render(64, 48, 6, 8, integrator, 5, sampler_length=length)
    ref, q = o.render(64, 48, 6, 8, integrator, 5, 0, sampler_length=length)
    err = rel_err(img, ref)
    log("micro-jittered, integrator %d, length %d: queries %d / %d, max rel err %g" %
        (integrator, length, st["queries"], q, float(err.max())))
    assert np.abs(ref).mean() > 0.01
    assert st["queries"] == q
    assert err.max() <= 1e-4
    rnd, _ = b.render(64, 48, 6, 8, integrator, 5)
    assert not np.array_equal(img, rnd)


def test_microjittered_jobs_in_a_batch(gpu):
    c = ctx()
    cb = c.scene(scenes.cornell_box()).build_bvh(16)
    specs = [(40, 32, 5, 8, 0, 3, 16), (32, 32, 4, 8, 1, 4, 0), (24, 40, 6, 8, 0, 5, 3)]
    jobs = [P.Job(cb, w, h, spp, d, i, sd, sampler_length=sl) for (w, h, spp, d, i, sd, sl) in specs]
    eng = P.Engine(c, max_paths=1 << 14, max_jobs=4)
    eng.render_batch(jobs, iterations_per_sync=4)
    for j, (w, h, spp, d, i, sd, sl) in zip(jobs, specs):
        plain, _ = cb.render(w, h, spp, d, i, sd, sampler_length=sl)
        assert bits_equal(j.result, plain), (w, h, sl)
    eng.close()


def test_bad_sampler_length_is_rejected(gpu):
    b = ctx().scene(scenes.cornell_box()).build_bvh(16)
    for bad in (-1, P.MAX_SAMPLER_LENGTH + 1):
        with pytest.raises(P.PtError):
            b.render(8, 8, 1, 4, 0, 1, sampler_length=bad)


def test_cpp_api_microjittered_render(gpu):
    sc = scenes.cornell_box()
    for integrator in (0, 1):
        _, img = cpp_api.run(sc, (48, 36, 4, 8, integrator, 5, 16))
        ref, _ = ctx().scene(sc).build_bvh(16).render(48, 36, 4, 8, integrator, 5, sampler_length=16)
        assert bits_equal(img, ref)
