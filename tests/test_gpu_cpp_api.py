"""The C++ host API renders through the C-ABI: a Cornell box built with Scene, Mesh,
Triangle, Lambert/Emission and Camera and rendered by Petrichor::Render equals the
Python binding's image of the same scene bit for bit (both integrators)."""
import numpy as np
import pytest

import cpp_api
from gpu_util import bits_equal, ctx, log
from petrichor_amd import scenes

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("integrator", [0, 1])
def test_cpp_render_equals_python(gpu, integrator):
    log("C++ API render, integrator", integrator)
    sc = scenes.cornell_box()
    size = (64, 48, 4, 8, integrator, 5)
    cam, img = cpp_api.run(sc, size)
    ref, _ = ctx().scene(sc).build_bvh(16).render(64, 48, 4, 8, integrator, 5)
    assert bits_equal(cam, np.asarray(sc["camera"], np.float32))
    assert img.mean() > 0
    assert bits_equal(img, ref)


def test_cpp_texture_save_png_and_hdr(gpu):
    # Texture2D::Save: the PNG holds the oracle's tonemapped bytes of the rendered image;
    # the HDR file decodes to it within RGBE precision (1/128 of a pixel's largest channel)
    import oracle_lib as O
    log("C++ API Texture2D::Save")
    saved = {}
    _, img = cpp_api.run(scenes.cornell_box(), (40, 24, 4, 8, 0, 3), saved)
    assert np.array_equal(saved["png"], O.tonemap_rgb8(img))
    peak = img.max(axis=2, keepdims=True)
    assert np.all(np.abs(saved["hdr"] - img) <= peak / 128 + 1e-30)


def test_cpp_render_fills_denoising_aovs(gpu):
    # DenoisingAlbedo / DenoisingNormal targets set on the Scene are rendered by
    # Petrichor::Render as pt_render_aov renders them
    import petrichor_amd as P
    log("C++ API denoising AOVs")
    saved = {}
    sc = scenes.cornell_box()
    cpp_api.run(sc, (40, 24, 4, 8, 0, 3), saved)
    b = ctx().scene(sc).build_bvh(16)
    alb, _ = b.render_aov(40, 24, 4, 8, P.AOV_DENOISING_ALBEDO, 3)
    nrm, _ = b.render_aov(40, 24, 4, 8, P.AOV_DENOISING_NORMAL, 3)
    assert bits_equal(saved["albedo"], alb) and bits_equal(saved["normal"], nrm)
    assert np.abs(nrm).mean() > 0.1
    # the scene's vertices carry no texture coordinates: the UV target is all zero
    assert not np.any(saved["uv"])


def test_cpp_render_with_material_maps(gpu):
    # Vertex uv, Lambert::SetTexAlbedo and GGX::SetF0Texture / SetRoughnessMap through
    # Petrichor::Render equal the Python binding's descriptor with the same maps, and the
    # maps change the image
    log("C++ API material maps")
    sc = dict(scenes.cornell_box())
    rng = np.random.default_rng(9)
    n = len(sc["tri_mat"])
    tri_mat = np.asarray(sc["tri_mat"], np.int32).copy()
    tri_mat[(tri_mat == 0) & (np.arange(n) % 3 == 0)] = 4
    mt = np.append(np.asarray(sc["mat_type"], np.int32), scenes.PT_MAT_GLOSSY).astype(np.int32)
    mp = np.vstack([sc["mat_param"], np.array([0.9, 0.8, 0.7, 0.3, 0, 0, 1, 0], np.float32)])
    assert mt[4] == scenes.PT_MAT_GLOSSY
    maps = -np.ones((5, 2), np.int32)
    maps[0, 0] = 0
    maps[4] = (0, 0)
    sc.update(tri_mat=tri_mat, mat_type=mt, mat_param=mp, tri_uv=rng.uniform(-0.3, 1.3, (n, 6)).astype(np.float32),
              textures=[rng.uniform(0.1, 1.0, (6, 11, 3)).astype(np.float32)], mat_maps=maps)
    for integrator in (0, 1):
        _, img = cpp_api.run(sc, (48, 36, 4, 8, integrator, 5))
        ref, _ = ctx().scene(sc).build_bvh(16).render(48, 36, 4, 8, integrator, 5)
        assert bits_equal(img, ref)
        plain = dict(sc)
        plain.pop("mat_maps")
        other, _ = ctx().scene(plain).build_bvh(16).render(48, 36, 4, 8, integrator, 5)
        assert not np.array_equal(img, other)
    # GGX::SetNormalMap, SetNormalMapStrength (clamped to [0, 1]) and SetRoughnessMapStrength
    nm = dict(sc, mat_nmaps=np.array([-1, -1, -1, -1, 0], np.int32),
              mat_map_strength=np.array([[1, 1]] * 4 + [[0.35, 0.8]], np.float32))
    for integrator in (0, 1):
        _, img = cpp_api.run(nm, (48, 36, 4, 8, integrator, 5))
        ref, _ = ctx().scene(nm).build_bvh(16).render(48, 36, 4, 8, integrator, 5)
        assert bits_equal(img, ref)
        other, _ = ctx().scene(sc).build_bvh(16).render(48, 36, 4, 8, integrator, 5)
        assert not np.array_equal(img, other)
