"""The oracle's material maps (Lambert albedo, GGX reflectance and roughness maps) against
what the mathematics fixes. Texture2D.cpp cannot be built here (it needs stb and spdlog),
so the textured-material path of the reference cannot run; these checks pin the oracle:
- maps whose every texel is 1 leave the image bit-identical to the untextured scene when
  every Lambert kd is 1 (textured GetAlbedo = kd x map, untextured 1; GGX GetF0's
  reflectance x 1: Lambert.cpp:117-126, GGX.h:126-139), and change it otherwise, since the
  textured BxDF is kd x kd x map / pi;
- a grey roughness map of 0.5 gives alpha = (1 x luminance(0.5, 0.5, 0.5))^2, and the
  luminance weights sum to 1 in float, so the image equals untextured roughness 0.5
  (GGX.h:57-75, GGX.cpp:16-20);
- a texture that is linear in its column (red) and row (green) index is read back
  bilinearly as a + b x (u w) and c + d x (v h) inside the image (Texture2D.cpp:287-335):
  the first-hit albedo pass then equals that closed form at the first-hit texture
  coordinates of the UV pass, which fixes the orientation of u and v and the scale by
  width and height."""
import numpy as np
import pytest

import oracle_lib as O
import petrichor_amd as P
from petrichor_amd import scenes

ALBEDO, UV = 2, 4


def spheres(seed=7):
    sc = dict(scenes.spheres_scene(n_spheres=6, seg=12, bands=7, terrain_n=16, env_w=32, env_h=16))
    rng = np.random.default_rng(seed)
    sc["tri_uv"] = rng.uniform(-0.5, 1.5, (len(sc["tri_mat"]), 6)).astype(np.float32)
    return sc


def render(sc, integrator, w=40, h=30, spp=2, depth=6, seed=3):
    return O.OracleScene(sc).build_bvh(16).render(w, h, spp, depth, integrator, seed, 0)[0]


@pytest.mark.parametrize("integrator", [0, 1, ALBEDO])
def test_unit_maps_change_nothing(integrator):
    sc = spheres()
    m = len(sc["mat_type"])
    lambert = sc["mat_type"] == P.PT_MAT_LAMBERT
    kd_one = sc["mat_param"].copy()
    kd_one[lambert, 0:3] = 1.0
    coloured = sc
    sc = dict(sc, mat_param=kd_one)
    textured = dict(sc, textures=[np.ones((3, 5, 3), np.float32)], mat_maps=-np.ones((m, 2), np.int32))
    for i, t in enumerate(sc["mat_type"]):
        if t in (P.PT_MAT_LAMBERT, P.PT_MAT_GLOSSY):
            textured["mat_maps"][i, 0] = 0
    assert (textured["mat_maps"][:, 0] == 0).sum() >= 2
    a, b = render(sc, integrator), render(textured, integrator)
    assert np.array_equal(a, b)
    assert not np.array_equal(render(coloured, integrator), render(dict(textured, mat_param=coloured["mat_param"]), integrator))


def test_grey_roughness_map_is_that_roughness():
    lum = np.float32(0.2126) * np.float32(0.5) + np.float32(0.7152) * np.float32(0.5)
    assert lum + np.float32(0.0722) * np.float32(0.5) == np.float32(0.5)
    sc = spheres()
    mp = sc["mat_param"].copy()
    glossy = sc["mat_type"] == P.PT_MAT_GLOSSY
    assert glossy.any()
    mp[glossy, 3] = 0.5
    plain = dict(sc, mat_param=mp)
    maps = -np.ones((len(mp), 2), np.int32)
    maps[glossy, 1] = 0
    # every glossy material keeps its other roughness: the map must replace it
    mapped = dict(sc, textures=[np.full((4, 6, 3), 0.5, np.float32)], mat_maps=maps)
    a, b = render(plain, 0), render(mapped, 0)
    assert np.array_equal(a, b)
    assert not np.array_equal(render(sc, 0), a)


def test_linear_texture_reads_back_at_first_hit_uv():
    sc = dict(scenes.cornell_box())
    m = len(sc["mat_type"])
    sc["mat_type"] = np.full(m, P.PT_MAT_LAMBERT, np.int32)
    mp = np.zeros((m, 8), np.float32)
    mp[:, 0:3] = 1.0
    mp[:, 6] = 1.0
    sc["mat_param"] = mp
    sc["light_tris"] = np.zeros(0, np.int32)
    rng = np.random.default_rng(11)
    tw, th = 7, 5
    sc["tri_uv"] = rng.uniform(0.0, 0.8, (len(sc["tri_mat"]), 6)).astype(np.float32)
    i, j = np.meshgrid(np.arange(tw), np.arange(th))
    tex = np.stack([0.1 + 0.1 * i, 0.2 + 0.15 * j, np.full(i.shape, 0.3)], axis=2).astype(np.float32)
    sc["textures"] = [tex]
    sc["mat_maps"] = np.tile(np.array([[0, -1]], np.int32), (m, 1))
    albedo = render(sc, ALBEDO, spp=1, depth=1)
    uv = render(sc, UV, spp=1, depth=1)
    # the blue channel is 0.3 on every texel: a miss sees the black environment instead
    hit = albedo[..., 2] != 0.0
    assert hit.mean() > 0.5
    u, v = uv[..., 0][hit].astype(np.float64), uv[..., 1][hit].astype(np.float64)
    assert u.std() > 0.1 and v.std() > 0.1
    # the camera ray's throughput (the lens weight, below 1 off axis) scales all three
    # channels alike; the constant blue channel measures it
    lens = albedo[..., 2][hit].astype(np.float64) / 0.3
    assert lens.min() > 0.3 and lens.max() <= 1.0 + 1e-6
    np.testing.assert_allclose(albedo[..., 0][hit] / lens, 0.1 + 0.1 * u * tw, rtol=2e-5, atol=2e-6)
    np.testing.assert_allclose(albedo[..., 1][hit] / lens, 0.2 + 0.15 * v * th, rtol=2e-5, atol=2e-6)
