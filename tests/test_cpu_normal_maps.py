"""The oracle's GGX normal maps and map strengths (GGX.h:84-119, Triangle.cpp:105-147)
against what the mathematics fixes; Texture2D.cpp cannot be built here, so these pin the
restatement:
- GGX::GetNormal's frame and weights: texels that decode to the frame's axes give U, V
  and (1 - s) W exactly; strength 1 leaves only l.y V; random inputs equal a float64
  evaluation of (1 - s) l.x U + l.y V + (1 - s) l.z W. Mutations that fail: U and V
  swapped, v = u x w, (1 - s) on the y term, l not normalised.
- Triangle::Interpolate's tangent: uv (0,0), (1,0), (0,1) gives e1, uv (0,0), (0,1), (1,0)
  gives -e2, no or constant texture coordinates give UnitX, and a smooth triangle's
  tangent loses its component along the normal. Mutations that fail: diffUV1 and diffUV2
  swapped, the orthogonalisation dropped, the UnitX fallback dropped.
- a flat normal map (texel (0.5, 0.5, 1), strength 0) leaves a scene whose normals are
  axis-aligned bit-identical (GetNormal = normalize(normal) = normal there), on both
  integrators; a tilted one changes it.
- SetRoughnessMapStrength: a grey 0.5 roughness map at strength 1.6 renders exactly as
  roughness 0.8 (0.5 x 1.6 is exact in float). Mutation that fails: the strength
  ignored."""
import numpy as np
import pytest

import oracle_lib as O
import petrichor_amd as P
from petrichor_amd import scenes

f32 = np.float32


@pytest.mark.parametrize("s", [0.0, 0.3, 1.0])
def test_get_normal_axes(s):
    n, t = [0.0, 0.0, 1.0], [1.0, 0.0, 0.0]  # W = z, U = x, V = z x x = y
    k = f32(1.0) - f32(s)
    assert np.array_equal(O.ggx_map_normal(n, t, [1.0, 0.5, 0.5], s), [k, 0.0, 0.0])
    assert np.array_equal(O.ggx_map_normal(n, t, [0.5, 1.0, 0.5], s), [0.0, 1.0, 0.0])
    assert np.array_equal(O.ggx_map_normal(n, t, [0.5, 0.5, 1.0], s), [0.0, 0.0, k])
    # a tangent along -y: U = -y, V = z x -y = +x
    assert np.array_equal(O.ggx_map_normal(n, [0.0, -2.0, 0.0], [0.5, 1.0, 0.5], s), [1.0, 0.0, 0.0])


def test_get_normal_strength_one_keeps_only_v():
    rng = np.random.default_rng(3)
    for _ in range(50):
        t = rng.uniform(0.05, 0.95, 3)
        g = O.ggx_map_normal([0.0, 0.0, 1.0], [1.0, 0.0, 0.0], t, 1.0)
        l = (2 * t.astype(f32) - 1).astype(np.float64)
        assert g[0] == 0.0 and g[2] == 0.0
        assert abs(g[1] - l[1] / np.linalg.norm(l)) < 1e-6


def test_get_normal_matches_float64():
    rng = np.random.default_rng(5)
    for _ in range(500):
        n = rng.normal(size=3)
        t = rng.normal(size=3)
        tex = rng.uniform(0, 1, 3)
        s = rng.uniform(0, 1)
        g = O.ggx_map_normal(n, t, tex, s).astype(np.float64)
        n, t, tex = n.astype(f32).astype(np.float64), t.astype(f32).astype(np.float64), tex.astype(f32).astype(np.float64)
        w = n / np.linalg.norm(n)
        u = t / np.linalg.norm(t)
        v = np.cross(w, u)
        l = 2 * tex - 1
        l /= np.linalg.norm(l)
        k = 1 - np.float64(f32(s))
        ref = k * l[0] * u + l[1] * v + k * l[2] * w
        assert np.abs(g - ref).max() < 2e-6, (g, ref)


TRI = [0.0, 0.0, 0.0, 2.0, 0.0, 0.0, 0.0, 3.0, 0.0]
RAY = [0.5, 0.5, 1.0, 0.0, 0.0, -1.0]


def test_interpolated_tangent_closed_forms():
    n, t = O.interpolate_frame(TRI, RAY, 1.0, uv6=[0, 0, 1, 0, 0, 1])
    assert np.array_equal(n, [0.0, 0.0, 1.0])
    assert np.array_equal(t, [2.0, 0.0, 0.0])  # e1
    _, t = O.interpolate_frame(TRI, RAY, 1.0, uv6=[0, 0, 0, 1, 1, 0])
    assert np.array_equal(t, [0.0, -3.0, 0.0])  # -e2
    _, t = O.interpolate_frame(TRI, RAY, 1.0)
    assert np.array_equal(t, [1.0, 0.0, 0.0])  # UnitX
    _, t = O.interpolate_frame(TRI, RAY, 1.0, uv6=[0.3, 0.7, 0.9, 0.7, 0.1, 0.7])
    assert np.array_equal(t, [1.0, 0.0, 0.0])  # constant v: zero tangent, UnitX
    # smooth: the tangent loses its component along the interpolated normal
    nv = np.array([0.0, 1.0, 1.0]) / np.sqrt(2.0)
    n, t = O.interpolate_frame(TRI, RAY, 1.0, nrm9=np.tile(nv, 3), uv6=[0, 0, 0, 1, 1, 0])
    assert np.abs(n - nv).max() < 1e-6
    assert np.abs(t - np.array([0.0, -1.5, 1.5])).max() < 1e-5
    assert abs(float(np.dot(n, t))) < 1e-5


def floor_scene(normal_map=None, strength=(1.0, 1.0)):
    """A GGX floor (normal exactly +Z) and a Lambert back wall (normal exactly -Y) under a
    sky; texture coordinates from the position, the floor's tangent +X."""
    P3 = lambda *v: np.array(v, np.float32)
    tris = [scenes._quad(P3(-2, -2, 0), P3(2, -2, 0), P3(2, 2, 0), P3(-2, 2, 0)),
            scenes._quad(P3(-2, 2, 0), P3(2, 2, 0), P3(2, 2, 3), P3(-2, 2, 3))]
    tri = np.concatenate([np.asarray(q, np.float32).reshape(-1, 9) for q in tris])
    mats = [scenes._material(P.PT_MAT_GLOSSY, (0.9, 0.8, 0.7), 0.35), scenes._material(P.PT_MAT_LAMBERT, (0.6, 0.5, 0.4))]
    tri_mat = np.array([0, 0, 1, 1], np.int32)
    cam = scenes.camera(pos=(0.3, -5.0, 2.0), look_at=(0.0, 0.5, 0.3), focal_length=30e-3, aperture=0.0)
    sc = scenes._finish(tri, mats, tri_mat, [], cam, scenes.sky_env(64, 32), (1, 1, 1, 0.3, 1))
    p = tri.reshape(-1, 3, 3)
    sc["tri_uv"] = np.stack([p[:, :, 0] * 0.25, p[:, :, 1] * 0.25], -1).reshape(-1, 6).astype(np.float32)
    if normal_map is not None:
        sc["textures"] = [np.asarray(normal_map, np.float32)]
        sc["mat_maps"] = -np.ones((2, 2), np.int32)
        sc["mat_nmaps"] = np.array([0, -1], np.int32)
        sc["mat_map_strength"] = np.array([strength, (1.0, 1.0)], np.float32)
    return sc


def render(sc, integrator, w=40, h=30, spp=3, depth=6, seed=3):
    return O.OracleScene(sc).build_bvh(16).render(w, h, spp, depth, integrator, seed, 0)[0]


@pytest.mark.parametrize("integrator", [0, 1])
def test_flat_normal_map_at_strength_zero_changes_nothing(integrator):
    flat = np.tile(np.array([0.5, 0.5, 1.0], np.float32), (4, 4, 1))
    a = render(floor_scene(), integrator)
    b = render(floor_scene(flat, (0.0, 1.0)), integrator)
    assert np.abs(a).mean() > 0.01
    assert np.array_equal(a, b)
    tilted = np.tile(np.array([0.7, 0.5, 0.9], np.float32), (4, 4, 1))
    assert not np.array_equal(a, render(floor_scene(tilted, (0.0, 1.0)), integrator))


def test_roughness_map_strength():
    sc = dict(scenes.spheres_scene(n_spheres=6, seg=12, bands=7, terrain_n=16, env_w=32, env_h=16))
    rng = np.random.default_rng(7)
    sc["tri_uv"] = rng.uniform(-0.5, 1.5, (len(sc["tri_mat"]), 6)).astype(np.float32)
    mp = sc["mat_param"].copy()
    glossy = sc["mat_type"] == P.PT_MAT_GLOSSY
    assert glossy.any()
    mp[glossy, 3] = 0.8
    plain = dict(sc, mat_param=mp)
    maps = -np.ones((len(mp), 2), np.int32)
    maps[glossy, 1] = 0
    strength = np.ones((len(mp), 2), np.float32)
    strength[:, 1] = 1.6
    assert f32(0.5) * f32(1.6) == f32(0.8)
    mapped = dict(sc, textures=[np.full((4, 6, 3), 0.5, np.float32)], mat_maps=maps, mat_map_strength=strength)
    a, b = render(plain, 0), render(mapped, 0)
    assert np.array_equal(a, b)
    assert not np.array_equal(render(dict(mapped, mat_map_strength=None), 0), a)
