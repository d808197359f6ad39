"""GGX normal maps and map strengths (GGX::GetNormal, SetNormalMapStrength,
SetRoughnessMapStrength: Core/Material/GGX.h:84-119, with Triangle::Interpolate's tangent,
Core/Geometry/Triangle.cpp:105-147) on the GPU against the oracle: renders on both
integrators, the denoising albedo and normal AOVs, flat and smooth triangles, texture
coordinates that give every tangent case, normal-map strengths 0, 0.4 and 1 and a
roughness-map strength; within the 1e-4 bar with equal query counts."""
import numpy as np
import pytest

import oracle_lib as O
import petrichor_amd as P
from gpu_util import ctx, log, rel_err
from petrichor_amd import scenes
from test_cpu_normal_maps import floor_scene

pytestmark = pytest.mark.gpu

REL_TOL = 1e-4


def normal_mapped_spheres():
    """The small spheres scene (smooth spheres, flat terrain) with texture coordinates (some
    triangles without a v gradient, for the UnitX tangent), a bumpy normal map on every GGX
    material at strengths 0, 0.4 and 1 in turn, and a roughness map at strength 0.7 on
    every third GGX."""
    sc = dict(scenes.spheres_scene(n_spheres=8, seg=16, bands=9, terrain_n=32, env_w=32, env_h=16))
    rng = np.random.default_rng(11)
    n, m = len(sc["tri_mat"]), len(sc["mat_type"])
    uv = rng.uniform(-0.25, 1.25, (n, 6)).astype(np.float32)
    flat_v = rng.uniform(0, 1, n) < 0.1
    uv[flat_v, 3] = uv[flat_v, 1]
    uv[flat_v, 5] = uv[flat_v, 1]
    sc["tri_uv"] = uv
    bump = np.empty((24, 32, 3), np.float32)
    bump[..., 0:2] = rng.uniform(0.3, 0.7, (24, 32, 2))
    bump[..., 2] = rng.uniform(0.75, 1.0, (24, 32))
    sc["textures"] = [bump, rng.uniform(0.1, 1.0, (7, 5, 3)).astype(np.float32)]
    maps = -np.ones((m, 2), np.int32)
    nmaps = -np.ones(m, np.int32)
    strength = np.ones((m, 2), np.float32)
    g = 0
    for i, t in enumerate(sc["mat_type"]):
        if t == P.PT_MAT_GLOSSY:
            nmaps[i] = 0
            strength[i, 0] = (0.0, 0.4, 1.0)[g % 3]
            if g % 3 == 2:
                maps[i, 1] = 1
                strength[i, 1] = 0.7
            g += 1
    assert g >= 3
    sc["mat_maps"], sc["mat_nmaps"], sc["mat_map_strength"] = maps, nmaps, strength
    return sc


CASES = {"spheres": normal_mapped_spheres,
         "floor": lambda: floor_scene(np.random.default_rng(2).uniform(0.3, 1.0, (8, 8, 3)).astype(np.float32),
                                      (0.4, 1.0))}


@pytest.mark.parametrize("name", list(CASES))
@pytest.mark.parametrize("mode", [0, 1, "albedo", "normal"])
def test_normal_maps_match_oracle(gpu, name, mode):
    sc = CASES[name]()
    b = ctx().scene(sc).build_bvh(16)
    o = O.OracleScene(sc).use_tree(b.export_nodes() | {"prim_ids": b.export_order()})
    if mode in ("albedo", "normal"):
        aov = P.AOV_DENOISING_ALBEDO if mode == "albedo" else P.AOV_DENOISING_NORMAL
        img, st = b.render_aov(64, 48, 4, 8, aov, 5)
        ref, q = o.render(64, 48, 4, 8, 2 + aov, 5, 0)
    else:
        img, st = b.render(64, 48, 4, 8, mode, 5)
        ref, q = o.render(64, 48, 4, 8, mode, 5, 0)
    err = rel_err(img, ref).max(axis=2)
    exact = np.all(img.view(np.uint32) == ref.view(np.uint32), axis=2)
    log("normal maps %s, %s: queries %d / %d, bit-identical %d of %d, max rel err %g" %
        (name, mode, st["queries"], q, int(exact.sum()), exact.size, float(err.max())))
    assert np.abs(ref).mean() > 0.01
    assert st["queries"] == q
    assert (err <= REL_TOL).all()
    if mode in (0, 1):
        # the normal maps change the image
        plain = dict(sc, mat_nmaps=None)
        img2 = ctx().scene(plain).build_bvh(16).render(64, 48, 4, 8, mode, 5)[0]
        assert not np.array_equal(img, img2)


def test_bad_strengths_are_rejected(gpu):
    sc = normal_mapped_spheres()
    st = sc["mat_map_strength"].copy()
    st[0, 0] = 1.5
    with pytest.raises(P.PtError):
        ctx().scene(dict(sc, mat_map_strength=st))
    nm = sc["mat_nmaps"].copy()
    nm[0] = 9
    with pytest.raises(P.PtError):
        ctx().scene(dict(sc, mat_nmaps=nm))
