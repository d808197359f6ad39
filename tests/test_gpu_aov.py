"""Denoising AOVs (pt_render_aov) against the oracle's restatement of
AOVDenoisingAlbedo / AOVDenoisingNormal (Core/AOV/*.cpp): same queries, every pixel
within 1e-4."""
import numpy as np
import pytest

import oracle_lib as O
import petrichor_amd as P
from gpu_util import ctx, log, rel_err
from petrichor_amd import scenes

pytestmark = pytest.mark.gpu
REL_TOL = 1e-4


def specular_spheres():
    """The small spheres scene with every GGX made perfectly smooth (alpha 0) on odd
    materials and a mix material, so that AOV paths continue through glass and smooth GGX."""
    sc = scenes.spheres_scene(n_spheres=8, seg=16, bands=9, terrain_n=32, env_w=32, env_h=16)
    sc = dict(sc)
    mt, mp = sc["mat_type"].copy(), sc["mat_param"].copy()
    for i in range(len(mt)):
        if mt[i] == P.PT_MAT_GLOSSY and i % 2 == 1:
            mp[i, 3] = 0.0
    mix = np.zeros(8, np.float32)
    mix[3], mix[4], mix[5], mix[6] = 0.5, 0, len(mt) - 1, 1
    sc["mat_type"] = np.concatenate([mt, [P.PT_MAT_MIX]]).astype(np.int32)
    sc["mat_param"] = np.concatenate([mp, mix[None]]).astype(np.float32)
    tm = sc["tri_mat"].copy()
    tm[::7] = len(mt)
    sc["tri_mat"] = tm
    return sc


CASES = {"cornell": scenes.cornell_box, "specular_spheres": specular_spheres}


@pytest.mark.parametrize("name", list(CASES))
@pytest.mark.parametrize("aov", [P.AOV_DENOISING_ALBEDO, P.AOV_DENOISING_NORMAL])
@pytest.mark.parametrize("width", [2, 4, 8])
def test_aov_matches_oracle(gpu, name, aov, width):
    sc = CASES[name]()
    b = ctx().scene(sc).build_bvh(16, width=width)
    img, st = b.render_aov(64, 48, 6, 8, aov, 5)
    o = O.OracleScene(sc).use_tree(b.export_nodes() | {"prim_ids": b.export_order()})
    ref, q = o.render(64, 48, 6, 8, 2 + aov, 5, 0)
    err = rel_err(img, ref).max(axis=2)
    exact = np.all(img.view(np.uint32) == ref.view(np.uint32), axis=2)
    log("aov %d %s width %d: queries %d / %d, bit-identical %d of %d, max rel err %g" %
        (aov, name, width, st["queries"], q, int(exact.sum()), exact.size, float(err.max())))
    assert np.abs(ref).mean() > 0.01
    if width == 2:
        assert st["queries"] == q
        assert (err <= REL_TOL).all()
    else:
        # the 8-wide walk finds the same nearest hits except on rays grazing a box
        assert exact.mean() > 0.99


@pytest.mark.parametrize("width", [2, 4, 8])
def test_uv_aov_matches_oracle(gpu, width):
    # texture coordinates per corner (some outside [0, 1]); the UV AOV averages the first
    # hit's interpolated (u, v, 0) over the camera rays
    sc = dict(specular_spheres())
    n = len(sc["tri_mat"])
    sc["tri_uv"] = np.random.default_rng(4).uniform(-0.5, 1.5, (n, 6)).astype(np.float32)
    b = ctx().scene(sc).build_bvh(16, width=width)
    img, st = b.render_aov(64, 48, 6, 8, P.AOV_UV, 5)
    o = O.OracleScene(sc).use_tree(b.export_nodes() | {"prim_ids": b.export_order()})
    ref, q = o.render(64, 48, 6, 8, 4, 5, 0)
    exact = np.all(img.view(np.uint32) == ref.view(np.uint32), axis=2)
    log("uv aov width %d: queries %d / %d, bit-identical %d of %d" % (width, st["queries"], q, int(exact.sum()),
                                                                     exact.size))
    assert st["queries"] == q == 64 * 48 * 6
    assert np.all(img[..., 2] == 0) and np.abs(ref[..., :2]).mean() > 0.1
    if width == 2:
        assert exact.all()
    else:
        assert exact.mean() > 0.99
    # a scene without texture coordinates gives zero
    b2 = ctx().scene(specular_spheres()).build_bvh(16, width=width)
    z, _ = b2.render_aov(16, 12, 2, 8, P.AOV_UV, 5)
    assert not np.any(z)


def test_aov_seeded_and_on_device(gpu):
    import torch
    b = ctx().scene(scenes.cornell_box()).build_bvh(16)
    a, _ = b.render_aov(40, 30, 4, 8, P.AOV_DENOISING_NORMAL, 3)
    out = torch.empty((30, 40, 3), dtype=torch.float32, device="cuda")
    b.render_aov(40, 30, 4, 8, P.AOV_DENOISING_NORMAL, 3, out=out)
    torch.cuda.synchronize()
    assert np.array_equal(out.cpu().numpy(), a)
    c, _ = b.render_aov(40, 30, 4, 8, P.AOV_DENOISING_NORMAL, 4)
    assert not np.array_equal(a, c)
    with pytest.raises(P.PtError):
        b.render_aov(40, 30, 4, 8, 7, 3)


def textured_spheres():
    """The small spheres scene with texture coordinates and material maps: an albedo map on
    every Lambert, a reflectance map on even and a roughness map on odd GGX materials."""
    sc = dict(scenes.spheres_scene(n_spheres=8, seg=16, bands=9, terrain_n=32, env_w=32, env_h=16))
    rng = np.random.default_rng(6)
    n, m = len(sc["tri_mat"]), len(sc["mat_type"])
    sc["tri_uv"] = rng.uniform(-0.25, 1.25, (n, 6)).astype(np.float32)
    sc["textures"] = [rng.uniform(0.05, 1.0, (9, 13, 3)).astype(np.float32),
                      rng.uniform(0.2, 1.0, (5, 7, 3)).astype(np.float32),
                      rng.uniform(0.0, 1.0, (16, 4, 3)).astype(np.float32)]
    maps = -np.ones((m, 2), np.int32)
    for i, t in enumerate(sc["mat_type"]):
        if t == P.PT_MAT_LAMBERT:
            maps[i, 0] = 0
        elif t == P.PT_MAT_GLOSSY:
            maps[i, 0 if i % 2 == 0 else 1] = 1 if i % 2 == 0 else 2
    sc["mat_maps"] = maps
    return sc


@pytest.mark.parametrize("integrator", [0, 1, "albedo"])
def test_material_maps_match_oracle(gpu, integrator):
    sc = textured_spheres()
    b = ctx().scene(sc).build_bvh(16)
    o = O.OracleScene(sc).use_tree(b.export_nodes() | {"prim_ids": b.export_order()})
    if integrator == "albedo":
        img, st = b.render_aov(64, 48, 4, 8, P.AOV_DENOISING_ALBEDO, 5)
        ref, q = o.render(64, 48, 4, 8, 2, 5, 0)
    else:
        img, st = b.render(64, 48, 4, 8, integrator, 5)
        ref, q = o.render(64, 48, 4, 8, integrator, 5, 0)
    err = rel_err(img, ref).max(axis=2)
    log("material maps, %s: queries %d / %d, max rel err %g" % (integrator, st["queries"], q, float(err.max())))
    assert st["queries"] == q
    assert (err <= REL_TOL).all()
    # the maps change the image
    plain = dict(sc)
    plain.pop("mat_maps")
    b2 = ctx().scene(plain).build_bvh(16)
    img2 = (b2.render_aov(64, 48, 4, 8, P.AOV_DENOISING_ALBEDO, 5) if integrator == "albedo"
            else b2.render(64, 48, 4, 8, integrator, 5))[0]
    assert not np.array_equal(img, img2)
