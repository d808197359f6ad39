"""A scene read from the reference's file formats (JSON scene, OBJ meshes with vertex
normals, a run-length Radiance environment map) renders on the GPU as the oracle renders
the same loaded scene."""
import json

import numpy as np
import pytest

import oracle_lib as O
import petrichor_amd as P
from gpu_util import ctx, log, rel_err
from petrichor_amd import loader
import cpp_api
from gpu_util import bits_equal
from test_cpu_loader import cornell_files, write_rle_hdr

pytestmark = pytest.mark.gpu
REL_TOL = 1e-4


def sphere_obj(path, cx, cy, cz, r, nu=16, nv=10):
    """A UV sphere with per-vertex normals given at twice unit length (the loader
    normalises them) and quads for the middle bands."""
    lines, idx = [], {}
    for j in range(nv + 1):
        th = np.pi * j / nv
        for i in range(nu):
            ph = 2 * np.pi * i / nu
            d = np.array([np.sin(th) * np.cos(ph), np.sin(th) * np.sin(ph), np.cos(th)])
            idx[(i, j)] = len(idx) + 1
            lines.append("v %r %r %r" % tuple(float(x) for x in (cx + r * d[0], cy + r * d[1], cz + r * d[2])))
            lines.append("vn %r %r %r" % tuple(float(x) for x in 2 * d))
    for j in range(nv):
        for i in range(nu):
            a, b = idx[(i, j)], idx[((i + 1) % nu, j)]
            c, d = idx[((i + 1) % nu, j + 1)], idx[(i, j + 1)]
            lines.append("f %d//%d %d//%d %d//%d" % (a, a, d, d, c, c))
            lines.append("f %d//%d %d//%d %d//%d" % (a, a, c, c, b, b))
    path.write_text("\n".join(lines) + "\n")


def loaded_scene(tmp_path):
    h, w = 32, 64
    yy, xx = np.mgrid[0:h, 0:w]
    sky = np.stack([0.2 + xx / w, 0.3 + 0.5 * yy / h, 0.8 - 0.3 * yy / h], -1).astype(np.float32)
    sky[4:6, 10:12] = 40.0  # a small bright sun
    P.save_hdr(str(tmp_path / "flat.hdr"), sky)
    raw = open(tmp_path / "flat.hdr", "rb").read()
    rgbe = np.frombuffer(raw[raw.index(b"+X %d\n" % w) + len(b"+X %d\n" % w):], np.uint8).reshape(h, w, 4)
    write_rle_hdr(tmp_path / "sky.hdr", rgbe)
    _, path = cornell_files(tmp_path, env_texture="sky.hdr")
    sphere_obj(tmp_path / "ball.obj", 0.35, 0.4, 0.25, 0.2)
    doc = json.loads(open(path).read())
    doc["env"]["base_color"] = [1, 1, 1]
    doc["assets"].append({"path": "ball.obj", "mat": "mx"})
    (tmp_path / "scene.json").write_text(json.dumps(doc))
    return loader.load_scene(str(tmp_path / "scene.json"))


def test_cpp_loaded_scene_equals_python_loaded(gpu, tmp_path):
    # the same files through Scene::LoadRenderSetting + Scene::LoadAssets + Petrichor::Render
    sc = loaded_scene(tmp_path)
    (tmp_path / "settings.json").write_text(json.dumps({"outputWidth": 48, "outputHeight": 32, "spp": 4,
                                                         "maxBounces": 8}))
    log("C++ loaded scene")
    cam, img = cpp_api.run_loaded(str(tmp_path / "scene.json"), str(tmp_path / "settings.json"), 0, 9)
    ref, _ = ctx().scene(sc).build_bvh(16).render(48, 32, 4, 8, 0, 9)
    assert bits_equal(cam, sc["camera"])
    assert img.mean() > 0.01
    assert bits_equal(img, ref)


@pytest.mark.parametrize("integrator", [0, 1])
def test_loaded_scene_renders_as_the_oracle(gpu, tmp_path, integrator):
    sc = loaded_scene(tmp_path)
    log("loaded scene: %d triangles, env %s, integrator %d" % (len(sc["tri_mat"]), sc["env_rgb"].shape, integrator))
    b = ctx().scene(sc).build_bvh(16)
    img, stats = b.render(64, 48, 8, 8, integrator, 3)
    o = O.OracleScene(sc).use_tree(b.export_nodes() | {"prim_ids": b.export_order()})
    ref, q = o.render(64, 48, 8, 8, integrator, 3, 0)
    err = rel_err(img, ref).max(axis=2)
    log("  max rel err", float(err.max()), "queries gpu/oracle", stats["queries"], q)
    assert ref.mean() > 0.01
    assert stats["queries"] == q
    assert (err <= REL_TOL).all()
