"""The C-ABI library loads and exports every declared symbol (no GPU calls), and the
integrator plumbing holds on cases with closed-form pixels."""
import ctypes
import os
import re

import numpy as np

import oracle_lib as O
import petrichor_amd as P
from petrichor_amd import scenes

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    src = open(os.path.join(ROOT, "include", "petrichor_hip.h")).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"\b(pt_[a-z_0-9]+)\s*\(", src)))


def test_library_exports_every_declared_symbol():
    L = ctypes.CDLL(P.LIB_PATH)
    names = declared_functions()
    assert len(names) >= 12
    for n in names:
        assert hasattr(L, n), n
    assert sorted(P.EXPORTED) == names


def test_host_builder_runs_without_device():
    h = P.build_bvh_host(scenes.cornell_box()["tri_pos"])
    assert len(h["leaf"]) == 23 and h["max_depth"] == 7


def _emitter_wall(le=3.0):
    """One emissive quad filling the camera's view."""
    P3 = lambda *v: np.array(v, np.float32)
    tris = scenes._quad(P3(-50, 5, -50), P3(50, 5, -50), P3(50, 5, 50), P3(-50, 5, 50))
    mats = [scenes._material(P.PT_MAT_EMISSION, (le, le * 0.5, le * 0.25))]
    cam = scenes.camera(pos=(0, 0, 0), look_at=(0, 1, 0), focal_length=35e-3, focus_target=(0, 5, 0), f_number=4)
    return scenes._finish(tris, mats, [0, 0], [], cam)


def test_primary_emission_pixel_is_camera_throughput_times_le():
    """PathTracing.cpp:62-69: a camera ray that hits an emitter adds T * Le, so with one
    sample the pixel is the camera throughput (Camera.cpp:80-85) times Le."""
    sc = _emitter_wall()
    o = O.OracleScene(sc).build_bvh(16)
    w, h = 16, 8
    for integ in (0, 1):
        img, q = o.render(w, h, 1, 8, integ, 7, 1)
        assert q == w * h
        for (x, y) in [(0, 0), (5, 3), (15, 7)]:
            pix = y * w + x
            r = O.sample_camera_ray(sc["camera"], x, y, w, h, 7)
            le = sc["mat_param"][0, :3]
            expect = (r[6:9] * le).astype(np.float32)
            assert np.array_equal(img[y, x].view(np.uint32), expect.view(np.uint32))


def test_empty_scene_pixel_is_throughput_times_base_colour():
    """A miss adds T * Environment::GetColor = T * baseColor without a texture."""
    sc = _emitter_wall()
    sc["tri_pos"] = np.zeros((0, 9), np.float32)
    sc["tri_mat"] = np.zeros(0, np.int32)
    sc["env_param"] = np.array([0.5, 0.25, 2.0, 0, 1], np.float32)
    o = O.OracleScene(sc).build_bvh(16)
    img, q = o.render(8, 8, 1, 4, 0, 3, 1)
    pix = 2 * 8 + 5
    r = O.sample_camera_ray(sc["camera"], 5, 2, 8, 8, 3)
    assert np.array_equal(img[2, 5], (r[6:9] * sc["env_param"][:3]).astype(np.float32))


def test_degenerate_env_importance_sampling_closed_form():
    """Environment::ImportanceSampling with never-loaded PDF/CDF textures (Environment.cpp:
    64-65, Texture2D.cpp:263-266): the u bisection always moves left and stops at 2^-12
    for W = 2048, so phi = 2*pi*(1 - 2^-12); theta follows the 1D CDF (j+1)/H; pdf ~ 1."""
    sc = scenes.spheres_scene(n_spheres=1, seg=8, bands=5, terrain_n=2, env_w=2048, env_h=1024)
    o = O.OracleScene(sc)
    phis, pdfs = [], []
    for s in range(200):
        d4 = o.env_sample(1000 + s, 5000 + s)
        d = d4[:3].astype(np.float64)
        phis.append(np.arctan2(d[1], d[0]) % (2 * np.pi))
        pdfs.append(d4[3])
    phi_expect = 2 * np.pi * (1 - 2.0 ** -12)
    assert np.allclose(phis, phi_expect, atol=2e-6)
    assert np.allclose(pdfs, 1.0, atol=1e-5)
    # GetImportanceSamplingPDF is ~1 everywhere
    rng = np.random.default_rng(0)
    for _ in range(50):
        v = rng.normal(size=3)
        v /= np.linalg.norm(v)
        assert abs(o.env_pdf(v.astype(np.float32)) - 1.0) < 1e-5
