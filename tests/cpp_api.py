"""Runs the C++ host API driver (tests/cpp/render_api.cpp, built next to the library)."""
import os
import subprocess
import tempfile

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
DRIVER = os.path.join(ROOT, "petrichor_amd", "_build", "render_api")
LOAD_DRIVER = os.path.join(ROOT, "petrichor_amd", "_build", "load_api")


def run(scene, size=None, saved=None):
    """Builds the scene through the C++ API; returns (camera state [16], image or None).
    size: (width, height, spp, depth, integrator, seed[, sampler_length]) to render. saved: a dict that
    receives the files Texture2D::Save wrote of the image ("png": uint8 [h, w, 3],
    "hdr": float32 [h, w, 3]) and the denoising AOVs Petrichor::Render filled ("albedo",
    "normal"). A scene with mat_maps also hands over tri_uv and its first texture, which
    the driver sets as the white Lambert's albedo texture and the GGX (material 4) F0
    texture and roughness map; with mat_nmaps also as that GGX's normal map, at the map
    strengths of mat_map_strength[4]."""
    n = len(scene["tri_mat"])
    with tempfile.TemporaryDirectory() as d:
        fin, fout = os.path.join(d, "in.bin"), os.path.join(d, "out.bin")
        with open(fin, "wb") as f:
            np.array([n, len(scene["light_tris"])], np.int32).tofile(f)
            np.ascontiguousarray(scene["tri_pos"], np.float32).tofile(f)
            np.ascontiguousarray(scene["tri_mat"], np.int32).tofile(f)
            np.ascontiguousarray(scene["light_tris"], np.int32).tofile(f)
            if scene.get("mat_maps") is not None:
                tex = np.ascontiguousarray(scene["textures"][0], np.float32)
                np.ascontiguousarray(scene["tri_uv"], np.float32).tofile(f)
                np.array([tex.shape[1], tex.shape[0]], np.int32).tofile(f)
                tex.tofile(f)
                if scene.get("mat_nmaps") is not None:
                    np.asarray(scene["mat_map_strength"][4], np.float32).tofile(f)
        args = [DRIVER, fin, fout] + ([str(v) for v in size] if size else [])
        r = subprocess.run(args, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True, timeout=300)
        assert r.returncode == 0, r.stdout
        data = np.fromfile(fout, np.float32)
        if size and saved is not None:
            import petrichor_amd as P
            saved["png"] = P.load_png_rgb8(fout + ".png")
            saved["hdr"] = P.load_hdr(fout + ".hdr")
    cam = data[:16]
    img = None
    if size:
        n = size[0] * size[1] * 3
        img = data[16:16 + n].reshape(size[1], size[0], 3)
        if saved is not None:
            saved["albedo"] = data[16 + n:16 + 2 * n].reshape(size[1], size[0], 3)
            saved["normal"] = data[16 + 2 * n:16 + 3 * n].reshape(size[1], size[0], 3)
            saved["uv"] = data[16 + 3 * n:16 + 4 * n].reshape(size[1], size[0], 3)
    return cam, img


def run_loaded(scene_json, settings_json, integrator, seed):
    """Loads and renders a scene file through the C++ API (Scene::LoadRenderSetting,
    Scene::LoadAssets, Petrichor::Render); returns (camera state [16], image)."""
    import json
    with open(settings_json) as f:
        rs = json.load(f)
    with tempfile.TemporaryDirectory() as d:
        fout = os.path.join(d, "out.bin")
        r = subprocess.run([LOAD_DRIVER, scene_json, settings_json, fout, str(integrator), str(seed)],
                           stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True, timeout=300)
        assert r.returncode == 0, r.stdout
        data = np.fromfile(fout, np.float32)
    return data[:16], data[16:].reshape(int(rs["outputHeight"]), int(rs["outputWidth"]), 3)
