"""The C++ host API (include/petrichor/Petrichor.h) on the host: the Camera set-up
operations (Core/Camera.cpp:14-50, Camera.h:71-83) must leave the same state as the
Python restatement used by every scene, bit for bit."""
import os

import numpy as np
import pytest

import cpp_api
from petrichor_amd import scenes


@pytest.mark.skipif(not os.path.exists(cpp_api.DRIVER), reason="C++ driver not built")
def test_camera_state_matches_python_restatement():
    sc = scenes.cornell_box()
    cam, _ = cpp_api.run(sc)
    assert np.array_equal(cam.view(np.uint32), np.asarray(sc["camera"], np.float32).view(np.uint32))


def test_cpp_scene_files_load_without_a_device(tmp_path):
    # Scene::LoadRenderSetting + Scene::LoadAssets (JSON, OBJ) run on the host: the driver
    # gets past loading (and renders where a GPU exists); malformed or unsupported files
    # stop at LoadAssets with the reason
    import json
    import subprocess

    import cpp_api
    from test_cpu_loader import cornell_files
    from test_cpu_png import encode
    sky = np.random.default_rng(5).integers(0, 256, (8, 16, 3), dtype=np.uint8)
    (tmp_path / "sky.png").write_bytes(encode(sky, 2))  # Environment::Load of a .png
    _, scene_json = cornell_files(tmp_path, env_texture="sky.png")
    settings = tmp_path / "settings.json"
    settings.write_text(json.dumps({"outputWidth": 16, "outputHeight": 8, "spp": 1, "maxBounces": 2}))

    def run(path):
        return subprocess.run([cpp_api.LOAD_DRIVER, str(path), str(settings), str(tmp_path / "out.bin"), "0", "1"],
                              stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True, timeout=120)

    r = run(scene_json)
    assert r.returncode == 0 or (r.returncode == 1 and "render failed" in r.stdout), r.stdout
    doc = json.loads(open(scene_json).read())
    # Texture2D::Load reads .png and .hdr paths; a .png path holding a BMP is read by its
    # header (as stbi_load reads it); another extension is reported
    import io
    from PIL import Image
    b = io.BytesIO()
    Image.fromarray(sky, "RGB").save(b, format="BMP")
    (tmp_path / "bmp_inside.png").write_bytes(b.getvalue())
    doc0 = json.loads(open(scene_json).read())
    doc0["materials"][0]["color_tex"] = "bmp_inside.png"
    (tmp_path / "bmp.json").write_text(json.dumps(doc0))
    r = run(tmp_path / "bmp.json")
    assert r.returncode == 0 or (r.returncode == 1 and "render failed" in r.stdout), r.stdout
    (tmp_path / "wood.jpg").write_bytes(b"\xff\xd8\xff")
    doc["materials"][0]["color_tex"] = "wood.jpg"
    (tmp_path / "tex.json").write_text(json.dumps(doc))
    r = run(tmp_path / "tex.json")
    assert r.returncode == 1 and "LoadAssets" in r.stdout and ".png" in r.stdout, r.stdout
    (tmp_path / "bad.json").write_text('{"materials": [ {"type": "lambert", ')
    r = run(tmp_path / "bad.json")
    assert r.returncode == 1 and "cannot read or parse" in r.stdout, r.stdout
