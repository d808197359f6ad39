"""Scene ingestion (petrichor_amd/loader.py) on the CPU: the OBJ restatement, the
render-setting and scene JSON files, and Radiance HDR environment maps."""
import json
import os

import numpy as np
import pytest

import petrichor_amd as P
from petrichor_amd import loader, scenes

f32 = np.float32


def test_obj_corners_indices_and_normals():
    quad = """# a quad, a triangle, relative indices, another object
v 0 0 0
v 1 0 0
v 1 1 0
v 0 1 0
vt 0 0
f 1 2 3 4
f -4 -3 -2
o second
usemtl other
f 2/1 3/1 4/1
"""
    p, n, uv = loader.parse_obj(quad)
    assert p.shape == (3, 9)
    # the quad keeps its first three corners (no triangulation)
    assert np.array_equal(p[0], [0, 0, 0, 1, 0, 0, 1, 1, 0])
    assert np.array_equal(p[1], p[0])  # -4 -3 -2 are vertices 1 2 3
    assert np.array_equal(p[2], [1, 0, 0, 1, 1, 0, 0, 1, 0])
    # no normal indices: the Vertex default +Z
    assert np.array_equal(n, np.tile(np.array([0, 0, 1], f32), (3, 3)))
    # texture coordinates where the face gives them (the last face: vt 1 on every corner)
    assert np.array_equal(uv[:2], np.zeros((2, 6), f32)) and np.array_equal(uv[2], np.zeros(6, f32))
    p, _, uv = loader.parse_obj("v 0 0 0\nv 1 0 0\nv 1 1 0\nvt 0.25 0.5\nvt 1 0.75 0\nf 1/1 2/2 3/-1\nf 1 2/1 3\n")
    assert np.array_equal(uv, np.array([[0.25, 0.5, 1, 0.75, 1, 0.75], [0, 0, 0.25, 0.5, 0, 0]], f32))
    with_normals = "v 0 0 0\nv 1 0 0\nv 1 1 0\nvn 0 0 2\nvn 3 4 0\nf 1//1 2//2 3//-2\n"
    p, n, _ = loader.parse_obj(with_normals)
    # normalised per corner: v / sqrt(dot(v, v))
    assert np.array_equal(n[0], np.array([0, 0, 1, 0.6, 0.8, 0, 0, 0, 1], f32))
    # homogeneous and coloured vertices
    p, _, _ = loader.parse_obj("v 2 4 6 2\nv 1 0 0 0.5 0.5 0.5\nv 0 1 0\nf 1 2 3\n")
    assert np.array_equal(p[0], [1, 2, 3, 1, 0, 0, 0, 1, 0])
    for bad in ("v 0 0 0\nv 1 0 0\nl 1 2\n", "v 0 0 0\nf 1 2 3\n", "v 0 0 0\nv 1 0 0\nf 1 2\n",
                "v 0 0 0\nv 1 0 0\nv 1 1 0\nvn 0 0 1\nf 1 2 3\nf 1//1 2//1 3//1\n", "v 0 0 x\n"):
        with pytest.raises(ValueError):
            loader.parse_obj(bad)


def test_obj_numbers_are_read_as_assimp_reads_them():
    # fast_atof: float(integer part) + float(digits * 10^-n in double), then * powf(10, e)
    rng = np.random.default_rng(8)
    ints = rng.integers(0, 5000, 3000)
    fracs = rng.integers(0, 10 ** 7, 3000)
    texts = ["%d.%07d" % (i, fr) for i, fr in zip(ints, fracs)]
    want = [f32(f32(i) + f32(float(fr) * 1e-7)) for i, fr in zip(ints, fracs)]
    texts += ["-0.5", "1e2", "2.5E-3", ".25", "7.", "3,5", "1.0000000000000000001"]
    want += [f32(-0.5), f32(100), f32(f32(2.5) * np.power(f32(10), f32(-3))), f32(0.25), f32(7), f32(3.5), f32(1)]
    lines = ["v %s 0 0" % t for t in texts] + ["f 1 2 3"]
    p, _, _ = loader.parse_obj("\n".join(lines) + "\n")
    assert p.shape == (1, 9)
    # read the numbers back through the three corners of as many faces as needed
    faces = "\n".join("f %d %d %d" % (i + 1, i + 1, i + 1) for i in range(len(texts)))
    p, _, _ = loader.parse_obj("\n".join(lines[:-1]) + "\n" + faces + "\n")
    got = p[:, 0]
    assert np.array_equal(got.view(np.uint32), np.array(want, f32).view(np.uint32))
    # and the rounding differs from a correctly rounded parse on some of them
    assert any(f32(float(t)) != w for t, w in zip(texts[:3000], want[:3000]))


def test_render_setting_file(tmp_path):
    path = tmp_path / "rs.json"
    path.write_text(json.dumps({"outputWidth": 640, "spp": 32.9, "maxBounces": 8}))
    rs = loader.load_render_setting(str(path))
    assert rs["outputWidth"] == 640 and rs["spp"] == 32 and rs["maxBounces"] == 8
    assert rs["outputHeight"] == 720 and rs["tileWidth"] == 16 and rs["denoising_spp"] == 128


def write_rle_hdr(path, rgbe):
    """New-style run-length Radiance scanlines (runs of equal bytes, else literals)."""
    h, w, _ = rgbe.shape
    out = bytearray(b"#?RADIANCE\nFORMAT=32-bit_rle_rgbe\n\n-Y %d +X %d\n" % (h, w))
    for y in range(h):
        out += bytes([2, 2, w >> 8, w & 255])
        for ch in range(4):
            row = rgbe[y, :, ch]
            x = 0
            while x < w:
                run = 1
                while x + run < w and run < 127 and row[x + run] == row[x]:
                    run += 1
                if run >= 3:
                    out += bytes([128 + run, int(row[x])])
                    x += run
                else:
                    n = 1
                    while x + n < w and n < 128 and not (x + n + 2 < w and row[x + n] == row[x + n + 1] == row[x + n + 2]):
                        n += 1
                    out += bytes([n]) + bytes(row[x:x + n].tolist())
                    x += n
    path.write_bytes(bytes(out))


def test_hdr_flat_and_run_length(tmp_path):
    rng = np.random.default_rng(4)
    img = (rng.random((6, 40, 3)) * 4).astype(f32)
    img[2, 5:30] = 0.25  # runs
    img[0, 0] = 0
    P.save_hdr(str(tmp_path / "flat.hdr"), img)
    flat = loader.load_hdr_file(str(tmp_path / "flat.hdr"))
    assert np.all(np.abs(flat - img) <= img.max(axis=2, keepdims=True) / 128 + 1e-30)
    # the same RGBE bytes, run-length encoded, decode to the same floats
    raw = open(tmp_path / "flat.hdr", "rb").read()
    rgbe = np.frombuffer(raw[raw.index(b"+X 40\n") + 6:], np.uint8).reshape(6, 40, 4)
    write_rle_hdr(tmp_path / "rle.hdr", rgbe)
    assert np.array_equal(loader.load_hdr_file(str(tmp_path / "rle.hdr")), flat)
    # decoding: byte * 2^(e - 136)
    assert flat[2, 5, 0] == f32(0.25)


def cornell_files(d, env_texture=None):
    """The Cornell box's walls and boxes as OBJ files (one per material) and a JSON scene."""
    sc = scenes.cornell_box()
    names = ["white", "red", "green", "light"]
    assets = []
    for m in range(3):
        tris = sc["tri_pos"][sc["tri_mat"] == m]
        lines = []
        for t in tris:
            for k in range(3):
                lines.append("v %r %r %r" % tuple(float(x) for x in t[3 * k:3 * k + 3]))
        for i in range(len(tris)):
            lines.append("f %d %d %d" % (3 * i + 1, 3 * i + 2, 3 * i + 3))
        (d / ("%s.obj" % names[m])).write_text("\n".join(lines) + "\n")
        assets.append({"path": "%s.obj" % names[m], "mat": names[m]})
    doc = {
        "materials": [
            {"type": "lambert", "name": "white", "base_color": [0.73, 0.73, 0.73]},
            {"type": "lambert", "name": "red", "base_color": [0.65, 0.05, 0.05]},
            {"type": "lambert", "name": "green", "base_color": [0.05, 0.65, 0.05]},
            {"type": "emission", "name": "light", "color": [15, 15, 15]},
            {"type": "glass", "name": "g", "color": [1, 1, 1], "ior": 1.5},
            {"type": "glossy", "name": "gl", "color": [0.9, 0.8, 0.7], "roughness": 0.3},
            {"type": "mix", "name": "mx", "mat0_name": "red", "mat1_name": "gl", "mix": 0.25},
        ],
        "camera": {"position": [0.5, -1.35, 0.5], "look_at": [0.5, 0.5, 0.5], "focus_pos": [0.5, 0.5, 0.3],
                   "f_number": 8.0, "focal_length": 35e-3},
        "env": {"base_color": [0.6, 0.7, 0.8]},
        "assets": assets,
    }
    if env_texture:
        doc["env"]["texture"] = env_texture
    (d / "scene.json").write_text(json.dumps(doc))
    return sc, str(d / "scene.json")


def test_scene_file(tmp_path):
    sc, path = cornell_files(tmp_path)
    got = loader.load_scene(path)
    # triangles: each material's OBJ in asset order, smooth with +Z normals
    want_pos = np.concatenate([sc["tri_pos"][sc["tri_mat"] == m] for m in range(3)])
    assert np.array_equal(got["tri_pos"], want_pos)
    assert np.array_equal(got["tri_smooth"], np.ones(len(want_pos), np.int32))
    assert np.array_equal(got["tri_nrm"][:, 2::3], np.ones((len(want_pos), 3), f32))
    # materials: "default" first; emission registers nothing
    assert list(got["mat_type"]) == [P.PT_MAT_LAMBERT] * 4 + [P.PT_MAT_GLASS, P.PT_MAT_GLOSSY, P.PT_MAT_MIX]
    assert np.array_equal(got["mat_param"][0, :3], [1, 1, 1])
    assert np.array_equal(got["mat_param"][1, :3], np.array([0.73, 0.73, 0.73], f32))
    assert got["mat_param"][4, 3] == f32(1.5) and got["mat_param"][5, 3] == f32(0.3)
    assert got["mat_param"][6, 3] == f32(0.25) and list(got["mat_param"][6, 4:6]) == [2, 5]
    assert np.array_equal(got["tri_mat"], np.repeat([1, 2, 3], [np.sum(sc["tri_mat"] == m) for m in range(3)]))
    assert len(got["light_tris"]) == 0
    # camera: SetFNumber with the default 55 mm lens, then SetFocalLength
    cam = got["camera"]
    ref = scenes.camera((0.5, -1.35, 0.5), (0.5, 0.5, 0.5), focal_length=35e-3, focus_target=(0.5, 0.5, 0.3),
                        f_number=8.0)
    assert np.array_equal(cam[:13], ref[:13]) and cam[13] == ref[13]
    h_default = f32(f32(24e-3) / f32(55e-3))
    assert cam[14] == f32(f32(f32(24e-3) / h_default) / f32(8.0))
    assert np.array_equal(got["env_param"], np.array([0.6, 0.7, 0.8, 0, 1], f32))


def test_scene_file_errors(tmp_path):
    _, path = cornell_files(tmp_path)
    doc = json.loads(open(path).read())
    doc["assets"][0]["mat"] = "light"  # an emission name was never registered
    (tmp_path / "bad.json").write_text(json.dumps(doc))
    with pytest.raises(ValueError, match="light"):
        loader.load_scene(str(tmp_path / "bad.json"))
    doc = json.loads(open(path).read())
    # a missing texture file leaves the material untextured; an existing non-PNG one is reported
    doc["materials"][0]["color_tex"] = "wood.png"
    (tmp_path / "tex.json").write_text(json.dumps(doc))
    assert "mat_maps" not in loader.load_scene(str(tmp_path / "tex.json"))
    (tmp_path / "wood.jpg").write_bytes(b"\xff\xd8\xff")
    doc["materials"][0]["color_tex"] = "wood.jpg"
    (tmp_path / "tex.json").write_text(json.dumps(doc))
    with pytest.raises(ValueError, match=".png"):
        loader.load_scene(str(tmp_path / "tex.json"))
