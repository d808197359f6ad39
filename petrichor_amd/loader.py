"""Scene ingestion: the reference's JSON scene and render-setting files and Wavefront OBJ
meshes, turned into the scene dict of pt_scene_desc (see scenes.py).

- load_scene(path) restates SceneLoaderJson::Load (Core/SceneLoader.cpp:17-296) with
  Scene::LoadModel (Core/Scene.cpp:28-51) and Mesh::Load (Core/Geometry/Mesh.cpp:18-125).
- load_render_setting(path) restates RenderSettingLoaderJson::Load
  (Core/RenderSettingLoader.cpp:15-64).
- parse_obj(text) calls pt_obj_parse (csrc/pt_obj.cpp), the restatement of what
  Assimp 4.1.0 (the reference's pinned dependency, conanfile.txt:2; not present here)
  hands Mesh::Load for an OBJ file read with no post-processing flags: one output vertex
  per face corner, faces in file order, no triangulation, Assimp's number parsing.
- Textures are read as Texture2D::Load reads them (Core/Texture2D.cpp:107-195), by the
  path's extension: a .hdr with load_hdr_file (Radiance RGBE, flat or run-length
  scanlines, as stbi_loadf decodes them, no gamma); a .png with decode_image
  (pt_image_decode), which, as stbi_load, tells the format by the file's header (PNG,
  BMP, TGA, binary PNM); any other extension is a file Texture2D::Load refuses.

Behaviour kept from the reference:
- "emission" materials are never registered (Core/SceneLoader.cpp:219-221), so a
  loaded scene has no area lights;
- the camera takes SetFNumber before SetFocalLength (Core/SceneLoader.cpp:247-257): the
  aperture comes from the default 55 mm lens;
- a polygon face of more than three corners keeps its first three (Mesh::Load reads
  mIndices[0..2]);
- a vertex without a normal has +Z (Core/Geometry/Vertex.h:19);
- every loaded triangle is smooth-shaded.

Material textures are read as Core/SceneLoader.cpp:106-189 reads them: a Lambert's and a
GGX's "color_tex" from the scene's folder as colour textures, a GGX's "roughness_tex"
from the path as written (relative to the working directory) as a non-colour texture. A
file that does not exist leaves the material untextured, and a texture registered under a
name already taken reuses the first one (Core/Scene.h:131-143).

Where the reference would dereference a missing material (an unknown material name, a
mix of unknown materials, a point or line element), or logs an error and goes on without
the texture (a texture path that is not .png or .hdr, an unreadable file, or a .png path
holding a JPEG, GIF or PSD image, whose pixels depend on stb's own decoder), these
functions raise ValueError instead.
"""
import json
import os

import numpy as np

from . import PT_MAT_GLASS, PT_MAT_GLOSSY, PT_MAT_LAMBERT, PT_MAT_MIX
from . import scenes

f32 = np.float32

# Core/RenderSetting.h:12-21
RENDER_SETTING_DEFAULTS = {"outputWidth": 1280, "outputHeight": 720, "spp": 128, "denoising_spp": 128,
                           "maxBounces": 16, "tileWidth": 16, "tileHeight": 16, "numThreads": 0}


def load_render_setting(path):
    """RenderSettingLoaderJson::Load: the keys present override the defaults (a number is
    converted to int by truncation, as static_cast<int> of a JSON value). Returns a dict
    with the JSON key names."""
    with open(path) as f:
        data = json.load(f)
    out = dict(RENDER_SETTING_DEFAULTS)
    for k in out:
        if k in data:
            out[k] = int(data[k])
    return out


def _vec3(obj, key):
    """loadVector3f (Core/SceneLoader.cpp:34-58): missing -> zero; up to three floats."""
    v = np.zeros(3, np.float32)
    if key in obj:
        vals = [f32(x) for x in obj[key]]
        for i in range(min(3, len(vals))):
            v[i] = vals[i]
    return v


def _normalized(v):
    """Vector3f::Normalized: v / sqrt(dot(v, v)) in float (Math/Vector3f.h:170-189)."""
    return scenes._normalized(v)


def parse_obj(text):
    """Triangles of an OBJ text as Mesh::Load builds them from Assimp's import
    (pt_obj_parse): returns (positions [F, 9], normals [F, 9], uvs [F, 6]) float32, one row
    per face (its first three corners), normals normalised per corner (+Z where the face
    gives none), texture coordinates per corner (zero where it gives none). Raises
    ValueError on malformed text."""
    import ctypes
    from . import lib, _ptr, _f, _l
    data = text.encode() if isinstance(text, str) else bytes(text)
    L = lib()
    n = np.zeros(1, np.int64)
    err = ctypes.create_string_buffer(256)
    rc = L.pt_obj_parse(data, len(data), 0, None, None, None, _ptr(n, _l), err, 256)
    if rc != 0:
        raise ValueError(err.value.decode())
    p = np.zeros((int(n[0]), 9), np.float32)
    nr = np.zeros((int(n[0]), 9), np.float32)
    uv = np.zeros((int(n[0]), 6), np.float32)
    rc = L.pt_obj_parse(data, len(data), int(n[0]), _ptr(p, _f), _ptr(nr, _f), _ptr(uv, _f), _ptr(n, _l), err, 256)
    if rc != 0:
        raise ValueError(err.value.decode())
    return p, nr, uv


def _decode_ldr(fn, data, color):
    import ctypes
    from . import lib, _ptr, _f, _i
    data = bytes(data)
    f = getattr(lib(), fn)
    w = np.zeros(1, np.int32)
    h = np.zeros(1, np.int32)
    err = ctypes.create_string_buffer(256)
    if f(data, len(data), int(color), _ptr(w, _i), _ptr(h, _i), None, err, 256) != 0:
        raise ValueError(err.value.decode())
    out = np.zeros((int(h[0]), int(w[0]), 3), np.float32)
    if f(data, len(data), int(color), _ptr(w, _i), _ptr(h, _i), _ptr(out, _f), err, 256) != 0:
        raise ValueError(err.value.decode())
    return out


def decode_png(data, color=True):
    """A PNG file's pixels as Texture2D::Load reads them (pt_png_decode): [H, W, 3] float32,
    bytes / 255, then powf(x, 2.2) for a colour texture."""
    return _decode_ldr("pt_png_decode", data, color)


def decode_image(data, color=True):
    """The pixels of the file at a .png path as Texture2D::Load reads them
    (pt_image_decode: PNG, BMP, TGA or binary PNM, told by the header, as stbi_load does):
    [H, W, 3] float32, bytes / 255, then powf(x, 2.2) for a colour texture."""
    return _decode_ldr("pt_image_decode", data, color)


def load_texture(path, color=True):
    """Texture2D::Load (Core/Texture2D.cpp:107-195) by the path's extension: .png through
    decode_image, .hdr through load_hdr_file (no gamma); ValueError for anything else."""
    if path.endswith(".png"):
        with open(path, "rb") as f:
            return decode_image(f.read(), color=color)
    if path.endswith(".hdr"):
        return load_hdr_file(path)
    raise ValueError("%s: Texture2D::Load reads .png and .hdr paths only" % path)


def load_obj(path):
    with open(path, "rb") as f:
        return parse_obj(f.read())


def _decode_rgbe(e):
    """stbi__hdr_convert: value = byte * 2^(exponent - 136), zero when the exponent is 0."""
    e = np.asarray(e, np.uint8)
    scale = np.where(e[..., 3:] > 0, np.ldexp(np.float32(1.0), e[..., 3:].astype(np.int32) - 136), 0)
    return (e[..., :3].astype(np.float32) * scale.astype(np.float32)).astype(np.float32)


def load_hdr_file(path):
    """A Radiance RGBE file as [H, W, 3] float32, as stbi_loadf reads it: header lines up to
    an empty line, '-Y H +X W', then flat pixels or new-style run-length scanlines."""
    with open(path, "rb") as f:
        data = f.read()
    if not (data.startswith(b"#?RADIANCE\n") or data.startswith(b"#?RGBE\n")):
        raise ValueError("%s: not a Radiance file" % path)
    pos = data.index(b"\n") + 1
    fmt_ok = False
    while True:
        end = data.index(b"\n", pos)
        line = data[pos:end]
        pos = end + 1
        if not line:
            break
        if line == b"FORMAT=32-bit_rle_rgbe":
            fmt_ok = True
    if not fmt_ok:
        raise ValueError("%s: unsupported Radiance format" % path)
    end = data.index(b"\n", pos)
    parts = data[pos:end].split()
    pos = end + 1
    if len(parts) != 4 or parts[0] != b"-Y" or parts[2] != b"+X":
        raise ValueError("%s: unsupported orientation" % path)
    h, w = int(parts[1]), int(parts[3])
    buf = np.frombuffer(data, np.uint8, offset=pos)
    if w < 8 or w >= 32768 or not (buf[0] == 2 and buf[1] == 2 and not buf[2] & 0x80):
        return _decode_rgbe(buf[:h * w * 4].reshape(h, w, 4))
    out = np.zeros((h, w, 4), np.uint8)
    i = 0
    for y in range(h):
        if buf[i] != 2 or buf[i + 1] != 2 or ((int(buf[i + 2]) << 8) | int(buf[i + 3])) != w:
            raise ValueError("%s: bad run-length scanline %d" % (path, y))
        i += 4
        for ch in range(4):
            x = 0
            while x < w:
                count = int(buf[i])
                i += 1
                if count > 128:
                    count -= 128
                    out[y, x:x + count, ch] = buf[i]
                    i += 1
                else:
                    out[y, x:x + count, ch] = buf[i:i + count]
                    i += count
                x += count
    return _decode_rgbe(out)


def load_scene(path):
    """A JSON scene file (materials, camera, env, assets) as a scene dict for
    Context.scene(): SceneLoaderJson::Load with Scene::LoadModel and Mesh::Load."""
    base = os.path.dirname(os.path.abspath(path))
    with open(path) as f:
        data = json.load(f)
    # ---- materials (Core/SceneLoader.cpp:81-240): "default" first, then in file order ----
    names, types, params = {}, [], []

    textures, tex_names, maps = [], {}, []

    def texture(path, key, color, name):
        # Texture2D::Load, then Scene::RegisterTexture (the first texture of a name stays)
        if not os.path.exists(path):
            return -1
        try:
            rgb = load_texture(path, color=color)
        except ValueError as e:
            raise ValueError("material '%s': %s" % (name, e))
        if key not in tex_names:
            tex_names[key] = len(textures)
            textures.append(rgb)
        return tex_names[key]

    def register(name, mtype, color, p3=0.0, sub=(0, 0), mmap=(-1, -1)):
        p = np.zeros(8, np.float32)
        p[0:3] = color
        p[3] = p3
        p[4], p[5] = sub
        p[6] = 1.0  # importance sampling on (the materials' default)
        if name in names:  # Scene::RegisterMaterial replaces an entry of the same name
            types[names[name]], params[names[name]], maps[names[name]] = mtype, p, mmap
        else:
            names[name] = len(types)
            types.append(mtype)
            params.append(p)
            maps.append(mmap)

    def material(name):
        if name not in names:
            raise ValueError("material '%s' is not registered" % name)
        return names[name]

    register("default", PT_MAT_LAMBERT, np.ones(3, np.float32))
    for m in data.get("materials", []):
        mtype, name = m.get("type", ""), m.get("name", "")
        if not mtype or not name:
            continue
        ctex, rtex = m.get("color_tex") or "", m.get("roughness_tex") or ""
        if mtype == "lambert":
            t = texture(os.path.join(base, ctex), os.path.join(base, ctex), True, name) if ctex else -1
            register(name, PT_MAT_LAMBERT, _vec3(m, "base_color"), mmap=(t, -1))
        elif mtype == "glass":
            register(name, PT_MAT_GLASS, _vec3(m, "color"), f32(m.get("ior", 1.0)))
        elif mtype == "glossy":
            t0 = texture(os.path.join(base, ctex), ctex, True, name) if ctex else -1
            t1 = texture(rtex, rtex, False, name) if rtex else -1
            register(name, PT_MAT_GLOSSY, _vec3(m, "color"), f32(m.get("roughness", 0.0)), mmap=(t0, t1))
        elif mtype == "mix":
            i0, i1 = material(m.get("mat0_name", "")), material(m.get("mat1_name", ""))
            register(name, PT_MAT_MIX, np.zeros(3, np.float32), f32(m.get("mix", 0.5)), (i0, i1))
        # "emission" registers nothing; other types are reported and skipped
    # ---- camera (Core/SceneLoader.cpp:230-260) ----
    cd = data.get("camera", {})
    position, look_at, focus = _vec3(cd, "position"), _vec3(cd, "look_at"), _vec3(cd, "focus_pos")
    f_number = f32(cd.get("f_number", f32(2.8)))
    focal_length = f32(cd.get("focal_length", f32(55e-3)))
    cam = scenes.camera(position, look_at, focal_length=None, focus_target=focus, f_number=f_number)
    cam[13] = f32(f32(24e-3) / focal_length)  # SetFocalLength after SetFNumber
    # ---- env (Core/SceneLoader.cpp:262-282) ----
    ed = data.get("env", {})
    env_param = np.zeros(5, np.float32)
    env_param[0:3] = _vec3(ed, "base_color")
    env_param[4] = 1.0
    env_rgb = None
    if ed.get("texture"):
        tex = os.path.join(base, ed["texture"])
        # Environment::Load reads its texture as a colour texture (Core/Environment.cpp:22)
        env_rgb = load_texture(tex, color=True)
    # ---- assets (Core/SceneLoader.cpp:284-300): smooth triangles of each OBJ ----
    tri_pos, tri_nrm, tri_uv, tri_mat = [], [], [], []
    for a in data.get("assets", []):
        mat = material(a["mat"]) if "mat" in a else names["default"]
        p, n, uv = load_obj(os.path.join(base, a["path"]))
        tri_pos.append(p)
        tri_nrm.append(n)
        tri_uv.append(uv)
        tri_mat.append(np.full(len(p), mat, np.int32))
    uv_all = np.concatenate(tri_uv) if tri_uv else np.zeros((0, 6), np.float32)
    tp = np.concatenate(tri_pos) if tri_pos else np.zeros((0, 9), np.float32)
    return {
        "tri_pos": tp,
        "tri_mat": np.concatenate(tri_mat) if tri_mat else np.zeros(0, np.int32),
        "tri_nrm": np.concatenate(tri_nrm) if tri_nrm else None,
        "tri_smooth": np.ones(len(tp), np.int32) if tri_pos else None,
        "tri_uv": uv_all if np.any(uv_all != 0) else None,
        "mat_type": np.asarray(types, np.int32),
        "mat_param": np.stack(params).astype(np.float32),
        "light_tris": np.zeros(0, np.int32),
        "camera": cam,
        "env_rgb": env_rgb,
        "env_param": env_param,
        **({"textures": textures, "mat_maps": np.asarray(maps, np.int32).reshape(-1, 2)}
           if any(t >= 0 for mm in maps for t in mm) else {}),
    }
