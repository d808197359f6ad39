"""petrichor_amd — MI355X-native rendering hot path of the Petrichor path tracer.

Python host layer over the C-ABI of include/petrichor_hip.h (libpetrichor_hip.so,
built in-tree by petrichor_amd.build). It mirrors LibPetrichor Core's objects:
a Scene (Core/Scene.h) is uploaded once, a BVH is built over it
(Core/Accel/BinnedSAHBVH.cpp), and render() replaces Core::Petrichor::Render
(Core/Petrichor.cpp:34-102). Device buffers for the trace entry points are torch
tensors on the GPU (torch is plumbing here: memory and streams).

The library fails loudly: there is no CPU fallback. If the shared library is
missing, importing works but every entry point raises.
"""
import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libpetrichor_hip.so")

PT_MAT_LAMBERT, PT_MAT_GLASS, PT_MAT_EMISSION, PT_MAT_GLOSSY, PT_MAT_MIX = 0, 1, 2, 3, 4
INTEGRATOR_MIS, INTEGRATOR_SIMPLE = 0, 1

_f = ctypes.POINTER(ctypes.c_float)
_i = ctypes.POINTER(ctypes.c_int32)
_l = ctypes.POINTER(ctypes.c_int64)


class PtError(RuntimeError):
    pass


class SceneDesc(ctypes.Structure):
    _fields_ = [
        ("num_triangles", ctypes.c_int32), ("positions", _f), ("normals", _f), ("material_ids", _i),
        ("smooth", _i), ("num_materials", ctypes.c_int32), ("material_types", _i), ("material_params", _f),
        ("num_lights", ctypes.c_int32), ("light_triangles", _i), ("camera", ctypes.c_float * 16),
        ("env_width", ctypes.c_int32), ("env_height", ctypes.c_int32), ("env_rgb", _f),
        ("env_base_color", ctypes.c_float * 3), ("env_z_rotation", ctypes.c_float),
        ("env_importance_sampling", ctypes.c_int32), ("uvs", _f),
        ("num_textures", ctypes.c_int32), ("texture_sizes", _i), ("textures", ctypes.POINTER(_f)),
        ("material_maps", _i), ("material_normal_maps", _i), ("material_map_strengths", _f),
        ("num_spheres", ctypes.c_int32), ("spheres", _f), ("sphere_material_ids", _i),
    ]


BUILD_DEVICE = 0
BUILD_HOST = 1
BUILD_HOST_REFERENCE_ORDER = 2  # the reference's sort at every node: its leaf order among equal keys


class BvhDesc(ctypes.Structure):
    _fields_ = [("num_bins", ctypes.c_int32), ("leaf_max", ctypes.c_int32), ("width", ctypes.c_int32),
                ("builder", ctypes.c_int32)]


class RenderDesc(ctypes.Structure):
    _fields_ = [("width", ctypes.c_int32), ("height", ctypes.c_int32), ("spp", ctypes.c_int32),
                ("max_bounces", ctypes.c_int32), ("integrator", ctypes.c_int32), ("seed", ctypes.c_uint32),
                ("flags", ctypes.c_int32), ("sampler_length", ctypes.c_int32)]


class JobDesc(ctypes.Structure):
    _fields_ = [("scene", ctypes.c_void_p), ("bvh", ctypes.c_void_p), ("render", RenderDesc),
                ("rgb_out", ctypes.c_void_p), ("out_on_device", ctypes.c_int32)]


class BatchStats(ctypes.Structure):
    _fields_ = [("queries", ctypes.c_int64), ("iterations", ctypes.c_int64), ("elapsed_ms", ctypes.c_double),
                ("wall_ms", ctypes.c_double), ("syncs", ctypes.c_int64)]


class Tuning(ctypes.Structure):
    _fields_ = [("iterations_per_sync", ctypes.c_int32), ("trace_blocks_per_cu", ctypes.c_int32),
                ("trace_chunk", ctypes.c_int32), ("shade_blocks_per_cu", ctypes.c_int32),
                ("admission", ctypes.c_int32), ("solo_log2_samples", ctypes.c_int32),
                ("threads_per_ray", ctypes.c_int32), ("leaf_batch", ctypes.c_int32)]


ADMIT_FIT = 0    # pt_render_batch admits any pending job whose slots are free
ADMIT_TAILS = 1  # ... only once every running job has handed out all its samples


RENDER_TIME_KERNELS = 1
RENDER_PLAIN_LAUNCHES = 2
RENDER_TRACE_STATS = 4
RENDER_TWO_STREAMS = 8
RENDER_LAUNCH_LOG = 16
RENDER_THREE_STREAMS = 32
MAX_SAMPLER_LENGTH = 65536  # PT_MAX_SAMPLER_LENGTH
AOV_DENOISING_ALBEDO, AOV_DENOISING_NORMAL, AOV_UV = 0, 1, 2


class ServerDesc(ctypes.Structure):
    _fields_ = [("device", ctypes.c_int32), ("max_jobs", ctypes.c_int32), ("hbm_budget_bytes", ctypes.c_int64),
                ("max_paths", ctypes.c_int64), ("iterations_per_sync", ctypes.c_int32),
                ("solo_log2_samples", ctypes.c_int32), ("solo_paths", ctypes.c_int64)]


class RequestInfo(ctypes.Structure):
    _fields_ = [("state", ctypes.c_int32), ("progress", ctypes.c_float), ("queue_ms", ctypes.c_double),
                ("latency_ms", ctypes.c_double)]


class ServerStats(ctypes.Structure):
    _fields_ = [(k, ctypes.c_int64) for k in ("submitted", "completed", "running", "queued", "resident_scenes",
                                             "evicted_scenes", "bytes_in_use", "budget_bytes", "queries",
                                             "iterations", "syncs")] + [("busy_ms", ctypes.c_double)]


REQ_QUEUED, REQ_RUNNING, REQ_DONE, REQ_FAILED = 0, 1, 2, 3


class RenderStats(ctypes.Structure):
    _fields_ = [("queries", ctypes.c_int64), ("iterations", ctypes.c_int64), ("elapsed_ms", ctypes.c_double),
                ("kernel_items", ctypes.c_int64 * 4), ("kernel_launches", ctypes.c_int64 * 4),
                ("kernel_ms", ctypes.c_double * 4), ("kernel_nodes", ctypes.c_int64 * 4),
                ("kernel_tris", ctypes.c_int64 * 4), ("log_launches", ctypes.c_int64), ("log_items", ctypes.c_int64),
                ("log_trace_ms", ctypes.c_double), ("log_nodes", ctypes.c_int64), ("log_tris", ctypes.c_int64),
                ("finish_launches", ctypes.c_int64), ("finish_rays", ctypes.c_int64), ("finish_ms", ctypes.c_double)]


_lib = None


def lib():
    """Loads libpetrichor_hip.so; raises if it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise PtError("libpetrichor_hip.so is not built (run python -m petrichor_amd.build)")
    L = ctypes.CDLL(LIB_PATH)
    vp = ctypes.c_void_p
    pvp = ctypes.POINTER(ctypes.c_void_p)
    L.pt_ctx_create.argtypes = [ctypes.c_int, pvp]
    L.pt_ctx_destroy.argtypes = [vp]
    L.pt_last_error.restype = ctypes.c_char_p
    L.pt_last_error.argtypes = [vp]
    L.pt_scene_create.argtypes = [vp, ctypes.POINTER(SceneDesc), pvp]
    L.pt_scene_destroy.argtypes = [vp]
    L.pt_bvh_build.argtypes = [vp, ctypes.POINTER(BvhDesc), pvp]
    L.pt_bvh_destroy.argtypes = [vp]
    L.pt_bvh_export_order.argtypes = [vp, _i]
    L.pt_bvh_info.argtypes = [vp, _i, _i, ctypes.POINTER(ctypes.c_double)]
    L.pt_bvh_export_nodes.argtypes = [vp, _f, _i, _i, _i]
    L.pt_bvh_build_host.argtypes = [_f, ctypes.c_int32, ctypes.POINTER(BvhDesc), _i, _f, _i, _i, _i, _i]
    L.pt_trace_closest.argtypes = [vp, ctypes.c_int64, vp, vp, vp, vp]
    L.pt_trace_any.argtypes = [vp, ctypes.c_int64, vp, vp, vp]
    L.pt_trace_closest_noalloc.argtypes = [vp, ctypes.c_int64, vp, vp, vp, vp]
    L.pt_bvh_set_trace_tuning.argtypes = [vp, ctypes.POINTER(Tuning)]
    L.pt_trace_closest_soa.argtypes = [vp, ctypes.c_int64, ctypes.POINTER(RaysSoa), ctypes.POINTER(HitsSoa), vp]
    L.pt_trace_any_soa.argtypes = [vp, ctypes.c_int64, ctypes.POINTER(RaysSoa), vp, vp]
    L.pt_trace_stats.argtypes = [vp, ctypes.c_int64, vp, vp, vp, _l]
    L.pt_render.argtypes = [vp, vp, ctypes.POINTER(RenderDesc), vp, ctypes.c_int32, ctypes.POINTER(RenderStats)]
    L.pt_render_aov.argtypes = [vp, vp, ctypes.c_int32, ctypes.POINTER(RenderDesc), vp, ctypes.c_int32,
                                ctypes.POINTER(RenderStats)]
    L.pt_engine_create.argtypes = [vp, ctypes.c_int64, ctypes.c_int32, ctypes.c_int32, pvp]
    L.pt_engine_destroy.argtypes = [vp]
    L.pt_render_batch.argtypes = [vp, ctypes.c_int32, ctypes.POINTER(JobDesc), ctypes.c_int32,
                                  ctypes.POINTER(BatchStats)]
    L.pt_engine_set_tuning.argtypes = [vp, ctypes.POINTER(Tuning)]
    L.pt_engine_get_tuning.argtypes = [vp, ctypes.POINTER(Tuning)]
    i64p = ctypes.POINTER(ctypes.c_int64)
    L.pt_server_create.argtypes = [ctypes.POINTER(ServerDesc), pvp]
    L.pt_server_destroy.argtypes = [vp]
    L.pt_server_last_error.restype = ctypes.c_char_p
    L.pt_server_last_error.argtypes = []
    L.pt_server_add_scene.argtypes = [vp, ctypes.POINTER(SceneDesc), ctypes.POINTER(BvhDesc), i64p]
    L.pt_server_remove_scene.argtypes = [vp, ctypes.c_int64]
    L.pt_server_scene_resident.argtypes = [vp, ctypes.c_int64]
    L.pt_server_submit.argtypes = [vp, ctypes.c_int64, ctypes.POINTER(RenderDesc), i64p]
    L.pt_server_poll.argtypes = [vp, ctypes.c_int64, ctypes.POINTER(RequestInfo)]
    L.pt_server_wait.argtypes = [vp, ctypes.c_int64, ctypes.c_int32, ctypes.POINTER(RequestInfo)]
    L.pt_server_preview.argtypes = [vp, ctypes.c_int64, _f, ctypes.POINTER(ctypes.c_float)]
    L.pt_server_fetch.argtypes = [vp, ctypes.c_int64, _f]
    L.pt_server_get_stats.argtypes = [vp, ctypes.POINTER(ServerStats)]
    L.pt_tonemap_rgb8.argtypes = [vp, ctypes.c_int64, vp, vp, ctypes.c_int32, vp]
    L.pt_obj_parse.argtypes = [ctypes.c_char_p, ctypes.c_int64, ctypes.c_int64, _f, _f, _f, _l, ctypes.c_char_p,
                               ctypes.c_int32]
    L.pt_png_decode.argtypes = [ctypes.c_char_p, ctypes.c_int64, ctypes.c_int32, _i, _i, _f, ctypes.c_char_p,
                                ctypes.c_int32]
    L.pt_image_decode.argtypes = [ctypes.c_char_p, ctypes.c_int64, ctypes.c_int32, _i, _i, _f, ctypes.c_char_p,
                                  ctypes.c_int32]
    _lib = L
    return L


# every symbol include/petrichor_hip.h declares
EXPORTED = ["pt_ctx_create", "pt_ctx_destroy", "pt_last_error", "pt_scene_create", "pt_scene_destroy",
            "pt_bvh_build", "pt_bvh_destroy", "pt_bvh_export_order", "pt_bvh_export_nodes", "pt_bvh_build_host", "pt_bvh_info", "pt_trace_closest", "pt_trace_closest_host",
            "pt_trace_any", "pt_trace_closest_noalloc", "pt_trace_closest_soa", "pt_trace_any_soa",
            "pt_bvh_set_trace_tuning", "pt_trace_stats", "pt_render", "pt_engine_create", "pt_engine_destroy",
            "pt_render_batch", "pt_engine_set_tuning", "pt_engine_get_tuning", "pt_server_create", "pt_server_destroy", "pt_server_last_error",
            "pt_server_add_scene", "pt_server_remove_scene", "pt_server_scene_resident", "pt_server_submit",
            "pt_server_poll", "pt_server_wait", "pt_server_preview", "pt_server_fetch", "pt_server_get_stats",
            "pt_tonemap_rgb8", "pt_obj_parse", "pt_render_aov", "pt_png_decode", "pt_image_decode"]


class RaysSoa(ctypes.Structure):
    _fields_ = [(k, ctypes.c_void_p) for k in ("ox", "oy", "oz", "dx", "dy", "dz", "dist_min", "dist_max")]


class HitsSoa(ctypes.Structure):
    _fields_ = [(k, ctypes.c_void_p) for k in ("t", "prim", "u", "v")]


def _ptr(a, t):
    return ctypes.cast(None, t) if a is None else a.ctypes.data_as(t)


def build_bvh_host(tri_pos, bins=16, leaf_max=4, builder=BUILD_HOST):
    """Binary BVH built on the host only (pt_bvh_build_host): the reference's node table
    and leaf order, without a device. builder=BUILD_HOST_REFERENCE_ORDER also keeps the
    reference's order among equal centroid keys."""
    tri_pos = np.ascontiguousarray(tri_pos, np.float32).reshape(-1, 9)
    n = tri_pos.shape[0]
    cap = max(2 * n - 1, 1)
    order = np.zeros(n, np.int32)
    boxes = np.zeros((cap, 6), np.float32)
    child = np.zeros((cap, 2), np.int32)
    rng = np.zeros((cap, 2), np.int32)
    leaf = np.zeros(cap, np.int32)
    depth = ctypes.c_int32()
    d = BvhDesc(bins, leaf_max, 2, builder)
    nn = lib().pt_bvh_build_host(_ptr(tri_pos, _f), n, ctypes.byref(d), _ptr(order, _i), _ptr(boxes, _f),
                                 _ptr(child, _i), _ptr(rng, _i), _ptr(leaf, _i), ctypes.byref(depth))
    if nn < 0:
        raise PtError("pt_bvh_build_host failed (%d)" % nn)
    return {"prim_ids": order, "boxes": boxes[:nn], "child": child[:nn], "range": rng[:nn], "leaf": leaf[:nn],
            "max_depth": depth.value}


def save_png(path, rgb8):
    """Writes an [H, W, 3] uint8 image as an 8-bit RGB PNG (the file Texture2D::Save
    writes for a .png path, Core/Texture2D.cpp:217-222, without stb): one zlib stream of
    unfiltered rows."""
    import struct
    import zlib
    a = np.ascontiguousarray(np.asarray(rgb8, dtype=np.uint8))
    assert a.ndim == 3 and a.shape[2] == 3
    h, w = a.shape[:2]
    raw = np.concatenate([np.zeros((h, 1), np.uint8), a.reshape(h, w * 3)], axis=1).tobytes()

    def chunk(kind, data):
        return struct.pack(">I", len(data)) + kind + data + struct.pack(">I", zlib.crc32(kind + data) & 0xFFFFFFFF)

    with open(path, "wb") as f:
        f.write(b"\x89PNG\r\n\x1a\n")
        f.write(chunk(b"IHDR", struct.pack(">IIBBBBB", w, h, 8, 2, 0, 0, 0)))
        f.write(chunk(b"IDAT", zlib.compress(raw, 6)))
        f.write(chunk(b"IEND", b""))


def load_png_rgb8(path):
    """Reads back an 8-bit RGB PNG whose rows are unfiltered, as save_png writes them."""
    import struct
    import zlib
    data = open(path, "rb").read()
    assert data[:8] == b"\x89PNG\r\n\x1a\n"
    pos, idat, w, h = 8, b"", 0, 0
    while pos < len(data):
        n = struct.unpack(">I", data[pos:pos + 4])[0]
        kind, body = data[pos + 4:pos + 8], data[pos + 8:pos + 8 + n]
        if kind == b"IHDR":
            w, h, depth, ctype = struct.unpack(">IIBB", body[:10])
            assert depth == 8 and ctype == 2
        elif kind == b"IDAT":
            idat += body
        pos += 12 + n
    rows = np.frombuffer(zlib.decompress(idat), np.uint8).reshape(h, 1 + 3 * w)
    if np.any(rows[:, 0] != 0):
        raise ValueError("filtered PNG rows are not supported")
    return rows[:, 1:].reshape(h, w, 3).copy()


def save_hdr(path, rgb):
    """Writes an [H, W, 3] float image as a Radiance RGBE file (the format Texture2D::Save
    writes for a .hdr path, Core/Texture2D.cpp:225-243, without stb): flat scanlines,
    each pixel's shared exponent from frexp of its largest component (Ward's encoding)."""
    a = np.ascontiguousarray(np.asarray(rgb, dtype=np.float32))
    assert a.ndim == 3 and a.shape[2] == 3
    h, w = a.shape[:2]
    mx = a.max(axis=2)
    mant, ex = np.frexp(mx)
    ok = mx >= np.float32(1e-32)
    scale = np.where(ok, (mant.astype(np.float32) * np.float32(256.0)) / np.where(ok, mx, 1), 0).astype(np.float32)
    rgbe = np.zeros((h, w, 4), np.uint8)
    rgbe[..., :3] = np.where(ok[..., None], (a * scale[..., None]), 0).astype(np.uint8)
    rgbe[..., 3] = np.where(ok, ex + 128, 0).astype(np.uint8)
    with open(path, "wb") as f:
        f.write(b"#?RADIANCE\nFORMAT=32-bit_rle_rgbe\n\n-Y %d +X %d\n" % (h, w))
        f.write(rgbe.tobytes())


def load_hdr(path):
    """Reads a flat (not run-length encoded) Radiance RGBE file as [H, W, 3] floats."""
    data = open(path, "rb").read()
    head, _, rest = data.partition(b"\n\n")
    line, _, pix = rest.partition(b"\n")
    parts = line.split()
    h, w = int(parts[1]), int(parts[3])
    e = np.frombuffer(pix, np.uint8, count=h * w * 4).reshape(h, w, 4)
    f = np.where(e[..., 3:] > 0, np.ldexp(1.0, e[..., 3:].astype(np.int32) - (128 + 8)), 0.0)
    return (e[..., :3] * f).astype(np.float32)


class Context:
    """One device + one HIP stream (pt_ctx_create)."""

    def __init__(self, device=0):
        self._lib = lib()
        h = ctypes.c_void_p()
        rc = self._lib.pt_ctx_create(device, ctypes.byref(h))
        if rc != 0:
            raise PtError("pt_ctx_create failed (%d)" % rc)
        self.handle = h

    def check(self, rc, what):
        if rc != 0:
            msg = self._lib.pt_last_error(self.handle)
            raise PtError("%s failed (%d): %s" % (what, rc, msg.decode() if msg else ""))

    def scene(self, scene):
        return Scene(self, scene)

    def tonemap_rgb8(self, rgb, out=None, stream=None):
        """The reference's PNG bytes of an RGB float image (pt_tonemap_rgb8, the pixel
        loop of Texture2D::Save): rgb is a float32 CUDA tensor [..., 3]; returns a uint8
        tensor of the same shape on the same device."""
        import torch
        assert rgb.is_cuda and rgb.dtype == torch.float32 and rgb.is_contiguous() and rgb.shape[-1] == 3
        if out is None:
            out = torch.empty(rgb.shape, dtype=torch.uint8, device=rgb.device)
        assert out.dtype == torch.uint8 and out.is_contiguous() and out.numel() == rgb.numel()
        st = ctypes.c_void_p(stream) if stream else None
        self.check(self._lib.pt_tonemap_rgb8(self.handle, rgb.numel() // 3, rgb.data_ptr(), out.data_ptr(), 1, st),
                   "pt_tonemap_rgb8")
        return out

    def close(self):
        if self.handle:
            self._lib.pt_ctx_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Scene:
    """A scene uploaded to the device (pt_scene_create). `scene` is a dict as produced
    by petrichor_amd.scenes."""

    def __init__(self, ctx, scene):
        self.ctx = ctx
        self.src = scene
        L = ctx._lib
        d, self._keep = scene_desc(scene)
        h = ctypes.c_void_p()
        ctx.check(L.pt_scene_create(ctx.handle, ctypes.byref(d), ctypes.byref(h)), "pt_scene_create")
        self.handle = h
        self.num_triangles = d.num_triangles
        self.num_primitives = d.num_triangles + d.num_spheres  # triangles, then spheres
        self._keep = []  # the library copied everything

    def build_bvh(self, bins=16, leaf_max=4, width=2, builder=BUILD_DEVICE):
        return Bvh(self, bins, leaf_max, width, builder)

    def close(self):
        if getattr(self, "handle", None):
            self.ctx._lib.pt_scene_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def scene_desc(scene):
    """pt_scene_desc of a scene dict (petrichor_amd.scenes); returns (desc, arrays that
    must stay alive while the descriptor is used)."""
    keep = []

    def k(a, dt):
        a = np.ascontiguousarray(a, dtype=dt)
        keep.append(a)
        return a

    d = SceneDesc()
    tri_pos = k(scene["tri_pos"], np.float32).reshape(-1, 9)
    d.num_triangles = tri_pos.shape[0]
    d.positions = _ptr(tri_pos, _f)
    d.normals = _ptr(k(scene["tri_nrm"], np.float32), _f) if scene.get("tri_nrm") is not None else None
    d.material_ids = _ptr(k(scene["tri_mat"], np.int32), _i)
    d.smooth = _ptr(k(scene["tri_smooth"], np.int32), _i) if scene.get("tri_smooth") is not None else None
    mt = k(scene["mat_type"], np.int32)
    d.num_materials = len(mt)
    d.material_types = _ptr(mt, _i)
    d.material_params = _ptr(k(scene["mat_param"], np.float32), _f)
    lt = k(scene.get("light_tris", np.zeros(0, np.int32)), np.int32)
    d.num_lights = len(lt)
    d.light_triangles = _ptr(lt, _i)
    cam = np.asarray(scene["camera"], np.float32)
    for i in range(16):
        d.camera[i] = float(cam[i])
    env = scene.get("env_rgb")
    if env is not None:
        env = k(env, np.float32)
        d.env_height, d.env_width = env.shape[0], env.shape[1]
        d.env_rgb = _ptr(env, _f)
    else:
        d.env_width = d.env_height = 0
        d.env_rgb = None
    ep = np.asarray(scene["env_param"], np.float32)
    for i in range(3):
        d.env_base_color[i] = float(ep[i])
    d.env_z_rotation = float(ep[3])
    d.env_importance_sampling = int(ep[4] != 0)
    d.uvs = _ptr(k(scene["tri_uv"], np.float32), _f) if scene.get("tri_uv") is not None else None
    if scene.get("mat_maps") is not None or scene.get("mat_nmaps") is not None:
        texs = [k(t, np.float32) for t in scene.get("textures", [])]
        sizes = k(np.array([[t.shape[1], t.shape[0]] for t in texs], np.int32).reshape(-1), np.int32)
        ptrs = (_f * max(len(texs), 1))(*[_ptr(t, _f) for t in texs])
        keep.append(ptrs)
        d.num_textures = len(texs)
        d.texture_sizes = _ptr(sizes, _i)
        d.textures = ctypes.cast(ptrs, ctypes.POINTER(_f))
        if scene.get("mat_maps") is not None:
            d.material_maps = _ptr(k(scene["mat_maps"], np.int32), _i)
        if scene.get("mat_nmaps") is not None:
            d.material_normal_maps = _ptr(k(scene["mat_nmaps"], np.int32), _i)
    if scene.get("mat_map_strength") is not None:
        d.material_map_strengths = _ptr(k(scene["mat_map_strength"], np.float32), _f)
    if scene.get("spheres") is not None and len(scene["spheres"]) > 0:
        sp = k(scene["spheres"], np.float32).reshape(-1, 4)
        d.num_spheres = sp.shape[0]
        d.spheres = _ptr(sp, _f)
        d.sphere_material_ids = _ptr(k(scene["sphere_mat"], np.int32), _i)
    return d, keep


class Bvh:
    """BVH over a scene (pt_bvh_build) with the trace and render entry points."""

    def __init__(self, scene, bins=16, leaf_max=4, width=2, builder=BUILD_DEVICE):
        self.scene = scene
        self.ctx = scene.ctx
        self.width = width
        d = BvhDesc(bins, leaf_max, width, builder)
        h = ctypes.c_void_p()
        self.ctx.check(self.ctx._lib.pt_bvh_build(scene.handle, ctypes.byref(d), ctypes.byref(h)), "pt_bvh_build")
        self.handle = h

    def info(self):
        nn = ctypes.c_int32()
        md = ctypes.c_int32()
        ms = ctypes.c_double()
        self.ctx.check(self.ctx._lib.pt_bvh_info(self.handle, ctypes.byref(nn), ctypes.byref(md), ctypes.byref(ms)),
                       "pt_bvh_info")
        return {"num_nodes": nn.value, "max_depth": md.value, "build_ms": ms.value}

    def export_order(self):
        out = np.zeros(self.scene.num_primitives, np.int32)
        self.ctx.check(self.ctx._lib.pt_bvh_export_order(self.handle, _ptr(out, _i)), "pt_bvh_export_order")
        return out

    def export_nodes(self):
        n = self.info()["num_nodes"]
        boxes = np.zeros((n, 6), np.float32)
        child = np.zeros((n, 2), np.int32)
        rng = np.zeros((n, 2), np.int32)
        leaf = np.zeros(n, np.int32)
        self.ctx.check(self.ctx._lib.pt_bvh_export_nodes(self.handle, _ptr(boxes, _f), _ptr(child, _i),
                                                         _ptr(rng, _i), _ptr(leaf, _i)), "pt_bvh_export_nodes")
        return {"boxes": boxes, "child": child, "range": rng, "leaf": leaf}

    def set_trace_tuning(self, trace_blocks_per_cu=0, trace_chunk=64, threads_per_ray=1, leaf_batch=0):
        """pt_bvh_set_trace_tuning: launch shapes of this BVH's traces and renders."""
        t = Tuning(8, trace_blocks_per_cu, trace_chunk, 8, 0, 0, threads_per_ray, leaf_batch)
        self.ctx.check(self.ctx._lib.pt_bvh_set_trace_tuning(self.handle, ctypes.byref(t)), "pt_bvh_set_trace_tuning")

    def trace_closest(self, rays, t_out=None, prim_out=None, stream=None):
        """rays: torch float32 CUDA tensor [n, 8]. Returns (t, prim) tensors."""
        import torch
        n = rays.shape[0]
        assert rays.is_cuda and rays.dtype == torch.float32 and rays.is_contiguous() and rays.shape[1] == 8
        if t_out is None:
            t_out = torch.empty(n, dtype=torch.float32, device=rays.device)
        if prim_out is None:
            prim_out = torch.empty(n, dtype=torch.int32, device=rays.device)
        st = ctypes.c_void_p(stream) if stream else None
        self.ctx.check(self.ctx._lib.pt_trace_closest(self.handle, n, rays.data_ptr(), t_out.data_ptr(),
                                                      prim_out.data_ptr(), st), "pt_trace_closest")
        return t_out, prim_out

    @staticmethod
    def _rays_soa(rays, default_window=False):
        """pt_rays_soa over the columns of an [n, 8] ray tensor's structure-of-arrays copy;
        default_window: null dist_min / dist_max (0 and FLT_MAX)."""
        cols = rays.t().contiguous()  # [8, n]: ox oy oz dist_min dx dy dz dist_max
        ptrs = [cols[k].data_ptr() for k in (0, 1, 2, 4, 5, 6, 3, 7)]
        if default_window:
            ptrs[6] = ptrs[7] = None
        return RaysSoa(*ptrs), cols

    def trace_closest_soa(self, rays, uv=True, stream=None, default_window=False):
        """pt_trace_closest_soa on the columns of `rays` ([n, 8] as trace_closest) given as
        separate arrays. Returns (t, prim, u, v) tensors (u, v None unless uv)."""
        import torch
        n = rays.shape[0]
        assert rays.is_cuda and rays.dtype == torch.float32 and rays.shape[1] == 8
        r, cols = self._rays_soa(rays, default_window)
        t = torch.empty(n, dtype=torch.float32, device=rays.device)
        p = torch.empty(n, dtype=torch.int32, device=rays.device)
        u = torch.empty(n, dtype=torch.float32, device=rays.device) if uv else None
        v = torch.empty(n, dtype=torch.float32, device=rays.device) if uv else None
        h = HitsSoa(t.data_ptr(), p.data_ptr(), u.data_ptr() if uv else None, v.data_ptr() if uv else None)
        st = ctypes.c_void_p(stream) if stream else None
        self.ctx.check(self.ctx._lib.pt_trace_closest_soa(self.handle, n, ctypes.byref(r), ctypes.byref(h), st),
                       "pt_trace_closest_soa")
        if stream is None:
            torch.cuda.synchronize()  # cols must outlive the kernels
        return t, p, u, v

    def trace_any_soa(self, rays, stream=None):
        """pt_trace_any_soa on the columns of `rays`; returns the int32 hit flags."""
        import torch
        n = rays.shape[0]
        r, cols = self._rays_soa(rays)
        out = torch.empty(n, dtype=torch.int32, device=rays.device)
        st = ctypes.c_void_p(stream) if stream else None
        self.ctx.check(self.ctx._lib.pt_trace_any_soa(self.handle, n, ctypes.byref(r), out.data_ptr(), st),
                       "pt_trace_any_soa")
        if stream is None:
            torch.cuda.synchronize()
        return out

    def trace_closest_noalloc(self, rays, t_out, prim_out, stream=None):
        """pt_trace_closest_noalloc: as trace_closest on caller-given outputs, on the GPU's
        fixed scratch (no allocation; at most 2^21 rays; overlapping calls on one GPU race)."""
        import torch
        n = rays.shape[0]
        assert rays.is_cuda and rays.dtype == torch.float32 and rays.is_contiguous() and rays.shape[1] == 8
        assert t_out.numel() >= n and prim_out.numel() >= n
        st = ctypes.c_void_p(stream) if stream else None
        self.ctx.check(self.ctx._lib.pt_trace_closest_noalloc(self.handle, n, rays.data_ptr(), t_out.data_ptr(),
                                                              prim_out.data_ptr(), st), "pt_trace_closest_noalloc")
        return t_out, prim_out

    def trace_any(self, rays, hit_out=None, stream=None):
        import torch
        n = rays.shape[0]
        assert rays.is_cuda and rays.dtype == torch.float32 and rays.is_contiguous() and rays.shape[1] == 8
        if hit_out is None:
            hit_out = torch.empty(n, dtype=torch.int32, device=rays.device)
        st = ctypes.c_void_p(stream) if stream else None
        self.ctx.check(self.ctx._lib.pt_trace_any(self.handle, n, rays.data_ptr(), hit_out.data_ptr(), st),
                       "pt_trace_any")
        return hit_out

    def trace_stats(self, rays):
        import torch
        n = rays.shape[0]
        t_out = torch.empty(n, dtype=torch.float32, device=rays.device)
        prim_out = torch.empty(n, dtype=torch.int32, device=rays.device)
        st = np.zeros(2, np.int64)
        self.ctx.check(self.ctx._lib.pt_trace_stats(self.handle, n, rays.data_ptr(), t_out.data_ptr(),
                                                    prim_out.data_ptr(), _ptr(st, _l)), "pt_trace_stats")
        return t_out, prim_out, st

    def render(self, width, height, spp, max_bounces, integrator=INTEGRATOR_MIS, seed=1, out=None,
               time_kernels=False, plain=False, trace_stats=False, two_streams=False, launch_log=False,
               sampler_length=0, three_streams=False):
        """Renders one image; returns (rgb [h, w, 3] float32 numpy array or the given device
        tensor, stats dict). time_kernels: every launch timed on its own; plain: plain
        launches with one synchronisation per iteration; two_streams: a large image in two
        halves on two streams (three_streams: three parts on three streams); launch_log: the
        production schedule with every trace launch
        logged (with trace_stats: and its work counted) (all give the same image).
        sampler_length > 0: the 2D sampler is MicroJitteredSampler over Halton sequences of
        that length instead of RandomSampler2D."""
        flags = (RENDER_TIME_KERNELS if (time_kernels or trace_stats) and not launch_log else 0) | \
            (RENDER_PLAIN_LAUNCHES if plain else 0) | (RENDER_TRACE_STATS if trace_stats else 0) | \
            (RENDER_TWO_STREAMS if two_streams or three_streams else 0) | \
            (RENDER_THREE_STREAMS if three_streams else 0) | (RENDER_LAUNCH_LOG if launch_log else 0)
        d = RenderDesc(width, height, spp, max_bounces, integrator, seed, flags, sampler_length)
        st = RenderStats()
        if out is None:
            host = np.zeros((height, width, 3), np.float32)
            self.ctx.check(self.ctx._lib.pt_render(self.scene.handle, self.handle, ctypes.byref(d),
                                                   host.ctypes.data, 0, ctypes.byref(st)), "pt_render")
            res = host
        else:
            self.ctx.check(self.ctx._lib.pt_render(self.scene.handle, self.handle, ctypes.byref(d),
                                                   out.data_ptr(), 1, ctypes.byref(st)), "pt_render")
            res = out
        kinds = ("extension", "shadow_any", "shadow_closest", "shade")
        return res, {"queries": st.queries, "iterations": st.iterations, "elapsed_ms": st.elapsed_ms,
                     "kernels": {k: {"items": st.kernel_items[i], "launches": st.kernel_launches[i],
                                     "ms": st.kernel_ms[i], "nodes": st.kernel_nodes[i], "tris": st.kernel_tris[i]}
                            for i, k in enumerate(kinds)},
                     "log": {"trace_launches": st.log_launches, "trace_rays": st.log_items,
                             "trace_ms": st.log_trace_ms, "trace_nodes": st.log_nodes, "trace_tris": st.log_tris,
                             "finish_launches": st.finish_launches, "finish_rays": st.finish_rays,
                             "finish_ms": st.finish_ms}}

    def render_aov(self, width, height, spp, max_bounces, aov=AOV_DENOISING_ALBEDO, seed=1, out=None):
        """A denoising AOV (pt_render_aov): AOV_DENOISING_ALBEDO or AOV_DENOISING_NORMAL,
        spp the denoising sample count. Returns (rgb, stats) as render()."""
        d = RenderDesc(width, height, spp, max_bounces, INTEGRATOR_MIS, seed, 0)
        st = RenderStats()
        if out is None:
            res = np.zeros((height, width, 3), np.float32)
            ptr, on_dev = res.ctypes.data, 0
        else:
            res, ptr, on_dev = out, out.data_ptr(), 1
        self.ctx.check(self.ctx._lib.pt_render_aov(self.scene.handle, self.handle, aov, ctypes.byref(d), ptr, on_dev,
                                                   ctypes.byref(st)), "pt_render_aov")
        return res, {"queries": st.queries, "elapsed_ms": st.elapsed_ms}

    def close(self):
        if getattr(self, "handle", None):
            self.ctx._lib.pt_bvh_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Job:
    """One render job of a batch: a BVH (and its scene) plus render settings."""

    def __init__(self, bvh, width, height, spp, max_bounces, integrator=INTEGRATOR_MIS, seed=1, out=None,
                 sampler_length=0):
        self.bvh, self.width, self.height = bvh, width, height
        self.spp, self.max_bounces, self.integrator, self.seed = spp, max_bounces, integrator, seed
        self.sampler_length = sampler_length
        self.out = out  # device tensor [h, w, 3] or None (host numpy result)
        self.result = None


class Engine:
    """A multi-job engine (pt_engine_create): renders batches of jobs that share every
    launch (pt_render_batch)."""

    def __init__(self, ctx, max_paths, max_jobs=256, multi=True):
        self.ctx = ctx
        h = ctypes.c_void_p()
        ctx.check(ctx._lib.pt_engine_create(ctx.handle, int(max_paths), int(max_jobs), 1 if multi else 0,
                                            ctypes.byref(h)), "pt_engine_create")
        self.handle = h

    def set_tuning(self, iterations_per_sync=8, trace_blocks_per_cu=0, trace_chunk=64, shade_blocks_per_cu=8,
                   admission=ADMIT_FIT, solo_log2_samples=0, threads_per_ray=1, leaf_batch=0):
        """pt_engine_set_tuning. solo_log2_samples = k > 0: render_batch renders each job of
        at least 2^k samples alone on the context's single-job engine, beside the others.
        threads_per_ray = 2: two lanes walk each ray of a binary BVH (the same hits).
        leaf_batch = k in 1..4 (0: the default, 3): the binary walk runs in wave-wide node
        and triangle phases; -1: every step runs both (the same hits either way)."""
        t = Tuning(iterations_per_sync, trace_blocks_per_cu, trace_chunk, shade_blocks_per_cu, admission,
                   solo_log2_samples, threads_per_ray, leaf_batch)
        self.ctx.check(self.ctx._lib.pt_engine_set_tuning(self.handle, ctypes.byref(t)), "pt_engine_set_tuning")

    def get_tuning(self):
        t = Tuning()
        self.ctx.check(self.ctx._lib.pt_engine_get_tuning(self.handle, ctypes.byref(t)), "pt_engine_get_tuning")
        return {k: getattr(t, k) for k, _ in Tuning._fields_}

    def render_batch(self, jobs, iterations_per_sync=0):
        arr = (JobDesc * len(jobs))()
        keep = []
        for i, j in enumerate(jobs):
            d = arr[i]
            d.scene = j.bvh.scene.handle
            d.bvh = j.bvh.handle
            d.render = RenderDesc(j.width, j.height, j.spp, j.max_bounces, j.integrator, j.seed, 0,
                                  getattr(j, "sampler_length", 0))
            if j.out is None:
                host = np.zeros((j.height, j.width, 3), np.float32)
                keep.append(host)
                j.result = host
                d.rgb_out = host.ctypes.data
                d.out_on_device = 0
            else:
                j.result = j.out
                d.rgb_out = j.out.data_ptr()
                d.out_on_device = 1
        st = BatchStats()
        self.ctx.check(self.ctx._lib.pt_render_batch(self.handle, len(jobs), arr, iterations_per_sync,
                                                     ctypes.byref(st)), "pt_render_batch")
        return {"queries": st.queries, "iterations": st.iterations, "elapsed_ms": st.elapsed_ms,
                "wall_ms": st.wall_ms, "syncs": st.syncs}

    def close(self):
        if getattr(self, "handle", None):
            self.ctx._lib.pt_engine_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Server:
    """Render server (pt_server_*): resident scenes, queued requests and a request loop
    with continuous batching."""

    def __init__(self, budget_bytes, device=0, max_paths=0, max_jobs=0, iterations_per_sync=0, solo_log2_samples=0,
                 solo_paths=0):
        """solo_log2_samples = k > 0: requests of at least 2^k samples render one at a time on a
        second, single-job engine of solo_paths path slots (0: 8M) beside the shared pool."""
        self._lib = lib()
        d = ServerDesc(device, max_jobs, int(budget_bytes), int(max_paths), iterations_per_sync, solo_log2_samples,
                       int(solo_paths))
        h = ctypes.c_void_p()
        self._check(self._lib.pt_server_create(ctypes.byref(d), ctypes.byref(h)), "pt_server_create")
        self.handle = h
        self._shapes = {}

    def _check(self, rc, what):
        if rc != 0:
            msg = self._lib.pt_server_last_error()
            raise PtError("%s failed (%d): %s" % (what, rc, msg.decode() if msg else ""))

    def add_scene(self, scene, bins=16, leaf_max=4, builder=BUILD_DEVICE):
        d, keep = scene_desc(scene)
        bd = BvhDesc(bins, leaf_max, 2, builder)
        sid = ctypes.c_int64()
        self._check(self._lib.pt_server_add_scene(self.handle, ctypes.byref(d), ctypes.byref(bd), ctypes.byref(sid)),
                    "pt_server_add_scene")
        del keep
        return sid.value

    def remove_scene(self, scene_id):
        self._check(self._lib.pt_server_remove_scene(self.handle, scene_id), "pt_server_remove_scene")

    def resident(self, scene_id):
        return bool(self._lib.pt_server_scene_resident(self.handle, scene_id))

    def submit(self, scene_id, width, height, spp, max_bounces, integrator=INTEGRATOR_MIS, seed=1, sampler_length=0):
        rd = RenderDesc(width, height, spp, max_bounces, integrator, seed, 0, sampler_length)
        rid = ctypes.c_int64()
        self._check(self._lib.pt_server_submit(self.handle, scene_id, ctypes.byref(rd), ctypes.byref(rid)),
                    "pt_server_submit")
        self._shapes[rid.value] = (height, width, 3)
        return rid.value

    @staticmethod
    def _info(i):
        return {"state": i.state, "progress": i.progress, "queue_ms": i.queue_ms, "latency_ms": i.latency_ms}

    def poll(self, rid):
        i = RequestInfo()
        self._check(self._lib.pt_server_poll(self.handle, rid, ctypes.byref(i)), "pt_server_poll")
        return self._info(i)

    def wait(self, rid, timeout_ms=-1):
        i = RequestInfo()
        self._check(self._lib.pt_server_wait(self.handle, rid, int(timeout_ms), ctypes.byref(i)), "pt_server_wait")
        return self._info(i)

    def preview(self, rid):
        img = np.zeros(self._shapes[rid], np.float32)
        pr = ctypes.c_float()
        self._check(self._lib.pt_server_preview(self.handle, rid, _ptr(img, _f), ctypes.byref(pr)),
                    "pt_server_preview")
        return img, pr.value

    def fetch(self, rid):
        img = np.zeros(self._shapes[rid], np.float32)
        self._check(self._lib.pt_server_fetch(self.handle, rid, _ptr(img, _f)), "pt_server_fetch")
        del self._shapes[rid]
        return img

    def stats(self):
        st = ServerStats()
        self._check(self._lib.pt_server_get_stats(self.handle, ctypes.byref(st)), "pt_server_get_stats")
        return {k: getattr(st, k) for k, _ in ServerStats._fields_}

    def close(self):
        if getattr(self, "handle", None):
            self._lib.pt_server_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
