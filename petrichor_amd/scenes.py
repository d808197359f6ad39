"""Procedural scenes and ray sets of the benchmark configurations.

A scene is a dict of arrays in the layout of pt_scene_desc (include/petrichor_hip.h):
  tri_pos [N, 9] f32, tri_mat [N] i32, tri_nrm / tri_smooth / tri_uv [N, 6] (optional),
  mat_type [M] i32, mat_param [M, 8] f32, light_tris [L] i32, camera [16] f32,
  env_rgb [H, W, 3] f32 or None, env_param [5] f32 (base rgb, z rotation, importance sampling);
  optional material maps: textures (list of [h, w, 3] f32) and mat_maps [M, 2] i32 (albedo or
  reflectance map, roughness map; -1 none), read at the texture coordinates tri_uv.
Every generator takes an explicit seed.
"""
import numpy as np

from . import PT_MAT_EMISSION, PT_MAT_GLASS, PT_MAT_GLOSSY, PT_MAT_LAMBERT

f32 = np.float32


# ---- Camera set-up, restating Core/Camera.cpp:14-50 in float32 -----------------------------
def _dot(a, b):
    return f32(f32(f32(a[0] * b[0]) + f32(a[1] * b[1])) + f32(a[2] * b[2]))


def _cross(a, b):
    return np.array([f32(a[1] * b[2]) - f32(a[2] * b[1]), f32(a[2] * b[0]) - f32(a[0] * b[2]),
                     f32(a[0] * b[1]) - f32(a[1] * b[0])], dtype=np.float32)


def _length(a):
    return f32(np.sqrt(_dot(a, a)))


def _normalized(a):
    a = np.asarray(a, np.float32)
    return (a / _length(a)).astype(np.float32)


def camera(pos, look_at, focal_length=55e-3, focus_target=None, f_number=None, aperture=None):
    """Camera state (16 floats) as Camera(pos, dir) + SetFocalLength + FocusTo +
    SetFNumber / SetApeture would leave it (Core/Camera.cpp:14-50, Camera.h:71-83)."""
    pos = np.asarray(pos, np.float32)
    look_at = np.asarray(look_at, np.float32)
    forward = _normalized(look_at - pos)                  # SetViewDir
    right = _cross(forward, np.array([0, 0, 1], np.float32))
    up = _cross(right, forward)
    sensor_h = f32(24e-3)
    h_perf = f32(sensor_h / f32(55e-3))                   # default 55 mm lens
    focus = f32(1.0)
    ap = f32(25e-3)
    if focal_length is not None:
        h_perf = f32(sensor_h / f32(focal_length))         # SetFocalLength
    if focus_target is not None:
        focus = _length(np.asarray(focus_target, np.float32) - pos)  # FocusTo
    if f_number is not None:
        ap = f32(f32(sensor_h / h_perf) / f32(f_number))   # SetFNumber
    if aperture is not None:
        ap = f32(aperture)                                 # SetApeture
    c = np.zeros(16, np.float32)
    c[0:3], c[3:6], c[6:9], c[9:12] = pos, right, up, forward
    c[12], c[13], c[14] = focus, h_perf, ap
    return c


def _material(mtype, color=(1, 1, 1), p3=0.0, importance_sampling=True):
    p = np.zeros(8, np.float32)
    p[0:3] = color
    p[3] = p3
    p[6] = 1.0 if importance_sampling else 0.0
    return mtype, p


def _quad(a, b, c, d):
    """Two triangles (a, b, c) and (a, c, d)."""
    return [np.concatenate([a, b, c]), np.concatenate([a, c, d])]


def _finish(tris, mats, tri_mat, lights, cam, env_rgb=None, env_param=(0, 0, 0, 0, 1)):
    mt = np.array([m[0] for m in mats], np.int32)
    mp = np.stack([m[1] for m in mats]).astype(np.float32)
    return {
        "tri_pos": np.asarray(tris, np.float32).reshape(-1, 9),
        "tri_mat": np.asarray(tri_mat, np.int32),
        "tri_nrm": None, "tri_smooth": None,
        "mat_type": mt, "mat_param": mp,
        "light_tris": np.asarray(lights, np.int32),
        "camera": cam,
        "env_rgb": env_rgb,
        "env_param": np.asarray(env_param, np.float32),
    }


# ---- Config 1: Cornell box, 32 triangles -------------------------------------------------
def cornell_box(emission=15.0):
    """Unit room [0,1]^3 (+Z up): 5 walls, a 0.3 cube and a 0.3x0.3x0.6 box rotated by
    +17 and -17 degrees (bottom faces omitted), a 0.25 x 0.25 quad light just under the
    ceiling. White 0.73, red (0.65,0.05,0.05), green (0.05,0.65,0.05), emission 15."""
    P = lambda *v: np.array(v, np.float32)
    tris, tri_mat = [], []

    def add(quads, m):
        for q in quads:
            tris.append(q)
            tri_mat.append(m)

    W, R, G, L = 0, 1, 2, 3
    add(_quad(P(0, 0, 0), P(1, 0, 0), P(1, 1, 0), P(0, 1, 0)), W)       # floor z=0
    add(_quad(P(0, 0, 1), P(0, 1, 1), P(1, 1, 1), P(1, 0, 1)), W)       # ceiling z=1
    add(_quad(P(0, 1, 0), P(1, 1, 0), P(1, 1, 1), P(0, 1, 1)), W)       # back y=1
    add(_quad(P(0, 0, 0), P(0, 1, 0), P(0, 1, 1), P(0, 0, 1)), R)       # left x=0
    add(_quad(P(1, 0, 0), P(1, 0, 1), P(1, 1, 1), P(1, 1, 0)), G)       # right x=1

    def box(cx, cy, sx, sy, h, deg):
        a = np.deg2rad(deg)
        ca, sa = np.cos(a), np.sin(a)
        corners = []
        for (dx, dy) in ((-1, -1), (1, -1), (1, 1), (-1, 1)):
            x, y = dx * sx / 2, dy * sy / 2
            corners.append((cx + ca * x - sa * y, cy + sa * x + ca * y))
        lo = [P(x, y, 0) for x, y in corners]
        hi = [P(x, y, h) for x, y in corners]
        quads = []
        for i in range(4):
            j = (i + 1) % 4
            quads += _quad(lo[i], lo[j], hi[j], hi[i])
        quads += _quad(hi[0], hi[1], hi[2], hi[3])
        add(quads, W)

    box(0.33, 0.35, 0.3, 0.3, 0.3, 17.0)
    box(0.67, 0.65, 0.3, 0.3, 0.6, -17.0)
    light_first = len(tris)
    z = f32(0.999)
    add(_quad(P(0.375, 0.375, z), P(0.375, 0.625, z), P(0.625, 0.625, z), P(0.625, 0.375, z)), L)
    mats = [_material(PT_MAT_LAMBERT, (0.73, 0.73, 0.73)), _material(PT_MAT_LAMBERT, (0.65, 0.05, 0.05)),
            _material(PT_MAT_LAMBERT, (0.05, 0.65, 0.05)), _material(PT_MAT_EMISSION, (emission,) * 3)]
    cam = camera(pos=(0.5, -1.35, 0.5), look_at=(0.5, 0.5, 0.5), focal_length=35e-3,
                 focus_target=(0.5, 0.5, 0.3), f_number=8.0)
    assert len(tris) == 32
    return _finish(tris, mats, tri_mat, [light_first, light_first + 1], cam)


# ---- Config 2: spheres over a terrain, equirect sky -----------------------------------------
_SPHERE_CACHE = {}


def _unit_sphere(seg, bands):
    key = (seg, bands)
    if key not in _SPHERE_CACHE:
        th = np.linspace(0.0, np.pi, bands + 1)
        ph = np.linspace(0.0, 2 * np.pi, seg + 1)[:-1]
        ring = lambda k: np.stack([np.sin(th[k]) * np.cos(ph), np.sin(th[k]) * np.sin(ph),
                                   np.full_like(ph, np.cos(th[k]))], -1)
        i = np.arange(seg)
        j = (i + 1) % seg
        top = np.broadcast_to(np.array([0, 0, 1.0]), (seg, 3))
        bot = np.broadcast_to(np.array([0, 0, -1.0]), (seg, 3))
        out = [np.concatenate([top, ring(1)[i], ring(1)[j]], 1)]
        for k in range(1, bands - 1):
            a, b = ring(k), ring(k + 1)
            q = np.stack([np.concatenate([a[i], b[i], b[j]], 1), np.concatenate([a[i], b[j], a[j]], 1)], 1)
            out.append(q.reshape(-1, 9))
        rl = ring(bands - 1)
        out.append(np.concatenate([bot, rl[j], rl[i]], 1))
        _SPHERE_CACHE[key] = np.concatenate(out)
    return _SPHERE_CACHE[key]


def uv_sphere(center, radius, seg=64, bands=33):
    """seg x bands UV sphere with triangle fans at the poles: 2*seg + 2*seg*(bands-2)
    triangles (4096 for 64 x 33), none degenerate."""
    t = _unit_sphere(seg, bands).reshape(-1, 3, 3) * radius + np.asarray(center, np.float64)
    return t.reshape(-1, 9).astype(np.float32)


def _fbm(x, y, rng, octaves=6):
    """Fractional Brownian motion from random-phase sinusoid octaves."""
    h = np.zeros_like(x)
    amp, freq = 0.6, 0.15
    for _ in range(octaves):
        for _k in range(3):
            a = rng.uniform(0, 2 * np.pi)
            ph = rng.uniform(0, 2 * np.pi)
            h += amp / 3 * np.sin(freq * (np.cos(a) * x + np.sin(a) * y) * 2 * np.pi + ph)
        amp *= 0.5
        freq *= 2.0
    return h


def terrain(n, half=10.0, seed=0):
    rng = np.random.default_rng(seed)
    g = np.linspace(-half, half, n + 1)
    X, Y = np.meshgrid(g, g, indexing="xy")
    Z = _fbm(X, Y, rng) - 0.5
    V = np.stack([X, Y, Z], -1)
    a, b, c, d = V[:-1, :-1], V[:-1, 1:], V[1:, 1:], V[1:, :-1]
    t1 = np.concatenate([a, b, c], -1).reshape(-1, 9)
    t2 = np.concatenate([a, c, d], -1).reshape(-1, 9)
    return np.stack([t1, t2], 1).reshape(-1, 9).astype(np.float32)


def sky_env(width=2048, height=1024, sun_elev_deg=35.0, sun_azim_deg=60.0, sun_sigma_deg=0.5, sun_peak=5e4):
    """Equirect sky: zenith/horizon gradient, dark ground and a Gaussian sun lobe. Texel
    (i, j) holds direction theta = pi*j/H, phi = 2*pi*(1 - i/W) (the lookup of
    Environment::GetColor, Environment.cpp:29-48, with zero rotation)."""
    j = np.arange(height, dtype=np.float64)[:, None]
    i = np.arange(width, dtype=np.float64)[None, :]
    theta = np.pi * j / height + 0 * i
    phi = 2 * np.pi * (1 - i / width) + 0 * j
    elev = np.pi / 2 - theta
    zen, hor, gnd = np.array([0.25, 0.45, 0.9]), np.array([0.85, 0.9, 1.0]), np.array([0.12, 0.1, 0.08])
    t = np.clip(elev / (np.pi / 2), 0, 1)[..., None]
    sky = np.where(elev[..., None] >= 0, hor * (1 - t) + zen * t, gnd)
    se, sa = np.deg2rad(sun_elev_deg), np.deg2rad(sun_azim_deg)
    sd = np.array([np.cos(se) * np.cos(sa), np.cos(se) * np.sin(sa), np.sin(se)])
    d = np.stack([np.sin(theta) * np.cos(phi), np.sin(theta) * np.sin(phi), np.cos(theta)], -1)
    ang = np.arccos(np.clip(d @ sd, -1, 1))
    sun = sun_peak * np.exp(-0.5 * (ang / np.deg2rad(sun_sigma_deg)) ** 2)
    return (sky + sun[..., None] * np.array([1.0, 0.95, 0.85])).astype(np.float32)


def spheres_scene(seed=2, n_spheres=256, seg=64, bands=33, terrain_n=1024, env_w=2048, env_h=1024, env=None):
    """Config 2 at full size: 256 spheres x 4096 triangles = 1,048,576 over a 1024 x 1024
    quad terrain (2,097,152 triangles), 3,145,728 triangles in all; radii
    lognormal(-2, 0.5), centres uniform in [-10,10]^2 x [0,4]; 40% Lambert, 40% GGX
    (roughness U(0.05, 0.8)), 20% glass (ior 1.5); 2048 x 1024 sky with a sun lobe."""
    rng = np.random.default_rng(seed)
    mats = [_material(PT_MAT_LAMBERT, (0.35, 0.4, 0.3))]  # terrain
    parts = [terrain(terrain_n, seed=seed + 1)]
    tri_mat = [np.zeros(len(parts[0]), np.int32)]
    radii = rng.lognormal(-2.0, 0.5, n_spheres)
    centers = np.stack([rng.uniform(-10, 10, n_spheres), rng.uniform(-10, 10, n_spheres),
                        rng.uniform(0, 4, n_spheres)], -1)
    kinds = rng.uniform(0, 1, n_spheres)
    for s in range(n_spheres):
        if kinds[s] < 0.4:
            mats.append(_material(PT_MAT_LAMBERT, rng.uniform(0.1, 0.9, 3)))
        elif kinds[s] < 0.8:
            mats.append(_material(PT_MAT_GLOSSY, rng.uniform(0.3, 1.0, 3), rng.uniform(0.05, 0.8)))
        else:
            mats.append(_material(PT_MAT_GLASS, (1, 1, 1), 1.5))
        t = uv_sphere(centers[s], radii[s], seg, bands)
        parts.append(t)
        tri_mat.append(np.full(len(t), len(mats) - 1, np.int32))
    cam = camera(pos=(0.0, -14.0, 3.5), look_at=(0.0, 0.0, 0.8), focal_length=28e-3,
                 focus_target=(0.0, 0.0, 0.8), f_number=8.0)
    if env is None:
        env = sky_env(env_w, env_h)
    return _finish(np.concatenate(parts), mats, np.concatenate(tri_mat), [], cam, env, (1, 1, 1, 0, 1))


def scaled_scene(n_tris, seed, env=None):
    """The single-large-scene generator scaled to about n_tris triangles: two thirds in
    the terrain grid, one third in 1 to 256 UV spheres (the batched and served workloads).
    env: a shared equirect map (default: the 2048 x 1024 sky)."""
    terrain_n = max(2, int(round(np.sqrt(n_tris / 3.0))))  # 2 n^2 triangles
    sphere_tris = max(n_tris - 2 * terrain_n ** 2, 64)
    n_spheres = int(np.clip(sphere_tris // 1024, 1, 256))
    per = sphere_tris / n_spheres  # a UV sphere has 2 * seg * (bands - 1) triangles
    seg = max(8, int(round(np.sqrt(per))))
    bands = max(3, int(round(per / (2 * seg))) + 1)
    return spheres_scene(seed=seed, n_spheres=n_spheres, seg=seg, bands=bands, terrain_n=terrain_n, env=env)


def sphere_primitives(seed=5, n_spheres=24, env_w=64, env_h=32):
    """Sphere primitives (Core/Geometry/Sphere.cpp) over a triangle floor under a sky: a
    4 x 4 Lambert floor (2 triangles), n_spheres spheres with centres in [-1.5, 1.5]^2 x
    [0.2, 1.2] and radii U(0.08, 0.35) of Lambert, GGX (roughness U(0.05, 0.6)) and glass
    materials, one emissive sphere light (radius 0.25, radiance 12) and one emissive quad
    (radiance 8), both registered as lights (the sphere first). Spheres are the primitives
    after the triangles: 2 floor + 2 quad triangles, then the spheres."""
    rng = np.random.default_rng(seed)
    P = lambda *v: np.array(v, np.float32)
    tris = [np.asarray(q, np.float32).reshape(-1, 9) for q in (
        _quad(P(-2, -2, 0), P(2, -2, 0), P(2, 2, 0), P(-2, 2, 0)),
        _quad(P(-0.6, 0.9, 2.2), P(-0.1, 0.9, 2.2), P(-0.1, 1.4, 2.2), P(-0.6, 1.4, 2.2)))]
    mats = [_material(PT_MAT_LAMBERT, (0.6, 0.55, 0.5)), _material(PT_MAT_EMISSION, (8.0, 8.0, 8.0)),
            _material(PT_MAT_EMISSION, (12.0, 11.0, 10.0))]
    tri_mat = np.array([0, 0, 1, 1], np.int32)
    sph, sph_mat = [], []
    for _ in range(n_spheres):
        c = (rng.uniform(-1.5, 1.5), rng.uniform(-1.5, 1.5), rng.uniform(0.2, 1.2))
        r = rng.uniform(0.08, 0.35)
        kind = rng.uniform()
        if kind < 0.4:
            mats.append(_material(PT_MAT_LAMBERT, rng.uniform(0.2, 0.9, 3)))
        elif kind < 0.8:
            mats.append(_material(PT_MAT_GLOSSY, rng.uniform(0.3, 1.0, 3), rng.uniform(0.05, 0.6)))
        else:
            mats.append(_material(PT_MAT_GLASS, (1, 1, 1), 1.5))
        sph.append((*c, r))
        sph_mat.append(len(mats) - 1)
    sph.append((0.9, -0.4, 1.9, 0.25))  # the sphere light
    sph_mat.append(2)
    cam = camera(pos=(0.2, -4.5, 2.2), look_at=(0.0, 0.0, 0.5), focal_length=30e-3, aperture=0.0)
    sc = _finish(np.concatenate(tris), mats, tri_mat, [4 + n_spheres, 2, 3], cam, sky_env(env_w, env_h),
                 (1, 1, 1, 0.3, 1))
    sc["spheres"] = np.asarray(sph, np.float32)
    sc["sphere_mat"] = np.asarray(sph_mat, np.int32)
    return sc


def deep_chain(n=120):
    """A scene whose binned-SAH tree with 2 bins is about n levels deep: n triangles with
    centroids x = 2^-i (i = 0..n-1), y = 1, z = 2^-i / 48 (distinct depths, so no two are
    hit at the same distance), each 2^-i / 2 wide, 3 tall and tilted so that its box is
    2^-i / 8 thick (no box is flat, which slab tests meet only within rounding). With
    two bins every split puts the rightmost centroid alone in the upper bin, so each level
    peels one triangle. Lambert 0.7 under a uniform 8 x 4 sky of radiance 1; the camera
    looks at the chain's wide end from z = -3."""
    x = np.exp2(-np.arange(n, dtype=np.float64))
    w = 0.25 * x
    z = x / 64
    y = np.zeros_like(x)
    t = np.stack([x - w, y, z, x + w, y, z, x, y + 3.0, z + 0.5 * w], 1)
    cam = camera(pos=(0.4, 1.0, -3.0), look_at=(0.4, 1.0, 0.0), focal_length=35e-3, aperture=0.0)
    env = np.ones((4, 8, 3), np.float32)
    return _finish(t, [_material(PT_MAT_LAMBERT, (0.7, 0.7, 0.7))], np.zeros(n, np.int32), [], cam, env,
                   (1, 1, 1, 0, 1))


def deep_chain_rays(tri_pos, seed=5):
    """Rays for deep_chain: one along +z through every triangle's centroid (the walk to
    the deepest leaves), one along -z from behind through each centroid offset by a
    quarter of its width, and 512 random rays from in front of the plane towards x in
    [2^-16, 1] (farther in, a ray from z = -2 meets the chain only within rounding)."""
    rng = np.random.default_rng(seed)
    t = tri_pos.reshape(-1, 3, 3).astype(np.float64)
    c = t.mean(1)
    w = (t[:, 1, 0] - t[:, 0, 0]) / 2
    r = []
    for k in range(len(c)):
        r.append([c[k, 0], c[k, 1], -1.0, 0.0, 0.0, 0.0, 1.0, np.finfo(np.float32).max])
        r.append([c[k, 0] + 0.25 * w[k], c[k, 1] + 0.5, 1.0, 0.0, 0.0, 0.0, -1.0, np.finfo(np.float32).max])
    o = np.stack([rng.uniform(-0.2, 1.4, 512), rng.uniform(-0.5, 3.5, 512), np.full(512, -2.0)], 1)
    tgt = np.stack([np.exp2(-rng.uniform(0, 16, 512)), rng.uniform(0, 3, 512), np.zeros(512)], 1)
    d = tgt - o
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    for k in range(512):
        r.append([*o[k], 0.0, *d[k], np.finfo(np.float32).max])
    return np.asarray(r, np.float32)


# ---- Config 3: random triangle soup + rays ---------------------------------------------------
def random_triangles(n, seed=3, mu=-6.0, sigma=1.0):
    """n triangles with centroids uniform in [0,1]^3; vertex offsets a, b, -(a+b) with
    lognormal(mu, sigma) lengths and uniform directions (the centroid is exact)."""
    rng = np.random.default_rng(seed)
    c = rng.uniform(0, 1, (n, 3))

    def vec():
        v = rng.normal(size=(n, 3))
        v /= np.linalg.norm(v, axis=1, keepdims=True)
        return v * rng.lognormal(mu, sigma, (n, 1))

    a, b = vec(), vec()
    t = np.concatenate([c + a, c + b, c - a - b], 1).astype(np.float32)
    mats = [_material(PT_MAT_LAMBERT, (0.5, 0.5, 0.5))]
    cam = camera(pos=(0.5, -1.2, 0.5), look_at=(0.5, 0.5, 0.5), focal_length=35e-3, aperture=0.0)
    return _finish(t, mats, np.zeros(n, np.int32), [], cam)


def primary_rays(cam, width, height, seed=4):
    """Pinhole rays through jittered pixel positions (8 floats per ray)."""
    rng = np.random.default_rng(seed)
    pos, right, up, fwd = cam[0:3], cam[3:6], cam[6:9], cam[9:12]
    focus, hp = float(cam[12]), float(cam[13])
    j, i = np.meshgrid(np.arange(height), np.arange(width), indexing="ij")
    u = (i.reshape(-1) + rng.uniform(0, 1, width * height)) / width - 0.5
    v = (j.reshape(-1) + rng.uniform(0, 1, width * height)) / height - 0.5
    aspect = width / height
    d = (right[None] * (u * hp * focus * aspect)[:, None] - up[None] * (v * hp * focus)[:, None]
         + fwd[None] * focus)
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    r = np.zeros((width * height, 8), np.float32)
    r[:, 0:3] = pos
    r[:, 3] = 0.0
    r[:, 4:7] = d
    r[:, 7] = np.finfo(np.float32).max
    return r


def surface_rays(tri_pos, n, seed=5):
    """Rays from uniform points on random triangles, uniform directions on the sphere."""
    rng = np.random.default_rng(seed)
    k = rng.integers(0, len(tri_pos), n)
    t = tri_pos[k].reshape(n, 3, 3).astype(np.float64)
    u, v = rng.uniform(0, 1, n), rng.uniform(0, 1, n)
    su = np.sqrt(u)
    p = t[:, 0] * (1 - su)[:, None] + t[:, 1] * (su * (1 - v))[:, None] + t[:, 2] * (su * v)[:, None]
    d = rng.normal(size=(n, 3))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    r = np.zeros((n, 8), np.float32)
    r[:, 0:3] = p
    r[:, 3] = 0.0
    r[:, 4:7] = d
    r[:, 7] = np.finfo(np.float32).max
    return r


# ---- ray sets and job mixes of the benchmark configurations --------------------------------
def primary_rays_device(cam, width, height, gen):
    """Pinhole rays through jittered pixel positions, built on the device (8 floats per ray)."""
    import torch
    dev = "cuda"
    pos, right, up, fwd = (torch.tensor(cam[a:a + 3], dtype=torch.float64, device=dev) for a in (0, 3, 6, 9))
    focus, hp = float(cam[12]), float(cam[13])
    n = width * height
    idx = torch.arange(n, device=dev)
    i, j = (idx % width).double(), (idx // width).double()
    u = (i + torch.rand(n, generator=gen, device=dev, dtype=torch.float64)) / width - 0.5
    v = (j + torch.rand(n, generator=gen, device=dev, dtype=torch.float64)) / height - 0.5
    d = (right[None] * (u * hp * focus * width / height)[:, None] - up[None] * (v * hp * focus)[:, None]
         + fwd[None] * focus)
    d = d / d.norm(dim=1, keepdim=True)
    r = torch.zeros((n, 8), dtype=torch.float32, device=dev)
    r[:, 0:3] = pos.float()
    r[:, 4:7] = d.float()
    r[:, 7] = float(np.finfo(np.float32).max)
    return r


def surface_rays_device(tri_pos_dev, n, gen, dist_min):
    """Rays from uniform points on random triangles with uniform directions on the sphere."""
    import torch
    dev = tri_pos_dev.device
    k = torch.randint(0, tri_pos_dev.shape[0], (n,), generator=gen, device=dev)
    t = tri_pos_dev[k].view(n, 3, 3).double()
    u = torch.rand(n, generator=gen, device=dev, dtype=torch.float64)
    v = torch.rand(n, generator=gen, device=dev, dtype=torch.float64)
    su = u.sqrt()
    p = t[:, 0] * (1 - su)[:, None] + t[:, 1] * (su * (1 - v))[:, None] + t[:, 2] * (su * v)[:, None]
    d = torch.randn((n, 3), generator=gen, device=dev, dtype=torch.float64)
    d = d / d.norm(dim=1, keepdim=True)
    r = torch.zeros((n, 8), dtype=torch.float32, device=dev)
    r[:, 0:3] = p.float()
    r[:, 3] = dist_min
    r[:, 4:7] = d.float()
    r[:, 7] = float(np.finfo(np.float32).max)
    return r


RESOLUTIONS = [(256, 256), (512, 512), (1280, 720), (1920, 1080)]


def job_specs(n, rng, tris_max, spp_max):
    """Render jobs of the batched and served workloads: triangles log-uniform in
    [1e3, tris_max], resolution from RESOLUTIONS, spp log-uniform in [4, spp_max],
    max depth uniform in 8..12."""
    out = []
    for _ in range(n):
        tris = int(round(np.exp(rng.uniform(np.log(1e3), np.log(tris_max)))))
        w, h = RESOLUTIONS[rng.integers(0, len(RESOLUTIONS))]
        spp = int(round(np.exp(rng.uniform(np.log(4), np.log(spp_max)))))
        depth = int(rng.integers(8, 13))
        out.append({"tris": tris, "width": w, "height": h, "spp": spp, "depth": depth})
    return out
