"""Checks that pin the oracle's integrator composition and environment lookup to what the
mathematics fixes, independently of the oracle's own code.

The reference's integrators (Core/Integrator/PathTracing.cpp, SimplePathTracing.cpp),
Environment.cpp and the AOV passes cannot be compiled in this image (they need spdlog, fmt
and stb), so their restatements in oracle/refcpu.cpp are pinned here by:
  * closed forms: pixels whose value the reference's arithmetic fixes exactly (the camera
    throughput times a constant), and images that quirks of the reference make zero;
  * expectations: the mean of many samples against a float64 quadrature of the same
    estimator's expected value, with a tolerance taken from the sample spread;
  * the platform C library for the double-precision sin/cos/acos/atan2 the reference binds.
Each test names in its docstring a one-line mutation of the oracle that makes it fail (each
was applied to oracle/refcpu.cpp or petrichor_amd/csrc/pt_libm.h and observed to fail).
"""
import math

import numpy as np
import pytest

import oracle_lib as O
import petrichor_amd as P
from petrichor_amd import scenes

KEPS = 1e-3  # Core/Constants.h:10


def v3(*a):
    return np.array(a, np.float32)


def floor_quad(r, z=0.0):
    """A square [-r, r]^2 at height z, counter-clockwise seen from above (normal +Z)."""
    return scenes._quad(v3(-r, -r, z), v3(r, -r, z), v3(r, r, z), v3(-r, r, z))


def ceiling_quad(r, z):
    """A square [-r, r]^2 at height z, normal -Z."""
    return scenes._quad(v3(-r, -r, z), v3(-r, r, z), v3(r, r, z), v3(r, -r, z))


def pinhole_down(height, focal=2.4):
    """A pinhole camera above the origin looking down (slightly tilted so that the world up
    +Z is not parallel to the view), with a field of view of about 0.01 rad."""
    return scenes.camera(pos=(0, -1e-2, height), look_at=(0, 0, 0), focal_length=focal, aperture=0.0)


def lambert(kd=1.0):
    return scenes._material(P.PT_MAT_LAMBERT, (kd, kd, kd))


def camera_rays(cam, w, h, job_seed):
    """The first camera ray of every pixel: (o, dir, throughput) rows, as the renderers draw
    them from the pixel's x and y streams (samplers seeded per pixel)."""
    out = np.zeros((h, w, 10), np.float32)
    for y in range(h):
        for x in range(w):
            pix = y * w + x
            out[y, x] = O.sample_camera_ray(cam, x, y, w, h, job_seed)
    return out


def gauss_square(f, half, n=160):
    """Integral of f(x, y) over [-half, half]^2 by n x n Gauss-Legendre nodes (float64)."""
    t, w = np.polynomial.legendre.leggauss(n)
    x, y = np.meshgrid(t * half, t * half, indexing="ij")
    return float((f(x, y) * np.outer(w, w)).sum() * half * half)


def mean_and_se(img):
    """Grand mean of the pixel values and its standard error from the spread of the pixels
    (each pixel is the mean of its own independent samples)."""
    v = img.reshape(-1, img.shape[-1]).astype(np.float64)
    return v.mean(0), v.std(0, ddof=1) / math.sqrt(len(v))


# ---- Environment::GetColor ---------------------------------------------------------------

def test_env_get_color_linear_map_with_rotation():
    """Environment::GetColor (Core/Environment.cpp:29-48) on an environment map linear in
    column and row: texel (i, j) = (0.5 + 0.01 i, 0.25 + 0.02 j, 1). With theta = acos(dz),
    phi = atan2(dy, dx), u = 1 - (phi - rot) / 2pi and v = theta / pi (both mod 1), the
    bilinear fetch at (u W, v H) (texel centres at integer coordinates, Texture2D.cpp:
    287-335) reads (0.5 + 0.01 u W, 0.25 + 0.02 v H, 1), times the base colour.
    Mutation that fails: `u = 1.0f + (phi - rot) * 0.5f * kInvPi` (or dropping `rot`)."""
    W, H = 64, 32
    jj, ii = np.meshgrid(np.arange(H), np.arange(W), indexing="ij")
    tex = np.stack([0.5 + 0.01 * ii, 0.25 + 0.02 * jj, np.ones_like(ii, float)], -1).astype(np.float32)
    base, rot = (2.0, 3.0, 0.5), 0.7
    sc = scenes._finish(floor_quad(1.0), [lambert()], [0, 0], [], pinhole_down(5.0), env_rgb=tex,
                        env_param=(*base, rot, 1))
    o = O.OracleScene(sc)
    rng = np.random.default_rng(4)
    checked = 0
    for _ in range(400):
        theta, phi = rng.uniform(0.2, math.pi - 0.2), rng.uniform(-math.pi, math.pi)
        d = v3(math.sin(theta) * math.cos(phi), math.sin(theta) * math.sin(phi), math.cos(theta))
        th = math.acos(float(d[2]))
        ph = math.atan2(float(d[1]), float(d[0]))
        u = (1.0 - (ph - rot) / (2 * math.pi)) % 1.0
        v = (th / math.pi) % 1.0
        x, y = u * W, v * H
        if not (0.5 < x < W - 1.5 and 0.5 < y < H - 1.5):
            continue  # away from the clamped last column and row
        expect = np.array([0.5 + 0.01 * x, 0.25 + 0.02 * y, 1.0]) * base
        got = o.env_color(d)
        assert np.allclose(got, expect, rtol=2e-5, atol=0), (theta, phi, got, expect)
        checked += 1
    assert checked > 300


# ---- the double-precision libm routines (pt_libm.h) against the platform C library ----------

@pytest.mark.parametrize("fn", ["sin", "cos", "acos", "atan2"])
def test_libm_matches_platform_libm_at_float(fn):
    """The reference's unqualified sin/cos/acos/atan2 on floats call the double C functions
    and round the result to float where it is stored (e.g. Environment.cpp:38-39,
    Lambert.cpp:69-72). The oracle and the device share pt_libm.h; on 1M float arguments
    per function (plus special values) its result rounded to float equals the platform's
    libm (Python's math module) rounded to float, and stays within 4 double ulps of it.
    Mutations that fail: a sign dropped in pt_sincos's quadrant table (`case 2: *s = sr`,
    `case 1: *c = sr`); pt_acos's upper branch taking asin(z) for asin(sqrt(z)); the first
    coefficient of pt_kernel_atan doubled."""
    rng = np.random.default_rng({"sin": 1, "cos": 2, "acos": 3, "atan2": 4}[fn])
    n = 1 << 20
    if fn in ("sin", "cos"):
        a = np.concatenate([rng.uniform(-2 * math.pi, 2 * math.pi, n // 2), rng.uniform(-200.0, 200.0, n // 2),
                            [0.0, -0.0, math.pi, -math.pi, math.pi / 2, 1e-30, 3.14159265358979]])
        a = a.astype(np.float32).astype(np.float64)
        b = None
        ref = np.array([(math.sin if fn == "sin" else math.cos)(x) for x in a.tolist()])
        got = O.libm(0 if fn == "sin" else 1, a)
    elif fn == "acos":
        a = np.concatenate([rng.uniform(-1.0, 1.0, n), [1.0, -1.0, 0.0, -0.0, 0.5, -0.5, 1 - 2 ** -24]])
        a = a.astype(np.float32).astype(np.float64)
        b = None
        ref = np.array([math.acos(x) for x in a.tolist()])
        got = O.libm(2, a)
    else:
        a = np.concatenate([rng.normal(size=n), [0.0, -0.0, 1.0, -1.0, 0.0, 1e-20]]).astype(np.float32)
        b = np.concatenate([rng.normal(size=n), [1.0, -1.0, 0.0, 0.0, -0.0, -1.0]]).astype(np.float32)
        a, b = a.astype(np.float64), b.astype(np.float64)
        ref = np.array([math.atan2(y, x) for y, x in zip(a.tolist(), b.tolist())])
        got = O.libm(3, a, b)
    assert np.array_equal(got.astype(np.float32).view(np.uint32), ref.astype(np.float32).view(np.uint32))
    ulp = np.abs(got.view(np.int64) - ref.view(np.int64))
    assert int(ulp.max()) <= 4


# ---- PathTracing::SampleLight --------------------------------------------------------------

def three_lights_scene():
    """Three emissive triangles of areas 0.5, 2 and 4.5 (right triangles with legs 1, 2, 3),
    registered in that order, plus a floor that is not a light."""
    tris = floor_quad(5.0)
    for k, s in enumerate((1.0, 2.0, 3.0)):
        z = 1.0 + k
        tris.append(np.concatenate([v3(0, 0, z), v3(s, 0, z), v3(0, s, z)]))
    mats = [lambert(), scenes._material(P.PT_MAT_EMISSION, (1, 1, 1))]
    return scenes._finish(tris, mats, [0, 0, 1, 1, 1], [2, 3, 4], pinhole_down(10.0)), (0.5, 2.0, 4.5)


def test_sample_light_pdf_slots_and_draws():
    """PathTracing::SampleLight (PathTracing.cpp:329-386): with the environment slot
    (CalcLightContribution's call) slot k of L + 1 is int(rnd (L + 1)), clamped to L, and
    slot L returns before any light is sampled (no 2D draw); otherwise every light draws one
    2D sample (L draws), the chosen light's point is returned and pdfArea = 1 / sum(area).
    Without the environment slot (the emitter-hit MIS call) the slot is int(rnd L) clamped
    to L - 1. Mutations that fail: `sum_area += pa` instead of `1.0f / pa`; `index = (int)(rnd
    * L)` in the environment branch."""
    sc, areas = three_lights_scene()
    o = O.OracleScene(sc)
    L = len(areas)
    tris = sc["tri_pos"][sc["light_tris"]].astype(np.float64).reshape(L, 3, 3)
    for rnd in np.linspace(0.0, 1.0, 2001).astype(np.float32):
        for want_env in (True, False):
            k, pos, nrm, pdf_area, draws = o.sample_light(float(rnd), 1234, 5678, want_env)
            slots = L + 1 if want_env else L
            expect = min(int(np.float32(rnd) * np.float32(slots)), slots - 1)
            if want_env and expect == L:
                assert k == -1 and draws == 0 and pdf_area == -1.0
                continue
            assert k == expect and draws == L
            assert abs(pdf_area - 1.0 / sum(areas)) <= 1e-6 / sum(areas)
            # the point lies on light k: inside its triangle and in its plane
            p0, p1, p2 = tris[k]
            e1, e2, q = p1 - p0, p2 - p0, pos.astype(np.float64) - p0
            assert abs(q[2]) < 1e-6
            bu, bv = np.linalg.solve(np.array([[e1[0], e2[0]], [e1[1], e2[1]]]), q[:2])
            assert -1e-6 <= bu and -1e-6 <= bv and bu + bv <= 1 + 1e-6
            assert np.allclose(np.abs(nrm), (0, 0, 1))
    # slot frequencies over uniform random values: 1 / (L + 1) each
    picks = np.array([o.sample_light(float(r), 1, 2, True)[0] for r in np.random.default_rng(0).random(40000)])
    freq = np.array([(picks == s).mean() for s in (0, 1, 2, -1)])
    assert np.allclose(freq, 0.25, atol=0.012)


# ---- PathTracing: area-light NEE, emitter hits and their MIS weights --------------------------

def area_light_expectation(S, H, L, le=1.0):
    """Expected radiance at the origin of a kd = 1 Lambert floor under an emissive square
    [-S, S]^2 at height H (L equal triangles registered as lights), for the MIS integrator
    at one bounce: the area-light estimator of CalcLightContribution (PathTracing.cpp:
    222-280), chosen with probability L / (L + 1) and not compensated for it, plus the
    BSDF-sampled hit of the emitter (PathTracing.cpp:97-137), each with the power heuristic.
    Shadow and bounce rays leave from the origin offset by kEps along the normal; l2 is
    measured from the shading point (NEE) or the ray origin (emitter hit), as there."""
    A = (2 * S) ** 2
    pl_area = 1.0 / A

    def nee(x, y):
        dz = H - KEPS
        r = np.sqrt(x * x + y * y + dz * dz)
        cos = dz / r              # at the floor (direction from the offset origin)
        cos_p = cos               # at the emitter (normal -Z)
        l2 = x * x + y * y + H * H  # from the shading point
        pb = (cos / np.pi) * cos_p / l2
        w = pl_area ** 2 / (pl_area ** 2 + pb ** 2)
        return w * le * (1 / np.pi) * cos * cos_p / l2  # estimator x its density 1 / A

    def bsdf(x, y):
        dz = H - KEPS
        l2 = x * x + y * y + dz * dz  # from the bounce ray's origin
        r = np.sqrt(l2)
        cos = dz / r
        cos_p = cos
        pb = cos / np.pi
        pl = l2 / cos_p * pl_area
        w = pb ** 2 / (pl ** 2 + pb ** 2)
        return (cos / np.pi) * (cos_p / l2) * w * le  # solid-angle density to area

    return L / (L + 1) * gauss_square(nee, S) + gauss_square(bsdf, S)


def test_area_light_mis_expectation():
    """A kd = 1 Lambert floor under a 20 x 20 emissive ceiling at height 10 (two triangles,
    both registered), MIS integrator, one bounce: the mean of 262,144 samples equals the
    quadrature of the estimator's expectation (area_light_expectation: NEE weighted by
    L / (L + 1) = 2/3, the power heuristic on both strategies, pdfArea = 1 / total area)
    within 5 standard errors. Mutations that fail: the power heuristic replaced by the
    balance heuristic in CalcLightContribution; the emitter-hit weight set to 1; `index =
    std::min(index, (int)(L - 1))` in the environment-slot branch (no environment slot)."""
    S, H = 10.0, 10.0
    tris = floor_quad(30.0) + ceiling_quad(S, H)
    mats = [lambert(), scenes._material(P.PT_MAT_EMISSION, (1, 1, 1))]
    sc = scenes._finish(tris, mats, [0, 0, 1, 1], [2, 3], pinhole_down(5.0))
    o = O.OracleScene(sc).build_bvh(16)
    img, q = o.render(64, 64, 64, 1, 0, 17, 0)
    m, se = mean_and_se(img)
    expect = area_light_expectation(S, H, 2)
    assert 0.2 < expect < 0.6
    # camera throughput: cos^4 of at most 0.006 rad off the axis
    assert np.all(np.abs(m - expect) < 5 * se + 2e-4 * expect), (m, se, expect)


def test_unregistered_emitter_is_weighted_zero_under_mis():
    """With no registered light, SampleLight's emitter-hit call computes index -1 and
    pdfArea = 1 / 0 = +inf (PathTracing.cpp:341-385), so the BSDF-sampled hit of an
    emitter gets MIS weight pdfBSDF^2 / (inf + pdfBSDF^2) = 0 (PathTracing.cpp:119-124):
    a Lambert floor lit only by an unregistered emissive ceiling renders exactly black under
    MIS, and the simple integrator, which has no MIS, sees the ceiling (SimplePathTracing.cpp:
    58-66). Mutation that fails: the emitter-hit weight `mis` set to 1."""
    tris = floor_quad(30.0) + ceiling_quad(10.0, 10.0)
    mats = [lambert(), scenes._material(P.PT_MAT_EMISSION, (1, 1, 1))]
    sc = scenes._finish(tris, mats, [0, 0, 1, 1], [], pinhole_down(5.0))
    o = O.OracleScene(sc).build_bvh(16)
    mis, _ = o.render(16, 16, 16, 4, 0, 3, 0)
    assert np.count_nonzero(mis) == 0
    simple, _ = o.render(16, 16, 16, 4, 1, 3, 0)
    assert simple.min() > 0


# ---- PathTracing: environment NEE and escaped bounces --------------------------------------

def env_mis_expectation(lights):
    """Expected radiance / (throughput x env) at a kd = 1 Lambert floor under a constant,
    importance-sampled environment, MIS integrator at one bounce, with `lights` registered
    lights that are never visible. The environment slot is chosen with probability
    1 / (lights + 1); its sample has pdfuv = 1 (the degenerate tables) and theta uniform on
    [0, pi], and weights the power heuristic against 2 pi^2 sin(theta) pdfBSDF
    (PathTracing.cpp:282-321); the escaped bounce weighs pdfBSDF against pdfEnv = 1 /
    (2 pi^2 sin) (PathTracing.cpp:141-174). With x = sin^2 theta both parts reduce to
    integrals over [0, 1] of w(x) = 1 / (1 + 4 pi^2 x (1 - x))."""
    t, wt = np.polynomial.legendre.leggauss(400)
    x = 0.5 * (t + 1.0)
    w_env = 1.0 / (1.0 + 4 * np.pi ** 2 * x * (1 - x))
    nee = 0.5 * float((wt * w_env).sum())              # int_0^1 w dx
    esc = 0.5 * float((wt * (1.0 - w_env)).sum())      # int_0^1 (1 - w) dx
    return nee / (lights + 1) + esc


def constant_env(c=(0.5, 1.0, 2.0), w=8, h=1024):
    return np.ones((h, w, 3), np.float32), c


@pytest.mark.parametrize("lights", [0, 1])
def test_environment_mis_expectation(lights):
    """A kd = 1 Lambert floor under a constant importance-sampled environment, MIS integrator
    at one bounce, with 0 or 1 registered (but occluded: below the floor) lights: the mean of
    262,144 samples / (env colour) equals env_mis_expectation within 5 standard errors. With
    no light the two weighted strategies sum to exactly the environment (1); with one light
    the uncompensated slot choice halves the NEE part. Mutations that fail: the balance
    heuristic in the environment NEE weight; `2.0f * kPi * kPi` dropped from pdfBSDF there."""
    tex, c = constant_env()
    tris = floor_quad(30.0)
    mats = [lambert(), scenes._material(P.PT_MAT_EMISSION, (1, 1, 1))]
    tri_mat = [0, 0]
    light = []
    if lights:
        tris.append(np.concatenate([v3(-0.5, -0.5, -5), v3(0.5, -0.5, -5), v3(0, 0.5, -5)]))
        tri_mat.append(1)
        light = [2]
    sc = scenes._finish(tris, mats, tri_mat, light, pinhole_down(5.0), env_rgb=tex, env_param=(*c, 0.0, 1))
    o = O.OracleScene(sc).build_bvh(16)
    img, _ = o.render(64, 64, 64, 1, 0, 29, 0)
    m, se = mean_and_se(img / np.array(c, np.float32))
    expect = env_mis_expectation(lights)
    if lights == 0:
        assert abs(expect - 1.0) < 1e-9
    assert np.all(np.abs(m - expect) < 5 * se + 2e-4), (m, se, expect)


def test_glass_gets_zero_weight_under_mis():
    """Glass has BxDF = PDF = 0 (Glass.cpp:11-27): its NEE contribution is 0 x Li, and an
    escaped bounce after glass has MIS weight 0^2 / (pdfEnv^2 + 0^2) = 0 (PathTracing.cpp:
    151-166), so a glass slab that fills the view renders exactly black under an
    importance-sampled environment, whatever the bounces. The view is oblique: an escaped
    ray with dir.x or dir.y zero, or with sin(theta) rounding to 0 (a vertical ray), takes
    weight 1 instead (PathTracing.cpp:146-167). Mutation that fails: the glass PDF
    returning 1 instead of 0."""
    tex, c = constant_env()
    tris = floor_quad(30.0, 1.0) + ceiling_quad(30.0, 0.8)
    mats = [scenes._material(P.PT_MAT_GLASS, (1, 1, 1), 1.5)]
    cam = scenes.camera(pos=(-3, -2.5, 4), look_at=(0, 0, 1), focal_length=0.2, aperture=0.0)
    sc = scenes._finish(tris, mats, [0, 0, 0, 0], [], cam, env_rgb=tex, env_param=(*c, 0.0, 1))
    o = O.OracleScene(sc).build_bvh(16)
    img, q = o.render(16, 16, 16, 8, 0, 5, 0)
    assert np.count_nonzero(img) == 0
    assert q > 16 * 16 * 16 * 3  # the paths went through the slab and out


def test_env_without_importance_sampling_ignores_kd():
    """Environment importance sampling off, no lights, a constant environment c over a
    Lambert floor: NEE picks the (empty) environment slot and adds nothing, the bounce keeps
    the throughput (Lambert::GetAlbedo is one when untextured, Lambert.cpp:77,117-126) and
    the escaped ray adds T c unweighted (PathTracing.cpp:176-181). So at one sample per pixel
    every pixel is exactly the camera throughput times c, whatever kd is, and at 16 samples
    the images for kd 0.2 and 0.9 are identical. Mutation that fails: the Lambert bounce
    multiplying the throughput by kd."""
    c = (0.25, 0.5, 0.75)
    cam = pinhole_down(5.0)

    def scene(kd):
        return scenes._finish(floor_quad(30.0), [lambert(kd)], [0, 0], [], cam, env_param=(*c, 0.0, 0))

    w, h = 12, 10
    img, _ = O.OracleScene(scene(0.2)).build_bvh(16).render(w, h, 1, 4, 0, 9, 1)
    rays = camera_rays(cam, w, h, 9)
    expect = (rays[..., 6:9] * np.array(c, np.float32)).astype(np.float32)
    assert np.array_equal(img.view(np.uint32), expect.view(np.uint32))
    a, _ = O.OracleScene(scene(0.2)).build_bvh(16).render(w, h, 16, 4, 0, 9, 1)
    b, _ = O.OracleScene(scene(0.9)).build_bvh(16).render(w, h, 16, 4, 0, 9, 1)
    assert np.array_equal(a.view(np.uint32), b.view(np.uint32))


# ---- SimplePathTracing: Russian roulette ----------------------------------------------------

def closed_box(r=1.0):
    P3 = v3
    q = []
    q += scenes._quad(P3(-r, -r, -r), P3(-r, r, -r), P3(r, r, -r), P3(r, -r, -r))
    q += scenes._quad(P3(-r, -r, r), P3(r, -r, r), P3(r, r, r), P3(-r, r, r))
    q += scenes._quad(P3(-r, -r, -r), P3(r, -r, -r), P3(r, -r, r), P3(-r, -r, r))
    q += scenes._quad(P3(-r, r, -r), P3(-r, r, r), P3(r, r, r), P3(r, r, -r))
    q += scenes._quad(P3(-r, -r, -r), P3(-r, -r, r), P3(-r, r, r), P3(-r, r, -r))
    q += scenes._quad(P3(r, -r, -r), P3(r, r, -r), P3(r, r, r), P3(r, -r, r))
    return q


def test_simple_russian_roulette_path_length():
    """SimplePathTracing (SimplePathTracing.cpp:45-92) inside a closed Lambert box: a path
    makes max_bounces + 1 queries, then survives each roulette step with probability 0.9
    (prob restarts at 1 on every Lambert bounce, Ray.h:24-40 and Lambert.cpp:79-83, so the
    0.1 floor never applies). The number of survivals K is geometric with P(K >= k) = 0.9^k,
    mean 9 and variance 90. Over 32,768 samples the mean number of queries per sample is
    max_bounces + 1 + 9 within 5 standard errors. Mutations that fail: `ray.bounce >=
    max_bounces` as the roulette condition; the roulette probability kept across bounces."""
    cam = scenes.camera(pos=(0.1, -0.2, 0.05), look_at=(0.5, 0.3, 0.2), aperture=0.0)
    sc = scenes._finish(closed_box(), [lambert(0.5)], [0] * 12, [], cam)
    o = O.OracleScene(sc).build_bvh(16)
    max_b = 3
    n = 32 * 32 * 32
    _, q = o.render(32, 32, 32, max_b, 1, 21, 0)
    k_mean = q / n - (max_b + 1)
    se = math.sqrt(90.0 / n)
    assert abs(k_mean - 9.0) < 5 * se, (k_mean, se)


def test_simple_russian_roulette_is_unbiased():
    """SimplePathTracing at max_bounces 0 over a Lambert floor under a constant environment
    c: every bounce ray enters the roulette, survives with probability 0.9 and then carries
    T / 0.9 (SimplePathTracing.cpp:82-88) into the sky, so a sample is T c / 0.9 or 0 and the
    mean over 131,072 samples is T c (T = 1 within 2e-5 here) within 5 standard errors.
    Mutation that fails: the survivor's throughput not divided by prob."""
    c = (0.3, 0.6, 0.9)
    sc = scenes._finish(floor_quad(30.0), [lambert()], [0, 0], [], pinhole_down(5.0), env_param=(*c, 0.0, 0))
    o = O.OracleScene(sc).build_bvh(16)
    img, _ = o.render(64, 64, 32, 0, 1, 13, 0)
    m, se = mean_and_se(img / np.array(c, np.float32))
    assert np.all(np.abs(m - 1.0) < 5 * se + 2e-5), (m, se)


# ---- denoising AOVs -------------------------------------------------------------------------

def test_aov_closed_forms():
    """AOVDenoisingAlbedo / AOVDenoisingNormal (Core/AOV/AOVDenoisingAlbedo.cpp:56-133,
    AOVDenoisingNormal.cpp:54-128) at one sample per pixel: over an untextured Lambert floor
    (kd 0.5) the albedo is the camera throughput T (GetAlbedo is one, not kd) and the normal
    is T times the flat normal (0, 0, 1); over a rough GGX floor the albedo is T x F0 with
    F0 = 0.16 r^2 (GGX.h:126-139); a miss gives T x the environment and T x (-dir).
    Mutation that fails: the albedo AOV of a Lambert taking kd."""
    cam = pinhole_down(5.0)
    w, h = 8, 6
    rays = camera_rays(cam, w, h, 4)
    T = rays[..., 6:9]
    for mat, expect_albedo in ((lambert(0.5), T),
                               (scenes._material(P.PT_MAT_GLOSSY, (0.9, 0.5, 0.2), 0.6),
                                T * (np.float32(0.16) * np.array([0.9, 0.5, 0.2], np.float32) ** 2))):
        sc = scenes._finish(floor_quad(30.0), [mat], [0, 0], [], cam, env_param=(1, 2, 3, 0, 0))
        o = O.OracleScene(sc).build_bvh(16)
        alb, _ = o.render(w, h, 1, 4, 2, 4, 1)
        nrm, _ = o.render(w, h, 1, 4, 3, 4, 1)
        assert np.allclose(alb, expect_albedo, rtol=1e-6, atol=0)
        assert np.array_equal(nrm[..., :2], np.zeros_like(nrm[..., :2]))
        assert np.array_equal(nrm[..., 2], T[..., 2])
    # misses: the floor is behind the camera
    sc = scenes._finish(floor_quad(1.0, 50.0), [lambert()], [0, 0], [], cam, env_param=(1, 2, 3, 0, 0))
    o = O.OracleScene(sc).build_bvh(16)
    alb, _ = o.render(w, h, 1, 4, 2, 4, 1)
    nrm, _ = o.render(w, h, 1, 4, 3, 4, 1)
    assert np.array_equal(alb, (T * np.array([1, 2, 3], np.float32)).astype(np.float32))
    assert np.array_equal(nrm, (T * -rays[..., 3:6]).astype(np.float32))


def test_aov_closed_forms_through_specular_and_emission():
    """The denoising AOVs' pass-through cases (Core/AOV/AOVDenoisingAlbedo.cpp:56-133,
    AOVDenoisingNormal.cpp:54-128), with a tilted pinhole camera over an infinite-looking
    floor under a constant environment c:
    - a smooth GGX floor (roughness 0) is followed: GGX::CreateNextRay with alpha 0 takes
      h = n, so the ray leaves as the mirror direction r with throughput x F,
      F = F0 + (1 - F0)(1 - cos)^5, F0 = 0.16 f0^2 (GGX.cpp:104-153, GGX.h:126-139), and
      then misses: albedo = T F c, normal = T F (-r);
    - a glass floor is followed through: its ray keeps the refracted direction t
      (Vector3f::Refracted with ior 1.5, Glass.cpp:28-95) and its throughput, in both of
      Glass's branches, and misses below: albedo = T c, normal = T (-t); the glass colour
      is not applied;
    - an Emission floor gives zero in both AOVs (the switch's default branch).
    Mutations that fail: dropping the Schlick term, taking the glass colour into the
    throughput, returning T x Le for an emitter."""
    cam = scenes.camera(pos=(0, -4.0, 1.5), look_at=(0, 0, 0), focal_length=0.05, aperture=0.0)
    w, h = 16, 12
    rays = camera_rays(cam, w, h, 4)
    T, d = rays[..., 6:9].astype(np.float64), rays[..., 3:6].astype(np.float64)
    c = np.array([1.0, 2.0, 3.0])
    env = (1, 2, 3, 0, 0)
    cos = -d[..., 2]
    assert cos.min() > 0.1 and cos.max() < 0.6  # grazing enough for the Schlick term to matter
    # smooth GGX
    f0 = np.array([0.9, 0.5, 0.2])
    F0 = 0.16 * f0 ** 2
    F = F0 + (1 - F0) * (1 - cos[..., None]) ** 5
    r = d * np.array([1, 1, -1])
    sc = scenes._finish(floor_quad(1e4), [scenes._material(P.PT_MAT_GLOSSY, tuple(f0), 0.0)], [0, 0], [], cam,
                        env_param=env)
    o = O.OracleScene(sc).build_bvh(16)
    alb, _ = o.render(w, h, 1, 4, 2, 4, 1)
    nrm, _ = o.render(w, h, 1, 4, 3, 4, 1)
    assert np.allclose(alb, T * F * c, rtol=2e-5, atol=0)
    assert np.allclose(nrm, T * F * -r, rtol=2e-5, atol=1e-7)
    # glass (coloured, to show the colour is not applied); the direction is
    # Vector3f::Refracted's, (d - n (-d.n + sqrt((d.n)^2 - 1 + eta^2))) / eta
    # (Math/Vector3f.h:209-234; the function itself is pinned by the reference's own glass
    # vectors), normalised
    # vectors), normalised. Glass's "reflect" branch keeps that direction and the ray's ior
    # and starts it above the floor, so the ray meets the floor again and is refracted once
    # more: after k such draws the ray leaves with t_(k+1) = refract(t_k).
    eta = 1.5

    def refract(v):
        dn = v[..., 2:3]
        t = (v - (-dn + np.sqrt(dn ** 2 - 1 + eta ** 2)) * np.array([0, 0, 1.0])) / eta
        return t / np.linalg.norm(t, axis=-1, keepdims=True)

    ts = [refract(d)]
    for _ in range(3):
        ts.append(refract(ts[-1]))
    sc = scenes._finish(floor_quad(1e4), [scenes._material(P.PT_MAT_GLASS, (0.5, 0.7, 0.9), 1.5)], [0, 0], [], cam,
                        env_param=env)
    o = O.OracleScene(sc).build_bvh(16)
    alb, _ = o.render(w, h, 1, 8, 2, 4, 1)
    nrm, _ = o.render(w, h, 1, 8, 3, 4, 1)
    assert np.array_equal(alb, (rays[..., 6:9] * np.float32([1, 2, 3])).astype(np.float32))
    match = np.stack([np.isclose(nrm, T * -tk, rtol=2e-5, atol=1e-7).all(axis=-1) for tk in ts])
    assert match.any(axis=0).all()
    assert match[0].mean() > 0.8  # most rays are transmitted at the first draw (R < 0.2)
    # emitter
    sc = scenes._finish(floor_quad(1e4), [scenes._material(P.PT_MAT_EMISSION, (5.0, 5.0, 5.0))], [0, 0], [], cam,
                        env_param=env)
    o = O.OracleScene(sc).build_bvh(16)
    for mode in (2, 3):
        img, _ = o.render(w, h, 1, 4, mode, 4, 1)
        assert not img.any()
