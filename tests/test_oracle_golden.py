"""The oracle against vectors produced by the reference's own translation units.

tests/golden/ref_vectors.npz is written by tests/golden/make_golden.py, which runs
oracle/_ref/ref_harness: XorShift128 (Random/XorShift.h), Triangle::Intersect and
SampleSurface (Core/Geometry/Triangle.cpp), Camera::GenerateRay (Core/Camera.cpp),
Lambert/GGX/Glass CreateNextRay, BxDF and PDF (Core/Material/*.cpp) and
OrthonormalBasis (Math/OrthonormalBasis.cpp), all compiled from the reference
sources. Every comparison is bit-exact.
"""
import os
import subprocess

import numpy as np
import pytest

import oracle_lib as O

HERE = os.path.dirname(os.path.abspath(__file__))
GOLD = np.load(os.path.join(HERE, "golden", "ref_vectors.npz"))


def bits(a):
    return np.ascontiguousarray(a, dtype=np.float32).view(np.uint32)


def test_xorshift_matches_reference():
    seeds = GOLD["xorshift_seeds"][:, 0]
    for k, s in enumerate(seeds):
        u, f = O.xorshift(int(s), 100)
        assert np.array_equal(u, GOLD["xorshift_u32"][k])
        assert np.array_equal(bits(f), bits(GOLD["xorshift_f32"][k]))


def test_xorshift_closed_form():
    # seed overwrites only w (XorShift.h:17-18); first output from the fixed x,y,z.
    x = 123456789
    t = (x ^ (x << 11)) & 0xFFFFFFFF
    for w in (0, 1, 0xDEADBEEF):
        expect = ((w ^ (w >> 19)) ^ (t ^ (t >> 8))) & 0xFFFFFFFF
        u, f = O.xorshift(w, 1)
        assert int(u[0]) == expect
        assert f[0] == np.float32(np.float32(expect) / np.float32(4294967296.0))
    # next() reaches exactly 1.0 for u32 >= 2^32 - 128 (rounding of the u32 -> f32 conversion)
    assert np.float32(np.float32(0xFFFFFFFF) / np.float32(4294967296.0)) == np.float32(1.0)


def test_triangle_intersect_matches_reference():
    tri, ray, hit, t = GOLD["tri_tri"], GOLD["tri_ray"], GOLD["tri_hit"][:, 0], GOLD["tri_t"][:, 0]
    assert hit.sum() > 1000
    for k in range(len(hit)):
        h, tk = O.triangle_intersect(tri[k], ray[k])
        assert h == hit[k], k
        if h:
            assert bits(np.float32(tk)) == bits(t[k]), k


def test_triangle_sample_matches_reference():
    tri, seed, out = GOLD["trisample_tri"], GOLD["trisample_seed"], GOLD["trisample_out"]
    for k in range(len(out)):
        o = O.triangle_sample(tri[k], int(seed[k, 0]), int(seed[k, 1]))
        assert np.array_equal(bits(o), bits(out[k])), k


def test_camera_matches_reference():
    cam, pix, seed, out = GOLD["camera_cam"], GOLD["camera_pix"], GOLD["camera_seed"], GOLD["camera_out"]
    assert (cam[:, 14] == 0).any() and (cam[:, 14] > 0).any()  # pinhole and thin lens
    for k in range(len(out)):
        o = O.camera_ray(cam[k], *[int(v) for v in pix[k]], int(seed[k, 0]), int(seed[k, 1]))
        assert np.array_equal(bits(o[:9]), bits(out[k])), k
        assert o[9] == 2  # two 2D samples per camera ray (Camera.cpp:58,69)


@pytest.mark.parametrize("kind", [0, 1, 3])
def test_material_sampling_matches_reference(kind):
    mt = GOLD["mat_type"][:, 0]
    idx = np.nonzero(mt == kind)[0]
    assert len(idx) > 500
    for k in idx:
        o = O.next_ray(int(mt[k]), GOLD["mat_param"][k], GOLD["mat_shading"][k], GOLD["mat_ray_in"][k],
                       int(GOLD["mat_seed"][k, 0]), int(GOLD["mat_seed"][k, 1]))
        assert np.array_equal(bits(o), bits(GOLD["mat_ray_out"][k])), k
        e = O.bsdf_eval(int(mt[k]), GOLD["mat_param"][k], GOLD["mat_shading"][k][3:6],
                        GOLD["mat_ray_in"][k][3:6], GOLD["mat_dir_out"][k])
        assert np.array_equal(bits(e), bits(GOLD["mat_eval"][k])), k


def test_lambert_uniform_branch_present():
    # half of the Lambert fixtures run with importance sampling disabled (Lambert.cpp:84-101)
    p = GOLD["mat_param"][GOLD["mat_type"][:, 0] == 0]
    assert (p[:, 6] == 0).sum() > 100 and (p[:, 6] == 1).sum() > 100


def test_onb_matches_reference():
    i, o = GOLD["onb_in"], GOLD["onb_out"]
    for k in range(len(o)):
        r = O.onb(i[k][:3], float(i[k][3]), float(i[k][4]))
        assert np.array_equal(bits(r), bits(o[k])), k


HARNESS = os.path.join(os.path.dirname(HERE), "oracle", "_ref", "ref_harness")


@pytest.mark.skipif(not os.path.exists(HARNESS), reason="reference harness not built here")
def test_full_harness_run_matches(tmp_path):
    """Where the reference sources exist, check every harness vector, not just the fixture subset."""
    subprocess.check_call([HARNESS, str(tmp_path)], stdout=subprocess.DEVNULL)
    rd = lambda n, t: np.fromfile(str(tmp_path / n), dtype=t)
    tri = rd("tri_tri.f32", np.float32).reshape(-1, 9)
    ray = rd("tri_ray.f32", np.float32).reshape(-1, 6)
    hit = rd("tri_hit.i32", np.int32)
    t = rd("tri_t.f32", np.float32)
    for k in range(len(hit)):
        h, tk = O.triangle_intersect(tri[k], ray[k])
        assert h == hit[k] and (not h or bits(np.float32(tk)) == bits(t[k])), k
    mt = rd("mat_type.i32", np.int32)
    mp = rd("mat_param.f32", np.float32).reshape(-1, 8)
    sh = rd("mat_shading.f32", np.float32).reshape(-1, 6)
    ri = rd("mat_ray_in.f32", np.float32).reshape(-1, 12)
    sd = rd("mat_seed.u32", np.uint32).reshape(-1, 2)
    ro = rd("mat_ray_out.f32", np.float32).reshape(-1, 12)
    for k in range(len(mt)):
        o = O.next_ray(int(mt[k]), mp[k], sh[k], ri[k], int(sd[k, 0]), int(sd[k, 1]))
        assert np.array_equal(bits(o), bits(ro[k])), k


def test_halton_matches_reference():
    for k, base in enumerate((2, 3, 5)):
        assert np.array_equal(bits(O.halton(1000, base)), bits(GOLD["halton"][k]))


def test_microjittered_matches_reference():
    for (seed, length), ref in zip(GOLD["mj_cfg"], GOLD["mj_out"]):
        got = O.microjittered(int(seed), int(length), 600)
        assert np.array_equal(bits(got.reshape(-1)), bits(ref)), (seed, length)
