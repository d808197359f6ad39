"""Writes tests/golden/ref_vectors.npz from the reference's own translation units.

Runs oracle/_ref/ref_harness (built by `make -C oracle` from the sources under
/root/reference, see oracle/Makefile) on its seeded inputs and keeps a subset of
inputs and outputs as fixtures, so the oracle can be checked against the reference
where the reference is absent. Usage: python tests/golden/make_golden.py
"""
import os
import subprocess
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
HARNESS = os.path.join(ROOT, "oracle", "_ref", "ref_harness")

# name -> (dtype, row width, rows kept)
FILES = {
    "xorshift_seeds.u32": (np.uint32, 1, 32),
    "xorshift_u32.u32": (np.uint32, 100, 32),
    "xorshift_f32.f32": (np.float32, 100, 32),
    "tri_tri.f32": (np.float32, 9, 6000),
    "tri_ray.f32": (np.float32, 6, 6000),
    "tri_hit.i32": (np.int32, 1, 6000),
    "tri_t.f32": (np.float32, 1, 6000),
    "trisample_tri.f32": (np.float32, 9, 1000),
    "trisample_seed.u32": (np.uint32, 2, 1000),
    "trisample_out.f32": (np.float32, 7, 1000),
    "camera_cam.f32": (np.float32, 16, 1200),
    "camera_pix.i32": (np.int32, 4, 1200),
    "camera_seed.u32": (np.uint32, 2, 1200),
    "camera_out.f32": (np.float32, 9, 1200),
    "mat_type.i32": (np.int32, 1, 4000),
    "mat_param.f32": (np.float32, 8, 4000),
    "mat_shading.f32": (np.float32, 6, 4000),
    "mat_ray_in.f32": (np.float32, 12, 4000),
    "mat_seed.u32": (np.uint32, 2, 4000),
    "mat_ray_out.f32": (np.float32, 12, 4000),
    "mat_dir_out.f32": (np.float32, 3, 4000),
    "mat_eval.f32": (np.float32, 4, 4000),
    "onb_in.f32": (np.float32, 5, 2000),
    "onb_out.f32": (np.float32, 12, 2000),
    "halton.f32": (np.float32, 1000, 3),
    "mj_cfg.u32": (np.uint32, 2, 24),
    "mj_out.f32": (np.float32, 1200, 24),
}


def run_harness(out_dir):
    if not os.path.exists(HARNESS):
        subprocess.check_call(["make", "-C", os.path.join(ROOT, "oracle")])
    subprocess.check_call([HARNESS, out_dir])
    arrays = {}
    for name, (dt, width, rows) in FILES.items():
        a = np.fromfile(os.path.join(out_dir, name), dtype=dt).reshape(-1, width)
        arrays[name.split(".")[0]] = a
    return arrays


def main():
    with tempfile.TemporaryDirectory() as d:
        arrays = run_harness(d)
    keep = {k: arrays[k][:FILES[[n for n in FILES if n.startswith(k + ".")][0]][2]] for k in arrays}
    np.savez_compressed(os.path.join(HERE, "ref_vectors.npz"), **keep)
    print("wrote", os.path.join(HERE, "ref_vectors.npz"))


if __name__ == "__main__":
    sys.exit(main())
