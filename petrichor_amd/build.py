"""Builds the gfx950 library libpetrichor_hip.so in-tree (and the test oracle).

python -m petrichor_amd.build   (also called by __graft_entry__.build())
"""
import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "libpetrichor_hip.so")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")

# Floating point: no FMA contraction (the reference's x86-64 build rounds every
# product and sum), correctly rounded f32 division and square root, denormals kept.
# Packed FP32 instructions are disabled for all device code.
FP_FLAGS = ["-ffp-contract=off", "-fhip-fp32-correctly-rounded-divide-sqrt", "-fno-fast-math"]
DEVICE_FLAGS = ["--offload-arch=gfx950", "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops"]
COMMON = ["-O3", "-std=c++17", "-fPIC", "-Wall", "-Wno-unused-function"] + FP_FLAGS

SOURCES = ["pt_kernels.hip", "pt_trace.hip", "pt_build.hip", "pt_wide.hip", "pt_post.hip", "pt_order.hip", "pt_obj.cpp", "pt_png.cpp", "pt_image.cpp", "pt_bvh_host.cpp", "pt_capi.cpp", "pt_engine.cpp", "pt_server.cpp"]


def _run(cmd):
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    # hipcc's host pass does not know the packed-fp32-ops feature; that warning is harmless
    out = "\n".join(l for l in r.stdout.splitlines() if "packed-fp32-ops" not in l)
    if r.returncode != 0:
        sys.stderr.write(out + "\n")
        raise RuntimeError("command failed: " + " ".join(cmd))
    return out


def build(verbose=False):
    bdir = os.path.join(HERE, "_build")
    os.makedirs(bdir, exist_ok=True)
    objs = [os.path.join(bdir, src + ".o") for src in SOURCES]

    def compile_one(i):
        cmd = [HIPCC] + COMMON + DEVICE_FLAGS + ["-x", "hip", "-c", os.path.join(CSRC, SOURCES[i]), "-o", objs[i]]
        return _run(cmd)

    workers = max(1, min(len(SOURCES), (os.cpu_count() or 1) // 2, 8))
    with ThreadPoolExecutor(workers) as pool:
        for out in pool.map(compile_one, range(len(SOURCES))):
            if verbose and out.strip():
                print(out)
    tmp = LIB + ".tmp"
    _run([HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", "-Wl,--no-undefined", "-o", tmp] + objs + ["-lz"])
    os.replace(tmp, LIB)
    # the C++ host API's test drivers (tests/cpp), linked against the library
    _run(["g++", "-std=c++17", "-O2", "-ffp-contract=off", "-Wall", "-I", os.path.join(ROOT, "include"),
          os.path.join(ROOT, "tests", "cpp", "render_api.cpp"), "-L", HERE, "-lpetrichor_hip",
          "-Wl,-rpath,$ORIGIN/..", "-Wl,-rpath-link,/opt/rocm/lib", "-o", os.path.join(bdir, "render_api")])
    _run(["g++", "-std=c++17", "-O2", "-ffp-contract=off", "-Wall", "-I", os.path.join(ROOT, "include"),
          os.path.join(ROOT, "tests", "cpp", "load_api.cpp"), "-L", HERE, "-lpetrichor_hip",
          "-Wl,-rpath,$ORIGIN/..", "-Wl,-rpath-link,/opt/rocm/lib", "-o", os.path.join(bdir, "load_api")])
    for drv in ("sampler_api", "sphere_api", "accel_api"):
        _run(["g++", "-std=c++17", "-O2", "-ffp-contract=off", "-Wall", "-I", os.path.join(ROOT, "include"),
              os.path.join(ROOT, "tests", "cpp", drv + ".cpp"), "-L", HERE, "-lpetrichor_hip",
              "-Wl,-rpath,$ORIGIN/..", "-Wl,-rpath-link,/opt/rocm/lib", "-o", os.path.join(bdir, drv)])
    # test oracle (CPU restatement; also builds oracle/_ref where the reference exists)
    subprocess.check_call(["make", "-s", "-C", os.path.join(ROOT, "oracle")])
    return LIB


def gen_tonemap_lut():
    """Regenerates csrc/pt_tonemap_lut.h from the host C library's powf (about 40 s on 8
    cores: every 32-bit input pattern is evaluated). build() uses the committed table."""
    bdir = os.path.join(HERE, "_build")
    os.makedirs(bdir, exist_ok=True)
    exe = os.path.join(bdir, "gen_tonemap_lut")
    _run(["g++", "-std=c++17", "-O2", "-ffp-contract=off", "-fopenmp", os.path.join(CSRC, "gen_tonemap_lut.cpp"),
          "-o", exe])
    out = os.path.join(CSRC, "pt_tonemap_lut.h")
    with open(out + ".tmp", "w") as f:
        subprocess.check_call([exe], stdout=f)
    os.replace(out + ".tmp", out)
    return out


if __name__ == "__main__":
    if "--tonemap-lut" in sys.argv[1:]:
        print(gen_tonemap_lut())
    print(build(verbose=True))
