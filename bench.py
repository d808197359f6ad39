"""Benchmark of the MI355X render path.

Default workload: the single large scene — 3,145,728 triangles (256 UV spheres of 4096
triangles over a 1024 x 1024-quad fBm terrain), 2048 x 1024 equirect sky with a sun lobe,
1920 x 1080, 64 spp, max depth 12, MIS integrator, per-pixel seeded samplers. One step is
one full frame. `value` is BVH queries (camera, bounce and shadow rays) per second over
the timed steps, in Mrays/s.

  python bench.py [--gpus 1] [--steps K] [--warmup W] [--cpu-seconds S]

Prints one JSON line. Only one GPU is supported; --gpus N > 1 exits with a message.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "Mrays/s on 1M-tri scene; batched/served jobs/s vs plain; p99 job latency"
HBM_PEAK_GBS = 8000.0

# Algorithmic bytes of one extension query on the binary BVH layout: ray in (2 x float4),
# path id (4), hit out (t + slot, 8); 32 B per node visit (an inner-node record of 64 B
# holds the boxes of both children, and each child box is one visit); 48 B per triangle
# test (p0, e1, e2, n).
RAY_IN_B, PATH_B, HIT_OUT_B, NODE_B, TRI_B = 32, 4, 8, 32, 48


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def cpu_baseline(sc, tree, width, height, spp, depth, seconds):
    """The CPU oracle (clean-room restatement of the reference renderer, tile pool of
    std::threads as Core/Petrichor.cpp) on a random pixel sample of the same frame."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import oracle_lib as O
    threads = min(os.cpu_count() or 1, 16)
    o = O.OracleScene(sc).use_tree(tree)
    rng = np.random.default_rng(12345)
    n = 2048
    total_q, total_t, total_px = 0, 0.0, 0
    work = np.zeros(9, np.int64)
    while True:
        pix = np.sort(rng.choice(width * height, n, replace=False)).astype(np.int32)
        w = np.zeros(9, np.int64)
        t0 = time.perf_counter()
        _, q = o.render(width, height, spp, depth, 0, 1, threads, pixels=pix, work=w)
        dt = time.perf_counter() - t0
        total_q += q
        total_t += dt
        total_px += n
        work += w
        if total_t >= seconds or total_px >= width * height:
            break
        n = int(min(max(n * 2, n * (seconds - total_t) / max(dt, 1e-3)), width * height - total_px, 1 << 20))
        n = max(n, 1)
    return {
        "value": total_q / total_t / 1e6, "unit": "Mrays/s", "cores": threads, "kind": "port",
        "sample": "%d random pixels of the 1920x1080 frame at full 64 spp / depth 12 (%.1f s, %d queries)"
                  % (total_px, total_t, total_q),
    }, work


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--width", type=int, default=1920)
    ap.add_argument("--height", type=int, default=1080)
    ap.add_argument("--spp", type=int, default=64)
    ap.add_argument("--depth", type=int, default=12)
    ap.add_argument("--mode", default="frame", choices=["frame", "copy"],
                    help="copy: a plain 1 GiB read + 1 GiB write elementwise kernel, the reference "
                         "for scaling rocprofv3 FETCH_SIZE / WRITE_SIZE")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py runs on one GPU only: this build has no multi-GPU path "
              "(pixels or jobs would shard across GPUs with the scene replicated).", file=sys.stderr)
        sys.exit(2)

    import torch
    import petrichor_amd as P
    from petrichor_amd import scenes

    assert torch.cuda.is_available(), "no GPU"
    if args.mode == "copy":
        x = torch.ones(1 << 28, dtype=torch.float32, device="cuda")
        for _ in range(4):
            x.add_(1.0)  # reads 1 GiB and writes 1 GiB per launch
        torch.cuda.synchronize()
        print(json.dumps({"mode": "copy", "launches": 4, "read_bytes_per_launch": 1 << 30,
                          "write_bytes_per_launch": 1 << 30}), flush=True)
        return
    log("generating scene")
    t0 = time.perf_counter()
    sc = scenes.spheres_scene(seed=2)
    log("scene: %d triangles in %.1f s" % (len(sc["tri_mat"]), time.perf_counter() - t0))
    ctx = P.Context(0)
    s = ctx.scene(sc)
    bvh = s.build_bvh(16)
    info = bvh.info()
    log("bvh:", info)
    W, H, SPP, D = args.width, args.height, args.spp, args.depth
    out = torch.empty((H, W, 3), dtype=torch.float32, device="cuda")
    for i in range(args.warmup):
        _, st = bvh.render(W, H, SPP, D, P.INTEGRATOR_MIS, 1, out=out)
        log("warmup", i, st["queries"], "queries", "%.1f ms" % st["elapsed_ms"], st["iterations"], "iterations")
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    stats = []
    for i in range(args.steps):
        _, st = bvh.render(W, H, SPP, D, P.INTEGRATOR_MIS, 1, out=out)
        stats.append(st)
        log("step", i, st["queries"], "queries", "%.1f ms" % st["elapsed_ms"], st["iterations"], "iterations")
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    queries = sum(s_["queries"] for s_ in stats)
    value = queries / wall / 1e6
    img = out.cpu().numpy()
    assert np.isfinite(img).all() and img.mean() > 0

    # One more frame with every launch on its own and timed by HIP events on the render
    # stream: the per-kernel durations of the roofline (not part of `value`).
    timed = torch.empty_like(out)
    _, kst = bvh.render(W, H, SPP, D, P.INTEGRATOR_MIS, 1, out=timed, time_kernels=True)
    torch.cuda.synchronize()
    assert torch.equal(timed, out), "the instrumented frame must match the production frames"
    log("instrumented frame: %.1f ms" % kst["elapsed_ms"])
    ext_items = kst["kernels"]["extension"]["items"]
    ext_launches = kst["kernels"]["extension"]["launches"]
    ext_ms = kst["kernels"]["extension"]["ms"]
    kernels = {k: {"items": v["items"], "launches": v["launches"], "ms": round(v["ms"], 3)}
               for k, v in kst["kernels"].items()}

    cpu = None
    n_node = n_tri = None
    if not args.no_cpu:
        log("cpu baseline")
        tree = bvh.export_nodes() | {"prim_ids": bvh.export_order()}
        cpu, work = cpu_baseline(sc, tree, W, H, SPP, D, args.cpu_seconds)
        n_node = work[1] / max(work[0], 1)
        n_tri = work[2] / max(work[0], 1)
        log("cpu:", cpu)
    roofline = None
    if n_node is not None and ext_ms > 0:
        b_ray = RAY_IN_B + PATH_B + HIT_OUT_B + n_node * NODE_B + n_tri * TRI_B
        achieved = ext_items * b_ray / (ext_ms * 1e-3) / 1e9
        roofline = {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                    "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": None,
                    "kernel": "k_trace_persistent (extension rays, instrumented frame)",
                    "bytes_per_query": round(b_ray, 1), "n_node": round(n_node, 2), "n_tri": round(n_tri, 3),
                    "s_node": NODE_B, "s_tri": TRI_B, "avg_launch_ms": round(ext_ms / max(ext_launches, 1), 4),
                    "queries_per_launch": round(ext_items / max(ext_launches, 1), 1)}
    res = {
        "metric": METRIC, "value": round(value, 2), "unit": "Mrays/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": round(wall / args.steps * 1e3, 2), "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f32", "data": "synthetic",
        "config": {"workload": "single large scene: %d triangles (256 spheres x 4096 + 1024^2-quad terrain), "
                               "2048x1024 env, %dx%d, %d spp, depth %d, MIS" % (len(sc["tri_mat"]), W, H, SPP, D),
                   "triangles": int(len(sc["tri_mat"])), "width": W, "height": H, "spp": SPP, "max_depth": D,
                   "integrator": "MIS", "bvh": "binned SAH, 16 bins, leaf<=4, binary (reference topology)"},
        "queries_per_step": int(queries / args.steps),
        "iterations_per_step": int(stats[-1]["iterations"]),
        "bvh_build_ms": round(info["build_ms"], 1),
        "kernels_instrumented_frame": kernels,
        "roofline": roofline,
        "cpu_baseline": cpu,
    }
    print(json.dumps(res), flush=True)


if __name__ == "__main__":
    main()
