"""Autotuner of the batched engine.

Chooses, for a given set of render jobs, the path-pool size (the capacity of every ray
queue), the graph replays per host synchronisation and the launch shapes of the trace
and shading kernels (pt_tuning). It times pt_render_batch on a calibration batch - the
same scenes and resolutions at one sample per pixel - with coordinate descent over the
space below, for each pool size. Every candidate renders the same images (tuning only
changes launch shapes and scheduling), so the fastest one is taken as is.
"""
import time

DEFAULT = {"iterations_per_sync": 8, "trace_blocks_per_cu": 0, "trace_chunk": 64, "shade_blocks_per_cu": 8}

SPACE = {
    "iterations_per_sync": [2, 4, 8, 16, 32],
    "trace_blocks_per_cu": [2, 3, 4, 6],  # 0 (default): as many as the LDS stacks allow
    "trace_chunk": [64, 128, 256, 512],
    "shade_blocks_per_cu": [4, 8, 16],
}


def calibration_jobs(jobs, spp=1):
    """The jobs at `spp` samples per pixel with host outputs dropped (device outputs kept)."""
    from . import Job
    return [Job(j.bvh, j.width, j.height, min(spp, j.spp), j.max_bounces, j.integrator, j.seed, j.out) for j in jobs]


def tune(ctx, jobs, pool_sizes=None, passes=1, log=None):
    """Returns {"max_paths", "tuning", "seconds", "trials"} for the fastest configuration
    found on the calibration batch of `jobs`."""
    from . import Engine
    need = max(j.width * j.height for j in jobs)
    if pool_sizes is None:
        pool_sizes = sorted({max(need, p) for p in (1 << 21, 1 << 22, 1 << 23)})
    calib = calibration_jobs(jobs)
    trials = []
    best = None
    for pool in pool_sizes:
        eng = Engine(ctx, pool, max_jobs=max(len(jobs), 1))

        def measure(t):
            eng.set_tuning(**t)
            t0 = time.perf_counter()
            eng.render_batch(calib)
            dt = time.perf_counter() - t0
            trials.append({"max_paths": pool, "tuning": dict(t), "seconds": dt})
            return dt

        measure(DEFAULT)  # first use: graph instantiation and caches
        cur, cur_t = dict(DEFAULT), measure(DEFAULT)
        for _ in range(passes):
            for key, values in SPACE.items():
                for v in values:
                    if v == cur[key]:
                        continue
                    cand = dict(cur, **{key: v})
                    t = measure(cand)
                    if t < cur_t:
                        cur, cur_t = cand, t
        if log:
            log("autotune: pool %d -> %s, %.1f ms" % (pool, cur, cur_t * 1e3))
        if best is None or cur_t < best["seconds"]:
            best = {"max_paths": pool, "tuning": cur, "seconds": cur_t}
        eng.close()
    best["trials"] = len(trials)
    return best
