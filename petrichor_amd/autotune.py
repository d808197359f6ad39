"""Autotuner of the batched engine.

Chooses, for a given set of render jobs, the BVH node width (when the jobs are given over
BVHs of several widths: 2, 4, 8), the path-pool size (the capacity of every ray queue), the graph
replays per host synchronisation, the lanes per ray and the phases of the binary walk and the launch
shapes of the trace and shading kernels (pt_tuning). It times pt_render_batch on a calibration batch - the same scenes and
resolutions at a few samples per pixel - first for each width and then for each pool size
with the default launch shapes, then with coordinate descent over the space below on the
best pool size. Last, on the full batch, it picks the size from which a job renders alone
on the context's single-job engine (the single-job kernels, beside the shared engine on a
second stream) rather than in the shared pool. Every candidate of one width renders the
same images (tuning only changes launch shapes and scheduling), so the fastest one is taken
as is.

Only the binary width walks the tree exactly as the reference (BinnedSAHBVH::Intersect,
Core/Accel/BinnedSAHBVH.cpp:272-374). The 4- and 8-wide walks can return another hit on
rays that graze a box within rounding of their hit (with 8-wide nodes 35 of 2,073,600
pixels of the single large scene differ by more than 1e-4), so they are candidates only
when the caller accepts that (exact_only=False).
"""
import time

DEFAULT = {"iterations_per_sync": 8, "trace_blocks_per_cu": 0, "trace_chunk": 64, "shade_blocks_per_cu": 8,
           "threads_per_ray": 1, "leaf_batch": 3}
# Thresholds tried for rendering the largest jobs alone on the single-job engine, in log2
# samples (0: every job in the shared engine). They are timed on the full batch, since the
# calibration batch's fewer samples per pixel would move every job across them.
SOLO = (0, 23, 25, 27)

SPACE = {
    # lanes per ray of the binary walk (the same walk and hits either way)
    "threads_per_ray": [1, 2],
    # phased binary walk: triangle steps once 1/2^(k-1) of a wave's rays are in a leaf (-1: off)
    "leaf_batch": [-1, 2, 3, 4],
    "iterations_per_sync": [2, 4, 8, 16, 32],
    "trace_blocks_per_cu": [2, 3, 4, 6],  # 0 (default): as many as the LDS stacks allow
    "trace_chunk": [64, 128, 256, 512],
    "shade_blocks_per_cu": [4, 8, 16],
}

# Path-pool sizes tried. Larger pools keep more jobs, and so more scenes, in flight at once,
# whose BVHs and path state then compete for L2 and the Infinity Cache (the 64-job batch:
# 2.71 s with 2M slots, 2.67 s with 4M, 3.38 s with 8M).
POOLS = (1 << 21, 1 << 22, 6 << 20)


def calibration_jobs(jobs, spp=2):
    """The jobs at `spp` samples per pixel with host outputs dropped (device outputs kept)."""
    from . import Job
    return [Job(j.bvh, j.width, j.height, min(spp, j.spp), j.max_bounces, j.integrator, j.seed, j.out) for j in jobs]


def tune(ctx, jobs, pool_sizes=None, passes=1, log=None, by_width=None, calib_spp=8, exact_only=True, solo=SOLO):
    """Returns {"width", "jobs", "max_paths", "tuning", "seconds", "trials"} for the fastest
    configuration found on the calibration batch of `jobs`. by_width (optional) maps a BVH
    width (2, 4, 8) to the same jobs over BVHs of that width; "jobs" is then the chosen list.
    exact_only: only binary BVHs (the reference's walk) are candidates. solo: thresholds
    (log2 samples) for rendering the largest jobs alone on the single-job engine, timed on
    the full batch last ("solo_seconds" then holds the full batch's time per threshold)."""
    from . import Engine
    by_width = dict(by_width) if by_width else {getattr(jobs[0].bvh, "width", 2): jobs}
    if exact_only:
        by_width = {w: j for w, j in by_width.items() if w == 2}
        if not by_width:
            raise ValueError("exact_only: the jobs need binary BVHs (width 2)")
    jobs = next(iter(by_width.values()))
    if pool_sizes is None:
        pool_sizes = list(POOLS)
    trials = []

    def timer(eng, calib, pool):
        def measure(t):
            eng.set_tuning(**t)
            t0 = time.perf_counter()
            eng.render_batch(calib)
            dt = time.perf_counter() - t0
            trials.append({"max_paths": pool, "tuning": dict(t), "seconds": dt})
            return dt
        return measure

    # node width: the default launch shapes on the middle pool size
    width = next(iter(by_width))
    if len(by_width) > 1:
        pool = pool_sizes[len(pool_sizes) // 2]
        eng = Engine(ctx, pool, max_jobs=max(len(jobs), 1))
        best_w = None
        for w, wj in sorted(by_width.items()):
            measure = timer(eng, calibration_jobs(wj, calib_spp), pool)
            measure(DEFAULT)  # first use of this width: graph instantiation and caches
            t = measure(DEFAULT)
            if log:
                log("autotune: width %d -> %.1f ms" % (w, t * 1e3))
            if best_w is None or t < best_w[1]:
                best_w = (w, t)
        eng.close()
        width = best_w[0]
    chosen = by_width[width]
    calib = calibration_jobs(chosen, calib_spp)
    # the pool size with the default launch shapes, then coordinate descent over the launch
    # shapes on the best pool
    best = None
    for pool in pool_sizes:
        eng = Engine(ctx, pool, max_jobs=max(len(chosen), 1))
        measure = timer(eng, calib, pool)
        measure(DEFAULT)  # first use: graph instantiation and caches
        t = measure(DEFAULT)
        if log:
            log("autotune: pool %d -> %.1f ms" % (pool, t * 1e3))
        if best is None or t < best["seconds"]:
            best = {"max_paths": pool, "tuning": dict(DEFAULT), "seconds": t}
        eng.close()
    pool = best["max_paths"]
    eng = Engine(ctx, pool, max_jobs=max(len(chosen), 1))
    measure = timer(eng, calib, pool)
    measure(DEFAULT)
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
    solo_seconds = {}
    if solo and len(solo) > 1:
        # first use of the single-job engine (allocation, graph capture) outside the timing
        eng.set_tuning(**dict(cur, solo_log2_samples=min(k for k in solo if k > 0)))
        eng.render_batch(calib)
        for k in solo:
            cand = dict(cur, solo_log2_samples=k)
            eng.set_tuning(**cand)
            t0 = time.perf_counter()
            eng.render_batch(chosen)
            solo_seconds[k] = time.perf_counter() - t0
            trials.append({"max_paths": pool, "tuning": cand, "seconds": solo_seconds[k], "full_batch": True})
            if log:
                log("autotune: solo 2^%d samples -> %.1f ms (full batch)" % (k, solo_seconds[k] * 1e3))
        cur = dict(cur, solo_log2_samples=min(solo_seconds, key=solo_seconds.get))
    eng.close()
    if log:
        log("autotune: pool %d -> %s, %.1f ms" % (pool, cur, cur_t * 1e3))
    best = {"max_paths": pool, "tuning": cur, "seconds": cur_t, "solo_seconds": solo_seconds}
    best["width"] = width
    best["jobs"] = chosen
    best["trials"] = len(trials)
    return best
