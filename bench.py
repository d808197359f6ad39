"""Benchmark of the MI355X render path.

Default workload: the single large scene — 3,145,728 triangles (256 UV spheres of 4096
triangles over a 1024 x 1024-quad fBm terrain), 2048 x 1024 equirect sky with a sun lobe,
1920 x 1080, 64 spp, max depth 12, MIS integrator, per-pixel seeded samplers, BVH built
on the device and walked in the reference's order. One step is one full frame. `value` is BVH
queries (camera, bounce and shadow rays) per second over the timed steps, in Mrays/s.
The oracle renders a random sample of the frame's pixels on the host cores: that is the
CPU baseline, and the same pixels check the GPU frame (per-pixel relative error).

  python bench.py [--gpus 1] [--steps K] [--warmup W] [--cpu-seconds S] [--bvh-width 2|4|8]
  python bench.py --mode trace    # BVH build and traversal microbenchmark (10M triangles)
  python bench.py --mode batch    # 64 uneven jobs: batched vs one by one
  python bench.py --mode serve    # Poisson request stream against the render server

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
# wide layouts: one record per node visit (quantised child boxes and references): 128 B for
# 8 children, 64 B for 4
NODE8_B = 128
NODE4_B = 64


def node_bytes(width):
    return {2: NODE_B, 4: NODE4_B, 8: NODE8_B}[width]


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def host_cpu():
    """Threads for the CPU baseline and a description of the host. The reference's pool
    takes std::thread::hardware_concurrency() workers (Core/Thread/ThreadPool.cpp:19-26,
    Core/RenderSetting.h:17-21); a shared host gives this process a share of its cores
    (OMP_NUM_THREADS names it where it is set, else the affinity mask), and the baseline
    uses that share."""
    nproc = os.cpu_count() or 1
    try:
        affinity = len(os.sched_getaffinity(0))
    except (AttributeError, OSError):
        affinity = nproc
    share = affinity
    env = os.environ.get("OMP_NUM_THREADS", "")
    if env.isdigit() and int(env) > 0:
        share = min(share, int(env))
    model = ""
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    return share, {"nproc": nproc, "affinity": affinity, "cpu_model": model,
                   "omp_num_threads": env or None}


def cpu_baseline(sc, tree, width, height, spp, depth, seconds):
    """The CPU oracle (clean-room restatement of the reference renderer, tile pool of
    std::threads as Core/Petrichor.cpp) on a random pixel sample of the same frame: one
    random permutation of the pixels, rendered in consecutive slices, so no pixel is
    rendered twice."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import oracle_lib as O
    threads, host = host_cpu()
    o = O.OracleScene(sc).use_tree(tree)
    order = np.random.default_rng(12345).permutation(width * height).astype(np.int32)
    n = 2048
    total_q, total_t, total_px = 0, 0.0, 0
    work = np.zeros(9, np.int64)
    pixels, rgbs = [], []
    while True:
        pix = np.sort(order[total_px:total_px + n])
        w = np.zeros(9, np.int64)
        t0 = time.perf_counter()
        rgb, q = o.render(width, height, spp, depth, 0, 1, threads, pixels=pix, work=w)
        dt = time.perf_counter() - t0
        pixels.append(pix)
        rgbs.append(rgb)
        total_q += q
        total_t += dt
        total_px += n
        work += w
        if total_t >= seconds or total_px >= width * height:
            break
        n = int(min(max(n * 2, n * (seconds - total_t) / max(dt, 1e-3)), width * height - total_px, 1 << 20))
        n = max(n, 1)
    pix_all = np.concatenate(pixels)
    assert len(np.unique(pix_all)) == len(pix_all)
    return {
        "value": total_q / total_t / 1e6, "unit": "Mrays/s", "cores": threads, "kind": "port",
        "sample": "%d distinct random pixels of the %dx%d frame at full %d spp / depth %d (%.1f s, %d queries)"
                  % (total_px, width, height, spp, depth, total_t, total_q),
        "host": host,
    }, work, pix_all, np.concatenate(rgbs)


def cpu_trace_baseline(sc, tree, rays_host, seconds):
    """The CPU oracle's closest hit (restatement of BinnedSAHBVH::Intersect) on a sample of
    the same rays through the same tree. ctypes releases the GIL during the call, so the
    chunks run in parallel, one worker thread per core."""
    from concurrent.futures import ThreadPoolExecutor
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import oracle_lib as O
    threads, host = host_cpu()
    o = O.OracleScene(sc).use_tree(tree)
    n = 1 << 14
    done, total_t = 0, 0.0
    pool = ThreadPoolExecutor(threads)
    ts, ps = [], []
    while True:
        batch = rays_host[done:done + n]
        chunks = np.array_split(batch, threads)
        t0 = time.perf_counter()
        for t_, p_, _ in pool.map(o.trace_closest, chunks):
            ts.append(t_)
            ps.append(p_)
        dt = time.perf_counter() - t0
        done += len(batch)
        total_t += dt
        if total_t >= seconds or done >= len(rays_host):
            break
        n = int(min(max(n * 2, n * (seconds - total_t) / max(dt, 1e-3)), len(rays_host) - done))
    pool.shutdown()
    return {"value": done / total_t / 1e6, "unit": "Mrays/s", "cores": threads, "kind": "port",
            "sample": "%d rays (a random half-primary, half-surface sample of the 2^26), closest hit, %.1f s"
                      % (done, total_t), "host": host}, np.concatenate(ts), np.concatenate(ps)


def trace_bench(args):
    """BVH build and traversal microbenchmark: random triangles with centroids uniform in the
    unit cube and lognormal(-6, 1) edge lengths, binned SAH with 32 bins and leaves of at
    most 4, half pinhole primary rays and half surface rays with uniform directions;
    closest hit and any hit timed separately."""
    import torch
    import petrichor_amd as P
    from petrichor_amd import scenes
    N = args.triangles
    width = args.bvh_width or 8
    log("generating %d triangles" % N)
    sc = scenes.random_triangles(N, seed=3, mu=-6.0, sigma=1.0)
    ctx = P.Context(0)
    s = ctx.scene(sc)
    builds = []
    bvh = None
    for i in range(args.warmup + args.steps):
        if bvh is not None:
            bvh.close()
        bvh = s.build_bvh(32, 4, width=width, builder=P.BUILD_DEVICE)
        if args.threads_per_ray != 1 or args.leaf_batch:
            bvh.set_trace_tuning(threads_per_ray=args.threads_per_ray, leaf_batch=args.leaf_batch)
        info = bvh.info()
        if i >= args.warmup:
            builds.append(info["build_ms"])
        log("device build %d: %.1f ms, %d nodes, depth %d" % (i, info["build_ms"], info["num_nodes"],
                                                              info["max_depth"]))
    build_ms = float(np.median(builds))
    nr = 1 << args.log2_rays
    gen = torch.Generator(device="cuda")
    gen.manual_seed(7)
    w = 1 << ((args.log2_rays - 1) // 2 + 1)
    h = (nr // 2) // w
    rays = torch.empty((nr, 8), dtype=torch.float32, device="cuda")
    rays[: nr // 2] = scenes.primary_rays_device(sc["camera"], w, h, gen)
    tri_dev = torch.from_numpy(sc["tri_pos"]).cuda()
    rays[nr // 2:] = scenes.surface_rays_device(tri_dev, nr - nr // 2, gen, 1e-5)
    del tri_dev
    t_out = torch.empty(nr, dtype=torch.float32, device="cuda")
    p_out = torch.empty(nr, dtype=torch.int32, device="cuda")
    hit_out = torch.empty(nr, dtype=torch.int32, device="cuda")
    stream = torch.cuda.current_stream()

    def timed(fn):
        for _ in range(args.warmup):
            fn()
        torch.cuda.synchronize()
        evs = []
        t0 = time.perf_counter()
        for _ in range(args.steps):
            a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            a.record(stream)
            fn()
            b.record(stream)
            evs.append((a, b))
        torch.cuda.synchronize()
        wall = time.perf_counter() - t0
        return wall, [a.elapsed_time(b) for a, b in evs]

    wall_c, ms_c = timed(lambda: bvh.trace_closest(rays, t_out, p_out, stream=stream.cuda_stream))
    wall_a, ms_a = timed(lambda: bvh.trace_any(rays, hit_out, stream=stream.cuda_stream))
    hits = int((p_out >= 0).sum().item())
    assert torch.equal(hit_out, (p_out >= 0).int()), "any hit must agree with closest hit"
    log("closest: %.2f ms/launch, any: %.2f ms/launch, %d hits" % (np.mean(ms_c), np.mean(ms_a), hits))
    _, _, st = bvh.trace_stats(rays)
    n_node, n_tri = st[0] / nr, st[1] / nr
    node_b = node_bytes(width)
    b_ray = RAY_IN_B + HIT_OUT_B + n_node * node_b + n_tri * TRI_B
    avg_ms = float(np.mean(ms_c))
    achieved = nr * b_ray / (avg_ms * 1e-3) / 1e9
    roofline = {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": None,
                "kernel": "k_trace_persistent (closest hit, one launch over all rays)",
                "bytes_per_query": round(b_ray, 1), "n_node": round(n_node, 2), "n_tri": round(n_tri, 3),
                "s_node": node_b, "s_tri": TRI_B, "avg_launch_ms": round(avg_ms, 3), "queries_per_launch": nr}
    cpu = None
    parity = None
    if not args.no_cpu:
        log("cpu baseline")
        tree = bvh.export_nodes() | {"prim_ids": bvh.export_order()}
        sel = torch.randperm(nr, generator=torch.Generator().manual_seed(11))[: 1 << 21].cuda()
        sample = rays[sel].cpu().numpy()
        cpu, t_ref, p_ref = cpu_trace_baseline(sc, tree, sample, args.cpu_seconds)
        log("cpu:", cpu)
        # the same rays against the oracle (reference walk order on the same binary tree)
        k = len(t_ref)
        t_gpu = t_out[sel[:k]].cpu().numpy()
        p_gpu = p_out[sel[:k]].cpu().numpy()
        parity = {"rays_checked": int(k),
                  "t_bit_identical": int((t_gpu.view(np.uint32) == t_ref.view(np.uint32)).sum()),
                  "prim_identical": int((p_gpu == p_ref).sum())}
        log("parity vs oracle:", parity)
    value = nr * args.steps / wall_c / 1e6
    res = {
        "metric": METRIC, "value": round(value, 2), "unit": "Mrays/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": round(wall_c / args.steps * 1e3, 2), "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f32", "data": "synthetic",
        "config": {"workload": "BVH microbenchmark: %d random triangles (centroids U[0,1]^3, edges lognormal(-6,1)), "
                               "2^%d rays (half pinhole primary, half surface rays with uniform directions), "
                               "closest hit" % (N, args.log2_rays),
                   "triangles": N, "rays": nr, "bvh": "binned SAH, 32 bins, leaf<=4, device builder, %s"
                                                   % ("collapsed to %d-wide nodes" % width if width != 2
                                                      else "binary (reference topology and walk order)"),
                   "threads_per_ray": args.threads_per_ray, "leaf_batch": args.leaf_batch or 3},
        "closest_mrays_s": round(value, 2),
        "any_mrays_s": round(nr * args.steps / wall_a / 1e6, 2),
        "closest_ms_per_launch": round(avg_ms, 3), "any_ms_per_launch": round(float(np.mean(ms_a)), 3),
        "hit_fraction": round(hits / nr, 4),
        "bvh_build_ms": round(build_ms, 1), "bvh_build_mprims_s": round(N / build_ms / 1e3, 1),
        "bvh_nodes": info["num_nodes"], "bvh_depth": info["max_depth"],
        "roofline": roofline,
        "cpu_baseline": cpu,
        "parity_vs_oracle": parity,
    }
    print(json.dumps(res), flush=True)


def run_batch(args):
    """Batched form: jobs of uneven scene size, resolution and spp in one graph-replayed
    multi-job engine (autotuned), against the same jobs rendered one by one with plain
    launches (one synchronisation per iteration) and with the single-job graph path.
    Returns the result line's dict."""
    import torch
    import petrichor_amd as P
    from petrichor_amd import autotune, scenes
    rng = np.random.default_rng(2024)
    specs = scenes.job_specs(args.jobs, rng, args.tris_max, args.spp_max)
    env = scenes.sky_env()
    ctx = P.Context(0)
    t0 = time.perf_counter()
    # the autotuner picks the node width unless --bvh-width fixes it
    # the binary walk is the reference's; the 8-wide one only on request (--bvh-width 8)
    widths = (args.bvh_width,) if args.bvh_width else ((2, 4, 8) if args.tune_widths else (2,))
    bvhs_by_width = {w: [] for w in widths}
    for i, sp in enumerate(specs):
        sc = scenes.scaled_scene(sp["tris"], seed=1000 + i, env=env)
        sp["tris"] = int(len(sc["tri_mat"]))
        s_ = ctx.scene(sc)
        for w in bvhs_by_width:
            bvhs_by_width[w].append(s_.build_bvh(16, width=w))
    log("%d scenes (%d triangles) generated, uploaded and built in %.1f s"
        % (len(specs), sum(sp["tris"] for sp in specs), time.perf_counter() - t0))

    def outs():
        return [torch.empty((sp["height"], sp["width"], 3), dtype=torch.float32, device="cuda") for sp in specs]

    def jobs_with(o, bvhs):
        return [P.Job(b, sp["width"], sp["height"], sp["spp"], sp["depth"], P.INTEGRATOR_MIS, 7 + i, out=o[i])
                for i, (b, sp) in enumerate(zip(bvhs, specs))]

    out_b = outs()
    if args.no_tune:
        w0 = widths[0]
        best = {"width": w0, "jobs": jobs_with(out_b, bvhs_by_width[w0]), "max_paths": args.pool,
                "tuning": dict(autotune.DEFAULT, admission=args.admission, solo_log2_samples=args.solo),
                "trials": 0, "seconds": 0.0}
    else:
        best = autotune.tune(ctx, None, log=log, by_width={w: jobs_with(out_b, bvhs_by_width[w]) for w in widths},
                             exact_only=widths == (2,))
    jobs = best["jobs"]
    bvhs = bvhs_by_width[best["width"]]
    eng = P.Engine(ctx, best["max_paths"], max_jobs=len(jobs))
    eng.set_tuning(**best["tuning"])
    eng.render_batch(autotune.calibration_jobs(jobs))  # warm-up
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    st = eng.render_batch(jobs)
    torch.cuda.synchronize()
    t_batch = time.perf_counter() - t0
    log("batched: %.2f s, %d queries, %d iterations, %d syncs" % (t_batch, st["queries"], st["iterations"],
                                                                   st["syncs"]))
    if args.batch_only:
        return {"batched_jobs_s": round(len(specs) / t_batch, 3), "batched_s": round(t_batch, 3),
                "queries": int(st["queries"])}
    out_p = outs()
    t0 = time.perf_counter()
    q_plain = 0
    per_plain, per_q = [], []
    for i, (b, sp) in enumerate(zip(bvhs, specs)):
        t1 = time.perf_counter()
        _, s1 = b.render(sp["width"], sp["height"], sp["spp"], sp["depth"], P.INTEGRATOR_MIS, 7 + i, out=out_p[i],
                         plain=True)
        torch.cuda.synchronize()
        per_plain.append(time.perf_counter() - t1)
        per_q.append(s1["queries"])
        q_plain += s1["queries"]
    t_plain = time.perf_counter() - t0
    log("plain launches, one job at a time: %.2f s" % t_plain)
    out_g = outs()
    t0 = time.perf_counter()
    per_graph = []
    for i, (b, sp) in enumerate(zip(bvhs, specs)):
        t1 = time.perf_counter()
        b.render(sp["width"], sp["height"], sp["spp"], sp["depth"], P.INTEGRATOR_MIS, 7 + i, out=out_g[i])
        torch.cuda.synchronize()
        per_graph.append(time.perf_counter() - t1)
    t_graph = time.perf_counter() - t0
    log("single-job graph path, one job at a time: %.2f s" % t_graph)
    # Amdahl bound of the mix: the largest jobs fill the GPU on their own, so batching gains
    # little on them; if every other job cost nothing, batching could not beat
    # plain_total / (time of the largest jobs alone at their single-job graph speed).
    order = np.argsort([-(sp["width"] * sp["height"] * sp["spp"]) for sp in specs])
    big = order[:4]
    amdahl = {
        "largest_4_jobs": [specs[k] | {"plain_s": round(per_plain[k], 3), "graph_s": round(per_graph[k], 3)}
                           for k in big],
        "largest_4_share_of_plain_time": round(sum(per_plain[k] for k in big) / t_plain, 3),
        "largest_4_share_of_queries": round(sum(per_q[k] for k in big) / q_plain, 3),
        "bound_if_other_jobs_were_free": round(t_plain / sum(per_graph[k] for k in big), 2),
        "bound_at_batched_throughput": round(t_plain / (sum(per_q[k] for k in big) / (st["queries"] / t_batch)), 2),
    }
    log("amdahl:", {k: v for k, v in amdahl.items() if k != "largest_4_jobs"})
    same_plain = all(torch.equal(a, b) for a, b in zip(out_b, out_p))
    same_graph = all(torch.equal(a, b) for a, b in zip(out_b, out_g))
    assert same_plain and same_graph, "batched images must equal the per-job images bit for bit"
    assert q_plain == st["queries"]
    n = len(specs)
    res = {
        "metric": METRIC, "value": round(n / t_batch, 3), "unit": "jobs/s", "n_gpus": 1, "steps": 1, "warmup": 1,
        "ms_per_step": round(t_batch * 1e3, 1), "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f32", "data": "synthetic",
        "config": {"workload": "batched form: %d jobs, scenes of 1e3-%.0e triangles (scaled single-large-scene "
                               "generator, 2048x1024 sky), resolutions 256^2/512^2/1280x720/1920x1080, spp "
                               "log-uniform 4-%d, depth 8-12, MIS" % (n, args.tris_max, args.spp_max),
                   "jobs": n, "triangles_total": int(sum(sp["tris"] for sp in specs)),
                   "samples_total": int(sum(sp["width"] * sp["height"] * sp["spp"] for sp in specs))},
        "batched_jobs_s": round(n / t_batch, 3),
        "plain_launch_jobs_s": round(n / t_plain, 3),
        "single_graph_jobs_s": round(n / t_graph, 3),
        "speedup_vs_plain_launches": round(t_plain / t_batch, 2),
        "speedup_vs_single_graph": round(t_graph / t_batch, 2),
        "batched_mrays_s": round(st["queries"] / t_batch / 1e6, 1),
        "queries": int(st["queries"]),
        "images_bit_identical": bool(same_plain and same_graph),
        "autotune": {"bvh_width": best["width"], "max_paths": best["max_paths"], "tuning": best["tuning"],
                     "trials": best["trials"],
                     "calibration_ms": round(best["seconds"] * 1e3, 1)},
        "amdahl": amdahl,
        "jobs_detail": specs,
    }
    return res


def run_serve(args):
    """Served form: Poisson arrivals of render requests (batched-form job distribution, a
    quarter reusing a resident scene) against a render server with an HBM budget; latency
    is arrival to final image on the host. Returns the result line's dict."""
    import threading
    from concurrent.futures import ThreadPoolExecutor
    import petrichor_amd as P
    from petrichor_amd import scenes
    rng = np.random.default_rng(99)
    n_req = int(rng.poisson(args.rate * args.duration))
    specs = scenes.job_specs(n_req, rng, args.tris_max, args.spp_max)
    arrivals = np.sort(rng.uniform(0, args.duration, n_req))
    reuse = rng.uniform(0, 1, n_req) < 0.25
    env = scenes.sky_env()
    # geometry of the new scenes is generated before the stream starts (it stands for
    # scene files a client would upload); upload and BVH build happen per request
    t0 = time.perf_counter()
    pool_n = min(n_req, 48)
    geo = [scenes.scaled_scene(specs[i]["tris"], seed=5000 + i, env=env) for i in range(pool_n)]
    log("%d requests over %.0f s; %d scene geometries generated in %.1f s" % (n_req, args.duration, pool_n,
                                                                            time.perf_counter() - t0))
    srv = P.Server(int(args.budget_gb * (1 << 30)), max_paths=1 << 22, max_jobs=128, iterations_per_sync=8,
                   solo_log2_samples=args.serve_solo)
    lock = threading.Lock()
    resident = []          # scene ids added so far (some may be evicted later)
    outstanding = {}       # request id -> (index, arrival, submission), seconds from the start
    lat, queue_ms, failed, evicted_retry = [], [], [0], [0]

    def submit(i, t_arr):
        sp = specs[i]
        sid = None
        with lock:
            live = [x for x in resident if srv.resident(x)]
            if reuse[i] and live:
                sid = live[int(rng.integers(0, len(live)))]
        if sid is None:
            sid = srv.add_scene(geo[i % pool_n])
            with lock:
                resident.append(sid)
        args_ = (sp["width"], sp["height"], sp["spp"], sp["depth"], P.INTEGRATOR_MIS, 11 + i)
        try:
            rid = srv.submit(sid, *args_)
        except P.PtError:
            # the scene was evicted between the check and the submission: upload it again
            with lock:
                evicted_retry[0] += 1
            sid = srv.add_scene(geo[i % pool_n])
            rid = srv.submit(sid, *args_)
        with lock:
            outstanding[rid] = (i, t_arr, time.perf_counter() - t_start)

    def collect(deadline):
        while True:
            with lock:
                items = list(outstanding.items())
            if not items and arrivals_done[0]:
                return
            if time.perf_counter() - t_start > deadline:
                with lock:
                    failed[0] += len(outstanding)
                return
            for rid, (i, t_arr, t_sub) in items:
                info = srv.poll(rid)
                if info["state"] == P.REQ_DONE:
                    srv.fetch(rid)
                    with lock:
                        del outstanding[rid]
                        lat.append(info["latency_ms"] + (t_sub - t_arr) * 1e3)
                        queue_ms.append(info["queue_ms"])
                elif info["state"] == P.REQ_FAILED:
                    with lock:
                        del outstanding[rid]
                        failed[0] += 1
            time.sleep(0.002)

    arrivals_done = [False]
    pool = ThreadPoolExecutor(16)
    t_start = time.perf_counter()
    collector = threading.Thread(target=collect, args=(args.duration + args.drain,))
    collector.start()
    futs = []
    next_log = 10.0
    for i in range(n_req):
        dt = arrivals[i] - (time.perf_counter() - t_start)
        if dt > 0:
            time.sleep(dt)
        futs.append(pool.submit(submit, i, arrivals[i]))
        if arrivals[i] > next_log:
            st = srv.stats()
            log("t=%.0f s: %d submitted, %d done, %d running, %d queued" % (arrivals[i], st["submitted"],
                                                                             st["completed"], st["running"],
                                                                             st["queued"]))
            next_log += 10.0
    for f in futs:
        f.result()
    arrivals_done[0] = True
    while collector.is_alive():
        collector.join(10.0)
        st = srv.stats()
        log("draining: %d done, %d running, %d queued" % (st["completed"], st["running"], st["queued"]))
    wall = time.perf_counter() - t_start
    pool.shutdown()
    st = srv.stats()
    srv.close()
    lat_a = np.array(lat)
    res = {
        "metric": METRIC, "value": round(len(lat) / wall, 3), "unit": "jobs/s", "n_gpus": 1, "steps": 1,
        "warmup": 0, "ms_per_step": round(wall * 1e3, 1), "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f32", "data": "synthetic",
        "config": {"workload": "served form: Poisson %.1f req/s for %.0f s, batched-form job distribution "
                               "(spp <= %d, scenes <= %.0e triangles), 25%% of requests on a resident scene, "
                               "%.0f GB budget" % (args.rate, args.duration, args.spp_max, args.tris_max,
                                                   args.budget_gb),
                   "requests": n_req},
        "completed": len(lat), "failed": failed[0],
        "p50_latency_ms": round(float(np.percentile(lat_a, 50)), 1) if len(lat_a) else None,
        "p99_latency_ms": round(float(np.percentile(lat_a, 99)), 1) if len(lat_a) else None,
        "p50_queue_ms": round(float(np.percentile(queue_ms, 50)), 1) if queue_ms else None,
        "mrays_s": round(st["queries"] / wall / 1e6, 1),
        "busy_fraction": round(st["busy_ms"] / (wall * 1e3), 3),
        "evicted_scenes": st["evicted_scenes"], "resubmitted_after_eviction": evicted_retry[0],
        "peak_note": "latency counts arrival to final image on the host, scene upload and BVH build included",
    }
    return res


def tonemap_bench(args):
    """The 8-bit output kernel (pt_tonemap_rgb8) on an image of --width x --height HDR
    pixels resident in HBM: Mpixels/s, HIP-event launch times on the launch stream, the
    HBM roofline (12 bytes read + 3 written per pixel) and the oracle on the host cores."""
    import torch
    import petrichor_amd as P
    W, H = args.width, args.height
    n = W * H
    gen = torch.Generator(device="cuda")
    gen.manual_seed(5)
    # log-normal HDR values around 0.3 (the range where the bytes change), plus exact zeros
    img = torch.exp(torch.randn((H, W, 3), generator=gen, device="cuda") * 1.5 - 1.2)
    img.view(-1)[::97] = 0.0
    ctx = P.Context(0)
    out = torch.empty((H, W, 3), dtype=torch.uint8, device="cuda")
    st = torch.cuda.current_stream()
    for _ in range(max(args.warmup, 1)):
        ctx.tonemap_rgb8(img, out, stream=st.cuda_stream)
    torch.cuda.synchronize()
    steps = max(args.steps, 1) * 10
    ev = [torch.cuda.Event(enable_timing=True) for _ in range(2)]
    t0 = time.perf_counter()
    ev[0].record(st)
    for _ in range(steps):
        ctx.tonemap_rgb8(img, out, stream=st.cuda_stream)
    ev[1].record(st)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    launch_ms = ev[0].elapsed_time(ev[1]) / steps
    alg = 15 * n
    achieved = alg / (launch_ms * 1e-3) / 1e9
    got = out.cpu().numpy().reshape(-1, 3)
    cpu = None
    parity = None
    if not args.no_cpu:
        sys.path.insert(0, os.path.join(ROOT, "tests"))
        import oracle_lib as O
        sample = min(n, 1 << 24)
        host = img.view(-1, 3)[:sample].cpu().numpy()
        threads, hinfo = host_cpu()
        t1 = time.perf_counter()
        ref = O.tonemap_rgb8(host, threads)
        cpu_s = time.perf_counter() - t1
        cpu = {"value": round(sample / cpu_s / 1e6, 2), "unit": "Mpixels/s", "cores": threads, "kind": "port",
               "sample": "the first %d pixels of the same image" % sample, "host": hinfo}
        parity = {"pixels_checked": sample, "bytes_equal": int((got[:sample] == ref).sum()),
                  "bytes_checked": int(ref.size)}
    traffic = None  # rocprofv3 FETCH_SIZE + WRITE_SIZE per launch, committed for one image size
    tf = os.path.join(ROOT, "profiles", "pmc_traffic_tonemap_%dx%d.json" % (W, H))
    if os.path.exists(tf):
        with open(tf) as f:
            traffic = json.load(f).get("hbm_bytes_per_launch")
    res = {"metric": "tonemap Mpixels/s (8-bit PNG bytes of an HDR image in HBM)",
           "value": round(n * steps / wall / 1e6, 1), "unit": "Mpixels/s", "n_gpus": 1, "steps": steps,
           "warmup": max(args.warmup, 1), "ms_per_step": round(wall / steps * 1e3, 4), "higher_is_better": True,
           "scaling": "weak", "vs_baseline": None, "dtype": "f32 in, u8 out", "data": "synthetic",
           "config": {"workload": "ACES + RTCamp7 + degamma to 8 bits of a %dx%d RGB float image" % (W, H),
                      "width": W, "height": H},
           "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                        "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic,
                        "kernel": "k_tonemap_rgb8", "algorithmic_bytes_per_launch": alg,
                        "avg_launch_ms": round(launch_ms, 5)},
           "cpu_baseline": cpu, "parity_vs_oracle": parity}
    print(json.dumps(res), flush=True)


def aov_bench(args):
    """The denoising AOVs (pt_render_aov) of the single large scene at --width x --height:
    --spp denoising samples per pixel, albedo then normal; Mrays/s and the oracle's rate on
    a pixel sample."""
    import torch
    import petrichor_amd as P
    from petrichor_amd import scenes
    sc = scenes.spheres_scene(seed=2)
    ctx = P.Context(0)
    bvh = ctx.scene(sc).build_bvh(16)
    W, H, SPP, D = args.width, args.height, args.spp, args.depth
    out = torch.empty((H, W, 3), dtype=torch.float32, device="cuda")
    res = {}
    for name, aov in (("albedo", P.AOV_DENOISING_ALBEDO), ("normal", P.AOV_DENOISING_NORMAL)):
        bvh.render_aov(W, H, SPP, D, aov, 1, out=out)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        q = 0
        for _ in range(max(args.steps, 1)):
            _, st = bvh.render_aov(W, H, SPP, D, aov, 1, out=out)
            q += st["queries"]
        torch.cuda.synchronize()
        wall = time.perf_counter() - t0
        res[name] = {"mrays_s": round(q / wall / 1e6, 1), "ms": round(wall / max(args.steps, 1) * 1e3, 2),
                     "queries": int(q / max(args.steps, 1))}
        log(name, res[name])
    cpu = None
    if not args.no_cpu:
        sys.path.insert(0, os.path.join(ROOT, "tests"))
        import oracle_lib as O
        o = O.OracleScene(sc).use_tree(bvh.export_nodes() | {"prim_ids": bvh.export_order()})
        rng = np.random.default_rng(3)
        pix = rng.choice(W * H, 4096, replace=False).astype(np.int32)
        threads, hinfo = host_cpu()
        t1 = time.perf_counter()
        ref, q = o.render(W, H, SPP, D, 2, 1, threads, pixels=pix)
        cpu_s = time.perf_counter() - t1
        img = bvh.render_aov(W, H, SPP, D, P.AOV_DENOISING_ALBEDO, 1)[0].reshape(-1, 3)[pix]
        cpu = {"value": round(q / cpu_s / 1e6, 2), "unit": "Mrays/s", "cores": threads, "kind": "port",
               "sample": "albedo AOV of 4096 random pixels", "host": hinfo, "pixels_bit_identical": int(np.all(img == ref, axis=1).sum())}
    print(json.dumps({"metric": "denoising AOV Mrays/s (albedo, normal)", "value": res["albedo"]["mrays_s"],
                      "unit": "Mrays/s", "n_gpus": 1, "steps": args.steps, "higher_is_better": True,
                      "dtype": "f32", "data": "synthetic",
                      "config": {"workload": "denoising AOVs of the single large scene, %dx%d, %d spp, depth %d"
                                             % (W, H, SPP, D)},
                      "aovs": res, "cpu_baseline": cpu}), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--compare-wide", action="store_true",
                    help="frame mode: also render the frame through the 4- and 8-wide trees")
    ap.add_argument("--separate-launches", action="store_true",
                    help="frame mode: also a frame with every launch on its own (per-kernel split)")
    ap.add_argument("--no-three", action="store_true", help="frame mode: skip the three-stream frames")
    ap.add_argument("--no-batch", action="store_true", help="frame mode: skip the batched form")
    ap.add_argument("--no-serve", action="store_true", help="frame mode: skip the served form")
    ap.add_argument("--serve-seconds", type=float, default=20.0,
                    help="frame mode: seconds of Poisson arrivals of the served form")
    ap.add_argument("--width", type=int, default=1920)
    ap.add_argument("--height", type=int, default=1080)
    ap.add_argument("--spp", type=int, default=64)
    ap.add_argument("--depth", type=int, default=12)
    ap.add_argument("--mode", default="frame", choices=["frame", "trace", "batch", "serve", "copy", "tonemap", "aov"],
                    help="frame: the single large scene (default); trace: the BVH build and "
                         "traversal microbenchmark (10M triangles, 2^26 rays); copy: a plain 1 GiB "
                         "read + 1 GiB write elementwise kernel, the reference for scaling rocprofv3 "
                         "FETCH_SIZE / WRITE_SIZE; tonemap: the 8-bit output kernel on a "
                         "--width x --height image; aov: the denoising AOVs of the single "
                         "large scene")
    ap.add_argument("--triangles", type=int, default=10_000_000, help="trace mode: triangle count")
    ap.add_argument("--bvh-width", type=int, default=None, choices=[2, 4, 8],
                    help="BVH width: 2 walks the binary tree in the reference's order (frame default), "
                         "8 the 8-wide tree (traversal microbenchmark default); batch mode: both are "
                         "built and the autotuner picks one unless this is given")
    ap.add_argument("--serve-solo", type=int, default=23,
                    help="serve mode: requests of at least 2^k samples render alone on the server's single-job "
                         "engine beside the shared pool (0: none)")
    ap.add_argument("--threads-per-ray", type=int, default=1, choices=[1, 2],
                    help="frame and trace modes: lanes per ray of the binary walk (pt_bvh_set_trace_tuning)")
    ap.add_argument("--trace-chunk", type=int, default=64,
                    help="frame mode: rays a trace wave takes from its part's cursor at a time (64..4096)")
    ap.add_argument("--trace-blocks-per-cu", type=int, default=0, choices=range(0, 9),
                    help="frame mode: persistent trace blocks per CU (0: as many as the LDS stacks allow)")
    ap.add_argument("--leaf-batch", type=int, default=0, choices=[-1, 0, 1, 2, 3, 4],
                    help="trace/frame modes: phased binary walk, triangle steps once 1/2^(k-1) of a wave's rays "
                         "are in a leaf (0: the library's default, 3; -1: off)")
    ap.add_argument("--tune-widths", action="store_true",
                    help="batch mode: the autotuner also tries the 4- and 8-wide trees (not the reference's walk)")
    ap.add_argument("--jobs", type=int, default=64, help="batch mode: number of jobs")
    ap.add_argument("--no-tune", action="store_true", help="batch mode: default launch shapes, --pool path slots")
    ap.add_argument("--pool", type=int, default=1 << 23, help="batch mode with --no-tune: path slots")
    ap.add_argument("--solo", type=int, default=0,
                    help="batch mode with --no-tune: jobs of at least 2^SOLO samples render alone on the "
                         "single-job engine beside the shared one (0: none)")
    ap.add_argument("--admission", type=int, default=0, choices=[0, 1],
                    help="batch mode with --no-tune: 0 admits any job that fits, 1 waits for the running "
                         "jobs to hand out their samples")
    ap.add_argument("--batch-only", action="store_true", help="batch mode: only the batched run (no per-job runs)")
    ap.add_argument("--tris-max", type=float, default=2e6, help="batch/serve modes: largest scene")
    ap.add_argument("--spp-max", type=int, default=256, help="batch/serve modes: largest spp")
    ap.add_argument("--rate", type=float, default=20.0, help="serve mode: Poisson arrival rate (requests/s)")
    ap.add_argument("--duration", type=float, default=120.0, help="serve mode: seconds of arrivals")
    ap.add_argument("--drain", type=float, default=600.0, help="serve mode: longest wait for the backlog")
    ap.add_argument("--budget-gb", type=float, default=200.0, help="serve mode: HBM budget")
    ap.add_argument("--log2-rays", type=int, default=26, help="trace mode: log2 of the ray count")
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
    if args.mode == "trace":
        return trace_bench(args)
    if args.mode == "batch":
        print(json.dumps(run_batch(args)), flush=True)
        return
    if args.mode == "serve":
        print(json.dumps(run_serve(args)), flush=True)
        return
    if args.mode == "tonemap":
        return tonemap_bench(args)
    if args.mode == "aov":
        return aov_bench(args)
    log("generating scene")
    t0 = time.perf_counter()
    sc = scenes.spheres_scene(seed=2)
    log("scene: %d triangles in %.1f s" % (len(sc["tri_mat"]), time.perf_counter() - t0))
    ctx = P.Context(0)
    s = ctx.scene(sc)
    width = args.bvh_width or 2
    bvh = s.build_bvh(16, width=width)
    if args.threads_per_ray != 1 or args.leaf_batch or args.trace_blocks_per_cu or args.trace_chunk != 64:
        bvh.set_trace_tuning(trace_blocks_per_cu=args.trace_blocks_per_cu, trace_chunk=args.trace_chunk,
                             threads_per_ray=args.threads_per_ray, leaf_batch=args.leaf_batch)
    info = bvh.info()
    log("bvh:", info)
    W, H, SPP, D = args.width, args.height, args.spp, args.depth
    out = torch.empty((H, W, 3), dtype=torch.float32, device="cuda")
    for i in range(args.warmup):
        _, st = bvh.render(W, H, SPP, D, P.INTEGRATOR_MIS, 1, out=out, two_streams=True)
        log("warmup", i, st["queries"], "queries", "%.1f ms" % st["elapsed_ms"], st["iterations"], "iterations")
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    stats = []
    for i in range(args.steps):
        _, st = bvh.render(W, H, SPP, D, P.INTEGRATOR_MIS, 1, out=out, two_streams=True)
        stats.append(st)
        log("step", i, st["queries"], "queries", "%.1f ms" % st["elapsed_ms"], st["iterations"], "iterations")
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    queries = sum(s_["queries"] for s_ in stats)
    value = queries / wall / 1e6
    img = out.cpu().numpy()
    assert np.isfinite(img).all() and img.mean() > 0

    # The roofline of the traversal kernel on the production schedule: the same frame
    # (graph replays, two halves on two streams) with a launch log, once for the span of
    # every trace launch on the device clock and its rays, once with the traversal counted
    # (node visits, triangle tests), since counting slows the launches. The rays left to
    # the path-completion launches are counted and timed apart.
    logged = torch.empty_like(out)
    _, lst = bvh.render(W, H, SPP, D, P.INTEGRATOR_MIS, 1, out=logged, two_streams=True, launch_log=True)
    torch.cuda.synchronize()
    assert torch.equal(logged, out), "the logged frame must match the production frames"
    _, cst = bvh.render(W, H, SPP, D, P.INTEGRATOR_MIS, 1, out=logged, two_streams=True, launch_log=True,
                        trace_stats=True)
    torch.cuda.synchronize()
    assert torch.equal(logged, out), "the counted frame must match the production frames"
    # The same frame in three parts on three streams (PT_RENDER_THREE_STREAMS): reported
    # beside the headline. Its trace launches overlap each other, so their spans no
    # longer time the kernel, and the roofline stays on the two-stream schedule.
    three = None
    if not args.no_three:
        bvh.render(W, H, SPP, D, P.INTEGRATOR_MIS, 1, out=logged, three_streams=True)
        torch.cuda.synchronize()
        t3 = time.perf_counter()
        q3 = 0
        for i in range(2):
            _, st3 = bvh.render(W, H, SPP, D, P.INTEGRATOR_MIS, 1, out=logged, three_streams=True)
            q3 += st3["queries"]
        torch.cuda.synchronize()
        w3 = time.perf_counter() - t3
        three = {"mrays_s": round(q3 / w3 / 1e6, 2), "ms_per_frame": round(w3 / 2 * 1e3, 2), "frames": 2,
                 "image_bit_identical": bool(torch.equal(logged, out)),
                 "render": "pt_render with PT_RENDER_TWO_STREAMS | PT_RENDER_THREE_STREAMS (three parts of the "
                           "pixels on three streams)"}
        log("three streams:", three)
    L, C = lst["log"], cst["log"]
    log("logged frame:", L)
    assert C["trace_rays"] == L["trace_rays"] and C["trace_launches"] == L["trace_launches"]
    assert L["trace_rays"] + L["finish_rays"] == lst["queries"]
    node_b = node_bytes(width)
    s_ray = RAY_IN_B + PATH_B + HIT_OUT_B
    alg_bytes = L["trace_rays"] * s_ray + C["trace_nodes"] * node_b + C["trace_tris"] * TRI_B
    n_launch = max(L["trace_launches"], 1)
    avg_ms = L["trace_ms"] / n_launch
    achieved = alg_bytes / (L["trace_ms"] * 1e-3) / 1e9 if L["trace_ms"] > 0 else 0.0
    traffic = None
    tf = os.path.join(ROOT, "profiles", "pmc_traffic_frame_production_w%d.json" % width)
    if os.path.exists(tf):
        with open(tf) as f:
            traffic = json.load(f).get("hbm_bytes_per_launch")
    counters = None  # L2 hit rate and VALU activity of the trace launches (rocprofv3 --pmc)
    cf = os.path.join(ROOT, "profiles", "pmc_l2_valu_frame_production_w%d.json" % width)
    if os.path.exists(cf):
        with open(cf) as f:
            c = json.load(f)["k_trace_persistent"]
        counters = {"l2_hit": c["l2_hit"], "valu_active": c["valu_active"], "source": os.path.basename(cf)}
    roofline = {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic, "counters": counters,
                "kernel": "k_trace_persistent, every trace launch of the production frame (graph replays, "
                          "two halves on two streams)",
                "timing": "span of each launch's blocks on the device clock (wall_clock64), logged live; "
                          "rocprofv3 --kernel-trace of the same schedule in profiles/",
                "algorithmic_bytes_per_launch": round(alg_bytes / n_launch, 1),
                "bytes_per_query": round(alg_bytes / max(L["trace_rays"], 1), 1),
                "n_node": round(C["trace_nodes"] / max(L["trace_rays"], 1), 2),
                "n_tri": round(C["trace_tris"] / max(L["trace_rays"], 1), 3),
                "s_node": node_b, "s_tri": TRI_B, "s_ray": s_ray,
                "avg_launch_ms": round(avg_ms, 4), "launches": L["trace_launches"],
                "queries_per_launch": round(L["trace_rays"] / n_launch, 1),
                "path_completion": {"kernel": "k_finish_paths", "launches": L["finish_launches"],
                                    "rays": L["finish_rays"], "ms": round(L["finish_ms"], 3)}}

    kernels = None
    if args.separate_launches:
        # a frame with every launch on its own, timed by HIP events: the split of the time
        # between the trace kinds and shading (not the production schedule)
        timed = torch.empty_like(out)
        _, kst = bvh.render(W, H, SPP, D, P.INTEGRATOR_MIS, 1, out=timed, trace_stats=True)
        torch.cuda.synchronize()
        assert torch.equal(timed, out), "the instrumented frame must match the production frames"
        kernels = {k: {"items": v["items"], "launches": v["launches"], "ms": round(v["ms"], 3),
                       "node_visits": v["nodes"], "triangle_tests": v["tris"]} for k, v in kst["kernels"].items()}

    cpu = None
    parity = None
    if not args.no_cpu:
        log("cpu baseline")
        tree = bvh.export_nodes() | {"prim_ids": bvh.export_order()}
        cpu, work, pix, rgb_ref = cpu_baseline(sc, tree, W, H, SPP, D, args.cpu_seconds)
        log("cpu:", cpu)
        # the GPU frame against the oracle on the sampled pixels (per-pixel relative error)
        got = img.reshape(-1, 3)[pix].astype(np.float64)
        ref = rgb_ref.astype(np.float64)
        rel = np.abs(got - ref) / np.maximum(np.abs(ref), 1e-30)
        parity = {"distinct_pixels_checked": int(len(pix)), "max_rel_err": float(rel.max()),
                  "pixels_bit_identical": int(np.all(got == ref, axis=1).sum()),
                  "pixels_over_1e-4": int((rel.max(axis=1) > 1e-4).sum())}
        log("parity vs oracle:", parity)
    # The same frame through the 4- and 8-wide trees, on request (walks that differ from the
    # reference's on rays grazing a box within rounding of their hit). The exact frame equals
    # the oracle's bit for bit, so differences from it are differences from the oracle.
    other = None
    if width == 2 and args.compare_wide:
        other = {}
        for ww in (4, 8):
            bw = s.build_bvh(16, width=ww)
            outw = torch.empty_like(out)
            bw.render(W, H, SPP, D, P.INTEGRATOR_MIS, 1, out=outw, two_streams=True)
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            qw = 0
            for _ in range(2):
                _, stw = bw.render(W, H, SPP, D, P.INTEGRATOR_MIS, 1, out=outw, two_streams=True)
                qw += stw["queries"]
            torch.cuda.synchronize()
            tw = time.perf_counter() - t0
            imgw = outw.cpu().numpy()
            relw = np.abs(imgw.astype(np.float64) - img) / np.maximum(np.abs(img.astype(np.float64)), 1e-30)
            other["width%d" % ww] = {
                "bvh": "collapsed to %d-wide nodes (not the reference's walk)" % ww,
                "mrays_s": round(qw / tw / 1e6, 2), "ms_per_frame": round(tw / 2 * 1e3, 2),
                "bvh_build_ms": round(bw.info()["build_ms"], 1),
                "pixels_identical_to_exact_frame": int(np.all(imgw == img, axis=2).sum()),
                "pixels_over_1e-4": int((relw.max(axis=2) > 1e-4).sum()),
                "max_rel_err": float(relw.max())}
            log("%d-wide:" % ww, other["width%d" % ww])
            bw.close()
    del bvh, s
    torch.cuda.empty_cache()
    batched = served = None
    if not args.no_batch:
        log("batched form")
        br = run_batch(args)
        batched = {k: br[k] for k in ("batched_jobs_s", "plain_launch_jobs_s", "single_graph_jobs_s",
                                      "speedup_vs_plain_launches", "speedup_vs_single_graph", "batched_mrays_s",
                                      "images_bit_identical", "amdahl")}
        batched["workload"] = br["config"]["workload"]
        batched["autotune"] = {k: br["autotune"][k] for k in ("bvh_width", "max_paths", "tuning")}
    if not args.no_serve:
        log("served form")
        a2 = argparse.Namespace(**vars(args))
        a2.duration = args.serve_seconds
        sr = run_serve(a2)
        served = {k: sr[k] for k in ("value", "completed", "failed", "p50_latency_ms", "p99_latency_ms",
                                     "p50_queue_ms", "mrays_s", "busy_fraction", "evicted_scenes")}
        served["jobs_s"] = served.pop("value")
        served["workload"] = sr["config"]["workload"]
    res = {
        "metric": METRIC, "value": round(value, 2), "unit": "Mrays/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": round(wall / args.steps * 1e3, 2), "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f32", "data": "synthetic",
        "config": {"workload": "single large scene: %d triangles (256 spheres x 4096 + 1024^2-quad terrain), "
                               "2048x1024 env, %dx%d, %d spp, depth %d, MIS" % (len(sc["tri_mat"]), W, H, SPP, D),
                   "triangles": int(len(sc["tri_mat"])), "width": W, "height": H, "spp": SPP, "max_depth": D,
                   "integrator": "MIS",
                   "bvh": "binned SAH, 16 bins, leaf<=4, device builder, " +
                          ("collapsed to %d-wide nodes" % width if width != 2
                           else "binary (reference topology and walk order)"),
                   "threads_per_ray": args.threads_per_ray, "leaf_batch": args.leaf_batch or 3,
                   "trace_blocks_per_cu": args.trace_blocks_per_cu or 8, "trace_chunk": args.trace_chunk,
                   "render": "pt_render with PT_RENDER_TWO_STREAMS (two halves of the pixels on two streams)"},
        "queries_per_step": int(queries / args.steps),
        "iterations_per_step": int(stats[-1]["iterations"]),
        "bvh_build_ms": round(info["build_ms"], 1),
        "roofline": roofline,
        "cpu_baseline": cpu,
        "parity_vs_oracle": parity,
        "three_streams": three,
        "batched": batched,
        "served": served,
    }
    if kernels is not None:
        res["kernels_separate_launches"] = kernels
    if other is not None:
        res["wide_frames"] = other
    print(json.dumps(res), flush=True)

if __name__ == "__main__":
    main()
