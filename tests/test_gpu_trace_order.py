"""Large trace launches (at least 2^20 rays) run in coherence order (pt_order.hip): their
hits equal those of the same rays traced in launches below the threshold, which trace in
the caller's order, and the oracle's on a sample."""
import numpy as np
import pytest

import oracle_lib as O
from gpu_util import ctx, log
from petrichor_amd import scenes

pytestmark = pytest.mark.gpu
N = (1 << 20) + 12345


def surface_rays(sc, n, seed):
    """Rays from random points inside the scene's box in uniform directions (incoherent),
    plus a block of rays from one point (their order is kept by the stable sort)."""
    t = sc["tri_pos"].reshape(-1, 3)
    rng = np.random.default_rng(seed)
    lo, hi = t.min(0), t.max(0)
    o = rng.uniform(lo, hi, (n, 3))
    o[: n // 8] = (lo + hi) / 2
    d = rng.normal(size=(n, 3))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    r = np.zeros((n, 8), np.float32)
    r[:, 0:3], r[:, 4:7] = o, d
    r[:, 3], r[:, 7] = 1e-5, np.finfo(np.float32).max
    r[::1000, 7] = 0.01  # short rays
    return r


@pytest.mark.parametrize("width", [2, 4, 8])
def test_large_launch_equals_small_launches_and_oracle(gpu, width):
    import torch
    sc = scenes.random_triangles(200000, seed=5, mu=-5.0)
    b = ctx().scene(sc).build_bvh(16, width=width)
    rays_h = surface_rays(sc, N, 11)
    rays = torch.from_numpy(rays_h).cuda()
    log("coherence-ordered launch of %d rays, width %d" % (N, width))
    t, p = b.trace_closest(rays)
    hit = b.trace_any(rays)
    t_small = torch.empty_like(t)
    p_small = torch.empty_like(p)
    hit_small = torch.empty_like(hit)
    step = 1 << 19  # below the threshold: caller's order
    for c in range(0, N, step):
        b.trace_closest(rays[c:c + step], t_small[c:c + step], p_small[c:c + step])
        b.trace_any(rays[c:c + step], hit_small[c:c + step])
    torch.cuda.synchronize()
    assert torch.equal(p, p_small) and torch.equal(t.view(torch.int32), t_small.view(torch.int32))
    assert torch.equal(hit, hit_small)
    assert 0.2 < (p >= 0).float().mean().item() < 1.0
    # the oracle on a sample (binary tree: the reference's walk)
    if width == 2:
        o = O.OracleScene(sc).use_tree(b.export_nodes() | {"prim_ids": b.export_order()})
        sel = np.random.default_rng(2).choice(N, 20000, replace=False)
        to, po, _ = o.trace_closest(rays_h[sel])
        assert np.array_equal(p.cpu().numpy()[sel], po)
        assert np.array_equal(t.cpu().numpy()[sel].view(np.uint32), to.view(np.uint32))
