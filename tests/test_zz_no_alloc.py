"""pt_trace_closest_noalloc (the allocation-free closest-hit entry point, on a fixed
scratch array per GPU) called through the C-ABI: its hits equal pt_trace_closest's bit
for bit on two calls in a row, the first large enough for the coherence order (which
uses the scratch's sort buffers), the second smaller; larger launches are refused."""
import ctypes

import numpy as np
import pytest

from gpu_util import ctx, log
from petrichor_amd import scenes

pytestmark = pytest.mark.gpu
MAX_RAYS = 1 << 21


def rays_in_box(sc, n, seed):
    t = sc["tri_pos"].reshape(-1, 3)
    rng = np.random.default_rng(seed)
    lo, hi = t.min(0), t.max(0)
    r = np.zeros((n, 8), np.float32)
    r[:, 0:3] = rng.uniform(lo, hi, (n, 3))
    d = rng.normal(size=(n, 3))
    r[:, 4:7] = d / np.linalg.norm(d, axis=1, keepdims=True)
    r[:, 3], r[:, 7] = 1e-5, np.finfo(np.float32).max
    return r


def test_noalloc_trace_equals_trace_closest(gpu):
    import torch
    sc = scenes.random_triangles(150000, seed=9, mu=-5.0)
    c = ctx()
    b = c.scene(sc).build_bvh(16)
    L = c._lib
    for n, seed in (((1 << 20) + 4321, 1), (200003, 2)):
        log("noalloc trace of %d rays" % n)
        rays = torch.from_numpy(rays_in_box(sc, n, seed)).cuda()
        t_ref, p_ref = b.trace_closest(rays)
        t = torch.full((n,), -1.0, dtype=torch.float32, device="cuda")
        p = torch.full((n,), -7, dtype=torch.int32, device="cuda")
        rc = L.pt_trace_closest_noalloc(b.handle, n, ctypes.c_void_p(rays.data_ptr()), ctypes.c_void_p(t.data_ptr()),
                                        ctypes.c_void_p(p.data_ptr()), None)
        assert rc == 0, L.pt_last_error(c.handle)
        torch.cuda.synchronize()
        assert torch.equal(p, p_ref)
        assert torch.equal(t.view(torch.int32), t_ref.view(torch.int32))
        assert 0.1 < (p >= 0).float().mean().item() < 1.0
    # above the scratch's limit: refused before any launch
    rc = L.pt_trace_closest_noalloc(b.handle, MAX_RAYS + 1, ctypes.c_void_p(rays.data_ptr()),
                                    ctypes.c_void_p(t.data_ptr()), ctypes.c_void_p(p.data_ptr()), None)
    assert rc == -1 and b"PT_NOALLOC_MAX_RAYS" in L.pt_last_error(c.handle)
