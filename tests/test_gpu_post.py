"""The 8-bit output kernel (pt_tonemap_rgb8) against the oracle's Texture2D::Save bytes
(Core/Texture2D.cpp:18-91, 199-215): every byte equal."""
import re

import numpy as np
import pytest

import oracle_lib as O
import petrichor_amd as P
from gpu_util import ctx, log
from petrichor_amd import scenes
from test_cpu_tonemap import load_table

pytestmark = pytest.mark.gpu


def gpu_bytes(c, values):
    import torch
    x = torch.from_numpy(np.ascontiguousarray(values, np.float32).reshape(-1, 3)).cuda()
    out = c.tonemap_rgb8(x)
    torch.cuda.synchronize()
    return out.cpu().numpy()


def test_random_bit_patterns_equal_oracle(gpu):
    # every kind of float: negatives, subnormals, infinities, NaNs with any payload
    rng = np.random.default_rng(31)
    bits = rng.integers(0, 2 ** 32, 3 * (1 << 20) + 3, dtype=np.uint64).astype(np.uint32)
    vals = bits.view(np.float32).reshape(-1, 3)
    log("tonemap: %d random pixels" % vals.shape[0])
    got = gpu_bytes(ctx(), vals)
    assert np.array_equal(got, O.tonemap_rgb8(vals))


def test_run_boundaries_equal_oracle(gpu):
    start, _ = load_table()
    edges = np.unique(np.concatenate([start.ravel().astype(np.int64) + d for d in range(-3, 4)]))
    edges = edges[(edges >= 0) & (edges < 2 ** 32)].astype(np.uint32).view(np.float32)
    vals = np.repeat(edges[:, None], 3, axis=1)  # every channel sees every boundary
    log("tonemap: %d boundary pixels" % vals.shape[0])
    assert np.array_equal(gpu_bytes(ctx(), vals), O.tonemap_rgb8(vals))


def test_rendered_image_and_png_round_trip(gpu, tmp_path):
    import torch
    c = ctx()
    b = c.scene(scenes.cornell_box()).build_bvh()
    img, _ = b.render(96, 64, 8, 8, out=torch.empty((64, 96, 3), dtype=torch.float32, device="cuda"))
    rgb8 = c.tonemap_rgb8(img)
    torch.cuda.synchronize()
    want = O.tonemap_rgb8(img.cpu().numpy())
    got = rgb8.cpu().numpy()
    assert np.array_equal(got, want)
    assert 0 < got.mean() < 254
    P.save_png(str(tmp_path / "c.png"), got)
    assert np.array_equal(P.load_png_rgb8(str(tmp_path / "c.png")), want)


@pytest.mark.parametrize("n", [1, 2, 3, 4, 5, 7, 8, 9, 13, 1000, 4097])
@pytest.mark.parametrize("in_off,out_off", [(0, 0), (1, 0), (0, 1), (3, 2)])
def test_sizes_and_alignments(gpu, n, in_off, out_off):
    import torch
    c = ctx()
    rng = np.random.default_rng(n * 7 + in_off)
    vals = (10.0 ** rng.uniform(-3, 1, 3 * n)).astype(np.float32)
    big_in = torch.zeros(3 * n + 8, dtype=torch.float32, device="cuda")
    big_in[in_off:in_off + 3 * n] = torch.from_numpy(vals).cuda()
    big_out = torch.full((3 * n + 8,), 77, dtype=torch.uint8, device="cuda")
    src, dst = big_in[in_off:in_off + 3 * n], big_out[out_off:out_off + 3 * n]
    c.check(c._lib.pt_tonemap_rgb8(c.handle, n, src.data_ptr(), dst.data_ptr(), 1, None), "pt_tonemap_rgb8")
    torch.cuda.synchronize()
    out = big_out.cpu().numpy()
    assert np.array_equal(out[out_off:out_off + 3 * n], O.tonemap_rgb8(vals.reshape(-1, 3)).ravel())
    # nothing outside the image is written
    assert np.all(out[:out_off] == 77) and np.all(out[out_off + 3 * n:] == 77)


def test_host_buffers(gpu):
    # on_device = 0: host pointers, staged through device memory
    rng = np.random.default_rng(3)
    vals = (10.0 ** rng.uniform(-3, 1, (1001, 3))).astype(np.float32)
    out = np.zeros((1001, 3), np.uint8)
    c = ctx()
    c.check(c._lib.pt_tonemap_rgb8(c.handle, 1001, vals.ctypes.data, out.ctypes.data, 0, None), "pt_tonemap_rgb8")
    assert np.array_equal(out, O.tonemap_rgb8(vals))


def test_bad_arguments_are_rejected(gpu):
    import torch
    c = ctx()
    x = torch.zeros(3, dtype=torch.float32, device="cuda")
    assert c._lib.pt_tonemap_rgb8(c.handle, -1, x.data_ptr(), x.data_ptr(), 1, None) == -1
    assert c._lib.pt_tonemap_rgb8(c.handle, 1, None, x.data_ptr(), 1, None) == -1
    assert c._lib.pt_tonemap_rgb8(c.handle, 0, None, None, 1, None) == 0
    assert re.search("tonemap", c._lib.pt_last_error(c.handle).decode())
