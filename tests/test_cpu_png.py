"""PNG decoding as Texture2D::Load reads textures (pt_png_decode): 8-bit RGB as stbi_load
gives it, bytes / 255, powf(x, 2.2) for colour textures. Files are encoded here with every
scanline filter, colour type and bit depth (1 to 16) the format allows."""
import ctypes
import struct
import zlib

import numpy as np
import pytest

from petrichor_amd import loader

LIBM = ctypes.CDLL("libm.so.6")
LIBM.powf.restype = ctypes.c_float
LIBM.powf.argtypes = [ctypes.c_float, ctypes.c_float]
# ApplyGamma of every byte value: powf(b / 255, 2.2) with the C library's powf
GAMMA = np.array([LIBM.powf(float(np.float32(b) / np.float32(255)), 2.2) for b in range(256)], np.float32)


def paeth(a, b, c):
    p = a + b - c
    pa, pb, pc = abs(p - a), abs(p - b), abs(p - c)
    return a if pa <= pb and pa <= pc else (b if pb <= pc else c)


def pack_row(values, depth):
    """One row of samples of 1, 2 or 4 bits, most significant bits first, padded to a byte."""
    out = bytearray((len(values) * depth + 7) // 8)
    for x, v in enumerate(values):
        bit = x * depth
        out[bit // 8] |= int(v) << (8 - depth - bit % 8)
    return out


def encode(samples, ctype, depth=8, plte=None, interlace=0):
    """PNG of samples [h, w, channels] (uint8 or uint16), row y filtered with type y % 5."""
    h, w, ch = samples.shape
    bpp = max(1, ch * depth // 8)
    raw = bytearray()
    prev = bytearray((w * ch * depth + 7) // 8)
    for y in range(h):
        if depth < 8:
            row = pack_row(samples[y, :, 0], depth)
        else:
            row = bytearray(samples[y].astype(">u2" if depth == 16 else np.uint8).tobytes())
        f = y % 5
        out = bytearray(len(row))
        for x in range(len(row)):
            a = row[x - bpp] if x >= bpp else 0
            b = prev[x]
            c = prev[x - bpp] if x >= bpp else 0
            pred = [0, a, b, (a + b) >> 1, paeth(a, b, c)][f]
            out[x] = (row[x] - pred) & 255
        raw += bytes([f]) + out
        prev = row

    def chunk(kind, data):
        return struct.pack(">I", len(data)) + kind + data + struct.pack(">I", zlib.crc32(kind + data) & 0xFFFFFFFF)

    png = b"\x89PNG\r\n\x1a\n" + chunk(b"IHDR", struct.pack(">IIBBBBB", w, h, depth, ctype, 0, 0, interlace))
    if plte is not None:
        png += chunk(b"PLTE", plte.astype(np.uint8).tobytes())
    return png + chunk(b"IDAT", zlib.compress(bytes(raw))) + chunk(b"IEND", b"")


def expected(rgb8, color):
    return GAMMA[rgb8] if color else rgb8.astype(np.float32) / np.float32(255)


@pytest.mark.parametrize("color", [True, False])
def test_colour_types_filters_and_depths(color):
    rng = np.random.default_rng(7)
    h, w = 11, 9
    rgb = rng.integers(0, 256, (h, w, 3), dtype=np.uint8)
    gray = rng.integers(0, 256, (h, w, 1), dtype=np.uint8)
    alpha = rng.integers(0, 256, (h, w, 1), dtype=np.uint8)
    cases = {
        "rgb": (encode(rgb, 2), rgb),
        "rgba": (encode(np.concatenate([rgb, alpha], 2), 6), rgb),
        "gray": (encode(gray, 0), np.repeat(gray, 3, 2)),
        "gray_alpha": (encode(np.concatenate([gray, alpha], 2), 4), np.repeat(gray, 3, 2)),
    }
    # 16-bit samples: stbi_load keeps the high byte
    rgb16 = rng.integers(0, 65536, (h, w, 3), dtype=np.uint16)
    cases["rgb16"] = (encode(rgb16, 2, depth=16), (rgb16 >> 8).astype(np.uint8))
    plte = rng.integers(0, 256, (16, 3), dtype=np.uint8)
    idx = rng.integers(0, 16, (h, w, 1), dtype=np.uint8)
    cases["palette"] = (encode(idx, 3, plte=plte), plte[idx[..., 0]])
    # packed samples: grey levels scaled to 8 bits (x255, x85, x17), palette indices unpacked
    for d in (1, 2, 4):
        g = rng.integers(0, 1 << d, (h, w, 1), dtype=np.uint8)
        cases["gray%d" % d] = (encode(g, 0, depth=d), np.repeat(g * (255 // ((1 << d) - 1)), 3, 2).astype(np.uint8))
        k = rng.integers(0, 1 << d, (h, w, 1), dtype=np.uint8)
        cases["palette%d" % d] = (encode(k, 3, depth=d, plte=plte[: 1 << d]), plte[k[..., 0]])
    for name, (png, want8) in cases.items():
        got = loader.decode_png(png, color=color)
        assert got.shape == (h, w, 3), name
        assert np.array_equal(got.view(np.uint32), expected(want8, color).view(np.uint32)), name


def test_rejected_files():
    rgb = np.zeros((2, 2, 3), np.uint8)
    for bad in (b"not a png", encode(rgb, 2, interlace=1), encode(rgb, 2)[:40]):
        with pytest.raises(ValueError):
            loader.decode_png(bad)
    # bit depth 4 is not valid for an RGB image; 8-bit data under a 4-bit grey header has
    # the wrong size
    for ctype, samples in ((2, rgb), (0, np.zeros((2, 2, 1), np.uint8))):
        png = bytearray(encode(samples, ctype))
        png[24] = 4
        with pytest.raises(ValueError):
            loader.decode_png(bytes(png))


def test_png_environment_in_a_scene_file(tmp_path):
    import json
    from test_cpu_loader import cornell_files
    rng = np.random.default_rng(2)
    sky = rng.integers(0, 256, (8, 16, 3), dtype=np.uint8)
    (tmp_path / "sky.png").write_bytes(encode(sky, 2))
    _, path = cornell_files(tmp_path, env_texture="sky.png")
    sc = loader.load_scene(path)
    assert np.array_equal(sc["env_rgb"].view(np.uint32), GAMMA[sky].view(np.uint32))
    doc = json.loads(open(path).read())
    doc["env"]["texture"] = "sky.tga"
    (tmp_path / "tga.json").write_text(json.dumps(doc))
    with pytest.raises(ValueError):
        loader.load_scene(str(tmp_path / "tga.json"))
