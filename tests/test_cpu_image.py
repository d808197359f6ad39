"""The image files Texture2D::Load reads at a .png path besides PNG (pt_image_decode):
Texture2D::Load (Core/Texture2D.cpp:107-195) takes .png and .hdr paths only, and reads a
.png path with stbi_load(..., 3), which tells the format by the file's header. BMP, TGA and
binary PNM files are written here by an independent encoder (Pillow) and, for the variants
Pillow does not write (16-bit BMP and TGA, run-length TGA with mixed packets), by hand
from the format specifications; the decoded floats must be the file's RGB bytes / 255,
with powf(x, 2.2) for a colour texture."""
import ctypes
import io
import os
import struct
import tempfile

import numpy as np
import pytest

from petrichor_amd import loader

PIL = pytest.importorskip("PIL.Image")

LIBM = ctypes.CDLL("libm.so.6")
LIBM.powf.restype = ctypes.c_float
LIBM.powf.argtypes = [ctypes.c_float, ctypes.c_float]
GAMMA = np.array([LIBM.powf(float(np.float32(b) / np.float32(255)), 2.2) for b in range(256)], np.float32)
LINEAR = (np.arange(256, dtype=np.float32) / np.float32(255)).astype(np.float32)


def pil_bytes(img, fmt, **kw):
    b = io.BytesIO()
    img.save(b, format=fmt, **kw)
    return b.getvalue()


def expect(rgb8, color):
    return (GAMMA if color else LINEAR)[rgb8]


def rgb_image(w, h, seed):
    return np.random.default_rng(seed).integers(0, 256, (h, w, 3), dtype=np.uint8)


@pytest.mark.parametrize("mode", ["RGB", "RGBA", "P", "L", "1"])
def test_bmp_written_by_pillow(mode):
    src = rgb_image(37, 23, 1)  # odd width: padded rows
    img = PIL.fromarray(src, "RGB")
    if mode == "RGBA":
        img = PIL.fromarray(np.dstack([src, src[:, :, 0]]), "RGBA")
    elif mode != "RGB":
        img = img.convert(mode)
    data = pil_bytes(img, "BMP")
    ref = np.asarray(PIL.open(io.BytesIO(data)).convert("RGB"))
    for color in (True, False):
        got = loader.decode_image(data, color=color)
        assert got.shape == ref.shape
        assert np.array_equal(got.view(np.uint32), expect(ref, color).view(np.uint32))


def bmp16(px, w, h, masks=None):
    """A 16-bit BMP (BI_RGB 5-5-5, or BI_BITFIELDS with the given masks), rows bottom-up."""
    stride = (w * 2 + 3) // 4 * 4
    rows = b""
    for y in range(h - 1, -1, -1):
        r = struct.pack("<%dH" % w, *px[y])
        rows += r + b"\0" * (stride - len(r))
    hdr = struct.pack("<IiiHHIIiiII", 40, w, h, 1, 16, 3 if masks else 0, len(rows), 2835, 2835, 0, 0)
    extra = struct.pack("<III", *masks) if masks else b""
    off = 14 + len(hdr) + len(extra)
    return b"BM" + struct.pack("<IHHI", off + len(rows), 0, 0, off) + hdr + extra + rows


def widen(v, bits):
    """A channel of `bits` bits widened to 8 bits by repeating its bit pattern."""
    out, have = 0, 0
    while have < 8:
        out = (out << bits) | v
        have += bits
    return (out >> (have - 8)) & 255


def test_bmp_16_bit_channels_are_widened_by_bit_replication():
    rng = np.random.default_rng(2)
    w, h = 9, 5
    px = rng.integers(0, 1 << 16, (h, w)).astype(np.uint32)
    for masks, layout in ((None, (10, 5, 5, 5, 0, 5)), ((0xF800, 0x07E0, 0x001F), (11, 5, 5, 6, 0, 5))):
        v = px & (0x7FFF if masks is None else 0xFFFF)
        data = bmp16(v.tolist(), w, h, masks)
        sr, br, sg, bg, sb, bb = layout
        ref = np.zeros((h, w, 3), np.uint8)
        for y in range(h):
            for x in range(w):
                q = int(v[y, x])
                ref[y, x] = [widen((q >> sr) & ((1 << br) - 1), br), widen((q >> sg) & ((1 << bg) - 1), bg),
                             widen((q >> sb) & ((1 << bb) - 1), bb)]
        got = loader.decode_image(data, color=False)
        assert np.array_equal(got.view(np.uint32), expect(ref, False).view(np.uint32))
    # spot values of the widening: 5 bits 16 -> 132, 6 bits 1 -> 4, 5 bits 31 -> 255
    assert (widen(16, 5), widen(1, 6), widen(31, 5), widen(1, 5)) == (132, 4, 255, 8)


def test_bmp_top_down_and_run_length_rejected():
    src = rgb_image(6, 4, 3)
    data = bytearray(pil_bytes(PIL.fromarray(src, "RGB"), "BMP"))
    h = struct.unpack_from("<i", data, 22)[0]
    off = struct.unpack_from("<I", data, 10)[0]
    stride = (6 * 3 + 3) // 4 * 4
    rows = [bytes(data[off + stride * i: off + stride * (i + 1)]) for i in range(h)]
    struct.pack_into("<i", data, 22, -h)
    data[off:off + stride * h] = b"".join(rows[::-1])  # top row first
    got = loader.decode_image(bytes(data), color=False)
    assert np.array_equal(got.view(np.uint32), expect(src, False).view(np.uint32))
    rle = bytearray(data)
    struct.pack_into("<I", rle, 30, 1)  # BI_RLE8
    with pytest.raises(ValueError, match="run-length"):
        loader.decode_image(bytes(rle))


@pytest.mark.parametrize("mode", ["RGB", "RGBA", "L", "P"])
@pytest.mark.parametrize("rle", [False, True])
@pytest.mark.parametrize("orientation", [-1, 1])
def test_tga_written_by_pillow(mode, rle, orientation):
    src = rgb_image(19, 11, 4)
    src[3:6, :] = src[3, 0]  # runs for the run-length packets
    img = PIL.fromarray(src, "RGB")
    if mode == "RGBA":
        img = PIL.fromarray(np.dstack([src, src[:, :, 1]]), "RGBA")
    elif mode != "RGB":
        img = img.convert(mode)
    kw = {"orientation": orientation}
    if rle:
        kw["compression"] = "tga_rle"
    data = pil_bytes(img, "TGA", **kw)
    ref = np.asarray(PIL.open(io.BytesIO(data)).convert("RGB"))
    for color in (True, False):
        got = loader.decode_image(data, color=color)
        assert np.array_equal(got.view(np.uint32), expect(ref, color).view(np.uint32))


def test_tga_16_bit_and_mixed_run_length_packets():
    """Type 10 (run-length true colour), 16-bit 5-5-5 pixels, bottom-up: each channel
    v * 255 // 31; a run packet, a raw packet and a packet that spans two rows."""
    w, h = 5, 3
    rng = np.random.default_rng(5)
    vals = rng.integers(0, 1 << 15, 7).tolist()
    # pixels in file order (bottom row first): run of 4 (vals[0]), raw 3 (vals[1..3]), run of 8 (vals[4])
    order = [vals[0]] * 4 + vals[1:4] + [vals[4]] * 8
    body = bytes([0x80 | 3]) + struct.pack("<H", vals[0]) + bytes([2]) + struct.pack("<3H", *vals[1:4]) + \
        bytes([0x80 | 7]) + struct.pack("<H", vals[4])
    hdr = struct.pack("<BBBHHBHHHHBB", 0, 0, 10, 0, 0, 0, 0, 0, w, h, 16, 0)
    got = loader.decode_image(hdr + body, color=False)
    ref = np.zeros((h, w, 3), np.uint8)
    for i, v in enumerate(order):
        y = h - 1 - i // w
        ref[y, i % w] = [((v >> 10) & 31) * 255 // 31, ((v >> 5) & 31) * 255 // 31, (v & 31) * 255 // 31]
    assert np.array_equal(got.view(np.uint32), expect(ref, False).view(np.uint32))


@pytest.mark.parametrize("mode", ["RGB", "L"])
def test_pnm_written_by_pillow(mode):
    src = rgb_image(13, 7, 6)
    img = PIL.fromarray(src, "RGB").convert(mode)
    data = pil_bytes(img, "PPM")
    assert data[:2] == (b"P6" if mode == "RGB" else b"P5")
    ref = np.asarray(PIL.open(io.BytesIO(data)).convert("RGB"))
    got = loader.decode_image(data, color=True)
    assert np.array_equal(got.view(np.uint32), expect(ref, True).view(np.uint32))
    # a comment in the header, and a maximum value below 255 (samples taken as they are)
    low = b"P5\n# c\n3 1\n# d\n100\n" + bytes([0, 50, 100])
    got = loader.decode_image(low, color=False)
    assert np.array_equal(got[0, :, 0], LINEAR[[0, 50, 100]])
    with pytest.raises(ValueError, match="255"):
        loader.decode_image(b"P5 1 1 65535\n\0\0")


def test_texture_paths_follow_texture2d_load():
    """A .png path holding a BMP reads as the BMP (stbi_load tells formats by header); a
    PNG matches pt_png_decode; a JPEG at a .png path and any other extension are refused."""
    src = rgb_image(8, 6, 7)
    with tempfile.TemporaryDirectory() as d:
        p = os.path.join(d, "t.png")
        with open(p, "wb") as f:
            f.write(pil_bytes(PIL.fromarray(src, "RGB"), "BMP"))
        assert np.array_equal(loader.load_texture(p, color=True).view(np.uint32), expect(src, True).view(np.uint32))
        with open(p, "wb") as f:
            f.write(pil_bytes(PIL.fromarray(src, "RGB"), "PNG"))
        png = open(p, "rb").read()
        assert np.array_equal(loader.load_texture(p, color=False), loader.decode_png(png, color=False))
        with open(p, "wb") as f:
            f.write(pil_bytes(PIL.fromarray(src, "RGB"), "JPEG"))
        with pytest.raises(ValueError, match="JPEG"):
            loader.load_texture(p)
        q = os.path.join(d, "t.bmp")
        with open(q, "wb") as f:
            f.write(pil_bytes(PIL.fromarray(src, "RGB"), "BMP"))
        with pytest.raises(ValueError, match=".png and .hdr"):
            loader.load_texture(q)
