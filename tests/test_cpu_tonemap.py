"""The 8-bit output path (Texture2D::Save's PNG bytes, Core/Texture2D.cpp:18-91,199-222)
on the CPU: the oracle's restatement against an independent double-precision evaluation
of the same formulas, and the kernel's run table (petrichor_amd/csrc/pt_tonemap_lut.h)
against the oracle."""
import os
import re

import numpy as np

import oracle_lib as O

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LUT = os.path.join(ROOT, "petrichor_amd", "csrc", "pt_tonemap_lut.h")


def load_table():
    text = open(LUT).read()
    runs = int(re.search(r"#define PT_TONEMAP_RUNS (\d+)", text).group(1))
    start_body = text[text.index("pt_tonemap_start"):text.index("pt_tonemap_byte")]
    byte_body = text[text.index("pt_tonemap_byte"):text.index("pt_tonemap_bucket")]
    start = np.array([int(h, 16) for h in re.findall(r"0x([0-9a-f]{8})u", start_body)], np.uint64)
    byte = np.array([int(b) for b in re.findall(r"\b(\d+)\b", byte_body.split("=", 1)[1])], np.uint8)
    assert start.size == 3 * runs and byte.size == 3 * runs
    return start.reshape(3, runs), byte.reshape(3, runs)


def load_index():
    """The bucket index of the run table: (shift, [3, buckets + 1] run bounds)."""
    text = open(LUT).read()
    shift = int(re.search(r"#define PT_TONEMAP_BUCKET_SHIFT (\d+)", text).group(1))
    nb = int(re.search(r"#define PT_TONEMAP_BUCKETS (\d+)", text).group(1))
    body = text[text.index("pt_tonemap_bucket"):].split("=", 1)[1]
    bucket = np.array([int(b) for b in re.findall(r"\b(\d+)\b", body)], np.int64)
    assert bucket.size == 3 * (nb + 1) and nb << shift == 2 ** 32
    return shift, bucket.reshape(3, nb + 1)


def table_bytes(start, byte, values, index=None):
    """What the kernel computes: the byte of the last run whose start is <= the value's
    bits; with index, found as the kernel finds it (five steps between the bounds of the
    value's bucket)."""
    u = np.ascontiguousarray(values, np.float32).view(np.uint32).astype(np.uint64)
    out = np.zeros(u.shape + (3,), np.uint8)
    for ch in range(3):
        if index is None:
            i = np.searchsorted(start[ch], u, side="right") - 1
        else:
            shift, bucket = index
            b = (u >> np.uint64(shift)).astype(np.int64)
            i, hi = bucket[ch][b], bucket[ch][b + 1]
            for step in (16, 8, 4, 2, 1):
                j = np.minimum(i + step, hi)
                i = np.where(start[ch][j] <= u, j, i)
        out[..., ch] = byte[ch][i]
    return out


def oracle_bytes(values):
    v = np.ascontiguousarray(values, np.float32)
    return O.tonemap_rgb8(np.repeat(v[:, None], 3, axis=1))


def double_bytes(x):
    """The same formulas in float64 with numpy's pow (no float roundings): the byte and the
    distance of the scaled value from the nearest integer (where the two may differ)."""
    x = x.astype(np.float64)
    aces = (x * (2.51 * x + 0.03)) / (x * (2.43 * x + 0.59) + 0.14)
    x0, x1 = 0.05 ** 2.2, 0.9 ** 2.2
    with np.errstate(invalid="ignore"):
        c = (np.power(aces, 1.2) - x0) / (x1 - x0)
        chans = [c, np.power(c, 1.3), np.power(c, 1.6)]
        out, margin = [], []
        for v in chans:
            v = np.power(v, 1.1)
            v = np.where(v < 0, 0.0, np.where(v > 1, 1.0, v))
            q = (255 - 1e-3) * np.power(v, 1 / 2.2)
            q = np.where(np.isnan(q), 0.0, q)
            out.append(np.floor(q))
            margin.append(np.abs(q - np.round(q)))
    return np.stack(out, -1).astype(np.uint8), np.stack(margin, -1)


def test_oracle_matches_the_formulas_in_double():
    # 400K values over the range where the bytes change (3e-3 .. 3) and exact zeros:
    # any wrong constant, exponent or stage order moves many bytes
    rng = np.random.default_rng(5)
    x = np.concatenate([10.0 ** rng.uniform(-2.5, 0.5, 400000), [0.0, 1.0, 0.5, 2.0]]).astype(np.float32)
    got = oracle_bytes(x)
    want, margin = double_bytes(x)
    far = margin > 1e-3  # float roundings cannot move a value this far from a byte boundary
    assert np.array_equal(got[far], want[far])
    interior = (want > 0) & (want < 254)  # the toe and the top sit exactly on 0 and 254.999
    assert interior.mean() > 0.6 and far[interior].mean() > 0.99
    # black is black: ACES(0) = 0 lies below the curve's toe, whose negative values give NaN -> 0
    assert np.array_equal(oracle_bytes(np.array([0.0, -0.0], np.float32)), np.zeros((2, 3), np.uint8))
    # NaN and both infinities (ACES(inf) = inf / inf) give 0 through the NaN conversion
    special = np.array([np.nan, np.inf, -np.inf], np.float32)
    assert np.array_equal(oracle_bytes(special), np.zeros((3, 3), np.uint8))
    # the curve's top: the byte never reaches 255 ((255 - 1e-3) * 1 truncates to 254)
    big = oracle_bytes(np.array([1e6, 3.0e38], np.float32))
    assert big.max() <= 254 and big[0].min() >= 250


def test_table_equals_oracle():
    start, byte = load_table()
    assert np.all(start[:, 0] == 0)
    assert np.all(np.diff(start.astype(np.int64), axis=1) >= 0)
    # every run start and its neighbours, where a misplaced boundary would show
    edges = np.unique(np.concatenate([start.ravel().astype(np.int64) + d for d in (-2, -1, 0, 1, 2)]))
    edges = edges[(edges >= 0) & (edges < 2 ** 32)].astype(np.uint32)
    rng = np.random.default_rng(9)
    rand_bits = rng.integers(0, 2 ** 32, 1 << 21, dtype=np.uint64).astype(np.uint32)
    mids = (10.0 ** rng.uniform(-3, 4, 1 << 20)).astype(np.float32).view(np.uint32)
    index = load_index()
    for bits in (edges, rand_bits, mids):
        vals = bits.view(np.float32)
        want = oracle_bytes(vals)
        assert np.array_equal(table_bytes(start, byte, vals), want)
        assert np.array_equal(table_bytes(start, byte, vals, index), want)


def test_bucket_index_bounds_every_search():
    # bucket b's first pattern lies in run bucket[b], its last in run bucket[b + 1] or the
    # one before it; the last bound is the last run; at most 5 search steps
    start, _ = load_table()
    shift, bucket = load_index()
    nb = bucket.shape[1] - 1
    for ch in range(3):
        real = int((start[ch] != 0xFFFFFFFF).sum())
        firsts = np.arange(nb, dtype=np.uint64) << np.uint64(shift)
        assert np.array_equal(bucket[ch][:nb], np.searchsorted(start[ch][:real], firsts, side="right") - 1)
        assert bucket[ch][nb] == real - 1
        assert np.all(np.diff(bucket[ch]) >= 0) and np.diff(bucket[ch]).max() < 32


def test_table_runs_increase_over_the_visible_range():
    # Between 0 and the overflow of x * (2.51 x + 0.03) each channel's bytes grow, except
    # where float rounding makes the reference's formulas step back by one byte for at most
    # 4 consecutive inputs at a byte boundary (44, 58 and 75 such blips in the three
    # channels): a search over 254 thresholds would get those inputs wrong, the runs do not.
    start, byte = load_table()
    top = np.float32(1e18).view(np.uint32)
    for ch in range(3):
        sel = start[ch] <= top
        b = byte[ch][sel].astype(np.int32)
        s = start[ch][sel].astype(np.int64)
        down = np.where(np.diff(b) < 0)[0]
        assert 0 < down.size < 100
        assert np.all(b[down + 1] == b[down] - 1) and np.all(b[down + 2] == b[down])
        assert np.all(s[down + 2] - s[down + 1] <= 4)
        up = np.delete(b, down + 1)
        assert np.all(np.diff(up) >= 0) and b[-1] >= 250
