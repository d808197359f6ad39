// pt_png.cpp — PNG images as the reference's Texture2D::Load reads them (host code).
//
// Texture2D::Load (Core/Texture2D.cpp:107-160) reads a .png with stbi_load(..., 3): 8-bit
// RGB whatever the file holds (grey replicated, alpha dropped, palettes expanded, 16-bit
// samples reduced to their high byte). Each byte becomes a float divided by 255, and a
// colour texture then goes through ApplyGamma, powf(x, 2.2) per channel
// (Core/Texture2D.cpp:44-56). stb is not part of this library: the PNG format is decoded
// here from its specification (zlib inflate, the five scanline filters), for
// non-interlaced images of every colour type and bit depth: samples of 1, 2 or 4 bits are
// unpacked (grey levels scaled to 8 bits by 255 / (2^depth - 1), palette indices as they
// are), 16-bit samples reduced to their high byte.
#include <zlib.h>

#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <new>
#include <string>
#include <vector>

#include "../../include/petrichor_hip.h"

namespace {

uint32_t be32(const uint8_t* p) { return (uint32_t)p[0] << 24 | (uint32_t)p[1] << 16 | (uint32_t)p[2] << 8 | p[3]; }

int paeth(int a, int b, int c)
{
    const int p = a + b - c, pa = std::abs(p - a), pb = std::abs(p - b), pc = std::abs(p - c);
    return pa <= pb && pa <= pc ? a : (pb <= pc ? b : c);
}

int decode(const uint8_t* data, int64_t len, int32_t color, int32_t* width, int32_t* height, float* rgb_out,
           char* err, int32_t err_len)
{
    auto fail = [&](const std::string& msg) {
        if (err && err_len > 0) std::snprintf(err, (size_t)err_len, "PNG: %s", msg.c_str());
        return PT_ERR_ARG;
    };
    static const uint8_t kSig[8] = {0x89, 'P', 'N', 'G', '\r', '\n', 0x1a, '\n'};
    if (!data || len < 8 || !width || !height || std::memcmp(data, kSig, 8) != 0) return fail("not a PNG file");
    int64_t pos = 8;
    uint32_t w = 0, h = 0;
    int depth = 0, ctype = -1, interlace = 0;
    std::vector<uint8_t> idat, plte;
    while (pos + 12 <= len) {
        const uint32_t n = be32(data + pos);
        const uint8_t* kind = data + pos + 4;
        const uint8_t* body = data + pos + 8;
        if ((int64_t)n > len - pos - 12) return fail("truncated chunk");
        if (std::memcmp(kind, "IHDR", 4) == 0 && n >= 13) {
            w = be32(body);
            h = be32(body + 4);
            depth = body[8];
            ctype = body[9];
            interlace = body[12];
        } else if (std::memcmp(kind, "PLTE", 4) == 0) {
            plte.assign(body, body + n);
        } else if (std::memcmp(kind, "IDAT", 4) == 0) {
            idat.insert(idat.end(), body, body + n);
        } else if (std::memcmp(kind, "IEND", 4) == 0) {
            break;
        }
        pos += 12 + (int64_t)n;
    }
    static const int kChannels[7] = {1, 0, 3, 1, 2, 0, 4}; // by colour type
    if (w == 0 || h == 0 || w > 65536 || h > 65536 || ctype < 0 || ctype > 6 || kChannels[ctype] == 0)
        return fail("bad header");
    if (interlace != 0) return fail("interlaced images are not supported");
    // the depths the format allows per colour type
    const bool depth_ok = (ctype == 0 && (depth == 1 || depth == 2 || depth == 4 || depth == 8 || depth == 16)) ||
                          (ctype == 3 && (depth == 1 || depth == 2 || depth == 4 || depth == 8)) ||
                          ((ctype == 2 || ctype == 4 || ctype == 6) && (depth == 8 || depth == 16));
    if (!depth_ok) return fail("bit depth " + std::to_string(depth) + " is not valid for colour type " +
                               std::to_string(ctype));
    if (ctype == 3 && plte.size() < 3) return fail("palette missing");
    *width = (int32_t)w;
    *height = (int32_t)h;
    if (!rgb_out) return PT_OK; // size query
    const int ch = kChannels[ctype];
    // bytes per row and the filters' byte distance to the previous pixel (at least 1)
    const size_t stride = ((size_t)w * ch * depth + 7) / 8;
    const int bpp = depth >= 8 ? ch * (depth / 8) : 1, bps = depth >= 8 ? depth / 8 : 0;
    if ((stride + 1) * (size_t)h > ((size_t)1 << 32)) return fail("image too large");
    std::vector<uint8_t> raw((stride + 1) * h);
    uLongf raw_len = (uLongf)raw.size();
    if (uncompress(raw.data(), &raw_len, idat.data(), (uLong)idat.size()) != Z_OK || raw_len != raw.size())
        return fail("bad image data");
    // undo the scanline filters in place (filter byte, then stride bytes per row)
    std::vector<uint8_t> prev(stride, 0);
    for (uint32_t y = 0; y < h; ++y) {
        uint8_t* row = &raw[(stride + 1) * y];
        const int f = row[0];
        uint8_t* cur = row + 1;
        for (size_t x = 0; x < stride; ++x) {
            const int a = x >= (size_t)bpp ? cur[x - bpp] : 0, b = prev[x], c = x >= (size_t)bpp ? prev[x - bpp] : 0;
            int pred = 0;
            switch (f) {
            case 0: pred = 0; break;
            case 1: pred = a; break;
            case 2: pred = b; break;
            case 3: pred = (a + b) >> 1; break;
            case 4: pred = paeth(a, b, c); break;
            default: return fail("bad filter");
            }
            cur[x] = (uint8_t)(cur[x] + pred);
        }
        std::memcpy(prev.data(), cur, stride);
    }
    for (uint32_t y = 0; y < h; ++y) {
        const uint8_t* cur = &raw[(stride + 1) * y + 1];
        for (uint32_t x = 0; x < w; ++x) {
            // 8-bit samples of this pixel: the high byte of 16-bit ones; a packed sample of
            // 1, 2 or 4 bits (most significant bits first) as it is
            auto sample = [&](int c) -> int {
                if (depth >= 8) return cur[(size_t)x * bpp + (size_t)c * bps];
                const size_t bit = (size_t)x * depth;
                return (cur[bit / 8] >> (8 - depth - (int)(bit % 8))) & ((1 << depth) - 1);
            };
            // grey levels of fewer than 8 bits scaled to 8 bits (stb's depth scale: x255,
            // x85, x17)
            const int grey_scale = depth == 1 ? 255 : depth == 2 ? 85 : depth == 4 ? 17 : 1;
            uint8_t rgb[3];
            if (ctype == 3) {
                const size_t i = 3 * (size_t)sample(0);
                if (i + 2 >= plte.size()) return fail("palette index out of range");
                rgb[0] = plte[i];
                rgb[1] = plte[i + 1];
                rgb[2] = plte[i + 2];
            } else if (ctype == 0 || ctype == 4) {
                rgb[0] = rgb[1] = rgb[2] = (uint8_t)(sample(0) * grey_scale);
            } else {
                rgb[0] = sample(0);
                rgb[1] = sample(1);
                rgb[2] = sample(2);
            }
            float* o = rgb_out + 3 * ((size_t)y * w + x);
            for (int c = 0; c < 3; ++c) {
                // Color3f(byte) /= 255, then ApplyGamma for colour textures
                const float v = (float)rgb[c] / 255.0f;
                o[c] = color ? powf(v, 2.2f) : v;
            }
        }
    }
    return PT_OK;
}

} // namespace

extern "C" int pt_png_decode(const uint8_t* data, int64_t len, int32_t color, int32_t* width, int32_t* height,
                             float* rgb_out, char* err, int32_t err_len)
{
    try {
        return decode(data, len, color, width, height, rgb_out, err, err_len);
    } catch (const std::bad_alloc&) {
        if (err && err_len > 0) std::snprintf(err, (size_t)err_len, "PNG: not enough host memory for the image");
        return PT_ERR_OOM;
    }
}
