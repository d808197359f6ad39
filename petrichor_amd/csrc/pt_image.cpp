// pt_image.cpp — the LDR image files Texture2D::Load reads besides PNG (host code).
//
// Texture2D::Load (Core/Texture2D.cpp:107-160) reads a non-HDR file with stbi_load(..., 3),
// which tries each format stb knows by its header, and gets 8-bit RGB whatever the file
// holds: grey replicated, alpha dropped, palettes expanded. Each byte then becomes
// byte / 255, and a colour texture goes through ApplyGamma, powf(x, 2.2) per channel
// (Core/Texture2D.cpp:44-56). pt_image_decode does the same for the lossless formats,
// decoded here from their specifications:
//   PNG   pt_png_decode (pt_png.cpp);
//   BMP   Windows bitmaps: core, info (40), V2/V3 (52/56), V4 (108) and V5 (124) headers;
//         1/4/8-bit palettes, 16-bit (5-5-5 by default) and 32-bit channels at their bit
//         masks, 24-bit BGR; rows bottom-up unless the height is negative, each padded to
//         4 bytes. Channels of fewer than 8 bits are widened by bit replication (5 bits:
//         v * 33 >> 2), as stb widens masked channels. Run-length (BI_RLE4/8) bitmaps are
//         rejected, as stb rejects them.
//   TGA   Truevision images of types 1, 2, 3 (colour-mapped, true-colour, grey) and their
//         run-length forms 9, 10, 11; 8-bit grey or indices, 15/16-bit 5-5-5 (each channel
//         v * 255 / 31, integer division, as stb's TGA reader scales them), 24/32-bit BGR(A);
//         rows bottom-up unless descriptor bit 5 is set. Colour maps are read as stb reads
//         them (the first-entry field skips that many bytes; indices address the map
//         directly, an index past its end reads entry 0); for the usual first entry 0 that
//         is the format's own reading.
//   PNM   binary P5 (grey) and P6 (RGB) with a maximum value of at most 255; samples are
//         taken as they are (stb does not rescale a maximum below 255).
// JPEG, GIF, PSD and PIC files are reported as unsupported: a JPEG's pixels depend on the
// decoder's IDCT and upsampling arithmetic, which is stb's own and not restated here.
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <new>
#include <string>
#include <vector>

#include "../../include/petrichor_hip.h"

namespace {

struct Img {
    int w = 0, h = 0;
    std::vector<uint8_t> rgb; // w * h * 3, top row first
};

struct Reader {
    const uint8_t* p;
    int64_t len;
    int64_t pos = 0;
    bool ok = true;
    int u8()
    {
        if (pos >= len) {
            ok = false;
            return 0;
        }
        return p[pos++];
    }
    int le16()
    {
        const int a = u8();
        return a | (u8() << 8);
    }
    uint32_t le32()
    {
        const uint32_t a = (uint32_t)le16();
        return a | ((uint32_t)le16() << 16);
    }
    void skip(int64_t n)
    {
        if (n < 0 || n > len - pos) ok = false;
        else pos += n;
    }
};

// A channel of `bits` bits widened to 8 by repeating its bit pattern.
int widen(uint32_t v, int bits)
{
    if (bits <= 0) return 0;
    if (bits >= 8) return (int)(v >> (bits - 8)) & 255;
    int out = 0, have = 0;
    while (have < 8) {
        out = (out << bits) | (int)v;
        have += bits;
    }
    return (out >> (have - 8)) & 255;
}

// Bit position and width of a contiguous mask.
void mask_shift(uint32_t m, int* shift, int* bits)
{
    *shift = 0;
    *bits = 0;
    if (!m) return;
    while (!(m & 1u)) {
        m >>= 1;
        ++*shift;
    }
    while (m & 1u) {
        m >>= 1;
        ++*bits;
    }
}

bool check_size(int64_t w, int64_t h, std::string* why)
{
    if (w <= 0 || h <= 0 || w > 65536 || h > 65536 || w * h > ((int64_t)1 << 28)) {
        *why = "bad image size";
        return false;
    }
    return true;
}

bool decode_bmp(const uint8_t* d, int64_t len, bool size_only, Img* im, std::string* why)
{
    Reader r{d, len};
    r.skip(10);
    const uint32_t data_off = r.le32();
    const uint32_t hsz = r.le32();
    int64_t w, h;
    int bpp, comp = 0, ncolors = 0;
    uint32_t mr = 0, mg = 0, mb = 0;
    if (hsz == 12) {
        w = r.le16();
        h = r.le16(); // unsigned in the core header: always bottom-up
        r.le16();
        bpp = r.le16();
    } else if (hsz == 40 || hsz == 52 || hsz == 56 || hsz == 108 || hsz == 124) {
        w = (int32_t)r.le32();
        h = (int32_t)r.le32();
        r.le16();
        bpp = r.le16();
        comp = (int)r.le32();
        r.skip(12);
        ncolors = (int)r.le32();
        r.skip(4);
        if (comp == 1 || comp == 2) {
            *why = "BMP: run-length encoded bitmaps are not supported";
            return false;
        }
        if (comp != 0 && comp != 3) {
            *why = "BMP: unsupported compression " + std::to_string(comp);
            return false;
        }
        if (comp == 3 || hsz >= 52) {
            // BI_BITFIELDS masks: after a 40-byte header, or inside a larger one
            mr = r.le32();
            mg = r.le32();
            mb = r.le32();
        }
        if (comp == 0) mr = mg = mb = 0; // plain bitmaps use the default layout
    } else {
        *why = "BMP: unsupported header size " + std::to_string(hsz);
        return false;
    }
    if (!r.ok) {
        *why = "BMP: truncated header";
        return false;
    }
    const bool top_down = h < 0;
    if (h < 0) h = -h;
    if (!check_size(w, h, why)) return false;
    if (bpp != 1 && bpp != 4 && bpp != 8 && bpp != 16 && bpp != 24 && bpp != 32) {
        *why = "BMP: unsupported bit depth " + std::to_string(bpp);
        return false;
    }
    im->w = (int)w;
    im->h = (int)h;
    if (size_only) return true;
    // palette: after the header (and the masks of a 40-byte BI_BITFIELDS header)
    std::vector<uint8_t> pal;
    if (bpp <= 8) {
        // as many entries as fit between the header and the pixels (stb's palette size),
        // at most 256
        const int64_t pal_off = 14 + (int64_t)hsz + (hsz == 40 && comp == 3 ? 12 : 0);
        const int entry = hsz == 12 ? 3 : 4;
        const int64_t fit = ((int64_t)data_off - pal_off) / entry;
        const int n = (int)(fit < 256 ? fit : 256);
        (void)ncolors;
        if (n <= 0 || pal_off + (int64_t)n * entry > len) {
            *why = "BMP: palette missing";
            return false;
        }
        pal.resize(3 * (size_t)n);
        for (int i = 0; i < n; ++i) { // B, G, R(, reserved)
            pal[3 * i] = d[pal_off + (int64_t)i * entry + 2];
            pal[3 * i + 1] = d[pal_off + (int64_t)i * entry + 1];
            pal[3 * i + 2] = d[pal_off + (int64_t)i * entry];
        }
    }
    if (bpp == 16 && !mr) {
        mr = 0x7c00u;
        mg = 0x03e0u;
        mb = 0x001fu;
    }
    if (bpp == 32 && !mr) {
        mr = 0x00ff0000u;
        mg = 0x0000ff00u;
        mb = 0x000000ffu;
    }
    int sr, br, sg, bg, sb, bb;
    mask_shift(mr, &sr, &br);
    mask_shift(mg, &sg, &bg);
    mask_shift(mb, &sb, &bb);
    const int64_t stride = ((w * bpp + 31) / 32) * 4;
    if ((int64_t)data_off + stride * h > len) {
        *why = "BMP: truncated pixel data";
        return false;
    }
    im->rgb.resize((size_t)(w * h * 3));
    for (int64_t row = 0; row < h; ++row) {
        const uint8_t* s = d + data_off + stride * row;
        const int64_t y = top_down ? row : h - 1 - row;
        uint8_t* o = &im->rgb[(size_t)(y * w * 3)];
        for (int64_t x = 0; x < w; ++x) {
            int c[3];
            if (bpp <= 8) {
                const int64_t bit = x * bpp;
                const int idx = (s[bit / 8] >> (8 - bpp - (int)(bit % 8))) & ((1 << bpp) - 1);
                if (3 * (size_t)idx + 2 >= pal.size()) {
                    *why = "BMP: palette index out of range";
                    return false;
                }
                c[0] = pal[3 * idx];
                c[1] = pal[3 * idx + 1];
                c[2] = pal[3 * idx + 2];
            } else if (bpp == 24) {
                c[0] = s[3 * x + 2];
                c[1] = s[3 * x + 1];
                c[2] = s[3 * x];
            } else {
                const uint32_t v = bpp == 16 ? (uint32_t)(s[2 * x] | s[2 * x + 1] << 8)
                                             : (uint32_t)s[4 * x] | (uint32_t)s[4 * x + 1] << 8 |
                                                   (uint32_t)s[4 * x + 2] << 16 | (uint32_t)s[4 * x + 3] << 24;
                c[0] = widen((v & mr) >> sr, br);
                c[1] = widen((v & mg) >> sg, bg);
                c[2] = widen((v & mb) >> sb, bb);
            }
            o[3 * x] = (uint8_t)c[0];
            o[3 * x + 1] = (uint8_t)c[1];
            o[3 * x + 2] = (uint8_t)c[2];
        }
    }
    return true;
}

// A TGA header stb accepts (the format has no signature): a known image type, a colour
// map where the type needs one, and a pixel depth that fits the type.
bool tga_header_ok(const uint8_t* d, int64_t len)
{
    if (len < 18) return false;
    const int cmap = d[1], type = d[2], bits = d[16], w = d[12] | d[13] << 8, h = d[14] | d[15] << 8;
    const int base = type & ~8;
    if (type != 1 && type != 2 && type != 3 && type != 9 && type != 10 && type != 11) return false;
    if (cmap > 1 || (base == 1) != (cmap == 1)) return false;
    if (w < 1 || h < 1) return false;
    if (base == 1) return bits == 8 && (d[7] == 15 || d[7] == 16 || d[7] == 24 || d[7] == 32);
    if (base == 3) return bits == 8 || bits == 16;
    return bits == 8 || bits == 15 || bits == 16 || bits == 24 || bits == 32; // 8-bit true colour reads as grey
}

// One TGA pixel of `bytes` bytes to RGB.
void tga_rgb(const uint8_t* px, int bytes, bool grey, uint8_t* o)
{
    if (grey) {
        o[0] = o[1] = o[2] = px[0]; // 16-bit grey + alpha: the grey byte
    } else if (bytes == 2) {
        const int v = px[0] | px[1] << 8;
        o[0] = (uint8_t)(((v >> 10) & 31) * 255 / 31);
        o[1] = (uint8_t)(((v >> 5) & 31) * 255 / 31);
        o[2] = (uint8_t)((v & 31) * 255 / 31);
    } else {
        o[0] = px[2];
        o[1] = px[1];
        o[2] = px[0];
    }
}

bool decode_tga(const uint8_t* d, int64_t len, bool size_only, Img* im, std::string* why)
{
    if (!tga_header_ok(d, len)) {
        *why = "TGA: unsupported header";
        return false;
    }
    const int id_len = d[0], type = d[2], bits = d[16], desc = d[17];
    const int cmap_first = d[3] | d[4] << 8, cmap_len = d[5] | d[6] << 8, cmap_bits = d[7];
    const int64_t w = d[12] | d[13] << 8, h = d[14] | d[15] << 8;
    if (!check_size(w, h, why)) return false;
    im->w = (int)w;
    im->h = (int)h;
    if (size_only) return true;
    const int base = type & ~8;
    const bool rle = (type & 8) != 0, grey = base == 3 || (base == 2 && bits == 8);
    int64_t pos = 18 + id_len;
    std::vector<uint8_t> pal;
    if (base == 1) {
        // stb's reading of the colour map: the first-entry field counts bytes skipped before
        // the map, indices address the map directly, and an index past its end reads entry 0
        const int eb = (cmap_bits + 7) / 8;
        pos += cmap_first;
        if (cmap_len < 1 || pos + (int64_t)cmap_len * eb > len) {
            *why = "TGA: truncated colour map";
            return false;
        }
        pal.resize(3 * (size_t)cmap_len);
        for (int i = 0; i < cmap_len; ++i) tga_rgb(d + pos + (int64_t)i * eb, eb, false, &pal[3 * (size_t)i]);
        pos += (int64_t)cmap_len * eb;
    }
    const int pb = (bits + 7) / 8;
    const bool top_down = (desc & 0x20) != 0;
    im->rgb.resize((size_t)(w * h * 3));
    int64_t run = 0;
    bool raw_run = false;
    const uint8_t* px = nullptr;
    for (int64_t i = 0; i < w * h; ++i) {
        if (rle) {
            if (run == 0) { // packet header: 1 + count raw pixels, or one pixel repeated 1 + count times
                if (pos >= len) {
                    *why = "TGA: truncated run-length data";
                    return false;
                }
                const int hdr = d[pos++];
                run = (hdr & 127) + 1;
                raw_run = !(hdr & 128);
                px = nullptr;
            }
            if (raw_run || !px) {
                if (pos + pb > len) {
                    *why = "TGA: truncated pixel data";
                    return false;
                }
                px = d + pos;
                pos += pb;
            }
            --run;
        } else {
            if (pos + pb > len) {
                *why = "TGA: truncated pixel data";
                return false;
            }
            px = d + pos;
            pos += pb;
        }
        const int64_t x = i % w, row = i / w, y = top_down ? row : h - 1 - row;
        uint8_t* o = &im->rgb[(size_t)((y * w + x) * 3)];
        if (base == 1) {
            const int idx = px[0] < cmap_len ? px[0] : 0;
            std::memcpy(o, &pal[3 * (size_t)idx], 3);
        } else {
            tga_rgb(px, pb, grey, o);
        }
    }
    return true;
}

bool decode_pnm(const uint8_t* d, int64_t len, bool size_only, Img* im, std::string* why)
{
    const bool rgb = d[1] == '6';
    int64_t pos = 2;
    int64_t v[3] = {0, 0, 0}; // width, height, maximum value
    for (int k = 0; k < 3; ++k) {
        for (;;) { // white space and comments
            while (pos < len && (d[pos] == ' ' || d[pos] == '\t' || d[pos] == '\n' || d[pos] == '\r' ||
                                 d[pos] == '\v' || d[pos] == '\f'))
                ++pos;
            if (pos < len && d[pos] == '#') {
                while (pos < len && d[pos] != '\n' && d[pos] != '\r') ++pos;
                continue;
            }
            break;
        }
        if (pos >= len || d[pos] < '0' || d[pos] > '9') {
            *why = "PNM: bad header";
            return false;
        }
        while (pos < len && d[pos] >= '0' && d[pos] <= '9' && v[k] < ((int64_t)1 << 31)) v[k] = v[k] * 10 + (d[pos++] - '0');
    }
    ++pos; // the single white-space byte before the samples
    if (v[2] < 1 || v[2] > 255) {
        *why = "PNM: maximum values above 255 (16-bit samples) are not supported";
        return false;
    }
    if (!check_size(v[0], v[1], why)) return false;
    im->w = (int)v[0];
    im->h = (int)v[1];
    if (size_only) return true;
    const int64_t n = v[0] * v[1] * (rgb ? 3 : 1);
    if (pos + n > len) {
        *why = "PNM: truncated pixel data";
        return false;
    }
    im->rgb.resize((size_t)(v[0] * v[1] * 3));
    for (int64_t i = 0; i < v[0] * v[1]; ++i)
        for (int c = 0; c < 3; ++c) im->rgb[3 * (size_t)i + c] = d[pos + (rgb ? 3 * i + c : i)];
    return true;
}

int decode(const uint8_t* data, int64_t len, int32_t color, int32_t* width, int32_t* height, float* rgb_out,
           char* err, int32_t err_len)
{
    auto fail = [&](int code, const std::string& msg) {
        if (err && err_len > 0) std::snprintf(err, (size_t)err_len, "%s", msg.c_str());
        return code;
    };
    if (!data || len < 2 || !width || !height) return fail(PT_ERR_ARG, "image: no data");
    static const uint8_t kPng[8] = {0x89, 'P', 'N', 'G', '\r', '\n', 0x1a, '\n'};
    if (len >= 8 && std::memcmp(data, kPng, 8) == 0)
        return pt_png_decode(data, len, color, width, height, rgb_out, err, err_len);
    if (len >= 3 && data[0] == 0xff && data[1] == 0xd8 && data[2] == 0xff)
        return fail(PT_ERR_UNSUPPORTED, "JPEG images are not supported (PNG, BMP, TGA and binary PNM are)");
    if (len >= 6 && (std::memcmp(data, "GIF87a", 6) == 0 || std::memcmp(data, "GIF89a", 6) == 0))
        return fail(PT_ERR_UNSUPPORTED, "GIF images are not supported (PNG, BMP, TGA and binary PNM are)");
    if (len >= 4 && std::memcmp(data, "8BPS", 4) == 0)
        return fail(PT_ERR_UNSUPPORTED, "PSD images are not supported (PNG, BMP, TGA and binary PNM are)");
    Img im;
    std::string why;
    bool ok;
    if (len >= 26 && data[0] == 'B' && data[1] == 'M') {
        ok = decode_bmp(data, len, rgb_out == nullptr, &im, &why);
    } else if (len >= 3 && data[0] == 'P' && (data[1] == '5' || data[1] == '6')) {
        ok = decode_pnm(data, len, rgb_out == nullptr, &im, &why);
    } else if (tga_header_ok(data, len)) {
        ok = decode_tga(data, len, rgb_out == nullptr, &im, &why);
    } else {
        return fail(PT_ERR_ARG, "image: unknown format (PNG, BMP, TGA and binary PNM are read)");
    }
    if (!ok) return fail(PT_ERR_ARG, why);
    *width = im.w;
    *height = im.h;
    if (!rgb_out) return PT_OK;
    const size_t n = (size_t)im.w * im.h * 3;
    for (size_t i = 0; i < n; ++i) {
        // Color3f(byte) / 255, then ApplyGamma for colour textures
        const float v = (float)im.rgb[i] / 255.0f;
        rgb_out[i] = color ? powf(v, 2.2f) : v;
    }
    return PT_OK;
}

} // namespace

extern "C" int pt_image_decode(const uint8_t* data, int64_t len, int32_t color, int32_t* width, int32_t* height,
                               float* rgb_out, char* err, int32_t err_len)
{
    try {
        return decode(data, len, color, width, height, rgb_out, err, err_len);
    } catch (const std::bad_alloc&) {
        if (err && err_len > 0) std::snprintf(err, (size_t)err_len, "image: not enough host memory");
        return PT_ERR_OOM;
    }
}
