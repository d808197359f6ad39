// gen_tonemap_lut.cpp — writes pt_tonemap_lut.h, the table behind the 8-bit output
// kernel (k_tonemap_rgb8 in pt_post.hip).
//
// The reference writes a PNG through ACES filmic -> "RTCamp7" tonemap -> degamma ->
// truncation to 8 bits (Core/Texture2D.cpp:18-91, 197-222), all in float, with the
// gamma curves taken by std::pow(float, float), i.e. the C library's powf. glibc's
// powf is not correctly rounded (on [0, 2] it differs from the correctly rounded value
// on about 0.05 % of the inputs for each of the six exponents used here), so a device
// pow cannot reproduce the reference's bytes. Each output byte, however, depends only
// on its own input float, and that function of the 2^32 input bit patterns is a short
// sequence of runs of equal bytes. This program evaluates the reference's formulas with
// the host's powf on every bit pattern and prints, per channel, the start of every run
// and its byte, and an index over the top 12 bits of the pattern: bucket b (the patterns
// b 2^20 .. (b + 1) 2^20 - 1) lies in the runs pt_tonemap_bucket[c][b] ..
// pt_tonemap_bucket[c][b + 1], at most 31 of them. The kernel reads the bucket's bounds
// and binary-searches the runs between them.
//
//   g++ -std=c++17 -O2 -ffp-contract=off -fopenmp gen_tonemap_lut.cpp -o gen_tonemap_lut
//   ./gen_tonemap_lut > pt_tonemap_lut.h        (python -m petrichor_amd.build --tonemap-lut)
//
// The table is a function of the host C library's powf (glibc 2.35 on x86-64 with FMA
// here); regenerate it on a host whose powf differs.
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <vector>

namespace {

// Core/Texture2D.cpp:18-27, evaluated one rounding at a time
float aces_film(float x)
{
    const float a = 2.51f, b = 0.03f, c = 2.43f, d = 0.59f, e = 0.14f;
    return (x * (a * x + b)) / (x * (c * x + d) + e);
}

// std::clamp(v, 0, 1) (Core/Texture2D.cpp:84): NaN passes through
float clamp01(float v) { return v < 0.0f ? 0.0f : (1.0f < v ? 1.0f : v); }

// Core/Texture2D.cpp:80-86
float degamma(float v)
{
    v = powf(v, 1.1f);
    v = clamp01(v);
    return powf(v, 1 / 2.2f);
}

// static_cast<uint8_t>((kColorDepth8Bit - kEps) * v) (Core/Texture2D.cpp:209-214,
// Core/Texture2D.h:117, Core/Constants.h:8). The product lies in [0, 255) or is NaN; the
// x86-64 conversion goes through a 32-bit integer, which a NaN turns into 0x80000000,
// whose low byte is 0.
uint8_t to_byte(float v)
{
    const float q = (255 - 1.0e-3f) * v;
    if (!(q == q)) return 0;
    return (uint8_t)(int32_t)q;
}

// The three output bytes of one input value x taken as the red, green and blue channel
// (Core/Texture2D.cpp:58-76: ApplyToneMapping = RTCamp7Tonemap(ACESFilm(c))).
void bytes_of(float x, uint8_t out[3])
{
    const float x0 = powf(0.05f * 1.0f, 2.2f);
    const float x1 = powf(0.9f * 1.0f, 2.2f);
    const float c = (powf(aces_film(x), 1.2f) - x0) / (x1 - x0);
    out[0] = to_byte(degamma(c));
    out[1] = to_byte(degamma(powf(c, 1.3f)));
    out[2] = to_byte(degamma(powf(c, 1.6f)));
}

struct Run {
    uint32_t start;
    uint8_t byte;
};

} // namespace

int main()
{
    const int kShift = 20, kBuckets = 1 << (32 - kShift);
    const int kChunks = 1024;
    const uint64_t kChunk = (1ull << 32) / kChunks;
    std::vector<std::vector<Run>> part[3];
    for (int ch = 0; ch < 3; ++ch) part[ch].resize(kChunks);
#pragma omp parallel for schedule(dynamic, 1)
    for (int k = 0; k < kChunks; ++k) {
        int last[3] = {-1, -1, -1};
        for (uint64_t u = k * kChunk; u < (k + 1) * kChunk; ++u) {
            const uint32_t bits = (uint32_t)u;
            float x;
            std::memcpy(&x, &bits, 4);
            uint8_t b[3];
            bytes_of(x, b);
            for (int ch = 0; ch < 3; ++ch) {
                if (b[ch] != last[ch]) {
                    part[ch][k].push_back(Run{bits, b[ch]});
                    last[ch] = b[ch];
                }
            }
        }
    }
    std::vector<Run> runs[3];
    size_t most = 0;
    for (int ch = 0; ch < 3; ++ch) {
        for (int k = 0; k < kChunks; ++k) {
            for (const Run& r : part[ch][k]) {
                if (runs[ch].empty() || runs[ch].back().byte != r.byte) runs[ch].push_back(r);
            }
        }
        most = runs[ch].size() > most ? runs[ch].size() : most;
    }
    // pad every channel to one power of two: padding runs start at 0xffffffff and repeat
    // the last byte, so a binary search over the padded table returns the same byte
    size_t cap = 1;
    while (cap < most) cap <<= 1;
    printf("// pt_tonemap_lut.h — generated by gen_tonemap_lut.cpp (do not edit).\n");
    printf("// Runs of equal output bytes of the reference's PNG tonemap over the 2^32 bit\n");
    printf("// patterns of an input float, per channel (red, green, blue): run i covers the bit\n");
    printf("// patterns from pt_tonemap_start[c][i] up to the next start. Runs per channel: %zu, %zu, %zu.\n",
           runs[0].size(), runs[1].size(), runs[2].size());
    printf("#pragma once\n#include <stdint.h>\n\n");
    printf("#ifndef PT_TONEMAP_STORAGE\n#define PT_TONEMAP_STORAGE static const\n#endif\n\n");
    printf("#define PT_TONEMAP_RUNS %zu\n", cap);
    printf("#define PT_TONEMAP_BUCKET_SHIFT %d\n#define PT_TONEMAP_BUCKETS %d\n", kShift, kBuckets);
    // every bucket spans fewer than 32 runs (bound[b + 1] - bound[b] < 32): the kernel's
    // search takes five steps
    printf("#define PT_TONEMAP_MAX_SPAN_POW2 32\n\n");
    printf("PT_TONEMAP_STORAGE uint32_t pt_tonemap_start[3][PT_TONEMAP_RUNS] = {\n");
    for (int ch = 0; ch < 3; ++ch) {
        printf("    {");
        for (size_t i = 0; i < cap; ++i) {
            const uint32_t s = i < runs[ch].size() ? runs[ch][i].start : 0xffffffffu;
            printf("%s0x%08xu", i % 8 == 0 ? "\n        " : " ", s);
            if (i + 1 < cap) printf(",");
        }
        printf("},\n");
    }
    printf("};\n\nPT_TONEMAP_STORAGE uint8_t pt_tonemap_byte[3][PT_TONEMAP_RUNS] = {\n");
    for (int ch = 0; ch < 3; ++ch) {
        printf("    {");
        for (size_t i = 0; i < cap; ++i) {
            const uint8_t b = i < runs[ch].size() ? runs[ch][i].byte : runs[ch].back().byte;
            printf("%s%u", i % 16 == 0 ? "\n        " : " ", b);
            if (i + 1 < cap) printf(",");
        }
        printf("},\n");
    }
    printf("};\n\nPT_TONEMAP_STORAGE uint16_t pt_tonemap_bucket[3][PT_TONEMAP_BUCKETS + 1] = {\n");
    for (int ch = 0; ch < 3; ++ch) {
        printf("    {");
        size_t r = 0, prev = 0; // the run holding the first pattern of bucket b
        for (int b = 0; b <= kBuckets; ++b) {
            if (b < kBuckets) {
                const uint32_t first = (uint32_t)b << kShift;
                while (r + 1 < runs[ch].size() && runs[ch][r + 1].start <= first) ++r;
            } else {
                r = runs[ch].size() - 1; // the run holding the last pattern
            }
            if (r - prev >= 32) {
                fprintf(stderr, "bucket %d of channel %d spans %zu runs: the kernel searches 31\n", b - 1, ch, r - prev);
                return 1;
            }
            prev = r;
            printf("%s%zu", b % 16 == 0 ? "\n        " : " ", r);
            if (b < kBuckets) printf(",");
        }
        printf("},\n");
    }
    printf("};\n");
    return 0;
}
