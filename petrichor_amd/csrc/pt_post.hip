// pt_post.hip — the output side of a render: the reference's 8-bit PNG tonemap
// (Core/Texture2D.cpp:18-91, 199-222) as one pass over the image in HBM.
//
// The reference maps every channel value through ACES filmic, the "RTCamp7" curve and a
// degamma, with the C library's powf, and truncates to a byte. The byte is a function of
// the input float alone, stored in pt_tonemap_lut.h as the runs of equal bytes over the
// 2^32 input bit patterns (at most PT_TONEMAP_RUNS per channel, generated on the host by
// gen_tonemap_lut.cpp), with an index over the top 12 bits of the pattern. A block copies
// the tables to LDS once (39 KB) and then converts pixels: per value one index read and a
// binary search of five steps over the runs of its bucket, against 12 bytes read
// and 3 written per pixel in HBM.
#include <hip/hip_runtime.h>

#include "pt_kernels.h"

#define PT_TONEMAP_STORAGE static __constant__
#include "pt_tonemap_lut.h"

static_assert((PT_TONEMAP_RUNS & (PT_TONEMAP_RUNS - 1)) == 0, "the run table must be a power of two long");
static_assert(PT_TONEMAP_MAX_SPAN_POW2 == 32, "five search steps per value");

namespace pt {

namespace {

constexpr int kPostBlock = 256;
// LDS per block: runs 3 * PT_TONEMAP_RUNS * 5 bytes + index 3 * (PT_TONEMAP_BUCKETS + 1) * 2
// bytes = 39.6 KB at 1024 runs and 4096 buckets: 4 blocks per CU
constexpr int kPostBlocksPerCu = 4;
constexpr int kBucketRow = PT_TONEMAP_BUCKETS + 1;

struct Tables {
    uint32_t start[3 * PT_TONEMAP_RUNS];
    uint16_t bucket[3 * kBucketRow];
    uint8_t byte[3 * PT_TONEMAP_RUNS];
};

// Largest i with start[i] <= u, searched between the bounds [lo, hi] of u's bucket
// (start[lo] <= u; a bucket spans at most 31 runs, which the generator checks) in five
// fixed steps, so that the twelve searches of a thread issue their LDS reads together:
// with the answer a in [i, i + 2s), step s moves i to min(i + s, hi) when that run starts
// at or below u.
__device__ __forceinline__ uint32_t run_byte(const Tables& t, int ch, uint32_t u)
{
    const uint16_t* bk = t.bucket + ch * kBucketRow;
    const uint32_t* st = t.start + ch * PT_TONEMAP_RUNS;
    const uint32_t b = u >> PT_TONEMAP_BUCKET_SHIFT;
    int i = bk[b];
    const int hi = bk[b + 1];
#pragma unroll
    for (int s = PT_TONEMAP_MAX_SPAN_POW2 / 2; s > 0; s >>= 1) {
        const int j = min(i + s, hi);
        if (st[j] <= u) i = j;
    }
    return t.byte[ch * PT_TONEMAP_RUNS + i];
}

__device__ __forceinline__ void convert_group(const Tables& t, const float4& a, const float4& b, const float4& c,
                                              uint32_t* dst)
{
    const float v[12] = {a.x, a.y, a.z, a.w, b.x, b.y, b.z, b.w, c.x, c.y, c.z, c.w};
    uint32_t w[3] = {0u, 0u, 0u};
#pragma unroll
    for (int k = 0; k < 12; ++k) {
        // value 12 g + k of the image is channel k % 3
        w[k / 4] |= run_byte(t, k % 3, __float_as_uint(v[k])) << (8 * (k % 4));
    }
    dst[0] = w[0];
    dst[1] = w[1];
    dst[2] = w[2];
}

// VEC: rgb is 16-byte and out 4-byte aligned; a thread converts 4 pixels (12 values: three
// 16-byte loads, three 4-byte stores) per turn of a grid-stride loop, with the next turn's
// loads issued before the current group is converted. The values after the last whole
// group of 4 pixels, and every value when !VEC, go one per thread.
template <bool VEC>
__global__ void __launch_bounds__(kPostBlock) k_tonemap_rgb8(const float* __restrict__ rgb, uint8_t* __restrict__ out,
                                                            int64_t num_values)
{
    __shared__ Tables t;
    for (int i = threadIdx.x; i < 3 * PT_TONEMAP_RUNS; i += kPostBlock) {
        t.start[i] = (&pt_tonemap_start[0][0])[i];
        t.byte[i] = (&pt_tonemap_byte[0][0])[i];
    }
    for (int i = threadIdx.x; i < 3 * kBucketRow; i += kPostBlock) t.bucket[i] = (&pt_tonemap_bucket[0][0])[i];
    __syncthreads();
    const int64_t stride = (int64_t)gridDim.x * kPostBlock;
    const int64_t first = (int64_t)blockIdx.x * kPostBlock + threadIdx.x;
    int64_t tail = 0;
    if (VEC) {
        const int64_t groups = num_values / 12;
        tail = groups * 12;
        const float4* src = reinterpret_cast<const float4*>(rgb);
        uint32_t* dst = reinterpret_cast<uint32_t*>(out);
        int64_t g = first;
        float4 a, b, c;
        if (g < groups) {
            a = src[3 * g];
            b = src[3 * g + 1];
            c = src[3 * g + 2];
        }
        while (g < groups) {
            const int64_t gn = g + stride;
            float4 na, nb, nc;
            if (gn < groups) {
                na = src[3 * gn];
                nb = src[3 * gn + 1];
                nc = src[3 * gn + 2];
            }
            convert_group(t, a, b, c, dst + 3 * g);
            a = na;
            b = nb;
            c = nc;
            g = gn;
        }
    }
    for (int64_t i = tail + first; i < num_values; i += stride) {
        out[i] = (uint8_t)run_byte(t, (int)(i % 3), __float_as_uint(rgb[i]));
    }
}

} // namespace

hipError_t launch_tonemap_rgb8(const float* rgb, uint8_t* out, int64_t num_pixels, hipStream_t st)
{
    if (num_pixels <= 0) return hipSuccess;
    const int64_t num_values = 3 * num_pixels;
    const bool vec = ((uintptr_t)rgb % 16) == 0 && ((uintptr_t)out % 4) == 0;
    const int64_t items = vec ? num_values / 12 + num_values % 12 : num_values;
    const int64_t want = (items + kPostBlock - 1) / kPostBlock;
    const int64_t cap = (int64_t)cu_count() * kPostBlocksPerCu;
    const dim3 grid((unsigned)(want < cap ? want : cap)), block(kPostBlock);
    if (vec) hipLaunchKernelGGL(k_tonemap_rgb8<true>, grid, block, 0, st, rgb, out, num_values);
    else hipLaunchKernelGGL(k_tonemap_rgb8<false>, grid, block, 0, st, rgb, out, num_values);
    return hipGetLastError();
}

} // namespace pt
