// pt_order.hip — coherence order for large launches of the trace entry points.
//
// Rays that start near each other walk the same subtrees. A launch of rays in arbitrary
// order (the traversal microbenchmark's surface rays start anywhere in the scene) keeps
// every wave's rays in different parts of the tree, so node reads miss the caches. The
// trace entry points therefore sort the ray indices of a large launch by the Morton code
// of the ray origin, quantised to 10 bits per axis over the BVH's root box (a stable
// radix sort, so rays from one point, such as a pinhole camera's, keep their order), and
// the trace kernel reads ray order[i] and writes its hit back at order[i]. Hits do not
// depend on the order in which rays are traced.
//
// The file also holds the fixed device scratch of the allocation-free trace entry point
// (pt_trace_closest_noalloc): its work cursors and the sort buffers of at most
// PT_NOALLOC_MAX_RAYS rays live in one static device array per GPU.
#include <hip/hip_runtime.h>
#include <rocprim/device/device_radix_sort.hpp>
#include <rocprim/device/device_scan.hpp>

#include <mutex>

#include "pt_kernels.h"

namespace pt {

namespace {

__device__ __forceinline__ uint32_t spread10(uint32_t x)
{
    x &= 0x3ffu;
    x = (x | (x << 16)) & 0x030000ffu;
    x = (x | (x << 8)) & 0x0300f00fu;
    x = (x | (x << 4)) & 0x030c30c3u;
    x = (x | (x << 2)) & 0x09249249u;
    return x;
}

__device__ __forceinline__ uint32_t quantise(float v, float lo, float scale)
{
    const float q = (v - lo) * scale;
    // NaN and values outside the box clamp to the box's faces
    return q >= 1023.0f ? 1023u : (q > 0.0f ? (uint32_t)q : 0u);
}

__global__ void __launch_bounds__(256) k_ray_keys(const float4* __restrict__ rays, int n, float3 lo, float3 scale,
                                                  uint32_t* __restrict__ keys, int32_t* __restrict__ idx)
{
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        const float4 o = rays[2 * (size_t)i];
        keys[i] = spread10(quantise(o.x, lo.x, scale.x)) | (spread10(quantise(o.y, lo.y, scale.y)) << 1) |
                  (spread10(quantise(o.z, lo.z, scale.z)) << 2);
        idx[i] = i;
    }
}

} // namespace

size_t ray_order_temp_bytes(int n)
{
    // the sort ping-pongs between the two halves of keys and idx, so its own temporary
    // storage holds only the digit histograms and the look-back state
    size_t bytes = 0;
    rocprim::double_buffer<uint32_t> k(nullptr, nullptr);
    rocprim::double_buffer<int32_t> v(nullptr, nullptr);
    if (rocprim::radix_sort_pairs(nullptr, bytes, k, v, n, 0, 30) != hipSuccess) return 0;
    return bytes;
}

hipError_t launch_ray_order(const float4* rays, int n, V3 box_lo, V3 box_hi, uint32_t* keys, int32_t* idx,
                            void* temp, size_t temp_bytes, hipStream_t st, int32_t** sorted)
{
    *sorted = idx;
    if (n <= 0) return hipSuccess;
    auto inv = [](float lo, float hi) { return hi > lo ? 1024.0f / (hi - lo) : 0.0f; };
    const float3 lo = make_float3(box_lo.x, box_lo.y, box_lo.z);
    const float3 scale = make_float3(inv(box_lo.x, box_hi.x), inv(box_lo.y, box_hi.y), inv(box_lo.z, box_hi.z));
    const int64_t want = ((int64_t)n + 255) / 256, cap = (int64_t)cu_count() * 8;
    hipLaunchKernelGGL(k_ray_keys, dim3((unsigned)(want < cap ? want : cap)), dim3(256), 0, st, rays, n, lo, scale,
                       keys, idx);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    // stable LSD radix sort over the 30 key bits; the sorted indices end in one of the two
    // halves of idx, which the number of passes (known on the host) decides
    size_t bytes = temp_bytes;
    rocprim::double_buffer<uint32_t> k(keys, keys + n);
    rocprim::double_buffer<int32_t> v(idx, idx + n);
    e = rocprim::radix_sort_pairs(temp, bytes, k, v, n, 0, 30, st);
    *sorted = v.current();
    return e;
}

template <typename T>
hipError_t exclusive_sum(void* temp, size_t& temp_bytes, const T* in, T* out, size_t n, hipStream_t st)
{
    return rocprim::exclusive_scan(temp, temp_bytes, in, out, T(0), n, rocprim::plus<T>(), st);
}
template hipError_t exclusive_sum<int>(void*, size_t&, const int*, int*, size_t, hipStream_t);
template hipError_t exclusive_sum<unsigned long long>(void*, size_t&, const unsigned long long*, unsigned long long*,
                                                      size_t, hipStream_t);

namespace {

// Scratch of pt_trace_closest_noalloc, one copy per GPU (the loader places the array in
// each device's memory when the code object is loaded there, and it stays until the
// process ends). Layout: the work cursors and the ray count (PT_CURSOR_INTS + 1 ints,
// padded to 4 KiB), then the keys and indices of the coherence order (2 x 2 x
// PT_NOALLOC_MAX_RAYS words), then the radix sort's temporary storage.
constexpr size_t kScratchHead = 4096;
constexpr size_t kScratchSortTemp = (size_t)4 << 20;
constexpr size_t kScratchBytes = kScratchHead + 16 * (size_t)PT_NOALLOC_MAX_RAYS + kScratchSortTemp;
static_assert((PT_CURSOR_INTS + 1) * sizeof(int) <= kScratchHead, "cursor block");

__device__ __align__(256) unsigned char scratch[kScratchBytes];

std::mutex g_scratch_mu;
void* g_scratch_ptr[PT_MAX_DEVICES] = {};

} // namespace

hipError_t noalloc_scratch_init(int device)
{
    if (device < 0 || device >= PT_MAX_DEVICES) return hipErrorInvalidDevice;
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (g_scratch_ptr[device]) return hipSuccess;
    int cur = -1;
    hipError_t e = hipGetDevice(&cur);
    if (e != hipSuccess) return e;
    if (cur != device && (e = hipSetDevice(device)) != hipSuccess) return e;
    void* p = nullptr;
    e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (cur != device) (void)hipSetDevice(cur);
    if (e != hipSuccess) return e;
    g_scratch_ptr[device] = p;
    return hipSuccess;
}

NoallocScratch noalloc_scratch(int device)
{
    NoallocScratch s{};
    if (device < 0 || device >= PT_MAX_DEVICES) return s;
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    unsigned char* base = (unsigned char*)g_scratch_ptr[device];
    if (!base) return s;
    s.cursor = (int*)base;
    s.keys = (uint32_t*)(base + kScratchHead);
    s.idx = (int32_t*)(base + kScratchHead + 8 * (size_t)PT_NOALLOC_MAX_RAYS);
    s.temp = base + kScratchHead + 16 * (size_t)PT_NOALLOC_MAX_RAYS;
    s.temp_bytes = kScratchSortTemp;
    return s;
}

namespace {

__global__ void __launch_bounds__(256) k_gather_rows(const float* __restrict__ src, const int32_t* __restrict__ idx,
                                                     int n, int row, float* __restrict__ dst)
{
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        const size_t from = (size_t)idx[i] * row, to = (size_t)i * row;
        for (int c = 0; c < row; ++c) dst[to + c] = src[from + c];
    }
}

} // namespace

hipError_t launch_gather_rows(const float* src, const int32_t* idx, int n, int row, float* dst, hipStream_t st)
{
    if (n <= 0) return hipSuccess;
    const int64_t want = ((int64_t)n + 255) / 256, cap = (int64_t)cu_count() * 8;
    hipLaunchKernelGGL(k_gather_rows, dim3((unsigned)(want < cap ? want : cap)), dim3(256), 0, st, src, idx, n, row,
                       dst);
    return hipGetLastError();
}

} // namespace pt
