// pt_device.h — per-ray arithmetic of the render path, for gfx950 kernels and the
// host builder.
//
// Every function restates a reference function (cited file:line, paths relative to
// the reference's LibPetrichor/) with the same float evaluation order; translation
// units including this are built with -ffp-contract=off and correctly rounded
// f32 division/sqrt, so results are bit-identical to the x86-64 reference build.
// Calls the reference makes to the double C library (sin/cos/acos/atan2 and a few
// sqrt whose result is not stored straight to a float) go through pt_libm.h.
#pragma once

#include "pt_libm.h"

#include <float.h>
#include <string.h>
#include <stdint.h>

namespace pt {

#define PT_EPS 1.0e-3f                  /* kEps, Core/Constants.h:8 */
#define PT_INFINITY (0.9f * FLT_MAX)    /* kInfinity, Core/Constants.h:10 */
#define PT_PI 3.14159265358979f         /* Math::kPi, Core/Constants.h:14 */
#define PT_INV_PI (1.0f / 3.14159265358979f)

struct V3 {
    float x, y, z;
};

PT_HD V3 v3(float x, float y, float z)
{
    V3 r;
    r.x = x;
    r.y = y;
    r.z = z;
    return r;
}
// Math/Vector3f.h:236-357 operators, element by element.
PT_HD V3 operator+(V3 a, V3 b) { return v3(a.x + b.x, a.y + b.y, a.z + b.z); }
PT_HD V3 operator-(V3 a, V3 b) { return v3(a.x - b.x, a.y - b.y, a.z - b.z); }
PT_HD V3 operator-(V3 a) { return v3(-a.x, -a.y, -a.z); }
PT_HD V3 operator*(V3 a, V3 b) { return v3(a.x * b.x, a.y * b.y, a.z * b.z); }
PT_HD V3 operator*(V3 a, float c) { return v3(a.x * c, a.y * c, a.z * c); }
PT_HD V3 operator*(float c, V3 a) { return v3(a.x * c, a.y * c, a.z * c); }
PT_HD V3 operator/(V3 a, float c) { return v3(a.x / c, a.y / c, a.z / c); }
PT_HD V3 operator/(V3 a, V3 b) { return v3(a.x / b.x, a.y / b.y, a.z / b.z); }
PT_HD float dot(V3 a, V3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; } // Vector3f.h:128-132
PT_HD V3 cross(V3 a, V3 b) // Vector3f.h:134-140
{
    return v3(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x);
}
PT_HD float sq_length(V3 a) { return dot(a, a); }
PT_HD float length(V3 a) { return __builtin_sqrtf(dot(a, a)); }
PT_HD V3 normalized(V3 a) { return a / length(a); }
PT_HD float smin(float a, float b) { return (b < a) ? b : a; } // std::min
PT_HD float smax(float a, float b) { return (a < b) ? b : a; } // std::max
PT_HD float min_elem(V3 a) { return smin(a.x, smin(a.y, a.z)); }
PT_HD float fabsf_(float x) { return __builtin_fabsf(x); }
PT_HD float copysignf_(float m, float s) { return __builtin_copysignf(m, s); }
PT_HD float floorf_(float x) { return __builtin_floorf(x); }
PT_HD float comp(V3 a, int i) { return i == 0 ? a.x : (i == 1 ? a.y : a.z); }
PT_HD V3 reflected(V3 d, V3 n) { return d - n * 2.0f * dot(d, n); } // Vector3f.h:203-208
PT_HD float pow2(float x) { return x * x; }
PT_HD float pow5(float x) { return x * (pow2(x) * pow2(x)); }          // MathUtils.h:38-56
PT_HD float fmodp(float x, float m) { return x - m * floorf_(x / m); } // MathUtils.h:30-35
PT_HD float lerpf(float a, float b, float t) { return (1.0f - t) * a + t * b; }

// ---- XorShift128 + samplers (Random/XorShift.h:36-52, Core/Sampler) ----------
struct XorShift {
    uint32_t x, y, z, w;
};
PT_HD XorShift xorshift_seeded(uint32_t seed)
{
    XorShift s;
    s.x = 123456789u;
    s.y = 362436069u;
    s.z = 521288629u;
    s.w = seed;
    return s;
}
PT_HD uint32_t xs_random(XorShift& s)
{
    const uint32_t t = s.x ^ (s.x << 11);
    s.x = s.y;
    s.y = s.z;
    s.z = s.w;
    s.w = (s.w ^ (s.w >> 19)) ^ (t ^ (t >> 8));
    return s.w;
}
// random() / float(UINT_MAX) == (float)u32 / 2^32
PT_HD float xs_next(XorShift& s) { return (float)xs_random(s) / 4294967296.0f; }

PT_HD float bits_to_float(uint32_t u)
{
    float f;
    memcpy(&f, &u, 4);
    return f;
}
PT_HD uint32_t float_to_bits(float f)
{
    uint32_t u;
    memcpy(&u, &f, 4);
    return u;
}

// Sampler state of one path: the 1D stream and the two streams of the 2D sampler
// (RandomSampler2D: x stream, y stream). With a micro-jittered 2D sampler `mj` is the
// job's Halton table (entry 0: the sequence length as bits and mu; then the pairs in bases
// 2 and 3), s2x is the sampler's offset stream and s2y holds its index and the two offsets
// (as bits).
struct Samplers {
    XorShift s1, s2x, s2y;
    const float2* mj = nullptr;
};
PT_HD float next1(Samplers& s) { return xs_next(s.s1); }
// MicroJitteredSampler::Next (Core/Sampler/MicroJitteredSampler.cpp:28-48): the Halton pair
// at the index plus the offsets, each taken Mod(x, 1) = x - 1 * floor(x / 1); after the
// last point of the sequence two new offsets mu * (next() - 0.5) and index 0.
PT_HD void mj_next(Samplers& s, float* a, float* b)
{
    const float2 h = s.mj[1 + s.s2y.x];
    float x = h.x + bits_to_float(s.s2y.y), y = h.y + bits_to_float(s.s2y.z);
    x = x - 1.0f * floorf(x / 1.0f);
    y = y - 1.0f * floorf(y / 1.0f);
    if (++s.s2y.x >= float_to_bits(s.mj[0].x)) {
        const float mu = s.mj[0].y;
        s.s2y.y = float_to_bits(mu * (xs_next(s.s2x) - 0.5f));
        s.s2y.z = float_to_bits(mu * (xs_next(s.s2x) - 0.5f));
        s.s2y.x = 0;
    }
    *a = x;
    *b = y;
}
PT_HD void next2(Samplers& s, float* a, float* b)
{
    if (s.mj != nullptr) {
        mj_next(s, a, b);
        return;
    }
    *a = xs_next(s.s2x);
    *b = xs_next(s.s2y);
}

// Per-sample seeds (this library's definition; the reference seeds one sampler pair per
// tile, Core/Petrichor.cpp:72-77): MurmurHash3 fmix32 over (job seed, pixel index,
// stream, sample index). Sample s of a pixel starts three fresh XorShift128 streams, so
// the samples of one pixel are independent and run in parallel; sample 0's seeds are the
// pixel seeds (fmix32(0) = 0). XorShift128(seed) sets only w (Random/XorShift.h:17-18),
// and the first outputs of such a stream are strongly correlated (the first two have
// correlation -0.88), so each stream is used from its PT_STREAM_WARMUP-th output on.
PT_HD uint32_t fmix32(uint32_t h)
{
    h ^= h >> 16;
    h *= 0x85ebca6bu;
    h ^= h >> 13;
    h *= 0xc2b2ae35u;
    h ^= h >> 16;
    return h;
}
PT_HD uint32_t sample_seed(uint32_t job_seed, uint32_t pixel, uint32_t sample, uint32_t stream)
{
    return fmix32(fmix32(fmix32(job_seed) ^ pixel) + 0x9e3779b9u * (stream + 1u) + fmix32(sample));
}
#define PT_STREAM_WARMUP 24
PT_HD XorShift xorshift_warmed(uint32_t seed)
{
    XorShift x = xorshift_seeded(seed);
    for (int i = 0; i < PT_STREAM_WARMUP; ++i) (void)xs_random(x);
    return x;
}
// With a micro-jittered 2D sampler (this library's definition, like the seeds): sample s
// starts at Halton point s mod length with offsets drawn from its x stream as at the end
// of a sequence, so the samples of a pixel take successive Halton points, each jittered
// on its own.
PT_HD Samplers samplers_for_sample(uint32_t job_seed, uint32_t pixel, uint32_t sample,
                                   const float2* mj = nullptr)
{
    Samplers s;
    s.s1 = xorshift_warmed(sample_seed(job_seed, pixel, sample, 0));
    s.s2x = xorshift_warmed(sample_seed(job_seed, pixel, sample, 1));
    s.s2y = xorshift_warmed(sample_seed(job_seed, pixel, sample, 2));
    if (mj != nullptr) {
        s.mj = mj;
        const float mu = mj[0].y;
        s.s2y.x = sample % float_to_bits(mj[0].x);
        s.s2y.y = float_to_bits(mu * (xs_next(s.s2x) - 0.5f));
        s.s2y.z = float_to_bits(mu * (xs_next(s.s2x) - 0.5f));
        s.s2y.w = 0;
    }
    return s;
}

// ---- rays ----------------------------------------------------------------------
struct Ray { // Core/Ray.h:20-49
    V3 o, dir, throughput;
    int bounce;
    float prob, ior;
};
PT_HD Ray make_ray(V3 o, V3 d)
{
    Ray r;
    r.o = o;
    r.dir = d;
    r.throughput = v3(1.0f, 1.0f, 1.0f);
    r.bounce = 0;
    r.prob = 1.0f;
    r.ior = 1.0f;
    return r;
}

// ---- camera (Core/Camera.cpp:52-91) --------------------------------------------
struct CameraDev {
    V3 pos, right, up, forward;
    float focus, h_perf, aperture;
};

PT_HD Ray generate_ray(const CameraDev& c, int i, int j, int w, int h, Samplers& s)
{
    float r0, r1;
    next2(s, &r0, &r1);
    const float aspect = (float)w / h;
    const float u = (i + r0) / w - 0.5f;
    const float v = (j + r1) / h - 0.5f;
    const V3 sensor = c.pos + c.right * u * c.h_perf * c.focus * aspect + -c.up * v * c.h_perf * c.focus +
                      c.forward * c.focus;
    next2(s, &r0, &r1);
    const float r = __builtin_sqrtf(r0);
    const float theta = 2.0f * PT_PI * r1;
    double st, ct;
    pt_sincos((double)theta, &st, &ct);
    const V3 lens = c.pos + c.right * 0.5f * c.aperture * r * (float)ct + -c.up * 0.5f * c.aperture * r * (float)st;
    Ray ray = make_ray(lens, normalized(sensor - lens));
    const float d = dot(ray.dir, c.forward);
    ray.throughput = v3(1.0f, 1.0f, 1.0f) * d * d * c.focus * c.focus / sq_length(sensor - lens);
    return ray;
}

// ---- triangles (Core/Geometry/Triangle.cpp) -------------------------------------
// Triangle record: p0, e1 = p1 - p0, e2 = p2 - p0 and n = cross(e1, e2), the values
// Triangle::Intersect forms first (Triangle.cpp:57-60), stored precomputed.
struct TriDev {
    V3 p0, e1, e2, n;
};

PT_HD TriDev tri_from_points(V3 p0, V3 p1, V3 p2)
{
    TriDev t;
    t.p0 = p0;
    t.e1 = p1 - p0;
    t.e2 = p2 - p0;
    t.n = cross(t.e1, t.e2);
    return t;
}

// Triangle::Intersect (Triangle.cpp:51-88): returns true and the distance on a hit.
PT_HD bool tri_intersect(const TriDev& tr, V3 o, V3 d, float* dist_out)
{
    const float inv_det = 1.0f / dot(-d, tr.n);
    const V3 vec = o - tr.p0;
    const float w1 = dot(vec, cross(d, tr.e2)) * inv_det;
    const float w2 = dot(vec, cross(tr.e1, d)) * inv_det;
    if (w1 < 0.0f || w1 >= 1.0f) return false;
    if (w2 < 0.0f || w2 >= 1.0f) return false;
    if (w1 + w2 >= 1.0f) return false;
    const float dist = dot(vec, tr.n) * inv_det;
    if (dist < 0.0f) return false;
    *dist_out = dist;
    return true;
}

// Triangle::Interpolate (Triangle.cpp:90-153): position and shading normal.
PT_HD void tri_shading(const TriDev& tr, bool smooth, V3 n0, V3 n1, V3 n2, V3 o, V3 d, float t, V3* pos,
                       V3* normal)
{
    *pos = o + d * t;
    if (!smooth) {
        *normal = normalized(tr.n);
    } else {
        const float inv_det = 1.0f / dot(-d, tr.n);
        const V3 vec = o - tr.p0;
        const float w1 = dot(vec, cross(d, tr.e2)) * inv_det;
        const float w2 = dot(vec, cross(tr.e1, d)) * inv_det;
        const V3 n = w1 * (n1 - n0) + w2 * (n2 - n0) + n0;
        *normal = n / length(n);
    }
}

// Triangle::SampleSurface (Triangle.cpp:155-182) with e0 = e1, e1 = e2 of the record.
PT_HD void tri_sample(const TriDev& tr, Samplers& s, V3* pos, V3* normal, float* pdf_area)
{
    float u0, u1;
    next2(s, &u0, &u1);
    const float su = __builtin_sqrtf(u0);
    *pos = tr.p0 + (1.0f - su) * tr.e1 + su * u1 * tr.e2;
    *normal = normalized(tr.n);
    *pdf_area = 1.0f / (0.5f * length(tr.n));
}

// ---- spheres (Core/Geometry/Sphere.cpp) --------------------------------------------
// A sphere's leaf slot holds (centre, radius) in its first float4 (TriDev: p0 and e1.x).

// Sphere::Intersect (Sphere.cpp:32-59): the quadratic in float, its roots with the double
// sqrt the unqualified call binds to, each rounded to float once; a ray whose two roots
// are both behind its origin misses; else min(max(l1, 0), max(l2, 0)), which is 0 for an
// origin inside the sphere.
PT_HD bool sphere_intersect(V3 c, float r, V3 o, V3 d, float* t)
{
    const V3 oc = o - c;
    const float a = dot(d, d);
    const float b2 = dot(d, oc);
    const float cc = dot(oc, oc) - r * r;
    const float d4 = b2 * b2 - a * cc;
    if (d4 < 0.0f) return false;
    const double sq = pt_sqrt_d((double)d4);
    const float l1 = (float)(((double)-b2 - sq) / (double)a);
    const float l2 = (float)(((double)-b2 + sq) / (double)a);
    if (l1 < 0.0f && l2 < 0.0f) return false;
    *t = smin(smax(l1, 0.0f), smax(l2, 0.0f));
    return true;
}

// Sphere::Interpolate (Sphere.cpp:61-88): the normal points away from the centre, or
// towards it when the ray starts inside (|o - c|^2 <= r^2).
PT_HD void sphere_shading(V3 c, float r, V3 o, V3 d, float t, V3* pos, V3* normal)
{
    *pos = o + d * t;
    const float l2 = sq_length(o - c);
    const V3 n = normalized(*pos - c);
    *normal = l2 > r * r ? n : -n;
}

// Sphere::SampleSurface (Sphere.cpp:90-115): a direction in the cone the sphere subtends
// from p, in the frame OrthonormalBasis::Build(p - c) (Duff et al. 2017), theta = acos(1 -
// u0 (1 - r / l)), phi = 2 pi u1, through GetDir; pdf l / (2 pi r^2 (l - r)).
PT_HD void sphere_sample(V3 c, float r, V3 p, Samplers& s, V3* pos, V3* normal, float* pdf_area)
{
    const V3 op = p - c;
    const V3 w = normalized(op);
    const float sign = copysignf_(1.0f, w.z);
    const float a = -1.0f / (sign + w.z);
    const float b = w.x * w.y * a;
    const V3 u = v3(1.0f + sign * w.x * w.x * a, sign * b, -sign * w.x);
    const V3 vv = v3(b, sign + w.y * w.y * a, -w.y);
    float u0, u1;
    next2(s, &u0, &u1);
    const float l = length(op);
    const float theta = (float)pt_acos((double)(1.0f - u0 * (1.0f - r / l)));
    const float phi = 2.0f * PT_PI * u1;
    double st, ct, sp, cp;
    pt_sincos((double)theta, &st, &ct);
    pt_sincos((double)phi, &sp, &cp);
    const float cx = (float)(st * cp), cy = (float)(st * sp), cz = (float)ct;
    const V3 dir = cx * u + cy * vv + cz * w;
    *pos = c + r * dir;
    *normal = dir;
    *pdf_area = l / (2.0f * PT_PI * r * r * (l - r));
}

// ---- materials (Core/Material) ----------------------------------------------------
enum { kMatLambert = 0, kMatGlass = 1, kMatEmission = 2, kMatGlossy = 3, kMatMix = 4 };

struct MaterialDev {
    int type;
    V3 color;     // Lambert kd, Glass colour, Emission Le, GGX reflectance
    float alpha;  // GGX roughness^2 (GGX.cpp:16-20)
    float ior;    // Glass
    float mix;    // Mix weight
    int mat0, mat1;
    int is_enabled;
    V3 albedo;    // Lambert::GetAlbedo: one, or kd x the albedo map at the hit (Lambert.cpp:117-126)
};

// A texture of a material map: row-major RGB floats.
struct TexDev {
    const float* rgb;
    int w, h;
};

// Texture2D::GetPixelByUV (Texture2D.cpp:325-333) with its default bilinear fetch: u and v
// wrapped into [0, 1), scaled by the texture's size.
PT_HD V3 tex_bilinear(const float* rgb, int w, int h, float x, float y);
PT_HD V3 tex_by_uv(const TexDev& t, float u, float v)
{
    u = fmodp(u, 1.0f);
    v = fmodp(v, 1.0f);
    return tex_bilinear(t.rgb, t.w, t.h, u * t.w, v * t.h);
}

// GetLuminance (Core/Color3f.cpp:7-11)
PT_HD float luminance(V3 c) { return 0.2126f * c.x + 0.7152f * c.y + 0.0722f * c.z; }

// GGX::GetF0 (GGX.h:126-139), metalness 0, base colour 1, no map.
PT_HD V3 ggx_f0(const MaterialDev& m)
{
    return 0.16f * (m.color * m.color) * (1.0f - 0.0f) + v3(1.0f, 1.0f, 1.0f) * 0.0f;
}

// GGX::Lambda (GGX.cpp:141-153); the final expression is evaluated in double.
PT_HD float ggx_lambda(V3 dir, V3 half, float alpha)
{
    const float alpha2 = alpha * alpha;
    const float c = fabsf_(dot(dir, half));
    const float tan2 = c > 0 ? 1.0f / (c * c) - 1.0f : PT_INFINITY;
    const float arg = 1.0f + alpha2 * tan2;
    return (float)((double)-0.5f + (double)0.5f * pt_sqrt_d((double)arg));
}

// Material::BxDF (Lambert.cpp:27-33, GGX.cpp:22-57, Glass.cpp:11-19).
PT_HD V3 bxdf(const MaterialDev& m, V3 din, V3 dout, V3 normal)
{
    if (m.type == kMatLambert) return m.color * m.albedo * PT_INV_PI;
    if (m.type == kMatGlossy) {
        const V3 half = normalized(-din + dout);
        const float hn = fabsf_(dot(half, normal));
        const float alpha = m.alpha;
        const float alpha2 = alpha * alpha;
        const float hn2 = hn * hn;
        const float k = (alpha2 - 1.0f) * hn2 + 1.0f;
        const float d_term = alpha2 / (PT_PI * k * k);
        const float li = ggx_lambda(dout, half, alpha);
        const float lo = ggx_lambda(-din, half, alpha);
        const float g_term = 1.0f / (1.0f + li + lo);
        const float hl = fabsf_(dot(half, dout));
        const V3 f0 = ggx_f0(m);
        const V3 f_term = f0 + (v3(1.0f, 1.0f, 1.0f) - f0) * pow5(1.0f - hl);
        const float ln = fabsf_(dot(dout, normal));
        const float vn = fabsf_(dot(-din, normal));
        return d_term * g_term * f_term / (4.0f * ln * vn);
    }
    return v3(0.0f, 0.0f, 0.0f);
}

// Material::PDF (Lambert.cpp:35-49, GGX.cpp:59-76, Glass.cpp:21-27).
PT_HD float bsdf_pdf(const MaterialDev& m, V3 din, V3 dout, V3 normal)
{
    if (m.type == kMatLambert) {
        if (m.is_enabled) return fabsf_(dot(dout, normal)) * PT_INV_PI;
        return 0.5f * PT_INV_PI;
    }
    if (m.type == kMatGlossy) {
        const V3 half = normalized(-din + dout);
        const float hn = fabsf_(dot(half, normal));
        const float alpha2 = m.alpha * m.alpha;
        const float hn2 = hn * hn;
        const float tmp = 1.0f - (1.0f - alpha2) * hn2;
        return tmp > 0 ? PT_INV_PI * alpha2 / (tmp * tmp) : PT_INFINITY;
    }
    return 0.0f;
}

// GGX::SampleGGXVNDF (GGX.cpp:155-208), Heitz 2017.
PT_HD V3 ggx_sample_vndf(const MaterialDev& m, V3 view, V3 n0, Samplers& s)
{
    const V3 normal = n0 * copysignf_(1.0f, dot(view, n0));
    const V3 t = (fabsf_(normal.z) < 0.9999f) ? cross(normal, v3(0.0f, 0.0f, 1.0f)) : v3(1.0f, 0.0f, 0.0f);
    const V3 bw = normalized(normal);
    const V3 bu = normalized(t);
    const V3 bv = cross(bw, bu);
    V3 v = v3(dot(view, bu), dot(view, bv), dot(view, bw));
    const float alpha = m.alpha;
    v = v * v3(alpha, alpha, 1.0f);
    v = v / length(v);
    const V3 t1 = (v.z < 0.9999f) ? normalized(cross(v, v3(0.0f, 0.0f, 1.0f))) : v3(1.0f, 0.0f, 0.0f);
    const V3 t2 = cross(t1, v);
    float u0, u1;
    next2(s, &u0, &u1);
    const float a = 1.0f / (1.0f + v.z);
    const float r = __builtin_sqrtf(u0);
    const float phi = (u1 < a) ? u1 / a * PT_PI : PT_PI + (u1 - a) / (1.0f - a) * PT_PI;
    double sp, cp;
    pt_sincos((double)phi, &sp, &cp);
    const float p1 = (float)((double)r * cp);
    const float p2 = (float)((double)r * sp * (double)((u1 < a) ? 1.0f : v.z));
    V3 n = p1 * t1 + p2 * t2 + __builtin_sqrtf(smax(0.0f, 1.0f - p1 * p1 - p2 * p2)) * v;
    n = v3(alpha * n.x, alpha * n.y, smax(0.0f, n.z));
    n = bu * n.x + bv * n.y + bw * n.z;
    return n / length(n);
}

// Material::CreateNextRay (Lambert.cpp:51-102, GGX.cpp:78-139, Glass.cpp:29-95).
PT_HD Ray next_ray(const MaterialDev& m, const Ray& in, V3 pos, V3 sn, Samplers& s)
{
    Ray out = make_ray(v3(0.0f, 0.0f, 0.0f), v3(0.0f, 0.0f, 0.0f));
    if (m.type == kMatLambert) {
        const V3 normal = sn * copysignf_(1.0f, -dot(in.dir, sn));
        // OrthonormalBasis::Build(w) (OrthonormalBasis.cpp:10-27)
        const V3 w = normalized(normal);
        const float sign = copysignf_(1.0f, w.z);
        const float a = -1.0f / (sign + w.z);
        const float b = w.x * w.y * a;
        const V3 u = v3(1.0f + sign * w.x * w.x * a, sign * b, -sign * w.x);
        const V3 vv = v3(b, sign + w.y * w.y * a, -w.y);
        float r0, r1;
        next2(s, &r0, &r1);
        out.o = pos + PT_EPS * normal;
        out.bounce = in.bounce + 1;
        if (m.is_enabled) {
            const float phi = 2.0f * PT_PI * r1;
            const float sin_theta = __builtin_sqrtf(r0);
            double sp, cp;
            pt_sincos((double)phi, &sp, &cp);
            const float cu = (float)((double)sin_theta * cp);
            const float cv = (float)((double)sin_theta * sp);
            const float cw = __builtin_sqrtf(1.0f - sin_theta * sin_theta);
            out.dir = cu * u + cv * vv + cw * w;
            out.throughput = in.throughput * m.albedo;
        } else {
            const float cos_theta = r0;
            const float sin_theta = __builtin_sqrtf(1.0f - cos_theta * cos_theta);
            const float phi = 2.0f * PT_PI * r1;
            double sp, cp;
            pt_sincos((double)phi, &sp, &cp);
            const float cu = (float)((double)sin_theta * cp);
            const float cv = (float)((double)sin_theta * sp);
            out.dir = cu * u + cv * vv + cos_theta * w;
            out.throughput = in.throughput;
            const V3 f = bxdf(m, in.dir, out.dir, sn);
            out.throughput = out.throughput * (f * cos_theta / bsdf_pdf(m, in.dir, out.dir, sn));
        }
        return out;
    }
    if (m.type == kMatGlossy) {
        const V3 normal = sn * copysignf_(1.0f, -dot(in.dir, sn));
        const V3 h = ggx_sample_vndf(m, -in.dir, sn, s);
        const V3 dir = reflected(in.dir, h);
        const float hl = fabsf_(dot(h, dir));
        const V3 f0 = ggx_f0(m);
        const V3 f_term = f0 + (v3(1.0f, 1.0f, 1.0f) - f0) * pow5(1.0f - hl);
        out.o = pos;
        out.dir = dir;
        out.throughput = in.throughput;
        out.bounce = in.bounce + 1;
        const float li = ggx_lambda(dir, h, m.alpha);
        const float lo = ggx_lambda(-in.dir, h, m.alpha);
        const float g2 = 1.0f / (1.0f + li + lo);
        const float g1 = 1.0f / (1.0f + li);
        out.throughput = out.throughput * (f_term * g2 / g1);
        out.o = out.o + PT_EPS * normal;
        return out;
    }
    if (m.type == kMatGlass) {
        const V3 normal = sn * copysignf_(1.0f, -dot(in.dir, sn));
        const float eta = m.ior / in.ior;
        // Vector3f::Refracted (Vector3f.h:210-234) with the un-flipped normal
        const float dn = dot(in.dir, sn);
        const float disc = pow2(dn) - 1 + pow2(eta);
        if (disc < 0) {
            out.o = pos + PT_EPS * normal;
            out.dir = normalized(reflected(in.dir, normal));
            out.throughput = v3(0.0f, 0.0f, 0.0f);
            out.bounce = 0 + 1;
            out.prob = in.prob;
            out.ior = in.ior;
            return out;
        }
        const float k = (float)((double)(-dn) + pt_sqrt_d((double)disc));
        const V3 rd = (1.0f / eta) * (in.dir - sn * k);
        out.dir = normalized(rd);
        const float ndo = dot(sn, out.dir);
        const float f0 = pow2((1.0f - eta) / (1.0f + eta));
        const float refl = f0 + (1.0f - f0) * pow5(1.0f - fabsf_(ndo));
        float rx, ry;
        next2(s, &rx, &ry);
        out.throughput = in.throughput;
        out.bounce = in.bounce + 1;
        out.prob = in.prob;
        if (rx <= refl) {
            out.o = pos + PT_EPS * normal; // keeps the refracted direction (Glass.cpp:70-80)
            out.ior = in.ior;
            return out;
        }
        out.o = pos - PT_EPS * normal;
        out.ior = (ndo < 0) ? 1.0f * eta : 1.0f / eta;
        return out;
    }
    return in;
}

// ---- environment (Core/Environment.cpp) -------------------------------------------
struct EnvDev {
    const float* rgb; // w*h RGB texels or null
    int w, h;
    V3 base;
    float rot;
    int use_is;
};

PT_HD bool env_valid(const EnvDev& e) { return e.rgb != nullptr && e.w > 0 && e.h > 0; }
PT_HD bool env_use_is(const EnvDev& e) { return env_valid(e) && e.use_is; }

PT_HD int clampi(int v, int lo, int hi) { return v < lo ? lo : (hi < v ? hi : v); }

// Texture2D::GetPixel(float, float, Bilinear) (Texture2D.cpp:287-317)
PT_HD V3 tex_texel(const float* rgb, int w, int i, int j)
{
    const float* p = rgb + 3 * ((int64_t)w * j + i);
    return v3(p[0], p[1], p[2]);
}
PT_HD V3 tex_bilinear(const float* rgb, int w, int h, float x, float y)
{
    int ix0 = (int)x;
    int ix1 = (int)(x + 1.0f);
    int iy0 = (int)y;
    int iy1 = (int)(y + 1.0f);
    const float wx0 = x - ix0;
    const float wy0 = y - iy0;
    const float wx1 = ix1 - x;
    const float wy1 = iy1 - y;
    ix0 = clampi(ix0, 0, w - 1);
    ix1 = clampi(ix1, 0, w - 1);
    iy0 = clampi(iy0, 0, h - 1);
    iy1 = clampi(iy1, 0, h - 1);
    const V3 cm0 = wx1 * tex_texel(rgb, w, ix0, iy0) + wx0 * tex_texel(rgb, w, ix1, iy0);
    const V3 cm1 = wx1 * tex_texel(rgb, w, ix0, iy1) + wx0 * tex_texel(rgb, w, ix1, iy1);
    return wy1 * cm0 + wy0 * cm1;
}
// .x channel of the bilinear fetch of a never-loaded texture: every texel reads as
// (1, 0, 1) (Texture2D.cpp:263-266), so only the weights' arithmetic remains.
PT_HD float invalid_tex_bilinear_x(float x, float y)
{
    const int ix0 = (int)x;
    const int ix1 = (int)(x + 1.0f);
    const int iy0 = (int)y;
    const int iy1 = (int)(y + 1.0f);
    const float wx0 = x - ix0;
    const float wy0 = y - iy0;
    const float wx1 = ix1 - x;
    const float wy1 = iy1 - y;
    const float cm0 = wx1 * 1.0f + wx0 * 1.0f;
    const float cm1 = wx1 * 1.0f + wx0 * 1.0f;
    return wy1 * cm0 + wy0 * cm1;
}

// Polar angles of a direction as Environment::GetColor and GetImportanceSamplingPDF
// compute them (Environment.cpp:31-32, 225-226): double acos / atan2, rounded to float.
struct EnvAngles {
    float theta, phi;
};
PT_HD EnvAngles env_angles(V3 dir)
{
    EnvAngles a;
    a.theta = (float)pt_acos((double)dir.z);
    a.phi = (float)pt_atan2((double)dir.y, (double)dir.x);
    return a;
}

// Environment::GetColor (Environment.cpp:29-48) for a valid map, from the angles
PT_HD V3 env_color_at(const EnvDev& e, EnvAngles a)
{
    const float theta = a.theta, phi = a.phi;
    float u = 1.0f - (phi - e.rot) * 0.5f * PT_INV_PI;
    float v = theta * PT_INV_PI;
    u = fmodp(u, 1.0f);
    v = fmodp(v, 1.0f);
    u = fmodp(u, 1.0f); // Texture2D::GetPixelByUV wraps again (Texture2D.cpp:327-335)
    v = fmodp(v, 1.0f);
    return e.base * tex_bilinear(e.rgb, e.w, e.h, u * e.w, v * e.h);
}

PT_HD V3 env_color(const EnvDev& e, V3 dir)
{
    if (!env_valid(e)) return e.base;
    return env_color_at(e, env_angles(dir));
}

// The 1D tables of Environment::PreCalcCumulativeDistTex (Environment.cpp:56-128). The
// reference builds them from its 2D tables, which are never-loaded textures reading
// (1, 0, 1), so every row weighs 1 whatever the map: m_pdf1D[j] = H / H = 1 and
// m_cdf1D[j] = (j + 1) / H (a running float sum of ones, exact below 2^24, then one
// correctly rounded division). They are evaluated here with the same rounding instead of
// stored, which keeps the bisection of env_sample free of dependent loads. The oracle
// builds the tables the reference's way, and the render parity tests compare.
PT_HD float env_pdf1(const EnvDev&, int) { return 1.0f; }
PT_HD float env_cdf1(const EnvDev& e, int j) { return (float)(j + 1) / (float)e.h; }

// Environment::ImportanceSampling (Environment.cpp:130-220). m_cdf2D / m_pdf2D are
// never-loaded textures, so their fetches go through invalid_tex_bilinear_x.
PT_HD V3 env_sample(const EnvDev& e, Samplers& s, float* pdf_xy)
{
    if (!env_use_is(e)) return v3(1.0f, 0.0f, 0.0f);
    const float texel_w = 1.0f / e.w;
    const float texel_h = 1.0f / e.h;
    float r0, r1;
    next2(s, &r0, &r1);
    float v0 = 0.0f, v1 = 1.0f;
    for (;;) {
        const float v = 0.5f * (v0 + v1);
        const float y = v * e.h;
        const int fy = (int)y;
        const int cy = fy + 1 < e.h - 1 ? fy + 1 : e.h - 1;
        const float c = lerpf(env_cdf1(e, fy), env_cdf1(e, cy), y - fy);
        const float diff = c - r0;
        if (v1 - v0 <= texel_h) {
            v0 = v1 = v;
            break;
        }
        if (diff > 0) v1 = v;
        else if (diff < 0) v0 = v;
        else {
            v0 = v1 = v;
            break;
        }
    }
    float u0 = 0.0f, u1 = 1.0f;
    for (;;) {
        const float u = 0.5f * (u0 + u1);
        const float uu = fmodp(u, 1.0f), vv = fmodp(v0, 1.0f);
        const float c = invalid_tex_bilinear_x(uu * e.w, vv * e.h);
        const float diff = c - r1;
        if (u1 - u0 <= texel_w) {
            u0 = u1 = u;
            break;
        }
        if (diff >= 0) u1 = u;
        else u0 = u;
    }
    {
        const float y = v0 * e.h;
        const int y0 = (int)y;
        const int y1 = y0 + 1 < e.h - 1 ? y0 + 1 : e.h - 1;
        const float py = lerpf(env_pdf1(e, y0), env_pdf1(e, y1), y - y0);
        const float uu = fmodp(u0, 1.0f), vv = fmodp(v0, 1.0f);
        const float px = invalid_tex_bilinear_x(uu * e.w, vv * e.h);
        *pdf_xy = py * px;
    }
    const float theta = v0 * PT_PI;
    const float phi = 2.0f * PT_PI * (1.0f - u0) + e.rot;
    // OrthonormalBasis::Build(UnitZ, UnitX) + GetDir (OrthonormalBasis.cpp:29-43)
    const V3 bw = normalized(v3(0.0f, 0.0f, 1.0f));
    const V3 bu = normalized(v3(1.0f, 0.0f, 0.0f));
    const V3 bv = cross(bw, bu);
    double st, ct, sp, cp;
    pt_sincos((double)theta, &st, &ct);
    pt_sincos((double)phi, &sp, &cp);
    const float cx = (float)(st * cp);
    const float cy = (float)(st * sp);
    const float cz = (float)ct;
    return cx * bu + cy * bv + cz * bw;
}

// Environment::GetImportanceSamplingPDF (Environment.cpp:222-244) with importance
// sampling on, from the angles
PT_HD float env_pdf_at(const EnvDev& e, EnvAngles a)
{
    const float theta = a.theta, phi = a.phi;
    const float u0 = 1.0f - (phi - e.rot) * 0.5f * PT_INV_PI;
    const float v0 = theta * PT_INV_PI;
    const float y = v0 * e.h;
    const int y0 = (int)y;
    const int y1 = y0 + 1 < e.h - 1 ? y0 + 1 : e.h - 1;
    const float py = lerpf(env_pdf1(e, y0), env_pdf1(e, y1), y - y0);
    const float uu = fmodp(u0, 1.0f), vv = fmodp(v0, 1.0f);
    const float px = invalid_tex_bilinear_x(uu * e.w, vv * e.h);
    return py * px;
}

PT_HD float env_pdf(const EnvDev& e, V3 dir)
{
    if (!env_use_is(e)) return 1.0f;
    return env_pdf_at(e, env_angles(dir));
}

// ---- BVH slab test (Core/Accel/BinnedSAHBVH.h:95-170) -------------------------------
// Box given as lower/upper; inv = 1/dir, sgn[i] = inv[i] < 0. Returns the node distance
// (tMin if 0 <= tMin, else tMax if 0 <= tMax) or false on a miss.
PT_HD bool slab(V3 o, V3 inv, int sx, int sy, int sz, V3 lo, V3 hi, float* dist)
{
    float tmin = ((sx ? hi.x : lo.x) - o.x) * inv.x;
    float tmax = ((sx ? lo.x : hi.x) - o.x) * inv.x;
    const float tymin = ((sy ? hi.y : lo.y) - o.y) * inv.y;
    const float tymax = ((sy ? lo.y : hi.y) - o.y) * inv.y;
    if (tymax < tmin || tmax < tymin) return false;
    if (tmin < tymin) tmin = tymin;
    if (tymax < tmax) tmax = tymax;
    const float tzmin = ((sz ? hi.z : lo.z) - o.z) * inv.z;
    const float tzmax = ((sz ? lo.z : hi.z) - o.z) * inv.z;
    if (tzmax < tmin || tmax < tzmin) return false;
    if (tmin < tzmin) tmin = tzmin;
    if (tzmax < tmax) tmax = tzmax;
    float d = FLT_MAX;
    const bool near_ok = (0.0f <= tmin && tmin < d);
    if (near_ok) d = tmin;
    const bool far_ok = (0.0f <= tmax && tmax < d);
    if (far_ok) d = tmax;
    if (!(near_ok || far_ok)) return false;
    *dist = d;
    return true;
}

// The same test without NaN: when lo <= hi on every axis, inv is finite and nonzero and o
// is finite, no product is NaN; (lo - o) * inv and (hi - o) * inv are then ordered by the
// sign of inv (rounding is monotone), so the near plane's distance is their minimum and
// the far plane's their maximum, and the reference's sequence of axis checks and updates
// reduces to max3 of the near distances against min3 of the far ones. Returns the same
// hit flag and distance as slab().
PT_HD bool slab_ordered(V3 o, V3 inv, V3 lo, V3 hi, float* dist)
{
    const float lx = (lo.x - o.x) * inv.x, hx = (hi.x - o.x) * inv.x;
    const float ly = (lo.y - o.y) * inv.y, hy = (hi.y - o.y) * inv.y;
    const float lz = (lo.z - o.z) * inv.z, hz = (hi.z - o.z) * inv.z;
    const float tmin = fmaxf(fmaxf(fminf(lx, hx), fminf(ly, hy)), fminf(lz, hz));
    const float tmax = fminf(fminf(fmaxf(lx, hx), fmaxf(ly, hy)), fmaxf(lz, hz));
    float d = FLT_MAX;
    const bool near_ok = (0.0f <= tmin && tmin < d);
    if (near_ok) d = tmin;
    const bool far_ok = (0.0f <= tmax && tmax < d);
    if (far_ok) d = tmax;
    *dist = d;
    return tmin <= tmax && (near_ok || far_ok);
}

// Child / stack references of the traversal: bits 0-26 node index or first leaf slot,
// bits 27-30 leaf size (0 = inner node, 15 = big leaf: bits 0-26 index the big-leaf
// table), bit 31 = the ray origin is strictly inside the box (stack entries only).
#define PT_REF_INDEX_MASK 0x07FFFFFFu
#define PT_REF_MAX_SMALL_LEAF 14
#define PT_REF_BIG 15u
#define PT_REF_INSIDE 0x80000000u
#define PT_REF_EMPTY 0xFFFFFFFFu
PT_HD uint32_t pt_ref_inner(uint32_t idx) { return idx; }
PT_HD uint32_t pt_ref_leaf(uint32_t first, uint32_t cnt) { return first | (cnt << 27); }
PT_HD uint32_t pt_ref_big_leaf(uint32_t i) { return i | (PT_REF_BIG << 27); }
PT_HD uint32_t pt_ref_cnt(uint32_t r) { return (r >> 27) & 0xFu; }
PT_HD uint32_t pt_ref_index(uint32_t r) { return r & PT_REF_INDEX_MASK; }

// AABB::Contains (AABB.h:58-68): strict inequalities.
PT_HD bool box_contains(V3 lo, V3 hi, V3 p)
{
    return lo.x < p.x && p.x < hi.x && lo.y < p.y && p.y < hi.y && lo.z < p.z && p.z < hi.z;
}

} // namespace pt
