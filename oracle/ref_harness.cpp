// ref_harness.cpp — runs the reference's own hot-path translation units on seeded
// inputs and writes inputs and outputs as raw little-endian arrays.
//
// TEST INFRASTRUCTURE ONLY (see oracle/Makefile). It is compiled against the
// reference sources where they lie (LibPetrichor/Math, Core/Geometry/Triangle,
// Core/Camera, Core/Material/{Lambert,GGX,Glass}, Core/Sampler, Math/Halton) — the units that
// build from their own sources with no third-party header. Units that need spdlog,
// fmt or stb (BinnedSAHBVH, Environment, Texture2D, the integrators) are not built.
// Private members are read through `#define private public`, applied after every
// standard header has been included.
#include <algorithm>
#include <array>
#include <cassert>
#include <cfloat>
#include <climits>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <filesystem>
#include <iostream>
#include <limits>
#include <memory>
#include <optional>
#include <random>
#include <string>
#include <tuple>
#include <vector>

#define private public
#define protected public
#include "Core/Camera.h"
#include "Core/Geometry/Triangle.h"
#include "Core/Geometry/Vertex.h"
#include "Core/HitInfo.h"
#include "Core/Material/GGX.h"
#include "Core/Material/Glass.h"
#include "Core/Material/Lambert.h"
#include "Core/Ray.h"
#include "Core/Sampler/MicroJitteredSampler.h"
#include "Core/Sampler/RandomSampler2D.h"
#include "Math/Halton.h"
#include "Math/OrthonormalBasis.h"
#include "Random/XorShift.h"
#undef private
#undef protected

using namespace Petrichor;
using namespace Petrichor::Core;
using Math::Vector3f;

namespace {

std::string g_dir;

template <typename T>
void dump(const char* name, const std::vector<T>& v)
{
    const std::string path = g_dir + "/" + name;
    FILE* f = std::fopen(path.c_str(), "wb");
    if (!f) {
        std::perror(path.c_str());
        std::exit(1);
    }
    std::fwrite(v.data(), sizeof(T), v.size(), f);
    std::fclose(f);
}

// Input generator: splitmix-style 64-bit LCG, independent of the code under test.
struct Gen {
    uint64_t s;
    explicit Gen(uint64_t seed) : s(seed) {}
    uint32_t u32()
    {
        s = s * 6364136223846793005ull + 1442695040888963407ull;
        return (uint32_t)(s >> 32);
    }
    float uni(float lo, float hi) { return lo + (hi - lo) * ((u32() >> 8) * (1.0f / 16777216.0f)); }
    Vector3f vec(float lo, float hi) { return Vector3f(uni(lo, hi), uni(lo, hi), uni(lo, hi)); }
    Vector3f unit()
    {
        for (;;) {
            Vector3f v = vec(-1.0f, 1.0f);
            const float l2 = v.SquaredLength();
            if (l2 > 1e-4f && l2 <= 1.0f) return v.Normalized();
        }
    }
};

void push3(std::vector<float>& o, const Vector3f& v)
{
    o.push_back(v.x);
    o.push_back(v.y);
    o.push_back(v.z);
}

void case_xorshift()
{
    std::vector<uint32_t> seeds = {0u, 1u, 2u, 123u, 0xffffffffu, 0x80000000u, 486213789u};
    Gen g(1);
    while (seeds.size() < 32) seeds.push_back(g.u32());
    std::vector<uint32_t> u;
    std::vector<float> f;
    for (uint32_t s : seeds) {
        Math::XorShift128 a(s), b(s);
        for (int i = 0; i < 100; ++i) {
            u.push_back(a.random());
            f.push_back(b.next());
        }
    }
    dump("xorshift_seeds.u32", seeds);
    dump("xorshift_u32.u32", u);
    dump("xorshift_f32.f32", f);
}

void case_triangle()
{
    Gen g(2);
    const int n = 20000;
    std::vector<float> tri, ray, t;
    std::vector<int32_t> hit;
    for (int i = 0; i < n; ++i) {
        const Vertex v0(g.uni(-1, 1), g.uni(-1, 1), g.uni(-1, 1));
        const Vertex v1(g.uni(-1, 1), g.uni(-1, 1), g.uni(-1, 1));
        const Vertex v2(g.uni(-1, 1), g.uni(-1, 1), g.uni(-1, 1));
        const Triangle tr(&v0, &v1, &v2);
        // aim most rays at a point inside the triangle's bounding region
        const Vector3f o = g.vec(-2.0f, 2.0f);
        const Vector3f target = (v0.pos + v1.pos + v2.pos) / 3.0f + g.vec(-0.3f, 0.3f);
        Vector3f d = (i % 7 == 0) ? g.unit() : (target - o).Normalized();
        if (i % 50 == 1) d = Vector3f(0.0f, 0.0f, (i % 100 == 1) ? 1.0f : -1.0f);
        const Ray r(o, d);
        const auto h = tr.Intersect(r);
        push3(tri, v0.pos);
        push3(tri, v1.pos);
        push3(tri, v2.pos);
        push3(ray, o);
        push3(ray, d);
        hit.push_back(h ? 1 : 0);
        t.push_back(h ? h->distance : 0.0f);
    }
    dump("tri_tri.f32", tri);
    dump("tri_ray.f32", ray);
    dump("tri_hit.i32", hit);
    dump("tri_t.f32", t);

    // SampleSurface
    std::vector<float> stri, sout;
    std::vector<uint32_t> sseed;
    for (int i = 0; i < 2000; ++i) {
        const Vertex v0(g.uni(-1, 1), g.uni(-1, 1), g.uni(-1, 1));
        const Vertex v1(g.uni(-1, 1), g.uni(-1, 1), g.uni(-1, 1));
        const Vertex v2(g.uni(-1, 1), g.uni(-1, 1), g.uni(-1, 1));
        const Triangle tr(&v0, &v1, &v2);
        const uint32_t s0 = g.u32(), s1 = g.u32();
        RandomSampler2D smp(s0, s1);
        PointData pd;
        float pdf = 0.0f;
        tr.SampleSurface(Vector3f(), smp, &pd, &pdf);
        push3(stri, v0.pos);
        push3(stri, v1.pos);
        push3(stri, v2.pos);
        sseed.push_back(s0);
        sseed.push_back(s1);
        push3(sout, pd.pos);
        push3(sout, pd.normal);
        sout.push_back(pdf);
    }
    dump("trisample_tri.f32", stri);
    dump("trisample_seed.u32", sseed);
    dump("trisample_out.f32", sout);
}

void case_camera()
{
    Gen g(3);
    std::vector<float> cam, out;
    std::vector<int32_t> pix;
    std::vector<uint32_t> seeds;
    for (int c = 0; c < 200; ++c) {
        const Vector3f pos = g.vec(-5.0f, 5.0f);
        Vector3f dir = g.unit();
        if (std::fabs(dir.z) > 0.95f) dir = Vector3f(dir.x, dir.y, 0.3f);
        Camera camera(pos, dir);
        camera.SetFocalLength(g.uni(18e-3f, 100e-3f));
        camera.FocusTo(pos + dir * g.uni(0.5f, 10.0f));
        if (c % 4 == 0) camera.SetApeture(0.0f);
        else if (c % 4 == 1) camera.SetFNumber(g.uni(1.4f, 16.0f));
        const float p[16] = {camera.m_pos.x, camera.m_pos.y, camera.m_pos.z,
                             camera.m_right.x, camera.m_right.y, camera.m_right.z,
                             camera.m_up.x, camera.m_up.y, camera.m_up.z,
                             camera.m_forward.x, camera.m_forward.y, camera.m_forward.z,
                             camera.m_focusDist, camera.m_hPerf, camera.m_apeture, 0.0f};
        for (int s = 0; s < 20; ++s) {
            const int w = (s % 2) ? 1920 : 512, h = (s % 2) ? 1080 : 512;
            const int i = (int)(g.u32() % (uint32_t)w), j = (int)(g.u32() % (uint32_t)h);
            const uint32_t s0 = g.u32(), s1 = g.u32();
            RandomSampler2D smp(s0, s1);
            const Ray r = camera.GenerateRay(i, j, w, h, smp);
            cam.insert(cam.end(), p, p + 16);
            pix.push_back(i);
            pix.push_back(j);
            pix.push_back(w);
            pix.push_back(h);
            seeds.push_back(s0);
            seeds.push_back(s1);
            push3(out, r.o);
            push3(out, r.dir);
            push3(out, r.throughput);
        }
    }
    dump("camera_cam.f32", cam);
    dump("camera_pix.i32", pix);
    dump("camera_seed.u32", seeds);
    dump("camera_out.f32", out);
}

// mat_param layout of the oracle: color[3], p3, p4, p5, is_enabled, 0
void case_materials()
{
    Gen g(4);
    std::vector<int32_t> type;
    std::vector<float> param, shading, rin, rout, dirs, evals;
    std::vector<uint32_t> seeds;
    for (int i = 0; i < 8000; ++i) {
        const int kind = i % 4; // 0 Lambert IS, 1 Lambert uniform, 2 GGX, 3 Glass
        std::unique_ptr<MaterialBase> m;
        float pr[8] = {0, 0, 0, 0, 0, 0, 1, 0};
        const Vector3f col = g.vec(0.05f, 0.95f);
        pr[0] = col.x;
        pr[1] = col.y;
        pr[2] = col.z;
        int t = 0;
        if (kind <= 1) {
            auto l = std::make_unique<Lambert>(col);
            if (kind == 1) {
                l->EnableImportanceSampling(false);
                pr[6] = 0.0f;
            }
            m = std::move(l);
            t = 0;
        } else if (kind == 2) {
            const float rough = g.uni(0.02f, 1.0f);
            pr[3] = rough;
            m = std::make_unique<GGX>(col, rough);
            t = 3;
        } else {
            const float ior = (i % 8 == 3) ? 1.5f : g.uni(1.05f, 2.5f);
            pr[3] = ior;
            m = std::make_unique<Glass>(col, ior);
            t = 1;
        }
        ShadingInfo si;
        si.normal = g.unit();
        si.pos = g.vec(-3.0f, 3.0f);
        si.material = m.get();
        Vector3f d = g.unit();
        Ray in(g.vec(-3.0f, 3.0f), d, RayTypes::Camera, g.vec(0.1f, 2.0f), (int)(g.u32() % 10),
               (i % 3 == 0) ? 0.9f : 1.0f, (kind == 3 && (i % 5 == 0)) ? 1.5f : 1.0f);
        const uint32_t s0 = g.u32(), s1 = g.u32();
        RandomSampler2D smp(s0, s1);
        const Ray o = m->CreateNextRay(in, si, smp);
        type.push_back(t);
        param.insert(param.end(), pr, pr + 8);
        push3(shading, si.pos);
        push3(shading, si.normal);
        push3(rin, in.o);
        push3(rin, in.dir);
        push3(rin, in.throughput);
        rin.push_back((float)in.bounce);
        rin.push_back(in.prob);
        rin.push_back(in.ior);
        seeds.push_back(s0);
        seeds.push_back(s1);
        push3(rout, o.o);
        push3(rout, o.dir);
        push3(rout, o.throughput);
        rout.push_back((float)o.bounce);
        rout.push_back(o.prob);
        rout.push_back(o.ior);
        // BxDF / PDF for an independent outgoing direction
        Ray out_ray(si.pos, g.unit());
        const Color3f f = m->BxDF(in, out_ray, si);
        const float p = m->PDF(in, out_ray, si);
        push3(dirs, out_ray.dir);
        push3(evals, f);
        evals.push_back(p);
    }
    dump("mat_type.i32", type);
    dump("mat_param.f32", param);
    dump("mat_shading.f32", shading);
    dump("mat_ray_in.f32", rin);
    dump("mat_seed.u32", seeds);
    dump("mat_ray_out.f32", rout);
    dump("mat_dir_out.f32", dirs);
    dump("mat_eval.f32", evals);
}

void case_onb()
{
    Gen g(5);
    std::vector<float> in, out;
    for (int i = 0; i < 5000; ++i) {
        const Vector3f w = g.vec(-2.0f, 2.0f);
        const float theta = g.uni(0.0f, 3.1415927f);
        const float phi = g.uni(-7.0f, 13.0f);
        Math::OrthonormalBasis onb;
        onb.Build(w);
        const Vector3f d = onb.GetDir(theta, phi);
        push3(in, w);
        in.push_back(theta);
        in.push_back(phi);
        push3(out, onb.GetU());
        push3(out, onb.GetV());
        push3(out, onb.GetW());
        push3(out, d);
    }
    dump("onb_in.f32", in);
    dump("onb_out.f32", out);
}

// Math::Halton tables and MicroJitteredSampler sequences: 600 draws (several sequence
// wraps for the short lengths) per (seed, length).
void case_microjittered()
{
    std::vector<float> halton;
    for (unsigned base : {2u, 3u, 5u}) {
        Math::Halton h;
        h.Initialize(1000, base);
        for (unsigned i = 0; i < 1000; ++i) halton.push_back(h.GetValue(i));
    }
    dump("halton.f32", halton);
    std::vector<uint32_t> cfg;
    std::vector<float> out;
    for (unsigned seed : {1u, 7u, 12345u, 0xdeadbeefu})
        for (int len : {1, 2, 3, 16, 100, 256}) {
            MicroJitteredSampler smp(seed);
            smp.Initialize(len);
            cfg.push_back(seed);
            cfg.push_back((uint32_t)len);
            for (int i = 0; i < 600; ++i) {
                const auto v = smp.Next();
                out.push_back(std::get<0>(v));
                out.push_back(std::get<1>(v));
            }
        }
    dump("mj_cfg.u32", cfg);
    dump("mj_out.f32", out);
}

} // namespace

int main(int argc, char** argv)
{
    if (argc != 2) {
        std::fprintf(stderr, "usage: %s OUT_DIR\n", argv[0]);
        return 2;
    }
    g_dir = argv[1];
    std::filesystem::create_directories(g_dir);
    case_xorshift();
    case_triangle();
    case_camera();
    case_materials();
    case_onb();
    case_microjittered();
    std::printf("ok\n");
    return 0;
}
