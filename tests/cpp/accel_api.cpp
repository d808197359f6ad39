// accel_api.cpp — the acceleration structure of the C++ host API as host code uses it:
// BinnedSAHBVH::Build over a Scene's triangles and spheres, then AccelBase::Intersect per
// ray (the three overloads) and IntersectBatch. Used by the tests to compare with the
// Python binding's pt_trace_closest.
//
//   accel_api IN OUT [reference]   (reference: BinnedSAHBVH::SetReferenceLeafOrder(true))
// IN:  int32 n_tris, n_spheres, n_rays, float pos[9 n_tris], float sphere[4 n_spheres],
//      float ray[8 n_rays] (origin, dist_min, direction, dist_max)
// OUT: per ray int32 primitive (triangles first, then spheres; -1 on a miss) and float t
//      from IntersectBatch; then the same for the first min(n_rays, 64) rays from
//      Intersect(ray, scene, dmin, dmax), and for ray 0 from Intersect(ray, scene) and
//      Intersect(ray, scene, dmin)
#include <cstdio>
#include <cstdlib>
#include <map>
#include <vector>

#include "petrichor/Petrichor.h"

using namespace Petrichor;
using namespace Petrichor::Core;
using Math::Vector3f;

int main(int argc, char** argv)
{
    if (argc != 3 && argc != 4) {
        std::fprintf(stderr, "usage: accel_api IN OUT [reference]\n");
        return 2;
    }
    FILE* in = std::fopen(argv[1], "rb");
    if (!in) return 2;
    int32_t hdr[3];
    bool ok = std::fread(hdr, 4, 3, in) == 3;
    const int nt = hdr[0], ns = hdr[1], nr = hdr[2];
    std::vector<float> pos(9 * (size_t)nt), sph(4 * (size_t)ns), rays(8 * (size_t)nr);
    ok = ok && std::fread(pos.data(), 4, pos.size(), in) == pos.size();
    ok = ok && std::fread(sph.data(), 4, sph.size(), in) == sph.size();
    ok = ok && std::fread(rays.data(), 4, rays.size(), in) == rays.size();
    std::fclose(in);
    if (!ok) return 2;

    Scene scene;
    Mesh mesh(3 * (size_t)nt);
    for (int i = 0; i < nt; ++i) {
        const float* p = &pos[9 * (size_t)i];
        const Vertex* v0 = mesh.AddVertex(Vertex(p[0], p[1], p[2]));
        const Vertex* v1 = mesh.AddVertex(Vertex(p[3], p[4], p[5]));
        const Vertex* v2 = mesh.AddVertex(Vertex(p[6], p[7], p[8]));
        mesh.AddTriangle(Triangle(v0, v1, v2));
    }
    scene.AppendMesh(mesh);
    std::vector<Sphere> spheres;
    spheres.reserve(ns);
    for (int i = 0; i < ns; ++i) spheres.emplace_back(Vector3f(sph[4 * i], sph[4 * i + 1], sph[4 * i + 2]), sph[4 * i + 3]);
    for (const Sphere& s : spheres) scene.AppendGeometry(&s);
    std::map<const GeometryBase*, int32_t> index; // geometry -> primitive index
    for (int i = 0; i < nt; ++i) index[&mesh.GetTriangles()[i]] = i;
    for (int i = 0; i < ns; ++i) index[&spheres[i]] = nt + i;

    BinnedSAHBVH bvh;
    bvh.SetReferenceLeafOrder(argc == 4);
    bvh.Build(scene);
    if (!bvh.GetLastError().empty()) {
        std::fprintf(stderr, "build failed: %s\n", bvh.GetLastError().c_str());
        return 1;
    }
    std::vector<Ray> rs;
    for (int i = 0; i < nr; ++i) {
        const float* r = &rays[8 * (size_t)i];
        rs.emplace_back(Vector3f(r[0], r[1], r[2]), Vector3f(r[4], r[5], r[6]));
    }
    std::vector<int32_t> prim;
    std::vector<float> t;
    auto put = [&](const std::optional<HitInfo>& h) {
        prim.push_back(h ? index.at(h->hitObj) : -1);
        t.push_back(h ? h->distance : 0.0f);
    };
    const float dmin = nr ? rays[3] : 0.0f, dmax = nr ? rays[7] : 0.0f;
    for (const auto& h : bvh.IntersectBatch(rs, dmin, dmax)) put(h);
    for (int i = 0; i < std::min(nr, 64); ++i) put(bvh.Intersect(rs[i], scene, rays[8 * i + 3], rays[8 * i + 7]));
    if (nr) {
        put(bvh.Intersect(rs[0], scene));
        put(bvh.Intersect(rs[0], scene, 0.5f));
    }
    if (!bvh.GetLastError().empty()) {
        std::fprintf(stderr, "trace failed: %s\n", bvh.GetLastError().c_str());
        return 1;
    }
    FILE* out = std::fopen(argv[2], "wb");
    if (!out) return 2;
    std::fwrite(prim.data(), 4, prim.size(), out);
    std::fwrite(t.data(), 4, t.size(), out);
    std::fclose(out);
    return 0;
}
