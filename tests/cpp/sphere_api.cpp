// sphere_api.cpp — a scene of triangles and Sphere geometries through the C++ host API
// (Scene::AppendGeometry / AppendLight of a Sphere, Petrichor::Render), as host code
// written against LibPetrichor Core builds one. Used by the tests to compare with the
// Python binding.
//
//   sphere_api IN OUT W H SPP DEPTH INTEGRATOR SEED
// IN:  int32 n_tris, n_spheres, n_lights, float pos[9 n_tris], int32 tri_material[n_tris],
//      float sphere[4 n_spheres] (centre, radius), int32 sphere_material[n_spheres],
//      int32 lights[n_lights] (triangles first, then spheres); materials 0 white, 1 red,
//      2 green, 3 light (Lambert 0.73 / 0.65,0.05,0.05 / 0.05,0.65,0.05, Emission 15),
//      4 GGX (0.9, 0.8, 0.7; roughness 0.3), 5 glass (ior 1.5)
// OUT: W*H*3 floats of the image
#include <cstdio>
#include <cstdlib>
#include <vector>

#include "petrichor/Petrichor.h"

using namespace Petrichor;
using namespace Petrichor::Core;
using Math::Vector3f;

int main(int argc, char** argv)
{
    if (argc != 9) {
        std::fprintf(stderr, "usage: sphere_api IN OUT W H SPP DEPTH INTEGRATOR SEED\n");
        return 2;
    }
    FILE* in = std::fopen(argv[1], "rb");
    if (!in) return 2;
    int32_t hdr[3];
    bool ok = std::fread(hdr, 4, 3, in) == 3;
    const int nt = hdr[0], ns = hdr[1], nl = hdr[2];
    std::vector<float> pos(9 * (size_t)nt), sph(4 * (size_t)ns);
    std::vector<int32_t> tmat(nt), smat(ns), lights(nl);
    ok = ok && std::fread(pos.data(), 4, pos.size(), in) == pos.size();
    ok = ok && std::fread(tmat.data(), 4, tmat.size(), in) == tmat.size();
    ok = ok && std::fread(sph.data(), 4, sph.size(), in) == sph.size();
    ok = ok && std::fread(smat.data(), 4, smat.size(), in) == smat.size();
    ok = ok && std::fread(lights.data(), 4, lights.size(), in) == lights.size();
    std::fclose(in);
    if (!ok) return 2;

    Scene scene;
    scene.RegisterMaterial("white", std::make_unique<Lambert>(Color3f(0.73f, 0.73f, 0.73f)));
    scene.RegisterMaterial("red", std::make_unique<Lambert>(Color3f(0.65f, 0.05f, 0.05f)));
    scene.RegisterMaterial("green", std::make_unique<Lambert>(Color3f(0.05f, 0.65f, 0.05f)));
    scene.RegisterMaterial("light", std::make_unique<Emission>(Color3f(15.0f, 15.0f, 15.0f)));
    scene.RegisterMaterial("ggx", std::make_unique<GGX>(Color3f(0.9f, 0.8f, 0.7f), 0.3f));
    scene.RegisterMaterial("glass", std::make_unique<Glass>(Color3f(1.0f, 1.0f, 1.0f), 1.5f));
    const char* names[6] = {"white", "red", "green", "light", "ggx", "glass"};
    Mesh mesh(3 * (size_t)nt);
    for (int i = 0; i < nt; ++i) {
        const float* p = &pos[9 * (size_t)i];
        const Vertex* v0 = mesh.AddVertex(Vertex(p[0], p[1], p[2]));
        const Vertex* v1 = mesh.AddVertex(Vertex(p[3], p[4], p[5]));
        const Vertex* v2 = mesh.AddVertex(Vertex(p[6], p[7], p[8]));
        Triangle t(v0, v1, v2);
        t.SetMaterial(scene.GetMaterial(names[tmat[i]]));
        mesh.AddTriangle(t);
    }
    scene.AppendMesh(mesh);
    std::vector<Sphere> spheres;
    spheres.reserve(ns);
    for (int i = 0; i < ns; ++i) {
        spheres.emplace_back(Vector3f(sph[4 * i], sph[4 * i + 1], sph[4 * i + 2]), sph[4 * i + 3]);
        spheres.back().SetMaterial(scene.GetMaterial(names[smat[i]]));
    }
    for (const Sphere& s : spheres) scene.AppendGeometry(&s);
    for (int32_t l : lights) {
        if (l < nt) scene.AppendLight(&mesh.GetTriangles()[l]);
        else scene.AppendLight(&spheres[l - nt]);
    }
    // the Cornell box camera of render_api.cpp
    const Vector3f eye(0.5f, -1.35f, 0.5f);
    auto camera = std::make_unique<Camera>(eye, Vector3f(0.5f, 0.5f, 0.5f) - eye);
    camera->SetFocalLength(35e-3f);
    camera->FocusTo(Vector3f(0.5f, 0.5f, 0.3f));
    camera->SetFNumber(8.0f);
    scene.SetMainCamera(std::move(camera));
    const int w = std::atoi(argv[3]), h = std::atoi(argv[4]);
    Texture2D target(w, h);
    scene.SetTargetTexture(Scene::AOVType::Rendered, &target);
    RenderSetting rs;
    rs.numSamplesPerPixel = std::atoi(argv[5]);
    rs.numMaxBounces = std::atoi(argv[6]);
    scene.SetRenderSetting(rs);
    Core::Petrichor renderer;
    renderer.SetIntegrator(std::atoi(argv[7]) == 0 ? Integrator::PathTracing : Integrator::SimplePathTracing);
    renderer.SetSeed((uint32_t)std::atoi(argv[8]));
    if (!renderer.Render(scene)) {
        std::fprintf(stderr, "render failed: %s\n", renderer.GetLastError().c_str());
        return 1;
    }
    FILE* out = std::fopen(argv[2], "wb");
    if (!out) return 2;
    std::fwrite(target.GetRawDataPtr(), 4, 3 * (size_t)w * h, out);
    std::fclose(out);
    return 0;
}
