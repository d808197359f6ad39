// render_api.cpp — drives the C++ host API (include/petrichor/Petrichor.h) the way host
// code written against LibPetrichor Core does: vertices, triangles and materials into a
// Scene, a Camera set up with the reference's calls, a target Texture2D, then
// Petrichor::Render. Used by the tests to compare with the Python binding.
//
//   render_api IN OUT [W H SPP DEPTH INTEGRATOR SEED [SAMPLER_LENGTH]]   (no size: camera
//   state only; SAMPLER_LENGTH > 0: Petrichor::SetMicroJitteredSampler)
// IN:  int32 n, int32 n_lights, float pos[9 n], int32 material[n] (0 white, 1 red,
//      2 green, 3 light, 4 GGX), int32 lights[n_lights]; optionally float uv[6 n], int32
//      tw, th, float texture[th][tw][3]: the vertices' uv, and the texture as the white
//      Lambert's albedo texture and the GGX's F0 texture and roughness map; optionally
//      then float normal_strength, roughness_strength: the texture is also the GGX's
//      normal map, with those map strengths
// OUT: float camera[16], then W*H*3 floats each of the image, the denoising albedo and
//      normal AOVs (numSppForDenoising = SPP) and the UV AOV
#include <cstdio>
#include <cstdlib>
#include <vector>

#include "petrichor/Petrichor.h"

using namespace Petrichor;
using namespace Petrichor::Core;
using Math::Vector3f;

int main(int argc, char** argv)
{
    if (argc != 3 && argc != 9 && argc != 10) {
        std::fprintf(stderr, "usage: render_api IN OUT [W H SPP DEPTH INTEGRATOR SEED [SAMPLER_LENGTH]]\n");
        return 2;
    }
    FILE* in = std::fopen(argv[1], "rb");
    if (!in) return 2;
    int32_t n = 0, nl = 0;
    bool ok = std::fread(&n, 4, 1, in) == 1 && std::fread(&nl, 4, 1, in) == 1;
    std::vector<float> pos(9 * (size_t)n);
    std::vector<int32_t> mat(n), lights(nl);
    ok = ok && std::fread(pos.data(), 4, pos.size(), in) == pos.size();
    ok = ok && std::fread(mat.data(), 4, mat.size(), in) == mat.size();
    ok = ok && std::fread(lights.data(), 4, lights.size(), in) == lights.size();
    std::vector<float> uv(6 * (size_t)n), texels;
    int32_t tsize[2] = {0, 0};
    const bool textured = ok && std::fread(uv.data(), 4, uv.size(), in) == uv.size() && n > 0;
    if (textured) {
        ok = std::fread(tsize, 4, 2, in) == 2 && tsize[0] > 0 && tsize[1] > 0;
        texels.resize(ok ? 3 * (size_t)tsize[0] * tsize[1] : 0);
        ok = ok && std::fread(texels.data(), 4, texels.size(), in) == texels.size();
    }
    float strengths[2] = {1.0f, 1.0f};
    const bool normal_mapped = textured && ok && std::fread(strengths, 4, 2, in) == 2;
    std::fclose(in);
    if (!ok) return 2;
    Texture2D tex(textured ? tsize[0] : 0, textured ? tsize[1] : 0);
    for (int j = 0; j < tsize[1]; ++j)
        for (int i = 0; i < tsize[0]; ++i) {
            const float* c = &texels[3 * ((size_t)j * tsize[0] + i)];
            tex.SetPixel(i, j, Color3f(c[0], c[1], c[2]));
        }

    Scene scene;
    auto white = std::make_unique<Lambert>(Color3f(0.73f, 0.73f, 0.73f));
    auto ggx = std::make_unique<GGX>(Color3f(0.9f, 0.8f, 0.7f), 0.3f);
    if (textured) {
        white->SetTexAlbedo(&tex);
        ggx->SetF0Texture(&tex);
        ggx->SetRoughnessMap(&tex);
    }
    if (normal_mapped) {
        ggx->SetNormalMap(&tex);
        ggx->SetNormalMapStrength(strengths[0]);
        ggx->SetRoughnessMapStrength(strengths[1]);
    }
    scene.RegisterMaterial("white", std::move(white));
    scene.RegisterMaterial("red", std::make_unique<Lambert>(Color3f(0.65f, 0.05f, 0.05f)));
    scene.RegisterMaterial("green", std::make_unique<Lambert>(Color3f(0.05f, 0.65f, 0.05f)));
    scene.RegisterMaterial("light", std::make_unique<Emission>(Color3f(15.0f, 15.0f, 15.0f)));
    scene.RegisterMaterial("ggx", std::move(ggx));
    const MaterialBase* mats[5] = {scene.GetMaterial("white"), scene.GetMaterial("red"), scene.GetMaterial("green"),
                                   scene.GetMaterial("light"), scene.GetMaterial("ggx")};
    Mesh mesh(3 * (size_t)n);
    for (int i = 0; i < n; ++i) {
        const float* p = &pos[9 * (size_t)i];
        const float* tc = &uv[6 * (size_t)i];
        const Vertex* v0 = mesh.AddVertex(Vertex(Vector3f(p[0], p[1], p[2]), Vector3f::UnitZ(), Vector3f(tc[0], tc[1], 0.0f)));
        const Vertex* v1 = mesh.AddVertex(Vertex(Vector3f(p[3], p[4], p[5]), Vector3f::UnitZ(), Vector3f(tc[2], tc[3], 0.0f)));
        const Vertex* v2 = mesh.AddVertex(Vertex(Vector3f(p[6], p[7], p[8]), Vector3f::UnitZ(), Vector3f(tc[4], tc[5], 0.0f)));
        Triangle t(v0, v1, v2);
        t.SetMaterial(mats[mat[i]]);
        mesh.AddTriangle(t);
    }
    scene.AppendMesh(mesh);
    for (int32_t l : lights) scene.AppendLight(&mesh.GetTriangles()[l]);

    // the Cornell box camera: Camera(pos, dir), SetFocalLength, FocusTo, SetFNumber
    const Vector3f eye(0.5f, -1.35f, 0.5f);
    auto camera = std::make_unique<Camera>(eye, Vector3f(0.5f, 0.5f, 0.5f) - eye);
    camera->SetFocalLength(35e-3f);
    camera->FocusTo(Vector3f(0.5f, 0.5f, 0.3f));
    camera->SetFNumber(8.0f);
    float state[16];
    camera->CopyState(state);
    scene.SetMainCamera(std::move(camera));

    FILE* out = std::fopen(argv[2], "wb");
    if (!out) return 2;
    std::fwrite(state, 4, 16, out);
    int rc = 0;
    if (argc >= 9) {
        const int w = std::atoi(argv[3]), h = std::atoi(argv[4]);
        Texture2D target(w, h), albedo(w, h), normal(w, h), uvtex(w, h);
        scene.SetTargetTexture(Scene::AOVType::Rendered, &target);
        scene.SetTargetTexture(Scene::AOVType::DenoisingAlbedo, &albedo);
        scene.SetTargetTexture(Scene::AOVType::DenoisingNormal, &normal);
        scene.SetTargetTexture(Scene::AOVType::UV, &uvtex);
        RenderSetting rs;
        rs.numSamplesPerPixel = std::atoi(argv[5]);
        rs.numSppForDenoising = std::atoi(argv[5]);
        rs.numMaxBounces = std::atoi(argv[6]);
        scene.SetRenderSetting(rs);
        Core::Petrichor renderer;
        renderer.SetIntegrator(std::atoi(argv[7]) == 0 ? Integrator::PathTracing : Integrator::SimplePathTracing);
        renderer.SetSeed((uint32_t)std::atoi(argv[8]));
        if (argc == 10) renderer.SetMicroJitteredSampler(std::atoi(argv[9]));
        float seconds = -1.0f;
        renderer.SetRenderCallback([&](const RenderingResult& r) { seconds = r.totalSec; });
        if (!renderer.Render(scene)) {
            std::fprintf(stderr, "render failed: %s\n", renderer.GetLastError().c_str());
            rc = 1;
        } else {
            std::fwrite(target.GetRawDataPtr(), 4, 3 * (size_t)w * h, out);
            std::fwrite(albedo.GetRawDataPtr(), 4, 3 * (size_t)w * h, out);
            std::fwrite(normal.GetRawDataPtr(), 4, 3 * (size_t)w * h, out);
            std::fwrite(uvtex.GetRawDataPtr(), 4, 3 * (size_t)w * h, out);
            // Texture2D::Save of the same image, next to the output file
            const std::string base = argv[2];
            if (!target.Save(base + ".png") || !target.Save(base + ".hdr")) {
                std::fprintf(stderr, "Texture2D::Save failed\n");
                rc = 1;
            }
            std::printf("rendered %dx%d in %.3f s, %u of %u tiles\n", w, h, seconds, renderer.GetNumRenderedTiles(),
                        renderer.GetNumTiles());
        }
    }
    std::fclose(out);
    return rc;
}
