// load_api.cpp — host code of the shipped application's kind: a scene and its render
// settings read from the reference's files (Scene::LoadRenderSetting, Scene::LoadAssets),
// rendered by Petrichor::Render. Used by the tests to compare with the Python loader.
//
//   load_api SCENE.json SETTINGS.json OUT INTEGRATOR SEED
// OUT: float camera[16], then W*H*3 floats of the image (W, H from the settings)
#include <cstdio>
#include <cstdlib>
#include <string>

#include "petrichor/Petrichor.h"

using namespace Petrichor;
using namespace Petrichor::Core;

int main(int argc, char** argv)
{
    if (argc != 6) {
        std::fprintf(stderr, "usage: load_api SCENE.json SETTINGS.json OUT INTEGRATOR SEED\n");
        return 2;
    }
    Scene scene;
    if (!scene.LoadRenderSetting(argv[2])) {
        std::fprintf(stderr, "cannot read %s\n", argv[2]);
        return 1;
    }
    std::string error;
    if (!scene.LoadAssets(argv[1], &error)) {
        std::fprintf(stderr, "LoadAssets: %s\n", error.c_str());
        return 1;
    }
    const RenderSetting& rs = scene.GetRenderSetting();
    Texture2D target(rs.outputWidth, rs.outputHeight);
    scene.SetTargetTexture(Scene::AOVType::Rendered, &target);
    Core::Petrichor renderer;
    renderer.SetIntegrator(std::atoi(argv[4]) == 0 ? Integrator::PathTracing : Integrator::SimplePathTracing);
    renderer.SetSeed((uint32_t)std::atoi(argv[5]));
    if (!renderer.Render(scene)) {
        std::fprintf(stderr, "render failed: %s\n", renderer.GetLastError().c_str());
        return 1;
    }
    float state[16];
    scene.GetMainCamera()->CopyState(state);
    FILE* out = std::fopen(argv[3], "wb");
    if (!out) return 2;
    std::fwrite(state, 4, 16, out);
    std::fwrite(target.GetRawDataPtr(), 4, 3 * (size_t)rs.outputWidth * rs.outputHeight, out);
    std::fclose(out);
    std::printf("%zu triangles rendered at %dx%d\n", scene.GetGeometries().size(), rs.outputWidth, rs.outputHeight);
    return 0;
}
