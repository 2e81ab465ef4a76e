// A caller written against the reference's class interface (src/source.cu:7-108 drives the
// same objects in the same order), compiled against include/cupbr.hpp and libcupbr.
//
//   render_example <scene.xml> <out.bmp> [--load-only]
//
// --load-only stops after the host-side scene load (runs without a GPU).
#include <cmath>
#include <cstdio>
#include <cstring>

#include "cupbr.hpp"

int main(int argc, char** argv) {
    if (argc < 3) {
        std::fprintf(stderr, "usage: %s scene.xml out.bmp [--load-only]\n", argv[0]);
        return 2;
    }
    cupbr::Scene scene = cupbr::SceneLoader::loadFromFile(argv[1]);
    if (!scene.handle) return 1;
    uint32_t n_objects = 0, n_triangles = 0, n_lights = 0;
    int32_t env = 0;
    cupbr_scene_get_counts(scene.handle, &n_objects, &n_triangles, &n_lights, &env);
    std::printf("scene: %u objects, %u triangles, %u lights, environment %d\n", n_objects, n_triangles, n_lights, env);
    if (argc > 3 && std::strcmp(argv[3], "--load-only") == 0) {
        cupbr::SceneLoader::destroyScene(scene);
        return 0;
    }

    const uint32_t width = 96, height = 64;
    cupbr::Camera camera(width, height);
    int rc = 0;
    {
        cupbr::PBRenderer renderer(cupbr::RenderingMethod::PATHTRACER);
        rc |= renderer.setOutputSize(width, height);
        rc |= renderer.registerScene(&scene);
        rc |= renderer.setMaxTraceDepth(5);
        for (int frame = 0; frame < 16; ++frame) rc |= renderer.render(&camera);

        cupbr::PostProcessor post;
        rc |= post.registerImage(renderer.getOutputImage());
        rc |= post.bloom(cupbr::computeThreshold(1.0f, 1.0f));
        cupbr::ToneMapper mapper(cupbr::ToneMappingType::REINHARD);
        rc |= mapper.registerImage(post.getPostProcessBuffer());
        rc |= mapper.toneMap();
        rc |= mapper.saveToFile(argv[2]);

        std::vector<float> hdr = renderer.toHost();
        double sum = 0;
        bool finite = true;
        for (float v : hdr) {
            sum += v;
            finite = finite && std::isfinite(v);
        }
        std::printf("frames 16, method %d, mean %.6f, finite %d, status %d\n", (int)renderer.getMethod(),
                    sum / (double)hdr.size(), (int)finite, rc);
        if (!finite) rc |= 1;
    }
    cupbr::SceneLoader::destroyScene(scene);
    return rc == 0 ? 0 : 1;
}
