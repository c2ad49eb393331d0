// The reference app's MDAS calls (optixMotionBlur.cpp:331-360, 473-491,
// 768-806) compiled against include/mdas_kdgrid.hpp and libmdas.so.  The
// OptiX trace is replaced by the analytic scene over the same buffers.
#include <cstdio>
#include <vector>

#include <hip/hip_runtime.h>

#include "mdas_kdgrid.hpp"

typedef mdas::Point5 Point;

int main() {
    const int width = 160, height = 90, max_samples = 1 << 21;
    const float samples_per_launch = 8.0f, scaleFactor = 1.0f;
    mdas::KDGrid<Point>* kdgrid = new mdas::KDGrid<Point>(
        max_samples, 4, 1, 3, 0.25f, scaleFactor * width, scaleFactor * height, nullptr);
    kdgrid->InitialSampling();
    float* coords = reinterpret_cast<float*>(kdgrid->GetSampleCoordinates().Data());
    (void)coords;
    mdas_scene_handle scene = nullptr;
    if (mdas_scene_create(MDAS_SCENE_SPHERES_MBDOF, 1, Point::DIM, kdgrid->GetScaleX(),
                          kdgrid->GetScaleY(), &scene) != MDAS_OK)
        return 2;
    mdas::check(mdas_eval_integrand(kdgrid->handle(), scene, 0, kdgrid->GetNumberOfSamples(), nullptr));
    const int total_samples = static_cast<int>(samples_per_launch * width * height);
    int passes = 0;
    while (kdgrid->GetNumberOfSamples() < std::min(total_samples, max_samples)) {
        kdgrid->SamplingPass();
        kdgrid->Validate();
        mdas::check(mdas_eval_new_samples(kdgrid->handle(), scene, nullptr));
        ++passes;
    }
    void* pixels = nullptr;
    void* bytes = nullptr;
    if (hipMalloc(&pixels, sizeof(float) * 4 * width * height) != hipSuccess) return 3;
    if (hipMalloc(&bytes, 4 * width * height) != hipSuccess) return 3;
    kdgrid->Integrate(pixels, width, height);
    kdgrid->ConvertToBytes(pixels, bytes, width, height);
    std::vector<float> density(4 * width * height);
    kdgrid->SamplingDensity(density.data(), width, height);
    std::vector<unsigned char> img(4 * width * height);
    if (hipMemcpy(img.data(), bytes, img.size(), hipMemcpyDeviceToHost) != hipSuccess) return 4;
    const bool valid = kdgrid->Validate();
    std::printf("passes %d samples %d nodes %d valid %d first pixel %d %d %d\n", passes,
                kdgrid->GetNumberOfSamples(), kdgrid->GetNumberOfNodes(), valid ? 1 : 0, img[0],
                img[1], img[2]);
    (void)hipFree(pixels);
    (void)hipFree(bytes);
    mdas_scene_destroy(scene);
    delete kdgrid;
    return valid && passes > 0 ? 0 : 1;
}
