// The reference app's MDAS code compiled against include/mdas_kdgrid.hpp and
// libmdas.so.  The statements marked "app" are the lines of
// SDK/optixMotionBlur/optixMotionBlur.cpp that use KDGrid, unchanged:
// initMdas (:331-360, minus the Environment reads that fill its locals),
// samplingPassMdas (:473-479), integrateMdas (:482-491, minus the
// CUDAOutputBuffer map/unmap) and the batch loop (:764-791).  LaunchParams
// stands in for whitted::LaunchParams (SDK/cuda/whitted.h:60-91) with the
// fields those lines write; the OptiX launch (launchSubframe) is replaced by
// the library's analytic integrand over params.sample_offset / sample_count,
// and a check that the values it wrote are visible through the float3*
// the app keeps.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdio>
#include <fstream>
#include <iostream>
#include <string>
#include <vector>

#include "mdas_kdgrid.hpp"

typedef mdas::Point5 Point;

namespace whitted {
struct LaunchParams {
    float2 scale;
    float* sample_coordinates;
    float3* sample_values;
    int sample_offset;
    int sample_count;
    int sample_dim;
};
}  // namespace whitted

// the app's globals (optixMotionBlur.cpp:81-91)
static int32_t width = 160;
static int32_t height = 90;
static int32_t max_samples = 1 << 21;
static float samples_per_launch = 8.0f;
static whitted::LaunchParams params = {};
static mdas::KDGrid<Point>* kdgrid = nullptr;
static mdas_scene_handle scene = nullptr;
static long long traced = 0;

// launchSubframe's MDAS branch: the integrand over the slice the app names
static void launchSubframe() {
    mdas::check(mdas_eval_integrand(kdgrid->handle(), scene, params.sample_offset,
                                    params.sample_count, nullptr));
    traced += params.sample_count;
}

void initMdas(std::ofstream* log = nullptr)
{
    float alpha = 0.25f, scaleFactor = 1.0f;
    int bitsPerDim = 1, extraImgBits = 3, candidatesNum = 4;

    // app: optixMotionBlur.cpp:342-359
    kdgrid = new mdas::KDGrid<Point>(
        max_samples,
        candidatesNum,
        bitsPerDim,
        extraImgBits,
        alpha,
        scaleFactor * width,
        scaleFactor * height,
        log
        );
    kdgrid->InitialSampling();

    params.scale = make_float2(kdgrid->GetScaleX(), kdgrid->GetScaleY());
    params.sample_coordinates = reinterpret_cast<float*>(kdgrid->GetSampleCoordinates().Data());
    params.sample_values = kdgrid->GetSampleValues().Data();
    params.sample_count = kdgrid->GetNumberOfSamples();
    params.sample_offset = 0;
    params.sample_dim = Point::DIM;
}

void samplingPassMdas()
{
    // app: optixMotionBlur.cpp:475-478
    kdgrid->SamplingPass();
    kdgrid->Validate();
    params.sample_count = kdgrid->GetNewSamples();
    params.sample_offset = kdgrid->GetNumberOfSamples() - kdgrid->GetNewSamples();
}

void integrateMdas(float4* result_buffer_data, uchar4* result_buffer_data_bytes)
{
    // app: optixMotionBlur.cpp:486-488
    kdgrid->Validate();
    kdgrid->Integrate(result_buffer_data, width, height);
    kdgrid->ConvertToBytes(result_buffer_data, result_buffer_data_bytes, width, height);
}

int main(int argc, char** argv)
{
    const std::string outfile = argc > 1 ? argv[1] : "drop_in.png";
    bool mdas_on = true;
    // app: optixMotionBlur.cpp:764-768
    std::ofstream log(outfile + ".log");
    log << "WIDTH\n" << width << std::endl;
    log << "HEIGHT\n" << height << std::endl;

    if (mdas_on) initMdas(&log);

    if (mdas_scene_create(MDAS_SCENE_SPHERES_MBDOF, 1, Point::DIM, params.scale.x, params.scale.y,
                          &scene) != MDAS_OK)
        return 2;
    launchSubframe();
    int passes = 0;
    if (mdas_on)
    {
        // app: optixMotionBlur.cpp:778-781
        int total_samples = static_cast<int>(samples_per_launch * width * height);
        std::cout << "Total samples " << total_samples << std::endl;
        std::cout << "Sampling..." << std::endl;
        while (kdgrid->GetNumberOfSamples() < std::min(total_samples, max_samples))
        {
            samplingPassMdas();
            launchSubframe();
            ++passes;
        }
    }
    float4* pixels = nullptr;
    uchar4* bytes = nullptr;
    if (hipMalloc(&pixels, sizeof(float4) * width * height) != hipSuccess) return 3;
    if (hipMalloc(&bytes, sizeof(uchar4) * width * height) != hipSuccess) return 3;
    integrateMdas(pixels, bytes);
    std::vector<float4> density((size_t)width * height);
    kdgrid->SamplingDensity(density.data(), width, height);

    // the integrand's values through the app's float3* equal the library's
    const int N = kdgrid->GetNumberOfSamples();
    std::vector<float3> via_params((size_t)N);
    std::vector<float> via_abi((size_t)N * 3);
    if (hipMemcpy(via_params.data(), params.sample_values, sizeof(float3) * N,
                  hipMemcpyDeviceToHost) != hipSuccess ||
        hipMemcpy(via_abi.data(), mdas_get_sample_values(kdgrid->handle()), sizeof(float) * 3 * N,
                  hipMemcpyDeviceToHost) != hipSuccess)
        return 4;
    bool same = true;
    for (int i = 0; i < N; ++i)
        same = same && via_params[i].x == via_abi[3 * i] && via_params[i].y == via_abi[3 * i + 1] &&
               via_params[i].z == via_abi[3 * i + 2];
    std::vector<uchar4> img((size_t)width * height);
    if (hipMemcpy(img.data(), bytes, sizeof(uchar4) * img.size(), hipMemcpyDeviceToHost) != hipSuccess)
        return 4;
    const bool valid = kdgrid->Validate();
    std::printf("passes %d samples %d nodes %d traced %lld valid %d values %d first pixel %d %d %d\n",
                passes, kdgrid->GetNumberOfSamples(), kdgrid->GetNumberOfNodes(), traced,
                valid ? 1 : 0, same ? 1 : 0, img[0].x, img[0].y, img[0].z);
    (void)hipFree(pixels);
    (void)hipFree(bytes);
    mdas_scene_destroy(scene);
    delete kdgrid;
    return valid && same && passes > 0 && traced == N ? 0 : 1;
}
