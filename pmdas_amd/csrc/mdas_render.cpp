// mdas_render: the sample apps' main loop (optixMotionBlur.cpp:755-806) over
// libmdas -- read a reference .env file, sample the image with MDAS (or the
// plain QMC sampler when Sampler.mdas is false), reconstruct, and write the
// image, the sampling density and the KEY/value log the experiment scripts
// parse.  The OptiX scene is replaced by an analytic one (Scene block of the
// .env file, or --scene / --seed / --dim).
//
//   mdas_render file.env [--out PREFIX] [--scene KIND] [--seed N] [--dim D]
//                        [--max-samples N]
//
// Writes PREFIX.pfm (float RGB), PREFIX.exr (uncompressed float RGBA, the
// format the reference apps save), PREFIX.ppm (8-bit sRGB), PREFIX-density.pfm
// (MDAS only) and PREFIX.log.  Exit status 0 on success.
#include <hip/hip_runtime.h>

#include <chrono>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <string>
#include <vector>

#include "image_io.hpp"
#include "mdas.h"

using mdas_io::write_exr;
using mdas_io::write_pfm;
using mdas_io::write_ppm;

namespace {

int fail(const char* what) {
    std::fprintf(stderr, "mdas_render: %s: %s\n", what, mdas_last_error());
    return 1;
}

}  // namespace

int main(int argc, char** argv) {
    if (argc < 2) {
        std::fprintf(stderr, "usage: %s file.env [--out PREFIX] [--scene KIND] [--seed N] "
                             "[--dim D] [--max-samples N]\n", argv[0]);
        return 2;
    }
    mdas_env env;
    if (mdas_env_read(argv[1], &env) != MDAS_OK) {
        std::fprintf(stderr, "mdas_render: %s\n", mdas_env_last_error());
        return 1;
    }
    std::string out = env.filename[0] ? std::string(env.filename) : std::string("mdas_render");
    if (out.size() > 4 && (out.substr(out.size() - 4) == ".exr" || out.substr(out.size() - 4) == ".png"))
        out = out.substr(0, out.size() - 4);
    uint32_t seed = env.scene_seed;
    int dim = env.dim;
    // without a Scene block: the analytic scene of the point dimension
    // (3 motion-blurred disks, 4 depth of field, 5 both, 6 ambient occlusion)
    const int by_dim[7] = {0, 0, 0, MDAS_SCENE_DISKS_MB, MDAS_SCENE_SPHERES_DOF,
                           MDAS_SCENE_SPHERES_MBDOF, MDAS_SCENE_AO};
    int kind = env.scene_kind >= 0 ? env.scene_kind : by_dim[dim >= 3 && dim <= 6 ? dim : 5];
    long long max_samples = 30000000;  // optixMotionBlur.cpp:81
    if ((argc - 2) % 2 != 0) {
        std::fprintf(stderr, "mdas_render: option %s has no value\n", argv[argc - 1]);
        return 2;
    }
    for (int i = 2; i + 1 < argc; i += 2) {
        const std::string k = argv[i];
        if (k == "--out") out = argv[i + 1];
        else if (k == "--scene") kind = std::atoi(argv[i + 1]);
        else if (k == "--seed") seed = (uint32_t)std::strtoul(argv[i + 1], nullptr, 10);
        else if (k == "--dim") {
            dim = std::atoi(argv[i + 1]);
            if (env.scene_kind < 0) kind = by_dim[dim >= 3 && dim <= 6 ? dim : 5];
        }
        else if (k == "--max-samples") max_samples = std::atoll(argv[i + 1]);
        else {
            std::fprintf(stderr, "mdas_render: unknown option %s\n", k.c_str());
            return 2;
        }
    }
    const int W = env.width, H = env.height;
    const float sx = env.scale_factor * (float)W, sy = env.scale_factor * (float)H;
    mdas_scene_handle scene = nullptr;
    if (mdas_scene_create(kind, seed, dim, sx, sy, &scene) != MDAS_OK) return fail("scene");
    float* pixels = nullptr;
    unsigned char* bytes = nullptr;
    if (hipMalloc((void**)&pixels, sizeof(float) * 4 * (size_t)W * H) != hipSuccess ||
        hipMalloc((void**)&bytes, 4 * (size_t)W * H) != hipSuccess) {
        std::fprintf(stderr, "mdas_render: device allocation failed\n");
        return 1;
    }
    const std::string log_path = out + ".log";
    const long long total = (long long)(env.samples * (float)W * (float)H);
    mdas_handle grid = nullptr;
    std::vector<float> host((size_t)4 * W * H);
    if (env.mdas) {
        mdas_config cfg;
        mdas_env_config(&env, max_samples, &cfg);
        cfg.dim = dim;
        cfg.log_path = log_path.c_str();  // created (truncated) by mdas_create
        if (mdas_create(&cfg, &grid) != MDAS_OK) return fail("grid");
        {
            std::ofstream log(log_path, std::ios::app);
            log << "WIDTH\n" << W << "\nHEIGHT\n" << H << std::endl;
        }
        if (mdas_run(grid, scene, total, nullptr) != MDAS_OK) return fail("sampling");
        std::printf("samples %d, nodes %d, passes %d\n", mdas_get_number_of_samples(grid),
                    mdas_get_number_of_nodes(grid), mdas_get_passes(grid));
        if (mdas_integrate(grid, pixels, W, H, nullptr) != MDAS_OK) return fail("integrate");
        std::vector<float> density((size_t)4 * W * H);
        if (mdas_sampling_density(grid, density.data(), W, H) != MDAS_OK) return fail("density");
        if (!write_pfm(out + "-density.pfm", density, W, H)) return fail("write density");
    } else {
        // the plain sampler: `samples` per pixel in one launch (whitted.cu:59-83)
        std::ofstream log(log_path, std::ios::trunc);
        log << "WIDTH\n" << W << "\nHEIGHT\n" << H << std::endl;
        const int spp = env.samples < 1.0f ? 1 : (int)env.samples;
        const auto t0 = std::chrono::steady_clock::now();
        if (mdas_render_qmc(scene, W, H, spp, 0, 0, 0, pixels, nullptr) != MDAS_OK) return fail("qmc");
        if (hipDeviceSynchronize() != hipSuccess) return fail("qmc");
        const std::chrono::duration<double, std::milli> t = std::chrono::steady_clock::now() - t0;
        log << "TRACE TIME\n" << t.count() << std::endl;
    }
    if (mdas_convert_to_bytes(grid, pixels, bytes, W, H, nullptr) != MDAS_OK) return fail("bytes");
    std::vector<unsigned char> b8((size_t)4 * W * H);
    if (hipMemcpy(host.data(), pixels, sizeof(float) * host.size(), hipMemcpyDeviceToHost) != hipSuccess ||
        hipMemcpy(b8.data(), bytes, b8.size(), hipMemcpyDeviceToHost) != hipSuccess) {
        std::fprintf(stderr, "mdas_render: copy failed\n");
        return 1;
    }
    if (!write_pfm(out + ".pfm", host, W, H) || !write_exr(out + ".exr", host, W, H) ||
        !write_ppm(out + ".ppm", b8, W, H)) {
        std::fprintf(stderr, "mdas_render: cannot write %s.*\n", out.c_str());
        return 1;
    }
    (void)hipFree(pixels);
    (void)hipFree(bytes);
    if (grid) mdas_destroy(grid);
    mdas_scene_destroy(scene);
    std::printf("wrote %s.pfm, %s.exr, %s.ppm, %s.log\n", out.c_str(), out.c_str(), out.c_str(),
                out.c_str());
    return 0;
}
