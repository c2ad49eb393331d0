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

#include "mdas.h"

namespace {

int fail(const char* what) {
    std::fprintf(stderr, "mdas_render: %s: %s\n", what, mdas_last_error());
    return 1;
}

// Rows bottom to top, as the sampler's y axis runs.
bool write_pfm(const std::string& path, const std::vector<float>& rgba, int w, int h) {
    FILE* f = std::fopen(path.c_str(), "wb");
    if (!f) return false;
    std::fprintf(f, "PF\n%d %d\n-1.0\n", w, h);
    std::vector<float> row(3 * (size_t)w);
    for (int y = 0; y < h; ++y) {
        for (int x = 0; x < w; ++x)
            for (int c = 0; c < 3; ++c) row[3 * x + c] = rgba[4 * ((size_t)y * w + x) + c];
        std::fwrite(row.data(), sizeof(float), row.size(), f);
    }
    return std::fclose(f) == 0;
}

// PPM rows run top to bottom.
bool write_ppm(const std::string& path, const std::vector<unsigned char>& rgba, int w, int h) {
    FILE* f = std::fopen(path.c_str(), "wb");
    if (!f) return false;
    std::fprintf(f, "P6\n%d %d\n255\n", w, h);
    std::vector<unsigned char> row(3 * (size_t)w);
    for (int y = h - 1; y >= 0; --y) {
        for (int x = 0; x < w; ++x)
            for (int c = 0; c < 3; ++c) row[3 * x + c] = rgba[4 * ((size_t)y * w + x) + c];
        std::fwrite(row.data(), 1, row.size(), f);
    }
    return std::fclose(f) == 0;
}

// Uncompressed scanline OpenEXR, FLOAT channels A, B, G, R (alphabetical, as
// the format requires), alpha 1; EXR rows run top to bottom.
bool write_exr(const std::string& path, const std::vector<float>& rgba, int w, int h) {
    std::string hdr;
    auto i32 = [&](int32_t v) { hdr.append((const char*)&v, 4); };
    auto f32 = [&](float v) { hdr.append((const char*)&v, 4); };
    auto attr = [&](const char* name, const char* type, int32_t size) {
        hdr.append(name).push_back('\0');
        hdr.append(type).push_back('\0');
        i32(size);
    };
    i32(20000630);  // magic
    i32(2);         // version 2, single-part scanline
    attr("channels", "chlist", 4 * 18 + 1);
    for (const char* c : {"A", "B", "G", "R"}) {
        hdr.append(c).push_back('\0');
        i32(2);                          // pixel type FLOAT
        hdr.append(4, '\0');             // pLinear + reserved
        i32(1);                          // x sampling
        i32(1);                          // y sampling
    }
    hdr.push_back('\0');
    attr("compression", "compression", 1);
    hdr.push_back('\0');                 // none
    attr("dataWindow", "box2i", 16);
    i32(0); i32(0); i32(w - 1); i32(h - 1);
    attr("displayWindow", "box2i", 16);
    i32(0); i32(0); i32(w - 1); i32(h - 1);
    attr("lineOrder", "lineOrder", 1);
    hdr.push_back('\0');                 // increasing y
    attr("pixelAspectRatio", "float", 4);
    f32(1.0f);
    attr("screenWindowCenter", "v2f", 8);
    f32(0.0f); f32(0.0f);
    attr("screenWindowWidth", "float", 4);
    f32(1.0f);
    hdr.push_back('\0');                 // end of header
    const int64_t line_bytes = 16 * (int64_t)w;
    const uint64_t first = hdr.size() + 8 * (uint64_t)h;
    FILE* f = std::fopen(path.c_str(), "wb");
    if (!f) return false;
    std::fwrite(hdr.data(), 1, hdr.size(), f);
    for (int y = 0; y < h; ++y) {
        const uint64_t off = first + (uint64_t)y * (8 + line_bytes);
        std::fwrite(&off, 8, 1, f);
    }
    std::vector<float> line(4 * (size_t)w);
    for (int y = 0; y < h; ++y) {
        const int32_t head[2] = {y, (int32_t)line_bytes};
        std::fwrite(head, 4, 2, f);
        const size_t row = (size_t)(h - 1 - y) * w;  // top row first
        for (int x = 0; x < w; ++x) {
            line[x] = 1.0f;                                  // A
            line[w + x] = rgba[4 * (row + x) + 2];           // B
            line[2 * (size_t)w + x] = rgba[4 * (row + x) + 1];  // G
            line[3 * (size_t)w + x] = rgba[4 * (row + x)];      // R
        }
        std::fwrite(line.data(), sizeof(float), line.size(), f);
    }
    return std::fclose(f) == 0;
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
