// Image writers of the .env front-end (host code): PFM (float RGB), PPM
// (8-bit RGB) and uncompressed scanline OpenEXR (float RGBA), from RGBA pixel
// arrays whose row 0 is the bottom row, as the sampler's y axis runs.
#pragma once
#include <cstdint>
#include <cstdio>
#include <string>
#include <vector>

namespace mdas_io {

// Rows bottom to top, as the sampler's y axis runs.
inline bool write_pfm(const std::string& path, const std::vector<float>& rgba, int w, int h) {
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
inline bool write_ppm(const std::string& path, const std::vector<unsigned char>& rgba, int w, int h) {
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
inline bool write_exr(const std::string& path, const std::vector<float>& rgba, int w, int h) {
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

}  // namespace mdas_io
