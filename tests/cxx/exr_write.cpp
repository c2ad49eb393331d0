// Writes a test image with the .env front-end's image writers
// (pmdas_amd/csrc/image_io.hpp): argv[1] = output prefix, argv[2] argv[3] =
// width, height.  Pixel (x, y), row 0 at the bottom: R = x + y/1000,
// G = -y, B = a power of two of x; the first pixels carry 0, -0, a subnormal,
// the largest float and infinity.
#include <cmath>
#include <cstdlib>
#include <limits>
#include <string>
#include <vector>

#include "image_io.hpp"

int main(int argc, char** argv) {
    if (argc < 4) return 2;
    const std::string out = argv[1];
    const int w = std::atoi(argv[2]), h = std::atoi(argv[3]);
    std::vector<float> rgba(4 * (size_t)w * h);
    std::vector<unsigned char> b8(4 * (size_t)w * h);
    for (int y = 0; y < h; ++y)
        for (int x = 0; x < w; ++x) {
            float* p = &rgba[4 * ((size_t)y * w + x)];
            p[0] = (float)x + (float)y / 1000.0f;
            p[1] = -(float)y;
            p[2] = std::ldexp(1.0f, (x % 40) - 20);
            p[3] = 0.0f;
            for (int c = 0; c < 4; ++c) b8[4 * ((size_t)y * w + x) + c] = (unsigned char)((x * 7 + y * 3 + c) & 255);
        }
    const float special[] = {0.0f, -0.0f, std::numeric_limits<float>::denorm_min(),
                             std::numeric_limits<float>::max(), std::numeric_limits<float>::infinity()};
    for (int k = 0; k < 5 && k < w * h; ++k) rgba[4 * (size_t)k] = special[k];
    const bool ok = mdas_io::write_exr(out + ".exr", rgba, w, h) &&
                    mdas_io::write_pfm(out + ".pfm", rgba, w, h) && mdas_io::write_ppm(out + ".ppm", b8, w, h);
    return ok ? 0 : 1;
}
