// Reader of OpenEXR files through the reference's own tinyexr -- TEST
// INFRASTRUCTURE, NOT PRODUCT CODE.  tinyexr.h (with its embedded miniz) is
// compiled where it lies in the reference (SDK/support/tinyexr/tinyexr.h,
// which the reference's sutil uses to load and save its images,
// SDK/sutil/sutil.cpp:311, 698-728) by `make -C oracle ref`; the library goes
// to oracle/_ref/.  A CPU test reads this build's EXR output with it.
#include <cstdlib>
#include <cstring>

#define TINYEXR_IMPLEMENTATION
#include "tinyexr.h"

// RGBA float pixels of `path` (rows as stored, top first) into out[cap];
// returns 0, or tinyexr's negative error code.
extern "C" int exr_ref_load(const char* path, float* out, long long cap, int* w, int* h) {
    float* rgba = nullptr;
    int W = 0, H = 0;
    const char* err = nullptr;
    const int rc = LoadEXR(&rgba, &W, &H, path, &err);
    if (rc != TINYEXR_SUCCESS) {
        if (err) FreeEXRErrorMessage(err);
        return rc < 0 ? rc : -1;
    }
    *w = W;
    *h = H;
    if ((long long)W * H * 4 <= cap) std::memcpy(out, rgba, sizeof(float) * 4 * (size_t)W * H);
    std::free(rgba);
    return 0;
}
