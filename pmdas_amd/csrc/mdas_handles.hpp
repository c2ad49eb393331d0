// The opaque handle types of the C ABI and helpers shared by the host files.
#pragma once
#include <string>

#include "mdas.h"
#include "mdas_internal.hpp"
#include "scenes.hpp"

struct mdas_grid {
    mdas::Grid g;
};

struct mdas_scene {
    mscene::Params host;
    void* dev = nullptr;
};

namespace mdas {
// Sets the calling thread's last error message and returns `code`.
int set_error(int code, const std::string& msg);
// Copies the device counters of a grid to its host mirror (synchronising).
int sync_grid(Grid& g, hipStream_t st);
}  // namespace mdas
