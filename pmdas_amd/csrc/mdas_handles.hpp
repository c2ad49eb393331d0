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
// The largest leaf extents of a fresh initial grid (after initial sampling).
void initial_extents(Grid& g);
// Appends the per-pass rows written since the last drain (g.host.passes must
// be current) to g.pstats_host; unless `all`, only once half the device ring
// is pending, so that a short run copies its rows once, at its end.
int drain_pass_stats(Grid& g, bool all);
}  // namespace mdas
