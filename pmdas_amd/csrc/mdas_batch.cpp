// Batched MDAS: many grids (images of uneven resolution, grouped by point
// dimension) advanced together.  Every pass of a group is one launch per
// kernel for all of its images (the kernels take a device array of grid
// descriptors and read their image from blockIdx.y), and a block of passes is
// captured once into a hipGraph and replayed until every image's device-side
// stopping rule has fired.  Each image follows exactly the single-image
// algorithm, so its result equals mdas_run on that image alone.
#include <hip/hip_runtime.h>
#include <stdint.h>
#include <string.h>

#include <algorithm>
#include <chrono>
#include <string>
#include <vector>

#include "mdas_handles.hpp"

namespace {

struct Group {
    int D = 0;
    std::vector<int> members;           // batch indices
    mdas::DevGrid* d_grids = nullptr;    // device copies of the descriptors
    const void** d_params = nullptr;     // device array of scene parameter blocks
    int* d_limits = nullptr;
    int* d_status = nullptr;             // (done, status, samples, passes) per member
    mdas::LearnDev* d_learn = nullptr;   // per member, when any member is in learned mode
    int blocks = 0;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
};

#define B_TRY(expr)                                                                            \
    do {                                                                                       \
        hipError_t e_ = (expr);                                                                \
        if (e_ != hipSuccess)                                                                  \
            return mdas::set_error(MDAS_ERR_HIP, std::string(#expr) + ": " + hipGetErrorString(e_)); \
    } while (0)

__global__ void k_batch_status(const mdas::DevGrid* grids, int n, int* out) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t < n) {
        out[4 * t] = grids[t].cnt->done;
        out[4 * t + 1] = grids[t].cnt->status;
        out[4 * t + 2] = grids[t].cnt->n_samples;
        out[4 * t + 3] = grids[t].cnt->passes;
    }
}

}  // namespace

struct mdas_batch {
    std::vector<mdas_handle> grids;
    std::vector<mdas_scene_handle> scenes;
    std::vector<Group> groups;
    // the stream a null-stream caller's graphs are captured and launched on
    // (the null stream cannot be captured), joined to the null stream on
    // both sides by the two events
    hipStream_t own_stream = nullptr;
    hipEvent_t own_in = nullptr, own_out = nullptr;
    int passes_per_block = 4;
    double last_ms = 0.0;
    int last_blocks = 0;
    int shape = 8192;  // workgroups per launch in all (mdas_batch_autotune)
    // the images' reconstructions as one captured graph (mdas_batch_integrate),
    // with every pointer and size its launches carry
    std::vector<unsigned long long> recon_key;
    hipGraph_t recon_graph = nullptr;
    hipGraphExec_t recon_exec = nullptr;
};

namespace {
// The stream a call runs on: the caller's, or for a null-stream caller that
// wants graphs the batch's own stream, made to wait for the work already on
// the null stream (a non-blocking stream does not by itself).
hipError_t enter_stream(mdas_batch* b, void* stream, bool use_graph, hipStream_t* st) {
    *st = (hipStream_t)stream;
    if (!use_graph || stream) return hipSuccess;
    hipError_t e = hipSuccess;
    if (!b->own_stream) e = hipStreamCreateWithFlags(&b->own_stream, hipStreamNonBlocking);
    if (e == hipSuccess && !b->own_in) e = hipEventCreateWithFlags(&b->own_in, hipEventDisableTiming);
    if (e == hipSuccess && !b->own_out) e = hipEventCreateWithFlags(&b->own_out, hipEventDisableTiming);
    if (e == hipSuccess) e = hipEventRecord(b->own_in, nullptr);
    if (e == hipSuccess) e = hipStreamWaitEvent(b->own_stream, b->own_in, 0);
    if (e == hipSuccess) *st = b->own_stream;
    return e;
}

// After a call that ran on the batch's own stream: later work on the null
// stream waits for it, as it would had the call run there.
hipError_t leave_stream(mdas_batch* b, hipStream_t st) {
    if (!b->own_stream || st != b->own_stream) return hipSuccess;
    hipError_t e = hipEventRecord(b->own_out, st);
    if (e == hipSuccess) e = hipStreamWaitEvent(nullptr, b->own_out, 0);
    return e;
}

void drop_recon_graph(mdas_batch* b) {
    if (b->recon_exec) (void)hipGraphExecDestroy(b->recon_exec);
    if (b->recon_graph) (void)hipGraphDestroy(b->recon_graph);
    b->recon_exec = nullptr;
    b->recon_graph = nullptr;
    b->recon_key.clear();
}
}  // namespace

extern "C" {

int mdas_batch_create(mdas_handle* grids, mdas_scene_handle* scenes, int n, mdas_batch_handle* out) {
    if (!grids || !scenes || !out || n <= 0) return mdas::set_error(MDAS_ERR_INVALID, "bad arguments");
    *out = nullptr;
    for (int i = 0; i < n; ++i) {
        if (!grids[i] || !scenes[i]) return mdas::set_error(MDAS_ERR_INVALID, "null grid or scene");
        if (scenes[i]->host.dim != grids[i]->g.D)
            return mdas::set_error(MDAS_ERR_INVALID, "scene and grid dimensions differ at " + std::to_string(i));
        for (int j = 0; j < i; ++j)
            if (grids[j] == grids[i]) return mdas::set_error(MDAS_ERR_INVALID, "a grid appears twice");
    }
    mdas_batch* b = new mdas_batch;
    b->grids.assign(grids, grids + n);
    b->scenes.assign(scenes, scenes + n);
    for (int D = 3; D <= 6; ++D) {
        Group gr;
        gr.D = D;
        for (int i = 0; i < n; ++i)
            if (grids[i]->g.D == D) gr.members.push_back(i);
        if (gr.members.empty()) continue;
        const int m = (int)gr.members.size();
        // workgroups per image: about 8192 per launch in all, at least 64
        gr.blocks = std::max(64, std::min(mdas::kChunkBlocks, 8192 / m));
        std::vector<mdas::DevGrid> dg(m);
        std::vector<const void*> dp(m);
        std::vector<mdas::LearnDev> dl(m);
        bool learned = false;
        for (int k = 0; k < m; ++k) {
            dg[k] = grids[gr.members[k]]->g.dev;
            dg[k].in_loop = 1;  // batches only run the device-side loop
            dp[k] = scenes[gr.members[k]]->dev;
            dl[k] = grids[gr.members[k]]->g.learn;  // feat == null: plain error for this member
            learned = learned || grids[gr.members[k]]->g.learn_mode != 0;
        }
        hipError_t e = hipMalloc(&gr.d_grids, sizeof(mdas::DevGrid) * m);
        if (e == hipSuccess) e = hipMalloc((void**)&gr.d_params, sizeof(void*) * m);
        if (e == hipSuccess) e = hipMalloc(&gr.d_limits, sizeof(int) * m);
        if (e == hipSuccess) e = hipMalloc(&gr.d_status, sizeof(int) * 4 * m);
        if (e == hipSuccess) e = hipMemcpy(gr.d_grids, dg.data(), sizeof(mdas::DevGrid) * m, hipMemcpyHostToDevice);
        if (e == hipSuccess) e = hipMemcpy(gr.d_params, dp.data(), sizeof(void*) * m, hipMemcpyHostToDevice);
        if (e == hipSuccess && learned) e = hipMalloc(&gr.d_learn, sizeof(mdas::LearnDev) * m);
        if (e == hipSuccess && learned)
            e = hipMemcpy(gr.d_learn, dl.data(), sizeof(mdas::LearnDev) * m, hipMemcpyHostToDevice);
        b->groups.push_back(gr);
        if (e != hipSuccess) {
            mdas_batch_destroy(b);
            return mdas::set_error(MDAS_ERR_HIP, std::string("batch allocation: ") + hipGetErrorString(e));
        }
    }
    *out = b;
    return MDAS_OK;
}

int mdas_batch_destroy(mdas_batch_handle b) {
    if (!b) return MDAS_OK;
    (void)hipDeviceSynchronize();
    for (Group& g : b->groups) {
        if (g.exec) (void)hipGraphExecDestroy(g.exec);
        if (g.graph) (void)hipGraphDestroy(g.graph);
        if (g.d_grids) (void)hipFree(g.d_grids);
        if (g.d_params) (void)hipFree((void*)g.d_params);
        if (g.d_limits) (void)hipFree(g.d_limits);
        if (g.d_status) (void)hipFree(g.d_status);
        if (g.d_learn) (void)hipFree(g.d_learn);
    }
    drop_recon_graph(b);
    if (b->own_in) (void)hipEventDestroy(b->own_in);
    if (b->own_out) (void)hipEventDestroy(b->own_out);
    if (b->own_stream) (void)hipStreamDestroy(b->own_stream);
    delete b;
    return MDAS_OK;
}

// Runs every image of the batch to its target: initial grids, integrand,
// then passes while N < min(target, max_samples) per image.  With use_graph
// the block of passes is a captured hipGraph (a stream is created when the
// caller passes the null stream, which cannot be captured).
int mdas_batch_run(mdas_batch_handle b, const long long* targets, void* stream, int use_graph) {
    if (!b || !targets) return mdas::set_error(MDAS_ERR_INVALID, "null argument");
    const int n = (int)b->grids.size();
    for (int i = 0; i < n; ++i)
        if (targets[i] < 0) return mdas::set_error(MDAS_ERR_INVALID, "negative target");
    hipStream_t st = nullptr;
    B_TRY(enter_stream(b, stream, use_graph != 0, &st));
    const auto t0 = std::chrono::steady_clock::now();
    const int B = std::max(1, b->passes_per_block);
    // initial grids and their integrand, one launch per kernel per group
    std::vector<std::vector<int>> limits(b->groups.size());
    for (size_t gi = 0; gi < b->groups.size(); ++gi) {
        Group& g = b->groups[gi];
        const int m = (int)g.members.size();
        limits[gi].resize(m);
        for (int k = 0; k < m; ++k) {
            const mdas::Grid& gg = b->grids[g.members[k]]->g;
            limits[gi][k] = (int)std::min<long long>(targets[g.members[k]], gg.max_samples);
        }
        B_TRY(hipMemcpyAsync(g.d_limits, limits[gi].data(), sizeof(int) * m, hipMemcpyHostToDevice, st));
        B_TRY(mdas::launch_initial_b(g.D, g.d_grids, m, g.blocks, g.d_limits, st));
        B_TRY(mdas::launch_eval_b(g.d_grids, g.d_params, m, g.blocks, st));
        for (int k = 0; k < m; ++k) {
            b->grids[g.members[k]]->g.initialized = true;
            mdas::initial_extents(b->grids[g.members[k]]->g);
            b->grids[g.members[k]]->g.pstats_host.clear();
        }
    }
    std::vector<char> finished(b->groups.size(), 0);
    std::vector<int> last_n(n, -1);  // samples of each image after the last block
    int blocks_run = 0;
    for (;;) {
        for (size_t gi = 0; gi < b->groups.size(); ++gi) {
            if (finished[gi]) continue;
            Group& g = b->groups[gi];
            const int m = (int)g.members.size();
            if (use_graph) {
                if (!g.exec) {
                    B_TRY(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
                    hipError_t ce = hipSuccess;
                    for (int p = 0; p < B && ce == hipSuccess; ++p) {
                        ce = mdas::launch_pass_b(g.D, g.d_grids, m, g.blocks, st, g.d_learn);
                        if (ce == hipSuccess) ce = mdas::launch_eval_b(g.d_grids, g.d_params, m, g.blocks, st);
                    }
                    B_TRY(hipStreamEndCapture(st, &g.graph));
                    B_TRY(ce);
                    B_TRY(hipGraphInstantiate(&g.exec, g.graph, nullptr, nullptr, 0));
                }
                B_TRY(hipGraphLaunch(g.exec, st));
            } else {
                for (int p = 0; p < B; ++p) {
                    B_TRY(mdas::launch_pass_b(g.D, g.d_grids, m, g.blocks, st, g.d_learn));
                    B_TRY(mdas::launch_eval_b(g.d_grids, g.d_params, m, g.blocks, st));
                }
            }
            hipLaunchKernelGGL(k_batch_status, dim3((m + 63) / 64), dim3(64), 0, st, g.d_grids, m, g.d_status);
        }
        ++blocks_run;
        B_TRY(hipStreamSynchronize(st));
        bool all = true;
        for (size_t gi = 0; gi < b->groups.size(); ++gi) {
            if (finished[gi]) continue;
            Group& g = b->groups[gi];
            const int m = (int)g.members.size();
            std::vector<int> s(4 * m);
            B_TRY(hipMemcpy(s.data(), g.d_status, sizeof(int) * 4 * m, hipMemcpyDeviceToHost));
            bool done = true;
            for (int k = 0; k < m; ++k) {
                if (s[4 * k + 1] != 0) {
                    finished[gi] = 1;
                    return mdas::set_error(s[4 * k + 1], "image " + std::to_string(g.members[k]) +
                                                             ": adaptive pass would exceed max_samples");
                }
                // every pass selects the leaf of maximum error, so a block of
                // passes that adds no sample means the errors are not finite
                int& prev = last_n[g.members[k]];
                if (s[4 * k] == 0 && s[4 * k + 2] == prev)
                    return mdas::set_error(MDAS_ERR_STATE, "image " + std::to_string(g.members[k]) +
                                                               ": a block of passes added no sample "
                                                               "(no leaf selected: non-finite or zero "
                                                               "maximum error)");
                prev = s[4 * k + 2];
                done = done && s[4 * k] != 0;
                mdas::Grid& gk = b->grids[g.members[k]]->g;  // rows before the ring wraps
                gk.host.passes = s[4 * k + 3];
                const int rc = mdas::drain_pass_stats(gk, false);
                if (rc) return rc;
            }
            finished[gi] = done ? 1 : 0;
            all = all && done;
        }
        if (all) break;
    }
    for (Group& g : b->groups)
        B_TRY(mdas::launch_mark_last_children_b(g.d_grids, (int)g.members.size(), st));
    for (int i = 0; i < n; ++i) {
        mdas::Grid& gg = b->grids[i]->g;
        int rc = mdas::sync_grid(gg, st);
        if (rc) return rc;
        gg.passes = gg.host.passes;
        rc = mdas::drain_pass_stats(gg, true);
        if (rc) return rc;
    }
    B_TRY(leave_stream(b, st));
    b->last_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    b->last_blocks = blocks_run;
    return MDAS_OK;
}

// Reconstructs every image into its own float4 image (kdgrid.cu:781-820 per
// image).  With use_graph the reconstructions of all images are captured once
// into a hipGraph and replayed while every image's size, gather tile, output
// and reconstruction buffers stay the same; the buffers are sized before the
// capture from each image's leaf count, so nothing in the graph allocates or
// waits for the host.  Stream-ordered: the caller synchronises.
int mdas_batch_integrate(mdas_batch_handle b, float* const* pixels, const int* widths,
                         const int* heights, void* stream, int use_graph) {
    if (!b || !pixels || !widths || !heights) return mdas::set_error(MDAS_ERR_INVALID, "null argument");
    const int n = (int)b->grids.size();
    for (int i = 0; i < n; ++i) {
        const mdas::Grid& g = b->grids[i]->g;
        if (!pixels[i]) return mdas::set_error(MDAS_ERR_INVALID, "null image " + std::to_string(i));
        if (widths[i] <= 0 || heights[i] <= 0 || widths[i] > 65535 || heights[i] > 65535)
            return mdas::set_error(MDAS_ERR_INVALID, "image size out of range at " + std::to_string(i));
        if (!g.initialized) return mdas::set_error(MDAS_ERR_STATE, "image " + std::to_string(i) + " has no samples");
    }
    hipStream_t st = nullptr;
    B_TRY(enter_stream(b, stream, use_graph != 0, &st));
    bool bounded = true;
    for (int i = 0; i < n; ++i) {
        mdas::Grid& g = b->grids[i]->g;
        B_TRY(mdas::recon_reserve(g, widths[i], heights[i]));
        bounded = bounded && g.recon_bounded;
    }
    if (!use_graph || !bounded) {  // one image after the other (may synchronise)
        for (int i = 0; i < n; ++i) {
            const int rc = mdas_integrate(b->grids[i], pixels[i], widths[i], heights[i], st);
            if (rc) return rc;
        }
        B_TRY(leave_stream(b, st));
        return MDAS_OK;
    }
    std::vector<unsigned long long> key;
    for (int i = 0; i < n; ++i) {
        const mdas::Grid& g = b->grids[i]->g;
        const unsigned long long k[] = {
            (unsigned long long)(uintptr_t)pixels[i], (unsigned long long)widths[i],
            (unsigned long long)heights[i], (unsigned long long)g.tile_w, (unsigned long long)g.tile_h,
            (unsigned long long)g.chunk_blocks, (unsigned long long)(uintptr_t)g.rb.rrec,
            (unsigned long long)(uintptr_t)g.rb.keys, (unsigned long long)(uintptr_t)g.rb.keys_alt,
            (unsigned long long)g.rb.keys_cap, (unsigned long long)(uintptr_t)g.rb.tile_start,
            (unsigned long long)(uintptr_t)g.rb.tile_end, (unsigned long long)(uintptr_t)g.rb.rs_hist};
        key.insert(key.end(), k, k + sizeof(k) / sizeof(k[0]));
    }
    if (!b->recon_exec || key != b->recon_key) {
        drop_recon_graph(b);
        std::vector<char> timing(n);
        for (int i = 0; i < n; ++i) {  // no timing events inside the graph
            timing[i] = b->grids[i]->g.timing;
            b->grids[i]->g.timing = false;
        }
        B_TRY(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
        hipError_t ce = hipSuccess;
        for (int i = 0; i < n && ce == hipSuccess; ++i) {
            mdas::Grid& g = b->grids[i]->g;
            ce = mdas::recon_prepare(g, widths[i], heights[i], st);
            if (ce == hipSuccess)
                ce = mdas::launch_integrate_gather(g, pixels[i], widths[i], heights[i], st, nullptr, nullptr);
        }
        hipGraph_t graph = nullptr;
        const hipError_t ee = hipStreamEndCapture(st, &graph);
        for (int i = 0; i < n; ++i) b->grids[i]->g.timing = timing[i] != 0;
        B_TRY(ce);
        B_TRY(ee);
        b->recon_graph = graph;
        B_TRY(hipGraphInstantiate(&b->recon_exec, b->recon_graph, nullptr, nullptr, 0));
        b->recon_key = key;
    }
    B_TRY(hipGraphLaunch(b->recon_exec, st));
    B_TRY(leave_stream(b, st));
    return MDAS_OK;
}

// The launch shape of a batch: workgroups per image for each dimension group
// (per launch about `total` in all, clamped to [64, kChunkBlocks] per image).
static void set_batch_shape(mdas_batch* b, int total) {
    for (Group& g : b->groups) {
        const int m = (int)g.members.size();
        g.blocks = std::max(64, std::min(mdas::kChunkBlocks, total / m));
        if (g.exec) (void)hipGraphExecDestroy(g.exec);
        if (g.graph) (void)hipGraphDestroy(g.graph);
        g.exec = nullptr;
        g.graph = nullptr;
    }
}

int mdas_batch_autotune(mdas_batch_handle b, const long long* targets, void* stream, int* chosen) {
    if (!b || !targets) return mdas::set_error(MDAS_ERR_INVALID, "null argument");
    static const int cand[] = {2048, 4096, 8192, 16384, 32768};
    int best = cand[0];
    double best_ms = 0.0;
    for (int c : cand) {
        set_batch_shape(b, c);
        const int rc = mdas_batch_run(b, targets, stream, 1);
        if (rc) return rc;
        if (best_ms == 0.0 || b->last_ms < best_ms) {
            best_ms = b->last_ms;
            best = c;
        }
    }
    set_batch_shape(b, best);
    b->shape = best;
    if (chosen) *chosen = best;
    return MDAS_OK;
}

int mdas_batch_info(mdas_batch_handle b, double* wall_ms, int* pass_blocks, int* groups) {
    if (!b) return mdas::set_error(MDAS_ERR_INVALID, "null handle");
    if (wall_ms) *wall_ms = b->last_ms;
    if (pass_blocks) *pass_blocks = b->last_blocks;
    if (groups) *groups = (int)b->groups.size();
    return MDAS_OK;
}

}  // extern "C"
