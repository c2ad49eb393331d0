// C ABI of libmdas (include/mdas.h): handle lifetime, the reference's KDGrid
// entry points, the device-side sampling loop, timing and state transfer.
#include "mdas.h"

#include <hip/hip_runtime.h>
#include <stdio.h>
#include <string.h>

#include <algorithm>
#include <cfloat>
#include <chrono>
#include <cmath>
#include <fstream>
#include <string>
#include <vector>

#include "mdas_handles.hpp"

using mdas::Grid;

namespace {

thread_local std::string t_err;

int fail(int code, const std::string& msg) {
    t_err = msg;
    return code;
}

#define HIP_TRY(expr)                                                                   \
    do {                                                                                \
        hipError_t e_ = (expr);                                                         \
        if (e_ != hipSuccess)                                                           \
            return fail(MDAS_ERR_HIP, std::string(#expr) + ": " + hipGetErrorString(e_)); \
    } while (0)

hipStream_t S(void* s) { return (hipStream_t)s; }

template <typename T>
hipError_t dalloc(T** p, size_t n) {
    return hipMalloc((void**)p, sizeof(T) * (n ? n : 1));
}

// The LCG x -> 1664525 x + 1013904223 (mod 2^32) applied k times is
// x -> A x + C.
void lcg_jump(unsigned k, unsigned& A, unsigned& C) {
    A = 1u;
    C = 0u;
    for (unsigned j = 0; j < k; ++j) {
        A = 1664525u * A;
        C = 1664525u * C + 1013904223u;
    }
}

void free_grid(Grid& g) {
    mdas::DevGrid& d = g.dev;
    mdas::ReconBufs& rb = g.rb;
    void* ptrs[] = {d.coords,   d.values,  d.leaf,        d.err,        d.seed,     d.flags,
                    d.xpts,     d.split_list, d.sdat,     d.sext,       d.part,     d.part_off,
                    d.epart,    d.cnt,     d.pstats,      rb.rrec,      rb.keys,    rb.keys_alt,
                    rb.tile_start, rb.tile_end, rb.rs_hist, g.d_scalar, g.learn.feat, g.learn.perr,
                    g.learn.rows};
    for (void* p : ptrs)
        if (p) (void)hipFree(p);
    for (hipEvent_t e : g.evpool) (void)hipEventDestroy(e);
    g.evpool.clear();
    d = mdas::DevGrid{};
    rb = mdas::ReconBufs{};
    g.keys_sorted = nullptr;
    g.d_scalar = nullptr;
    g.learn = mdas::LearnDev{};
    g.learn_mode = 0;
}

int sync_counters(Grid& g, hipStream_t st) {
    HIP_TRY(hipMemcpyAsync(&g.host, g.dev.cnt, sizeof(mdas::DevCounters), hipMemcpyDeviceToHost, st));
    HIP_TRY(hipStreamSynchronize(st));
    g.gmax_host = 0.0f;
    memcpy(&g.gmax_host, &g.host.gmax_bits, 4);
    return MDAS_OK;
}

// One "KEY\nvalue" pair as the reference writes it (kdgrid.cu:670-671,
// 713, 769-770, 816-817): sample counts as integers, times as the float that
// hipEventElapsedTime gives, in the stream's default format.
void log_line(Grid& g, const char* key, double v, bool count) {
    if (g.log_fn) g.log_fn(g.log_user, key, v, count ? 1 : 0);
    if (g.log_path.empty()) return;
    std::ofstream f(g.log_path, std::ios::app);
    if (count)
        f << key << "\n" << (long long)v << std::endl;
    else
        f << key << "\n" << (float)v << std::endl;
}
void log_count(Grid& g, const char* key, long long n) { log_line(g, key, (double)n, true); }
void log_time(Grid& g, const char* key, double ms) { log_line(g, key, ms, false); }

float elapsed(hipEvent_t a, hipEvent_t b) {
    float ms = 0.0f;
    if (hipEventElapsedTime(&ms, a, b) != hipSuccess) return 0.0f;
    return ms;
}

// Events bracketing one API call for its log lines, created only when the
// grid logs or times, and destroyed on every return path.
struct CallEvents {
    hipEvent_t e[4] = {nullptr, nullptr, nullptr, nullptr};
    CallEvents() = default;
    CallEvents(const CallEvents&) = delete;
    CallEvents& operator=(const CallEvents&) = delete;
    ~CallEvents() {
        for (hipEvent_t x : e)
            if (x) (void)hipEventDestroy(x);
    }
    hipError_t create(int n) {
        for (int i = 0; i < n; ++i) {
            const hipError_t r = hipEventCreate(&e[i]);
            if (r != hipSuccess) return r;
        }
        return hipSuccess;
    }
};

// Fold the recorded kernel events into per-kernel totals and phase times.
// Folds the timed launches into per-kernel totals; seq (optional) receives
// every launch's (name, ms) in launch order.
void collect_kernel_times(Grid& g, std::vector<std::pair<std::string, double>>* seq = nullptr) {
    if (g.krecs.empty()) return;
    if (g.krecs.back().b) (void)hipEventSynchronize(g.krecs.back().b);
    for (const auto& r : g.krecs) {
        const double ms = (r.a && r.b) ? elapsed(r.a, r.b) : 0.0;
        if (seq) seq->emplace_back(r.name, ms);
        bool found = false;
        for (auto& k : g.kstats)
            if (k.name == r.name) {
                k.launches++;
                k.ms += ms;
                found = true;
                break;
            }
        if (!found) g.kstats.push_back(Grid::KStat{r.name, 1, ms});
    }
    g.krecs.clear();
    g.evnext = 0;
    auto get = [&](const char* n) {
        for (auto& k : g.kstats)
            if (k.name == n) return k.ms;
        return 0.0;
    };
    g.t_initial = get("k_initial");
    g.t_errors = get("k_errors");
    g.t_adaptive = get("k_select") + get("k_scan_counts") + get("k_place") + get("k_mark_last");
    g.t_integrand = get("k_eval");
}

// The initial cell's x / y extents, the largest any leaf has until an upload
// (kdgrid.cu:158-164: 2^-(b+e) of the sample space along x and y).
void set_initial_extents(Grid& g) {
    g.ext_x = std::ldexp((double)g.sX, -(g.b + g.e));
    g.ext_y = std::ldexp((double)g.sY, -(g.b + g.e));
}

}  // namespace

namespace mdas {
int set_error(int code, const std::string& msg) { return fail(code, msg); }
void initial_extents(Grid& g) { set_initial_extents(g); }
int sync_grid(Grid& g, hipStream_t st) { return sync_counters(g, st); }
int drain_pass_stats(Grid& g, bool all) {
    const int have = (int)g.pstats_host.size(), done = g.host.passes;
    if (done <= have || (!all && done - have < kMaxPassStats / 2)) return MDAS_OK;
    g.pstats_host.resize((size_t)done);
    for (int p = have; p < done;) {  // contiguous stretches of the ring
        const int r = p & (kMaxPassStats - 1);
        const int n = std::min(done - p, kMaxPassStats - r);
        HIP_TRY(hipMemcpy(&g.pstats_host[(size_t)p], g.dev.pstats + r, sizeof(PassStat) * (size_t)n,
                          hipMemcpyDeviceToHost));
        p += n;
    }
    return MDAS_OK;
}
}  // namespace mdas

extern "C" {

const char* mdas_last_error(void) { return t_err.c_str(); }

int mdas_device_synchronize(void) {
    HIP_TRY(hipDeviceSynchronize());
    return MDAS_OK;
}

int mdas_create(const mdas_config* cfg, mdas_handle* out) {
    if (!cfg || !out) return fail(MDAS_ERR_INVALID, "null argument");
    *out = nullptr;
    const int D = cfg->dim;
    if (D < 3 || D > 6) return fail(MDAS_ERR_INVALID, "dim must be 3..6");
    if (cfg->bits_per_dim < 0 || cfg->extra_img_bits < 0)
        return fail(MDAS_ERR_INVALID, "bitsPerDim and extraImgBits must be >= 0");
    if (cfg->candidates_num < 1) return fail(MDAS_ERR_INVALID, "candidatesNum must be >= 1");
    if (!(cfg->alpha >= 0.0f)) return fail(MDAS_ERR_INVALID, "alpha must be >= 0");
    if (!(cfg->scale_x > 0.0f) || !(cfg->scale_y > 0.0f))
        return fail(MDAS_ERR_INVALID, "scale_x and scale_y must be > 0");
    if (cfg->max_samples < 8 || cfg->max_samples > 0x7fffffffLL)
        return fail(MDAS_ERR_INVALID, "max_samples must be in [8, 2^31-1]");
    const int bits = cfg->bits_per_dim * D + 2 * cfg->extra_img_bits;
    if (bits > 28) return fail(MDAS_ERR_INVALID, "initial grid too large (bitsPerDim*dim + 2*extraImgBits > 28)");
    const long long L0 = 1LL << bits;
    if (4 * L0 > cfg->max_samples)
        return fail(MDAS_ERR_INVALID, "max_samples is smaller than the initial sample count 4*2^(b*dim+2e)");

    mdas_grid* h = new mdas_grid;
    Grid& g = h->g;
    g.D = D;
    g.cand = cfg->candidates_num;
    g.b = cfg->bits_per_dim;
    g.e = cfg->extra_img_bits;
    g.alpha = cfg->alpha;
    g.sX = cfg->scale_x;
    g.sY = cfg->scale_y;
    g.max_samples = cfg->max_samples;
    g.cap_s = (int)cfg->max_samples;
    g.cap_l = g.cap_s / 2 + 1;  // every leaf holds >= 2 samples, so L <= N/2
    if (cfg->log_path) g.log_path = cfg->log_path;
    mdas::DevGrid& d = g.dev;
    hipError_t e = hipSuccess;
    const size_t cs = (size_t)g.cap_s, cl = (size_t)g.cap_l;
    if (e == hipSuccess) e = dalloc(&d.coords, cs * D);
    if (e == hipSuccess) e = dalloc(&d.values, cs * 3);
    if (e == hipSuccess) e = dalloc(&d.leaf, cl * (mdas::kLineFloats / 4));
    if (e == hipSuccess && mdas::xpts_stride(D) > 0) e = dalloc(&d.xpts, cl * mdas::xpts_stride(D));
    if (e == hipSuccess) e = dalloc(&d.err, cl);
    if (e == hipSuccess) e = dalloc(&d.seed, cl);
    if (e == hipSuccess) e = dalloc(&d.flags, cl);
    if (e == hipSuccess) e = dalloc(&d.split_list, cl);
    if (e == hipSuccess) e = dalloc(&d.sdat, 2 * (cl / 2 + 1));  // splits per pass <= cap_l / 2
    if (e == hipSuccess) e = dalloc(&d.part, (size_t)mdas::kChunkBlocks);
    if (e == hipSuccess) e = dalloc(&d.part_off, (size_t)mdas::kChunkBlocks);
    if (e == hipSuccess) e = dalloc(&d.epart, (size_t)mdas::kErrBlocks);
    if (e == hipSuccess) e = dalloc(&d.cnt, 1);
    if (e == hipSuccess) e = dalloc(&d.pstats, (size_t)mdas::kMaxPassStats);
    if (e == hipSuccess) e = dalloc(&g.d_scalar, 64);
    if (e != hipSuccess) {
        free_grid(g);
        delete h;
        return fail(MDAS_ERR_HIP, std::string("device allocation failed: ") + hipGetErrorString(e));
    }
    e = hipMemset(d.cnt, 0, sizeof(mdas::DevCounters));
    if (e == hipSuccess) e = hipMemset(d.values, 0, sizeof(float) * cs * 3);
    // overflow rows start as -1 slots (a stale slot is then -1 or an index
    // written before, never garbage)
    if (e == hipSuccess && d.xpts) e = hipMemset(d.xpts, 0xff, sizeof(float) * cl * mdas::xpts_stride(D));
    if (e == hipSuccess) e = hipDeviceSynchronize();
    if (e != hipSuccess) {
        free_grid(g);
        delete h;
        return fail(MDAS_ERR_HIP, std::string("device init failed: ") + hipGetErrorString(e));
    }
    d.D = g.D;
    d.cap_s = g.cap_s;
    d.cap_l = g.cap_l;
    d.cand = g.cand;
    lcg_jump((unsigned)(g.cand * g.D), d.jmul, d.jadd);
    d.b = g.b;
    d.e = g.e;
    d.alpha = g.alpha;
    d.sX = g.sX;
    d.sY = g.sY;
    if (!g.log_path.empty()) std::ofstream(g.log_path, std::ios::trunc);
    *out = h;
    return MDAS_OK;
}

int mdas_destroy(mdas_handle h) {
    if (!h) return MDAS_OK;
    hipDeviceSynchronize();
    free_grid(h->g);
    delete h;
    return MDAS_OK;
}

int mdas_initial_sampling(mdas_handle h, void* stream) {
    if (!h) return fail(MDAS_ERR_INVALID, "null handle");
    Grid& g = h->g;
    hipStream_t st = S(stream);
    const int L0 = 1 << (g.b * g.D + 2 * g.e);
    CallEvents ev;
    const bool logging = g.logging();
    if (logging) {
        HIP_TRY(ev.create(2));
        HIP_TRY(hipEventRecord(ev.e[0], st));
    }
    // The reference zeroes every seed here (kdgrid.cu:637) so that a leaf
    // created later starts from tea4(index, 0); a split writes seed 0 into
    // the new leaf's slot, which gives the same result without the fill.
    HIP_TRY(mdas::launch_initial(g, L0, st));
    g.initialized = true;
    set_initial_extents(g);
    g.host.n_leaves = L0;
    g.host.n_samples = 4 * L0;
    g.host.new_samples = 4 * L0;
    g.host.new_nodes = 0;
    g.host.passes = 0;
    if (logging) {
        HIP_TRY(hipEventRecord(ev.e[1], st));
        HIP_TRY(hipEventSynchronize(ev.e[1]));
        log_count(g, "INITIAL SAMPLES", 4LL * L0);
        log_time(g, "INITIAL SAMPLING TIME", elapsed(ev.e[0], ev.e[1]));
    }
    return MDAS_OK;
}

int mdas_compute_errors(mdas_handle h, void* stream) {
    if (!h) return fail(MDAS_ERR_INVALID, "null handle");
    Grid& g = h->g;
    if (!g.initialized) return fail(MDAS_ERR_STATE, "InitialSampling has not been called");
    hipStream_t st = S(stream);
    CallEvents ev;
    const bool logging = g.logging();
    if (logging) {
        HIP_TRY(ev.create(2));
        HIP_TRY(hipEventRecord(ev.e[0], st));
    }
    HIP_TRY(mdas::launch_reset_errors(g, 1, st));
    HIP_TRY(mdas::launch_errors(g, st));
    if (logging) {
        HIP_TRY(hipEventRecord(ev.e[1], st));
        HIP_TRY(hipEventSynchronize(ev.e[1]));
        log_time(g, "COMPUTE ERRORS TIME", elapsed(ev.e[0], ev.e[1]));
    }
    return MDAS_OK;
}

int mdas_adaptive_sampling(mdas_handle h, void* stream) {
    if (!h) return fail(MDAS_ERR_INVALID, "null handle");
    Grid& g = h->g;
    if (!g.initialized) return fail(MDAS_ERR_STATE, "InitialSampling has not been called");
    hipStream_t st = S(stream);
    CallEvents ev;
    const bool logging = g.logging();
    if (logging) {
        HIP_TRY(ev.create(2));
        HIP_TRY(hipEventRecord(ev.e[0], st));
    }
    HIP_TRY(mdas::launch_arm(g, st));
    HIP_TRY(mdas::launch_adaptive(g, st));
    int rc = sync_counters(g, st);  // kdgrid.cu:753-759 reads the counters back
    if (rc) return rc;
    if (g.host.status != 0) {
        int s = g.host.status;
        // clear the device error so that the grid stays usable
        mdas::DevCounters c = g.host;
        c.status = 0;
        c.done = 0;
        HIP_TRY(hipMemcpy(g.dev.cnt, &c, sizeof(c), hipMemcpyHostToDevice));
        g.host = c;
        if (s == MDAS_ERR_CAPACITY)
            return fail(MDAS_ERR_CAPACITY, "adaptive pass would exceed max_samples (" +
                                               std::to_string(g.max_samples) + ")");
        return fail(s, "device error");
    }
    if (logging) {
        HIP_TRY(hipEventRecord(ev.e[1], st));
        HIP_TRY(hipEventSynchronize(ev.e[1]));
        log_count(g, "SAMPLES", g.host.new_samples);
        log_time(g, "ADAPTIVE SAMPLING TIME", elapsed(ev.e[0], ev.e[1]));
    }
    return MDAS_OK;
}

int mdas_sampling_pass(mdas_handle h, void* stream) {
    int rc = mdas_compute_errors(h, stream);
    if (rc) return rc;
    return mdas_adaptive_sampling(h, stream);
}

static int check_image(Grid& g, int w, int hgt) {
    if (w <= 0 || hgt <= 0) return fail(MDAS_ERR_INVALID, "image size must be positive");
    if ((long long)w * hgt > 0x7fffffffLL || w > 65535 || hgt > 65535)
        return fail(MDAS_ERR_INVALID, "image too large (at most 65535 pixels per side)");
    if (!g.initialized) return fail(MDAS_ERR_STATE, "InitialSampling has not been called");
    return MDAS_OK;
}

int mdas_integrate(mdas_handle h, float* pixels, int w, int hgt, void* stream) {
    if (!h || !pixels) return fail(MDAS_ERR_INVALID, "null argument");
    Grid& g = h->g;
    int rc = check_image(g, w, hgt);
    if (rc) return rc;
    hipStream_t st = S(stream);
    const bool logging = g.logging();
    const bool timed = g.timing || logging;
    CallEvents ev;
    if (timed) {
        HIP_TRY(ev.create(4));
        HIP_TRY(hipEventRecord(ev.e[0], st));
    }
    HIP_TRY(mdas::recon_prepare(g, w, hgt, st));
    HIP_TRY(mdas::launch_integrate_gather(g, pixels, w, hgt, st, ev.e[1], ev.e[2]));
    if (timed) {
        HIP_TRY(hipEventRecord(ev.e[3], st));
        HIP_TRY(hipEventSynchronize(ev.e[3]));
        g.t_recon = elapsed(ev.e[0], ev.e[3]);
        g.t_sort = elapsed(ev.e[1], ev.e[2]);
        int rc2 = sync_counters(g, st);  // the key count, for the byte model
        if (rc2) return rc2;
        if (g.host.recon_overflow)
            return fail(MDAS_ERR_CAPACITY, "reconstruction keys exceed their buffer");
        collect_kernel_times(g);
    }
    if (logging) {
        log_count(g, "TOTAL SAMPLES", g.host.n_samples);
        log_time(g, "CONSTRUCT TIME", g.t_sort);
        log_time(g, "INTEGRATE TIME", g.t_recon);
    }
    return MDAS_OK;
}

int mdas_integrate_scatter(mdas_handle h, float* pixels, int w, int hgt, void* stream) {
    if (!h || !pixels) return fail(MDAS_ERR_INVALID, "null argument");
    Grid& g = h->g;
    int rc = check_image(g, w, hgt);
    if (rc) return rc;
    HIP_TRY(mdas::launch_integrate_scatter(g, pixels, w, hgt, S(stream)));
    return MDAS_OK;
}

int mdas_convert_to_bytes(mdas_handle h, const float* pixels, unsigned char* bytes, int w, int hgt,
                          void* stream) {
    if (!pixels || !bytes) return fail(MDAS_ERR_INVALID, "null argument");
    if (w <= 0 || hgt <= 0) return fail(MDAS_ERR_INVALID, "image size must be positive");
    (void)h;
    HIP_TRY(mdas::launch_bytes(pixels, bytes, (long long)w * hgt, S(stream)));
    return MDAS_OK;
}

int mdas_sampling_density(mdas_handle h, float* host_pixels, int w, int hgt) {
    if (!h || !host_pixels) return fail(MDAS_ERR_INVALID, "null argument");
    Grid& g = h->g;
    int rc = check_image(g, w, hgt);
    if (rc) return rc;
    rc = sync_counters(g, nullptr);
    if (rc) return rc;
    const long long N = g.host.n_samples;
    std::vector<float> c((size_t)N * g.D);
    HIP_TRY(hipMemcpy(c.data(), g.dev.coords, sizeof(float) * c.size(), hipMemcpyDeviceToHost));
    memset(host_pixels, 0, sizeof(float) * 4 * (size_t)w * hgt);
    const float dens = 0.05f;
    for (long long i = 0; i < N; ++i) {
        int x = (int)(c[(size_t)i * g.D] / g.sX * (float)w);
        int y = (int)(c[(size_t)i * g.D + 1] / g.sY * (float)hgt);
        x = std::min(x, w - 1);
        y = std::min(y, hgt - 1);
        float* p = host_pixels + 4 * ((size_t)y * w + x);
        for (int k = 0; k < 4; ++k) p[k] += dens;
    }
    for (long long i = 0; i < (long long)w * hgt; ++i) host_pixels[4 * i + 3] = 1.0f;
    return MDAS_OK;
}

int mdas_download_state(mdas_handle h, float* coords, float* values, int* idx, float* box,
                        float* err, uint32_t* seed) {
    if (!h) return fail(MDAS_ERR_INVALID, "null handle");
    Grid& g = h->g;
    int rc = sync_counters(g, nullptr);
    if (rc) return rc;
    const size_t N = (size_t)g.host.n_samples, L = (size_t)g.host.n_leaves;
    const int D = g.D;
    if (coords) HIP_TRY(hipMemcpy(coords, g.dev.coords, sizeof(float) * N * D, hipMemcpyDeviceToHost));
    if (values) HIP_TRY(hipMemcpy(values, g.dev.values, sizeof(float) * N * 3, hipMemcpyDeviceToHost));
    if (err) HIP_TRY(hipMemcpy(err, g.dev.err, sizeof(float) * L, hipMemcpyDeviceToHost));
    if (seed) HIP_TRY(hipMemcpy(seed, g.dev.seed, sizeof(uint32_t) * L, hipMemcpyDeviceToHost));
    if (idx || box) {  // slots and box from each leaf's line and overflow row
        const size_t LF = mdas::kLineFloats, XS = (size_t)mdas::xpts_stride(D);
        std::vector<float> rec(L * LF), xr(L * XS);
        std::vector<unsigned char> fl(L);
        HIP_TRY(hipMemcpy(rec.data(), g.dev.leaf, sizeof(float) * rec.size(), hipMemcpyDeviceToHost));
        if (XS) HIP_TRY(hipMemcpy(xr.data(), g.dev.xpts, sizeof(float) * xr.size(), hipMemcpyDeviceToHost));
        HIP_TRY(hipMemcpy(fl.data(), g.dev.flags, L, hipMemcpyDeviceToHost));
        for (size_t i = 0; i < L; ++i) {
            const float* r = &rec[i * LF];
            if (idx) {  // slots past the count read as the reference's empty marker ~D (kdgrid.h:27-30)
                const int cnt = fl[i] & 7;
                for (int q = 0; q < 4; ++q) {
                    const int at = mdas::slot_at(D, q);
                    int v = ~D;
                    if (q < cnt) memcpy(&v, at >= 0 ? r + at : &xr[i * XS + (size_t)(-1 - at)], 4);
                    idx[4 * i + q] = v;
                }
            }
            if (box) memcpy(box + i * 2 * D, r + 4, sizeof(float) * 2 * D);
        }
    }
    return MDAS_OK;
}

int mdas_upload_state(mdas_handle h, int n, int l, const float* coords, const float* values,
                      const int* idx, const float* box, const float* err, const uint32_t* seed,
                      float max_error) {
    if (!h) return fail(MDAS_ERR_INVALID, "null handle");
    Grid& g = h->g;
    if (n < 0 || l < 0 || n > g.cap_s || l > g.cap_l) return fail(MDAS_ERR_INVALID, "state exceeds capacity");
    const int D = g.D;
    if ((idx != nullptr) != (box != nullptr))
        return fail(MDAS_ERR_INVALID, "slots and boxes are uploaded together");
    // what the kernels rely on, checked before anything is copied: a slot
    // index below n, the occupied slots first (a pass fills slots in order
    // and a split writes both children from slot 0)
    for (int i = 0; idx && i < l; ++i) {
        int c = 0;
        for (int q = 0; q < 4; ++q) {
            const int k = idx[4 * (size_t)i + q];
            if (k >= n)
                return fail(MDAS_ERR_INVALID, "leaf " + std::to_string(i) + ": sample index beyond the samples");
            if (k >= 0 && c != q)
                return fail(MDAS_ERR_INVALID, "leaf " + std::to_string(i) + ": an occupied slot after an empty one");
            c += k >= 0 ? 1 : 0;
        }
    }
    HIP_TRY(hipDeviceSynchronize());
    if (coords) HIP_TRY(hipMemcpy(g.dev.coords, coords, sizeof(float) * (size_t)n * D, hipMemcpyHostToDevice));
    if (values) HIP_TRY(hipMemcpy(g.dev.values, values, sizeof(float) * (size_t)n * 3, hipMemcpyHostToDevice));
    if (err) HIP_TRY(hipMemcpy(g.dev.err, err, sizeof(float) * (size_t)l, hipMemcpyHostToDevice));
    if (seed) {
        std::vector<uint32_t> s((size_t)g.cap_l, 0u);
        memcpy(s.data(), seed, sizeof(uint32_t) * (size_t)l);
        HIP_TRY(hipMemcpy(g.dev.seed, s.data(), sizeof(uint32_t) * s.size(), hipMemcpyHostToDevice));
    }
    if (idx) {  // lines and overflow rows with slots and box; the points follow from the coordinates
        const size_t LF = mdas::kLineFloats, XS = (size_t)mdas::xpts_stride(D);
        std::vector<float> rec((size_t)l * LF, 0.0f);
        std::vector<int> xr((size_t)l * XS, -1);
        std::vector<unsigned char> f((size_t)l);
        for (int i = 0; i < l; ++i) {
            int c = 0;
            float* r = &rec[(size_t)i * LF];
            for (int q = 0; q < 4; ++q) {
                const int k = idx[4 * (size_t)i + q];
                c += k >= 0 ? 1 : 0;
                const int at = mdas::slot_at(D, q);
                if (at >= 0) memcpy(r + at, &k, 4);
                else xr[(size_t)i * XS + (size_t)(-1 - at)] = k;
            }
            memcpy(r + 4, box + (size_t)i * 2 * D, sizeof(float) * 2 * D);
            f[i] = (unsigned char)c;
        }
        HIP_TRY(hipMemcpy(g.dev.leaf, rec.data(), sizeof(float) * rec.size(), hipMemcpyHostToDevice));
        if (XS) HIP_TRY(hipMemcpy(g.dev.xpts, xr.data(), sizeof(int) * xr.size(), hipMemcpyHostToDevice));
        HIP_TRY(hipMemcpy(g.dev.flags, f.data(), f.size(), hipMemcpyHostToDevice));
        g.ext_x = g.ext_y = 0.0;
        for (int i = 0; i < l; ++i) {
            const float* bx = box + (size_t)i * 2 * D;
            g.ext_x = std::max(g.ext_x, (double)(bx[D] - bx[0]));
            g.ext_y = std::max(g.ext_y, (double)(bx[D + 1] - bx[1]));
        }
    }
    if (idx || coords) {
        HIP_TRY(mdas::launch_build_pts(g, l, 0));
        HIP_TRY(hipDeviceSynchronize());
    }
    mdas::DevCounters c{};
    c.n_samples = n;
    c.n_leaves = l;
    c.limit = 0x7fffffff;
    memcpy(&c.gmax_bits, &max_error, 4);
    HIP_TRY(hipMemcpy(g.dev.cnt, &c, sizeof(c), hipMemcpyHostToDevice));
    g.host = c;
    g.initialized = true;
    return MDAS_OK;
}

int mdas_set_values(mdas_handle h, int offset, int count, const float* hv) {
    if (!h || !hv) return fail(MDAS_ERR_INVALID, "null argument");
    Grid& g = h->g;
    if (offset < 0 || count < 0 || (long long)offset + count > g.cap_s)
        return fail(MDAS_ERR_INVALID, "range exceeds capacity");
    HIP_TRY(hipMemcpy(g.dev.values + (size_t)offset * 3, hv, sizeof(float) * 3 * (size_t)count,
                      hipMemcpyHostToDevice));
    return MDAS_OK;
}

int mdas_get_recon_keys(mdas_handle h, uint64_t* out, long long capacity, long long* n) {
    if (!h) return fail(MDAS_ERR_INVALID, "null handle");
    Grid& g = h->g;
    int rc = sync_counters(g, nullptr);
    if (rc) return rc;
    HIP_TRY(hipDeviceSynchronize());
    if (g.host.recon_overflow) return fail(MDAS_ERR_CAPACITY, "reconstruction keys exceeded their buffer");
    const long long nk = g.host.n_keys;
    if (n) *n = nk;
    if (out && g.keys_sorted) {
        long long m = std::min(capacity, nk);
        HIP_TRY(hipMemcpy(out, g.keys_sorted, sizeof(uint64_t) * (size_t)m, hipMemcpyDeviceToHost));
    }
    return MDAS_OK;
}

int mdas_validate(mdas_handle h, int* valid) {
    if (!h || !valid) return fail(MDAS_ERR_INVALID, "null argument");
    Grid& g = h->g;
    *valid = 0;
    int rc = sync_counters(g, nullptr);
    if (rc) return rc;
    const int D = g.D;
    const long long N = g.host.n_samples, L = g.host.n_leaves;
    std::vector<float> coords((size_t)N * D), box((size_t)L * 2 * D);
    std::vector<int> idx((size_t)L * 4);
    rc = mdas_download_state(h, coords.data(), nullptr, idx.data(), box.data(), nullptr, nullptr);
    if (rc) return rc;
    // the reference's four conditions (kdgrid.cu:876-956) and a slot index
    // below N (the reference would read past its arrays)
    bool ok = true;
    std::vector<int> hist((size_t)N, 0);
    double vol = 0.0;
    long long total = 0;
    for (long long i = 0; i < L; ++i) {
        const float* mn = &box[(size_t)i * 2 * D];
        const float* mx = mn + D;
        for (int q = 0; q < 4; ++q) {
            const int k = idx[(size_t)i * 4 + q];
            if (k < 0) continue;  // an empty slot
            if (k >= N) {
                ok = false;
                continue;
            }
            for (int d = 0; d < D; ++d) {  // aabb.h:57-62 Contains
                const float c = coords[(size_t)k * D + d];
                if (c < mn[d] || c > mx[d]) ok = false;
            }
            hist[k]++;
            ++total;
        }
        double v = 1.0;
        for (int d = 0; d < D; ++d) v *= (double)(mx[d] - mn[d]);
        vol += v;
    }
    const double root = (double)g.sX * (double)g.sY;
    if (std::fabs(vol - root) > 1.0e-2 * root) ok = false;  // kdgrid.cu:909-914
    if (total != N) ok = false;                              // kdgrid.cu:916-920
    for (long long k = 0; k < N && ok; ++k)
        if (hist[k] != 1) ok = false;                        // kdgrid.cu:922-956
    *valid = ok ? 1 : 0;
    return MDAS_OK;
}

float* mdas_get_sample_coordinates(mdas_handle h) { return h ? h->g.dev.coords : nullptr; }
float* mdas_get_sample_values(mdas_handle h) { return h ? h->g.dev.values : nullptr; }
int mdas_get_dim(mdas_handle h) { return h ? h->g.D : 0; }
int mdas_get_candidates_num(mdas_handle h) { return h ? h->g.cand : 0; }
int mdas_get_max_node_size(mdas_handle h) { return h ? 4 : 0; }
int mdas_get_new_samples(mdas_handle h) { return h ? h->g.host.new_samples : 0; }
int mdas_get_new_nodes(mdas_handle h) { return h ? h->g.host.new_nodes : 0; }
int mdas_get_number_of_samples(mdas_handle h) { return h ? h->g.host.n_samples : 0; }
int mdas_get_number_of_nodes(mdas_handle h) { return h ? h->g.host.n_leaves : 0; }
int mdas_get_max_samples(mdas_handle h) { return h ? (int)h->g.max_samples : 0; }
float mdas_get_scale_x(mdas_handle h) { return h ? h->g.sX : 0.0f; }
float mdas_get_scale_y(mdas_handle h) { return h ? h->g.sY : 0.0f; }
float mdas_get_max_error(mdas_handle h) {
    if (!h) return 0.0f;
    if (sync_counters(h->g, nullptr)) return 0.0f;
    return h->g.gmax_host;
}
int mdas_get_passes(mdas_handle h) { return h ? h->g.passes : 0; }

// ------------------------------------------------------------------ scenes --
int mdas_scene_create(int kind, uint32_t seed, int dim, float sx, float sy, mdas_scene_handle* out) {
    if (!out) return fail(MDAS_ERR_INVALID, "null argument");
    *out = nullptr;
    if (kind < 0 || kind > 5) return fail(MDAS_ERR_INVALID, "unknown scene kind");
    static const int need[6] = {2, 3, 4, 5, 6, 2};
    if (dim < need[kind] || dim > 6) return fail(MDAS_ERR_INVALID, "scene needs more dimensions");
    mdas_scene* s = new mdas_scene;
    mscene::make_scene(&s->host, kind, seed, dim, sx, sy);
    // the device copy: the parameters followed by the sphere groups
    std::vector<mscene::SceneBlock> blk(1);
    blk[0].p = s->host;
    mscene::build_accel(s->host, &blk[0].accel);
    hipError_t e = hipMalloc(&s->dev, sizeof(mscene::SceneBlock));
    if (e == hipSuccess)
        e = hipMemcpy(s->dev, blk.data(), sizeof(mscene::SceneBlock), hipMemcpyHostToDevice);
    if (e != hipSuccess) {
        if (s->dev) hipFree(s->dev);
        delete s;
        return fail(MDAS_ERR_HIP, std::string("scene upload failed: ") + hipGetErrorString(e));
    }
    *out = s;
    return MDAS_OK;
}

int mdas_scene_destroy(mdas_scene_handle s) {
    if (!s) return MDAS_OK;
    hipDeviceSynchronize();
    if (s->dev) hipFree(s->dev);
    delete s;
    return MDAS_OK;
}

long long mdas_scene_params(mdas_scene_handle s, void* out, long long cap) {
    if (!s) return 0;
    const long long n = (long long)sizeof(mscene::Params);
    if (out && cap >= n) memcpy(out, &s->host, (size_t)n);
    return n;
}

int mdas_eval_integrand(mdas_handle h, mdas_scene_handle s, int offset, int count, void* stream) {
    if (!h || !s) return fail(MDAS_ERR_INVALID, "null argument");
    Grid& g = h->g;
    if (s->host.dim != g.D) return fail(MDAS_ERR_INVALID, "scene and grid dimensions differ");
    if (offset < 0 || count < 0 || (long long)offset + count > g.cap_s)
        return fail(MDAS_ERR_INVALID, "sample range exceeds capacity");
    if (count == 0) return MDAS_OK;
    HIP_TRY(mdas::launch_eval(g, s->dev, offset, count, false, S(stream)));
    return MDAS_OK;
}

int mdas_eval_new_samples(mdas_handle h, mdas_scene_handle s, void* stream) {
    if (!h || !s) return fail(MDAS_ERR_INVALID, "null argument");
    Grid& g = h->g;
    if (s->host.dim != g.D) return fail(MDAS_ERR_INVALID, "scene and grid dimensions differ");
    HIP_TRY(mdas::launch_eval(g, s->dev, 0, 0, true, S(stream)));
    return MDAS_OK;
}

int mdas_render_qmc(mdas_scene_handle s, int w, int hgt, int spp, int first_sample, uint32_t frame,
                    int fp64, void* out, void* stream) {
    if (!s || !out) return fail(MDAS_ERR_INVALID, "null argument");
    if (w <= 0 || hgt <= 0 || spp <= 0 || first_sample < 0)
        return fail(MDAS_ERR_INVALID, "image size and spp must be positive");
    HIP_TRY(mdas::launch_qmc_image(s->dev, w, hgt, spp, first_sample, frame, fp64 != 0, out, S(stream)));
    return MDAS_OK;
}

// ---------------------------------------------------------------- the loop --
int mdas_set_log_sink(mdas_handle h, mdas_log_fn fn, void* user) {
    if (!h) return fail(MDAS_ERR_INVALID, "null handle");
    h->g.log_fn = fn;
    h->g.log_user = user;
    return MDAS_OK;
}

int mdas_set_timing(mdas_handle h, int enable) {
    if (!h) return fail(MDAS_ERR_INVALID, "null handle");
    h->g.timing = enable != 0;
    return MDAS_OK;
}

int mdas_run(mdas_handle h, mdas_scene_handle s, long long target, void* stream) {
    if (!h || !s) return fail(MDAS_ERR_INVALID, "null argument");
    Grid& g = h->g;
    if (s->host.dim != g.D) return fail(MDAS_ERR_INVALID, "scene and grid dimensions differ");
    if (target < 0) return fail(MDAS_ERR_INVALID, "target must be >= 0");
    hipStream_t st = S(stream);
    const int L0 = 1 << (g.b * g.D + 2 * g.e);
    const long long limit = std::min<long long>(target, g.max_samples);
    const int B = std::max(1, g.passes_per_sync);
    // the log needs per-pass times: logging times the run like mdas_set_timing
    const bool logging = g.logging();
    const bool timing_was = g.timing;
    if (logging) g.timing = true;
    struct Restore {
        Grid& g;
        bool v;
        ~Restore() { g.timing = v; }
    } restore{g, timing_was};
    g.kstats.clear();
    g.krecs.clear();
    g.evnext = 0;
    g.t_recon = g.t_sort = 0.0;
    g.pstats_host.clear();
    mdas::kbeg(g, st, "k_initial");
    HIP_TRY(mdas::launch_initial(g, L0, st));
    HIP_TRY(mdas::launch_set_limit(g, (int)limit, st));
    mdas::kend(g, st);
    g.initialized = true;
    set_initial_extents(g);
    mdas::kbeg(g, st, "k_eval");
    HIP_TRY(mdas::launch_eval(g, s->dev, 0, 4 * L0, false, st));
    mdas::kend(g, st);
    long long last_n = -1;  // samples after the last block of passes
    // every timed launch in order (name, ms); folded after each block, so the
    // event pool holds one block's events however many passes the run takes
    std::vector<std::pair<std::string, double>> seq;
    struct LoopMode {  // the pass kernels of this loop skip the stale markers
        mdas::DevGrid& d;
        explicit LoopMode(mdas::DevGrid& x) : d(x) { d.in_loop = 1; }
        ~LoopMode() { d.in_loop = 0; }
    } loop_mode{g.dev};
    for (;;) {
        for (int b = 0; b < B; ++b) {
            mdas::kbeg(g, st, "k_errors");
            HIP_TRY(g.learn_mode ? mdas::launch_errors_learned(g, st)
                                 : mdas::launch_errors_stale(g, st));
            mdas::kend(g, st);
            mdas::kbeg(g, st, "k_select");
            HIP_TRY(mdas::launch_select(g, st));
            mdas::kend(g, st);
            mdas::kbeg(g, st, "k_scan_counts");
            HIP_TRY(mdas::launch_scan(g, st));
            mdas::kend(g, st);
            mdas::kbeg(g, st, "k_place");
            HIP_TRY(mdas::launch_place(g, st));
            mdas::kend(g, st);
            mdas::kbeg(g, st, "k_eval");
            HIP_TRY(mdas::launch_eval(g, s->dev, 0, 0, true, st));
            mdas::kend(g, st);
        }
        int rc = sync_counters(g, st);
        if (rc == MDAS_OK) rc = mdas::drain_pass_stats(g, false);
        if (rc) return rc;
        if (g.timing) collect_kernel_times(g, logging ? &seq : nullptr);
        if (g.host.done || g.host.status) break;
        // every pass selects the leaf of maximum error (its ratio is 1), so
        // each block adds samples and the loop ends at the limit; a block that
        // adds none means no leaf passed the selection test: the maximum error
        // is not finite (e.g. a NaN integrand) or zero (scores that underflow)
        if (g.host.n_samples == last_n)
            return fail(MDAS_ERR_STATE, "a block of passes added no sample (no leaf selected: "
                                        "non-finite or zero maximum error)");
        last_n = g.host.n_samples;
    }
    mdas::kbeg(g, st, "k_mark_last");
    HIP_TRY(mdas::launch_mark_last_children(g, st));
    mdas::kend(g, st);
    g.passes = g.host.passes;
    collect_kernel_times(g, logging ? &seq : nullptr);
    {
        int rc = mdas::drain_pass_stats(g, true);
        if (rc) return rc;
    }
    if (logging && seq.size() >= 2) {
        // the app's and the library's keys in the order they write them
        // (kdgrid.cu:670-671, 713, 769-770; optixMotionBlur.cpp:774, 788):
        // the initial grid, its integrand launch, then per pass the error
        // estimate, the new samples, the adaptive step and the integrand
        log_count(g, "INITIAL SAMPLES", 4LL * L0);
        log_time(g, "INITIAL SAMPLING TIME", seq[0].second);
        log_time(g, "TRACE TIME", seq[1].second);
        size_t k = 2;
        for (int p = 0; p < g.passes && k + 4 < seq.size(); ++p, k += 5) {
            log_time(g, "COMPUTE ERRORS TIME", seq[k].second);
            log_count(g, "SAMPLES", p < (int)g.pstats_host.size() ? g.pstats_host[p].sel : 0);
            log_time(g, "ADAPTIVE SAMPLING TIME", seq[k + 1].second + seq[k + 2].second + seq[k + 3].second);
            log_time(g, "TRACE TIME", seq[k + 4].second);
        }
    }
    if (g.host.status) {
        int s2 = g.host.status;
        return fail(s2, s2 == MDAS_ERR_CAPACITY ? "adaptive pass would exceed max_samples" : "device error");
    }
    return MDAS_OK;
}

int mdas_get_pass_stats(mdas_handle h, int* out, int max_rows) {
    if (!h || !out) return fail(MDAS_ERR_INVALID, "null argument");
    Grid& g = h->g;
    int n = std::min<int>((int)g.pstats_host.size(), max_rows);
    for (int i = 0; i < n; ++i) {
        const mdas::PassStat& p = g.pstats_host[i];
        const int row[8] = {p.L, p.N, p.sel, p.split, p.selcnt, p.stale, p.stale_vals, 0};
        memcpy(out + 8 * i, row, sizeof(row));
    }
    return n;
}

// Algorithmic bytes of each kernel over the last run + reconstruction: the
// bytes the algorithm must move with this layout, each counted once: FP32
// coordinates (4*D per sample) and values (12), a leaf's slots and box
// (SB = 16 + 8*D bytes of its line) and its points (4*D each, in the line or
// the overflow row), 4-byte errors and LCG states, 1-byte flags, 8-byte keys,
// 32-byte reconstruction records, float4 pixels.  Line padding is not
// counted.
static double kernel_bytes(const Grid& g, const std::string& k) {
    const double D = g.D;
    const double SB = 16.0 + 8.0 * D;
    const double L0 = (double)(1LL << (g.b * g.D + 2 * g.e));
    double b = 0.0;
    if (k == "k_initial")  // coordinates, slots + box, the four points, seed, error, flags
        return L0 * (16.0 * D + SB + 16.0 * D + 4.0 + 4.0 + 1.0);
    if (k == "k_eval") {
        double n = (double)g.host.n_samples;
        return n * (4.0 * D + 12.0);
    }
    for (const auto& p : g.pstats_host) {
        if (k == "k_errors") {
            if (&p == &g.pstats_host[0])  // first pass: every leaf's slots + box, values, error
                b += (double)p.stale * (SB + 4.0) + 12.0 * p.stale_vals;
            else  // per split: list entry, split entry, two errors; the values
                b += 0.5 * (double)p.stale * (4.0 + 32.0 + 8.0) + 12.0 * p.stale_vals;
        }
        else if (k == "k_select")  // error, state, flags read; flags written
            b += 10.0 * p.L;
        else if (k == "k_scan_counts")
            b += (16.0 + 8.0) * g.chunk_blocks;
        else if (k == "k_place")  // compaction and placement
            b += 10.0 * p.L +                                    // flags, states read and written
                 (double)p.sel * (SB + 4.0 * D) +                // slots + box, the new sample
                 4.0 * D * p.selcnt +                            // the leaf's points
                 (double)(p.sel - p.split) * (4.0 + 4.0 * D) +   // slot entry + point of an insert
                 (double)p.split * ((SB + 8.0 * D) + (SB + 12.0 * D) +  // both children's lines
                                    4.0 + 1.0 + 4.0 + 32.0);     // list, flags, state, split entry
    }
    const double L = (double)g.host.n_leaves, N = (double)g.host.n_samples;
    const double K = (double)g.host.n_keys;
    if (k == "k_recon_prep") return L * (SB + 32.0) + 12.0 * N;
    if (k == "k_scan_keys") return (16.0 + 8.0) * g.chunk_blocks;
    if (k == "k_recon_emit") return L * 16.0 + K * 8.0;
    if (k == "k_rs_hist") return K * 8.0 * g.recon_sort_passes;
    if (k == "k_rs_scan") return 3.0 * 4.0 * (1 << g.recon_sort_bits) * 1024 * g.recon_sort_passes;
    if (k == "k_rs_scatter") return K * 16.0 * g.recon_sort_passes;
    if (k == "k_tile_ranges") return K * 8.0;
    if (k == "k_gather") return K * (8.0 + 32.0) + 16.0 * (double)g.recon_pixels;
    return b;
}

int mdas_kernel_stats(mdas_handle h, int index, char* name, int name_cap, int* launches,
                      double* total_ms, double* bytes) {
    if (!h) return fail(MDAS_ERR_INVALID, "null handle");
    Grid& g = h->g;
    const int n = (int)g.kstats.size();
    if (index < 0 || index >= n) return n;
    const auto& k = g.kstats[(size_t)index];
    if (name && name_cap > 0) {
        strncpy(name, k.name.c_str(), (size_t)name_cap - 1);
        name[name_cap - 1] = 0;
    }
    if (launches) *launches = k.launches;
    if (total_ms) *total_ms = k.ms;
    if (bytes) *bytes = kernel_bytes(g, k.name);
    return n;
}

int mdas_get_phase_times(mdas_handle h, mdas_phase_times* o) {
    if (!h || !o) return fail(MDAS_ERR_INVALID, "null argument");
    Grid& g = h->g;
    memset(o, 0, sizeof(*o));
    o->initial_ms = g.t_initial;
    o->errors_ms = g.t_errors;
    o->adaptive_ms = g.t_adaptive;
    o->integrand_ms = g.t_integrand;
    o->recon_ms = g.t_recon;
    o->sort_ms = g.t_sort;
    o->passes = g.passes;
    o->final_samples = g.host.n_samples;
    o->final_leaves = g.host.n_leaves;
    o->bytes_initial = kernel_bytes(g, "k_initial");
    o->bytes_errors = kernel_bytes(g, "k_errors");
    o->bytes_adaptive = kernel_bytes(g, "k_select") + kernel_bytes(g, "k_scan_counts") +
                        kernel_bytes(g, "k_place");
    double br = 0.0;
    for (const char* k : {"k_recon_prep", "k_scan_keys", "k_recon_emit", "k_rs_hist", "k_rs_scan",
                          "k_rs_scatter", "k_tile_ranges", "k_gather"})
        br += kernel_bytes(g, k);
    o->bytes_recon = br;
    return MDAS_OK;
}

// Times the reconstruction of the current state with every candidate tile
// shape (after one untimed warm-up each) and keeps the fastest.  Results are
// identical for every shape; only the speed differs.
int mdas_autotune_recon(mdas_handle h, float* pixels, int w, int hgt, void* stream, int* best_tw,
                        int* best_th) {
    if (!h || !pixels) return fail(MDAS_ERR_INVALID, "null argument");
    Grid& g = h->g;
    int rc = check_image(g, w, hgt);
    if (rc) return rc;
    static const int cand[][2] = {{8, 8}, {16, 4}, {4, 4}, {8, 2}, {16, 1}};
    hipStream_t st = S(stream);
    // timing off while tuning; on any return the grid keeps its timing flag
    // and the best tile found so far (its own tile if none was timed)
    struct Restore {
        Grid& g;
        bool timing;
        int tw, th;
        ~Restore() {
            g.timing = timing;
            g.tile_w = tw;
            g.tile_h = th;
        }
    } restore{g, g.timing, g.tile_w, g.tile_h};
    g.timing = false;
    CallEvents ev;
    HIP_TRY(ev.create(2));
    float best = 1e30f;
    for (const auto& c : cand) {
        g.tile_w = c[0];
        g.tile_h = c[1];
        HIP_TRY(mdas::recon_prepare(g, w, hgt, st));
        HIP_TRY(mdas::launch_integrate_gather(g, pixels, w, hgt, st, nullptr, nullptr));
        HIP_TRY(hipEventRecord(ev.e[0], st));
        HIP_TRY(mdas::recon_prepare(g, w, hgt, st));
        HIP_TRY(mdas::launch_integrate_gather(g, pixels, w, hgt, st, nullptr, nullptr));
        HIP_TRY(hipEventRecord(ev.e[1], st));
        HIP_TRY(hipEventSynchronize(ev.e[1]));
        const float ms = elapsed(ev.e[0], ev.e[1]);
        if (ms < best) {
            best = ms;
            restore.tw = c[0];
            restore.th = c[1];
        }
    }
    if (best_tw) *best_tw = restore.tw;
    if (best_th) *best_th = restore.th;
    return MDAS_OK;
}

int mdas_set_launch_shape(mdas_handle h, int chunk_blocks, int err_blocks) {
    if (!h) return fail(MDAS_ERR_INVALID, "null handle");
    if (chunk_blocks < 64 || chunk_blocks > mdas::kChunkBlocks || err_blocks < 64 ||
        err_blocks > mdas::kErrBlocks)
        return fail(MDAS_ERR_INVALID, "launch shape out of range");
    HIP_TRY(hipDeviceSynchronize());
    h->g.chunk_blocks = chunk_blocks;
    h->g.err_blocks = err_blocks;
    return MDAS_OK;
}

int mdas_set_recon_tile(mdas_handle h, int tw, int th) {
    if (!h) return fail(MDAS_ERR_INVALID, "null handle");
    if (tw <= 0 || th <= 0 || (tw * th != 64 && tw * th != 16))
        return fail(MDAS_ERR_INVALID, "a reconstruction tile holds 16 or 64 pixels");
    h->g.tile_w = tw;
    h->g.tile_h = th;
    return MDAS_OK;
}

}  // extern "C"
