// Internal state of one MDAS grid (one mdas_handle) and the kernel launchers.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <string>
#include <vector>

namespace mdas {

// Device-resident counters of one grid.  Every kernel reads N / L from here,
// so a whole sampling loop can be enqueued without host round trips.
struct DevCounters {
    int n_samples;       // N
    int n_leaves;        // L
    int n_samples_old;   // N before the last adaptive pass (new slice = [old, N))
    int n_leaves_old;    // L before the last adaptive pass
    unsigned gmax_bits;  // global max leaf error (float bits, errors are >= 0)
    int new_samples;
    int new_nodes;
    int done;            // stopping rule reached or error: pass kernels return early
    int status;          // 0 or an MDAS_ERR_* code raised on the device
    int passes;          // adaptive passes performed
    int limit;           // stopping limit min(target, max_samples)
    int n_keys;          // reconstruction keys emitted
    int sel_cnt;         // samples held by the leaves selected in this pass
    int stale;           // leaves whose error was recomputed in this pass
    int stale_vals;      // sample values read for them
    unsigned carry_bits; // max error of the leaves the last adaptive pass did not split
    int recon_overflow;  // keys counted by the last reconstruction when they exceeded the
                         //   key buffer (then nothing was emitted), else 0
    int guard_cap;       // device-side loop: no pass starts once N + L > guard_cap (the
                         //   sample capacity when the target is at or beyond it; 0 = no
                         //   guard: a smaller target, standalone calls)
};

// Per-pass statistics, one row per adaptive pass (for algorithmic byte counts):
// leaves and samples before the pass, selected and split leaves, the samples
// held by the selected leaves, the stale leaves whose error was recomputed
// and the sample values those read.
struct PassStat {
    int L, N, sel, split, selcnt, stale, stale_vals, pad;
};

// One timed kernel launch (HIP events on the launching stream).
struct KRec {
    const char* name;
    hipEvent_t a, b;
};

static const int kChunkBlocks = 32768;  // max workgroups of the chunked leaf kernels
static const int kErrBlocks = 8192;     // max workgroups of the error kernels
static const int kMaxPassStats = 4096;  // device ring of per-pass rows (a power of two); the
                                        // host drains it before it wraps

// Leaf lines.  Each leaf owns one aligned 128-byte line (kLineFloats floats):
// its sample slots, the box (mn[0..D-1], mx[0..D-1]) and the points of its
// first line_pts(D) slots, D floats each, so that the placement reads and
// writes a selected leaf in one line.  The points of the remaining slots (D = 5:
// slot 3; D = 6: slots 2 and 3) sit in a per-leaf overflow row of
// xpts_stride(D) floats (32 / 64 bytes), touched only by the last insert and
// the split.  Float offsets in a line: slots at slot_at(D, q) (0..3 for
// D <= 4), mn at 4, mx at 4 + D, point q at 4 + 2 D + q D.
//
// A leaf's sample count is the low 3 bits of its flags byte; slots at or past
// it hold stale indices and are never read.  For D >= 5 the slots that an
// insert fills sit in the same 32-byte sector as their points, so that an
// insert writes one sector: D = 5 keeps slot 2 at line float 29 (the line's
// last sector, with point 2 at 24..28) and slot 3 at overflow float 5 (point 3
// at 0..4); D = 6 keeps slots 2 and 3 at overflow floats 6 and 14 (points 2
// and 3 at 0..5 and 8..13).  Slots 0 and 1 are at line floats 0 and 1.
static const int kLineFloats = 32;
__host__ __device__ constexpr int line_pts(int D) { return D <= 4 ? 4 : (D == 5 ? 3 : 2); }
__host__ __device__ constexpr int line_pt(int D, int q) { return 4 + 2 * D + q * D; }
__host__ __device__ constexpr int xpts_stride(int D) { return D <= 4 ? 0 : (D == 5 ? 8 : 16); }
// overflow-row float of point q (q >= line_pts(D))
__host__ __device__ constexpr int xpt_at(int D, int q) { return D == 5 ? 0 : (q == 2 ? 0 : 8); }
// slot q: a line float (>= 0), or overflow float f encoded as -1 - f
__host__ __device__ constexpr int slot_at(int D, int q) {
    return D <= 4 || q < 2 ? q : (D == 5 ? (q == 2 ? 29 : -1 - 5) : (q == 2 ? -1 - 6 : -1 - 14));
}
// float4s of a line holding the slots and the box
__host__ __device__ constexpr int line_box_quads(int D) { return (4 + 2 * D + 3) / 4; }

// Plain-pointer view of the device state passed to kernels by value.
struct DevGrid {
    float* coords;         // [cap_s * D]  AoS (the integrand contract)
    float* values;         // [cap_s * 3]  float3 AoS
    float4* leaf;          // [cap_l * 8]  leaf lines: slots, box, first points
    float* xpts;           // [cap_l * xpts_stride(D)]  the points of the other slots (D >= 5)
    float* err;            // [cap_l]      < 0 = stale
    unsigned* seed;        // [cap_l]      LCG state, 0 = unset
    unsigned char* flags;  // [cap_l]      bits 0-2 count, 3 selected, 4 split
    int* split_list;       // [cap_l]      leaves split by the last pass, in leaf order
    int4* sdat;            // [cap_l/2+1][2] per split of the last pass: the children's sample
                           //   indices (left 2, right 3) and box volumes, in split order
    int4* sext;            // [cap_l/2+1][2] learned mode only (else null): per split the
                           //   parent's box extents with the left child's at the split axis,
                           //   the right child's extent at the axis, the axis
    int4* part;            // [kChunkBlocks] per-workgroup (selected, split, samples, carry)
    int2* part_off;        // [kChunkBlocks] exclusive prefix of (selected, split)
    uint4* epart;          // [kErrBlocks] per-workgroup (max error bits, stale, values)
    unsigned char* rcnt;   // [cap_l]      reconstruction tiles per leaf
    DevCounters* cnt;
    PassStat* pstats;
    int D;
    int cap_s, cap_l;
    int cand;
    unsigned jmul, jadd;  // LCG jump over candidatesNum * D draws
    int in_loop;                // 1 inside the device-side loop: the placement leaves the
                                //   children's stale markers to the next error pass
    int b, e;
    float alpha, sX, sY;
};

// Buffers of one reconstruction: a 32-byte record per leaf, the keys (two
// buffers for the sort), every tile's key range and the sort's histograms.
struct ReconBufs {
    float4* rrec = nullptr;                 // [rrec_cap * 2]
    long long rrec_cap = 0;                 // leaves the records can hold
    unsigned long long* keys = nullptr;     // [keys_cap]
    unsigned long long* keys_alt = nullptr; // [keys_cap]
    long long keys_cap = 0;
    int* tile_start = nullptr;              // [tiles_cap]
    int* tile_end = nullptr;                // [tiles_cap]
    long long tiles_cap = 0;
    unsigned* rs_hist = nullptr;            // [kSortHistWords]
};
static const long long kSortHistWords = 1024ll * 1024 + 1024;  // 10-bit digits x 1024 blocks + sums

// One MLP of the learned proposal (16 -> 64 -> 64 -> 1).  p is the FP32
// parameter blob (layout in mlp_device.hpp); the others are the bf16 forms
// the forward reads (W1 packed [64][32] with zero k >= 16, W2 [64][64]) and
// the bf16-rounded W2 and w3 as floats.
struct MlpDev {
    const float* p;
    const unsigned short* w1p;
    const unsigned short* w2p;
    const float* w2r;
    const float* w3r;
};

static const int kFeat = 16;  // features per leaf

// Learned-proposal state of one grid; unused by the plain method.
struct LearnDev {
    float* feat;         // [cap_l][16] features of each leaf at its last error estimate
    float* perr;         // [cap_l] plain error of each leaf (err holds the score)
    float* X;            // [cap_rows][16] collected rows: features of split leaves
    float* Y;            // [cap_rows] their labels; X == null: no collection
    long long* rows;     // rows offered so far (index >= cap_rows: dropped)
    long long cap_rows;
    MlpDev mlp;          // mlp.p == null: the score is the plain error
    float ylo, yhi;      // clamp of the network output
    int label_mode;      // 0: relative error reduction, score e 2^y; 1: image change, score 2^y
    float kappa_e, kappa_y;  // score = e^kappa_e * 2^(kappa_y * y)
    int keep;            // 1: full feature rows and plain errors; 0: only each leaf's depth
};

struct Grid {
    int D = 0;
    int cand = 4, b = 0, e = 0;
    float alpha = 0.25f, sX = 1.0f, sY = 1.0f;
    int cap_s = 0, cap_l = 0;
    long long max_samples = 0;
    bool initialized = false;
    DevGrid dev{};
    DevCounters host{};  // mirror after the last synchronising call
    float gmax_host = 0.0f;
    std::string log_path;
    // log sink (mdas_set_log_sink): each "KEY", value pair the reference
    // writes to its std::ofstream; is_count marks integral values
    void (*log_fn)(void* user, const char* key, double value, int is_count) = nullptr;
    void* log_user = nullptr;
    bool logging() const { return !log_path.empty() || log_fn != nullptr; }
    // reconstruction buffers (allocated at first use, grown with the tree)
    ReconBufs rb;
    unsigned long long* keys_sorted = nullptr;  // rb.keys or rb.keys_alt: the last sort's result
    int tile_w = 8, tile_h = 8;  // reconstruction tile (mdas_set_recon_tile / autotune)
    bool recon_bounded = false;   // the key buffers hold the bound of the last reservation
    // largest leaf extent along x and y (sample-space units): the initial cell's,
    // which splits only shrink, or the largest of an uploaded state's boxes
    double ext_x = 0.0, ext_y = 0.0;
    int* d_scalar = nullptr;      // small device scratch
    // timing: per-kernel events when enabled
    bool timing = false;
    std::vector<KRec> krecs;
    std::vector<hipEvent_t> evpool;
    size_t evnext = 0;
    struct KStat {
        std::string name;
        int launches;
        double ms;
    };
    std::vector<KStat> kstats;
    double t_initial = 0, t_errors = 0, t_adaptive = 0, t_integrand = 0, t_recon = 0, t_sort = 0;
    int passes = 0;
    int passes_per_sync = 4;
    std::vector<PassStat> pstats_host;
    long long recon_pixels = 0;
    int recon_sort_passes = 0;
    int recon_sort_bits = 8;  // digit width of the last reconstruction sort
    // launch shape of the per-pass kernels (mdas_set_launch_shape)
    int chunk_blocks = 4096;   // leaf chunks of selection / compaction / reconstruction
    int err_blocks = 1024;     // workgroups of the error pass
    // learned proposal (mode bits: 1 collect rows, 2 score with the MLP)
    int learn_mode = 0;
    LearnDev learn{};
};

// Kernel timing brackets: no-ops unless g.timing.
inline hipEvent_t pool_event(Grid& g) {
    if (g.evnext == g.evpool.size()) {
        hipEvent_t e = nullptr;
        if (hipEventCreate(&e) != hipSuccess) return nullptr;
        g.evpool.push_back(e);
    }
    return g.evpool[g.evnext++];
}
inline void kbeg(Grid& g, hipStream_t st, const char* name) {
    if (!g.timing) return;
    KRec r{name, pool_event(g), pool_event(g)};
    if (r.a) (void)hipEventRecord(r.a, st);
    g.krecs.push_back(r);
}
inline void kend(Grid& g, hipStream_t st) {
    if (!g.timing || g.krecs.empty()) return;
    if (g.krecs.back().b) (void)hipEventRecord(g.krecs.back().b, st);
}

// Launchers (k_sampling.hip)
hipError_t launch_initial(Grid& g, int L0, hipStream_t st);
hipError_t launch_reset_errors(Grid& g, int standalone, hipStream_t st);
hipError_t launch_errors(Grid& g, hipStream_t st);
hipError_t launch_errors_stale(Grid& g, hipStream_t st);
hipError_t launch_adaptive(Grid& g, hipStream_t st);
hipError_t launch_select(Grid& g, hipStream_t st);
hipError_t launch_scan(Grid& g, hipStream_t st);
hipError_t launch_place(Grid& g, hipStream_t st);
hipError_t launch_set_limit(Grid& g, int limit, hipStream_t st);
hipError_t launch_arm(Grid& g, hipStream_t st);
hipError_t launch_build_pts(Grid& g, int n_leaves, hipStream_t st);
hipError_t launch_mark_last_children(Grid& g, hipStream_t st);
hipError_t launch_mark_last_children_b(const DevGrid* grids, int n, hipStream_t st);
hipError_t launch_fill_i32(int* p, int v, long long n, hipStream_t st);
hipError_t launch_fill_f32(float* p, float v, long long n, hipStream_t st);

// Batched forms over a device array of n grids of dimension D (k_sampling.hip)
hipError_t launch_initial_b(int D, const DevGrid* grids, int n, int blocks, const int* limits,
                            hipStream_t st);
// learns: device array of n LearnDev (the learned error pass) or null (plain)
hipError_t launch_pass_b(int D, const DevGrid* grids, int n, int blocks, hipStream_t st,
                         const LearnDev* learns = nullptr);

// Integrand (k_integrand.hip)
hipError_t launch_eval(Grid& g, const void* dparams, int first, int count, bool dev_range,
                       hipStream_t st);
// Batched integrand over the new slice of every grid; params[i] is image i's
// device parameter block.
hipError_t launch_eval_b(const DevGrid* grids, const void* const* params, int n, int blocks,
                         hipStream_t st);
// Per-pixel QMC image (float) or FP64 ground truth (double), float4/double4 out.
hipError_t launch_qmc_image(const void* dparams, int W, int H, int spp, int first_sample,
                            unsigned frame, bool fp64, void* out, hipStream_t st);

// Reconstruction (k_recon.hip)
hipError_t launch_integrate_gather(Grid& g, float* pixels, int W, int H, hipStream_t st,
                                   hipEvent_t ev_sort0, hipEvent_t ev_sort1);
// The whole reconstruction on the given buffers (sized by the caller for the
// bound; no allocation, no host round trip); sets g.keys_sorted.
hipError_t recon_on_buffers(Grid& g, ReconBufs& rb, float* pixels, int W, int H, hipStream_t st);
// Bytes of buffers a reconstruction of the current tree into a W x H image
// needs with its key bound (-1 if the bound is not finite in 32-bit keys).
long long recon_bytes_needed(const Grid& g, int W, int H);
// Its parts: the key bound and the tile-range entries.
long long recon_key_bound(const Grid& g, int W, int H);
long long recon_tiles_cap(const Grid& g, int W, int H);
hipError_t launch_integrate_scatter(Grid& g, float* pixels, int W, int H, hipStream_t st);
hipError_t launch_bytes(const float* pixels, unsigned char* out, long long n, hipStream_t st);

// learned proposal (k_learned.hip)
hipError_t launch_errors_learned(Grid& g, hipStream_t st);
hipError_t launch_errors_learned_b(int D, const DevGrid* grids, const LearnDev* ls, int n, int blocks,
                                   hipStream_t st);
hipError_t launch_mlp_pack(float* p, unsigned short* w1p, unsigned short* w2p, float* w2r,
                           float* w3r, hipStream_t st);
hipError_t launch_mlp_forward(const MlpDev& M, const float* X, long long n, float* Y,
                              hipStream_t st);
// Per-wave partial gradients of mean((y - t)^2) over rows perm(q),
// q in [start, start + count), perm(q) = (pa * q + pc) mod n_rows.  Returns
// the number of wave partials written to gpart / lpart.
int launch_mlp_grad(const MlpDev& M, const float* X, const float* T, long long n_rows,
                    long long start, int count, unsigned long long pa, unsigned long long pc,
                    float* gpart, float* lpart, hipStream_t st, hipError_t* err);
hipError_t launch_mlp_sum(const float* gpart, const float* lpart, int nparts, float* gsum,
                          float* loss, hipStream_t st);
hipError_t launch_mlp_adam(float* p, float* m, float* v, const float* gpart, const float* lpart,
                           int nparts, float lr, float bc1, float bc2, float* loss_acc,
                           unsigned short* w1p, unsigned short* w2p, float* w2r, float* w3r,
                           hipStream_t st);
hipError_t launch_feat_stats(const float* X, long long n, double* part, int nblocks,
                             hipStream_t st);
static const int kMlpGradBlocks = 256;  // workgroups of the gradient kernel (4 waves each)
hipError_t recon_reserve(Grid& g, int W, int H);
hipError_t recon_prepare(Grid& g, int W, int H, hipStream_t st);

}  // namespace mdas
