/*
 * mdas.h -- C ABI of the MI355X-native MDAS sampler (libmdas.so).
 *
 * Drop-in boundary for the reference's C++ class mdas::KDGrid<Point>
 * (SDK/lib/mdas/kdgrid.h:10-107, explicitly instantiated for Point3..Point6 at
 * SDK/lib/mdas/kdgrid.cu:25-104) as the sample apps call it
 * (SDK/optixMotionBlur/optixMotionBlur.cpp:331-360, 473-491, 768-806).
 * Plain pointers and sizes only; a HIP stream is passed as `void*`
 * (NULL = the null stream, as the reference uses stream 0).
 *
 * Conventions (differences from the reference are deliberate and listed):
 *  - every entry point returns an int status (MDAS_OK = 0); the message of the
 *    last failure on the calling thread is mdas_last_error().  The reference
 *    returns void and only prints CUDA errors (buffer.h:6-16).
 *  - all state is per handle (the reference keeps g_error, g_warpCounter0/1 and
 *    a texture reference in device globals, kdgrid.cu:19-23).
 *  - buffers are device memory (hipMalloc); the reference uses managed memory
 *    (buffer.h:44).  Coordinates are AoS float[dim] per sample and values are
 *    float3 per sample, exactly as the reference's LaunchParams expect
 *    (SDK/cuda/whitted.h:76-81).
 *  - a pass that would exceed max_samples returns MDAS_ERR_CAPACITY and leaves
 *    samples and tree unchanged (the reference writes out of bounds).
 *  - indices are allocated in leaf order by a device-wide scan, so results are
 *    deterministic (the reference's warp atomics are one of many orders).
 */
#ifndef MDAS_H
#define MDAS_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define MDAS_OK 0
#define MDAS_ERR_INVALID (-1)   /* bad argument or configuration */
#define MDAS_ERR_CAPACITY (-2)  /* a pass would exceed max_samples */
#define MDAS_ERR_HIP (-3)       /* HIP runtime error */
#define MDAS_ERR_STATE (-4)     /* call out of order (e.g. no initial sampling) */

typedef struct mdas_grid* mdas_handle;
typedef struct mdas_scene* mdas_scene_handle;

/* Constructor arguments of KDGrid (kdgrid.h:42-51; kdgrid.cu:600-631) and the
 * app's Mdas.* options (optixMotionBlur.cpp:118-125, 336-350). */
typedef struct mdas_config {
    int dim;                /* Point::DIM, 3..6 */
    long long max_samples;  /* sample capacity and stopping cap (<= 2^31-1) */
    int candidates_num;     /* Mdas.candidatesNum (default 4) */
    int bits_per_dim;       /* Mdas.bitsPerDim */
    int extra_img_bits;     /* Mdas.extraImgBits */
    float alpha;            /* Mdas.alpha */
    float scale_x;          /* Mdas.scaleFactor * Film.width */
    float scale_y;          /* Mdas.scaleFactor * Film.height */
    const char* log_path;   /* optional "KEY\nvalue" log (kdgrid.cu:670-817); NULL = off */
} mdas_config;

/* KDGrid::KDGrid (kdgrid.h:42-51) */
int mdas_create(const mdas_config* cfg, mdas_handle* out);
/* KDGrid::~KDGrid (buffers freed; buffer.h:27-29) */
int mdas_destroy(mdas_handle h);

/* The constructor's std::ofstream* log (kdgrid.h:50, kdgrid.cu:609-630) for C
 * callers: fn(user, key, value, is_count) receives every "KEY", value pair the
 * reference writes to that stream (kdgrid.cu:670-671, 713, 769-770, 816-817;
 * is_count = 1 for sample counts, 0 for times in ms), in the same order, in
 * addition to the config's log_path file.  fn = NULL turns it off.
 * include/mdas_kdgrid.hpp uses it to write to the app's std::ofstream. */
typedef void (*mdas_log_fn)(void* user, const char* key, double value, int is_count);
int mdas_set_log_sink(mdas_handle h, mdas_log_fn fn, void* user);

/* KDGrid::InitialSampling (kdgrid.h:53; kdgrid.cu:633-674) */
int mdas_initial_sampling(mdas_handle h, void* stream);
/* KDGrid::ComputeErrors (kdgrid.h:55; kdgrid.cu:676-716) */
int mdas_compute_errors(mdas_handle h, void* stream);
/* KDGrid::AdaptiveSampling (kdgrid.h:56; kdgrid.cu:718-773).  Like the
 * reference it synchronises to read the new sample / node counts. */
int mdas_adaptive_sampling(mdas_handle h, void* stream);
/* KDGrid::SamplingPass (kdgrid.h:58; kdgrid.cu:775-779) */
int mdas_sampling_pass(mdas_handle h, void* stream);
/* KDGrid::Integrate (kdgrid.h:59; kdgrid.cu:781-820).  `pixels` is a device
 * float4[width*height]; every pixel is written (xyz = estimate, w = 0).
 * Stream-ordered, like the reference: the buffers are sized from the leaf
 * count the host already holds (a call allocates only when they grow), except
 * for trees whose key bound exceeds 8 GiB (initial cells spanning many
 * pixel tiles), where the key count is read back first. */
int mdas_integrate(mdas_handle h, float* pixels, int width, int height, void* stream);
/* No counterpart in the reference: mdas_integrate for callers that cannot
 * allocate device memory (code under hipGraph capture, tight serving loops,
 * harnesses that forbid hipMalloc).  The same kernels and pixels as
 * mdas_integrate, on a scratch array of MDAS_NOALLOC_SCRATCH_BYTES that the
 * library holds on each GPU it runs on until the process ends; the call makes
 * no allocation and no host round trip.  A tree and image whose
 * reconstruction buffers (leaf records, keys at their bound, tile ranges,
 * sort histograms) exceed the array return MDAS_ERR_CAPACITY before anything
 * runs.  Every call on one GPU shares that GPU's array: calls that overlap
 * on one GPU (other streams, other threads) race; serialise them. */
#define MDAS_NOALLOC_SCRATCH_BYTES (256u << 20)
int mdas_integrate_noalloc(mdas_handle h, float* pixels, int width, int height, void* stream);
/* Reference-style reconstruction with float atomics (kdgrid.cu:483-582);
 * same result up to summation order.  Kept for comparison. */
int mdas_integrate_scatter(mdas_handle h, float* pixels, int width, int height, void* stream);
/* KDGrid::ConvertToBytes (kdgrid.h:61; kdgrid.cu:839-846) -> uchar4 sRGB.
 * Uses no grid state: h may be NULL (e.g. for a plain QMC image). */
int mdas_convert_to_bytes(mdas_handle h, const float* pixels, unsigned char* bytes, int width,
                          int height, void* stream);
/* KDGrid::SamplingDensity (kdgrid.h:60; kdgrid.cu:822-837): host float4 image */
int mdas_sampling_density(mdas_handle h, float* host_pixels, int width, int height);
/* KDGrid::Validate (kdgrid.h:63; kdgrid.cu:848-961), always enabled; *valid = 0/1 */
int mdas_validate(mdas_handle h, int* valid);

/* Getters (kdgrid.h:65-77).  Pointers are device pointers.  After
 * InitialSampling, mdas_get_new_samples is the initial sample count (the
 * slice the first integrand launch covers); the reference leaves the
 * previous pass's count there, and its apps read GetNumberOfSamples at that
 * point (optixMotionBlur.cpp:352-358). */
float* mdas_get_sample_coordinates(mdas_handle h);
float* mdas_get_sample_values(mdas_handle h);
int mdas_get_candidates_num(mdas_handle h);
int mdas_get_max_node_size(mdas_handle h);
int mdas_get_new_samples(mdas_handle h);
int mdas_get_new_nodes(mdas_handle h);
int mdas_get_number_of_samples(mdas_handle h);
int mdas_get_number_of_nodes(mdas_handle h);
int mdas_get_max_samples(mdas_handle h);
float mdas_get_scale_x(mdas_handle h);
float mdas_get_scale_y(mdas_handle h);
/* Additions: the point dimension, and g_error as the last error pass left it
 * (kdgrid.cu:21, 286; adaptive passes do not change it). */
int mdas_get_dim(mdas_handle h);
float mdas_get_max_error(mdas_handle h);

/* ---- additions without a reference counterpart ---------------------------
 * Analytic integrands replacing the OptiX raygen programs
 * (SDK/cuda/whitted.cu:164-293, ao.cu:112-152) over the same buffer contract. */
#define MDAS_SCENE_CONSTANT 0
#define MDAS_SCENE_DISKS_MB 1
#define MDAS_SCENE_SPHERES_DOF 2
#define MDAS_SCENE_SPHERES_MBDOF 3
#define MDAS_SCENE_AO 4
#define MDAS_SCENE_RAMP 5
int mdas_scene_create(int kind, uint32_t seed, int dim, float scale_x, float scale_y,
                      mdas_scene_handle* out);
int mdas_scene_destroy(mdas_scene_handle s);
/* host copy of the scene parameter block; returns its size in bytes */
long long mdas_scene_params(mdas_scene_handle s, void* out, long long capacity);
/* Integrand over samples [offset, offset+count) -- what optixLaunch(width =
 * sample_count) does with sample_offset (optixMotionBlur.cpp:458-467). */
int mdas_eval_integrand(mdas_handle h, mdas_scene_handle s, int offset, int count, void* stream);
/* Integrand over the slice of the last adaptive pass, read on the device. */
int mdas_eval_new_samples(mdas_handle h, mdas_scene_handle s, void* stream);

/* The reference's plain (non-MDAS) sampler over the same scene
 * (SDK/cuda/sampler.h:4-28, whitted.cu:59-83): per pixel, Halton points
 * first_sample .. first_sample+spp-1 with TEA-seeded Cranley-Patterson
 * offsets (seed tea4(pixel, frame)).  `out` is a device float4 image, or a
 * double4 image evaluated in FP64 when fp64 != 0 (ground truth). */
int mdas_render_qmc(mdas_scene_handle s, int width, int height, int spp, int first_sample,
                    uint32_t frame, int fp64, void* out, void* stream);

/* Whole sampling loop without per-pass host round trips: initial sampling,
 * integrand, then passes while N < min(target_samples, max_samples)
 * (optixMotionBlur.cpp:768-791).  With a target at or beyond max_samples the
 * loop also keeps N + L <= max_samples, the capacity guard of the app's
 * interactive loop (optixMotionBlur.cpp:723-724), so such a run ends cleanly
 * (MDAS_OK) short of the capacity instead of overflowing; with a smaller
 * target a pass that would overflow returns MDAS_ERR_CAPACITY.  The stopping
 * rule is evaluated on the device; the host checks it once every few
 * passes. */
int mdas_run(mdas_handle h, mdas_scene_handle s, long long target_samples, void* stream);
int mdas_get_passes(mdas_handle h);

/* Per-phase GPU time of the last mdas_run / mdas_integrate (ms), from HIP
 * events on the caller's stream: initial, errors, adaptive, integrand,
 * reconstruction (incl. sort), bytes.  Enabled by mdas_set_timing(h, 1). */
typedef struct mdas_phase_times {
    double initial_ms, errors_ms, adaptive_ms, integrand_ms, recon_ms, sort_ms;
    int passes;
    long long final_samples, final_leaves;
    double bytes_initial, bytes_errors, bytes_adaptive, bytes_recon;  /* algorithmic */
} mdas_phase_times;
int mdas_set_timing(mdas_handle h, int enable);
int mdas_get_phase_times(mdas_handle h, mdas_phase_times* out);

/* Per-pass device counters of the last run, 8 ints per row: leaves and
 * samples before the pass, selected, split, samples held by the selected
 * leaves, stale leaves re-estimated, values they read, 0.  Returns the rows. */
int mdas_get_pass_stats(mdas_handle h, int* out, int max_rows);

/* Per-kernel GPU time of the last mdas_run + mdas_integrate when timing is
 * enabled: entry `index` gives the kernel name, launches, total ms and its
 * algorithmic bytes.  Returns the number of entries. */
int mdas_kernel_stats(mdas_handle h, int index, char* name, int name_cap, int* launches,
                      double* total_ms, double* bytes);

/* Tile of the reconstruction gather: tile_w x tile_h pixels, 16 or 64 in
 * all (default 8 x 8).  The result does not depend on it. */
int mdas_set_recon_tile(mdas_handle h, int tile_w, int tile_h);
/* Workgroups of the per-pass kernels: leaf chunks of the selection,
 * compaction and reconstruction kernels (64..32768, default 4096) and of
 * the error pass (64..8192, default 1024).  Results do not depend on them
 * (allocation follows leaf order); only the time does. */
int mdas_set_launch_shape(mdas_handle h, int chunk_blocks, int err_blocks);
/* Reconstructs the current state into `pixels` with each candidate tile
 * shape, keeps the fastest and returns it. */
int mdas_autotune_recon(mdas_handle h, float* pixels, int width, int height, void* stream,
                        int* tile_w, int* tile_h);

/* State transfer for tests and tools (synchronous).
 * idx[L*4], box[L*2*dim] (per leaf: mn[dim] then mx[dim]), err[L], seed[L],
 * coords[N*dim], values[N*3]; NULL pointers are skipped.  A leaf's occupied
 * slots come first; an upload takes any negative index as empty, a download
 * writes the reference's empty marker ~dim (kdgrid.h:27-30) there. */
int mdas_download_state(mdas_handle h, float* coords, float* values, int* idx, float* box,
                        float* err, uint32_t* seed);
int mdas_upload_state(mdas_handle h, int n_samples, int n_leaves, const float* coords,
                      const float* values, const int* idx, const float* box, const float* err,
                      const uint32_t* seed, float max_error);
int mdas_set_values(mdas_handle h, int offset, int count, const float* host_values);
/* Sorted reconstruction keys (morton2(tile_x, tile_y) << 32 | leaf) of the
 * last mdas_integrate; *n = number of keys. */
int mdas_get_recon_keys(mdas_handle h, uint64_t* out, long long capacity, long long* n);

/* ---- batches ------------------------------------------------------------------
 * Many grids (images of uneven resolution; dimensions are grouped) advanced
 * together: one launch per kernel and pass for every image of a dimension,
 * the block of passes captured once into a hipGraph and replayed until every
 * image meets its stopping rule.  Each image's result equals mdas_run on it
 * alone.  The grids and scenes must outlive the batch.  A grid's learned
 * mode (mdas_set_learned / _exponent) is taken at mdas_batch_create: the
 * batched error pass then computes features and network scores for the
 * learned members and the plain error for the others.  With use_graph and a
 * NULL stream (which cannot be captured) the batch runs on a stream of its
 * own that waits for earlier work on the null stream and that later work on
 * it waits for. */
typedef struct mdas_batch* mdas_batch_handle;
int mdas_batch_create(mdas_handle* grids, mdas_scene_handle* scenes, int n, mdas_batch_handle* out);
int mdas_batch_destroy(mdas_batch_handle b);
/* targets[i]: sample target of image i; use_graph: replay a captured graph */
int mdas_batch_run(mdas_batch_handle b, const long long* targets, void* stream, int use_graph);
/* Reconstruction of every image i (kdgrid.cu:781-820) into its device float4
 * image pixels[i] of widths[i] x heights[i], each with its grid's gather
 * tile; the same pixels as mdas_integrate on each grid.  With use_graph the
 * reconstructions of all images are one captured hipGraph, replayed while the
 * images' sizes, tiles, outputs and buffers stay the same (the buffers are
 * sized before the capture, from each image's leaf count).  Stream-ordered;
 * the caller synchronises. */
int mdas_batch_integrate(mdas_batch_handle b, float* const* pixels, const int* widths,
                         const int* heights, void* stream, int use_graph);
/* wall time of the last run (ms), blocks of passes run, dimension groups */
int mdas_batch_info(mdas_batch_handle b, double* wall_ms, int* pass_blocks, int* groups);
/* Runs the batch (with graph replay) once per candidate launch shape --
 * 2048 .. 32768 workgroups per launch in all, shared by a dimension group's
 * images -- and keeps the fastest; *chosen receives it.  The images' results
 * do not depend on the shape. */
int mdas_batch_autotune(mdas_batch_handle b, const long long* targets, void* stream, int* chosen);

/* ---- configuration files ----------------------------------------------------
 * Reader of the reference's ".env" files (SDK/sutil/Environment.cpp:593-792)
 * for the Sampler / Mdas / Film options (optixMotionBlur.cpp:107-134) plus a
 * `Scene { kind <constant|disks|dof|mbdof|ao|ramp> seed <n> dim <d> }` block.
 * Defaults are the reference's registered defaults. */
typedef struct mdas_env {
    int mdas;             /* Sampler.mdas */
    float samples;        /* Sampler.samples (spp target) */
    float scale_factor;   /* Mdas.scaleFactor */
    float alpha;          /* Mdas.alpha */
    int bits_per_dim;     /* Mdas.bitsPerDim */
    int extra_img_bits;   /* Mdas.extraImgBits */
    int candidates_num;   /* Mdas.candidatesNum */
    int width, height;    /* Film.width / Film.height */
    char filename[512];   /* Film.filename */
    int scene_kind;       /* Scene.kind, -1 if absent */
    uint32_t scene_seed;  /* Scene.seed */
    int dim;              /* Scene.dim, else the scene's dimension */
} mdas_env;
int mdas_env_read(const char* path, mdas_env* out);
const char* mdas_env_last_error(void);
/* KDGrid arguments of the app (optixMotionBlur.cpp:336-350). */
int mdas_env_config(const mdas_env* env, long long max_samples, mdas_config* cfg);

/* ------------------------------------------------------------------------
 * Learned proposal (no reference counterpart: a build addition beside the
 * plain method).  A 16 -> 64 -> 64 -> 1 MLP with GELU (tanh form), BF16
 * weights and layer inputs on MFMA, FP32 accumulation, predicts from 16 leaf
 * features the log2 of the relative error reduction a split achieves,
 * r = max(e_parent - e_left - e_right, 2^-10 e_parent) / e_parent, measured
 * by the plain error estimate.  With scoring on, selection uses
 * e * 2^clamp(prediction, -10, 0) in place of the error e in
 * pow(score / max score, alpha) >= u; the plain error remains the fallback
 * (no network or a non-finite output) and is still computed for every leaf.
 *
 * Parameter blob (MDAS_MLP_FLOATS floats): W1[64][16], b1[64], W2[64][64],
 * b2[64], w3[64], b3 -- the MDAS_MLP_PARAMS trained values -- then the input
 * normalisation mu[16] and 1/sigma[16] (x' = (x - mu) / sigma).
 * Features (MDAS_FEATURES floats per leaf): log2 extent per dimension over
 * the domain extent (6 values, 0 above D), sample count, mean, variance,
 * min and max luminance (0.2126 R + 0.7152 G + 0.0722 B), depth (splits
 * since the initial grid), log2 error, log2 error / volume, log2 volume, D.
 * All row and label pointers are device memory. */
typedef struct mdas_mlp* mdas_mlp_handle;
#define MDAS_MLP_FLOATS 5345
#define MDAS_MLP_PARAMS 5313
#define MDAS_FEATURES 16
#define MDAS_LEARN_COLLECT 1 /* write (parent features, label) rows of every split */
#define MDAS_LEARN_SCORE 2   /* select by the network's score */
/* Second label form: log2 of the squared image change a split makes,
 * sum over children c of (m_c - m_p)^2 v_c^2 / a_c (m mean luminance, v box
 * volume, a xy footprint area); scored as 2^clamp(prediction, -120, 120). */
#define MDAS_LEARN_IMAGE_LABEL 4
/* Keep every leaf's full feature row and plain error for
 * mdas_get_leaf_features (always kept while collecting); without it a
 * scoring run keeps only each leaf's depth, the one feature a split carries
 * over to its children. */
#define MDAS_LEARN_KEEP_FEATURES 8

/* Seeded default initialisation: U(-1/sqrt(fan_in), 1/sqrt(fan_in)). */
int mdas_mlp_create(uint32_t seed, mdas_mlp_handle* out);
int mdas_mlp_destroy(mdas_mlp_handle m);
/* Replace / read the blob; setting it also resets the Adam state. */
int mdas_mlp_set_params(mdas_mlp_handle m, const float* host, int n);
int mdas_mlp_get_params(mdas_mlp_handle m, float* host, int n);
/* mu and 1/sigma of the n rows (FP64 sums). */
int mdas_mlp_fit_norm(mdas_mlp_handle m, const float* rows, long long n, void* stream);
/* out[i] = network(rows[i]) for n rows of MDAS_FEATURES raw features. */
int mdas_mlp_forward(mdas_mlp_handle m, const float* rows, long long n, float* out, void* stream);
/* Gradient of mean((out - target)^2) over the n rows, summed over rows in a
 * fixed order; grad_host receives MDAS_MLP_PARAMS floats, loss_host the
 * loss.  Synchronises. */
int mdas_mlp_grad(mdas_mlp_handle m, const float* rows, const float* targets, long long n,
                  float* grad_host, float* loss_host, void* stream);
/* Adam (beta 0.9 / 0.999, eps 1e-8) over `epochs` passes of the rows in
 * minibatches of `batch`, each epoch in a seeded permutation
 * q -> (a q + c) mod n; epoch_loss_host (optional, `epochs` floats) gets the
 * mean loss seen during each epoch.  Synchronises. */
int mdas_mlp_train(mdas_mlp_handle m, const float* rows, const float* targets, long long n,
                   int epochs, int batch, float lr, uint32_t seed, float* epoch_loss_host,
                   void* stream);
/* Learned mode of a grid for mdas_run (0 = the plain method).  COLLECT
 * writes, for the j-th split leaf of each pass, its features and label to
 * rows / labels at index (splits of earlier passes) + j while that index is
 * below cap_rows; SCORE needs `m`, which must outlive the grid's runs.
 * Resets the row counter. */
int mdas_set_learned(mdas_handle h, int mode, mdas_mlp_handle m, float* rows, float* labels,
                     long long cap_rows);
/* Score exponent kappa in (0, 16] (default 1, reset by mdas_set_learned):
 * the network's score s becomes s^kappa, i.e. selection tests
 * pow(s / max s, kappa * alpha) >= u -- the learned form's own flattening. */
int mdas_set_learned_exponent(mdas_handle h, float kappa);
/* Separate exponents: score = e^kappa_e * 2^(kappa_y * prediction), each in
 * [0, 16], not both 0 (kappa_y = 0: the network is ignored; the plain error
 * with the effective exponent kappa_e * alpha).  Reset by mdas_set_learned. */
int mdas_set_learned_exponents(mdas_handle h, float kappa_e, float kappa_y);
/* Rows offered since mdas_set_learned (those beyond cap_rows were dropped). */
int mdas_get_learned_rows(mdas_handle h, long long* offered);
/* Features and plain errors of leaves [0, n) after a learned-mode run that
 * kept them (MDAS_LEARN_COLLECT or MDAS_LEARN_KEEP_FEATURES). */
int mdas_get_leaf_features(mdas_handle h, float* feat_host, float* perr_host, long long n);

const char* mdas_last_error(void);
int mdas_device_synchronize(void);

#ifdef __cplusplus
}
#endif
#endif /* MDAS_H */
