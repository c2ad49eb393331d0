/*
 * CPU oracle for the MDAS sampler -- TEST INFRASTRUCTURE, NOT PRODUCT CODE.
 *
 * A deterministic OpenMP C++ restatement of the reference's KDGrid
 * (SDK/lib/mdas/kdgrid.cu, kdgrid.h, aabb.h, point.h, cuda/random.h,
 * cuda/helpers.h).  Only tests/, __graft_entry__.smoke() and bench.py's
 * cpu_baseline leg may load it, and only to check results and as the CPU
 * baseline.
 *
 * Semantics fixed here (the reference is nondeterministic and fast-math):
 *   - IEEE-754 FP32, correctly rounded + - * / sqrt, no FP contraction;
 *   - powf is replaced by orc_powf (double-precision log/exp series rounded
 *     once to float), the same definition the GPU kernels use;
 *   - new sample and leaf indices are allocated by an exclusive scan in leaf
 *     index order (one legal schedule of the reference's warp atomics,
 *     kdgrid.cu:365-378, 426-440);
 *   - pixels accumulate contributions in leaf index order (the reference's
 *     float atomics, kdgrid.cu:562-564, have no fixed order).
 */
#ifndef MDAS_ORACLE_H
#define MDAS_ORACLE_H
#include <stdint.h>
#ifdef __cplusplus
extern "C" {
#endif

typedef struct orc_grid orc_grid;

/* kdgrid.cu:600-631 (constructor); capacities are sized by max_samples. */
orc_grid* orc_create(int dim, long long max_samples, int candidates_num, int bits_per_dim,
                     int extra_img_bits, float alpha, float scale_x, float scale_y);
void orc_destroy(orc_grid* g);
void orc_set_threads(int n);

int orc_initial_sampling(orc_grid* g);            /* kdgrid.cu:633-674, 121-202 */
int orc_compute_errors(orc_grid* g);              /* kdgrid.cu:676-716, 204-288 */
int orc_adaptive_sampling(orc_grid* g);           /* kdgrid.cu:718-773, 290-481 */
int orc_integrate(orc_grid* g, float* pixels, int w, int h); /* kdgrid.cu:781-820, 483-582 */
void orc_convert_to_bytes(const float* pixels, unsigned char* out, long long n); /* :584-598 */
void orc_sampling_density(orc_grid* g, float* pixels, int w, int h); /* kdgrid.cu:822-837 */
int orc_validate(orc_grid* g);                    /* kdgrid.cu:848-961, always on */

/* Whole run with the sample app's stopping rule (optixMotionBlur.cpp:768-791)
 * plus, for a target at or beyond max_samples, its interactive loop's capacity
 * guard (no pass once N + L > max_samples, :723-724), using the analytic scene `params` (mdas_scene::Params layout) as
 * integrand. */
int orc_run(orc_grid* g, const void* scene_params, long long target_samples, int* passes);
/* Scene parameter block (same generator as the product's mdas_scene_create). */
long long orc_make_scene(int kind, uint32_t seed, int dim, float scale_x, float scale_y, void* out,
                         long long capacity);
void orc_eval_scene(const void* scene_params, const float* coords, int dim, long long first,
                    long long count, float* values);
/* Integrand in place over samples [first, first+count) of the grid. */
void orc_eval_range(orc_grid* g, const void* scene_params, long long first, long long count);
void orc_eval_scene_f64(const void* scene_params, const double* coords, int dim,
                        long long count, double* values);
/* The same integrand through the product's sphere groups (mscene::build_accel,
 * the pruned loops of mscene::trace / occluded), compiled here so that a CPU
 * test can hold it against the plain loops above: float coords/values
 * (fp64 = 0) or double (fp64 = 1).  A check of the product's pruning, not a
 * restatement of the reference. */
void orc_eval_scene_grouped(const void* scene_params, const void* coords, int dim,
                            long long count, int fp64, void* values);
/* The plain QMC sampler's image (cuda/sampler.h:4-28, random.h:74-89,
 * whitted.cu:59-69): float4 pixels (fp64 = 0) or double4 (fp64 = 1). */
void orc_qmc_image(const void* scene_params, int w, int h, int spp, int first_sample,
                   uint32_t frame, int fp64, void* out);
float orc_halton(int dim, int i);

/* Reconstruction keys: (morton2(tile_x, tile_y) << 32) | leaf, sorted. */
long long orc_recon_keys(orc_grid* g, int w, int h, int tile_w, int tile_h, uint64_t* out,
                         long long cap);

/* state access */
long long orc_num_samples(orc_grid* g);
long long orc_num_leaves(orc_grid* g);
long long orc_new_samples(orc_grid* g);
long long orc_new_nodes(orc_grid* g);
float orc_max_error(orc_grid* g);
/* Copy state out / in.  Leaves: idx[L*4], box[L*2*dim] (mn then mx), err[L],
 * seed[L]; samples: coords[N*dim], values[N*3].  NULL pointers are skipped. */
void orc_get_state(orc_grid* g, float* coords, float* values, int* idx, float* box, float* err,
                   uint32_t* seed);
int orc_set_state(orc_grid* g, long long n_samples, long long n_leaves, const float* coords,
                  const float* values, const int* idx, const float* box, const float* err,
                  const uint32_t* seed, float gmax);
void orc_set_values(orc_grid* g, long long first, long long count, const float* values);

/* Scalar building blocks, exported for known-answer tests. */
uint32_t orc_tea(uint32_t v0, uint32_t v1, int rounds, const uint32_t* key, uint32_t* v1_out);
uint32_t orc_tea4(uint32_t v0, uint32_t v1);            /* random.h:31-46 */
uint32_t orc_lcg(uint32_t* state);                      /* random.h:49-55 */
float orc_rnd(uint32_t* state);                         /* random.h:64-67 */
float orc_powf(float x, float a);
void orc_morton_cell(long long leaf, int dim, int bits_per_dim, int extra_img_bits,
                     uint32_t* xq, float* offset, float* extent);  /* kdgrid.cu:141-164 */
int orc_largest_axis(const float* diag, int dim);       /* aabb.h:34-45 */
void orc_sort5(const float* keys, int* order);          /* kdgrid.cu:409-417 */
float orc_leaf_error(const float* vals, int count, float volume); /* kdgrid.cu:237-266 */
void orc_make_color(const float* rgb, unsigned char* out); /* helpers.h:35-66 */

#ifdef __cplusplus
}
#endif
#endif
