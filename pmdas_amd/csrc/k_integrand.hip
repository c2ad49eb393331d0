// Analytic integrand launch: the replacement of optixLaunch over the new
// sample slice (optixMotionBlur.cpp:458-467 with sample_offset/sample_count,
// SDK/cuda/whitted.cu:164-293).  The scene's object table (<= 12 KB) and its
// sphere groups (2.5 KB) are staged in LDS once per workgroup; every thread
// then evaluates samples grid-stride.
#include "mdas_device.hpp"
#include "mdas_internal.hpp"
#include "scenes.hpp"

#pragma clang fp contract(off)

namespace mdas {

// The scene's object table and groups into LDS (every thread of the
// workgroup takes part).
__device__ __forceinline__ void stage_scene(const mscene::SceneBlock* __restrict__ B, float* obj,
                                            mscene::Accel* acc) {
    const int nobj = B->p.n_obj;
    for (int k = threadIdx.x; k < nobj * mscene::OBJ_STRIDE; k += blockDim.x) obj[k] = B->p.obj[k];
    const int* src = reinterpret_cast<const int*>(&B->accel);
    int* dst = reinterpret_cast<int*>(acc);
    for (int k = threadIdx.x; k < (int)(sizeof(mscene::Accel) / 4); k += blockDim.x) dst[k] = src[k];
    __syncthreads();
}

__device__ __forceinline__ void eval_body(const float* __restrict__ coords,
                                          float* __restrict__ values, int dim,
                                          const mscene::SceneBlock* __restrict__ B, int first,
                                          int count, const DevCounters* cnt, int dev_range, int blk,
                                          int nblk) {
    __shared__ float obj[mscene::MAX_OBJ * mscene::OBJ_STRIDE];
    __shared__ mscene::Accel acc;
    int beg = first, end = first + count;
    if (dev_range) {
        if (cnt->done) return;
        beg = cnt->n_samples_old;
        end = cnt->n_samples;
    }
    if (beg >= end) return;
    stage_scene(B, obj, &acc);
    const mscene::Params* P = &B->p;
    for (int i = beg + blk * 256 + threadIdx.x; i < end; i += nblk * 256) {
        float x[6];
        const float* c = coords + (size_t)i * dim;
        for (int d = 0; d < dim; ++d) x[d] = c[d];
        mscene::V3<float> v = mscene::eval<float>(*P, obj, x, &acc);
        float* o = values + (size_t)i * 3;
        o[0] = v.x;
        o[1] = v.y;
        o[2] = v.z;
    }
}

__global__ __launch_bounds__(256) void k_eval(const float* __restrict__ coords,
                                              float* __restrict__ values, int dim,
                                              const mscene::SceneBlock* __restrict__ B,
                                              int first, int count, const DevCounters* cnt,
                                              int dev_range) {
    eval_body(coords, values, dim, B, first, count, cnt, dev_range, blockIdx.x, gridDim.x);
}

__global__ __launch_bounds__(256) void k_eval_b(const DevGrid* grids, const void* const* params) {
    const DevGrid& a = grids[blockIdx.y];
    eval_body(a.coords, a.values, a.D, (const mscene::SceneBlock*)params[blockIdx.y], 0, 0, a.cnt,
              1, blockIdx.x, gridDim.x);
}

// ------------------------------------------------- QMC and ground truth --
// The reference's non-MDAS sampler (SDK/cuda/sampler.h:4-28, random.h:75-89,
// whitted.cu:59-83): per pixel a TEA-seeded set of 6 Cranley-Patterson
// offsets, sample s takes the Halton point s (bases 2, 3, 5, 7, 11, 13)
// shifted by the offsets and wrapped when above 1; dims 0 and 1 jitter inside
// the pixel.
// `real` = float gives the MC/QMC baseline images, double the FP64
// high-sample-count ground truth.
template <typename real>
__device__ __forceinline__ real halton_r(int dim, int i) {
    const int primes[6] = {2, 3, 5, 7, 11, 13};
    const int base = primes[dim];
    real r = 0, f = (real)1 / (real)base;
    int j = i;
    while (j > 0) {
        r += f * (real)(j % base);
        j /= base;
        f /= (real)base;
    }
    return r;
}

template <typename real>
__global__ __launch_bounds__(256) void k_qmc_image(const mscene::SceneBlock* __restrict__ B,
                                                   int W, int H, int spp, int first_sample,
                                                   unsigned frame, real* out) {
    __shared__ float obj[mscene::MAX_OBJ * mscene::OBJ_STRIDE];
    __shared__ mscene::Accel acc;
    stage_scene(B, obj, &acc);
    const mscene::Params* P = &B->p;
    const int dim = P->dim;
    for (int pix = blockIdx.x * 256 + threadIdx.x; pix < W * H; pix += gridDim.x * 256) {
        const int px = pix % W, py = pix / W;
        unsigned seed = tea4((unsigned)pix, frame);
        real off[6];
        for (int d = 0; d < 6; ++d) off[d] = (real)rnd(seed);
        real ax = 0, ay = 0, az = 0;
        for (int s = first_sample; s < first_sample + spp; ++s) {
            real x[6];
            for (int d = 0; d < dim; ++d) {
                real v = halton_r<real>(d, s) + off[d];
                if (v > (real)1) v -= (real)1;
                x[d] = v;
            }
            x[0] = ((real)px + x[0]) * (real)P->scale_x / (real)W;
            x[1] = ((real)py + x[1]) * (real)P->scale_y / (real)H;
            const mscene::V3<real> v = mscene::eval<real>(*P, obj, x, &acc);
            ax += v.x;
            ay += v.y;
            az += v.z;
        }
        real* o = out + (size_t)pix * 4;
        const real inv = (real)1 / (real)spp;
        o[0] = ax * inv;
        o[1] = ay * inv;
        o[2] = az * inv;
        o[3] = 0;
    }
}

hipError_t launch_qmc_image(const void* dparams, int W, int H, int spp, int first_sample,
                            unsigned frame, bool fp64, void* out, hipStream_t st) {
    long long blocks = ((long long)W * H + 255) / 256;
    if (blocks > 8192) blocks = 8192;
    if (fp64)
        hipLaunchKernelGGL(k_qmc_image<double>, dim3((unsigned)blocks), dim3(256), 0, st,
                           (const mscene::SceneBlock*)dparams, W, H, spp, first_sample, frame,
                           (double*)out);
    else
        hipLaunchKernelGGL(k_qmc_image<float>, dim3((unsigned)blocks), dim3(256), 0, st,
                           (const mscene::SceneBlock*)dparams, W, H, spp, first_sample, frame,
                           (float*)out);
    return hipGetLastError();
}

hipError_t launch_eval(Grid& g, const void* dparams, int first, int count, bool dev_range,
                       hipStream_t st) {
    long long n = dev_range ? (long long)g.cap_s : (long long)count;
    long long blocks = (n + 255) / 256;
    if (blocks > 8192) blocks = 8192;
    if (blocks < 1) blocks = 1;
    hipLaunchKernelGGL(k_eval, dim3((unsigned)blocks), dim3(256), 0, st, g.dev.coords,
                       g.dev.values, g.D, (const mscene::SceneBlock*)dparams, first, count,
                       (const DevCounters*)g.dev.cnt, dev_range ? 1 : 0);
    return hipGetLastError();
}

hipError_t launch_eval_b(const DevGrid* grids, const void* const* params, int n, int blocks,
                         hipStream_t st) {
    hipLaunchKernelGGL(k_eval_b, dim3(blocks, n), dim3(256), 0, st, grids, params);
    return hipGetLastError();
}

}  // namespace mdas
