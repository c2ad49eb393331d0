// Analytic integrand launch: the replacement of optixLaunch over the new
// sample slice (optixMotionBlur.cpp:458-467 with sample_offset/sample_count,
// SDK/cuda/whitted.cu:164-293).  The scene's object table (<= 12 KB) is staged
// in LDS once per workgroup; every thread then evaluates samples grid-stride.
#include "mdas_device.hpp"
#include "mdas_internal.hpp"
#include "scenes.hpp"

#pragma clang fp contract(off)

namespace mdas {

__device__ __forceinline__ void eval_body(const float* __restrict__ coords,
                                          float* __restrict__ values, int dim,
                                          const mscene::Params* __restrict__ P, int first,
                                          int count, const DevCounters* cnt, int dev_range, int blk,
                                          int nblk) {
    __shared__ float obj[mscene::MAX_OBJ * mscene::OBJ_STRIDE];
    int beg = first, end = first + count;
    if (dev_range) {
        if (cnt->done) return;
        beg = cnt->n_samples_old;
        end = cnt->n_samples;
    }
    if (beg >= end) return;
    const int nobj = P->n_obj;
    for (int k = threadIdx.x; k < nobj * mscene::OBJ_STRIDE; k += 256) obj[k] = P->obj[k];
    __syncthreads();
    for (int i = beg + blk * 256 + threadIdx.x; i < end; i += nblk * 256) {
        float x[6];
        const float* c = coords + (size_t)i * dim;
        for (int d = 0; d < dim; ++d) x[d] = c[d];
        mscene::V3<float> v = mscene::eval<float>(*P, obj, x);
        float* o = values + (size_t)i * 3;
        o[0] = v.x;
        o[1] = v.y;
        o[2] = v.z;
    }
}

__global__ __launch_bounds__(256) void k_eval(const float* __restrict__ coords,
                                              float* __restrict__ values, int dim,
                                              const mscene::Params* __restrict__ P,
                                              int first, int count, const DevCounters* cnt,
                                              int dev_range) {
    eval_body(coords, values, dim, P, first, count, cnt, dev_range, blockIdx.x, gridDim.x);
}

__global__ __launch_bounds__(256) void k_eval_b(const DevGrid* grids, const void* const* params) {
    const DevGrid& a = grids[blockIdx.y];
    eval_body(a.coords, a.values, a.D, (const mscene::Params*)params[blockIdx.y], 0, 0, a.cnt, 1,
              blockIdx.x, gridDim.x);
}

hipError_t launch_eval(Grid& g, const void* dparams, int first, int count, bool dev_range,
                       hipStream_t st) {
    long long n = dev_range ? (long long)g.cap_s : (long long)count;
    long long blocks = (n + 255) / 256;
    if (blocks > 8192) blocks = 8192;
    if (blocks < 1) blocks = 1;
    hipLaunchKernelGGL(k_eval, dim3((unsigned)blocks), dim3(256), 0, st, g.dev.coords,
                       g.dev.values, g.D, (const mscene::Params*)dparams, first, count,
                       (const DevCounters*)g.dev.cnt, dev_range ? 1 : 0);
    return hipGetLastError();
}

hipError_t launch_eval_b(const DevGrid* grids, const void* const* params, int n, int blocks,
                         hipStream_t st) {
    hipLaunchKernelGGL(k_eval_b, dim3(blocks, n), dim3(256), 0, st, grids, params);
    return hipGetLastError();
}

}  // namespace mdas
