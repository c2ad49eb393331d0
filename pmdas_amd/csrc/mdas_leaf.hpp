// Access to the per-leaf records: one int4 of sample slots followed by the
// box (mn[0..D-1], mx[0..D-1]), leaf_quads(D) float4 per leaf.
#pragma once
#include "mdas_device.hpp"
#include "mdas_internal.hpp"

namespace mdas {

template <int D>
__device__ __forceinline__ void load_leaf(const DevGrid& a, int i, int4& id, float* mn, float* mx) {
    constexpr int Q = leaf_quads(D);
    const float4* p = a.leaf + (size_t)i * Q;
    const float4 q0 = p[0];
    id = make_int4(__float_as_int(q0.x), __float_as_int(q0.y), __float_as_int(q0.z),
                   __float_as_int(q0.w));
    float b[4 * (Q - 1)];
#pragma unroll
    for (int k = 1; k < Q; ++k) {
        const float4 v = p[k];
        b[4 * (k - 1) + 0] = v.x;
        b[4 * (k - 1) + 1] = v.y;
        b[4 * (k - 1) + 2] = v.z;
        b[4 * (k - 1) + 3] = v.w;
    }
#pragma unroll
    for (int d = 0; d < D; ++d) {
        mn[d] = b[d];
        mx[d] = b[D + d];
    }
}

template <int D>
__device__ __forceinline__ void load_box(const DevGrid& a, int i, float* mn, float* mx) {
    constexpr int Q = leaf_quads(D);
    const float4* p = a.leaf + (size_t)i * Q;
    float b[4 * (Q - 1)];
#pragma unroll
    for (int k = 1; k < Q; ++k) {
        const float4 v = p[k];
        b[4 * (k - 1) + 0] = v.x;
        b[4 * (k - 1) + 1] = v.y;
        b[4 * (k - 1) + 2] = v.z;
        b[4 * (k - 1) + 3] = v.w;
    }
#pragma unroll
    for (int d = 0; d < D; ++d) {
        mn[d] = b[d];
        mx[d] = b[D + d];
    }
}

__device__ __forceinline__ int4 load_slots(const DevGrid& a, int i, int Q) {
    const float4 q0 = a.leaf[(size_t)i * Q];
    return make_int4(__float_as_int(q0.x), __float_as_int(q0.y), __float_as_int(q0.z),
                     __float_as_int(q0.w));
}

__device__ __forceinline__ void store_slots(const DevGrid& a, int i, int Q, int4 id) {
    a.leaf[(size_t)i * Q] = make_float4(__int_as_float(id.x), __int_as_float(id.y),
                                        __int_as_float(id.z), __int_as_float(id.w));
}

template <int D>
__device__ __forceinline__ void store_leaf(const DevGrid& a, int i, int4 id, const float* mn,
                                           const float* mx) {
    constexpr int Q = leaf_quads(D);
    float4* p = a.leaf + (size_t)i * Q;
    float b[4 * (Q - 1)];
#pragma unroll
    for (int d = 0; d < D; ++d) {
        b[d] = mn[d];
        b[D + d] = mx[d];
    }
#pragma unroll
    for (int k = 2 * D; k < 4 * (Q - 1); ++k) b[k] = 0.0f;
    p[0] = make_float4(__int_as_float(id.x), __int_as_float(id.y), __int_as_float(id.z),
                       __int_as_float(id.w));
#pragma unroll
    for (int k = 1; k < Q; ++k)
        p[k] = make_float4(b[4 * (k - 1)], b[4 * (k - 1) + 1], b[4 * (k - 1) + 2], b[4 * (k - 1) + 3]);
}

}  // namespace mdas
