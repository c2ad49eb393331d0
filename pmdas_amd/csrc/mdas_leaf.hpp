// Access to the leaf lines (layout in mdas_internal.hpp): the sample slots,
// the box (mn[0..D-1], mx[0..D-1]) and the first line_pts(D) points in one
// aligned 128-byte line per leaf; the other points, and for D >= 5 the slots
// beside them, in the overflow rows.  A leaf's count (its flags byte) says
// which slots hold samples.
#pragma once
#include "mdas_device.hpp"
#include "mdas_internal.hpp"

#pragma clang fp contract(off)

namespace mdas {

// The whole line of leaf i (8 float4 loads of one 128-byte line).
__device__ __forceinline__ void load_line(const DevGrid& a, int i, float* f) {
    const float4* p = a.leaf + (size_t)i * (kLineFloats / 4);
#pragma unroll
    for (int k = 0; k < kLineFloats / 4; ++k) {
        const float4 v = p[k];
        f[4 * k] = v.x;
        f[4 * k + 1] = v.y;
        f[4 * k + 2] = v.z;
        f[4 * k + 3] = v.w;
    }
}

// The leaf's sample count (the low bits of its flags byte).
__device__ __forceinline__ int leaf_count(const DevGrid& a, int i) { return (int)(a.flags[i] & 7u); }

// Slot q of leaf i (q < its count) where slot_at(D, q) keeps it.
template <int D>
__device__ __forceinline__ int load_slot(const DevGrid& a, int i, int q) {
    const int at = slot_at(D, q);
    return at >= 0 ? reinterpret_cast<const int*>(a.leaf)[(size_t)i * kLineFloats + at]
                   : reinterpret_cast<const int*>(a.xpts)[(size_t)i * xpts_stride(D) + (-1 - at)];
}

// Slots (the first cnt; -1 after them) and box: the line's first
// line_box_quads(D) float4, and the slots kept past them (D >= 5) only when
// the count reaches them.
template <int D>
__device__ __forceinline__ void load_leaf(const DevGrid& a, int i, int cnt, int4& id, float* mn,
                                          float* mx) {
    constexpr int Q = line_box_quads(D);
    const float4* p = a.leaf + (size_t)i * (kLineFloats / 4);
    float b[4 * Q];
#pragma unroll
    for (int k = 0; k < Q; ++k) {
        const float4 v = p[k];
        b[4 * k] = v.x;
        b[4 * k + 1] = v.y;
        b[4 * k + 2] = v.z;
        b[4 * k + 3] = v.w;
    }
    int s[4] = {__float_as_int(b[0]), __float_as_int(b[1]), __float_as_int(b[2]), __float_as_int(b[3])};
    if constexpr (D >= 5) {
        s[2] = cnt > 2 ? load_slot<D>(a, i, 2) : -1;
        s[3] = cnt > 3 ? load_slot<D>(a, i, 3) : -1;
    }
#pragma unroll
    for (int q = 0; q < 4; ++q)
        if (q >= cnt) s[q] = -1;
    id = make_int4(s[0], s[1], s[2], s[3]);
#pragma unroll
    for (int d = 0; d < D; ++d) {
        mn[d] = b[4 + d];
        mx[d] = b[4 + D + d];
    }
}

template <int D>
__device__ __forceinline__ void load_box(const DevGrid& a, int i, float* mn, float* mx) {
    int4 id;
    load_leaf<D>(a, i, 0, id, mn, mx);
}

// The slots of leaf i (the first cnt; -1 after them).
template <int D>
__device__ __forceinline__ int4 load_slots(const DevGrid& a, int i, int cnt) {
    int s[4];
#pragma unroll
    for (int q = 0; q < 4; ++q) s[q] = q < cnt ? load_slot<D>(a, i, q) : -1;
    return make_int4(s[0], s[1], s[2], s[3]);
}

// Address of point q (slot order) of leaf i: in its line or its overflow row.
template <int D>
__device__ __forceinline__ float* point_addr(const DevGrid& a, int i, int q) {
    return q < line_pts(D) ? reinterpret_cast<float*>(a.leaf) + (size_t)i * kLineFloats + line_pt(D, q)
                           : a.xpts + (size_t)i * xpts_stride(D) + xpt_at(D, q);
}

// An insert (kdgrid.cu:392-396): point pt and sample index k into slot c of
// leaf i.  For D >= 5 and c >= 2 the slot shares the point's 32-byte sector
// (slot_at), written as two whole float4 with the sector's spare floats
// zeroed; otherwise the slot and the point are written where they lie.
template <int D>
__device__ __forceinline__ void store_point_slot(const DevGrid& a, int i, int c, const float* pt, int k) {
    if constexpr (D >= 5) {
        if (c >= 2) {
            float w[8];
#pragma unroll
            for (int t = 0; t < 8; ++t) w[t] = 0.0f;
#pragma unroll
            for (int d = 0; d < D; ++d) w[d] = pt[d];
            w[D] = __int_as_float(k);
            // D = 5: line floats 24..31 (c = 2) or overflow floats 0..7 (c = 3);
            // D = 6: overflow floats 0..7 (c = 2) or 8..15 (c = 3)
            float* base = c < line_pts(D) ? reinterpret_cast<float*>(a.leaf) + (size_t)i * kLineFloats + line_pt(D, c)
                                          : a.xpts + (size_t)i * xpts_stride(D) + xpt_at(D, c);
            float4* q4 = reinterpret_cast<float4*>(base);
            q4[0] = make_float4(w[0], w[1], w[2], w[3]);
            q4[1] = make_float4(w[4], w[5], w[6], w[7]);
            return;
        }
    }
    reinterpret_cast<int*>(a.leaf)[(size_t)i * kLineFloats + c] = k;
    float* dst = point_addr<D>(a, i, c);
#pragma unroll
    for (int d = 0; d < D; ++d) dst[d] = pt[d];
}

// Writes a whole line: the slots kept in it, box and the first points
// (pts[q] for q < min(npts, line_pts(D))), zeros elsewhere, as 8 float4
// stores.  Slots kept in the overflow row (D >= 5) are the caller's.
template <int D>
__device__ __forceinline__ void store_line(const DevGrid& a, int i, int4 id, const float* mn,
                                           const float* mx, const float (*pts)[D], int npts) {
    float f[kLineFloats];
    f[0] = __int_as_float(id.x);
    f[1] = __int_as_float(id.y);
    f[2] = __int_as_float(D <= 4 ? id.z : -1);
    f[3] = __int_as_float(D <= 4 ? id.w : -1);
#pragma unroll
    for (int d = 0; d < D; ++d) {
        f[4 + d] = mn[d];
        f[4 + D + d] = mx[d];
    }
#pragma unroll
    for (int k = 4 + 2 * D; k < kLineFloats; ++k) f[k] = 0.0f;
#pragma unroll
    for (int q = 0; q < line_pts(D); ++q)
#pragma unroll
        for (int d = 0; d < D; ++d)
            if (q < npts) f[line_pt(D, q) + d] = pts[q][d];
    if constexpr (D == 5) f[slot_at(5, 2)] = __int_as_float(id.z);  // beside point 2
    float4* p = a.leaf + (size_t)i * (kLineFloats / 4);
#pragma unroll
    for (int k = 0; k < kLineFloats / 4; ++k)
        p[k] = make_float4(f[4 * k], f[4 * k + 1], f[4 * k + 2], f[4 * k + 3]);
}

// Writes the line's first 4 + 4 D floats: slots, box and the first two
// points (a left child keeps two), whole float4s; the rest of the line keeps
// its old contents, past the leaf's count of 2.
template <int D>
__device__ __forceinline__ void store_line_head2(const DevGrid& a, int i, int4 id, const float* mn,
                                                 const float* mx, const float* p0, const float* p1) {
    constexpr int NF = 4 + 4 * D;  // 16, 20, 24, 28 floats: whole float4s
    static_assert(NF % 4 == 0 && line_pts(D) >= 2, "two points in the line head");
    float f[NF];
    f[0] = __int_as_float(id.x);
    f[1] = __int_as_float(id.y);
    f[2] = __int_as_float(D <= 4 ? id.z : -1);
    f[3] = __int_as_float(D <= 4 ? id.w : -1);
#pragma unroll
    for (int d = 0; d < D; ++d) {
        f[4 + d] = mn[d];
        f[4 + D + d] = mx[d];
        f[line_pt(D, 0) + d] = p0[d];
        f[line_pt(D, 1) + d] = p1[d];
    }
    float4* p = a.leaf + (size_t)i * (kLineFloats / 4);
#pragma unroll
    for (int k = 0; k < NF / 4; ++k)
        p[k] = make_float4(f[4 * k], f[4 * k + 1], f[4 * k + 2], f[4 * k + 3]);
}

// Stopping rule of the device-side loop (optixMotionBlur.cpp:781): no
// further pass once N >= min(target, max_samples).  When the limit is the
// capacity itself (target >= max_samples) also none once a pass could
// overflow it, N + L > max_samples (every leaf selected adds L samples; the
// guard of the app's interactive loop, optixMotionBlur.cpp:723-724), so that
// such a run ends cleanly instead of with a capacity error.
// The error kernels of a pass test it (every workgroup reads the same
// counters) and the reduction after them raises `done` for the rest of the
// pass.
__device__ __forceinline__ bool loop_limit_reached(const DevCounters* c) {
    return c->n_samples >= c->limit ||
           (c->guard_cap != 0 && (long long)c->n_samples + c->n_leaves > (long long)c->guard_cap);
}
__device__ __forceinline__ bool pass_stopped(const DevCounters* c) {
    return c->done || loop_limit_reached(c);
}

struct LeafSamples {
    float vx[4], vy[4], vz[4];
    int n;
    float vol;
};

// The values of a leaf's samples by slot (indices < 0 = empty).  The four
// loads go out together: an empty slot reads sample 0 (always present) and
// is zeroed after, so no load waits behind a branch.
__device__ __forceinline__ void load_values(const DevGrid& a, const int* ids, LeafSamples& ls) {
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        const float* v = a.values + (size_t)(ids[q] >= 0 ? ids[q] : 0) * 3;
        ls.vx[q] = v[0];
        ls.vy[q] = v[1];
        ls.vz[q] = v[2];
    }
#pragma unroll
    for (int q = 0; q < 4; ++q)
        if (ids[q] < 0) ls.vx[q] = ls.vy[q] = ls.vz[q] = 0.0f;
}

// Relative-L1 error of a leaf from its loaded slot values (the occupied slots
// come first) and its box volume (kdgrid.cu:238-266); sets ls.n and ls.vol.
__device__ __forceinline__ float error_of_values(const int* ids, float vol, LeafSamples& ls) {
    float sx = 0.0f, sy = 0.0f, sz = 0.0f;
    int n = 0;
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        if (ids[q] >= 0) {
            sx += ls.vx[q];
            sy += ls.vy[q];
            sz += ls.vz[q];
            ++n;
        }
    }
    // float3 /= float is a reciprocal multiply (sutil/vec_math.h:492-496)
    const float inv = 1.0f / (float)n;
    const float ax = sx * inv, ay = sy * inv, az = sz * inv;
    float dx = 0.0f, dy = 0.0f, dz = 0.0f;
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        if (q < n) {
            dx += fabsf(ls.vx[q] - ax);
            dy += fabsf(ls.vy[q] - ay);
            dz += fabsf(ls.vz[q] - az);
        }
    }
    float e = 0.0f;
    if (ax != 0.0f) e += dx / ax;
    if (ay != 0.0f) e += dy / ay;
    if (az != 0.0f) e += dz / az;
    e /= (float)n;
    e += 1.0e-5f;
    e *= vol;
    ls.n = n;
    ls.vol = vol;
    return e;
}

// Both in one: the error of a leaf from its sample indices and volume.
__device__ __forceinline__ float leaf_error_ids(const DevGrid& a, const int* ids, float vol,
                                                LeafSamples& ls) {
    load_values(a, ids, ls);
    return error_of_values(ids, vol, ls);
}

// Box volume: extents multiplied in dimension order from 1 (aabb.h:20-25).
template <int D>
__device__ __forceinline__ float box_volume(const float* mn, const float* mx) {
    float vol = 1.0f;
#pragma unroll
    for (int d = 0; d < D; ++d) vol *= (mx[d] - mn[d]);
    return vol;
}

// kdgrid.cu:237-266: relative-L1 error of a leaf's samples times its volume.
// Also returns the leaf's sample values by slot, its box and volume for
// callers that derive more statistics from the same reads.  Occupied slots
// are always 0..n-1: slots fill in order and a split writes both children
// from slot 0.
template <int D>
__device__ __forceinline__ float leaf_error_s(const DevGrid& a, int i, LeafSamples& ls, float* mn,
                                              float* mx) {
    int4 id;
    load_leaf<D>(a, i, leaf_count(a, i), id, mn, mx);
    const int ids[4] = {id.x, id.y, id.z, id.w};
    return leaf_error_ids(a, ids, box_volume<D>(mn, mx), ls);
}

// *nvals returns the number of sample values read.
template <int D>
__device__ __forceinline__ float leaf_error(const DevGrid& a, int i, int* nvals) {
    LeafSamples ls;
    float mn[D], mx[D];
    const float e = leaf_error_s<D>(a, i, ls, mn, mx);
    *nvals = ls.n;
    return e;
}

}  // namespace mdas
