// Device-side arithmetic of the MDAS sampler: RNG, pow, Morton cell decode,
// wave64 helpers.  Every float operation here is a single IEEE-754 operation
// (the library is built with -ffp-contract=off and correctly rounded f32
// division and square root), so the results are defined independently of the
// compiler's scheduling.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#pragma clang fp contract(off)

namespace mdas {

// cuda/random.h:31-46 -- TEA, 4 rounds, seeded with (leaf index, 0).
__device__ __forceinline__ unsigned tea4(unsigned v0, unsigned v1) {
    unsigned s0 = 0;
#pragma unroll
    for (int n = 0; n < 4; ++n) {
        s0 += 0x9e3779b9u;
        v0 += ((v1 << 4) + 0xa341316cu) ^ (v1 + s0) ^ ((v1 >> 5) + 0xc8013ea4u);
        v1 += ((v0 << 4) + 0xad90777du) ^ (v0 + s0) ^ ((v0 >> 5) + 0x7e95761eu);
    }
    return v0;
}

// cuda/random.h:49-55, 64-67 -- LCG step, 24 random bits -> [0, 1).
__device__ __forceinline__ float rnd(unsigned& s) {
    s = 1664525u * s + 1013904223u;
    return (float)(s & 0x00FFFFFFu) * (1.0f / 16777216.0f);
}

// powf(x, a) as exp(a ln x) in double precision, rounded once to float.
// ln: x = m 2^e, m in [sqrt(1/2), sqrt(2)), ln m = 2 atanh((m-1)/(m+1)) by its
// series to t^19 (Abramowitz & Stegun 4.1.27); exp: y = k ln2 + r, Taylor
// series to r^13 (A&S 4.2.1).  The double result is within ~1e-16 relative.
__device__ __forceinline__ double ln_pos(double x) {
    long long bits = __double_as_longlong(x);
    int e = (int)((bits >> 52) & 0x7ff) - 1023;
    bits = (bits & 0x000fffffffffffffll) | 0x3ff0000000000000ll;
    double m = __longlong_as_double(bits);
    if (m > 1.4142135623730951) {
        m = m * 0.5;
        e += 1;
    }
    double t = (m - 1.0) / (m + 1.0);
    double t2 = t * t;
    double s = 1.0 / 19.0;
    s = s * t2 + 1.0 / 17.0;
    s = s * t2 + 1.0 / 15.0;
    s = s * t2 + 1.0 / 13.0;
    s = s * t2 + 1.0 / 11.0;
    s = s * t2 + 1.0 / 9.0;
    s = s * t2 + 1.0 / 7.0;
    s = s * t2 + 1.0 / 5.0;
    s = s * t2 + 1.0 / 3.0;
    s = s * t2 + 1.0;
    return (double)e * 0.6931471805599453094 + 2.0 * t * s;
}

__device__ __forceinline__ double exp_dbl(double y) {
    if (y < -110.0) return 0.0;
    if (y > 100.0) return __longlong_as_double(0x7ff0000000000000ll);
    double kd = __builtin_floor(y * 1.4426950408889634074 + 0.5);
    int k = (int)kd;
    double r = y - kd * 0.6931471805599453094;
    double p = 1.0 / 6227020800.0;
    p = p * r + 1.0 / 479001600.0;
    p = p * r + 1.0 / 39916800.0;
    p = p * r + 1.0 / 3628800.0;
    p = p * r + 1.0 / 362880.0;
    p = p * r + 1.0 / 40320.0;
    p = p * r + 1.0 / 5040.0;
    p = p * r + 1.0 / 720.0;
    p = p * r + 1.0 / 120.0;
    p = p * r + 1.0 / 24.0;
    p = p * r + 1.0 / 6.0;
    p = p * r + 0.5;
    p = p * r + 1.0;
    p = p * r + 1.0;
    double sc = __longlong_as_double((long long)(k + 1023) << 52);
    return p * sc;
}

__device__ __forceinline__ float powf_mdas(float x, float a) {
    const float inf = __int_as_float(0x7f800000);
    if (x != x || a != a) return __int_as_float(0x7fc00000);
    if (a == 0.0f || x == 1.0f) return 1.0f;
    if (x == 0.0f) return a > 0.0f ? 0.0f : inf;
    if (x < 0.0f) return __int_as_float(0x7fc00000);
    if (x == inf) return a > 0.0f ? inf : 0.0f;
    double y = (double)a * ln_pos((double)x);
    return (float)exp_dbl(y);
}

// The selection test powf_mdas(x, a) >= u (kdgrid.cu:326) with a fast path:
// p = exp2(a * log2 x) from the hardware log / exp (relative error below
// ~1e-5 for |a log2 x| < 64) decides whenever it is more than 2^-10 away
// from u relative -- then the exactly rounded pow lies on the same side of u
// -- and only the rest evaluates powf_mdas.  Same result bit for bit.
__device__ __forceinline__ bool select_test(float x, float a, float u) {
    if (x >= 1.1754944e-38f && x <= 1.0f && a > 0.0f && u > 0.0f) {
        const float t = a * __builtin_amdgcn_logf(x);
        if (t > -64.0f) {
            const float p = __builtin_amdgcn_exp2f(t);
            if (p >= u * 1.0009765625f) return true;
            if (p <= u * 0.9990234375f) return false;
        }
    }
    return powf_mdas(x, a) >= u;
}

// kdgrid.cu:141-164 -- leaf index -> cell offset / extent of the initial grid.
// Bit k < b of coordinate d is index bit D*k + (D-1-d); for the image
// dimensions (d < 2) bit k in [b, b+e) is index bit D*b + 2(k-b) + 1 - d.
template <int D>
__device__ __forceinline__ void morton_cell(unsigned i, int b, int e, float sX, float sY,
                                            float* off, float* ext) {
#pragma unroll
    for (int d = 0; d < D; ++d) {
        unsigned xq = 0;
        unsigned inv = 1u << b;
        for (int k = 0; k < b; ++k) xq |= ((i >> (D * k + (D - d - 1))) & 1u) << k;
        if (d < 2) {
            for (int k = b; k < b + e; ++k) xq |= ((i >> (D * b + 2 * (k - b) + 1 - d)) & 1u) << k;
            inv <<= e;
        }
        ext[d] = 1.0f / (float)inv;
        off[d] = (float)xq * ext[d];
    }
    off[0] *= sX;
    off[1] *= sY;
    ext[0] *= sX;
    ext[1] *= sY;
}

// Interleave the low 16 bits of x and y (x in even bit positions).
__device__ __forceinline__ unsigned morton2(unsigned x, unsigned y) {
    auto spread = [](unsigned v) {
        v &= 0xffffu;
        v = (v | (v << 8)) & 0x00ff00ffu;
        v = (v | (v << 4)) & 0x0f0f0f0fu;
        v = (v | (v << 2)) & 0x33333333u;
        v = (v | (v << 1)) & 0x55555555u;
        return v;
    };
    return spread(x) | (spread(y) << 1);
}

// ---- wave64 / workgroup helpers ------------------------------------------
__device__ __forceinline__ int lane_id() { return __lane_id(); }

// number of set bits of `mask` below this lane
__device__ __forceinline__ int lanes_below(unsigned long long mask) {
    return __builtin_amdgcn_mbcnt_hi((unsigned)(mask >> 32),
                                     __builtin_amdgcn_mbcnt_lo((unsigned)mask, 0u));
}

__device__ __forceinline__ int wave_sum(int v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}

__device__ __forceinline__ float wave_max(float v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v = fmaxf(v, __shfl_xor(v, o, 64));
    return v;
}

// Exclusive scan of two 0/1 predicates over a 256-thread workgroup in thread
// order; returns the rank of this thread and the workgroup totals.
// `lds` needs 8 ints.  Contains two __syncthreads().
__device__ __forceinline__ void block_rank2(bool p0, bool p1, int* lds, int& r0, int& r1,
                                            int& t0, int& t1) {
    const int w = threadIdx.x >> 6;
    unsigned long long m0 = __ballot(p0);
    unsigned long long m1 = __ballot(p1);
    int l0 = lanes_below(m0), l1 = lanes_below(m1);
    if ((threadIdx.x & 63) == 0) {
        lds[w] = __popcll(m0);
        lds[4 + w] = __popcll(m1);
    }
    __syncthreads();
    int a0 = 0, a1 = 0;
    t0 = 0;
    t1 = 0;
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        int c0 = lds[k], c1 = lds[4 + k];
        if (k < w) {
            a0 += c0;
            a1 += c1;
        }
        t0 += c0;
        t1 += c1;
    }
    r0 = a0 + l0;
    r1 = a1 + l1;
    __syncthreads();
}

// Exclusive scan of an int over a 256-thread workgroup; `lds` needs 4 ints.
__device__ __forceinline__ int block_exscan(int v, int* lds, int& total) {
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    int x = v;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        int y = __shfl_up(x, o, 64);
        if (lane >= o) x += y;
    }
    if (lane == 63) lds[w] = x;
    __syncthreads();
    int pre = 0;
    total = 0;
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        int c = lds[k];
        if (k < w) pre += c;
        total += c;
    }
    __syncthreads();
    return pre + x - v;
}

// Chunk of leaves handled by workgroup `g` of `G` (multiple of 256 leaves).
__device__ __forceinline__ void chunk_range(int n, int G, int g, int& beg, int& end) {
    int c = (n + G - 1) / G;
    c = (c + 255) & ~255;
    long long b = (long long)g * c;
    beg = b < n ? (int)b : n;
    long long e = b + c;
    end = e < n ? (int)e : n;
}

}  // namespace mdas
