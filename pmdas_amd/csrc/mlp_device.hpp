// Device side of the learned proposal: a 16 -> 64 -> 64 -> 1 MLP with GELU
// (tanh form), BF16 weights and layer inputs on MFMA, FP32 accumulation.
//
// One wave evaluates 16 rows at a time with v_mfma_f32_16x16x32_bf16
// (gfx950).  A is 16 x 32 (lane l holds row l%16, k = 8*(l/16) .. +7), B is
// 32 x 16 (lane l holds column l%16, the same k range) and D is 16 x 16 (lane
// l holds rows 4*(l/16) + i, i = 0..3, column l%16).  A and B use the same
// k assignment, so the product does not depend on how the unit orders k.
// Layer 1 pads K = 16 to 32 with zeros; layer 2 is two K chunks of 32;
// layer 3 is a dot product of the D-layout activations, reduced across the
// 16 lanes that hold the same rows.
#pragma once
#include <hip/hip_runtime.h>

#include "mdas_internal.hpp"

namespace mdas {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));
typedef unsigned short u16x8 __attribute__((ext_vector_type(8)));

static const int kMlpIn = 16, kMlpH = 64;
// Parameter blob (floats): W1 [64][16], b1 [64], W2 [64][64], b2 [64],
// w3 [64], b3 [1] -- the trained parameters -- then the feature
// normalisation mu [16], 1/sigma [16].
static const int kOffW1 = 0, kOffB1 = 1024, kOffW2 = 1088, kOffB2 = 5184, kOffW3 = 5248,
                 kOffB3 = 5312, kMlpParams = 5313, kOffMu = 5313, kOffIsd = 5329,
                 kMlpFloats = 5345;

// Round-to-nearest-even float -> bf16 bits (finite inputs).
__host__ __device__ inline unsigned short f2bf(float f) {
    unsigned u;
    __builtin_memcpy(&u, &f, 4);
    u += 0x7fffu + ((u >> 16) & 1u);
    return (unsigned short)(u >> 16);
}
__host__ __device__ inline float bf2f(unsigned short h) {
    unsigned u = (unsigned)h << 16;
    float f;
    __builtin_memcpy(&f, &u, 4);
    return f;
}
__host__ __device__ inline float bfround(float f) { return bf2f(f2bf(f)); }

// GELU, tanh form: 0.5 x (1 + tanh(sqrt(2/pi) (x + 0.044715 x^3))), with
// tanh(z) = 1 - 2 / (exp(2z) + 1) (exact limits +-1 for large |z|).  The
// quotient uses the hardware reciprocal (about 1 ulp) instead of the
// correctly rounded division the build asks for elsewhere: the network's
// activations are rounded to bf16 right after, and the division was the
// largest part of the activation's cost.
__device__ __forceinline__ float gelu_tanh(float x) {
    const float z = 0.7978845608f * (x + 0.044715f * x * x * x);
    return 1.0f - 2.0f * __builtin_amdgcn_rcpf(__expf(2.0f * z) + 1.0f);
}
__device__ __forceinline__ float gelu(float x) { return 0.5f * x * (1.0f + gelu_tanh(x)); }
__device__ __forceinline__ float gelu_grad(float x) {
    const float t = gelu_tanh(x);
    return 0.5f * (1.0f + t) +
           0.5f * x * (1.0f - t * t) * 0.7978845608f * (1.0f + 3.0f * 0.044715f * x * x);
}

__device__ __forceinline__ bf16x8 load_b8(const unsigned short* p) {
    const u16x8 v = *(const u16x8*)p;
    return __builtin_bit_cast(bf16x8, v);
}

__device__ __forceinline__ bf16x8 pack_b8(const float* f) {
    u16x8 v;
#pragma unroll
    for (int k = 0; k < 8; ++k) v[k] = f2bf(f[k]);
    return __builtin_bit_cast(bf16x8, v);
}

// Orders the LDS accesses of one wave (all lanes of the wave take part).
__device__ __forceinline__ void wave_sync_lds() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Normalised, bf16-rounded feature row r of the tile into x (LDS [16][16]).
__device__ __forceinline__ void mlp_stage_row(const MlpDev& M, const float* f, float* xrow) {
#pragma unroll
    for (int k = 0; k < kMlpIn; ++k)
        xrow[k] = bfround((f[k] - M.p[kOffMu + k]) * M.p[kOffIsd + k]);
}

// Forward pass of 16 rows by one wave (every lane calls it).  x (LDS,
// [16][16]) holds the staged rows, act (LDS, [16][64]) is scratch.  Returns
// in y[i] the output of row 4*(lane/16)+i.  When save is non-null (LDS,
// [16][256]) it receives per row the pre-activations Z1 (0..63), the bf16
// layer-2 inputs A1 (64..127), Z2 (128..191) and the bf16 A2 (192..255).
__device__ __forceinline__ void mlp16(const MlpDev& M, const float* x, float* act, float y[4],
                                      float* save) {
    const int lane = threadIdx.x & 63;
    const int col = lane & 15, kg = lane >> 4;
    wave_sync_lds();
    float av[8];
#pragma unroll
    for (int k = 0; k < 8; ++k) {
        const int kk = 8 * kg + k;
        av[k] = kk < kMlpIn ? x[col * kMlpIn + kk] : 0.0f;
    }
    const bf16x8 a1 = pack_b8(av);
    f32x4 z[4];
#pragma unroll
    for (int t = 0; t < 4; ++t) {
        const bf16x8 b = load_b8(M.w1p + (16 * t + col) * 32 + 8 * kg);
        const f32x4 c = {0.0f, 0.0f, 0.0f, 0.0f};
        z[t] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a1, b, c, 0, 0, 0);
    }
#pragma unroll
    for (int t = 0; t < 4; ++t) {
        const int n = 16 * t + col;
        const float bias = M.p[kOffB1 + n];
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            const int r = 4 * kg + i;
            const float zz = z[t][i] + bias;
            const float a = bfround(gelu(zz));
            act[r * kMlpH + n] = a;
            if (save) {
                save[r * 256 + n] = zz;
                save[r * 256 + 64 + n] = a;
            }
        }
    }
    wave_sync_lds();
    bf16x8 a2[2];
#pragma unroll
    for (int kc = 0; kc < 2; ++kc) {
        float v[8];
#pragma unroll
        for (int k = 0; k < 8; ++k) v[k] = act[col * kMlpH + 32 * kc + 8 * kg + k];
        a2[kc] = pack_b8(v);
    }
    float part[4] = {0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
    for (int t = 0; t < 4; ++t) {
        f32x4 c = {0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
        for (int kc = 0; kc < 2; ++kc) {
            const bf16x8 b = load_b8(M.w2p + (16 * t + col) * 64 + 32 * kc + 8 * kg);
            c = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a2[kc], b, c, 0, 0, 0);
        }
        const int n = 16 * t + col;
        const float bias = M.p[kOffB2 + n], w3 = M.w3r[n];
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            const float zz = c[i] + bias;
            const float a = bfround(gelu(zz));
            part[i] += a * w3;
            if (save) {
                const int r = 4 * kg + i;
                save[r * 256 + 128 + n] = zz;
                save[r * 256 + 192 + n] = a;
            }
        }
    }
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        float s = part[i];
#pragma unroll
        for (int o = 1; o < 16; o <<= 1) s += __shfl_xor(s, o, 64);
        y[i] = s + M.p[kOffB3];
    }
    wave_sync_lds();
}

// The same forward pass from rows staged as bf16 bits (x: LDS [16][16]) with
// the bf16 layer-1 activations kept as bf16 bits (act: LDS [16][64]): the
// values are those of mlp16 (its rows and activations are rounded to bf16
// before use), in half the LDS; the bf16 weights are read from the
// workgroup's LDS copies w1s [64][32] and w2s [64][64], and b1, b2, the
// bf16-rounded w3 and b3 from its LDS copy bl [64 + 64 + 64 + 1].
// Inference only.
__device__ __forceinline__ void mlp16_bf(const unsigned short* w1s, const unsigned short* w2s,
                                         const float* bl, const unsigned short* x,
                                         unsigned short* act, float y[4]) {
    const int lane = threadIdx.x & 63;
    const int col = lane & 15, kg = lane >> 4;
    wave_sync_lds();
    u16x8 av;
#pragma unroll
    for (int k = 0; k < 8; ++k) {
        const int kk = 8 * kg + k;
        av[k] = kk < kMlpIn ? x[col * kMlpIn + kk] : (unsigned short)0;
    }
    const bf16x8 a1 = __builtin_bit_cast(bf16x8, av);
    f32x4 z[4];
#pragma unroll
    for (int t = 0; t < 4; ++t) {
        const bf16x8 b = load_b8(w1s + (16 * t + col) * 32 + 8 * kg);
        const f32x4 c = {0.0f, 0.0f, 0.0f, 0.0f};
        z[t] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a1, b, c, 0, 0, 0);
    }
#pragma unroll
    for (int t = 0; t < 4; ++t) {
        const int n = 16 * t + col;
        const float bias = bl[n];
#pragma unroll
        for (int i = 0; i < 4; ++i) act[(4 * kg + i) * kMlpH + n] = f2bf(gelu(z[t][i] + bias));
    }
    wave_sync_lds();
    bf16x8 a2[2];
#pragma unroll
    for (int kc = 0; kc < 2; ++kc) {
        u16x8 v;
#pragma unroll
        for (int k = 0; k < 8; ++k) v[k] = act[col * kMlpH + 32 * kc + 8 * kg + k];
        a2[kc] = __builtin_bit_cast(bf16x8, v);
    }
    float part[4] = {0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
    for (int t = 0; t < 4; ++t) {
        f32x4 c = {0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
        for (int kc = 0; kc < 2; ++kc) {
            const bf16x8 b = load_b8(w2s + (16 * t + col) * 64 + 32 * kc + 8 * kg);
            c = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a2[kc], b, c, 0, 0, 0);
        }
        const int n = 16 * t + col;
        const float bias = bl[64 + n], w3 = bl[128 + n];
#pragma unroll
        for (int i = 0; i < 4; ++i) part[i] += bfround(gelu(c[i] + bias)) * w3;
    }
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        float s = part[i];
#pragma unroll
        for (int o = 1; o < 16; o <<= 1) s += __shfl_xor(s, o, 64);
        y[i] = s + bl[192];
    }
    wave_sync_lds();
}

}  // namespace mdas
