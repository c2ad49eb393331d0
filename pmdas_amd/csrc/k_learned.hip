// Learned proposal: per-leaf features and labels, MLP scoring fused into the
// error pass, and the MLP's training kernels (forward on MFMA, FP32
// backward, Adam).
//
// Features of a leaf (computed whenever its plain error is, from the same
// reads): log2 extent per dimension relative to the domain (6, zero above
// D), sample count, mean / variance / min / max luminance of its samples,
// depth (splits since the initial grid), log2 error, log2 error per volume,
// log2 volume, D.  Label of a split leaf: log2 of its relative error
// reduction, r = max(e_parent - e_left - e_right, 2^-10 e_parent) / e_parent,
// measured by the plain error estimate of the two children one pass later.
// Score used for selection: e * 2^clamp(y), y the network's prediction of
// log2 r; the plain error e is the fallback when the network is absent or its
// output is not finite.  Second label form: log2 of the squared image change
// the split makes (image_label below), scored as 2^clamp(y).  In general the
// score is e^kappa_e * 2^(kappa_y y) -- (1, 1) for the first form, (0, 1)
// for the second; equal exponents kappa scale the selection test
// pow(score / max, alpha) to the effective exponent kappa * alpha.
#include "mdas_leaf.hpp"
#include "mlp_device.hpp"

#pragma clang fp contract(off)

namespace mdas {

__device__ __forceinline__ float log2_floor(float x) { return __log2f(fmaxf(x, 1.0e-30f)); }

template <int D>
__device__ __forceinline__ void leaf_features_ext(const LeafSamples& ls, const float* ext, float e,
                                                  float depth, float sX, float sY, float* f) {
#pragma unroll
    for (int d = 0; d < 6; ++d) {
        if (d < D) {
            const float dom = d == 0 ? sX : (d == 1 ? sY : 1.0f);
            f[d] = log2_floor(ext[d] / dom);
        } else {
            f[d] = 0.0f;
        }
    }
    const int n = ls.n;
    float lum[4], s = 0.0f, lo = 3.402823466e+38f, hi = -3.402823466e+38f;
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        lum[q] = 0.2126f * ls.vx[q] + 0.7152f * ls.vy[q] + 0.0722f * ls.vz[q];
        if (q < n) {
            s += lum[q];
            lo = fminf(lo, lum[q]);
            hi = fmaxf(hi, lum[q]);
        }
    }
    const float mean = s / (float)n;
    float var = 0.0f;
#pragma unroll
    for (int q = 0; q < 4; ++q)
        if (q < n) var += (lum[q] - mean) * (lum[q] - mean);
    var /= (float)n;
    f[6] = (float)n;
    f[7] = mean;
    f[8] = var;
    f[9] = lo;
    f[10] = hi;
    f[11] = depth;
    f[12] = log2_floor(e);
    f[13] = log2_floor(e / ls.vol);
    f[14] = log2_floor(ls.vol);
    f[15] = (float)D;
}

// The same from the leaf's box (extents mx - mn).
template <int D>
__device__ __forceinline__ void leaf_features(const LeafSamples& ls, const float* mn,
                                              const float* mx, float e, float depth, float sX,
                                              float sY, float* f) {
    float ext[6];
#pragma unroll
    for (int d = 0; d < 6; ++d) ext[d] = d < D ? mx[d] - mn[d] : 1.0f;
    leaf_features_ext<D>(ls, ext, e, depth, sX, sY, f);
}

__device__ __forceinline__ float split_label(float ep, float el, float er) {
    if (!(ep > 0.0f)) return -10.0f;
    float d = ep - el;
    d = d - er;
    const float r = fmaxf(d, ep * 0.0009765625f) / ep;
    return __log2f(r);
}

// Image-change label: the reconstruction spreads a leaf's mean times its
// volume over its xy footprint, so refining parent p into children c moves
// the image by (m_c - m_p) v_c / a_c over each child's footprint a_c; the
// label is log2 of the squared change integrated over the footprints,
// sum_c (m_c - m_p)^2 v_c^2 / a_c (m = mean luminance, v = volume).
__device__ __forceinline__ float image_label(float mp, const float* mc, const float* vc,
                                             const float* ac) {
    float s = 0.0f;
#pragma unroll
    for (int h = 0; h < 2; ++h) {
        const float d = mc[h] - mp;
        const float w = vc[h] * vc[h] / ac[h];
        s += d * d * w;
    }
    return __log2f(s + 1.0e-30f);
}

// ---------------------------------------------------- fused error pass --
// Same leaves as the plain stale-error pass (every leaf on the first pass,
// afterwards the two children of each split, handled here as one item so the
// split's label sees both), same plain error.  Writes perr and the features;
// with collection on, the label row of each split (parent features, which
// are still at the left child's slot); err receives the score and the
// workgroup max of the scores feeds the selection's global max.
template <int D>
__device__ __forceinline__ void errors_learned_body(const DevGrid& a, const LearnDev& l, int blk,
                                                    int nblk) {
    __shared__ float wmax[4];
    __shared__ int wst[8];
    __shared__ unsigned short xsb[4][128 * kMlpIn];  // a wave's staged rows (bf16)
    __shared__ unsigned short actb[4][16 * kMlpH];   // layer-1 activations (bf16)
    __shared__ float ys[4][16];
    __shared__ __align__(16) unsigned short w1s[64 * 32], w2s[64 * 64];  // the bf16 weights
    __shared__ float bl[3 * 64 + 1];  // b1, b2, bf16-rounded w3, b3
    if (pass_stopped(a.cnt)) return;
    const bool first = a.cnt->passes == 0;
    const int n_items = first ? a.cnt->n_leaves : a.cnt->new_nodes;
    const int L_old = a.cnt->n_leaves_old;
    const long long rows_base = (first || l.X == nullptr) ? 0 : *l.rows;
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const bool feats = l.feat != nullptr;  // false: a plain member of a batch
    const bool model = feats && l.mlp.p != nullptr;
    if (model) {  // the weights once per workgroup, for every wave's tiles
        for (int k = threadIdx.x; k < 64 * 32 / 8; k += 256)
            ((u16x8*)w1s)[k] = ((const u16x8*)l.mlp.w1p)[k];
        for (int k = threadIdx.x; k < 64 * 64 / 8; k += 256)
            ((u16x8*)w2s)[k] = ((const u16x8*)l.mlp.w2p)[k];
        if (threadIdx.x < 64) {
            bl[threadIdx.x] = l.mlp.p[kOffB1 + threadIdx.x];
            bl[64 + threadIdx.x] = l.mlp.p[kOffB2 + threadIdx.x];
            bl[128 + threadIdx.x] = l.mlp.w3r[threadIdx.x];
        }
        if (threadIdx.x == 0) bl[192] = l.mlp.p[kOffB3];
        __syncthreads();
    }
    float m = 0.0f;
    int n_stale = 0, n_vals = 0;
    for (int t0 = blk * 256; t0 < n_items; t0 += nblk * 256) {
        const int t = t0 + threadIdx.x;
        const bool act = t < n_items;
        int c[2] = {-1, -1};
        if (act) {
            c[0] = first ? t : a.split_list[t];
            c[1] = first ? -1 : L_old + t;
        }
        // the parent's row is still in the left child's slot; a collected
        // row takes it now, the label later needs only its error and mean
        float depth = 0.0f, pe = 0.0f, pm = 0.0f;
        if (act && !first && feats) {
            const float* prow = l.feat + (size_t)c[0] * kFeat;
            if (l.keep) {
                pe = l.perr[c[0]];
                pm = prow[7];
                if (l.X != nullptr && rows_base + t < l.cap_rows) {
                    const float4* pr = (const float4*)prow;
                    float4* xo = (float4*)(l.X + (rows_base + t) * kFeat);
#pragma unroll
                    for (int k = 0; k < 4; ++k) xo[k] = pr[k];
                }
            }
            depth = prow[11] + 1.0f;  // scoring only: all a split carries over
        }
        float e[2] = {0.0f, 0.0f}, vol[2] = {1.0f, 1.0f}, axy[2] = {1.0f, 1.0f}, mc[2] = {0.0f, 0.0f};
        const int hmax = first ? 1 : 2;
        // one child at a time: its features are stored and staged as the
        // network's input row (row 64 h + lane) as soon as they exist, so a
        // single feature row is live
        auto finish = [&](int h, const float* f) {
            mc[h] = f[7];
            if (c[h] >= 0 && feats) {
                if (l.keep) {
                    float4* fo = (float4*)(l.feat + (size_t)c[h] * kFeat);
#pragma unroll
                    for (int k = 0; k < 4; ++k)
                        fo[k] = make_float4(f[4 * k], f[4 * k + 1], f[4 * k + 2], f[4 * k + 3]);
                    l.perr[c[h]] = e[h];
                } else {
                    l.feat[(size_t)c[h] * kFeat + 11] = f[11];
                }
            }
            if (model && h < hmax) {
                unsigned short* xr = &xsb[w][(64 * h + lane) * kMlpIn];
#pragma unroll
                for (int k = 0; k < kMlpIn; ++k)
                    xr[k] = f2bf((f[k] - l.mlp.p[kOffMu + k]) * l.mlp.p[kOffIsd + k]);
            }
        };
        if (act && !first && a.sext != nullptr) {
            // both children from the split's entries: sample indices, volumes
            // and extents (the parent's, the axis's two halves), not records
            // (a plain member of a batch has no extents and reads records)
            const int4 s0 = a.sdat[2 * t], s1 = a.sdat[2 * t + 1];
            const int4 x0 = a.sext[2 * t], x1 = a.sext[2 * t + 1];
            const int ids[2][4] = {{s0.x, s0.y, -1, -1}, {s0.z, s0.w, s1.x, -1}};
            LeafSamples ls[2];
            load_values(a, ids[0], ls[0]);  // the five value loads together
            load_values(a, ids[1], ls[1]);
            vol[0] = __int_as_float(s1.y);
            vol[1] = __int_as_float(s1.z);
            const float exl[6] = {__int_as_float(x0.x), __int_as_float(x0.y), __int_as_float(x0.z),
                                  __int_as_float(x0.w), __int_as_float(x1.x), __int_as_float(x1.y)};
#pragma unroll
            for (int h = 0; h < 2; ++h) {
                float ex[6];
#pragma unroll
                for (int d = 0; d < 6; ++d) ex[d] = (h == 1 && d == x1.w) ? __int_as_float(x1.z) : exl[d];
                float f[kFeat];
                e[h] = error_of_values(ids[h], vol[h], ls[h]);
                leaf_features_ext<D>(ls[h], ex, e[h], depth, a.sX, a.sY, f);
                axy[h] = ex[0] * ex[1];
                n_stale += 1;
                n_vals += ls[h].n;
                finish(h, f);
            }
        } else {
#pragma unroll
            for (int h = 0; h < 2; ++h) {
                float f[kFeat];
                if (c[h] >= 0) {
                    LeafSamples ls;
                    float mn[D], mx[D];
                    e[h] = leaf_error_s<D>(a, c[h], ls, mn, mx);
                    leaf_features<D>(ls, mn, mx, e[h], depth, a.sX, a.sY, f);
                    vol[h] = ls.vol;
                    axy[h] = (mx[0] - mn[0]) * (mx[1] - mn[1]);
                    n_stale += 1;
                    n_vals += ls.n;
                } else {
#pragma unroll
                    for (int k = 0; k < kFeat; ++k) f[k] = 0.0f;
                }
                finish(h, f);
            }
        }
        if (act && !first && feats && l.X != nullptr && rows_base + t < l.cap_rows)
            l.Y[rows_base + t] = l.label_mode == 0 ? split_label(pe, e[0], e[1])
                                                   : image_label(pm, mc, vol, axy);
        // scores: the wave's (up to) 128 leaves as 8 tiles of 16 rows
        float q[2] = {e[0], e[1]};
        if (model) {
            for (int tile = 0; tile < 4 * hmax; ++tile) {
                float y[4];
                mlp16_bf(w1s, w2s, bl, &xsb[w][16 * tile * kMlpIn], actb[w], y);
                if ((lane & 15) == 0) {
#pragma unroll
                    for (int i = 0; i < 4; ++i) ys[w][4 * (lane >> 4) + i] = y[i];
                }
                wave_sync_lds();
                const int h = tile >> 2;
                if ((lane >> 4) == (tile & 3)) {
#pragma unroll
                    for (int hh = 0; hh < 2; ++hh) {
                        if (hh != h || c[hh] < 0) continue;
                        const float yv = ys[w][lane & 15];
                        if (fabsf(yv) < 3.0e38f) {  // false for NaN and +-inf
                            const float yc = fminf(fmaxf(yv, l.ylo), l.yhi);
                            float s;
                            if (l.kappa_e == 1.0f && l.kappa_y == 1.0f)
                                s = e[hh] * exp2f(yc);
                            else if (l.kappa_e == 0.0f && l.kappa_y == 1.0f)
                                s = exp2f(yc);
                            else  // e^kappa_e * 2^(kappa_y y)
                                s = exp2f(l.kappa_e * log2_floor(e[hh]) + l.kappa_y * yc);
                            if (s < 3.0e38f) q[hh] = s;
                        }
                    }
                }
                wave_sync_lds();
            }
        }
#pragma unroll
        for (int h = 0; h < 2; ++h) {
            if (c[h] >= 0) {
                a.err[c[h]] = q[h];
                m = fmaxf(m, q[h]);
            }
        }
    }
    // (max, stale count, values read) per workgroup, as the plain pass
    m = wave_max(m);
    n_stale = wave_sum(n_stale);
    n_vals = wave_sum(n_vals);
    if (lane == 0) {
        wmax[w] = m;
        wst[w] = n_stale;
        wst[4 + w] = n_vals;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        const float b = fmaxf(fmaxf(wmax[0], wmax[1]), fmaxf(wmax[2], wmax[3]));
        a.epart[blk] = make_uint4(__float_as_uint(b), (unsigned)(wst[0] + wst[1] + wst[2] + wst[3]),
                                  (unsigned)(wst[4] + wst[5] + wst[6] + wst[7]), 0u);
    }
}

// 4 waves per SIMD: the feature phase and the network fit the register file
// at that occupancy without spilling (the default allocation gave 3).
template <int D>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(4, 4))) void k_errors_learned(
    DevGrid a, LearnDev l) {
    errors_learned_body<D>(a, l, blockIdx.x, gridDim.x);
}
template <int D>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(4, 4))) void k_errors_learned_b(
    const DevGrid* grids, const LearnDev* ls) {
    errors_learned_body<D>(grids[blockIdx.y], ls[blockIdx.y], blockIdx.x, gridDim.x);
}

// Folds the partials like the plain pass, then advances the row counter by
// the splits whose rows this pass offered.
__device__ __forceinline__ void err_reduce_learned_body(const DevGrid& a, const LearnDev& l, int G) {
    __shared__ unsigned sm[16];
    __shared__ int ss[16], sv[16];
    if (a.cnt->done) return;
    if (loop_limit_reached(a.cnt)) {  // the loop stops before this pass
        if (threadIdx.x == 0) a.cnt->done = 1;
        return;
    }
    unsigned mx = 0;
    int st = 0, va = 0;
    for (int k = threadIdx.x; k < G; k += 1024) {
        const uint4 p = a.epart[k];
        mx = max(mx, p.x);
        st += (int)p.y;
        va += (int)p.z;
    }
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
        mx = max(mx, (unsigned)__shfl_xor((int)mx, o, 64));
        st += __shfl_xor(st, o, 64);
        va += __shfl_xor(va, o, 64);
    }
    if ((threadIdx.x & 63) == 0) {
        sm[threadIdx.x >> 6] = mx;
        ss[threadIdx.x >> 6] = st;
        sv[threadIdx.x >> 6] = va;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        unsigned mm = a.cnt->carry_bits;  // the loop's pass: from the scan's carried max
        int s1 = 0, s2 = 0;
        for (int k = 0; k < 16; ++k) {
            mm = max(mm, sm[k]);
            s1 += ss[k];
            s2 += sv[k];
        }
        a.cnt->gmax_bits = mm;
        a.cnt->stale += s1;
        a.cnt->stale_vals += s2;
        if (l.X != nullptr && l.feat != nullptr && a.cnt->passes > 0) *l.rows += a.cnt->new_nodes;
    }
}
__global__ __launch_bounds__(1024) void k_err_reduce_learned(DevGrid a, LearnDev l, int G) {
    err_reduce_learned_body(a, l, G);
}
__global__ __launch_bounds__(1024) void k_err_reduce_learned_b(const DevGrid* grids,
                                                                const LearnDev* ls, int G) {
    err_reduce_learned_body(grids[blockIdx.x], ls[blockIdx.x], G);
}

hipError_t launch_errors_learned(Grid& g, hipStream_t st) {
    switch (g.D) {
    case 3: hipLaunchKernelGGL(k_errors_learned<3>, dim3(g.err_blocks), dim3(256), 0, st, g.dev, g.learn); break;
    case 4: hipLaunchKernelGGL(k_errors_learned<4>, dim3(g.err_blocks), dim3(256), 0, st, g.dev, g.learn); break;
    case 5: hipLaunchKernelGGL(k_errors_learned<5>, dim3(g.err_blocks), dim3(256), 0, st, g.dev, g.learn); break;
    case 6: hipLaunchKernelGGL(k_errors_learned<6>, dim3(g.err_blocks), dim3(256), 0, st, g.dev, g.learn); break;
    default: return hipErrorInvalidValue;
    }
    hipLaunchKernelGGL(k_err_reduce_learned, dim3(1), dim3(1024), 0, st, g.dev, g.learn, g.err_blocks);
    return hipGetLastError();
}

hipError_t launch_errors_learned_b(int D, const DevGrid* grids, const LearnDev* ls, int n, int blocks,
                                   hipStream_t st) {
    const int eb = blocks < 1024 ? blocks : 1024;
    switch (D) {
    case 3: hipLaunchKernelGGL(k_errors_learned_b<3>, dim3(eb, n), dim3(256), 0, st, grids, ls); break;
    case 4: hipLaunchKernelGGL(k_errors_learned_b<4>, dim3(eb, n), dim3(256), 0, st, grids, ls); break;
    case 5: hipLaunchKernelGGL(k_errors_learned_b<5>, dim3(eb, n), dim3(256), 0, st, grids, ls); break;
    case 6: hipLaunchKernelGGL(k_errors_learned_b<6>, dim3(eb, n), dim3(256), 0, st, grids, ls); break;
    default: return hipErrorInvalidValue;
    }
    hipLaunchKernelGGL(k_err_reduce_learned_b, dim3(n), dim3(1024), 0, st, grids, ls, eb);
    return hipGetLastError();
}

// ------------------------------------------------------------ the network --
// bf16 forms of the weights from the FP32 blob.
__global__ __launch_bounds__(256) void k_mlp_pack(const float* __restrict__ p, unsigned short* w1p,
                                                  unsigned short* w2p, float* w2r, float* w3r) {
    const int t = blockIdx.x * 256 + threadIdx.x;
    if (t < 64 * 32) {
        const int n = t >> 5, k = t & 31;
        w1p[t] = k < kMlpIn ? f2bf(p[kOffW1 + n * kMlpIn + k]) : (unsigned short)0;
    }
    if (t < 64 * 64) {
        const unsigned short b = f2bf(p[kOffW2 + t]);
        w2p[t] = b;
        w2r[t] = bf2f(b);
    }
    if (t < 64) w3r[t] = bfround(p[kOffW3 + t]);
}

hipError_t launch_mlp_pack(float* p, unsigned short* w1p, unsigned short* w2p, float* w2r,
                           float* w3r, hipStream_t st) {
    hipLaunchKernelGGL(k_mlp_pack, dim3(16), dim3(256), 0, st, p, w1p, w2p, w2r, w3r);
    return hipGetLastError();
}

// Inference over n rows of raw features: one wave per 16-row tile.
__global__ __launch_bounds__(256) void k_mlp_forward(MlpDev M, const float* __restrict__ X,
                                                     long long n, float* __restrict__ Y) {
    __shared__ float xs[4][16 * kMlpIn];
    __shared__ float acts[4][16 * kMlpH];
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const long long tiles = (n + 15) / 16;
    for (long long tile = (long long)blockIdx.x * 4 + w; tile < tiles; tile += (long long)gridDim.x * 4) {
        const long long r0 = tile * 16;
        if (lane < 16) {
            float f[kMlpIn];
            const long long r = r0 + lane;
#pragma unroll
            for (int k = 0; k < kMlpIn; ++k) f[k] = r < n ? X[r * kMlpIn + k] : 0.0f;
            mlp_stage_row(M, f, &xs[w][lane * kMlpIn]);
        }
        float y[4];
        mlp16(M, xs[w], acts[w], y, nullptr);
        if ((lane & 15) == 0) {
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                const long long r = r0 + 4 * (lane >> 4) + i;
                if (r < n) Y[r] = y[i];
            }
        }
    }
}

hipError_t launch_mlp_forward(const MlpDev& M, const float* X, long long n, float* Y,
                              hipStream_t st) {
    long long blocks = ((n + 15) / 16 + 3) / 4;
    if (blocks > 4096) blocks = 4096;
    if (blocks < 1) blocks = 1;
    hipLaunchKernelGGL(k_mlp_forward, dim3((unsigned)blocks), dim3(256), 0, st, M, X, n, Y);
    return hipGetLastError();
}

// Gradient of mean((y - t)^2) over a batch.  Each wave runs the MFMA forward
// of its 16-row tiles (keeping Z1, A1, Z2, A2 in LDS), then the backward in
// FP32: lane k owns hidden unit k of both layers, i.e. column k of dW2,
// row k of dW1, and entry k of db1, db2, dw3; W2's column k stays in
// registers for dA1 = W2^T dZ2.  Each wave writes its partial gradient and
// squared-error sum; the sums over waves run in a fixed order afterwards.
__global__ __launch_bounds__(256) void k_mlp_grad(MlpDev M, const float* __restrict__ X,
                                                  const float* __restrict__ T, long long n_rows,
                                                  long long start, int count,
                                                  unsigned long long pa, unsigned long long pc,
                                                  float* __restrict__ gpart,
                                                  float* __restrict__ lpart) {
    __shared__ float xs[4][16 * kMlpIn];
    __shared__ float sv[4][16 * 256];
    __shared__ float dz[4][16 * kMlpH];
    __shared__ float dys[4][16];
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const int k = lane;
    float w2c[kMlpH];
#pragma unroll
    for (int n = 0; n < kMlpH; ++n) w2c[n] = M.w2r[n * kMlpH + k];
    const float w3k = M.w3r[k];
    float g2[kMlpH], g1[kMlpIn];
#pragma unroll
    for (int n = 0; n < kMlpH; ++n) g2[n] = 0.0f;
#pragma unroll
    for (int j = 0; j < kMlpIn; ++j) g1[j] = 0.0f;
    float gb1 = 0.0f, gb2 = 0.0f, gw3 = 0.0f, gb3 = 0.0f, loss = 0.0f;
    const float scale = 2.0f / (float)count;
    const int tiles = (count + 15) / 16;
    for (int tile = blockIdx.x * 4 + w; tile < tiles; tile += gridDim.x * 4) {
        const int q0 = tile * 16;
        if (lane < 16) {
            float f[kMlpIn];
            const int q = q0 + lane;
            if (q < count) {
                const long long r = (long long)((pa * (unsigned long long)(start + q) + pc) %
                                                (unsigned long long)n_rows);
#pragma unroll
                for (int j = 0; j < kMlpIn; ++j) f[j] = X[r * kMlpIn + j];
            } else {
#pragma unroll
                for (int j = 0; j < kMlpIn; ++j) f[j] = 0.0f;
            }
            mlp_stage_row(M, f, &xs[w][lane * kMlpIn]);
        }
        float y[4];
        mlp16(M, xs[w], dz[w], y, sv[w]);
        if ((lane & 15) == 0) {
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                const int q = q0 + 4 * (lane >> 4) + i;
                float dy = 0.0f;
                if (q < count) {
                    const long long r = (long long)((pa * (unsigned long long)(start + q) + pc) %
                                                    (unsigned long long)n_rows);
                    const float diff = y[i] - T[r];
                    loss += diff * diff;
                    dy = scale * diff;
                }
                dys[w][4 * (lane >> 4) + i] = dy;
            }
        }
        wave_sync_lds();
        // layer 3 and dZ2 (lane k = hidden unit k of layer 2)
#pragma unroll 4
        for (int r = 0; r < 16; ++r) {
            const float dy = dys[w][r];
            const float* s = &sv[w][r * 256];
            gb3 += dy;
            gw3 += dy * s[192 + k];
            const float d2 = dy * w3k * gelu_grad(s[128 + k]);
            dz[w][r * kMlpH + k] = d2;
            gb2 += d2;
        }
        wave_sync_lds();
        // dW2 column k, dA1 / dZ1 of unit k, dW1 row k
        for (int r = 0; r < 16; ++r) {
            const float* d2 = &dz[w][r * kMlpH];
            const float* s = &sv[w][r * 256];
            const float a1 = s[64 + k];
            float da = 0.0f;
#pragma unroll
            for (int n = 0; n < kMlpH; ++n) {
                const float dn = d2[n];
                g2[n] += dn * a1;
                da += w2c[n] * dn;
            }
            const float d1 = da * gelu_grad(s[k]);
            gb1 += d1;
            const float* xr = &xs[w][r * kMlpIn];
#pragma unroll
            for (int j = 0; j < kMlpIn; ++j) g1[j] += d1 * xr[j];
        }
        wave_sync_lds();
    }
    // partial of this wave, parameter layout of the blob
    const int wid = blockIdx.x * 4 + w;
    float* o = gpart + (size_t)wid * kMlpParams;
#pragma unroll
    for (int j = 0; j < kMlpIn; ++j) o[kOffW1 + k * kMlpIn + j] = g1[j];
    o[kOffB1 + k] = gb1;
#pragma unroll
    for (int n = 0; n < kMlpH; ++n) o[kOffW2 + n * kMlpH + k] = g2[n];
    o[kOffB2 + k] = gb2;
    o[kOffW3 + k] = gw3;
    // gb3 and the loss are the same in every lane / in lanes 0, 16, 32, 48
    loss += __shfl_xor(loss, 16, 64);
    loss += __shfl_xor(loss, 32, 64);
    if (lane == 0) {
        o[kOffB3] = gb3;
        lpart[wid] = loss;
    }
}

int launch_mlp_grad(const MlpDev& M, const float* X, const float* T, long long n_rows,
                    long long start, int count, unsigned long long pa, unsigned long long pc,
                    float* gpart, float* lpart, hipStream_t st, hipError_t* err) {
    int blocks = ((count + 15) / 16 + 3) / 4;
    if (blocks > kMlpGradBlocks) blocks = kMlpGradBlocks;
    if (blocks < 1) blocks = 1;
    hipLaunchKernelGGL(k_mlp_grad, dim3(blocks), dim3(256), 0, st, M, X, T, n_rows, start, count,
                       pa, pc, gpart, lpart);
    *err = hipGetLastError();
    return blocks * 4;
}

// Sum of the wave partials in wave order (deterministic).
__global__ __launch_bounds__(256) void k_mlp_sum(const float* __restrict__ gpart,
                                                 const float* __restrict__ lpart, int nparts,
                                                 float* gsum, float* loss) {
    const int t = blockIdx.x * 256 + threadIdx.x;
    if (t < kMlpParams) {
        float s = 0.0f;
        for (int w = 0; w < nparts; ++w) s += gpart[(size_t)w * kMlpParams + t];
        gsum[t] = s;
    }
    if (t == 0) {
        float s = 0.0f;
        for (int w = 0; w < nparts; ++w) s += lpart[w];
        *loss = s;
    }
}

hipError_t launch_mlp_sum(const float* gpart, const float* lpart, int nparts, float* gsum,
                          float* loss, hipStream_t st) {
    hipLaunchKernelGGL(k_mlp_sum, dim3((kMlpParams + 255) / 256), dim3(256), 0, st, gpart, lpart,
                       nparts, gsum, loss);
    return hipGetLastError();
}

// Adam (beta1 0.9, beta2 0.999, eps 1e-8) on the summed gradient, FP32
// master weights, then the bf16 forms of the updated weight.  bc1 / bc2 are
// the bias corrections 1 - beta^t.  loss_acc (if non-null) accumulates the
// batch's squared-error sum.
__global__ __launch_bounds__(256) void k_mlp_adam(float* p, float* m, float* v,
                                                  const float* __restrict__ gpart,
                                                  const float* __restrict__ lpart, int nparts,
                                                  float lr, float bc1, float bc2, float* loss_acc,
                                                  unsigned short* w1p, unsigned short* w2p,
                                                  float* w2r, float* w3r) {
    const int t = blockIdx.x * 256 + threadIdx.x;
    if (t < kMlpParams) {
        float g = 0.0f;
        for (int w = 0; w < nparts; ++w) g += gpart[(size_t)w * kMlpParams + t];
        const float mt = 0.9f * m[t] + 0.1f * g;
        const float vt = 0.999f * v[t] + 0.001f * g * g;
        m[t] = mt;
        v[t] = vt;
        const float mh = mt / bc1, vh = vt / bc2;
        const float nw = p[t] - lr * mh / (sqrtf(vh) + 1.0e-8f);
        p[t] = nw;
        if (t < kOffB1) {
            const int n = t / kMlpIn, kk = t % kMlpIn;
            w1p[n * 32 + kk] = f2bf(nw);
        } else if (t >= kOffW2 && t < kOffB2) {
            const unsigned short b = f2bf(nw);
            w2p[t - kOffW2] = b;
            w2r[t - kOffW2] = bf2f(b);
        } else if (t >= kOffW3 && t < kOffB3) {
            w3r[t - kOffW3] = bfround(nw);
        }
    }
    if (t == 0 && loss_acc != nullptr) {
        float s = 0.0f;
        for (int w = 0; w < nparts; ++w) s += lpart[w];
        *loss_acc += s;
    }
}

hipError_t launch_mlp_adam(float* p, float* m, float* v, const float* gpart, const float* lpart,
                           int nparts, float lr, float bc1, float bc2, float* loss_acc,
                           unsigned short* w1p, unsigned short* w2p, float* w2r, float* w3r,
                           hipStream_t st) {
    hipLaunchKernelGGL(k_mlp_adam, dim3((kMlpParams + 255) / 256), dim3(256), 0, st, p, m, v, gpart,
                       lpart, nparts, lr, bc1, bc2, loss_acc, w1p, w2p, w2r, w3r);
    return hipGetLastError();
}

// Per-workgroup sums and sums of squares (double) of the 16 features.
__global__ __launch_bounds__(256) void k_feat_stats(const float* __restrict__ X, long long n,
                                                    double* part) {
    __shared__ double red[2 * kMlpIn][4];
    const int j = threadIdx.x & 15;
    double s = 0.0, s2 = 0.0;
    for (long long r = (long long)blockIdx.x * 16 + (threadIdx.x >> 4); r < n;
         r += (long long)gridDim.x * 16) {
        const double x = X[r * kMlpIn + j];
        s += x;
        s2 += x * x;
    }
    // 16 rows per pass: lanes j, j+16, j+32, j+48 of a wave share feature j
    s += __shfl_xor(s, 16, 64);
    s += __shfl_xor(s, 32, 64);
    s2 += __shfl_xor(s2, 16, 64);
    s2 += __shfl_xor(s2, 32, 64);
    const int w = threadIdx.x >> 6;
    if ((threadIdx.x & 63) < 16) {
        red[j][w] = s;
        red[kMlpIn + j][w] = s2;
    }
    __syncthreads();
    if (threadIdx.x < 2 * kMlpIn) {
        const double* r = red[threadIdx.x];
        part[(size_t)blockIdx.x * 2 * kMlpIn + threadIdx.x] = ((r[0] + r[1]) + r[2]) + r[3];
    }
}

hipError_t launch_feat_stats(const float* X, long long n, double* part, int nblocks,
                             hipStream_t st) {
    hipLaunchKernelGGL(k_feat_stats, dim3(nblocks), dim3(256), 0, st, X, n, part);
    return hipGetLastError();
}

}  // namespace mdas
