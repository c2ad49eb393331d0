// Reconstruction: box-filter of the leaves into pixels (kdgrid.cu:483-582),
// done as a deterministic gather instead of the reference's float atomics.
//
//  1. k_recon_count / k_scan_keys / k_recon_emit: every leaf emits one 64-bit
//     key per pixel tile its pixel span touches,
//         key = morton2(tile_x, tile_y) << 32 | leaf index,
//     in leaf order (a workgroup-chunked exclusive scan gives the offsets).
//  2. radix_sort_keys: stable LSD radix sort (8-bit digits, LDS-staged
//     scatter) on the Morton bits only; since keys are emitted in leaf order
//     the result is the total 64-bit order.
//  3. k_tile_ranges: start / end of every tile's key segment.
//  4. k_gather: one wave per 64-pixel tile walks the tile's leaves in index
//     order (staged 64 at a time in LDS) and every lane accumulates its
//     pixel.  Each contribution is computed with the reference's own
//     expressions, including its accumulated pixel bounds (pMax += step), so a
//     pixel is the sum of its contributions in leaf index order.
#include "mdas_device.hpp"
#include "mdas_internal.hpp"

#pragma clang fp contract(off)

namespace mdas {

static const int kSortBlocks = 1024;

struct ReconGeo {
    int W, H;
    float fW, fH;
    float sX, sY;
    float stepX, stepY;   // scale / width (kdgrid.cu:501-503)
    float invPixArea;     // 1 / ((sX*sY) / float(W*H)) (kdgrid.cu:532, vec_math.h:483-487)
    int TW, TH, tilesX, tilesY;
    int tbits;            // Morton bits per tile axis
};

// kdgrid.cu:524-529: pixel span [x0, x1) x [y0, y1) of a leaf box
__device__ __forceinline__ void pixel_span(float mn0, float mx0, float mn1, float mx1,
                                           const ReconGeo& G, int& x0, int& x1, int& y0,
                                           int& y1) {
    x0 = (int)(mn0 * G.fW / G.sX);
    int xx = (int)(mx0 * G.fW / G.sX);
    y0 = (int)(mn1 * G.fH / G.sY);
    int yy = (int)(mx1 * G.fH / G.sY);
    x1 = min(xx + 1, G.W);
    y1 = min(yy + 1, G.H);
}

__device__ __forceinline__ int tiles_of(int x0, int x1, int y0, int y1, const ReconGeo& G) {
    if (x0 >= x1 || y0 >= y1) return 0;
    return ((x1 - 1) / G.TW - x0 / G.TW + 1) * ((y1 - 1) / G.TH - y0 / G.TH + 1);
}

__global__ __launch_bounds__(256) void k_recon_count(DevGrid a, ReconGeo G) {
    __shared__ int red[4];
    const int L = a.cnt->n_leaves;
    const int cap = a.cap_l;
    int beg, end;
    chunk_range(L, gridDim.x, blockIdx.x, beg, end);
    int c = 0;
    for (int i = beg + threadIdx.x; i < end; i += 256) {
        int x0, x1, y0, y1;
        pixel_span(a.bmn[i], a.bmx[i], a.bmn[(size_t)cap + i], a.bmx[(size_t)cap + i], G, x0, x1,
                   y0, y1);
        c += tiles_of(x0, x1, y0, y1, G);
    }
    c = wave_sum(c);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = c;
    __syncthreads();
    if (threadIdx.x == 0) a.part[blockIdx.x] = make_int2(red[0] + red[1] + red[2] + red[3], 0);
}

// Exclusive scan of part[].x into part_off[].x; total -> cnt->n_keys.
__global__ __launch_bounds__(1024) void k_scan_keys(DevGrid a, int G) {
    __shared__ int ls[16];
    const int per = (G + 1023) / 1024;
    const int g0 = threadIdx.x * per;
    int s = 0;
    for (int k = 0; k < per; ++k)
        if (g0 + k < G) s += a.part[g0 + k].x;
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    int x = s;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        int y = __shfl_up(x, o, 64);
        if (lane >= o) x += y;
    }
    if (lane == 63) ls[w] = x;
    __syncthreads();
    int p = 0, t = 0;
    for (int k = 0; k < 16; ++k) {
        if (k < w) p += ls[k];
        t += ls[k];
    }
    int r = p + x - s;
    for (int k = 0; k < per; ++k)
        if (g0 + k < G) {
            int c = a.part[g0 + k].x;
            a.part_off[g0 + k] = make_int2(r, 0);
            r += c;
        }
    if (threadIdx.x == 0) a.cnt->n_keys = t;
}

__global__ __launch_bounds__(256) void k_recon_emit(DevGrid a, ReconGeo G,
                                                    unsigned long long* keys, long long cap_keys) {
    __shared__ int lds[4];
    const int L = a.cnt->n_leaves;
    const int cap = a.cap_l;
    if ((long long)a.cnt->n_keys > cap_keys) return;  // host grows the buffer first
    int beg, end;
    chunk_range(L, gridDim.x, blockIdx.x, beg, end);
    int base = a.part_off[blockIdx.x].x;
    for (int t0 = beg; t0 < end; t0 += 256) {
        const int i = t0 + threadIdx.x;
        int x0 = 0, x1 = 0, y0 = 0, y1 = 0, c = 0;
        if (i < end) {
            pixel_span(a.bmn[i], a.bmx[i], a.bmn[(size_t)cap + i], a.bmx[(size_t)cap + i], G, x0,
                       x1, y0, y1);
            c = tiles_of(x0, x1, y0, y1, G);
        }
        int total;
        int off = base + block_exscan(c, lds, total);
        if (c > 0) {
            for (int ty = y0 / G.TH; ty <= (y1 - 1) / G.TH; ++ty)
                for (int tx = x0 / G.TW; tx <= (x1 - 1) / G.TW; ++tx)
                    keys[off++] = ((unsigned long long)morton2(tx, ty) << 32) | (unsigned)i;
        }
        base += total;
    }
}

// ------------------------------------------------------------ radix sort --
// Stable LSD radix sort of 64-bit keys over bits [shift, shift+8) per pass.
// Reduce-then-scan: per-workgroup digit histograms, one exclusive scan over
// (digit, workgroup), then a scatter in which every workgroup ranks a tile of
// 256 keys by digit with wave ballots, stages the tile in LDS in digit order
// and writes each digit run contiguously.
__global__ __launch_bounds__(256) void k_rs_hist(const unsigned long long* keys, const int* n_ptr,
                                                 int shift, unsigned* hist) {
    __shared__ unsigned h[256];
    const int n = *n_ptr;
    h[threadIdx.x] = 0;
    __syncthreads();
    int beg, end;
    chunk_range(n, gridDim.x, blockIdx.x, beg, end);
    for (int i = beg + threadIdx.x; i < end; i += 256)
        atomicAdd(&h[(unsigned)(keys[i] >> shift) & 255u], 1u);
    __syncthreads();
    hist[threadIdx.x * gridDim.x + blockIdx.x] = h[threadIdx.x];
}

__global__ __launch_bounds__(1024) void k_rs_scan(unsigned* hist, int total) {
    __shared__ unsigned ls[16];
    const int per = (total + 1023) / 1024;
    const int g0 = threadIdx.x * per;
    unsigned s = 0;
    for (int k = 0; k < per; ++k)
        if (g0 + k < total) s += hist[g0 + k];
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    unsigned x = s;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        unsigned y = __shfl_up(x, o, 64);
        if (lane >= o) x += y;
    }
    if (lane == 63) ls[w] = x;
    __syncthreads();
    unsigned p = 0;
    for (int k = 0; k < w; ++k) p += ls[k];
    unsigned r = p + x - s;
    for (int k = 0; k < per; ++k)
        if (g0 + k < total) {
            unsigned c = hist[g0 + k];
            hist[g0 + k] = r;
            r += c;
        }
}

__global__ __launch_bounds__(256) void k_rs_scatter(const unsigned long long* in,
                                                    unsigned long long* out, const int* n_ptr,
                                                    int shift, const unsigned* hist) {
    __shared__ unsigned base[256];
    __shared__ int wcnt[4][256];
    __shared__ int toff[256];
    __shared__ int ttot[256];
    __shared__ int scr[4];
    __shared__ unsigned long long stage[256];
    const int n = *n_ptr;
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    base[threadIdx.x] = hist[threadIdx.x * gridDim.x + blockIdx.x];
    int beg, end;
    chunk_range(n, gridDim.x, blockIdx.x, beg, end);
    for (int t0 = beg; t0 < end; t0 += 256) {
        const int i = t0 + threadIdx.x;
        const bool valid = i < end;
        const unsigned long long k = valid ? in[i] : 0ull;
        const unsigned d = (unsigned)(k >> shift) & 255u;
        for (int q = 0; q < 4; ++q) wcnt[q][threadIdx.x] = 0;
        __syncthreads();
        unsigned long long m = __ballot(valid);
#pragma unroll
        for (int b = 0; b < 8; ++b) {
            const bool bit = (d >> b) & 1u;
            const unsigned long long bal = __ballot(bit);
            m &= bit ? bal : ~bal;
        }
        const int rw = lanes_below(m);
        if (valid && rw == 0) wcnt[w][d] = __popcll(m);
        __syncthreads();
        {
            const int t = threadIdx.x;
            int tot = wcnt[0][t] + wcnt[1][t] + wcnt[2][t] + wcnt[3][t];
            int all;
            toff[t] = block_exscan(tot, scr, all);
            ttot[t] = tot;
        }
        __syncthreads();
        int nvalid = min(256, end - t0);
        if (valid) {
            int pre = 0;
            for (int q = 0; q < w; ++q) pre += wcnt[q][d];
            stage[toff[d] + pre + rw] = k;
        }
        __syncthreads();
        if (threadIdx.x < nvalid) {
            const unsigned long long kk = stage[threadIdx.x];
            const unsigned dd = (unsigned)(kk >> shift) & 255u;
            out[base[dd] + (threadIdx.x - toff[dd])] = kk;
        }
        __syncthreads();
        base[threadIdx.x] += ttot[threadIdx.x];
        __syncthreads();
    }
    (void)lane;
}

// Sorts keys[0, n) by bits [lo, hi); returns the buffer holding the result.
static unsigned long long* radix_sort_keys(Grid& g, int lo, int hi, hipStream_t st) {
    unsigned long long* a = g.keys;
    unsigned long long* b = g.keys_alt;
    const int* n_ptr = &g.dev.cnt->n_keys;
    g.recon_sort_passes = 0;
    for (int sh = lo; sh < hi; sh += 8) {
        kbeg(g, st, "k_rs_hist");
        hipLaunchKernelGGL(k_rs_hist, dim3(kSortBlocks), dim3(256), 0, st, a, n_ptr, sh, g.rs_hist);
        kend(g, st);
        kbeg(g, st, "k_rs_scan");
        hipLaunchKernelGGL(k_rs_scan, dim3(1), dim3(1024), 0, st, g.rs_hist, 256 * kSortBlocks);
        kend(g, st);
        kbeg(g, st, "k_rs_scatter");
        hipLaunchKernelGGL(k_rs_scatter, dim3(kSortBlocks), dim3(256), 0, st, a, b, n_ptr, sh,
                           g.rs_hist);
        kend(g, st);
        g.recon_sort_passes++;
        unsigned long long* t = a;
        a = b;
        b = t;
    }
    return a;
}

__global__ __launch_bounds__(256) void k_tile_ranges(const unsigned long long* keys,
                                                     const int* n_ptr, int* start, int* end) {
    const int n = *n_ptr;
    for (int p = blockIdx.x * 256 + threadIdx.x; p < n; p += gridDim.x * 256) {
        const unsigned t = (unsigned)(keys[p] >> 32);
        if (p == 0 || (unsigned)(keys[p - 1] >> 32) != t) start[t] = p;
        if (p == n - 1 || (unsigned)(keys[p + 1] >> 32) != t) end[t] = p + 1;
    }
}

// ---------------------------------------------------------------- gather --
template <int D>
__global__ __launch_bounds__(256) void k_gather(DevGrid a, ReconGeo G,
                                                const unsigned long long* __restrict__ keys,
                                                const int* __restrict__ tstart,
                                                const int* __restrict__ tend, float4* pixels) {
    __shared__ int rx0[4][64], rx1[4][64], ry0[4][64], ry1[4][64];
    __shared__ float rmn0[4][64], rmx0[4][64], rmn1[4][64], rmx1[4][64];
    __shared__ float rvol[4][64], rdxy[4][64], rax[4][64], ray[4][64], raz[4][64];
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const int tile = blockIdx.x * 4 + w;
    if (tile >= G.tilesX * G.tilesY) return;
    const int tx = tile % G.tilesX, ty = tile / G.tilesX;
    const int px = tx * G.TW + lane % G.TW;
    const int py = ty * G.TH + lane / G.TW;
    const unsigned code = morton2(tx, ty);
    const int s0 = tstart[code], s1 = tend[code];
    const int cap = a.cap_l;
    float accx = 0.0f, accy = 0.0f, accz = 0.0f;
    for (int base = s0; base < s1; base += 64) {
        const int nrec = min(64, s1 - base);
        if (lane < nrec) {
            const int i = (int)(unsigned)(keys[base + lane] & 0xffffffffull);
            float mn[D], mx[D];
#pragma unroll
            for (int d = 0; d < D; ++d) {
                mn[d] = a.bmn[(size_t)d * cap + i];
                mx[d] = a.bmx[(size_t)d * cap + i];
            }
            int x0, x1, y0, y1;
            pixel_span(mn[0], mx[0], mn[1], mx[1], G, x0, x1, y0, y1);
            const int4 id4 = a.idx[i];
            const int ids[4] = {id4.x, id4.y, id4.z, id4.w};
            float sx = 0.0f, sy = 0.0f, sz = 0.0f;
            int n = 0;
#pragma unroll
            for (int q = 0; q < 4; ++q)
                if (ids[q] >= 0) {
                    const float* v = a.values + (size_t)ids[q] * 3;
                    sx += v[0];
                    sy += v[1];
                    sz += v[2];
                    ++n;
                }
            const float inv = 1.0f / (float)n;
            float vol = 1.0f;
#pragma unroll
            for (int d = 0; d < D; ++d) vol *= (mx[d] - mn[d]);
            rx0[w][lane] = x0;
            rx1[w][lane] = x1;
            ry0[w][lane] = y0;
            ry1[w][lane] = y1;
            rmn0[w][lane] = mn[0];
            rmx0[w][lane] = mx[0];
            rmn1[w][lane] = mn[1];
            rmx1[w][lane] = mx[1];
            rvol[w][lane] = vol;
            rdxy[w][lane] = (mx[0] - mn[0]) * (mx[1] - mn[1]);
            rax[w][lane] = sx * inv;
            ray[w][lane] = sy * inv;
            raz[w][lane] = sz * inv;
        }
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
        for (int j = 0; j < nrec; ++j) {
            const int x0 = rx0[w][j], x1 = rx1[w][j], y0 = ry0[w][j], y1 = ry1[w][j];
            if (px >= x0 && px < x1 && py >= y0 && py < y1) {
                // accumulated pixel bounds exactly as kdgrid.cu:536-576 walks them
                float pMinX = (float)x0 * G.sX / G.fW;
                float pMaxX = pMinX + G.stepX;
                for (int q = x0; q < px; ++q) {
                    pMinX = pMaxX;
                    pMaxX += G.stepX;
                }
                float pMinY = (float)y0 * G.sY / G.fH;
                float pMaxY = pMinY + G.stepY;
                for (int q = y0; q < py; ++q) {
                    pMinY = pMaxY;
                    pMaxY += G.stepY;
                }
                const float clox = fmaxf(pMinX, rmn0[w][j]);
                const float chix = fminf(pMaxX, rmx0[w][j]);
                const float cloy = fmaxf(pMinY, rmn1[w][j]);
                const float chiy = fminf(pMaxY, rmx1[w][j]);
                if (clox < chix && cloy < chiy) {
                    const float v = (chix - clox) * (chiy - cloy) * rvol[w][j] / rdxy[w][j];
                    accx += (rax[w][j] * v) * G.invPixArea;
                    accy += (ray[w][j] * v) * G.invPixArea;
                    accz += (raz[w][j] * v) * G.invPixArea;
                }
            }
        }
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    }
    if (px < G.W && py < G.H) pixels[(size_t)py * G.W + px] = make_float4(accx, accy, accz, 0.0f);
}

// -------------------------------------------------------- scatter (atomics) --
// kdgrid.cu:483-582 as written: one thread per leaf, float atomics per pixel.
template <int D>
__global__ __launch_bounds__(256) void k_scatter(DevGrid a, ReconGeo G, float4* pixels) {
    const int L = a.cnt->n_leaves;
    const int cap = a.cap_l;
    for (int i = blockIdx.x * 256 + threadIdx.x; i < L; i += gridDim.x * 256) {
        float mn[D], mx[D];
#pragma unroll
        for (int d = 0; d < D; ++d) {
            mn[d] = a.bmn[(size_t)d * cap + i];
            mx[d] = a.bmx[(size_t)d * cap + i];
        }
        const int4 id4 = a.idx[i];
        const int ids[4] = {id4.x, id4.y, id4.z, id4.w};
        float sx = 0.0f, sy = 0.0f, sz = 0.0f;
        int n = 0;
#pragma unroll
        for (int q = 0; q < 4; ++q)
            if (ids[q] >= 0) {
                const float* v = a.values + (size_t)ids[q] * 3;
                sx += v[0];
                sy += v[1];
                sz += v[2];
                ++n;
            }
        const float inv = 1.0f / (float)n;
        sx *= inv;
        sy *= inv;
        sz *= inv;
        float vol = 1.0f;
#pragma unroll
        for (int d = 0; d < D; ++d) vol *= (mx[d] - mn[d]);
        int x0, x1, y0, y1;
        pixel_span(mn[0], mx[0], mn[1], mx[1], G, x0, x1, y0, y1);
        float pMinY = (float)y0 * G.sY / G.fH;
        float pMaxY = pMinY + G.stepY;
        for (int y = y0; y < y1; ++y) {
            float pMinX = (float)x0 * G.sX / G.fW;
            float pMaxX = pMinX + G.stepX;
            for (int x = x0; x < x1; ++x) {
                const float clox = fmaxf(pMinX, mn[0]);
                const float chix = fminf(pMaxX, mx[0]);
                const float cloy = fmaxf(pMinY, mn[1]);
                const float chiy = fminf(pMaxY, mx[1]);
                if (clox < chix && cloy < chiy) {
                    const float v = (chix - clox) * (chiy - cloy) * vol / ((mx[0] - mn[0]) * (mx[1] - mn[1]));
                    float* p = (float*)(pixels + (size_t)y * G.W + x);
                    atomicAdd(p, (sx * v) * G.invPixArea);
                    atomicAdd(p + 1, (sy * v) * G.invPixArea);
                    atomicAdd(p + 2, (sz * v) * G.invPixArea);
                }
                pMinX = pMaxX;
                pMaxX += G.stepX;
            }
            pMinY = pMaxY;
            pMaxY += G.stepY;
        }
    }
}

// ----------------------------------------------------------------- bytes --
// helpers.h:35-66: clamp, sRGB (pow with exponent 1/2.4), 8-bit quantise.
__device__ __forceinline__ unsigned char quantize8(float x) {
    x = fmaxf(0.0f, fminf(x, 1.0f));
    unsigned v = (unsigned)(x * 256.0f);
    return (unsigned char)min(v, 255u);
}

__global__ __launch_bounds__(256) void k_bytes(const float4* px, uchar4* out, long long n) {
    const float invGamma = 1.0f / 2.4f;
    for (long long i = blockIdx.x * 256ll + threadIdx.x; i < n; i += (long long)gridDim.x * 256) {
        const float4 c = px[i];
        const float in[3] = {c.x, c.y, c.z};
        unsigned char o[3];
#pragma unroll
        for (int ch = 0; ch < 3; ++ch) {
            const float x = fmaxf(0.0f, fminf(in[ch], 1.0f));
            const float p = powf_mdas(x, invGamma);
            const float s = x < 0.0031308f ? 12.92f * x : 1.055f * p - 0.055f;
            o[ch] = quantize8(s);
        }
        out[i] = make_uchar4(o[0], o[1], o[2], 255);
    }
}

// ------------------------------------------------------------ launchers --
static ReconGeo make_geo(const Grid& g, int W, int H) {
    ReconGeo G;
    G.W = W;
    G.H = H;
    G.fW = (float)W;
    G.fH = (float)H;
    G.sX = g.sX;
    G.sY = g.sY;
    G.stepX = g.sX / (float)W;
    G.stepY = g.sY / (float)H;
    const float pixArea = (g.sX * g.sY) / (float)(W * H);
    G.invPixArea = 1.0f / pixArea;
    G.TW = g.tile_w;
    G.TH = g.tile_h;
    G.tilesX = (W + G.TW - 1) / G.TW;
    G.tilesY = (H + G.TH - 1) / G.TH;
    int m = G.tilesX > G.tilesY ? G.tilesX : G.tilesY;
    int bits = 0;
    while ((1 << bits) < m) ++bits;
    G.tbits = bits;
    return G;
}

#define MDAS_DISPATCH_D(D_, STMT)                      \
    switch (D_) {                                      \
    case 3: { constexpr int DD = 3; STMT; } break;     \
    case 4: { constexpr int DD = 4; STMT; } break;     \
    case 5: { constexpr int DD = 5; STMT; } break;     \
    case 6: { constexpr int DD = 6; STMT; } break;     \
    default: return hipErrorInvalidValue;              \
    }

// Count + scan of the reconstruction keys, then (synchronising) make sure the
// key buffers hold them.
hipError_t recon_prepare(Grid& g, int W, int H, hipStream_t st) {
    ReconGeo G = make_geo(g, W, H);
    g.recon_pixels = (long long)W * H;
    kbeg(g, st, "k_recon_count");
    hipLaunchKernelGGL(k_recon_count, dim3(kChunkBlocks), dim3(256), 0, st, g.dev, G);
    kend(g, st);
    kbeg(g, st, "k_scan_keys");
    hipLaunchKernelGGL(k_scan_keys, dim3(1), dim3(1024), 0, st, g.dev, kChunkBlocks);
    kend(g, st);
    int nk = 0;
    hipError_t e = hipMemcpyAsync(&nk, &g.dev.cnt->n_keys, sizeof(int), hipMemcpyDeviceToHost, st);
    if (e != hipSuccess) return e;
    e = hipStreamSynchronize(st);
    if (e != hipSuccess) return e;
    g.last_n_keys = nk;
    if (nk > g.keys_cap) {
        if (g.keys) hipFree(g.keys);
        if (g.keys_alt) hipFree(g.keys_alt);
        g.keys = g.keys_alt = nullptr;
        long long cap = (long long)nk + nk / 4 + 1024;
        e = hipMalloc(&g.keys, sizeof(unsigned long long) * cap);
        if (e != hipSuccess) return e;
        e = hipMalloc(&g.keys_alt, sizeof(unsigned long long) * cap);
        if (e != hipSuccess) return e;
        g.keys_cap = cap;
    }
    long long tiles = 1ll << (2 * G.tbits);
    if (tiles > g.tiles_cap) {
        if (g.tile_start) hipFree(g.tile_start);
        if (g.tile_end) hipFree(g.tile_end);
        e = hipMalloc(&g.tile_start, sizeof(int) * tiles);
        if (e != hipSuccess) return e;
        e = hipMalloc(&g.tile_end, sizeof(int) * tiles);
        if (e != hipSuccess) return e;
        g.tiles_cap = tiles;
    }
    if (!g.rs_hist) {
        e = hipMalloc(&g.rs_hist, sizeof(unsigned) * 256 * kSortBlocks);
        if (e != hipSuccess) return e;
    }
    return hipSuccess;
}

hipError_t launch_integrate_gather(Grid& g, float* pixels, int W, int H, hipStream_t st,
                                   hipEvent_t ev_sort0, hipEvent_t ev_sort1) {
    ReconGeo G = make_geo(g, W, H);
    kbeg(g, st, "k_recon_emit");
    hipLaunchKernelGGL(k_recon_emit, dim3(kChunkBlocks), dim3(256), 0, st, g.dev, G, g.keys,
                       g.keys_cap);
    kend(g, st);
    if (ev_sort0) (void)hipEventRecord(ev_sort0, st);
    unsigned long long* sorted = radix_sort_keys(g, 32, 32 + 2 * G.tbits, st);
    if (ev_sort1) (void)hipEventRecord(ev_sort1, st);
    if (sorted != g.keys) {  // keep the sorted keys in g.keys for inspection
        unsigned long long* t = g.keys;
        g.keys = g.keys_alt;
        g.keys_alt = t;
    }
    const long long tiles = 1ll << (2 * G.tbits);
    kbeg(g, st, "k_tile_ranges");
    launch_fill_i32(g.tile_start, 0, tiles, st);
    launch_fill_i32(g.tile_end, 0, tiles, st);
    const long long nk = g.last_n_keys;
    int blocks = (int)std::min<long long>((nk + 255) / 256, 8192);
    if (blocks < 1) blocks = 1;
    hipLaunchKernelGGL(k_tile_ranges, dim3(blocks), dim3(256), 0, st, g.keys, &g.dev.cnt->n_keys,
                       g.tile_start, g.tile_end);
    kend(g, st);
    const int ntiles = G.tilesX * G.tilesY;
    kbeg(g, st, "k_gather");
    MDAS_DISPATCH_D(g.D, hipLaunchKernelGGL(k_gather<DD>, dim3((ntiles + 3) / 4), dim3(256), 0, st,
                                            g.dev, G, g.keys, g.tile_start, g.tile_end,
                                            (float4*)pixels));
    kend(g, st);
    return hipGetLastError();
}

hipError_t launch_integrate_scatter(Grid& g, float* pixels, int W, int H, hipStream_t st) {
    ReconGeo G = make_geo(g, W, H);
    launch_fill_f32(pixels, 0.0f, 4ll * W * H, st);
    int blocks = (int)std::min<long long>(((long long)g.cap_l + 255) / 256, 8192);
    MDAS_DISPATCH_D(g.D, hipLaunchKernelGGL(k_scatter<DD>, dim3(blocks), dim3(256), 0, st, g.dev,
                                            G, (float4*)pixels));
    return hipGetLastError();
}

hipError_t launch_bytes(const float* pixels, unsigned char* out, long long n, hipStream_t st) {
    int blocks = (int)std::min<long long>((n + 255) / 256, 8192);
    if (blocks < 1) blocks = 1;
    hipLaunchKernelGGL(k_bytes, dim3(blocks), dim3(256), 0, st, (const float4*)pixels,
                       (uchar4*)out, n);
    return hipGetLastError();
}

}  // namespace mdas
