// Reconstruction: box-filter of the leaves into pixels (kdgrid.cu:483-582),
// done as a deterministic gather instead of the reference's float atomics.
//
//  1. k_recon_prep: one coalesced pass over the leaves writes a 32-byte
//     record per leaf (xy box, mean value, volume) and counts the pixel tiles
//     each leaf's span touches.
//  2. k_scan_keys / k_recon_emit: every leaf emits one 64-bit key per tile,
//         key = morton2(tile_x, tile_y) << 32 | leaf index,
//     in leaf order (a workgroup-chunked exclusive scan gives the offsets).
//  3. radix_sort_keys: stable LSD radix sort (digits of up to 10 bits,
//     LDS-staged scatter) on the Morton bits only; since keys are emitted in leaf order
//     the result is the total 64-bit order.
//  4. k_tile_ranges: start / end of every tile's key segment.
//  5. k_gather: a group of 16 or 64 lanes owns a tile of 16 or 64 pixels,
//     walks the tile's leaf records in index order (staged in LDS) and every
//     lane accumulates its pixel.  Each contribution is computed with the
//     reference's own expressions, including its accumulated pixel bounds
//     (pMax += step), so a pixel is the sum of its contributions in leaf
//     index order.
#include <algorithm>
#include <cmath>
#include <type_traits>

#include "mdas_leaf.hpp"

#pragma clang fp contract(off)

namespace mdas {

static const int kSortBlocks = 1024;
static const int kSortMaxBits = 10;  // digit width limit (histograms of 1024 bins)
static const int kSortKeysPerThread = 16;  // scatter tile: 4096 keys staged in LDS (8192: slower)

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

// The span with the start clamped to the image (a span starting past the
// image is empty), from a record's xy box.
__device__ __forceinline__ void record_span(const float4& q0, const ReconGeo& G, int& x0, int& x1,
                                            int& y0, int& y1) {
    pixel_span(q0.x, q0.y, q0.z, q0.w, G, x0, x1, y0, y1);
    x0 = min(x0, G.W);
    y0 = min(y0, G.H);
}

__device__ __forceinline__ int tiles_of(int x0, int x1, int y0, int y1, const ReconGeo& G) {
    if (x0 >= x1 || y0 >= y1) return 0;
    return ((x1 - 1) / G.TW - x0 / G.TW + 1) * ((y1 - 1) / G.TH - y0 / G.TH + 1);
}

// Record of leaf i: q0 = (mn0, mx0, mn1, mx1), q1 = (mean value, volume).
// The mean is the slot-order sum times 1/n (kdgrid.cu:510-518,
// vec_math.h:492-496); the pixel span and the xy area are recomputed from q0
// by whoever needs them, with the same operations.
template <int D>
__global__ __launch_bounds__(256) void k_recon_prep(DevGrid a, ReconGeo G, float4* rrec) {
    __shared__ int red[4];
    const int L = a.cnt->n_leaves;
    int beg, end;
    chunk_range(L, gridDim.x, blockIdx.x, beg, end);
    int c = 0;
    for (int i = beg + threadIdx.x; i < end; i += 256) {
        int4 id4;
        float mn[D], mx[D];
        load_leaf<D>(a, i, leaf_count(a, i), id4, mn, mx);
        const int ids[4] = {id4.x, id4.y, id4.z, id4.w};
        // all four loads in flight at once; an empty slot reads sample 0
        float vx[4], vy[4], vz[4];
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            const float* v = a.values + (size_t)(ids[q] >= 0 ? ids[q] : 0) * 3;
            vx[q] = v[0];
            vy[q] = v[1];
            vz[q] = v[2];
        }
        float sx = 0.0f, sy = 0.0f, sz = 0.0f;
        int n = 0;
#pragma unroll
        for (int q = 0; q < 4; ++q)
            if (ids[q] >= 0) {
                sx += vx[q];
                sy += vy[q];
                sz += vz[q];
                ++n;
            }
        const float inv = 1.0f / (float)n;
        float vol = 1.0f;
#pragma unroll
        for (int d = 0; d < D; ++d) vol *= (mx[d] - mn[d]);
        const float4 q0 = make_float4(mn[0], mx[0], mn[1], mx[1]);
        int x0, x1, y0, y1;
        record_span(q0, G, x0, x1, y0, y1);
        c += tiles_of(x0, x1, y0, y1, G);
        float4* r = rrec + (size_t)i * 2;
        r[0] = q0;
        r[1] = make_float4(sx * inv, sy * inv, sz * inv, vol);
    }
    c = wave_sum(c);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = c;
    __syncthreads();
    if (threadIdx.x == 0) a.part[blockIdx.x] = make_int4(red[0] + red[1] + red[2] + red[3], 0, 0, 0);
}

// Exclusive scan of part[].x into part_off[].x; total -> cnt->n_keys.  A
// total beyond the key buffer (cap_keys) is recorded in recon_overflow and
// leaves n_keys = 0, so that no later kernel reads or writes past the buffer.
__global__ __launch_bounds__(1024) void k_scan_keys(DevGrid a, int G, long long cap_keys) {
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
    if (threadIdx.x == 0) {
        const bool over = (long long)t > cap_keys;
        a.cnt->n_keys = over ? 0 : t;
        a.cnt->recon_overflow = over ? t : 0;
    }
}

__global__ __launch_bounds__(256) void k_recon_emit(DevGrid a, ReconGeo G, const float4* rrec,
                                                    unsigned long long* keys) {
    __shared__ int lds[4];
    const int L = a.cnt->n_leaves;
    if (a.cnt->recon_overflow != 0) return;  // the keys do not fit the buffer
    int beg, end;
    chunk_range(L, gridDim.x, blockIdx.x, beg, end);
    int base = a.part_off[blockIdx.x].x;
    for (int t0 = beg; t0 < end; t0 += 256) {
        const int i = t0 + threadIdx.x;
        int x0 = 0, x1 = 0, y0 = 0, y1 = 0, c = 0;
        if (i < end) {
            record_span(rrec[(size_t)i * 2], G, x0, x1, y0, y1);
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

// ---------------------------------------------------------- device scan --
// Exclusive scan of n <= 1024*1024 unsigned ints in place: block sums, a
// scan of the block sums, then the block-local scans with their offsets.
__device__ __forceinline__ unsigned block_scan_u32(unsigned v, unsigned* ls, unsigned& total) {
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    unsigned x = v;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        unsigned y = (unsigned)__shfl_up((int)x, o, 64);
        if (lane >= o) x += y;
    }
    if (lane == 63) ls[w] = x;
    __syncthreads();
    unsigned pre = 0;
    total = 0;
    for (int k = 0; k < 16; ++k) {
        if (k < w) pre += ls[k];
        total += ls[k];
    }
    __syncthreads();
    return pre + x - v;
}

__global__ __launch_bounds__(1024) void k_scan_reduce(const unsigned* data, int n, unsigned* sums) {
    __shared__ unsigned ls[16];
    const int i = blockIdx.x * 1024 + threadIdx.x;
    unsigned total;
    (void)block_scan_u32(i < n ? data[i] : 0u, ls, total);
    if (threadIdx.x == 0) sums[blockIdx.x] = total;
}

__global__ __launch_bounds__(1024) void k_scan_top(unsigned* sums, int nb) {
    __shared__ unsigned ls[16];
    unsigned total;
    const unsigned v = threadIdx.x < (unsigned)nb ? sums[threadIdx.x] : 0u;
    const unsigned ex = block_scan_u32(v, ls, total);
    if (threadIdx.x < (unsigned)nb) sums[threadIdx.x] = ex;
}

__global__ __launch_bounds__(1024) void k_scan_down(unsigned* data, int n, const unsigned* sums) {
    __shared__ unsigned ls[16];
    const int i = blockIdx.x * 1024 + threadIdx.x;
    unsigned total;
    const unsigned ex = block_scan_u32(i < n ? data[i] : 0u, ls, total);
    if (i < n) data[i] = ex + sums[blockIdx.x];
}

// ------------------------------------------------------------ radix sort --
// Stable LSD radix sort of 64-bit keys over bits [shift, shift+BITS) per
// pass, BITS chosen per sort so the Morton bits take as few passes as
// possible (at most 10 bits a pass).  Reduce-then-scan: per-workgroup digit
// histograms, one exclusive scan over (digit, workgroup), then a scatter in
// which every workgroup ranks a tile of 4096 keys by digit with wave ballots
// (16 sub-tiles of 256), stages the tile in LDS in digit order and writes each
// digit run contiguously -- runs of 8 keys on average at 512 digits, where a
// 256-key tile would leave most keys alone in their line.
template <int BITS>
__global__ __launch_bounds__(256) void k_rs_hist(const unsigned long long* keys, const int* n_ptr,
                                                 int shift, unsigned* hist) {
    constexpr int NB = 1 << BITS;
    __shared__ unsigned h[NB];
    const int n = *n_ptr;
    for (int d = threadIdx.x; d < NB; d += 256) h[d] = 0;
    __syncthreads();
    int beg, end;
    chunk_range(n, gridDim.x, blockIdx.x, beg, end);
    for (int i = beg + threadIdx.x; i < end; i += 256)
        atomicAdd(&h[(unsigned)(keys[i] >> shift) & (unsigned)(NB - 1)], 1u);
    __syncthreads();
    for (int d = threadIdx.x; d < NB; d += 256) hist[d * gridDim.x + blockIdx.x] = h[d];
}

template <int BITS>
__global__ __launch_bounds__(256) void k_rs_scatter(const unsigned long long* in,
                                                    unsigned long long* out, const int* n_ptr,
                                                    int shift, const unsigned* hist) {
    constexpr int NB = 1 << BITS;
    constexpr int R = NB > 256 ? NB / 256 : 1;  // digits per thread in the tile scan
    constexpr int KPT = kSortKeysPerThread;
    constexpr int TILE = 256 * KPT;
    __shared__ unsigned base[NB];
    __shared__ int wcnt[4][NB];
    __shared__ int run[NB];   // keys of each digit in the tile's earlier sub-tiles
    __shared__ int toff[NB];
    __shared__ int scr[4];
    __shared__ unsigned long long stage[TILE];
    const int n = *n_ptr;
    const int w = threadIdx.x >> 6;
    for (int d = threadIdx.x; d < NB; d += 256) base[d] = hist[d * gridDim.x + blockIdx.x];
    int beg, end;
    chunk_range(n, gridDim.x, blockIdx.x, beg, end);
    for (int t0 = beg; t0 < end; t0 += TILE) {
        // sub-tile s holds keys t0 + 256 s + thread; their loads go out together
        unsigned long long k[KPT];
#pragma unroll
        for (int s = 0; s < KPT; ++s) {
            const int i = t0 + s * 256 + threadIdx.x;
            k[s] = i < end ? in[i] : 0ull;
        }
        for (int d = threadIdx.x; d < NB; d += 256) run[d] = 0;
        // rank of each key among the tile's keys of its digit, in input
        // order: earlier sub-tiles, then lower waves, then lower lanes
        int rk[KPT];
#pragma unroll
        for (int s = 0; s < KPT; ++s) {
            const bool valid = t0 + s * 256 + (int)threadIdx.x < end;
            const unsigned d = (unsigned)(k[s] >> shift) & (unsigned)(NB - 1);
            for (int q = 0; q < 4; ++q)
                for (int dd = threadIdx.x; dd < NB; dd += 256) wcnt[q][dd] = 0;
            __syncthreads();
            unsigned long long m = __ballot(valid);
#pragma unroll
            for (int b = 0; b < BITS; ++b) {
                const bool bit = (d >> b) & 1u;
                const unsigned long long bal = __ballot(bit);
                m &= bit ? bal : ~bal;
            }
            const int rw = lanes_below(m);
            if (valid && rw == 0) wcnt[w][d] = __popcll(m);
            __syncthreads();
            int pre = run[d];
            for (int q = 0; q < w; ++q) pre += wcnt[q][d];
            rk[s] = pre + rw;
            __syncthreads();
            for (int dd = threadIdx.x; dd < NB; dd += 256)
                run[dd] += wcnt[0][dd] + wcnt[1][dd] + wcnt[2][dd] + wcnt[3][dd];
        }
        __syncthreads();
        {
            int loc[R], sum = 0;
#pragma unroll
            for (int r = 0; r < R; ++r) {
                const int dd = (int)threadIdx.x * R + r;
                loc[r] = dd < NB ? run[dd] : 0;
                sum += loc[r];
            }
            int all;
            int ex = block_exscan(sum, scr, all);
#pragma unroll
            for (int r = 0; r < R; ++r) {
                const int dd = (int)threadIdx.x * R + r;
                if (dd < NB) {
                    toff[dd] = ex;
                    ex += loc[r];
                }
            }
        }
        __syncthreads();
#pragma unroll
        for (int s = 0; s < KPT; ++s) {
            if (t0 + s * 256 + (int)threadIdx.x < end) {
                const unsigned d = (unsigned)(k[s] >> shift) & (unsigned)(NB - 1);
                stage[toff[d] + rk[s]] = k[s];
            }
        }
        __syncthreads();
        // the tile in digit order: each digit's run goes to its own range
        const int nvalid = min(TILE, end - t0);
        for (int p = threadIdx.x; p < nvalid; p += 256) {
            const unsigned long long kk = stage[p];
            const unsigned dd = (unsigned)(kk >> shift) & (unsigned)(NB - 1);
            out[base[dd] + (p - toff[dd])] = kk;
        }
        __syncthreads();
        for (int dd = threadIdx.x; dd < NB; dd += 256) base[dd] += run[dd];
        __syncthreads();
    }
}

static void scan_u32(unsigned* data, int n, unsigned* sums, hipStream_t st) {
    const int nb = (n + 1023) / 1024;
    hipLaunchKernelGGL(k_scan_reduce, dim3(nb), dim3(1024), 0, st, data, n, sums);
    hipLaunchKernelGGL(k_scan_top, dim3(1), dim3(1024), 0, st, sums, nb);
    hipLaunchKernelGGL(k_scan_down, dim3(nb), dim3(1024), 0, st, data, n, sums);
}

template <int BITS>
static void sort_pass(Grid& g, unsigned* rs_hist, const unsigned long long* a, unsigned long long* b,
                      int sh, hipStream_t st) {
    const int* n_ptr = &g.dev.cnt->n_keys;
    const int nh = (1 << BITS) * kSortBlocks;
    kbeg(g, st, "k_rs_hist");
    hipLaunchKernelGGL(k_rs_hist<BITS>, dim3(kSortBlocks), dim3(256), 0, st, a, n_ptr, sh, rs_hist);
    kend(g, st);
    kbeg(g, st, "k_rs_scan");
    scan_u32(rs_hist, nh, rs_hist + nh, st);
    kend(g, st);
    kbeg(g, st, "k_rs_scatter");
    hipLaunchKernelGGL(k_rs_scatter<BITS>, dim3(kSortBlocks), dim3(256), 0, st, a, b, n_ptr, sh,
                       rs_hist);
    kend(g, st);
}

// Sorts keys[0, n) by bits [lo, hi) in ceil((hi - lo) / 10) passes of equal
// digit width; returns the buffer holding the result.
static unsigned long long* radix_sort_keys(Grid& g, ReconBufs& rb, int lo, int hi, hipStream_t st) {
    unsigned long long* a = rb.keys;
    unsigned long long* b = rb.keys_alt;
    g.recon_sort_passes = 0;
    const int total = hi - lo;
    if (total <= 0) return a;
    const int passes = (total + kSortMaxBits - 1) / kSortMaxBits;
    const int bits = max(4, (total + passes - 1) / passes);
    for (int sh = lo; sh < hi; sh += bits) {
        switch (bits) {
        case 4: sort_pass<4>(g, rb.rs_hist, a, b, sh, st); break;
        case 5: sort_pass<5>(g, rb.rs_hist, a, b, sh, st); break;
        case 6: sort_pass<6>(g, rb.rs_hist, a, b, sh, st); break;
        case 7: sort_pass<7>(g, rb.rs_hist, a, b, sh, st); break;
        case 8: sort_pass<8>(g, rb.rs_hist, a, b, sh, st); break;
        case 9: sort_pass<9>(g, rb.rs_hist, a, b, sh, st); break;
        default: sort_pass<10>(g, rb.rs_hist, a, b, sh, st); break;
        }
        g.recon_sort_passes++;
        g.recon_sort_bits = bits;
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
// GS = pixels per tile (16 or 64); a wave holds 64/GS tiles.
template <int GS>
__global__ __launch_bounds__(256) void k_gather(ReconGeo G, const float4* __restrict__ rrec,
                                                const unsigned long long* __restrict__ keys,
                                                const int* __restrict__ tstart,
                                                const int* __restrict__ tend, float4* pixels) {
    __shared__ float4 r0[4][64], r1[4][64], r2[4][64];
    __shared__ float2 r3[4][64];
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const int grp = lane / GS, gl = lane % GS;      // tile within the wave, lane within the tile
    const int tile = (blockIdx.x * 4 + w) * (64 / GS) + grp;
    const int ntiles = G.tilesX * G.tilesY;
    const bool live = tile < ntiles;
    const int tx = live ? tile % G.tilesX : 0, ty = live ? tile / G.tilesX : 0;
    const int px = tx * G.TW + gl % G.TW;
    const int py = ty * G.TH + gl / G.TW;
    int s0 = 0, s1 = 0;
    if (live) {
        const unsigned code = morton2(tx, ty);
        s0 = tstart[code];
        s1 = tend[code];
    }
    // rounds of GS records until the longest tile of the wave is done; the
    // records of round rd + 1 and the keys of round rd + 2 are loaded while
    // round rd is accumulated
    int rounds = (s1 - s0 + GS - 1) / GS;
#pragma unroll
    for (int o = GS; o < 64; o <<= 1) rounds = max(rounds, __shfl_xor(rounds, o, 64));
    float accx = 0.0f, accy = 0.0f, accz = 0.0f;
    const int gbase = grp * GS;
    auto nrec_of = [&](int rd) { return max(0, min(GS, s1 - (s0 + rd * GS))); };
    float4 c0 = make_float4(0.0f, 0.0f, 0.0f, 0.0f), c1 = c0;  // records of the round to stage
    if (gl < nrec_of(0)) {
        const int i = (int)(unsigned)(keys[s0 + gl] & 0xffffffffull);
        c0 = rrec[(size_t)i * 2];
        c1 = rrec[(size_t)i * 2 + 1];
    }
    unsigned long long knext = gl < nrec_of(1) ? keys[s0 + GS + gl] : 0ull;
    for (int rd = 0; rd < rounds; ++rd) {
        const int nrec = nrec_of(rd);
        unsigned long long own = 0ull;  // 64-pixel tiles: the pixels this lane's record covers
        if (gl < nrec) {
            // per record, once: span, xy area, the span's first pixel bounds
            // (kdgrid.cu:532-536, the same operations every lane would do)
            // and the mask of the tile's pixels inside the span
            int x0, x1, y0, y1;
            record_span(c0, G, x0, x1, y0, y1);
            unsigned mask = 0u;  // tiles of up to 32 pixels; larger tiles test the span
            if (GS <= 32) {
                for (int k = 0; k < GS; ++k) {
                    const int qx = tx * G.TW + k % G.TW, qy = ty * G.TH + k / G.TW;
                    if (qx >= x0 && qx < x1 && qy >= y0 && qy < y1) mask |= 1u << k;
                }
            } else {
                // the span clipped to the tile is a rectangle of bits: columns
                // [cx0, cx1) of every row in [cy0, cy1), row r at bits TW r..
                const int cx0 = min(max(x0 - tx * G.TW, 0), G.TW), cx1 = min(max(x1 - tx * G.TW, 0), G.TW);
                const int cy0 = min(max(y0 - ty * G.TH, 0), G.TH), cy1 = min(max(y1 - ty * G.TH, 0), G.TH);
                auto low = [](int n) { return n >= 64 ? ~0ull : (1ull << n) - 1ull; };
                const unsigned long long rep = G.TW >= 64 ? 1ull : ~0ull / low(G.TW);
                const unsigned long long cols = low(cx1) & ~low(cx0);
                const unsigned long long rows = rep & low(G.TW * cy1) & ~low(G.TW * cy0);
                own = cols * rows;  // no carries: the column field is TW bits wide
            }
            r0[w][gbase + gl] = c0;
            r1[w][gbase + gl] = c1;
            r2[w][gbase + gl] = make_float4((c0.y - c0.x) * (c0.w - c0.z),
                                            __uint_as_float((unsigned)x0 | ((unsigned)x1 << 16)),
                                            __uint_as_float((unsigned)y0 | ((unsigned)y1 << 16)),
                                            __uint_as_float(mask));
            r3[w][gbase + gl] = make_float2((float)x0 * G.sX / G.fW, (float)y0 * G.sY / G.fH);
        }
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
        if (gl < nrec_of(rd + 1)) {
            const int i = (int)(unsigned)(knext & 0xffffffffull);
            c0 = rrec[(size_t)i * 2];
            c1 = rrec[(size_t)i * 2 + 1];
        }
        knext = gl < nrec_of(rd + 2) ? keys[s0 + (rd + 2) * GS + gl] : 0ull;
        // the round's records that cover this lane's pixel, then only those,
        // in index order (the others add nothing): the loop runs as often as
        // the most-covered pixel of the wave needs, not once per record
        using CovT = typename std::conditional<(GS > 32), unsigned long long, unsigned>::type;
        CovT cov = 0;
        if constexpr (GS > 32) {
            // one tile per wave, record j in lane j: the wave's coverage
            // matrix transposed by ballots, column p delivered to lane p
#pragma unroll
            for (int p = 0; p < 64; ++p) {
                const unsigned long long b = __ballot(((own >> p) & 1ull) != 0ull);
                cov = lane == p ? (CovT)b : cov;
            }
        } else {
            for (int j = 0; j < nrec; ++j) {
                const float4 q2 = r2[w][gbase + j];
                cov |= (CovT)((__float_as_uint(q2.w) >> gl) & 1u) << j;
            }
        }
        while (cov != 0) {
            const int j = GS > 32 ? __builtin_ctzll((unsigned long long)cov) : __builtin_ctz((unsigned)cov);
            cov &= cov - 1;
            {
                const float4 q2 = r2[w][gbase + j];
                const unsigned spx = __float_as_uint(q2.y), spy = __float_as_uint(q2.z);
                const int x0 = (int)(spx & 0xffffu), y0 = (int)(spy & 0xffffu);
                const float4 q0 = r0[w][gbase + j];
                const float4 q1 = r1[w][gbase + j];
                const float2 q3 = r3[w][gbase + j];
                // accumulated pixel bounds exactly as kdgrid.cu:536-576 walks them
                float pMinX = q3.x;
                float pMaxX = pMinX + G.stepX;
                for (int q = x0; q < px; ++q) {
                    pMinX = pMaxX;
                    pMaxX += G.stepX;
                }
                float pMinY = q3.y;
                float pMaxY = pMinY + G.stepY;
                for (int q = y0; q < py; ++q) {
                    pMinY = pMaxY;
                    pMaxY += G.stepY;
                }
                const float clox = fmaxf(pMinX, q0.x);
                const float chix = fminf(pMaxX, q0.y);
                const float cloy = fmaxf(pMinY, q0.z);
                const float chiy = fminf(pMaxY, q0.w);
                if (clox < chix && cloy < chiy) {
                    const float v = (chix - clox) * (chiy - cloy) * q1.w / q2.x;
                    accx += (q1.x * v) * G.invPixArea;
                    accy += (q1.y * v) * G.invPixArea;
                    accz += (q1.z * v) * G.invPixArea;
                }
            }
        }
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    }
    if (live && px < G.W && py < G.H) pixels[(size_t)py * G.W + px] = make_float4(accx, accy, accz, 0.0f);
}

// -------------------------------------------------------- scatter (atomics) --
// kdgrid.cu:483-582 as written: one thread per leaf, float atomics per pixel.
template <int D>
__global__ __launch_bounds__(256) void k_scatter(DevGrid a, ReconGeo G, float4* pixels) {
    const int L = a.cnt->n_leaves;
    for (int i = blockIdx.x * 256 + threadIdx.x; i < L; i += gridDim.x * 256) {
        int4 id4;
        float mn[D], mx[D];
        load_leaf<D>(a, i, leaf_count(a, i), id4, mn, mx);
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

static hipError_t grow_keys(ReconBufs& rb, long long cap) {
    if (rb.keys) (void)hipFree(rb.keys);
    if (rb.keys_alt) (void)hipFree(rb.keys_alt);
    rb.keys = rb.keys_alt = nullptr;
    rb.keys_cap = 0;
    hipError_t e = hipMalloc(&rb.keys, sizeof(unsigned long long) * (size_t)cap);
    if (e == hipSuccess) e = hipMalloc(&rb.keys_alt, sizeof(unsigned long long) * (size_t)cap);
    if (e != hipSuccess) return e;
    rb.keys_cap = cap;
    return hipSuccess;
}

// Upper bound on the keys one leaf emits: the tiles its pixel span can touch.
// A leaf's box is at most g.ext_x by g.ext_y in x and y (splits only shrink
// boxes), and a span of extent ext over n pixels of a sample space of size s
// holds at most floor(ext n / s) + 2 pixels (floor(mx n/s) - floor(mn n/s) + 1,
// with a margin for the rounding of the float products); p pixels starting
// anywhere in a tile of T touch at most (p + T - 2) / T + 1 tiles.
static long long tiles_per_leaf(const Grid& g, const ReconGeo& G) {
    auto axis = [](double ext, int n, double s, int T, int tiles) {
        long long p = (long long)std::floor(ext * (double)n / s * (1.0 + 1e-6) + 1e-3) + 2;
        p = std::min<long long>(p, n);
        return std::min<long long>((p + T - 2) / T + 1, tiles);
    };
    return axis(g.ext_x, G.W, G.sX, G.TW, G.tilesX) * axis(g.ext_y, G.H, G.sY, G.TH, G.tilesY);
}

static long long key_bound(const Grid& g, const ReconGeo& G) {
    return tiles_per_leaf(g, G) * std::max<long long>(g.host.n_leaves, 1);
}

// Key buffers of at most this many bytes are sized from the bound up front
// (no host round trip in the reconstruction); larger bounds (tiny initial
// grids, whose leaves span many tiles) size them from the counted keys.
static const long long kKeyBoundBytes = 8ll << 30;

static long long align256(long long b) { return (b + 255) / 256 * 256; }

long long recon_key_bound(const Grid& g, int W, int H) { return key_bound(g, make_geo(g, W, H)); }
long long recon_tiles_cap(const Grid& g, int W, int H) { return 1ll << (2 * make_geo(g, W, H).tbits); }

long long recon_bytes_needed(const Grid& g, int W, int H) {
    const ReconGeo G = make_geo(g, W, H);
    const long long kb = key_bound(g, G);
    if (kb > 0x7fffffffLL) return -1;
    const long long L = std::max<long long>(g.host.n_leaves, 1);
    return align256(sizeof(unsigned) * kSortHistWords) + 2 * align256(sizeof(int) << (2 * G.tbits)) +
           align256(sizeof(float4) * 2 * L) + 2 * align256(sizeof(unsigned long long) * kb);
}

// Sizes the grid's reconstruction buffers for the current leaf count (g.host
// is current after every call that changes the tree) and image; allocates
// only when they grow.  Sets g.recon_bounded when the key buffers hold the
// bound, i.e. the reconstruction needs no host round trip.
hipError_t recon_reserve(Grid& g, int W, int H) {
    ReconGeo G = make_geo(g, W, H);
    ReconBufs& rb = g.rb;
    const long long L = g.host.n_leaves;
    if (L > rb.rrec_cap) {
        if (rb.rrec) (void)hipFree(rb.rrec);
        rb.rrec = nullptr;
        rb.rrec_cap = 0;
        const long long cap = std::min<long long>((long long)g.cap_l, L + L / 4 + 1024);
        hipError_t e = hipMalloc(&rb.rrec, sizeof(float4) * 2 * (size_t)cap);
        if (e != hipSuccess) return e;
        rb.rrec_cap = cap;
    }
    const long long kb = key_bound(g, G);
    g.recon_bounded = kb <= 0x7fffffffLL && kb * 16 <= kKeyBoundBytes;
    if (g.recon_bounded && kb > rb.keys_cap) {
        hipError_t e = grow_keys(rb, kb);
        if (e != hipSuccess) return e;
    }
    const long long tiles = 1ll << (2 * G.tbits);
    if (tiles > rb.tiles_cap) {
        if (rb.tile_start) (void)hipFree(rb.tile_start);
        if (rb.tile_end) (void)hipFree(rb.tile_end);
        rb.tile_start = rb.tile_end = nullptr;
        rb.tiles_cap = 0;
        hipError_t e = hipMalloc(&rb.tile_start, sizeof(int) * tiles);
        if (e == hipSuccess) e = hipMalloc(&rb.tile_end, sizeof(int) * tiles);
        if (e != hipSuccess) return e;
        rb.tiles_cap = tiles;
    }
    if (!rb.rs_hist) {
        hipError_t e = hipMalloc(&rb.rs_hist, sizeof(unsigned) * kSortHistWords);
        if (e != hipSuccess) return e;
    }
    return hipSuccess;
}

// Per-leaf records, tile counts and their scan.  Stream-ordered when the key
// buffers hold the bound; otherwise the key count is read back (a
// synchronisation) and the buffers grown to it.
hipError_t recon_prepare(Grid& g, int W, int H, hipStream_t st) {
    hipError_t e = recon_reserve(g, W, H);
    if (e != hipSuccess) return e;
    ReconGeo G = make_geo(g, W, H);
    g.recon_pixels = (long long)W * H;
    kbeg(g, st, "k_recon_prep");
    MDAS_DISPATCH_D(g.D, hipLaunchKernelGGL(k_recon_prep<DD>, dim3(g.chunk_blocks), dim3(256), 0, st,
                                            g.dev, G, g.rb.rrec));
    kend(g, st);
    kbeg(g, st, "k_scan_keys");
    hipLaunchKernelGGL(k_scan_keys, dim3(1), dim3(1024), 0, st, g.dev, g.chunk_blocks,
                       g.recon_bounded ? g.rb.keys_cap : (long long)0x7fffffff);
    kend(g, st);
    if (!g.recon_bounded) {
        int nk = 0;
        e = hipMemcpyAsync(&nk, &g.dev.cnt->n_keys, sizeof(int), hipMemcpyDeviceToHost, st);
        if (e == hipSuccess) e = hipStreamSynchronize(st);
        if (e != hipSuccess) return e;
        if (nk > g.rb.keys_cap) {
            e = grow_keys(g.rb, (long long)nk + nk / 4 + 1024);
            if (e != hipSuccess) return e;
        }
    }
    return hipGetLastError();
}

// Keys in leaf order, their sort, the tiles' key ranges and the gather, on
// the buffers of rb (records already written by k_recon_prep).
static hipError_t recon_finish(Grid& g, ReconBufs& rb, float* pixels, int W, int H, hipStream_t st,
                               hipEvent_t ev_sort0, hipEvent_t ev_sort1) {
    ReconGeo G = make_geo(g, W, H);
    kbeg(g, st, "k_recon_emit");
    hipLaunchKernelGGL(k_recon_emit, dim3(g.chunk_blocks), dim3(256), 0, st, g.dev, G, rb.rrec, rb.keys);
    kend(g, st);
    if (ev_sort0) (void)hipEventRecord(ev_sort0, st);
    // the sorted keys end in keys or keys_alt by the parity of the pass
    // count; the buffers keep their roles, so a captured graph stays valid
    unsigned long long* sorted = radix_sort_keys(g, rb, 32, 32 + 2 * G.tbits, st);
    g.keys_sorted = sorted;
    if (ev_sort1) (void)hipEventRecord(ev_sort1, st);
    const long long tiles = 1ll << (2 * G.tbits);
    kbeg(g, st, "k_tile_ranges");
    launch_fill_i32(rb.tile_start, 0, tiles, st);
    launch_fill_i32(rb.tile_end, 0, tiles, st);
    int blocks = (int)std::min<long long>((rb.keys_cap + 255) / 256, 8192);  // grid-stride over
    if (blocks < 1) blocks = 1;                                              // the device count
    hipLaunchKernelGGL(k_tile_ranges, dim3(blocks), dim3(256), 0, st, sorted, &g.dev.cnt->n_keys,
                       rb.tile_start, rb.tile_end);
    kend(g, st);
    const int ntiles = G.tilesX * G.tilesY;
    const int gs = G.TW * G.TH;
    const int tiles_per_block = 4 * (64 / gs);
    kbeg(g, st, "k_gather");
    if (gs == 64)
        hipLaunchKernelGGL(k_gather<64>, dim3((ntiles + tiles_per_block - 1) / tiles_per_block),
                           dim3(256), 0, st, G, rb.rrec, sorted, rb.tile_start, rb.tile_end,
                           (float4*)pixels);
    else
        hipLaunchKernelGGL(k_gather<16>, dim3((ntiles + tiles_per_block - 1) / tiles_per_block),
                           dim3(256), 0, st, G, rb.rrec, sorted, rb.tile_start, rb.tile_end,
                           (float4*)pixels);
    kend(g, st);
    return hipGetLastError();
}

hipError_t launch_integrate_gather(Grid& g, float* pixels, int W, int H, hipStream_t st,
                                   hipEvent_t ev_sort0, hipEvent_t ev_sort1) {
    return recon_finish(g, g.rb, pixels, W, H, st, ev_sort0, ev_sort1);
}

hipError_t recon_on_buffers(Grid& g, ReconBufs& rb, float* pixels, int W, int H, hipStream_t st) {
    ReconGeo G = make_geo(g, W, H);
    g.recon_pixels = (long long)W * H;
    MDAS_DISPATCH_D(g.D, hipLaunchKernelGGL(k_recon_prep<DD>, dim3(g.chunk_blocks), dim3(256), 0, st,
                                            g.dev, G, rb.rrec));
    hipLaunchKernelGGL(k_scan_keys, dim3(1), dim3(1024), 0, st, g.dev, g.chunk_blocks, rb.keys_cap);
    return recon_finish(g, rb, pixels, W, H, st, nullptr, nullptr);
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
