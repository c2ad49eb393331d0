// CPU oracle for the MDAS sampler -- TEST INFRASTRUCTURE, NOT PRODUCT CODE.
// See mdas_oracle.h for what it restates and which semantics it fixes.
// Build: oracle/Makefile (g++ -O3 -fopenmp -ffp-contract=off -fno-fast-math).
#include "mdas_oracle.h"

#include <omp.h>
#include <stdint.h>
#include <string.h>

#include <algorithm>
#include <cfloat>
#include <cmath>
#include <limits>
#include <vector>

#include "scenes.hpp"  // analytic integrands (input generator shared with the GPU path)

namespace {

// ------------------------------------------------------------------- RNG --
// cuda/random.h:31-46: TEA with the key words 0xa341316c, 0xc8013ea4,
// 0xad90777d, 0x7e95761e and N rounds; random.h:49-55 LCG (Numerical
// Recipes constants); random.h:64-67 rnd = (state & 0xFFFFFF) / 2^24.
const uint32_t kTeaKey[4] = {0xa341316cu, 0xc8013ea4u, 0xad90777du, 0x7e95761eu};

inline uint32_t tea_n(uint32_t v0, uint32_t v1, int rounds, const uint32_t* k, uint32_t* v1o) {
    uint32_t s0 = 0;
    for (int n = 0; n < rounds; ++n) {
        s0 += 0x9e3779b9u;
        v0 += ((v1 << 4) + k[0]) ^ (v1 + s0) ^ ((v1 >> 5) + k[1]);
        v1 += ((v0 << 4) + k[2]) ^ (v0 + s0) ^ ((v0 >> 5) + k[3]);
    }
    if (v1o) *v1o = v1;
    return v0;
}
inline uint32_t tea4(uint32_t a, uint32_t b) { return tea_n(a, b, 4, kTeaKey, nullptr); }
inline uint32_t lcg(uint32_t& s) {
    s = 1664525u * s + 1013904223u;
    return s & 0x00FFFFFFu;
}
inline float rnd(uint32_t& s) { return (float)lcg(s) / (float)0x01000000; }

// ------------------------------------------------------------------- pow --
// powf(x, a) = exp(a * ln x) evaluated in double and rounded once to float.
// ln: x = m 2^e with m in [sqrt(1/2), sqrt(2)), ln m = 2 atanh(t),
//     t = (m-1)/(m+1), atanh series to t^19 (Abramowitz & Stegun 4.1.27).
// exp: y = k ln2 + r, |r| <= ln2/2, Taylor series to r^13/13! (A&S 4.2.1).
// Relative error of the double result is ~1e-16, so the float result is the
// correctly rounded powf except where the exact value lies within ~1e-16 of a
// rounding boundary.  The GPU kernels evaluate the identical operation
// sequence, so both sides agree bit for bit.
const double kLn2 = 0.6931471805599453094;
const double kInvLn2 = 1.4426950408889634074;

inline double ln_pos(double x) {
    uint64_t bits;
    memcpy(&bits, &x, 8);
    int e = (int)((bits >> 52) & 0x7ff) - 1023;
    bits = (bits & 0x000fffffffffffffull) | 0x3ff0000000000000ull;
    double m;
    memcpy(&m, &bits, 8);
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
    return (double)e * kLn2 + 2.0 * t * s;
}

inline double exp_dbl(double y) {
    if (y < -110.0) return 0.0;
    if (y > 100.0) return HUGE_VAL;
    double kd = std::floor(y * kInvLn2 + 0.5);
    int k = (int)kd;
    double r = y - kd * kLn2;
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
    uint64_t sb = (uint64_t)(k + 1023) << 52;
    double sc;
    memcpy(&sc, &sb, 8);
    return p * sc;
}

inline float powf_mdas(float x, float a) {
    if (x != x || a != a) return std::numeric_limits<float>::quiet_NaN();
    if (a == 0.0f || x == 1.0f) return 1.0f;
    if (x == 0.0f) return a > 0.0f ? 0.0f : HUGE_VALF;
    if (x < 0.0f) return std::numeric_limits<float>::quiet_NaN();
    if (x == HUGE_VALF) return a > 0.0f ? HUGE_VALF : 0.0f;
    double y = (double)a * ln_pos((double)x);
    return (float)exp_dbl(y);
}

// ------------------------------------------------------------ grid state --
struct Grid {
    int D;
    long long cap_s, cap_l;
    int cand, b, e;
    float alpha, sX, sY;
    long long N = 0, L = 0, new_s = 0, new_l = 0;
    float gmax = 0.0f;
    std::vector<float> coords, values;  // [cap_s*D], [cap_s*3]
    std::vector<int> idx;               // [cap_l*4]
    std::vector<float> box;             // [cap_l*2D]: mn[D], mx[D]
    std::vector<float> err;             // [cap_l]
    std::vector<uint32_t> seed;         // [cap_l]
    std::vector<uint8_t> flags;
    std::vector<int> rank_s, rank_l;
};

int g_threads = 0;
inline int nthreads() { return g_threads > 0 ? g_threads : omp_get_max_threads(); }

// kdgrid.cu:141-164: leaf index -> cell offset and extent.  Bit k < b of
// coordinate d is index bit D*k + (D-1-d); for d < 2 bit k in [b, b+e) is
// index bit D*b + 2(k-b) + 1 - d.
template <int D>
inline void morton_cell(uint32_t i, int b, int e, float sX, float sY, float* off, float* ext,
                        uint32_t* xq_out = nullptr) {
    for (int d = 0; d < D; ++d) {
        uint32_t xq = 0;
        uint32_t inv = 1u << b;
        for (int k = 0; k < b; ++k) {
            int bit = D * k + (D - d - 1);
            xq |= ((i >> bit) & 1u) << k;
        }
        if (d < 2) {
            for (int k = b; k < b + e; ++k) {
                int bit = D * b + 2 * (k - b) + 1 - d;
                xq |= ((i >> bit) & 1u) << k;
            }
            inv <<= e;
        }
        ext[d] = 1.0f / (float)inv;
        off[d] = (float)xq * ext[d];
        if (xq_out) xq_out[d] = xq;
    }
    off[0] *= sX;
    off[1] *= sY;
    ext[0] *= sX;
    ext[1] *= sY;
}

// aabb.h:20-25
template <int D>
inline float volume(const float* mn, const float* mx) {
    float v = 1.0f;
    for (int d = 0; d < D; ++d) v *= (mx[d] - mn[d]);
    return v;
}

// aabb.h:34-45: scan from the last axis down, strict '<' (ties -> highest index)
inline int largest_axis(const float* diag, int D) {
    int axis = -1;
    float best = -FLT_MAX;
    for (int d = D - 1; d >= 0; --d)
        if (best < diag[d]) {
            best = diag[d];
            axis = d;
        }
    return axis;
}

// kdgrid.cu:237-266 (with the float3 /= semantics of sutil/vec_math.h:492-496)
inline float leaf_error(const float* v, int n, float vol) {
    float sx = 0.0f, sy = 0.0f, sz = 0.0f;
    for (int s = 0; s < n; ++s) {
        sx += v[3 * s];
        sy += v[3 * s + 1];
        sz += v[3 * s + 2];
    }
    float inv = 1.0f / (float)n;
    float ax = sx * inv, ay = sy * inv, az = sz * inv;
    float dx = 0.0f, dy = 0.0f, dz = 0.0f;
    for (int s = 0; s < n; ++s) {
        dx += std::fabs(v[3 * s] - ax);
        dy += std::fabs(v[3 * s + 1] - ay);
        dz += std::fabs(v[3 * s + 2] - az);
    }
    float e = 0.0f;
    if (ax != 0.0f) e += dx / ax;
    if (ay != 0.0f) e += dy / ay;
    if (az != 0.0f) e += dz / az;
    e /= (float)n;
    e += 1.0e-5f;
    e *= vol;
    return e;
}

template <int D>
int initial_sampling(Grid& g) {
    long long L0 = (1LL << (g.b * D)) << (g.e << 1);
    long long N0 = 4 * L0;
    if (N0 > g.cap_s || L0 > g.cap_l || N0 > 0x7fffffffLL) return -3;
    std::fill(g.seed.begin(), g.seed.end(), 0u);
    g.L = L0;
    g.N = N0;
    g.new_s = g.new_l = 0;
#pragma omp parallel for schedule(static) num_threads(nthreads())
    for (long long i = 0; i < L0; ++i) {
        float off[D], ext[D];
        morton_cell<D>((uint32_t)i, g.b, g.e, g.sX, g.sY, off, ext);
        uint32_t s = tea4((uint32_t)i, 0);
        for (int j = 0; j < 4; ++j) {
            float r[D];
            for (int d = 0; d < D; ++d) r[d] = rnd(s);
            long long k = 4 * i + j;
            for (int d = 0; d < D; ++d) g.coords[k * D + d] = off[d] + r[d] * ext[d];
            g.idx[4 * i + j] = (int)k;
        }
        g.seed[i] = s;
        g.err[i] = -1.0f;
        float* bx = &g.box[i * 2 * D];
        for (int d = 0; d < D; ++d) {
            bx[d] = off[d];
            bx[D + d] = off[d] + ext[d];
        }
    }
    return 0;
}

template <int D>
int compute_errors(Grid& g) {
    float gmax = 0.0f;
    long long L = g.L;
#pragma omp parallel for schedule(static) reduction(max : gmax) num_threads(nthreads())
    for (long long i = 0; i < L; ++i) {
        float e = g.err[i];
        if (e < 0.0f) {
            float v[12];
            int n = 0;
            for (int s = 0; s < 4; ++s) {
                int k = g.idx[4 * i + s];
                if (k >= 0) {
                    v[3 * n] = g.values[3LL * k];
                    v[3 * n + 1] = g.values[3LL * k + 1];
                    v[3 * n + 2] = g.values[3LL * k + 2];
                    ++n;
                }
            }
            const float* bx = &g.box[i * 2 * D];
            e = leaf_error(v, n, volume<D>(bx, bx + D));
            g.err[i] = e;
        }
        gmax = std::max(gmax, e);
    }
    g.gmax = gmax;
    return 0;
}

template <int D>
int adaptive_sampling(Grid& g) {
    const long long L = g.L, N = g.N;
    const float gmax = g.gmax, alpha = g.alpha;
    g.flags.assign(L, 0);
    // 1. selection (kdgrid.cu:323-326): one draw per leaf, always consumed
#pragma omp parallel for schedule(static) num_threads(nthreads())
    for (long long i = 0; i < L; ++i) {
        uint32_t s = g.seed[i];
        if (s == 0) s = tea4((uint32_t)i, 0);
        float u = rnd(s);
        bool sel = powf_mdas(g.err[i] / gmax, alpha) >= u;
        bool full = g.idx[4 * i + 3] >= 0;
        g.flags[i] = (uint8_t)((sel ? 1 : 0) | ((sel && full) ? 2 : 0));
    }
    // 2. allocation: exclusive scan in leaf order
    g.rank_s.resize(L);
    g.rank_l.resize(L);
    long long ns = 0, nl = 0;
    for (long long i = 0; i < L; ++i) {
        g.rank_s[i] = (int)ns;
        g.rank_l[i] = (int)nl;
        ns += g.flags[i] & 1;
        nl += (g.flags[i] >> 1) & 1;
    }
    if (N + ns > g.cap_s || L + nl > g.cap_l) return -2;
    // 3. placement and split (kdgrid.cu:328-477)
#pragma omp parallel for schedule(static) num_threads(nthreads())
    for (long long i = 0; i < L; ++i) {
        uint32_t s = g.seed[i];
        if (s == 0) s = tea4((uint32_t)i, 0);
        (void)rnd(s);
        if (g.flags[i] & 1) {
            float* bx = &g.box[i * 2 * D];
            int* id = &g.idx[4 * i];
            float diag[D];
            for (int d = 0; d < D; ++d) diag[d] = bx[D + d] - bx[d];
            // best candidate (kdgrid.cu:332-361)
            float maxDist = -1.0f;
            float best[D];
            for (int d = 0; d < D; ++d) best[d] = 0.0f;
            for (int j = 0; j < g.cand; ++j) {
                float c[D];
                for (int d = 0; d < D; ++d) c[d] = rnd(s);
                for (int d = 0; d < D; ++d) c[d] = bx[d] + c[d] * diag[d];
                float minDist = FLT_MAX;
                for (int q = 0; q < 4; ++q) {
                    if (id[q] >= 0) {
                        const float* p = &g.coords[(long long)id[q] * D];
                        float acc = 0.0f;
                        for (int d = 0; d < D; ++d) {
                            float t = c[d] - p[d];
                            acc += t * t;
                        }
                        float dist = std::sqrt(acc);
                        if (minDist > dist) minDist = dist;
                    }
                }
                if (maxDist < minDist) {
                    maxDist = minDist;
                    for (int d = 0; d < D; ++d) best[d] = c[d];
                }
            }
            long long k = N + g.rank_s[i];
            for (int d = 0; d < D; ++d) g.coords[k * D + d] = best[d];
            int cnt = 0;
            int five[5];
            for (int q = 0; q < 4; ++q)
                if (id[q] >= 0) five[cnt++] = id[q];
            if (cnt < 4) {
                id[cnt] = (int)k;  // kdgrid.cu:393-396: insert, error kept
            } else {
                five[cnt++] = (int)k;
                int a = largest_axis(diag, D);
                // stable bubble sort by coordinate a (kdgrid.cu:409-417)
                for (int p = 0; p < cnt - 1; ++p)
                    for (int q = 0; q < cnt - p - 1; ++q)
                        if (g.coords[(long long)five[q] * D + a] >
                            g.coords[(long long)five[q + 1] * D + a])
                            std::swap(five[q], five[q + 1]);
                int md = cnt >> 1;
                float pos = 0.5f * (g.coords[(long long)five[md - 1] * D + a] +
                                    g.coords[(long long)five[md] * D + a]);
                long long m = L + g.rank_l[i];
                int* idm = &g.idx[4 * m];
                for (int q = 0; q < 4; ++q) id[q] = q < md ? five[q] : ~D;
                for (int q = 0; q < 4; ++q) idm[q] = q < cnt - md ? five[md + q] : ~D;
                float* bm = &g.box[m * 2 * D];
                for (int d = 0; d < 2 * D; ++d) bm[d] = bx[d];
                bx[D + a] = pos;
                bm[a] = pos;
                g.err[i] = -1.0f;
                g.err[m] = -1.0f;
            }
        }
        g.seed[i] = s;
    }
    g.N = N + ns;
    g.L = L + nl;
    g.new_s = ns;
    g.new_l = nl;
    return 0;
}

// Pixel span of a leaf (kdgrid.cu:524-529)
inline void pixel_span(const float* mn, const float* mx, int W, int H, float sX, float sY,
                       int* x0, int* x1, int* y0, int* y1) {
    *x0 = (int)(mn[0] * (float)W / sX);
    int xx = (int)(mx[0] * (float)W / sX);
    *y0 = (int)(mn[1] * (float)H / sY);
    int yy = (int)(mx[1] * (float)H / sY);
    *x1 = std::min(xx + 1, W);
    *y1 = std::min(yy + 1, H);
}

template <int D>
int integrate(Grid& g, float* px, int W, int H) {
    memset(px, 0, sizeof(float) * 4 * (size_t)W * H);
    const float stepX = g.sX / (float)W, stepY = g.sY / (float)H;
    const float pixArea = (g.sX * g.sY) / (float)(W * H);
    const float invPA = 1.0f / pixArea;
    const int T = nthreads();
    const long long L = g.L;
#pragma omp parallel num_threads(T)
    {
        int t = omp_get_thread_num(), nt = omp_get_num_threads();
        int r0 = (int)((long long)H * t / nt), r1 = (int)((long long)H * (t + 1) / nt);
        for (long long i = 0; i < L; ++i) {
            const float* mn = &g.box[i * 2 * D];
            const float* mx = mn + D;
            int x0, x1, y0, y1;
            pixel_span(mn, mx, W, H, g.sX, g.sY, &x0, &x1, &y0, &y1);
            if (y1 <= r0 || y0 >= r1) continue;
            float ax = 0.0f, ay = 0.0f, az = 0.0f;
            int n = 0;
            for (int s = 0; s < 4; ++s) {
                int k = g.idx[4 * i + s];
                if (k >= 0) {
                    ax += g.values[3LL * k];
                    ay += g.values[3LL * k + 1];
                    az += g.values[3LL * k + 2];
                    ++n;
                }
            }
            float inv = 1.0f / (float)n;
            ax *= inv;
            ay *= inv;
            az *= inv;
            const float vol = volume<D>(mn, mx);
            float pMinY = (float)y0 * g.sY / (float)H;
            float pMaxY = pMinY + stepY;
            for (int y = y0; y < y1; ++y) {
                if (y >= r0 && y < r1) {
                    float pMinX = (float)x0 * g.sX / (float)W;
                    float pMaxX = pMinX + stepX;
                    for (int x = x0; x < x1; ++x) {
                        const float clox = std::fmax(pMinX, mn[0]);
                        const float chix = std::fmin(pMaxX, mx[0]);
                        const float cloy = std::fmax(pMinY, mn[1]);
                        const float chiy = std::fmin(pMaxY, mx[1]);
                        if (clox < chix && cloy < chiy) {
                            float v = (chix - clox) * (chiy - cloy) * vol /
                                      ((mx[0] - mn[0]) * (mx[1] - mn[1]));
                            float* p = px + 4 * ((size_t)y * W + x);
                            p[0] += (ax * v) * invPA;
                            p[1] += (ay * v) * invPA;
                            p[2] += (az * v) * invPA;
                        }
                        pMinX = pMaxX;
                        pMaxX += stepX;
                    }
                }
                pMinY = pMaxY;
                pMaxY += stepY;
            }
        }
    }
    return 0;
}

inline uint32_t spread16(uint32_t v) {
    v &= 0xffffu;
    v = (v | (v << 8)) & 0x00ff00ffu;
    v = (v | (v << 4)) & 0x0f0f0f0fu;
    v = (v | (v << 2)) & 0x33333333u;
    v = (v | (v << 1)) & 0x55555555u;
    return v;
}

template <int D>
long long recon_keys(Grid& g, int W, int H, int TW, int TH, uint64_t* out, long long cap) {
    std::vector<uint64_t> keys;
    for (long long i = 0; i < g.L; ++i) {
        const float* mn = &g.box[i * 2 * D];
        int x0, x1, y0, y1;
        pixel_span(mn, mn + D, W, H, g.sX, g.sY, &x0, &x1, &y0, &y1);
        if (x0 >= x1 || y0 >= y1) continue;
        for (int ty = y0 / TH; ty <= (y1 - 1) / TH; ++ty)
            for (int tx = x0 / TW; tx <= (x1 - 1) / TW; ++tx)
                keys.push_back(((uint64_t)(spread16(tx) | (spread16(ty) << 1)) << 32) |
                               (uint64_t)i);
    }
    std::sort(keys.begin(), keys.end());
    long long n = (long long)keys.size();
    if (out) memcpy(out, keys.data(), sizeof(uint64_t) * (size_t)std::min(n, cap));
    return n;
}

// helpers.h:35-66
inline unsigned char quantize8(float x) {
    x = std::fmax(0.0f, std::fmin(x, 1.0f));
    unsigned v = (unsigned)(x * 256.0f);
    return (unsigned char)std::min(v, 255u);
}
inline void make_color(const float* c, unsigned char* out) {
    const float invGamma = 1.0f / 2.4f;
    for (int ch = 0; ch < 3; ++ch) {
        float x = std::fmax(0.0f, std::fmin(c[ch], 1.0f));
        float p = powf_mdas(x, invGamma);
        float s = x < 0.0031308f ? 12.92f * x : 1.055f * p - 0.055f;
        out[ch] = quantize8(s);
    }
    out[3] = 255;
}

// kdgrid.cu:848-961: the reference's four conditions -- every sample inside
// its leaf's box, the boxes' volumes summing to scaleX * scaleY (1e-2
// relative), as many slots filled as samples, every sample referenced once
// -- plus a slot index below N (the reference would read past the arrays).
template <int D>
int validate(Grid& g) {
    bool ok = true;
    std::vector<int> hist(g.N, 0);
    double totalVol = 0.0;
    long long total = 0;
    for (long long i = 0; i < g.L; ++i) {
        const float* mn = &g.box[i * 2 * D];
        const float* mx = mn + D;
        for (int s = 0; s < 4; ++s) {
            int k = g.idx[4 * i + s];
            if (k < 0) continue;  // an empty slot (kdgrid.cu:877)
            if (k >= g.N) {
                ok = false;
                continue;
            }
            for (int d = 0; d < D; ++d) {
                float c = g.coords[(long long)k * D + d];
                if (c < mn[d] || c > mx[d]) ok = false;  // aabb.h:Contains
            }
            hist[k]++;
            ++total;
        }
        totalVol += (double)volume<D>(mn, mx);
    }
    double root = (double)g.sX * (double)g.sY;
    if (std::fabs(totalVol - root) > 1.0e-2 * root) ok = false;  // kdgrid.cu:909-914
    if (total != g.N) ok = false;                                 // :916-920
    for (long long k = 0; k < g.N; ++k)
        if (hist[k] != 1) ok = false;  // :922-956
    return ok ? 1 : 0;
}

}  // namespace

struct orc_grid {
    Grid g;
};

#define ORC_DISPATCH(D_, CALL)                      \
    switch (D_) {                                   \
    case 3: { const int DD = 3; return CALL; }      \
    case 4: { const int DD = 4; return CALL; }      \
    case 5: { const int DD = 5; return CALL; }      \
    case 6: { const int DD = 6; return CALL; }      \
    default: return -1;                             \
    }

// cuda/random.h:74-89: radical inverse of i in base primes[dim]; the
// reference divides f by the base each digit and takes floor(j / base).
template <typename real>
real halton(int dim, int i) {
    static const int primes[6] = {2, 3, 5, 7, 11, 13};
    const int base = primes[dim];
    real r = 0, f = (real)1 / (real)base;
    for (int j = i; j > 0; j /= base) {
        r += f * (real)(j % base);
        f /= (real)base;
    }
    return r;
}

// cuda/sampler.h:4-28 + whitted.cu:59-69: seed = tea4(pixel, frame), six
// offsets from rnd(seed), sample s of dim d = halton(d, s) + offset[d], minus
// 1 when above 1; dims 0/1 are the jitter inside pixel (px, py), mapped to
// the MDAS image coordinates [0, scale_x) x [0, scale_y).  Pixel value = sum
// of the spp values in sample order times 1/spp.
template <typename real>
void qmc_image(const mscene::Params& P, int W, int H, int spp, int first, uint32_t frame,
               real* out) {
    const int dim = P.dim;
#pragma omp parallel for schedule(dynamic, 64) num_threads(nthreads())
    for (long long pix = 0; pix < (long long)W * H; ++pix) {
        const int px = (int)(pix % W), py = (int)(pix / W);
        uint32_t seed = tea4((uint32_t)pix, frame);
        real off[6];
        for (int d = 0; d < 6; ++d) off[d] = (real)rnd(seed);
        real acc[3] = {0, 0, 0};
        for (int s = first; s < first + spp; ++s) {
            real x[6];
            for (int d = 0; d < dim; ++d) {
                real v = halton<real>(d, s) + off[d];
                if (v > (real)1) v -= (real)1;
                x[d] = v;
            }
            x[0] = ((real)px + x[0]) * (real)P.scale_x / (real)W;
            x[1] = ((real)py + x[1]) * (real)P.scale_y / (real)H;
            const mscene::V3<real> v = mscene::eval<real>(P, P.obj, x);
            acc[0] += v.x;
            acc[1] += v.y;
            acc[2] += v.z;
        }
        const real inv = (real)1 / (real)spp;
        real* o = out + pix * 4;
        o[0] = acc[0] * inv;
        o[1] = acc[1] * inv;
        o[2] = acc[2] * inv;
        o[3] = 0;
    }
}

extern "C" {

orc_grid* orc_create(int dim, long long max_samples, int candidates_num, int bits_per_dim,
                     int extra_img_bits, float alpha, float scale_x, float scale_y) {
    if (dim < 3 || dim > 6 || max_samples < 8) return nullptr;
    orc_grid* h = new orc_grid;
    Grid& g = h->g;
    g.D = dim;
    g.cap_s = max_samples;
    g.cap_l = max_samples / 2 + 1;
    g.cand = candidates_num;
    g.b = bits_per_dim;
    g.e = extra_img_bits;
    g.alpha = alpha;
    g.sX = scale_x;
    g.sY = scale_y;
    g.coords.assign((size_t)g.cap_s * dim, 0.0f);
    g.values.assign((size_t)g.cap_s * 3, 0.0f);
    g.idx.assign((size_t)g.cap_l * 4, ~dim);
    g.box.assign((size_t)g.cap_l * 2 * dim, 0.0f);
    g.err.assign((size_t)g.cap_l, 0.0f);
    g.seed.assign((size_t)g.cap_l, 0u);
    return h;
}

void orc_destroy(orc_grid* h) { delete h; }
void orc_set_threads(int n) { g_threads = n; }

int orc_initial_sampling(orc_grid* h) { ORC_DISPATCH(h->g.D, initial_sampling<DD>(h->g)); }
int orc_compute_errors(orc_grid* h) { ORC_DISPATCH(h->g.D, compute_errors<DD>(h->g)); }
int orc_adaptive_sampling(orc_grid* h) { ORC_DISPATCH(h->g.D, adaptive_sampling<DD>(h->g)); }
int orc_integrate(orc_grid* h, float* px, int w, int hgt) {
    ORC_DISPATCH(h->g.D, integrate<DD>(h->g, px, w, hgt));
}
int orc_validate(orc_grid* h) { ORC_DISPATCH(h->g.D, validate<DD>(h->g)); }
long long orc_recon_keys(orc_grid* h, int w, int hgt, int tw, int th, uint64_t* out,
                         long long cap) {
    ORC_DISPATCH(h->g.D, recon_keys<DD>(h->g, w, hgt, tw, th, out, cap));
}

void orc_convert_to_bytes(const float* px, unsigned char* out, long long n) {
#pragma omp parallel for schedule(static) num_threads(nthreads())
    for (long long i = 0; i < n; ++i) make_color(px + 4 * i, out + 4 * i);
}

void orc_sampling_density(orc_grid* h, float* px, int W, int H) {
    Grid& g = h->g;
    memset(px, 0, sizeof(float) * 4 * (size_t)W * H);
    for (long long i = 0; i < g.N; ++i) {
        int x = (int)(g.coords[i * g.D] / g.sX * (float)W);
        int y = (int)(g.coords[i * g.D + 1] / g.sY * (float)H);
        x = std::min(x, W - 1);
        y = std::min(y, H - 1);
        float* p = px + 4 * ((size_t)y * W + x);
        for (int c = 0; c < 4; ++c) p[c] += 0.05f;
    }
    for (long long i = 0; i < (long long)W * H; ++i) px[4 * i + 3] = 1.0f;
}

long long orc_make_scene(int kind, uint32_t seed, int dim, float sx, float sy, void* out,
                         long long cap) {
    const long long n = (long long)sizeof(mscene::Params);
    if (out && cap >= n) mscene::make_scene((mscene::Params*)out, kind, seed, dim, sx, sy);
    return n;
}

void orc_eval_scene(const void* params, const float* coords, int dim, long long first,
                    long long count, float* values) {
    const mscene::Params& P = *(const mscene::Params*)params;
#pragma omp parallel for schedule(dynamic, 4096) num_threads(nthreads())
    for (long long k = first; k < first + count; ++k) {
        mscene::V3<float> v = mscene::eval<float>(P, P.obj, coords + k * dim);
        values[3 * k] = v.x;
        values[3 * k + 1] = v.y;
        values[3 * k + 2] = v.z;
    }
}

void orc_eval_scene_f64(const void* params, const double* coords, int dim, long long count,
                        double* values) {
    const mscene::Params& P = *(const mscene::Params*)params;
#pragma omp parallel for schedule(dynamic, 1024) num_threads(nthreads())
    for (long long k = 0; k < count; ++k) {
        mscene::V3<double> v = mscene::eval<double>(P, P.obj, coords + k * dim);
        values[3 * k] = v.x;
        values[3 * k + 1] = v.y;
        values[3 * k + 2] = v.z;
    }
}

void orc_eval_scene_grouped(const void* params, const void* coords, int dim, long long count,
                            int fp64, void* values) {
    const mscene::Params& P = *(const mscene::Params*)params;
    mscene::Accel A;
    mscene::build_accel(P, &A);
#pragma omp parallel for schedule(dynamic, 1024) num_threads(nthreads())
    for (long long k = 0; k < count; ++k) {
        if (fp64) {
            const mscene::V3<double> v =
                mscene::eval<double>(P, P.obj, (const double*)coords + k * dim, &A);
            double* o = (double*)values + 3 * k;
            o[0] = v.x;
            o[1] = v.y;
            o[2] = v.z;
        } else {
            const mscene::V3<float> v =
                mscene::eval<float>(P, P.obj, (const float*)coords + k * dim, &A);
            float* o = (float*)values + 3 * k;
            o[0] = v.x;
            o[1] = v.y;
            o[2] = v.z;
        }
    }
}

float orc_halton(int dim, int i) { return halton<float>(dim, i); }

void orc_qmc_image(const void* params, int w, int h, int spp, int first, uint32_t frame, int fp64,
                   void* out) {
    const mscene::Params& P = *(const mscene::Params*)params;
    if (fp64)
        qmc_image<double>(P, w, h, spp, first, frame, (double*)out);
    else
        qmc_image<float>(P, w, h, spp, first, frame, (float*)out);
}

void orc_eval_range(orc_grid* h, const void* params, long long first, long long count) {
    Grid& g = h->g;
    orc_eval_scene(params, g.coords.data(), g.D, first, count, g.values.data());
}

int orc_run(orc_grid* h, const void* params, long long target, int* passes) {
    Grid& g = h->g;
    int rc = orc_initial_sampling(h);
    if (rc) return rc;
    orc_eval_scene(params, g.coords.data(), g.D, 0, g.N, g.values.data());
    long long limit = std::min(target, g.cap_s);
    int p = 0;
    // optixMotionBlur.cpp:781 (N < min(target, max_samples)); when the limit
    // is the capacity, also no pass that could overflow it: N + L <=
    // max_samples, the guard of the app's interactive loop (:723-724)
    const bool guard = target >= g.cap_s;
    while (g.N < limit && (!guard || g.N + g.L <= g.cap_s)) {
        orc_compute_errors(h);
        rc = orc_adaptive_sampling(h);
        if (rc) return rc;
        orc_eval_scene(params, g.coords.data(), g.D, g.N - g.new_s, g.new_s, g.values.data());
        ++p;
    }
    if (passes) *passes = p;
    return 0;
}

long long orc_num_samples(orc_grid* h) { return h->g.N; }
long long orc_num_leaves(orc_grid* h) { return h->g.L; }
long long orc_new_samples(orc_grid* h) { return h->g.new_s; }
long long orc_new_nodes(orc_grid* h) { return h->g.new_l; }
float orc_max_error(orc_grid* h) { return h->g.gmax; }

void orc_get_state(orc_grid* h, float* coords, float* values, int* idx, float* box, float* err,
                   uint32_t* seed) {
    Grid& g = h->g;
    if (coords) memcpy(coords, g.coords.data(), sizeof(float) * g.N * g.D);
    if (values) memcpy(values, g.values.data(), sizeof(float) * g.N * 3);
    if (idx) memcpy(idx, g.idx.data(), sizeof(int) * g.L * 4);
    if (box) memcpy(box, g.box.data(), sizeof(float) * g.L * 2 * g.D);
    if (err) memcpy(err, g.err.data(), sizeof(float) * g.L);
    if (seed) memcpy(seed, g.seed.data(), sizeof(uint32_t) * g.L);
}

int orc_set_state(orc_grid* h, long long n, long long l, const float* coords, const float* values,
                  const int* idx, const float* box, const float* err, const uint32_t* seed,
                  float gmax) {
    Grid& g = h->g;
    if (n > g.cap_s || l > g.cap_l) return -2;
    g.N = n;
    g.L = l;
    g.gmax = gmax;
    if (coords) memcpy(g.coords.data(), coords, sizeof(float) * n * g.D);
    if (values) memcpy(g.values.data(), values, sizeof(float) * n * 3);
    if (idx) memcpy(g.idx.data(), idx, sizeof(int) * l * 4);
    if (box) memcpy(g.box.data(), box, sizeof(float) * l * 2 * g.D);
    if (err) memcpy(g.err.data(), err, sizeof(float) * l);
    if (seed) {
        memcpy(g.seed.data(), seed, sizeof(uint32_t) * l);
        std::fill(g.seed.begin() + l, g.seed.end(), 0u);
    }
    return 0;
}

void orc_set_values(orc_grid* h, long long first, long long count, const float* values) {
    memcpy(h->g.values.data() + 3 * first, values, sizeof(float) * 3 * count);
}

uint32_t orc_tea(uint32_t v0, uint32_t v1, int rounds, const uint32_t* key, uint32_t* v1_out) {
    return tea_n(v0, v1, rounds, key, v1_out);
}
uint32_t orc_tea4(uint32_t v0, uint32_t v1) { return tea4(v0, v1); }
uint32_t orc_lcg(uint32_t* s) { return lcg(*s); }
float orc_rnd(uint32_t* s) { return rnd(*s); }
float orc_powf(float x, float a) { return powf_mdas(x, a); }

void orc_morton_cell(long long leaf, int dim, int b, int e, uint32_t* xq, float* off, float* ext) {
    switch (dim) {
    case 3: morton_cell<3>((uint32_t)leaf, b, e, 1.0f, 1.0f, off, ext, xq); break;
    case 4: morton_cell<4>((uint32_t)leaf, b, e, 1.0f, 1.0f, off, ext, xq); break;
    case 5: morton_cell<5>((uint32_t)leaf, b, e, 1.0f, 1.0f, off, ext, xq); break;
    case 6: morton_cell<6>((uint32_t)leaf, b, e, 1.0f, 1.0f, off, ext, xq); break;
    }
}

int orc_largest_axis(const float* diag, int dim) { return largest_axis(diag, dim); }

void orc_sort5(const float* keys, int* order) {
    for (int i = 0; i < 5; ++i) order[i] = i;
    for (int p = 0; p < 4; ++p)
        for (int q = 0; q < 4 - p; ++q)
            if (keys[order[q]] > keys[order[q + 1]]) std::swap(order[q], order[q + 1]);
}

float orc_leaf_error(const float* vals, int count, float vol) { return leaf_error(vals, count, vol); }

void orc_make_color(const float* rgb, unsigned char* out) { make_color(rgb, out); }

}  // extern "C"
