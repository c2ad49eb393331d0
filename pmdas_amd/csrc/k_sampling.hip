// Sampling kernels of the MDAS grid: initial Morton grid, per-leaf error with
// the global max, and the adaptive pass (selection, best-candidate placement,
// median split) with indices allocated by a deterministic device-wide scan.
//
// Layout (HBM): sample coordinates AoS float[D] and values float3 (the
// integrand contract); one 128-byte line per leaf (slots, box and its first
// points) plus an overflow row for the other points (D >= 5), so a leaf
// touched at a random index costs one line; errors, LCG states and a one-byte
// flag word (the sample count) per leaf as separate streams for the per-pass
// sweep over all leaves.  Cross-workgroup reductions go through per-workgroup
// partials (no same-address atomics).
#include "mdas_leaf.hpp"

#pragma clang fp contract(off)

namespace mdas {

#define MDAS_ERR_CAPACITY_DEV (-2)

// ----------------------------------------------------------------- initial --
// kdgrid.cu:121-202.  One thread per initial leaf; four samples per leaf are
// drawn from tea4(leaf, 0) in dimension order.  The 4*D coordinates of a
// workgroup's 256 leaves are one contiguous range, written through LDS so
// that every store instruction covers 256 contiguous bytes.
template <int D>
__device__ __forceinline__ void initial_body(const DevGrid& a, int L0, int blk, int nblk) {
    __shared__ float stage[256 * 4 * D];
    for (int base = blk * 256; base < L0; base += nblk * 256) {
        const int i = base + threadIdx.x;
        if (i < L0) {
            float off[D], ext[D];
            morton_cell<D>((unsigned)i, a.b, a.e, a.sX, a.sY, off, ext);
            unsigned s = tea4((unsigned)i, 0u);
            float pts[4][D];
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                float r[D];
#pragma unroll
                for (int d = 0; d < D; ++d) r[d] = rnd(s);
#pragma unroll
                for (int d = 0; d < D; ++d) {
                    pts[j][d] = off[d] + r[d] * ext[d];
                    stage[threadIdx.x * 4 * D + j * D + d] = pts[j][d];
                }
            }
            float mx[D];
#pragma unroll
            for (int d = 0; d < D; ++d) mx[d] = off[d] + ext[d];
            // the leaf's line (slots, box, first points) and its overflow row
            store_line<D>(a, i, make_int4(4 * i, 4 * i + 1, 4 * i + 2, 4 * i + 3), off, mx, pts, 4);
            if constexpr (line_pts(D) < 4) {
                // the overflow row: the points past the line and the slots
                // kept beside them
                constexpr int P = line_pts(D), XS = xpts_stride(D);
                float x[XS];
#pragma unroll
                for (int k = 0; k < XS; ++k) x[k] = 0.0f;
#pragma unroll
                for (int q = P; q < 4; ++q)
#pragma unroll
                    for (int d = 0; d < D; ++d) x[xpt_at(D, q) + d] = pts[q][d];
#pragma unroll
                for (int q = 2; q < 4; ++q)
                    if (slot_at(D, q) < 0) x[-1 - slot_at(D, q)] = __int_as_float(4 * i + q);
                float4* xr = reinterpret_cast<float4*>(a.xpts + (size_t)i * XS);
#pragma unroll
                for (int k = 0; k < XS / 4; ++k) xr[k] = make_float4(x[4 * k], x[4 * k + 1], x[4 * k + 2], x[4 * k + 3]);
            }
            a.seed[i] = s;
            a.err[i] = -1.0f;
            a.flags[i] = 4;
        }
        __syncthreads();
        // leaf i's samples are 4i..4i+3: the workgroup's coordinates are one
        // contiguous range, stored from LDS in full lines
        const int nl = min(256, L0 - base);
        float* out = a.coords + (size_t)base * 4 * D;
        for (int k = threadIdx.x; k < nl * 4 * D; k += 256) out[k] = stage[k];
        __syncthreads();
    }
}

template <int D>
__global__ __launch_bounds__(256) void k_initial(DevGrid a, int L0) {
    initial_body<D>(a, L0, blockIdx.x, gridDim.x);
}

// Batched form: image blockIdx.y of the set, initial leaf count from its
// configuration.
template <int D>
__global__ __launch_bounds__(256) void k_initial_b(const DevGrid* grids) {
    const DevGrid& a = grids[blockIdx.y];
    initial_body<D>(a, 1 << (a.b * D + 2 * a.e), blockIdx.x, gridDim.x);
}

// Counter set-up of InitialSampling (kdgrid.cu:639-642).
__device__ __forceinline__ void initial_counters_body(DevCounters* c, int L0) {
    c->n_leaves = L0;
    c->n_samples = 4 * L0;
    c->n_leaves_old = 0;
    c->n_samples_old = 0;
    c->new_samples = 4 * L0;  // the first integrand launch covers all of them
    c->new_nodes = 0;
    c->gmax_bits = 0;
    c->done = 0;
    c->status = 0;
    c->passes = 0;
    c->limit = 0x7fffffff;
    c->sel_cnt = 0;
    c->stale = 0;
    c->stale_vals = 0;
    c->carry_bits = 0;
    c->guard_cap = 0;
}
__global__ void k_initial_counters(DevCounters* c, int L0) { initial_counters_body(c, L0); }

// ComputeErrors prologue (kdgrid.cu:684-685): g_error = 0.  A standalone call
// also re-arms the pass kernels and lifts the sampling loop's limit.
__global__ void k_reset_errors(DevCounters* c, int standalone) {
    if (standalone) {
        c->done = 0;
        c->limit = 0x7fffffff;
        c->guard_cap = 0;
    }
    c->gmax_bits = 0;
}

__global__ void k_set_limit(DevCounters* c, int limit, int guard_cap) {
    c->limit = limit;
    c->guard_cap = guard_cap;
}

// Batched counter kernels: one thread per image.
__global__ void k_initial_counters_b(const DevGrid* grids, int n, const int* limits) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n) return;
    const DevGrid& a = grids[t];
    DevCounters* c = a.cnt;
    initial_counters_body(c, 1 << (a.b * a.D + 2 * a.e));
    c->limit = limits[t];
    c->guard_cap = limits[t] >= a.cap_s ? a.cap_s : 0;
}

// Re-arms the pass kernels for a standalone AdaptiveSampling call.
__global__ void k_arm(DevCounters* c) {
    c->done = 0;
    c->status = 0;
}

// ------------------------------------------------------------------ errors --
// Workgroup reduction of (max, stale count, values read) into epart[block].
__device__ __forceinline__ void errors_epilogue(const DevGrid& a, float m, int n_stale, int n_vals,
                                                float* wmax, int* wst, int blk) {
    m = wave_max(m);
    n_stale = wave_sum(n_stale);
    n_vals = wave_sum(n_vals);
    if ((threadIdx.x & 63) == 0) {
        wmax[threadIdx.x >> 6] = m;
        wst[threadIdx.x >> 6] = n_stale;
        wst[4 + (threadIdx.x >> 6)] = n_vals;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        const float b = fmaxf(fmaxf(wmax[0], wmax[1]), fmaxf(wmax[2], wmax[3]));
        a.epart[blk] = make_uint4(__float_as_uint(b), (unsigned)(wst[0] + wst[1] + wst[2] + wst[3]),
                                         (unsigned)(wst[4] + wst[5] + wst[6] + wst[7]), 0u);
    }
}

// kdgrid.cu:204-288 as the reference runs it: every leaf with error < 0 is
// re-estimated and every leaf takes part in the global max.  The max of
// non-negative floats is the max of their bit patterns, so the reduction is
// exact in any order.
template <int D>
__global__ __launch_bounds__(256) void k_errors(DevGrid a) {
    __shared__ float wmax[4];
    __shared__ int wst[8];
    if (a.cnt->done) return;
    const int L = a.cnt->n_leaves;
    float m = 0.0f;
    int n_stale = 0, n_vals = 0;
    for (int i = blockIdx.x * 256 + threadIdx.x; i < L; i += gridDim.x * 256) {
        float e = a.err[i];
        if (e < 0.0f) {
            int n;
            e = leaf_error<D>(a, i, &n);
            a.err[i] = e;
            n_stale += 1;
            n_vals += n;
        }
        m = fmaxf(m, e);
    }
    errors_epilogue(a, m, n_stale, n_vals, wmax, wst, blockIdx.x);
}

// The same result inside the sampling loop without the sweep over all
// leaves: the only stale leaves are the two children of every leaf the last
// pass split (split_list[t] and L_old + t, whose sample indices and volumes
// the placement left in the split's entry; inside the loop it does not mark
// them -1, since they are re-estimated here); the max over the other leaves
// was carried by the selection kernel into gmax by the allocation scan.  The
// first pass after the initial grid re-estimates every leaf.
template <int D>
__device__ __forceinline__ void errors_stale_body(const DevGrid& a, int blk, int nblk) {
    __shared__ float wmax[4];
    __shared__ int wst[8];
    if (pass_stopped(a.cnt)) return;
    const bool first = a.cnt->passes == 0;
    float m = 0.0f;
    int n_stale = 0, n_vals = 0;
    if (first) {
        const int L = a.cnt->n_leaves;
        for (int i = blk * 256 + threadIdx.x; i < L; i += nblk * 256) {
            int n;
            const float e = leaf_error<D>(a, i, &n);
            a.err[i] = e;
            n_stale += 1;
            n_vals += n;
            m = fmaxf(m, e);
        }
    } else {
        // one thread per split: both children from the split's entry
        // (sample indices and volumes), not from their leaf records
        const int nn = a.cnt->new_nodes;
        const int L_old = a.cnt->n_leaves_old;
        for (int t = blk * 256 + threadIdx.x; t < nn; t += nblk * 256) {
            const int4 s0 = a.sdat[2 * t], s1 = a.sdat[2 * t + 1];
            const int il = a.split_list[t];
            LeafSamples lsl, lsr;
            const int idl[4] = {s0.x, s0.y, -1, -1};
            const int idr[4] = {s0.z, s0.w, s1.x, -1};
            load_values(a, idl, lsl);  // the five value loads together
            load_values(a, idr, lsr);
            const float el = error_of_values(idl, __int_as_float(s1.y), lsl);
            const float er = error_of_values(idr, __int_as_float(s1.z), lsr);
            a.err[il] = el;
            a.err[L_old + t] = er;
            n_stale += 2;
            n_vals += lsl.n + lsr.n;
            m = fmaxf(m, el);
            m = fmaxf(m, er);
        }
    }
    errors_epilogue(a, m, n_stale, n_vals, wmax, wst, blk);
}

template <int D>
__global__ __launch_bounds__(256) void k_errors_stale(DevGrid a) {
    errors_stale_body<D>(a, blockIdx.x, gridDim.x);
}
template <int D>
__global__ __launch_bounds__(256) void k_errors_stale_b(const DevGrid* grids) {
    errors_stale_body<D>(grids[blockIdx.y], blockIdx.x, gridDim.x);
}

// Folds the error partials into gmax and the per-pass statistics.  A full
// sweep (ComputeErrors, after its reset) starts from gmax = 0; the loop's
// stale-list pass, which re-estimates only the last pass's children, starts
// from the max the allocation scan carried over (the leaves that did not
// split).
__device__ __forceinline__ void err_reduce_body(const DevGrid& a, int G, int from_carry) {
    __shared__ unsigned sm[16];
    __shared__ int ss[16], sv[16];
    if (a.cnt->done) return;
    if (loop_limit_reached(a.cnt)) {  // the loop stops before this pass
        if (threadIdx.x == 0) a.cnt->done = 1;
        return;
    }
    unsigned m = 0;
    int st = 0, va = 0;
    for (int k = threadIdx.x; k < G; k += 1024) {
        const uint4 p = a.epart[k];
        m = max(m, p.x);
        st += (int)p.y;
        va += (int)p.z;
    }
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
        m = max(m, (unsigned)__shfl_xor((int)m, o, 64));
        st += __shfl_xor(st, o, 64);
        va += __shfl_xor(va, o, 64);
    }
    if ((threadIdx.x & 63) == 0) {
        sm[threadIdx.x >> 6] = m;
        ss[threadIdx.x >> 6] = st;
        sv[threadIdx.x >> 6] = va;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        unsigned mm = from_carry ? a.cnt->carry_bits : a.cnt->gmax_bits;
        int s1 = 0, s2 = 0;
        for (int k = 0; k < 16; ++k) {
            mm = max(mm, sm[k]);
            s1 += ss[k];
            s2 += sv[k];
        }
        a.cnt->gmax_bits = mm;
        a.cnt->stale += s1;
        a.cnt->stale_vals += s2;
    }
}
__global__ __launch_bounds__(1024) void k_err_reduce(DevGrid a, int G, int from_carry) {
    err_reduce_body(a, G, from_carry);
}
__global__ __launch_bounds__(1024) void k_err_reduce_b(const DevGrid* grids, int G) {
    err_reduce_body(grids[blockIdx.x], G, 1);
}

// --------------------------------------------------------------- selection --
// kdgrid.cu:316-326.  Per leaf: the seed (tea4(i,0) when unset), one draw u,
// selected iff pow(err / g_error, alpha) >= u.  A selected full leaf splits.
// Writes the flags, and per workgroup the (selected, split) counts for the
// allocation scan, the samples held by the selected leaves, and the max error
// of the leaves that do not split.  The LCG states are not written here: the
// compaction stores them once the scan has accepted the pass, so that a pass
// refused for capacity leaves every state as it was.
// Chunk blk of nblk (the compaction's partition of the leaves).
__device__ __forceinline__ void select_chunk(const DevGrid& a, int blk, int nblk) {
    __shared__ int red[12];
    __shared__ float rmax[4];
    const int L = a.cnt->n_leaves;
    const float gmax = __uint_as_float(a.cnt->gmax_bits);
    const float alpha = a.alpha;
    int beg, end;
    chunk_range(L, nblk, blk, beg, end);
    int ns = 0, nl = 0, nc = 0;
    float carry = 0.0f;
    // four leaves per thread and step, their loads issued together
    constexpr int U = 4;
    for (int i0 = beg + threadIdx.x; i0 < end; i0 += U * 256) {
        float ev[U];
        unsigned sv[U], fv[U];
#pragma unroll
        for (int r = 0; r < U; ++r) {
            const int i = i0 + r * 256;
            if (i < end) {
                ev[r] = a.err[i];
                sv[r] = a.seed[i];
                fv[r] = a.flags[i];
            }
        }
#pragma unroll
        for (int r = 0; r < U; ++r) {
            const int i = i0 + r * 256;
            if (i >= end) continue;
            const float e = ev[r];
            unsigned s = sv[r];
            if (s == 0u) s = tea4((unsigned)i, 0u);
            const float u = rnd(s);
            const bool sel = select_test(e / gmax, alpha, u);
            const unsigned cnt = fv[r] & 7u;
            const bool split = sel && cnt == 4u;
            a.flags[i] = (unsigned char)(cnt | (sel ? 8u : 0u) | (split ? 16u : 0u));
            ns += sel ? 1 : 0;
            nl += split ? 1 : 0;
            nc += sel ? (int)cnt : 0;
            if (!split) carry = fmaxf(carry, e);
        }
    }
    ns = wave_sum(ns);
    nl = wave_sum(nl);
    nc = wave_sum(nc);
    carry = wave_max(carry);
    if ((threadIdx.x & 63) == 0) {
        red[threadIdx.x >> 6] = ns;
        red[4 + (threadIdx.x >> 6)] = nl;
        red[8 + (threadIdx.x >> 6)] = nc;
        rmax[threadIdx.x >> 6] = carry;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        const float cm = fmaxf(fmaxf(rmax[0], rmax[1]), fmaxf(rmax[2], rmax[3]));
        a.part[blk] = make_int4(red[0] + red[1] + red[2] + red[3], red[4] + red[5] + red[6] + red[7],
                                red[8] + red[9] + red[10] + red[11], (int)__float_as_uint(cm));
    }
    __syncthreads();  // red / rmax are reused by the workgroup's next chunk
}
// Workgroup blk takes chunks [blk * cpw, (blk + 1) * cpw) of the compaction's
// nblk * cpw: the placement balances better over many small chunks, the
// selection's streaming sweep runs best with fewer, longer ones.
__device__ __forceinline__ void select_body(const DevGrid& a, int blk, int nblk, int cpw) {
    if (a.cnt->done) return;
    for (int c = blk * cpw; c < (blk + 1) * cpw; ++c) select_chunk(a, c, nblk * cpw);
}
__global__ __launch_bounds__(256) void k_select(DevGrid a, int cpw) {
    select_body(a, blockIdx.x, gridDim.x, cpw);
}
__global__ __launch_bounds__(256) void k_select_b(const DevGrid* grids, int cpw) {
    select_body(grids[blockIdx.y], blockIdx.x, gridDim.x, cpw);
}

// Chunks per selection workgroup for `chunks` compaction chunks: as many as
// keep the selection at 4096 workgroups (chunks is then a multiple of it).
static int select_cpw(int chunks) { return chunks > 4096 && chunks % 4096 == 0 ? chunks / 4096 : 1; }

// Exclusive scan of the per-workgroup counts (one workgroup of 1024 threads),
// then the capacity check and the counter update (kdgrid.cu:753-759).
__device__ __forceinline__ void scan_counts_body(const DevGrid& a, int G) {
    __shared__ int ls[16], ll[16], lc[16];
    __shared__ unsigned lm[16];
    DevCounters* cn = a.cnt;
    if (cn->done) return;
    // thread 0's counter reads go out together with the partials' loads
    // (no other kernel runs between the error pass and this one)
    int N = 0, L = 0, P = 0, st = 0, sv = 0;
    if (threadIdx.x == 0) {
        N = cn->n_samples;
        L = cn->n_leaves;
        P = cn->passes;
        st = cn->stale;
        sv = cn->stale_vals;
    }
    const int per = (G + 1023) / 1024;
    const int g0 = threadIdx.x * per;
    // up to R partials per thread (G <= 8192 chunks) stay in registers for
    // the offsets written below; more are read again there
    constexpr int R = 8;
    int4 pr[R];
    int s = 0, l = 0, c = 0;
    unsigned mx = 0;
    if (per <= R) {
#pragma unroll
        for (int k = 0; k < R; ++k)
            pr[k] = k < per && g0 + k < G ? a.part[g0 + k] : make_int4(0, 0, 0, 0);
#pragma unroll
        for (int k = 0; k < R; ++k) {
            s += pr[k].x;
            l += pr[k].y;
            c += pr[k].z;
            mx = max(mx, (unsigned)pr[k].w);
        }
    } else {
        for (int k = 0; k < per; ++k)
            if (g0 + k < G) {
                int4 p = a.part[g0 + k];
                s += p.x;
                l += p.y;
                c += p.z;
                mx = max(mx, (unsigned)p.w);
            }
    }
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    int xs = s, xl = l, xc = c;
    unsigned xm = mx;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        int ys = __shfl_up(xs, o, 64), yl = __shfl_up(xl, o, 64), yc = __shfl_up(xc, o, 64);
        unsigned ym = (unsigned)__shfl_up((int)xm, o, 64);
        if (lane >= o) {
            xs += ys;
            xl += yl;
            xc += yc;
            xm = max(xm, ym);
        }
    }
    if (lane == 63) {
        ls[w] = xs;
        ll[w] = xl;
        lc[w] = xc;
        lm[w] = xm;
    }
    __syncthreads();
    int ps = 0, pl = 0, ts = 0, tl = 0, tc = 0;
    unsigned tm = 0;
    for (int k = 0; k < 16; ++k) {
        if (k < w) {
            ps += ls[k];
            pl += ll[k];
        }
        ts += ls[k];
        tl += ll[k];
        tc += lc[k];
        tm = max(tm, lm[k]);
    }
    int rs = ps + xs - s, rl = pl + xl - l;
    if (per <= R) {
#pragma unroll
        for (int k = 0; k < R; ++k)
            if (k < per && g0 + k < G) {
                a.part_off[g0 + k] = make_int2(rs, rl);
                rs += pr[k].x;
                rl += pr[k].y;
            }
    } else {
        for (int k = 0; k < per; ++k)
            if (g0 + k < G) {
                int4 p = a.part[g0 + k];
                a.part_off[g0 + k] = make_int2(rs, rl);
                rs += p.x;
                rl += p.y;
            }
    }
    if (threadIdx.x == 0) {
        if ((long long)N + ts > a.cap_s || (long long)L + tl > a.cap_l) {
            cn->status = MDAS_ERR_CAPACITY_DEV;
            cn->done = 1;
        } else {
            cn->n_samples_old = N;
            cn->n_leaves_old = L;
            cn->n_samples = N + ts;
            cn->n_leaves = L + tl;
            cn->new_samples = ts;
            cn->new_nodes = tl;
            // the next stale-list error pass starts from it; g_error itself
            // stays until the next error pass (kdgrid.cu:684-716)
            cn->carry_bits = tm;
            a.pstats[P & (kMaxPassStats - 1)] = PassStat{L, N, ts, tl, tc, st, sv, 0};
            cn->passes = P + 1;
        }
        cn->sel_cnt = 0;
        cn->stale = 0;
        cn->stale_vals = 0;
    }
}
__global__ __launch_bounds__(1024) void k_scan_counts(DevGrid a, int G) { scan_counts_body(a, G); }
__global__ __launch_bounds__(1024) void k_scan_counts_b(const DevGrid* grids, int G) {
    scan_counts_body(grids[blockIdx.x], G);
}

// ----------------------------------------------------- placement and split --
// kdgrid.cu:328-477 for one selected leaf: best candidate among
// candidatesNum uniform points of the box (max over candidates of the min
// Euclidean distance to the leaf's samples, earliest wins ties), appended at
// sample index k, the draws starting from LCG state s; a full leaf is split
// at the median of its 5 samples along the largest box axis, keeping the left
// half in place and the right half at leaf index m.
template <int D>
__device__ __forceinline__ void place_leaf(const DevGrid& a, int i, int k, int m, unsigned cnt,
                                           int L, unsigned s) {
    constexpr int P = line_pts(D);
    // the leaf's line: slots, box and its first points, one 128-byte read
    float f[kLineFloats];
    load_line(a, i, f);
    // the overflow row (points past the line, the slots beside them) only
    // when the leaf holds those points
    constexpr int XS = xpts_stride(D) > 0 ? xpts_stride(D) : 4;
    float x[XS];
#pragma unroll
    for (int t = 0; t < XS; ++t) x[t] = 0.0f;
    if constexpr (P < 4) {
        if (cnt > (unsigned)P) {
            const float4* xr = reinterpret_cast<const float4*>(a.xpts + (size_t)i * xpts_stride(D));
#pragma unroll
            for (int t = 0; t < XS / 4; ++t) {
                const float4 v = xr[t];
                x[4 * t] = v.x;
                x[4 * t + 1] = v.y;
                x[4 * t + 2] = v.z;
                x[4 * t + 3] = v.w;
            }
        }
    }
    // slots past the leaf's count are stale: -1
    int ids[4];
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        const int at = slot_at(D, q);
        const int v = at >= 0 ? __float_as_int(f[at]) : __float_as_int(x[-1 - at]);
        ids[q] = q < (int)cnt ? v : -1;
    }
    float mn[D], mx[D], diag[D];
#pragma unroll
    for (int d = 0; d < D; ++d) {
        mn[d] = f[4 + d];
        mx[d] = f[4 + D + d];
        diag[d] = mx[d] - mn[d];
    }
    // the points by slot (the floats of empty slots are replaced, never used
    // in arithmetic); those past the line only when the leaf holds them
    float p[4][D];
#pragma unroll
    for (int q = 0; q < 4; ++q)
#pragma unroll
        for (int d = 0; d < D; ++d)
            p[q][d] = ids[q] < 0 ? 0.0f : (q < P ? f[line_pt(D, q) + d] : x[xpt_at(D, q) + d]);
    float maxDist = -1.0f;
    float best[D];
#pragma unroll
    for (int d = 0; d < D; ++d) best[d] = 0.0f;
    for (int j = 0; j < a.cand; ++j) {
        float c[D];
#pragma unroll
        for (int d = 0; d < D; ++d) c[d] = rnd(s);
#pragma unroll
        for (int d = 0; d < D; ++d) c[d] = mn[d] + c[d] * diag[d];
        // min over samples of sqrt(d2) = sqrt(min over samples of d2), since
        // sqrt is monotone and correctly rounded; one sqrt per candidate.
        float minD2 = __int_as_float(0x7f800000);
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            if (ids[q] >= 0) {
                float acc = 0.0f;
#pragma unroll
                for (int d = 0; d < D; ++d) {
                    float t = c[d] - p[q][d];
                    acc += t * t;
                }
                if (minD2 > acc) minD2 = acc;
            }
        }
        const float minDist = minD2 < __int_as_float(0x7f800000) ? __builtin_sqrtf(minD2)
                                                                 : 3.402823466e+38f;
        if (maxDist < minDist) {
            maxDist = minDist;
#pragma unroll
            for (int d = 0; d < D; ++d) best[d] = c[d];
        }
    }
    float* out = a.coords + (size_t)k * D;
#pragma unroll
    for (int d = 0; d < D; ++d) out[d] = best[d];

    if (cnt < 4u) {
        // kdgrid.cu:393-396: insert into the first free slot; the error is kept
        store_point_slot<D>(a, i, (int)cnt, best, k);
    } else {
        // kdgrid.cu:399-471: split.  LargestAxis scans from the last axis down
        // with a strict '<' (aabb.h:34-45).
        int axis = -1;
        float ext = -3.402823466e+38f;
#pragma unroll
        for (int d = D - 1; d >= 0; --d)
            if (ext < diag[d]) {
                ext = diag[d];
                axis = d;
            }
        float key[5];
        int v[5] = {ids[0], ids[1], ids[2], ids[3], k};
        int perm[5] = {0, 1, 2, 3, 4};  // which of p[0..3], best sits at each position
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            float kv = p[q][0];
#pragma unroll
            for (int d = 1; d < D; ++d)
                if (d == axis) kv = p[q][d];
            key[q] = kv;
        }
        {
            float kv = best[0];
#pragma unroll
            for (int d = 1; d < D; ++d)
                if (d == axis) kv = best[d];
            key[4] = kv;
        }
        // stable bubble sort, strict '>' swaps (kdgrid.cu:409-417)
#pragma unroll
        for (int r = 0; r < 4; ++r) {
#pragma unroll
            for (int q = 0; q < 4 - r; ++q) {
                if (key[q] > key[q + 1]) {
                    float tk = key[q];
                    key[q] = key[q + 1];
                    key[q + 1] = tk;
                    int tv = v[q];
                    v[q] = v[q + 1];
                    v[q + 1] = tv;
                    tv = perm[q];
                    perm[q] = perm[q + 1];
                    perm[q + 1] = tv;
                }
            }
        }
        const float pos = 0.5f * (key[1] + key[2]);
        const int E = ~D;
        float cmn[D], cmx[D];
#pragma unroll
        for (int d = 0; d < D; ++d) {
            cmn[d] = d == axis ? pos : mn[d];
            cmx[d] = d == axis ? pos : mx[d];
        }
        // the points by sorted position: 0, 1 -> left child, 2..4 -> right
        float sp[5][D];
#pragma unroll
        for (int j = 0; j < 5; ++j)
#pragma unroll
            for (int d = 0; d < D; ++d) {
                float x = best[d];
#pragma unroll
                for (int q = 0; q < 4; ++q)
                    if (perm[j] == q) x = p[q][d];
                sp[j][d] = x;
            }
        // the left child in the parent's line (slots, box, two points: whole
        // float4s, the rest stays, past the child's count of 2), the right child's
        // whole line, and its third point in its overflow row when the line
        // holds two (D = 6)
        store_line_head2<D>(a, i, make_int4(v[0], v[1], E, E), mn, cmx, sp[0], sp[1]);
        store_line<D>(a, m, make_int4(v[2], v[3], v[4], E), cmn, mx, &sp[2], 3);
        if constexpr (P < 3) store_point_slot<D>(a, m, 2, sp[4], v[4]);  // D = 6: overflow row
        // what the next error pass needs of both children, coalesced by
        // split rank: their sample indices and box volumes
        a.sdat[2 * (m - L)] = make_int4(v[0], v[1], v[2], v[3]);
        a.sdat[2 * (m - L) + 1] = make_int4(v[4], __float_as_int(box_volume<D>(mn, cmx)),
                                            __float_as_int(box_volume<D>(cmn, mx)), 0);
        if (a.sext) {  // learned mode: the children's extents for their features
            float ex[6];
#pragma unroll
            for (int d = 0; d < 6; ++d) ex[d] = d < D ? (d == axis ? cmx[d] - mn[d] : diag[d]) : 0.0f;
            a.sext[2 * (m - L)] = make_int4(__float_as_int(ex[0]), __float_as_int(ex[1]),
                                            __float_as_int(ex[2]), __float_as_int(ex[3]));
            float right = 0.0f;
#pragma unroll
            for (int d = 0; d < D; ++d)
                if (d == axis) right = mx[d] - cmn[d];
            a.sext[2 * (m - L) + 1] = make_int4(__float_as_int(ex[4]), __float_as_int(ex[5]),
                                                __float_as_int(right), axis);
        }
        if (!a.in_loop) {  // in the loop the next error pass overwrites both
            a.err[i] = -1.0f;
            a.err[m] = -1.0f;
        }
        a.seed[m] = 0u;
        a.flags[m] = 3;
        a.split_list[m - L] = i;  // the split leaves in leaf order (right child m)
    }
    // the leaf's own state (already jumped) and count were written by the
    // selection and compaction kernels
}

// Compaction and placement in one kernel.  Each workgroup ranks the selected
// leaves of its chunk (the selection's partition) in leaf order by wave
// ballots: a leaf of global rank r (r = the rank among all selected leaves)
// gets the new sample N_old + r, a split leaf of split rank q the right child
// L_old + q.  The entries (leaf | count << 30, right child or -1; leaf
// indices are below 2^30, a split leaf's count is 4) and the LCG states
// they start from queue up in LDS and are placed a round of 256 at a time,
// one per thread.  Every leaf's count and LCG state are rewritten in the
// same sweep, as they stand after the placement, in full coalesced lines.
template <int D>
__device__ __forceinline__ void compact_place_body(const DevGrid& a, int blk, int nblk) {
    constexpr int U = 4;                   // tiles whose flags and states are loaded together
    constexpr int QCAP = U * 256 + 256;    // one step's entries plus an unplaced remainder
    __shared__ int lds[8];
    __shared__ int2 qe[QCAP];
    __shared__ unsigned qs[QCAP];
    if (a.cnt->done) return;
    const int L = a.cnt->n_leaves_old;
    const int N0 = a.cnt->n_samples_old;
    int beg, end;
    chunk_range(L, nblk, blk, beg, end);
    int2 base = a.part_off[blk];
    int jq = base.x;  // global rank of the entry at the front of the queue
    int qn = 0;       // queued entries
    auto place_round = [&](int p0, int count) {
        const int p = p0 + (int)threadIdx.x;
        if ((int)threadIdx.x < count) {
            const int2 q = qe[p];
            const unsigned cnt = q.y >= 0 ? 4u : ((unsigned)q.x >> 30);
            place_leaf<D>(a, q.x & 0x3fffffff, N0 + jq + p, q.y, cnt, L, qs[p]);
        }
    };
    for (int t00 = beg; t00 < end; t00 += U * 256) {
        unsigned fv[U], sv[U];
#pragma unroll
        for (int r = 0; r < U; ++r) {
            const int i = t00 + r * 256 + threadIdx.x;
            fv[r] = i < end ? a.flags[i] : 0u;
            sv[r] = i < end ? a.seed[i] : 0u;
        }
#pragma unroll
        for (int r = 0; r < U; ++r) {
            const int t0 = t00 + r * 256;
            if (t0 >= end) break;  // uniform across the workgroup
            const int i = t0 + threadIdx.x;
            const unsigned f = fv[r];
            const bool sel = (f & 8u) != 0u, split = (f & 16u) != 0u;
            const unsigned cnt = f & 7u;
            int rs, rl, ts, tl;
            block_rank2(sel, split, lds, rs, rl, ts, tl);
            // the state after the selection's draw (kdgrid.cu:323-326): the
            // placement starts from it, and a selected leaf's stored state is
            // the one after the candidatesNum * D candidate draws (the LCG's
            // affine jump-ahead)
            unsigned s1 = sv[r];
            if (s1 == 0u) s1 = tea4((unsigned)i, 0u);
            s1 = 1664525u * s1 + 1013904223u;
            if (sel) {
                qe[qn + rs] = make_int2(i | (int)((cnt & 3u) << 30), split ? L + base.y + rl : -1);
                qs[qn + rs] = s1;
            }
            if (i < end) {
                a.flags[i] = (unsigned char)(split ? 2u : (sel ? cnt + 1u : cnt));
                a.seed[i] = sel ? a.jmul * s1 + a.jadd : s1;
            }
            qn += ts;
            base.y += tl;
        }
        __syncthreads();
        int placed = 0;
        for (; qn - placed >= 256; placed += 256) place_round(placed, 256);
        if (placed > 0) {  // the remainder (< 256) to the front of the queue
            const int rem = qn - placed;
            int2 e = make_int2(0, 0);
            unsigned st = 0u;
            if ((int)threadIdx.x < rem) {
                e = qe[placed + threadIdx.x];
                st = qs[placed + threadIdx.x];
            }
            __syncthreads();
            if ((int)threadIdx.x < rem) {
                qe[threadIdx.x] = e;
                qs[threadIdx.x] = st;
            }
            jq += placed;
            qn = rem;
        }
        __syncthreads();
    }
    place_round(0, qn);
}

template <int D>
__global__ __launch_bounds__(256) void k_compact_place(DevGrid a) {
    compact_place_body<D>(a, blockIdx.x, gridDim.x);
}
template <int D>
__global__ __launch_bounds__(256) void k_compact_place_b(const DevGrid* grids) {
    compact_place_body<D>(grids[blockIdx.y], blockIdx.x, gridDim.x);
}

// After the device-side loop: when it stopped at its limit, no error pass
// followed the last placement, so its split children get the stale marker
// (error -1, kdgrid.cu:466-469) the placement left out.
__device__ __forceinline__ void mark_last_children_body(const DevGrid& a, int blk, int nblk) {
    const DevCounters* c = a.cnt;
    if (c->status != 0 || !loop_limit_reached(c)) return;
    const int nn = c->new_nodes, L_old = c->n_leaves_old;
    for (int t = blk * 256 + threadIdx.x; t < nn; t += nblk * 256) {
        a.err[a.split_list[t]] = -1.0f;
        a.err[L_old + t] = -1.0f;
    }
}
__global__ __launch_bounds__(256) void k_mark_last_children(DevGrid a) {
    mark_last_children_body(a, blockIdx.x, gridDim.x);
}
__global__ __launch_bounds__(256) void k_mark_last_children_b(const DevGrid* grids) {
    mark_last_children_body(grids[blockIdx.y], blockIdx.x, gridDim.x);
}

// The leaves' points in their lines and overflow rows from the sample
// coordinates (after a state upload).
template <int D>
__global__ __launch_bounds__(256) void k_build_pts(DevGrid a, int L) {
    for (int i = blockIdx.x * 256 + threadIdx.x; i < L; i += gridDim.x * 256) {
        const int4 id = load_slots<D>(a, i, leaf_count(a, i));
        const int ids[4] = {id.x, id.y, id.z, id.w};
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            if (ids[q] >= 0 && ids[q] < a.cap_s) {
                float* dst = point_addr<D>(a, i, q);
#pragma unroll
                for (int d = 0; d < D; ++d) dst[d] = a.coords[(size_t)ids[q] * D + d];
            }
        }
    }
}

// --------------------------------------------------------------- fills --
__global__ void k_fill_i32(int* p, int v, long long n) {
    for (long long i = blockIdx.x * 256ll + threadIdx.x; i < n; i += (long long)gridDim.x * 256) p[i] = v;
}
__global__ void k_fill_f32(float* p, float v, long long n) {
    for (long long i = blockIdx.x * 256ll + threadIdx.x; i < n; i += (long long)gridDim.x * 256) p[i] = v;
}

// ------------------------------------------------------------ launchers --
static int grid_for(long long n, int cap_blocks = 8192) {
    long long b = (n + 255) / 256;
    if (b < 1) b = 1;
    if (b > cap_blocks) b = cap_blocks;
    return (int)b;
}

#define MDAS_DISPATCH_D(D_, STMT)                      \
    switch (D_) {                                      \
    case 3: { constexpr int DD = 3; STMT; } break;     \
    case 4: { constexpr int DD = 4; STMT; } break;     \
    case 5: { constexpr int DD = 5; STMT; } break;     \
    case 6: { constexpr int DD = 6; STMT; } break;     \
    default: return hipErrorInvalidValue;              \
    }

hipError_t launch_initial(Grid& g, int L0, hipStream_t st) {
    const int blocks = grid_for(L0, 4096);
    MDAS_DISPATCH_D(g.D, hipLaunchKernelGGL(k_initial<DD>, dim3(blocks), dim3(256), 0, st, g.dev, L0));
    hipLaunchKernelGGL(k_initial_counters, dim3(1), dim3(1), 0, st, g.dev.cnt, L0);
    return hipGetLastError();
}

hipError_t launch_reset_errors(Grid& g, int standalone, hipStream_t st) {
    hipLaunchKernelGGL(k_reset_errors, dim3(1), dim3(1), 0, st, g.dev.cnt, standalone);
    return hipGetLastError();
}

hipError_t launch_set_limit(Grid& g, int limit, hipStream_t st) {
    // the capacity guard applies when the limit is the capacity (a target at
    // or beyond max_samples); below it a pass that would overflow is an error
    hipLaunchKernelGGL(k_set_limit, dim3(1), dim3(1), 0, st, g.dev.cnt, limit,
                       limit >= g.cap_s ? g.cap_s : 0);
    return hipGetLastError();
}

hipError_t launch_arm(Grid& g, hipStream_t st) {
    hipLaunchKernelGGL(k_arm, dim3(1), dim3(1), 0, st, g.dev.cnt);
    return hipGetLastError();
}

hipError_t launch_errors(Grid& g, hipStream_t st) {
    MDAS_DISPATCH_D(g.D, hipLaunchKernelGGL(k_errors<DD>, dim3(g.err_blocks), dim3(256), 0, st, g.dev));
    hipLaunchKernelGGL(k_err_reduce, dim3(1), dim3(1024), 0, st, g.dev, g.err_blocks, 0);
    return hipGetLastError();
}

hipError_t launch_errors_stale(Grid& g, hipStream_t st) {
    MDAS_DISPATCH_D(g.D, hipLaunchKernelGGL(k_errors_stale<DD>, dim3(g.err_blocks), dim3(256), 0, st, g.dev));
    hipLaunchKernelGGL(k_err_reduce, dim3(1), dim3(1024), 0, st, g.dev, g.err_blocks, 1);
    return hipGetLastError();
}

hipError_t launch_select(Grid& g, hipStream_t st) {
    const int cpw = select_cpw(g.chunk_blocks);
    hipLaunchKernelGGL(k_select, dim3(g.chunk_blocks / cpw), dim3(256), 0, st, g.dev, cpw);
    return hipGetLastError();
}

hipError_t launch_scan(Grid& g, hipStream_t st) {
    hipLaunchKernelGGL(k_scan_counts, dim3(1), dim3(1024), 0, st, g.dev, g.chunk_blocks);
    return hipGetLastError();
}

hipError_t launch_place(Grid& g, hipStream_t st) {
    MDAS_DISPATCH_D(g.D, hipLaunchKernelGGL(k_compact_place<DD>, dim3(g.chunk_blocks), dim3(256), 0, st,
                                            g.dev));
    return hipGetLastError();
}

hipError_t launch_adaptive(Grid& g, hipStream_t st) {
    hipError_t e = launch_select(g, st);
    if (e == hipSuccess) e = launch_scan(g, st);
    if (e == hipSuccess) e = launch_place(g, st);
    return e;
}

hipError_t launch_mark_last_children(Grid& g, hipStream_t st) {
    hipLaunchKernelGGL(k_mark_last_children, dim3(1024), dim3(256), 0, st, g.dev);
    return hipGetLastError();
}

hipError_t launch_mark_last_children_b(const DevGrid* grids, int n, hipStream_t st) {
    hipLaunchKernelGGL(k_mark_last_children_b, dim3(256, n), dim3(256), 0, st, grids);
    return hipGetLastError();
}

hipError_t launch_build_pts(Grid& g, int n_leaves, hipStream_t st) {
    if (n_leaves <= 0) return hipSuccess;
    const int blocks = grid_for(n_leaves, 4096);
    MDAS_DISPATCH_D(g.D, hipLaunchKernelGGL(k_build_pts<DD>, dim3(blocks), dim3(256), 0, st, g.dev, n_leaves));
    return hipGetLastError();
}

hipError_t launch_fill_i32(int* p, int v, long long n, hipStream_t st) {
    hipLaunchKernelGGL(k_fill_i32, dim3(grid_for(n)), dim3(256), 0, st, p, v, n);
    return hipGetLastError();
}

hipError_t launch_fill_f32(float* p, float v, long long n, hipStream_t st) {
    hipLaunchKernelGGL(k_fill_f32, dim3(grid_for(n)), dim3(256), 0, st, p, v, n);
    return hipGetLastError();
}

// ------------------------------------------------------ batched launchers --
// `grids` is a device array of n DevGrid of one dimension D; every chunked
// kernel gives each image `blocks` workgroups (<= kChunkBlocks).
hipError_t launch_initial_b(int D, const DevGrid* grids, int n, int blocks, const int* limits,
                            hipStream_t st) {
    MDAS_DISPATCH_D(D, hipLaunchKernelGGL(k_initial_b<DD>, dim3(blocks, n), dim3(256), 0, st, grids));
    hipLaunchKernelGGL(k_initial_counters_b, dim3((n + 63) / 64), dim3(64), 0, st, grids, n, limits);
    return hipGetLastError();
}

hipError_t launch_pass_b(int D, const DevGrid* grids, int n, int blocks, hipStream_t st,
                         const LearnDev* learns) {
    const int eb = blocks < 1024 ? blocks : 1024;
    if (learns) {
        const hipError_t e = launch_errors_learned_b(D, grids, learns, n, blocks, st);
        if (e != hipSuccess) return e;
    } else {
        MDAS_DISPATCH_D(D, hipLaunchKernelGGL(k_errors_stale_b<DD>, dim3(eb, n), dim3(256), 0, st, grids));
        hipLaunchKernelGGL(k_err_reduce_b, dim3(n), dim3(1024), 0, st, grids, eb);
    }
    const int cpw = select_cpw(blocks);
    hipLaunchKernelGGL(k_select_b, dim3(blocks / cpw, n), dim3(256), 0, st, grids, cpw);
    hipLaunchKernelGGL(k_scan_counts_b, dim3(n), dim3(1024), 0, st, grids, blocks);
    MDAS_DISPATCH_D(D, hipLaunchKernelGGL(k_compact_place_b<DD>, dim3(blocks, n), dim3(256), 0, st, grids));
    return hipGetLastError();
}

}  // namespace mdas
