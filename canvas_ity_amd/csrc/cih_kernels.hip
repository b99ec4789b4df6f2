// cih_kernels.hip -- the MI355X (gfx950) kernels of the canvas pipeline.
//
// Per frame (a list of draws over one or more canvases):
//   k_reset        clears the arena counters (a kernel, so frames can be
//                  captured in HIP graphs without memset nodes)
//   k_geometry     one wave per draw: flatten (one thread per cubic),
//                  dash and stroke (one thread per subpath) -> closed polygons
//   k_runs         one workgroup per draw and pass (main, and shadow when the
//                  frame has shadows): clip (hpp:2205-2238), signed-area runs
//                  (hpp:2109-2170) bucketed per pixel row; the shadow pass
//                  also sizes its image; each pass appends its rows to the
//                  coalesce work list
//   k_coalesce     a fixed grid of waves over the work list's rows: each
//                  row's runs sorted, coalesced and summed along the row with
//                  the reference's float arithmetic (hpp:2239-2252) ->
//                  coverage steps (x, running sum) in x order
//   k_clip         clip draws: the path's coverage merged with the parent
//                  mask into a new mask, one thread per canvas row
//                  (hpp:3057-3099)
//   k_shadow_*     shadow alpha raster, then the extended-box blur
//                  (hpp:2453-2503) with one thread per row / per column
//   k_fine         one wave per (canvas, row, 256-pixel segment): the fp32
//                  pixels stay in registers while every draw of the canvas is
//                  applied in painter's order -- coverage from the row's steps
//                  (a 64-way search, then LDS for steps inside the segment),
//                  paint (hpp:2265-2377) and composite (hpp:2576-2592);
//                  k_fine<true> also applies clip-mask visibility
//   k_readback     get_image_data's unpremultiply/sRGB/dither (hpp:3363-3380)
// None of this is a dense contraction: the pipeline is VALU/LDS and HBM work.
#include <hip/hip_runtime.h>

#include "cih_device.hpp"
#include "cih_geometry.hpp"
#include "cih_color.hpp"
#include "cih_blur.hpp"
#include "cih_abi.h"

namespace cih {

#define GEO_THREADS 256 // k_runs: one workgroup of 4 waves per draw and pass
// k_geometry: one wave per draw.  Most draws have a few cubics and subpaths
// (one thread each), so a wave keeps as many lanes busy as four would, and a
// CU holds four times as many draws (its LDS use is a few words).
#define GEOM_THREADS 64
#define FINE_THREADS 64 // one wave: segment_sums relies on it
#define FINE_PX 4
#define FINE_SEGMENT (FINE_THREADS * FINE_PX)
#define LDS_ROWS 4096

__constant__ float c_bayer[16] = {0.03125f, 0.53125f, 0.15625f, 0.65625f,
                                  0.78125f, 0.28125f, 0.90625f, 0.40625f,
                                  0.21875f, 0.71875f, 0.09375f, 0.59375f,
                                  0.96875f, 0.46875f, 0.84375f, 0.34375f};

static const float kThreshold = 1.0f / 8160.0f; // hpp:1289, 2429, 2573

// ---------------------------------------------------------------------------
// workgroup helpers

__device__ inline bool overflowed(const frame &F) {
    return __atomic_load_n(&F.counters[COUNTER_OVERFLOW], __ATOMIC_RELAXED) != 0;
}

// LDS ordering inside a workgroup of one wave: the wave issues its LDS
// operations in order, so once they have completed every lane sees what the
// others wrote.  Unlike __syncthreads, this leaves global loads in flight.
__device__ inline void wave_lds_sync() { asm volatile("s_waitcnt lgkmcnt(0)" ::: "memory"); }

// Bump allocation from an arena; returns ~0 (and flags overflow) when full.
__device__ inline uint64_t arena_alloc(const frame &F, int which, uint64_t n, uint64_t cap) {
    unsigned long long at = atomicAdd(&F.counters[which], (unsigned long long)n);
    if (at + n > cap) {
        atomicMax(&F.counters[COUNTER_OVERFLOW], 1ull);
        return ~0ull;
    }
    return at;
}

// Exclusive scan of one value per thread over the workgroup (of up to 16 waves).
__device__ inline uint32_t block_exclusive(uint32_t v, uint32_t *lds, uint32_t &total) {
    int t = threadIdx.x;
    int lane = t & 63, wave = t >> 6;
    uint32_t incl = v;
    for (int d = 1; d < 64; d <<= 1) {
        uint32_t n = __shfl_up(incl, d, 64);
        if (lane >= d)
            incl += n;
    }
    if (lane == 63)
        lds[wave] = incl;
    __syncthreads();
    uint32_t before = 0;
    uint32_t sum = 0;
    for (int w = 0; w < (int)(blockDim.x >> 6); ++w) {
        if (w < wave)
            before += lds[w];
        sum += lds[w];
    }
    __syncthreads();
    total = sum;
    return before + incl - v;
}

// In-place exclusive scan of a[0..n) in global memory; a[n] = total.
__device__ inline uint32_t scan_array(uint32_t *a, uint32_t n, uint32_t *lds) {
    uint32_t carry = 0;
    for (uint32_t base = 0; base < n; base += blockDim.x) {
        uint32_t i = base + threadIdx.x;
        uint32_t v = i < n ? a[i] : 0;
        uint32_t total;
        uint32_t e = block_exclusive(v, lds, total);
        if (i < n)
            a[i] = carry + e;
        carry += total;
    }
    if (threadIdx.x == 0)
        a[n] = carry;
    __syncthreads();
    return carry;
}

// Thread 0 allocates, everyone receives the result.
__device__ inline uint64_t block_alloc(const frame &F, int which, uint64_t n, uint64_t cap,
                                       uint64_t *slot) {
    if (threadIdx.x == 0)
        *slot = n ? arena_alloc(F, which, n, cap) : 0;
    __syncthreads();
    uint64_t r = *slot;
    __syncthreads();
    return r;
}

// Largest s in [0, n) with prefix[s] <= g (prefix ascending, prefix[0] = 0).
__device__ inline uint32_t find_prefix(const uint32_t *prefix, uint32_t n, uint32_t g) {
    uint32_t lo = 0, hi = n; // answer in [lo, hi)
    while (hi - lo > 1) {
        uint32_t mid = (lo + hi) >> 1;
        if (prefix[mid] <= g)
            lo = mid;
        else
            hi = mid;
    }
    return lo;
}

// ---------------------------------------------------------------------------
// sinks

struct count_points {
    uint32_t points, subpaths;
    __device__ count_points() : points(0), subpaths(0) {}
    __device__ void point(f2) { ++points; }
    __device__ void subpath(uint32_t, bool) { ++subpaths; }
    __device__ void close_last() {}
    __device__ void merge_last(uint32_t, uint32_t) { --subpaths; }
};

// Writes a thread's polylines into its own region [points, ...) and
// [subpaths, ...).  The newest subpath entry is held back until the next one
// starts (or finish()): dashing may still close it or merge it into the
// subpath's first dash (hpp:1916-1931), and the region was sized by a count
// pass that saw the merged result, so writing the entry early could land in
// the next thread's region.
struct write_points {
    f2 *pts;
    uint32_t *sub; // (count, closed) pairs
    uint32_t points, subpaths; // absolute indices
    bool pending = false;
    uint32_t pending_count = 0;
    bool pending_closed = false;
    __device__ void point(f2 p) { pts[points++] = p; }
    __device__ void commit() {
        if (pending) {
            sub[2 * (subpaths - 1)] = pending_count;
            sub[2 * (subpaths - 1) + 1] = pending_closed ? 1u : 0u;
            pending = false;
        }
    }
    __device__ void subpath(uint32_t c, bool closed) {
        commit();
        pending = true;
        pending_count = c;
        pending_closed = closed;
        ++subpaths;
    }
    __device__ void close_last() {
        if (pending)
            pending_closed = true;
        else
            sub[2 * (subpaths - 1) + 1] = 1u;
    }
    __device__ void reverse(uint32_t a, uint32_t b) { // [a, b)
        while (a + 1 < b) {
            f2 t = pts[a];
            pts[a] = pts[b - 1];
            pts[b - 1] = t;
            ++a, --b;
        }
    }
    // std::rotate of the last dash to the front of the subpath's first dash.
    __device__ void merge_last(uint32_t merge_point, uint32_t merge_subpath) {
        uint32_t c = pending ? pending_count : sub[2 * (subpaths - 1)];
        pending = false;
        uint32_t middle = points - c;
        reverse(merge_point, middle);
        reverse(middle, points);
        reverse(merge_point, points);
        sub[2 * merge_subpath] += c;
        --subpaths;
    }
    __device__ void finish() { commit(); }
};

struct write_flat {
    f2 *pts;
    uint32_t at;
    __device__ void point(f2 p) { pts[at++] = p; }
};

// ---------------------------------------------------------------------------
// k_geometry: path -> polygons for one draw per workgroup.

struct line_path {
    const f2 *pts;
    const uint32_t *sub; // (count, closed) pairs
    uint32_t npts, nsub;
};

__device__ line_path flatten(const frame &F, const draw_desc &D, bool stroking,
                             uint32_t *lds, uint64_t *slot) {
    line_path out = {0, 0, 0, 0};
    const f2 *in = (const f2 *)(F.payload + D.points_at);
    const uint32_t *in_sub = (const uint32_t *)(F.payload + D.subpaths_at);
    uint32_t nsub = D.nsubpaths;
    float angular = angular_limit(stroking, D.line_width);
    // a start point then 3 points per cubic, or (segments) 4 per cubic
    const bool segments = (D.flags & CIH_DRAW_FLAG_SEGMENTS) != 0;
    const uint32_t stride = segments ? 4u : 3u;
    // point and cubic prefixes over the subpaths
    uint64_t w = block_alloc(F, COUNTER_WORDS, 2 * (uint64_t)(nsub + 1), F.words_cap, slot);
    if (w == ~0ull)
        return out;
    uint32_t *sp_pts = F.words + w;
    uint32_t *sp_cub = sp_pts + nsub + 1;
    for (uint32_t s = threadIdx.x; s < nsub; s += blockDim.x) {
        uint32_t c = in_sub[2 * s];
        sp_pts[s] = c;
        sp_cub[s] = segments ? c / 4 : c >= 1 ? (c - 1) / 3 : 0;
    }
    __syncthreads();
    scan_array(sp_pts, nsub, lds);
    uint32_t cubics = scan_array(sp_cub, nsub, lds);
    uint64_t cw = block_alloc(F, COUNTER_WORDS, (uint64_t)cubics + 1, F.words_cap, slot);
    if (cw == ~0ull)
        return out;
    uint32_t *cnt = F.words + cw;
    for (uint32_t g = threadIdx.x; g < cubics; g += blockDim.x) {
        uint32_t s = find_prefix(sp_cub, nsub, g);
        uint32_t base = sp_pts[s] + stride * (g - sp_cub[s]);
        count_points c;
        add_bezier(c, in[base], in[base + 1], in[base + 2], in[base + 3], angular);
        cnt[g] = c.points;
    }
    __syncthreads();
    uint32_t total = scan_array(cnt, cubics, lds);
    uint64_t p = block_alloc(F, COUNTER_POINTS, (uint64_t)total + nsub, F.points_cap, slot);
    uint64_t q = block_alloc(F, COUNTER_WORDS, 2 * (uint64_t)nsub, F.words_cap, slot);
    if (p == ~0ull || q == ~0ull)
        return out;
    f2 *pts = (f2 *)F.points + p;
    uint32_t *sub = F.words + q;
    for (uint32_t s = threadIdx.x; s < nsub; s += blockDim.x) {
        uint32_t first = cnt[sp_cub[s]];
        uint32_t last = cnt[sp_cub[s + 1]];
        pts[first + s] = in[sp_pts[s]];
        sub[2 * s] = 1 + last - first;
        sub[2 * s + 1] = in_sub[2 * s + 1];
    }
    for (uint32_t g = threadIdx.x; g < cubics; g += blockDim.x) {
        uint32_t s = find_prefix(sp_cub, nsub, g);
        uint32_t base = sp_pts[s] + stride * (g - sp_cub[s]);
        write_flat wf = {pts, cnt[g] + s + 1};
        add_bezier(wf, in[base], in[base + 1], in[base + 2], in[base + 3], angular);
    }
    __syncthreads();
    out.pts = pts;
    out.sub = sub;
    out.npts = total + nsub;
    out.nsub = nsub;
    return out;
}

// Stroke expansion with one thread per join.  One thread per (subpath, side)
// runs stroke_walk, which only decides the kept points, and records each join
// as three point indices (a cap as two and ~0u); every join's points depend
// on those indices alone, so the joins are then counted and written in
// parallel, in the order the sequential walk emits them.
struct join_recorder {
    uint32_t *items;
    uint32_t n;
    __device__ void join(uint32_t prev, uint32_t at, uint32_t next) {
        items[3 * n] = prev;
        items[3 * n + 1] = at;
        items[3 * n + 2] = next;
        ++n;
    }
    __device__ void cap(uint32_t prev, uint32_t at) { join(prev, at, ~0u); }
};

template <class Sink>
__device__ void stroke_item(Sink &out, const uint32_t *item, const f2 *pts, const draw_desc &D) {
    if (item[2] == ~0u)
        stroke_cap(out, pts, item[0], item[1], D.line_width, D.line_cap, D.forward, D.inverse);
    else
        stroke_join(out, pts, item[0], item[1], item[2], D.line_width, D.miter_limit,
                    D.line_join, D.forward, D.inverse);
}

// Stroke outline of every subpath (hpp:2070-2100): per subpath, the forward
// side, then the backward side; an open subpath gives one closed loop, a
// closed one an outer and an inner loop.
__device__ line_path stroke_all(const frame &F, const draw_desc &D, line_path in,
                                uint32_t *lds, uint64_t *slot) {
    line_path out = {0, 0, 0, 0};
    uint32_t nsub = in.nsub;
    // at: point offset per subpath; room: item capacity offset per subpath;
    // items: item count per (subpath, side); ns: output subpaths per subpath
    uint64_t w = block_alloc(F, COUNTER_WORDS, 5 * (uint64_t)(nsub + 1), F.words_cap, slot);
    if (w == ~0ull)
        return out;
    uint32_t *at = F.words + w;
    uint32_t *room = at + nsub + 1;
    uint32_t *items = room + nsub + 1; // 2 * nsub + 1 entries
    uint32_t *ns = items + 2 * nsub + 1;
    for (uint32_t s = threadIdx.x; s < nsub; s += blockDim.x) {
        uint32_t count = in.sub[2 * s];
        at[s] = count;
        // a side makes at most one join per visit (< 2 * count visits) and a cap
        room[s] = count >= 2 ? 2 * (2 * count + 2) : 0;
        ns[s] = count >= 2 ? (in.sub[2 * s + 1] ? 2u : 1u) : 0u;
    }
    __syncthreads();
    scan_array(at, nsub, lds);
    uint32_t capacity = scan_array(room, nsub, lds);
    uint32_t nout = scan_array(ns, nsub, lds);
    uint64_t iw = block_alloc(F, COUNTER_WORDS, 3 * (uint64_t)capacity, F.words_cap, slot);
    if (iw == ~0ull)
        return out;
    uint32_t *rec = F.words + iw;
    for (uint32_t t = threadIdx.x; t < 2 * nsub; t += blockDim.x) {
        uint32_t s = t >> 1, side = t & 1;
        uint32_t count = in.sub[2 * s];
        join_recorder r = {rec + 3 * (room[s] + side * (2 * count + 2)), 0};
        if (count >= 2)
            stroke_walk(r, in.pts + at[s], side ? count - 1 : 0, side ? 0 : count - 1,
                        in.sub[2 * s + 1] != 0, D.inverse);
        items[t] = r.n;
    }
    __syncthreads();
    uint32_t total = scan_array(items, 2 * nsub, lds);
    uint64_t cw = block_alloc(F, COUNTER_WORDS, (uint64_t)total + 1, F.words_cap, slot);
    if (cw == ~0ull)
        return out;
    uint32_t *cnt = F.words + cw;
    for (uint32_t g = threadIdx.x; g < total; g += blockDim.x) {
        uint32_t t = find_prefix(items, 2 * nsub, g);
        uint32_t s = t >> 1, side = t & 1;
        uint32_t count = in.sub[2 * s];
        const uint32_t *item = rec + 3 * (room[s] + side * (2 * count + 2) + (g - items[t]));
        count_points c;
        stroke_item(c, item, in.pts + at[s], D);
        cnt[g] = c.points;
    }
    __syncthreads();
    uint32_t tp = scan_array(cnt, total, lds);
    uint64_t p = block_alloc(F, COUNTER_POINTS, tp, F.points_cap, slot);
    uint64_t q = block_alloc(F, COUNTER_WORDS, 2 * (uint64_t)nout, F.words_cap, slot);
    if (p == ~0ull || q == ~0ull)
        return out;
    f2 *pts = (f2 *)F.points + p;
    uint32_t *sub = F.words + q;
    for (uint32_t g = threadIdx.x; g < total; g += blockDim.x) {
        uint32_t t = find_prefix(items, 2 * nsub, g);
        uint32_t s = t >> 1, side = t & 1;
        uint32_t count = in.sub[2 * s];
        const uint32_t *item = rec + 3 * (room[s] + side * (2 * count + 2) + (g - items[t]));
        write_flat wf = {pts, cnt[g]};
        stroke_item(wf, item, in.pts + at[s], D);
    }
    for (uint32_t s = threadIdx.x; s < nsub; s += blockDim.x) {
        if (ns[s + 1] == ns[s])
            continue;
        uint32_t a = cnt[items[2 * s]], b = cnt[items[2 * s + 1]], e = cnt[items[2 * s + 2]];
        uint32_t *o = sub + 2 * ns[s];
        if (in.sub[2 * s + 1]) {
            o[0] = b - a;
            o[1] = 1u;
            o[2] = e - b;
            o[3] = 1u;
        } else {
            o[0] = e - a;
            o[1] = 1u;
        }
    }
    __syncthreads();
    out.pts = pts;
    out.sub = sub;
    out.npts = tp;
    out.nsub = nout;
    return out;
}

// Dash every subpath of `in` (hpp:1858-1934), one thread each.
__device__ line_path dash_all(const frame &F, const draw_desc &D, line_path in, uint32_t *lds,
                              uint64_t *slot) {
    line_path out = {0, 0, 0, 0};
    uint32_t nsub = in.nsub;
    uint64_t w = block_alloc(F, COUNTER_WORDS, 3 * (uint64_t)(nsub + 1), F.words_cap, slot);
    if (w == ~0ull)
        return out;
    uint32_t *at = F.words + w; // point offset of each input subpath
    uint32_t *np = at + nsub + 1;
    uint32_t *ns = np + nsub + 1;
    const float *dash = (const float *)(F.payload + D.dash_at);
    dash_setup setup = dash_prepare(dash, (int)D.ndash, D.dash_offset);
    for (uint32_t s = threadIdx.x; s < nsub; s += blockDim.x)
        at[s] = in.sub[2 * s];
    __syncthreads();
    scan_array(at, nsub, lds);
    for (uint32_t s = threadIdx.x; s < nsub; s += blockDim.x) {
        count_points c;
        const f2 *pts = in.pts + at[s];
        uint32_t count = in.sub[2 * s];
        bool closed = in.sub[2 * s + 1] != 0;
        dash_subpath(c, pts, count, closed, dash, (int)D.ndash, setup, D.inverse);
        np[s] = c.points;
        ns[s] = c.subpaths;
    }
    __syncthreads();
    uint32_t tp = scan_array(np, nsub, lds);
    uint32_t ts = scan_array(ns, nsub, lds);
    uint64_t p = block_alloc(F, COUNTER_POINTS, tp, F.points_cap, slot);
    uint64_t q = block_alloc(F, COUNTER_WORDS, 2 * (uint64_t)ts, F.words_cap, slot);
    if (p == ~0ull || q == ~0ull)
        return out;
    f2 *pts = (f2 *)F.points + p;
    uint32_t *sub = F.words + q;
    for (uint32_t s = threadIdx.x; s < nsub; s += blockDim.x) {
        write_points wp = {pts, sub, np[s], ns[s]};
        const f2 *src = in.pts + at[s];
        uint32_t count = in.sub[2 * s];
        bool closed = in.sub[2 * s + 1] != 0;
        dash_subpath(wp, src, count, closed, dash, (int)D.ndash, setup, D.inverse);
        wp.finish();
    }
    __syncthreads();
    out.pts = pts;
    out.sub = sub;
    out.npts = tp;
    out.nsub = ts;
    return out;
}

__global__ __launch_bounds__(GEOM_THREADS) void k_geometry(frame F) {
    __shared__ uint32_t lds[16];
    __shared__ uint64_t slot[1];
    int d = blockIdx.x;
    if (d >= F.ndraws || overflowed(F))
        return;
    const draw_desc &D = F.draws[d];
    line_path lp;
    bool stroke = D.kind == CIH_DRAW_STROKE_PATH || D.kind == CIH_DRAW_STROKE_LINES;
    draw_state &S = F.state[d];
    if (D.kind == CIH_DRAW_FILL_PATH || D.kind == CIH_DRAW_STROKE_PATH ||
        D.kind == CIH_DRAW_CLIP) {
        lp = flatten(F, D, stroke, lds, slot);
        if (!lp.pts && D.nsubpaths)
            return;
        if (threadIdx.x == 0) {
            S.flat_points = (uint64_t)(uintptr_t)lp.pts;
            S.flat_counts = (uint64_t)(uintptr_t)lp.sub;
            S.flat_npoints = lp.npts;
            S.flat_nsub = lp.nsub;
        }
    } else {
        lp.pts = (const f2 *)(F.payload + D.points_at);
        lp.sub = (const uint32_t *)(F.payload + D.subpaths_at);
        lp.npts = D.npoints;
        lp.nsub = D.nsubpaths;
    }
    if (stroke) {
        if (D.ndash) {
            lp = dash_all(F, D, lp, lds, slot);
            if (!lp.pts)
                return;
            if (threadIdx.x == 0) {
                S.dash_points = (uint64_t)(uintptr_t)lp.pts;
                S.dash_counts = (uint64_t)(uintptr_t)lp.sub;
                S.dash_npoints = lp.npts;
                S.dash_nsub = lp.nsub;
            }
        }
        lp = stroke_all(F, D, lp, lds, slot);
        if (!lp.pts)
            return;
    }
    if (threadIdx.x == 0) {
        S.poly_points = (uint64_t)(uintptr_t)lp.pts;
        S.poly_counts = (uint64_t)(uintptr_t)lp.sub;
        S.poly_npoints = lp.npts;
        S.poly_nsub = lp.nsub;
        atomicAdd(&F.counters[COUNTER_GEOMETRY_DONE], 1ull);
    }
}


// ---------------------------------------------------------------------------
// k_runs: polygons -> clipped edges -> runs bucketed by pixel row.
//
// Sinks drop runs outside the row range or the 16-bit x range, which only
// malformed (non-finite) geometry can produce; counting and writing apply the
// same test so bucket sizes always match.

// Writes one (edge, row) item's runs contiguously and keeps the statistics
// the shadow bounds need.  Runs with x beyond 16 bits (only an extended
// canvas wider than 65535 pixels makes them) are dropped, in the count too.
struct piece_writer {
    struct run *out;
    uint32_t at;
    int minx, maxx;
    int nzminx, nzmaxx, nzminy, nzmaxy;
    unsigned long long minkey;
    __device__ void run(int x, int yy, float delta) {
        if (x < 0 || x > 0xffff)
            return;
        struct run r;
        r.x = x;
        r.delta = delta;
        out[at++] = r;
        minx = min(minx, x);
        maxx = max(maxx, x);
        if (delta != 0.0f) {
            nzminx = min(nzminx, x);
            nzmaxx = max(nzmaxx, x);
            nzminy = min(nzminy, yy);
            nzmaxy = max(nzmaxy, yy);
        }
        // sort key of the reference's run order (y, x, |delta|), hpp:2172-2181
        unsigned long long key = ((unsigned long long)(uint32_t)yy << 48) |
                                 ((unsigned long long)(uint32_t)x << 32) |
                                 (unsigned long long)__float_as_uint(fabsf(delta));
        minkey = key < minkey ? key : minkey;
    }
};

// Runs of row k of an edge that fit 16-bit x: the row's runs sit at the
// consecutive pixels first_x .. first_x + steps + 1.
__device__ inline int item_runs(const edge_rows &E, int k) {
    edge_row r = edge_row_at(E, k);
    int first_x = (int)(E.pixel.x + (float)r.first);
    return min(r.steps + 2, max(0, 0x10000 - first_x));
}

// Endpoints of edge e of the draw's clipped polygons (with the pass offset).
struct edge_source {
    const f2 *pts, *clip;
    const uint32_t *sub, *at, *need, *edges;
    uint32_t nsub;
    f2 offset;
    __device__ void get(uint32_t e, f2 &from, f2 &to) const {
        uint32_t s = find_prefix(edges, nsub, e);
        uint32_t m = edges[s + 1] - edges[s];
        uint32_t k = e - edges[s];
        if (need[s + 1] != need[s]) {
            const f2 *poly = clip + need[s] + 8 * sub[2 * s];
            from = poly[(k ? k : m) - 1];
            to = poly[k];
        } else {
            const f2 *poly = pts + at[s];
            from = add(offset, poly[(k ? k : m) - 1]);
            to = add(offset, poly[k]);
        }
    }
};

// Appends a (draw, pass) with `rows` row buckets to the coalesce work list.
// One atomic hands out both the entry's index and its first row in the list,
// so entries in index order have increasing first rows.  The pass's S.*
// fields are written before the entry (k_coalesce reads them after the
// kernel boundary).
__device__ inline void enqueue_rows(const frame &F, int d, int pass, int rows) {
    unsigned long long old =
        atomicAdd(&F.counters[COUNTER_QUEUE], (1ull << 40) | (unsigned long long)rows);
    uint32_t index = (uint32_t)(old >> 40);
    unsigned long long first = old & ((1ull << 40) - 1);
    F.queue[index] = (first << 20) | (unsigned long long)(2 * d + pass);
}

// k_runs runs `passes` workgroups per draw: the main pass, and the shadow pass
// when the frame has shadows.
__global__ __launch_bounds__(GEO_THREADS) void k_runs(frame F, int passes) {
    __shared__ uint32_t lds[16];
    __shared__ uint64_t slot[1];
    __shared__ int s_ymin, s_ymax;
    __shared__ uint32_t s_rows[LDS_ROWS];
    __shared__ int s_b[6];
    __shared__ unsigned long long s_key;
    int d = (int)blockIdx.x / passes;
    int variant = (int)blockIdx.x % passes;
    if (d >= F.ndraws || overflowed(F))
        return;
    const draw_desc &D = F.draws[d];
    draw_state &S = F.state[d];
    if (variant && !D.shadow)
        return;
    const canvas_desc &C = F.canvases[D.canvas];
    int pad = variant ? 2 * D.border : 0;
    f2 offset = variant ? mk((float)D.border + D.shadow_offset_x,
                             (float)D.border + D.shadow_offset_y)
                        : mk(0.0f, 0.0f);
    float width = (float)(C.width + pad);
    float height = (float)(C.height + pad);
    const f2 *pts = (const f2 *)(uintptr_t)S.poly_points;
    const uint32_t *sub = (const uint32_t *)(uintptr_t)S.poly_counts;
    uint32_t nsub = S.poly_nsub;

    // Per subpath: point offset, clip scratch, edge count.
    uint64_t w = block_alloc(F, COUNTER_WORDS, 3 * (uint64_t)(nsub + 1), F.words_cap, slot);
    if (w == ~0ull)
        return;
    uint32_t *at = F.words + w;
    uint32_t *need = at + nsub + 1;
    uint32_t *edges = need + nsub + 1;
    for (uint32_t s = threadIdx.x; s < nsub; s += blockDim.x)
        at[s] = sub[2 * s];
    __syncthreads();
    scan_array(at, nsub, lds);
    for (uint32_t s = threadIdx.x; s < nsub; s += blockDim.x) {
        uint32_t n = sub[2 * s];
        bool inside = true;
        for (uint32_t k = 0; k < n && inside; ++k)
            inside = clip_inside(add(offset, pts[at[s] + k]), width, height);
        need[s] = inside ? 0u : 24u * n;
    }
    __syncthreads();
    uint32_t scratch = scan_array(need, nsub, lds);
    uint64_t sp = block_alloc(F, COUNTER_POINTS, scratch, F.points_cap, slot);
    if (sp == ~0ull)
        return;
    f2 *clip = (f2 *)F.points + sp;
    // Sutherland-Hodgman (hpp:2208-2229), one thread per crossing subpath;
    // stage k writes at most 2^(k+1) n points: a holds 8n, b holds 16n.
    for (uint32_t s = threadIdx.x; s < nsub; s += blockDim.x) {
        uint32_t n = sub[2 * s];
        if (need[s + 1] == need[s]) {
            edges[s] = n;
            continue;
        }
        struct buf {
            f2 *p;
            uint32_t n;
            __device__ void point(f2 v) { p[n++] = v; }
        } x = {clip + need[s], 0}, y = {clip + need[s] + 8 * n, 0};
        const f2 *src = pts + at[s];
        for (uint32_t k = 0; k < n; ++k)
            clip_edge(x, add(offset, src[(k ? k : n) - 1]), add(offset, src[k]), 0, width,
                      height);
        for (int e = 1; e < 4; ++e) {
            buf &i = (e & 1) ? x : y;
            buf &o = (e & 1) ? y : x;
            o.n = 0;
            for (uint32_t k = 0; k < i.n; ++k)
                clip_edge(o, i.p[(k ? k : i.n) - 1], i.p[k], e, width, height);
        }
        edges[s] = y.n;
    }
    __syncthreads();
    uint32_t nedges = scan_array(edges, nsub, lds);
    edge_source src = {pts, clip, sub, at, need, edges, nsub, offset};

    // Row range of the clamped edges (clamped coordinates are >= 0, so their
    // float bits order like ints).
    if (threadIdx.x == 0) {
        s_ymin = 0x7fffffff;
        s_ymax = -1;
        s_b[0] = 0x7fffffff, s_b[1] = -1, s_b[2] = 0x7fffffff, s_b[3] = -1;
        s_b[4] = 0x7fffffff, s_b[5] = -1;
        s_key = ~0ull;
    }
    __syncthreads();
    int ymin = 0x7fffffff, ymax = -1;
    for (uint32_t e = threadIdx.x; e < nedges; e += blockDim.x) {
        f2 from, to;
        src.get(e, from, to);
        float y = clamp_point(to, width, height).y;
        if (y == y) { // not NaN
            int bits = __float_as_int(y == 0.0f ? 0.0f : y);
            ymin = min(ymin, bits);
            ymax = max(ymax, bits);
        }
    }
    atomicMin(&s_ymin, ymin);
    atomicMax(&s_ymax, ymax);
    __syncthreads();
    int row0 = 0, rows = 0;
    if (s_ymax >= 0) {
        row0 = (int)floorf(__int_as_float(s_ymin));
        rows = (int)floorf(__int_as_float(s_ymax)) - row0 + 1;
    }
    uint64_t ro = block_alloc(F, COUNTER_WORDS, (uint64_t)rows + 1, F.words_cap, slot);
    if (ro == ~0ull)
        return;
    uint32_t *rowoff = F.words + ro;
    uint32_t *counter = s_rows;
    if (rows > LDS_ROWS) {
        uint64_t rc = block_alloc(F, COUNTER_WORDS, (uint64_t)rows, F.words_cap, slot);
        if (rc == ~0ull)
            return;
        counter = F.words + rc;
    }
    // Work items: (edge, pixel row) pairs, the rows of each edge in walk
    // order; thread t takes a contiguous range of items.
    uint64_t iw = block_alloc(F, COUNTER_WORDS, (uint64_t)nedges + 1, F.words_cap, slot);
    if (iw == ~0ull)
        return;
    uint32_t *erows = F.words + iw;
    for (uint32_t e = threadIdx.x; e < nedges; e += blockDim.x) {
        f2 from, to;
        src.get(e, from, to);
        edge_rows E = edge_rows_setup(clamp_point(from, width, height),
                                      clamp_point(to, width, height));
        erows[e] = E.live ? (uint32_t)E.rows : 0u;
    }
    for (int r = threadIdx.x; r < rows; r += blockDim.x)
        counter[r] = 0;
    __syncthreads();
    uint32_t nitems = scan_array(erows, nedges, lds);
    uint32_t item_lo = (uint32_t)((uint64_t)nitems * threadIdx.x / blockDim.x);
    uint32_t item_hi = (uint32_t)((uint64_t)nitems * (threadIdx.x + 1) / blockDim.x);
    // count
    {
        uint32_t e = item_lo < item_hi ? find_prefix(erows, nedges, item_lo) : 0, ecur = ~0u;
        edge_rows E;
        for (uint32_t g = item_lo; g < item_hi; ++g) {
            while (g >= erows[e + 1])
                ++e;
            if (e != ecur) {
                f2 from, to;
                src.get(e, from, to);
                E = edge_rows_setup(clamp_point(from, width, height),
                                    clamp_point(to, width, height));
                ecur = e;
            }
            int k = (int)(g - erows[e]);
            int row = (int)(E.pixel.y + (float)k * E.y_step) - row0;
            if (row >= 0 && row < rows)
                atomicAdd(&counter[row], (uint32_t)item_runs(E, k));
        }
    }
    __syncthreads();
    for (int r = threadIdx.x; r < rows; r += blockDim.x)
        rowoff[r] = counter[r];
    __syncthreads();
    uint32_t nruns = scan_array(rowoff, (uint32_t)rows, lds);
    uint64_t rb = block_alloc(F, COUNTER_RUNS, nruns, F.runs_cap, slot);
    if (rb == ~0ull)
        return;
    for (int r = threadIdx.x; r < rows; r += blockDim.x)
        counter[r] = rowoff[r];
    __syncthreads();
    piece_writer wr;
    wr.out = F.runs + rb;
    wr.minx = 0x7fffffff, wr.maxx = -1;
    wr.nzminx = 0x7fffffff, wr.nzmaxx = -1, wr.nzminy = 0x7fffffff, wr.nzmaxy = -1;
    wr.minkey = ~0ull;
    {
        uint32_t e = item_lo < item_hi ? find_prefix(erows, nedges, item_lo) : 0, ecur = ~0u;
        edge_rows E;
        for (uint32_t g = item_lo; g < item_hi; ++g) {
            while (g >= erows[e + 1])
                ++e;
            if (e != ecur) {
                f2 from, to;
                src.get(e, from, to);
                E = edge_rows_setup(clamp_point(from, width, height),
                                    clamp_point(to, width, height));
                ecur = e;
            }
            int k = (int)(g - erows[e]);
            int row = (int)(E.pixel.y + (float)k * E.y_step) - row0;
            if (row < 0 || row >= rows)
                continue;
            wr.at = atomicAdd(&counter[row], (uint32_t)item_runs(E, k));
            add_runs_row(wr, E, k);
        }
    }
    atomicMin(&s_b[0], wr.minx);
    atomicMax(&s_b[1], wr.maxx);
    atomicMin(&s_b[2], wr.nzminx);
    atomicMax(&s_b[3], wr.nzmaxx);
    atomicMin(&s_b[4], wr.nzminy);
    atomicMax(&s_b[5], wr.nzmaxy);
    atomicMin(&s_key, wr.minkey);
    __syncthreads();
    if (threadIdx.x != 0)
        return;
    if (!variant) {
        S.clast = -1; // raised by k_coalesce for clip draws
        if (rows > 0)
            enqueue_rows(F, d, 0, rows);
        S.row0 = row0;
        S.rows = rows;
        S.rowoff = ro;
        S.runs = rb;
        S.minx = s_b[0];
        S.maxx = s_b[1];
        atomicAdd(&F.counters[COUNTER_RUNS_DONE], 1ull);
        return;
    }
    S.srow0 = row0;
    S.srows = rows;
    S.srowoff = ro;
    S.sruns = rb;
    S.slast = -1; // raised by k_coalesce
    if (rows > 0)
        enqueue_rows(F, d, 1, rows);
    // Shadow bounds (hpp:2409-2426) over the runs that survive coalescing:
    // those with a non-zero delta, plus the first run in sort order.
    int b = D.border;
    int left = s_b[2], right = s_b[3], top = s_b[4], bottom = s_b[5];
    bool any = right >= 0;
    if (s_key != ~0ull && (uint32_t)s_key == 0u) {
        int kx = (int)((s_key >> 32) & 0xffff), ky = (int)(s_key >> 48);
        left = any ? min(left, kx) : kx;
        right = any ? max(right, kx) : kx;
        top = any ? min(top, ky) : ky;
        bottom = any ? max(bottom, ky) : ky;
        any = true;
    }
    if (!any) {
        left = C.width + 2 * b;
        right = 0;
        top = C.height + 2 * b;
        bottom = 0;
    }
    left = max(left - b, 0);
    right = min(right + b, C.width + 2 * b) + 1;
    top = max(top - b, 0);
    bottom = min(bottom + b, C.height + 2 * b);
    S.left = left;
    S.top = top;
    S.width = max(right - left, 0);
    S.height = max(bottom - top, 0);
    uint64_t size = (uint64_t)S.width * (uint64_t)S.height;
    uint64_t img = size ? arena_alloc(F, COUNTER_SHADOW, 2 * size, F.shadow_cap) : 0;
    if (img == ~0ull) {
        S.width = S.height = 0;
        img = 0;
    } else {
        atomicAdd(&F.counters[COUNTER_RUNS_DONE], 1ull);
    }
    S.image = img;
    S.image2 = img + size;
}

// ---------------------------------------------------------------------------
// Coverage: runs coalesced into per-row steps, read back per row segment.
//
// The reference sorts a draw's runs by (y, x, |delta|) (hpp:2172-2181,
// hpp:2241), adds up the deltas of one pixel in that order (hpp:2242-2250)
// and then keeps a float running sum along each row (hpp:2565-2604,
// hpp:2430-2451): a pixel's coverage is min(|running sum after the last run
// at or left of it|, 1).  k_coalesce repeats exactly that arithmetic for each
// row bucket: the runs are sorted by the same key (equal magnitudes: the
// negative delta first), each pixel's deltas are added in sorted order, and
// the row's sum is carried from pixel to pixel one float addition at a time.
// A row's result is a list of steps (x, running sum after x) in x order,
// stored in the step arena at the row's run offsets and padded to the
// bucket's size with sentinels (x = STEP_SENTINEL, sum = the bits of the
// row's last step x, or of -1).  Pixels whose deltas are all zero give no
// step: the reference's coalescing drops those runs, and they change no sum.

#define STEP_SENTINEL 0x7fffffff
#define COALESCE_CAP 1024 // runs sorted at once in LDS

// Sort key: x, then |delta|, then the negative delta first.
__device__ inline unsigned long long run_key(run r) {
    uint32_t bits = __float_as_uint(r.delta);
    return ((unsigned long long)(uint32_t)r.x << 32) |
           ((unsigned long long)(bits & 0x7fffffffu) << 1) | (unsigned long long)(1u - (bits >> 31));
}
__device__ inline int key_x(unsigned long long k) { return (int)(uint32_t)(k >> 32); }
__device__ inline float key_delta(unsigned long long k) {
    uint32_t mag = (uint32_t)(k >> 1) & 0x7fffffffu;
    return __uint_as_float(mag | ((uint32_t)(~k & 1ull) << 31));
}

__device__ inline unsigned long long shfl_xor64(unsigned long long v, int mask) {
    uint32_t lo = (uint32_t)__shfl_xor((int)(uint32_t)v, mask, 64);
    uint32_t hi = (uint32_t)__shfl_xor((int)(uint32_t)(v >> 32), mask, 64);
    return ((unsigned long long)hi << 32) | lo;
}

__device__ inline unsigned long long wave_min64(unsigned long long v) {
    for (int m = 32; m >= 1; m >>= 1) {
        unsigned long long o = shfl_xor64(v, m);
        v = o < v ? o : v;
    }
    return v;
}

__device__ inline int wave_min(int v) {
    for (int m = 32; m >= 1; m >>= 1)
        v = min(v, __shfl_xor(v, m, 64));
    return v;
}

__device__ inline int wave_max(int v) {
    for (int m = 32; m >= 1; m >>= 1)
        v = max(v, __shfl_xor(v, m, 64));
    return v;
}

// Bitonic sort of one key per lane over the wave, ascending in lane order.
__device__ inline unsigned long long wave_sort(unsigned long long key) {
    int lane = threadIdx.x & 63;
#pragma unroll
    for (int k = 2; k <= 64; k <<= 1)
#pragma unroll
        for (int j = k >> 1; j > 0; j >>= 1) {
            unsigned long long o = shfl_xor64(key, j);
            bool take_min = ((lane & j) == 0) == ((lane & k) == 0);
            key = take_min ? (o < key ? o : key) : (o > key ? o : key);
        }
    return key;
}

// Bitonic sort of buf[0, n), n a power of two, by a one-wave workgroup.
__device__ void block_sort(unsigned long long *buf, uint32_t n) {
    for (uint32_t k = 2; k <= n; k <<= 1)
        for (uint32_t j = k >> 1; j > 0; j >>= 1) {
            for (uint32_t i = threadIdx.x; i < n; i += 64) {
                uint32_t p = i ^ j;
                if (p > i) {
                    unsigned long long a = buf[i], b = buf[p];
                    if ((a > b) == ((i & k) == 0)) {
                        buf[i] = b;
                        buf[p] = a;
                    }
                }
            }
            __syncthreads();
        }
}

// Turns a row's runs, fed in key order, into its steps.  All members are
// wave-uniform except bx/bs, the step that lane (index mod 64) buffers for a
// coalesced store.
struct step_writer {
    step *out;
    uint32_t count;
    int last_x;
    int open_x; // pixel whose deltas may continue in the next chunk, or -1
    float open_sum;
    bool open_nz;
    float sum; // the row's running sum
    int bx;
    float bs;
    bool deltas; // write each pixel's summed delta instead of the running sum

    __device__ void init(step *o, bool d) {
        out = o;
        deltas = d;
        count = 0;
        last_x = -1;
        open_x = -1;
        open_sum = 0.0f;
        open_nz = false;
        sum = 0.0f;
        bx = 0;
        bs = 0.0f;
    }
    __device__ void emit(int x, float s) {
        int lane = threadIdx.x & 63;
        if (lane == (int)(count & 63)) {
            bx = x;
            bs = s;
        }
        ++count;
        last_x = x;
        if ((count & 63) == 0) {
            step v = {bx, bs};
            out[count - 64 + lane] = v;
        }
    }
    __device__ void close_open() {
        if (open_x >= 0 && open_nz) {
            sum = sum + open_sum;
            emit(open_x, deltas ? open_sum : sum);
        }
        open_x = -1;
    }
    // One delta of pixel x (runs of one pixel arrive consecutively).
    __device__ void feed_one(int x, float d) {
        if (x != open_x) {
            close_open();
            open_x = x;
            open_sum = d;
            open_nz = d != 0.0f;
        } else {
            open_sum = open_sum + d;
            open_nz = open_nz || d != 0.0f;
        }
    }
    // cnt (1..64) keys in ascending lane order; `more`: later keys follow.
    __device__ void chunk(unsigned long long key, int cnt, bool more) {
        int lane = threadIdx.x & 63;
        bool valid = lane < cnt;
        int x = valid ? key_x(key) : -2;
        float d = valid ? key_delta(key) : 0.0f;
        int x_first = __shfl(x, 0, 64);
        if (open_x >= 0 && x_first != open_x)
            close_open();
        bool cont = open_x >= 0; // lane 0 continues the open pixel
        int prev_x = __shfl_up(x, 1, 64);
        bool head = valid && (lane == 0 || x != prev_x);
        float m = (lane == 0 && cont) ? open_sum + d : d;
        bool nz = d != 0.0f || (lane == 0 && cont && open_nz);
        for (int t = 1; t < 64; ++t) {
            float nd = __shfl_down(d, t, 64);
            int nx = __shfl_down(x, t, 64);
            bool same = head && lane + t < cnt && nx == x;
            if (same) {
                m = m + nd;
                nz = nz || nd != 0.0f;
            }
            if (__ballot(same) == 0)
                break;
        }
        unsigned long long heads = __ballot(head);
        int h = 63 - __clzll(heads); // the chunk's last pixel
        open_x = -1;
        if (more) {
            open_x = __shfl(x, h, 64);
            open_sum = __shfl(m, h, 64);
            open_nz = __shfl((int)nz, h, 64) != 0;
        }
        unsigned long long kept = __ballot(head && nz && !(more && lane == h));
        while (kept) {
            int l = __ffsll((long long)kept) - 1;
            kept &= kept - 1;
            float ml = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(m), l));
            int xl = __builtin_amdgcn_readlane(x, l);
            sum = sum + ml;
            emit(xl, deltas ? ml : sum);
        }
    }
    // Closes the row: flushes buffered steps and pads to n with sentinels.
    __device__ uint32_t finish(uint32_t n) {
        close_open();
        int lane = threadIdx.x & 63;
        uint32_t r = count & 63;
        if ((uint32_t)lane < r) {
            step v = {bx, bs};
            out[count - r + lane] = v;
        }
        step pad = {STEP_SENTINEL, __int_as_float(last_x)};
        for (uint32_t i = count + lane; i < n; i += 64)
            out[i] = pad;
        return count;
    }
};

// Coalesces one row bucket in[0, n) into out[0, n); returns the step count.
__device__ uint32_t coalesce_row(const run *in, uint32_t n, step *out, unsigned long long *buf,
                                 uint32_t &s_count, bool deltas) {
    int lane = threadIdx.x & 63;
    step_writer w;
    w.init(out, deltas);
    if (n <= 64) {
        unsigned long long key = (uint32_t)lane < n ? run_key(in[lane]) : ~0ull;
        w.chunk(wave_sort(key), (int)n, false);
        return w.finish(n);
    }
    if (n <= COALESCE_CAP) {
        uint32_t N = 128;
        while (N < n)
            N <<= 1;
        for (uint32_t i = lane; i < N; i += 64)
            buf[i] = i < n ? run_key(in[i]) : ~0ull;
        __syncthreads();
        block_sort(buf, N);
        for (uint32_t c = 0; c < n; c += 64)
            w.chunk((c + lane < n) ? buf[c + lane] : ~0ull, (int)min(64u, n - c), c + 64 < n);
        __syncthreads();
        return w.finish(n);
    }
    // Longer rows: windows of whole pixels holding at most COALESCE_CAP runs.
    int xa = 0x7fffffff, xe = -1;
    for (uint32_t i = lane; i < n; i += 64) {
        int x = in[i].x;
        xa = min(xa, x);
        xe = max(xe, x);
    }
    xa = wave_min(xa);
    xe = wave_max(xe) + 1;
    uint32_t *hist = (uint32_t *)buf; // 1024 counters, aliasing the keys
    while (xa < xe) {
        for (int i = lane; i < 1024; i += 64)
            hist[i] = 0;
        __syncthreads();
        for (uint32_t i = lane; i < n; i += 64) {
            int x = in[i].x;
            if (x >= xa && x - xa < 1024)
                atomicAdd(&hist[x - xa], 1u);
        }
        __syncthreads();
        uint32_t local = 0;
        for (int j = 0; j < 16; ++j)
            local += hist[16 * lane + j];
        uint32_t incl = local;
        for (int dd = 1; dd < 64; dd <<= 1) {
            uint32_t o = (uint32_t)__shfl_up((int)incl, dd, 64);
            if (lane >= dd)
                incl += o;
        }
        int take = 0;
        uint32_t c = incl - local;
        if (c <= COALESCE_CAP)
            for (int j = 0; j < 16; ++j) {
                c += hist[16 * lane + j];
                if (c <= COALESCE_CAP)
                    take = 16 * lane + j + 1;
            }
        take = wave_max(take);
        __syncthreads(); // histogram read before the keys overwrite it
        if (take == 0) {
            // one pixel with more than COALESCE_CAP runs: feed them in key
            // order by repeated selection
            unsigned long long last = 0;
            bool started = false;
            for (;;) {
                unsigned long long best = ~0ull;
                for (uint32_t i = lane; i < n; i += 64) {
                    run r = in[i];
                    unsigned long long k = run_key(r);
                    if (r.x == xa && (!started || k > last) && k < best)
                        best = k;
                }
                best = wave_min64(best);
                if (best == ~0ull)
                    break;
                int same = 0;
                for (uint32_t i = lane; i < n; i += 64)
                    same += (in[i].x == xa && run_key(in[i]) == best) ? 1 : 0;
                for (int m = 32; m >= 1; m >>= 1)
                    same += __shfl_xor(same, m, 64);
                for (int k = 0; k < same; ++k)
                    w.feed_one(xa, key_delta(best));
                last = best;
                started = true;
            }
            xa += 1;
            continue;
        }
        int xb = xa + take;
        if (lane == 0)
            s_count = 0;
        __syncthreads();
        for (uint32_t i = lane; i < n; i += 64) {
            run r = in[i];
            if (r.x >= xa && r.x < xb)
                buf[atomicAdd(&s_count, 1u)] = run_key(r);
        }
        __syncthreads();
        uint32_t cnt = s_count;
        if (cnt) {
            uint32_t N = 64;
            while (N < cnt)
                N <<= 1;
            for (uint32_t i = cnt + lane; i < N; i += 64)
                buf[i] = ~0ull;
            __syncthreads();
            block_sort(buf, N);
            for (uint32_t c2 = 0; c2 < cnt; c2 += 64)
                w.chunk((c2 + lane < cnt) ? buf[c2 + lane] : ~0ull, (int)min(64u, cnt - c2),
                        c2 + 64 < cnt);
        }
        __syncthreads(); // keys read before the next histogram
        xa = xb;
    }
    return w.finish(n);
}

// The work list entry holding list row u: the last of the n entries whose
// first row is <= u (entry 0 starts at row 0), by a 64-way search.
__device__ inline uint32_t find_entry(const unsigned long long *queue, uint32_t n, uint64_t u) {
    int lane = threadIdx.x & 63;
    uint32_t lo = 0, hi = n; // answer in [lo, hi); queue[lo] starts at or before u
    for (int level = 0; level < 8 && hi - lo > 1; ++level) { // 64^8 > any list
        uint32_t stride = (hi - lo + 63) / 64;
        uint32_t p = lo + (uint32_t)lane * stride;
        bool before = p < hi && (queue[p] >> 20) <= u;
        int c = __popcll(__ballot(before)); // >= 1: lane 0 probes lo
        if (c == 0)
            break;
        lo = lo + (uint32_t)(c - 1) * stride;
        hi = min(hi, lo + stride);
    }
    return lo;
}

// A row of at most SMALL_ROW runs, coalesced by one lane: the keys sorted by
// a bitonic network in registers, then the same pixel sums and running sum as
// step_writer, in the same order.  Returns the number of steps.
#define SMALL_ROW 32

__device__ inline uint32_t coalesce_small(const run *in, uint32_t n, step *out, bool deltas) {
    unsigned long long k[SMALL_ROW];
#pragma unroll
    for (int i = 0; i < SMALL_ROW; ++i)
        k[i] = (uint32_t)i < n ? run_key(in[i]) : ~0ull;
#pragma unroll
    for (int size = 2; size <= SMALL_ROW; size <<= 1)
#pragma unroll
        for (int j = size >> 1; j > 0; j >>= 1)
#pragma unroll
            for (int i = 0; i < SMALL_ROW; ++i) {
                int p = i ^ j;
                if (p > i) {
                    unsigned long long lo = k[i] < k[p] ? k[i] : k[p];
                    unsigned long long hi = k[i] < k[p] ? k[p] : k[i];
                    bool up = (i & size) == 0;
                    k[i] = up ? lo : hi;
                    k[p] = up ? hi : lo;
                }
            }
    uint32_t count = 0;
    int last_x = -1, open_x = -1;
    float sum = 0.0f, open_sum = 0.0f;
    bool open_nz = false;
#pragma unroll
    for (int i = 0; i < SMALL_ROW; ++i) {
        if ((uint32_t)i < n) {
            int x = key_x(k[i]);
            float d = key_delta(k[i]);
            if (x != open_x) {
                if (open_x >= 0 && open_nz) {
                    sum = sum + open_sum;
                    step v = {open_x, deltas ? open_sum : sum};
                    out[count++] = v;
                    last_x = open_x;
                }
                open_x = x;
                open_sum = d;
                open_nz = d != 0.0f;
            } else {
                open_sum = open_sum + d;
                open_nz = open_nz || d != 0.0f;
            }
        }
    }
    if (open_x >= 0 && open_nz) {
        sum = sum + open_sum;
        step v = {open_x, deltas ? open_sum : sum};
        out[count++] = v;
        last_x = open_x;
    }
    step pad = {STEP_SENTINEL, __int_as_float(last_x)};
    for (uint32_t i = count; i < n; ++i)
        out[i] = pad;
    return count;
}

// The work list's rows, taken `per` at a time: row (u, its draw, pass, row
// index and run bucket) for lane t < per of a wave, chunk c holding list rows
// [c * per, c * per + per).
struct list_row {
    bool valid;
    int d, pass, r;
    uint32_t b, n;
    uint64_t base;
};

__device__ inline list_row work_row(const frame &F, uint32_t nentries, uint64_t total,
                                    uint64_t chunk, int per) {
    int lane = threadIdx.x & 63;
    uint64_t u0 = chunk * (uint64_t)per;
    uint32_t i = find_entry(F.queue, nentries, u0);
    uint64_t u = u0 + (uint64_t)lane;
    list_row w;
    w.valid = lane < per && u < total;
    // this lane's entry: a walk forward from the chunk's first
    while (w.valid && i + 1 < nentries && (F.queue[i + 1] >> 20) <= u)
        ++i;
    w.d = w.pass = w.r = 0;
    w.b = w.n = 0;
    w.base = 0;
    if (w.valid) {
        unsigned long long e = F.queue[i];
        int dp = (int)(e & 0xfffff);
        w.d = dp >> 1;
        w.pass = dp & 1;
        uint64_t first = e >> 20;
        const draw_state &S = F.state[w.d < F.ndraws ? w.d : 0];
        uint64_t rowoff = w.pass ? S.srowoff : S.rowoff;
        int rows = w.pass ? S.srows : S.rows;
        w.base = w.pass ? S.sruns : S.runs;
        // a row the list and the draw's state agree on, inside the arenas
        w.valid = w.d < F.ndraws && u >= first && u - first < (uint64_t)rows &&
                  rowoff + (uint64_t)rows < F.words_cap;
        if (w.valid) {
            w.r = (int)(u - first);
            const uint32_t *off = F.words + rowoff;
            w.b = off[w.r];
            w.n = off[w.r + 1] - w.b;
            w.valid = w.base + w.b + w.n <= F.runs_cap;
        }
        if (!w.valid)
            atomicMax(&F.counters[COUNTER_OVERFLOW], 1ull); // the host re-runs and reports
    }
    return w;
}

// Coalescing runs in two kernels over the work list, so each keeps its own
// register budget: k_coalesce gives each row of at most SMALL_ROW runs one
// lane (64 rows per wave); k_coalesce_wide gives each longer row the whole
// wave (coalesce_row), taking the list WIDE_CHUNK rows at a time so long rows
// spread over many waves.  Workgroup g takes chunks g, g + G, ...
#define COALESCE_GRID (256 * 16)
#define WIDE_CHUNK 8

__global__ __launch_bounds__(64) void k_coalesce(frame F) {
    if (overflowed(F))
        return;
    unsigned long long q = __atomic_load_n(&F.counters[COUNTER_QUEUE], __ATOMIC_RELAXED);
    uint32_t nentries = (uint32_t)(q >> 40);
    uint64_t total = q & ((1ull << 40) - 1);
    uint64_t nchunks = (total + 63) / 64;
    for (uint64_t chunk = blockIdx.x; chunk < nchunks; chunk += gridDim.x) {
        list_row w = work_row(F, nentries, total, chunk, 64);
        if (w.valid && w.n > 0 && w.n <= SMALL_ROW) {
            bool clip = F.draws[w.d].kind == CIH_DRAW_CLIP;
            uint32_t count =
                coalesce_small(F.runs + w.base + w.b, w.n, F.steps + w.base + w.b, clip);
            if (w.pass && count)
                atomicMax(&F.state[w.d].slast, F.state[w.d].srow0 + w.r);
            if (clip && count)
                atomicMax(&F.state[w.d].clast, F.state[w.d].row0 + w.r);
        }
    }
}

__global__ __launch_bounds__(64) void k_coalesce_wide(frame F) {
    __shared__ unsigned long long buf[COALESCE_CAP];
    __shared__ uint32_t s_count;
    if (overflowed(F))
        return;
    int lane = threadIdx.x & 63;
    unsigned long long q = __atomic_load_n(&F.counters[COUNTER_QUEUE], __ATOMIC_RELAXED);
    uint32_t nentries = (uint32_t)(q >> 40);
    uint64_t total = q & ((1ull << 40) - 1);
    uint64_t nchunks = (total + WIDE_CHUNK - 1) / WIDE_CHUNK;
    for (uint64_t chunk = blockIdx.x; chunk < nchunks; chunk += gridDim.x) {
        list_row w = work_row(F, nentries, total, chunk, WIDE_CHUNK);
        unsigned long long wide = __ballot(w.valid && w.n > SMALL_ROW);
        while (wide) {
            int l = __ffsll((long long)wide) - 1;
            wide &= wide - 1;
            int d = __builtin_amdgcn_readlane(w.d, l), pass = __builtin_amdgcn_readlane(w.pass, l);
            int r = __builtin_amdgcn_readlane(w.r, l);
            uint32_t b = (uint32_t)__builtin_amdgcn_readlane((int)w.b, l);
            uint32_t n = (uint32_t)__builtin_amdgcn_readlane((int)w.n, l);
            uint64_t base =
                ((uint64_t)(uint32_t)__builtin_amdgcn_readlane((int)(w.base >> 32), l) << 32) |
                (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)w.base, l);
            bool clip = F.draws[d].kind == CIH_DRAW_CLIP;
            uint32_t count =
                coalesce_row(F.runs + base + b, n, F.steps + base + b, buf, s_count, clip);
            if (pass && count && lane == 0)
                atomicMax(&F.state[d].slast, F.state[d].srow0 + r);
            if (clip && count && lane == 0)
                atomicMax(&F.state[d].clast, F.state[d].row0 + r);
        }
    }
}

// ---------------------------------------------------------------------------
// Clip masks (hpp:3057-3099).  The reference merges the clip path's coalesced
// runs with the current mask's runs in (y, x, |delta|) order (a tie takes the
// mask's run first), keeps a running sum of each list per row, and appends a
// piece wherever visibility = min(|sum_1|, 1) * min(|sum_2|, 1) changes,
// adding to the previous piece when it has the same pixel.  It stops when
// either list runs out, so rows after the last row of either are empty and
// that last row ends at the exhausted list's last run.  Rows are independent,
// so k_clip gives each row of the canvas one thread; the clip draw's steps
// hold each pixel's summed delta (k_coalesce's delta mode).  The constructor's
// mask (hpp:2636-2643) is mask 0: pieces (0, +1) and (width, -1) on every row.

__global__ __launch_bounds__(64) void k_clip(frame F, int d) {
    if (overflowed(F))
        return;
    const draw_desc &D = F.draws[d];
    const draw_state &S = F.state[d];
    const canvas_desc &C = F.canvases[D.canvas];
    const mask_desc *T = (const mask_desc *)(uintptr_t)C.masks;
    int y = (int)(blockIdx.x * 64 + threadIdx.x);
    if (!T || y >= C.height)
        return;
    uint32_t cap = (uint32_t)C.width + 2;
    const mask_desc &M = T[D.image];
    run *out = (run *)(uintptr_t)M.pieces + (size_t)y * cap;
    step *vis = (step *)(uintptr_t)M.steps + (size_t)y * cap;
    int parent = D.clip_mask;
    int last1 = S.clast;
    int last2 = parent ? T[parent].last_row : C.height - 1;
    uint32_t n = 0;
    if (last1 >= 0 && last2 >= 0 && y <= min(last1, last2)) {
        // the clip path's entries (x, summed delta) in this row
        const step *pe = nullptr;
        uint32_t n1 = 0;
        if (S.rows > 0 && y >= S.row0 && y < S.row0 + S.rows) {
            const uint32_t *ro = F.words + S.rowoff + (y - S.row0);
            pe = F.steps + S.runs + ro[0];
            uint32_t slots = ro[1] - ro[0];
            while (n1 < slots && pe[n1].x != STEP_SENTINEL)
                ++n1;
        }
        // the parent mask's pieces in this row
        run whole[2];
        whole[0].x = 0, whole[0].delta = 1.0f;
        whole[1].x = C.width, whole[1].delta = -1.0f;
        const run *pp = whole;
        uint32_t n2 = 2;
        if (parent) {
            pp = (const run *)(uintptr_t)T[parent].pieces + (size_t)y * cap;
            n2 = ((const uint32_t *)(uintptr_t)T[parent].counts)[y];
        }
        bool stop1 = y == last1, stop2 = y == last2;
        float last = 0.0f, sum_1 = 0.0f, sum_2 = 0.0f;
        uint32_t i1 = 0, i2 = 0;
        for (;;) {
            bool has1 = i1 < n1, has2 = i2 < n2;
            if ((!has1 && !has2) || (!has1 && stop1) || (!has2 && stop2))
                break;
            bool first = has1 && (!has2 || pe[i1].x < pp[i2].x ||
                                  (pe[i1].x == pp[i2].x && fabsf(pe[i1].sum) < fabsf(pp[i2].delta)));
            int x;
            if (first) {
                x = pe[i1].x;
                sum_1 = sum_1 + pe[i1].sum;
                ++i1;
            } else {
                x = pp[i2].x;
                sum_2 = sum_2 + pp[i2].delta;
                ++i2;
            }
            float visibility = smin(fabsf(sum_1), 1.0f) * smin(fabsf(sum_2), 1.0f);
            if (visibility == last)
                continue;
            if (n && out[n - 1].x == x) {
                out[n - 1].delta = out[n - 1].delta + (visibility - last);
            } else {
                if (n == cap) { // cannot happen: pieces have distinct x in [0, width + 1]
                    atomicMax(&F.counters[COUNTER_OVERFLOW], 1ull);
                    break;
                }
                out[n].x = x;
                out[n].delta = visibility - last;
                ++n;
            }
            last = visibility;
        }
    }
    // the running visibility sum render_main keeps (hpp:2603)
    float sum = 0.0f;
    for (uint32_t i = 0; i < n; ++i) {
        sum = sum + out[i].delta;
        vis[i].x = out[i].x;
        vis[i].sum = sum;
    }
    ((uint32_t *)(uintptr_t)M.counts)[y] = n;
}

// The new mask's last row with a piece, for the next clip's merge.
__global__ __launch_bounds__(64) void k_clip_finish(frame F, int d) {
    if (overflowed(F))
        return;
    const draw_desc &D = F.draws[d];
    const canvas_desc &C = F.canvases[D.canvas];
    mask_desc *T = (mask_desc *)(uintptr_t)C.masks;
    if (!T)
        return;
    const uint32_t *counts = (const uint32_t *)(uintptr_t)T[D.image].counts;
    int last = -1;
    for (int y = threadIdx.x; y < C.height; y += 64)
        if (counts[y])
            last = y;
    last = wave_max(last);
    if (threadIdx.x == 0)
        T[D.image].last_row = last;
}

// Reading steps.  A fine workgroup is one wave owning FINE_SEGMENT pixels
// [x0, x0 + FINE_SEGMENT) of a row; lane t owns FINE_PX consecutive pixels.

template <int PX> // pixels per lane: the segment is 64 * PX pixels
struct seg_lds_t {
    float sum[64 * PX];
    uint32_t mask[2 * PX]; // pixels with a step
};
typedef seg_lds_t<FINE_PX> seg_lds;

__device__ inline float coverage(float sum, bool evenodd) {
    if (evenodd) {
        // extension: triangle wave of the winding sum, period 2
        float a = fmodf(fabsf(sum), 2.0f);
        return smin(a > 1.0f ? 2.0f - a : a, 1.0f);
    }
    return smin(fabsf(sum), 1.0f);
}

template <int PX>
__device__ inline void scatter_step(seg_lds_t<PX> &L, step s, int x0) {
    int p = s.x - x0;
    L.sum[p] = s.sum;
    atomicOr(&L.mask[p >> 5], 1u << (p & 31));
}

// segment_sums: running sums of this lane's pixels in the segment starting at
// x0, for a row whose steps are st[0, n), n > 0.  `before` gets the sum left of the
// segment.  Returns false (uniformly) when no step lies inside the segment:
// then every pixel's sum is `before` and sums[] is not written.  L.mask is
// all zero on entry and on return.
//
// segment_window is the second half: every step before index lo has x < x0,
// every step from index hi on has x >= x0 (hi - lo <= 64), and lane t holds
// st[lo + t] in s when lo + t < hi.
template <int PX>
__device__ bool segment_window(const step *st, uint32_t n, uint32_t lo, uint32_t hi, step s,
                               int x0, seg_lds_t<PX> &L, float &before, float sums[PX]);

// Scatters the steps st[p, n), 64 at a time, for as long as whole chunks lie
// inside the segment; returns whether any was inside.
template <int PX>
__device__ bool scatter_rest(const step *st, uint32_t n, uint32_t p, int x0, seg_lds_t<PX> &L) {
    int lane = threadIdx.x & 63;
    int x1 = x0 + (64 * PX);
    bool any = false, go = p < n;
    while (go) {
        bool v = p + (uint32_t)lane < n;
        step s = {STEP_SENTINEL, 0.0f};
        if (v)
            s = st[p + lane];
        bool in = v && s.x < x1;
        if (in)
            scatter_step(L, s, x0);
        unsigned long long m = __ballot(in);
        any = any || m != 0;
        go = __popcll(m) == 64 && p + 64 < n;
        p += 64;
    }
    return any;
}

// Each pixel's sum is that of the last scattered step at or left of it, or
// `before`; leaves L.mask zero.
template <int PX>
__device__ void segment_lookup(seg_lds_t<PX> &L, float before, float sums[PX]) {
    int lane = threadIdx.x & 63;
    wave_lds_sync();
    uint32_t mw[(64 * PX) / 32];
#pragma unroll
    for (int k = 0; k < (64 * PX) / 32; ++k)
        mw[k] = L.mask[k];
    int base = lane * PX, wi = base >> 5;
    int prev = -1;
    uint32_t mine = 0;
#pragma unroll
    for (int k = 0; k < (64 * PX) / 32; ++k) {
        if (k < wi && mw[k])
            prev = k * 32 + 31 - __builtin_clz(mw[k]);
        if (k == wi)
            mine = mw[k];
    }
#pragma unroll
    for (int k = 0; k < PX; ++k) {
        int bit = (base + k) & 31;
        uint32_t mk = mine & (0xffffffffu >> (31 - bit));
        int pos = mk ? wi * 32 + 31 - __builtin_clz(mk) : prev;
        sums[k] = pos >= 0 ? L.sum[pos] : before;
    }
    wave_lds_sync();
    if (lane < (64 * PX) / 32)
        L.mask[lane] = 0;
    wave_lds_sync();
}

template <int PX>
__device__ bool segment_window(const step *st, uint32_t n, uint32_t lo, uint32_t hi, step s,
                               int x0, seg_lds_t<PX> &L, float &before, float sums[PX]) {
    int lane = threadIdx.x & 63;
    int x1 = x0 + (64 * PX);
    bool valid = lo + (uint32_t)lane < hi;
    bool lt = valid && s.x < x0;
    int c = __popcll(__ballot(lt));
    uint32_t first = lo + (uint32_t)c;
    float left = __shfl(s.sum, c > 0 ? c - 1 : 0, 64);
    before = first == 0 ? 0.0f : (c > 0 ? left : st[first - 1].sum);
    bool in = valid && !lt && s.x < x1;
    if (in)
        scatter_step(L, s, x0);
    unsigned long long m = __ballot(in);
    bool any = m != 0;
    if (__popcll(m) == (int)(hi - lo) - c)
        any = scatter_rest(st, n, hi, x0, L) || any;
    if (!any)
        return false;
    segment_lookup(L, before, sums);
    return true;
}

// Like segment_window when the first step at or right of x0 is known: index
// `first`, with lane t holding st[first + t] (a sentinel past n), and the sum
// before the segment given.
template <int PX>
__device__ bool segment_fill(const step *st, uint32_t n, uint32_t first, step s, int x0,
                             seg_lds_t<PX> &L, float before, float sums[PX]) {
    int x1 = x0 + (64 * PX);
    bool in = s.x < x1;
    if (in)
        scatter_step(L, s, x0);
    unsigned long long m = __ballot(in);
    bool any = m != 0;
    if (__popcll(m) == 64)
        any = scatter_rest(st, n, first + 64, x0, L) || any;
    if (!any)
        return false;
    segment_lookup(L, before, sums);
    return true;
}

template <int PX>
__device__ bool segment_sums(const step *st, uint32_t n, int x0, seg_lds_t<PX> &L, float &before,
                             float sums[PX]) {
    int lane = threadIdx.x & 63;
    // first step with x >= x0: 64-way search down to one wave-wide window
    uint32_t lo = 0, hi = n;
    while (hi - lo > 64) {
        uint32_t stride = (hi - lo + 63) / 64;
        uint32_t p = lo + (uint32_t)lane * stride;
        bool lt = p < hi && st[p].x < x0;
        int c = __popcll(__ballot(lt));
        if (c == 0) {
            hi = lo;
            break;
        }
        uint32_t nhi = lo + (uint32_t)c * stride;
        lo = lo + (uint32_t)(c - 1) * stride + 1;
        hi = nhi < hi ? nhi : hi;
    }
    step s = {STEP_SENTINEL, 0.0f};
    if (lo + (uint32_t)lane < hi)
        s = st[lo + lane];
    return segment_window(st, n, lo, hi, s, x0, L, before, sums);
}

// ---------------------------------------------------------------------------
// Paint (hpp:2265-2377) and composite (hpp:2576-2592).

__device__ inline float4 load4(const float *p) { return make_float4(p[0], p[1], p[2], p[3]); }

__device__ inline float4 premultiplied(float4 c) {
    return make_float4(c.x * c.w, c.y * c.w, c.z * c.w, c.w);
}

__device__ inline float cubic_weight(float v) { // Keys, a = -0.5 (hpp:2300-2302)
    return v < 1.0f ? (1.5f * v - 2.5f) * v * v + 1.0f
                    : ((-0.5f * v + 2.5f) * v - 4.0f) * v + 2.0f;
}

__device__ float4 paint_pattern(const frame &F, const draw_desc &D, f2 point) {
    const image_desc &I = F.images[D.image];
    const float4 *texels = (const float4 *)(uintptr_t)I.texels;
    int bw = D.image_width, bh = D.image_height;
    float width = (float)bw, height = (float)bh;
    if (((D.repetition & 2) && (point.x < 0.0f || width <= point.x)) ||
        ((D.repetition & 1) && (point.y < 0.0f || height <= point.y)))
        return make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    const float *inv = D.paint_inverse;
    float scale_x = fabsf(inv[0]) + fabsf(inv[2]);
    float scale_y = fabsf(inv[1]) + fabsf(inv[3]);
    scale_x = smax(1.0f, smin(scale_x, width * 0.25f));
    scale_y = smax(1.0f, smin(scale_y, height * 0.25f));
    float reciprocal_x = 1.0f / scale_x;
    float reciprocal_y = 1.0f / scale_y;
    point = sub(point, mk(0.5f, 0.5f));
    int left = (int)ceilf(point.x - scale_x * 2.0f);
    int top = (int)ceilf(point.y - scale_y * 2.0f);
    int right = (int)ceilf(point.x + scale_x * 2.0f);
    int bottom = (int)ceilf(point.y + scale_y * 2.0f);
    bool clamp = D.brush == CIH_BRUSH_IMAGE;
    float4 total = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    float total_weight = 0.0f;
    // repeat wraps by the remainder (hpp:2303-2305, 2316-2318), kept
    // incrementally over consecutive taps
    int wy_start = top % bh, wx_start = left % bw;
    if (wy_start < 0)
        wy_start += bh;
    if (wx_start < 0)
        wx_start += bw;
    int wrap_y = wy_start;
    for (int py = top; py < bottom; ++py, wrap_y = wrap_y + 1 == bh ? 0 : wrap_y + 1) {
        float y = fabsf(reciprocal_y * ((float)py - point.y));
        float weight_y = cubic_weight(y);
        int wy = clamp ? min(max(py, 0), bh - 1) : wrap_y;
        const float4 *row = texels + (size_t)wy * (size_t)bw;
        int wrap_x = wx_start;
        for (int px = left; px < right; ++px, wrap_x = wrap_x + 1 == bw ? 0 : wrap_x + 1) {
            float x = fabsf(reciprocal_x * ((float)px - point.x));
            float weight_x = cubic_weight(x);
            int wx = clamp ? min(max(px, 0), bw - 1) : wrap_x;
            float weight = weight_x * weight_y;
            float4 c = row[wx];
            total.x = total.x + c.x * weight;
            total.y = total.y + c.y * weight;
            total.z = total.z + c.z * weight;
            total.w = total.w + c.w * weight;
            total_weight += weight;
        }
    }
    float r = 1.0f / total_weight;
    return make_float4(total.x * r, total.y * r, total.z * r, total.w * r);
}

// Gradient colour at `point` (hpp:2331-2376); `stops` holds the ascending
// stop offsets and `colors` four floats (linear, straight alpha) per stop.
__device__ float4 paint_gradient(const draw_desc &D, const float *stops, const float *colors,
                                 f2 point) {
    const float4 zero = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    f2 start = mk(D.start_x, D.start_y), end = mk(D.end_x, D.end_y);
    f2 relative = sub(point, start);
    f2 line = sub(end, start);
    float gradient = dot(relative, line);
    float span = dot(line, line);
    float offset;
    if (D.brush == CIH_BRUSH_LINEAR) {
        if (span == 0.0f)
            return zero;
        offset = gradient / span;
    } else {
        float initial = D.start_radius;
        float change = D.end_radius - initial;
        float a = span - change * change;
        float b = -2.0f * (gradient + initial * change);
        float c = dot(relative, relative) - initial * initial;
        float discriminant = b * b - 4.0f * a * c;
        if (discriminant < 0.0f || (span == 0.0f && change == 0.0f))
            return zero;
        float root = sqrtf(discriminant);
        float reciprocal = 1.0f / (2.0f * a);
        float offset_1 = (-b - root) * reciprocal;
        float offset_2 = (-b + root) * reciprocal;
        float radius_1 = initial + change * offset_1;
        float radius_2 = initial + change * offset_2;
        if (radius_2 >= 0.0f)
            offset = offset_2;
        else if (radius_1 >= 0.0f)
            offset = offset_1;
        else
            return zero;
    }
    // std::upper_bound over ascending stops = the number of stops not
    // greater than offset (all of them for a NaN offset)
    uint32_t n = D.nstops;
    uint32_t index = 0;
    for (uint32_t k = 0; k < n; ++k)
        index += offset < stops[k] ? 0u : 1u;
    if (index == 0)
        return premultiplied(load4(colors));
    if (index == n)
        return premultiplied(load4(colors + 4 * (n - 1)));
    float mix = (offset - stops[index - 1]) / (stops[index] - stops[index - 1]);
    float4 c0 = load4(colors + 4 * (index - 1)), c1 = load4(colors + 4 * index);
    float4 c = make_float4(c0.x + (c1.x - c0.x) * mix, c0.y + (c1.y - c0.y) * mix,
                           c0.z + (c1.z - c0.z) * mix, c0.w + (c1.w - c0.w) * mix);
    return premultiplied(c);
}

// paint_pixel (hpp:2265-2377) at point (already offset to the pixel centre),
// with the brush's stops and colours at the given addresses.
__device__ inline float4 paint(const frame &F, const draw_desc &D, const float *stops,
                               const float *colors, f2 point) {
    if (D.brush == CIH_BRUSH_PATTERN || D.brush == CIH_BRUSH_IMAGE)
        return paint_pattern(F, D, apply(D.paint_inverse, point));
    if (D.ncolors == 0)
        return make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    if (D.brush == CIH_BRUSH_COLOR)
        return load4(colors);
    return paint_gradient(D, stops, colors, apply(D.paint_inverse, point));
}

__device__ inline float4 paint(const frame &F, const draw_desc &D, f2 point) {
    return paint(F, D, (const float *)(F.payload + D.stops_at),
                 (const float *)(F.payload + D.colors_at), point);
}

__device__ inline void composite(float4 &back, float4 fore, int op, float visibility) {
    float mix_fore = (op & 1) ? back.w : 0.0f;
    if (op & 2)
        mix_fore = 1.0f - mix_fore;
    float mix_back = (op & 4) ? fore.w : 0.0f;
    if (op & 8)
        mix_back = 1.0f - mix_back;
    float4 blend = make_float4(fore.x * mix_fore + back.x * mix_back,
                               fore.y * mix_fore + back.y * mix_back,
                               fore.z * mix_fore + back.z * mix_back,
                               fore.w * mix_fore + back.w * mix_back);
    blend.w = smin(blend.w, 1.0f);
    float keep = 1.0f - visibility;
    back = make_float4(blend.x * visibility + back.x * keep, blend.y * visibility + back.y * keep,
                       blend.z * visibility + back.z * keep, blend.w * visibility + back.w * keep);
}

__device__ inline float4 scaled(float4 c, float s) {
    return make_float4(c.x * s, c.y * s, c.z * s, c.w * s);
}

// ---------------------------------------------------------------------------
// Shadows.

// Alpha raster of the shadow pass over its padded bounding box (hpp:2427-2452).
// A row is painted from its first step through its last step's pixel, except
// that the reference never paints the pixel of the last run of all (its loop
// paints a run's pixels only when the next run arrives).
__global__ __launch_bounds__(FINE_THREADS) void k_shadow_raster(frame F) {
    __shared__ seg_lds L;
    int d = blockIdx.x;
    if (d >= F.ndraws || overflowed(F))
        return;
    const draw_desc &D = F.draws[d];
    if (!D.shadow)
        return;
    const draw_state &S = F.state[d];
    if (threadIdx.x < FINE_SEGMENT / 32)
        L.mask[threadIdx.x] = 0;
    __syncthreads();
    float *image = F.shadow + S.image;
    const step *steps = F.steps + S.sruns;
    const uint32_t *rowoff = F.words + S.srowoff;
    bool evenodd = (D.flags & CIH_DRAW_FLAG_EVENODD) != 0;
    f2 offset = mk((float)D.border + D.shadow_offset_x, (float)D.border + D.shadow_offset_y);
    for (int r = blockIdx.y; r < S.height; r += gridDim.y) {
        int y = S.top + r;
        bool hit = y >= S.srow0 && y < S.srow0 + S.srows;
        uint32_t begin = hit ? rowoff[y - S.srow0] : 0, end = hit ? rowoff[y - S.srow0 + 1] : 0;
        int first = -1, stop = -2;
        if (begin != end) {
            step a = steps[begin], z = steps[end - 1];
            if (a.x != STEP_SENTINEL) {
                first = a.x;
                stop = z.x != STEP_SENTINEL ? z.x : __float_as_int(z.sum);
                if (y == S.slast)
                    --stop;
            }
        }
        for (int x0 = S.left; x0 < S.left + S.width; x0 += FINE_SEGMENT) {
            float before = 0.0f, sums[FINE_PX];
            bool seg_hit = first >= 0 && first < x0 + FINE_SEGMENT && stop >= x0 && first <= stop;
            bool inside = false;
            if (seg_hit)
                inside = segment_sums(steps + begin, end - begin, x0, L, before, sums);
            for (int k = 0; k < FINE_PX; ++k) {
                int x = x0 + threadIdx.x * FINE_PX + k;
                if (x >= S.left + S.width)
                    continue;
                float alpha = 0.0f;
                if (seg_hit && x >= first && x <= stop) {
                    float cov = coverage(inside ? sums[k] : before, evenodd);
                    if (cov >= kThreshold) {
                        f2 p = sub(mk((float)x + 0.5f, (float)y + 0.5f), offset);
                        alpha = cov * paint(F, D, p).w;
                    }
                }
                image[(size_t)r * S.width + (x - S.left)] = alpha;
            }
        }
    }
}

// The blur of hpp:2453-2503: three extended-box passes along every row, then
// three along every column, each pass a sequential running sum (box_line).
// One thread runs one line, so a wave's lines must lie side by side in memory:
// the column passes read the image as the raster wrote it (row-major), and the
// row passes run on a transposed copy, made and undone by k_shadow_transpose.
//
//   raster -> image (rows)  --transpose-->  image2 (columns)
//   row passes: image2 -> image -> image2 -> image (columns)
//   image (columns)  --transpose-->  image2 (rows)
//   column passes: image2 -> image -> image2 -> image (rows), read by k_fine

// Copies each shadowed draw's image into image2 transposed: with to_columns,
// image holds S.height rows of S.width samples; otherwise S.width rows of
// S.height.  64x64 tiles through LDS, so both sides are coalesced.
__global__ __launch_bounds__(256) void k_shadow_transpose(frame F, int to_columns) {
    __shared__ float tile[64][65];
    int d = blockIdx.x;
    if (d >= F.ndraws || overflowed(F))
        return;
    const draw_desc &D = F.draws[d];
    if (!D.shadow)
        return;
    const draw_state &S = F.state[d];
    int rows = to_columns ? S.height : S.width, cols = to_columns ? S.width : S.height;
    const float *src = F.shadow + S.image;
    float *dst = F.shadow + S.image2;
    int tile_cols = (cols + 63) / 64, tiles = ((rows + 63) / 64) * tile_cols;
    int lx = threadIdx.x & 63, ly = threadIdx.x >> 6;
    for (int t = blockIdx.y; t < tiles; t += gridDim.y) {
        int r0 = (t / tile_cols) * 64, c0 = (t % tile_cols) * 64;
        for (int i = ly; i < 64; i += 4)
            if (r0 + i < rows && c0 + lx < cols)
                tile[i][lx] = src[(size_t)(r0 + i) * cols + (c0 + lx)];
        __syncthreads();
        for (int i = ly; i < 64; i += 4)
            if (c0 + i < cols && r0 + lx < rows)
                dst[(size_t)(c0 + i) * rows + (r0 + lx)] = tile[lx][i];
        __syncthreads();
    }
}

// Three passes along every row, on the transposed image (sample x of row r at
// x * S.height + r): image2 -> image -> image2 -> image.
__global__ __launch_bounds__(64) void k_blur_rows(frame F) {
    int d = blockIdx.x;
    if (d >= F.ndraws || overflowed(F))
        return;
    const draw_desc &D = F.draws[d];
    const draw_state &S = F.state[d];
    if (!D.shadow)
        return;
    size_t step = (size_t)S.height;
    for (int r = blockIdx.y * 64 + threadIdx.x; r < S.height; r += gridDim.y * 64) {
        float *a = F.shadow + S.image2 + r;
        float *b = F.shadow + S.image + r;
        box_line(a, b, S.width, step, D.radius, D.weight_1, D.weight_2);
        box_line(b, a, S.width, step, D.radius, D.weight_1, D.weight_2);
        box_line(a, b, S.width, step, D.radius, D.weight_1, D.weight_2);
    }
}

// Three passes along every column (one thread per column): image2 -> image
// -> image2 -> image.  The result is in `image`.
__global__ __launch_bounds__(64) void k_blur_cols(frame F) {
    int d = blockIdx.x;
    if (d >= F.ndraws || overflowed(F))
        return;
    const draw_desc &D = F.draws[d];
    const draw_state &S = F.state[d];
    if (!D.shadow)
        return;
    size_t step = (size_t)S.width;
    for (int c = blockIdx.y * 64 + threadIdx.x; c < S.width; c += gridDim.y * 64) {
        float *a = F.shadow + S.image + c;
        float *b = F.shadow + S.image2 + c;
        box_line(b, a, S.height, step, D.radius, D.weight_1, D.weight_2);
        box_line(a, b, S.height, step, D.radius, D.weight_1, D.weight_2);
        box_line(b, a, S.height, step, D.radius, D.weight_1, D.weight_2);
    }
}

// ---------------------------------------------------------------------------
// k_fine: every draw of a canvas over one row segment, pixels in registers.

__device__ inline uint32_t find_canvas(const uint32_t *prefix, uint32_t n, uint32_t item) {
    return find_prefix(prefix, n, item);
}

// Work items go to the XCDs in contiguous ranges: workgroup b runs on XCD
// b % FINE_XCDS (round-robin dispatch), and item = (b % FINE_XCDS) * per +
// b / FINE_XCDS keeps the segments of one row, which read the same steps, in
// one L2.
#define FINE_XCDS 8
#define FINE_PERSISTENT_GRID (256 * 16) // persistent fine waves: 16 per CU

__device__ inline bool fine_item(const frame &F, uint32_t &item) {
    uint32_t per = (F.nitems + FINE_XCDS - 1) / FINE_XCDS;
    item = (blockIdx.x % FINE_XCDS) * per + blockIdx.x / FINE_XCDS;
    return item < F.nitems;
}

__device__ inline int lane_int(int v, int l) { return __builtin_amdgcn_readlane(v, l); }
__device__ inline float lane_float(float v, int l) {
    return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), l));
}

// The draws of a canvas are taken 64 at a time: lane t reads the metadata of
// draw d0 + t (row range, operator, brush colour) and binary-searches the
// draw's steps in this row for the segment: the first step at or right of x0
// and the running sum before it.  The draws with work in the segment are then
// run in order; the next one's steps are loaded while the current one is
// composited.
#define STAGE_STOPS 64

// A non-solid draw's descriptor, stops and colours, staged for its pixels.
struct paint_lds {
    uint32_t words[sizeof(draw_desc) / 4]; // read back as a draw_desc
    float stops[STAGE_STOPS];
    float colors[4 * STAGE_STOPS];
};
static_assert(sizeof(draw_desc) % 4 == 0, "draw_desc is staged word by word");
static_assert(alignof(draw_desc) == 4, "the staged words are aligned for a draw_desc");

// Gradient terms that do not depend on the pixel (hpp:2336-2360), evaluated
// once per draw with the reference's float operations.
struct gradient_setup {
    float inv[6];
    f2 start, line;
    float span, initial, change, initial_change, initial_initial, a4, reciprocal;
    bool linear, empty;
    uint32_t n;
};

__device__ inline gradient_setup gradient_prepare(const draw_desc &D) {
    gradient_setup G;
    for (int k = 0; k < 6; ++k)
        G.inv[k] = D.paint_inverse[k];
    G.start = mk(D.start_x, D.start_y);
    G.line = sub(mk(D.end_x, D.end_y), G.start);
    G.span = dot(G.line, G.line);
    G.initial = D.start_radius;
    G.change = D.end_radius - G.initial;
    G.initial_change = G.initial * G.change;
    G.initial_initial = G.initial * G.initial;
    float a = G.span - G.change * G.change;
    G.a4 = 4.0f * a;
    G.reciprocal = 1.0f / (2.0f * a);
    G.linear = D.brush == CIH_BRUSH_LINEAR;
    G.n = D.nstops;
    G.empty = D.ncolors == 0 || G.n == 0 || (G.linear && G.span == 0.0f) ||
              (!G.linear && G.span == 0.0f && G.change == 0.0f);
    return G;
}

// paint_gradient without branches on the pixel: every lane runs the same
// instructions and the cases the reference returns early from are selected
// at the end.  index == 0 and index == n take colour 0 and n - 1 through a
// mix of 0, which leaves the colour unchanged.
__device__ __forceinline__ float4 gradient_pixel(const gradient_setup &G, const float *stops,
                                        const float *colors, f2 point) {
    f2 p = apply(G.inv, point);
    f2 relative = sub(p, G.start);
    float gradient = dot(relative, G.line);
    float offset;
    bool zero = false;
    if (G.linear) {
        offset = gradient / G.span;
    } else {
        float b = -2.0f * (gradient + G.initial_change);
        float c = dot(relative, relative) - G.initial_initial;
        float discriminant = b * b - G.a4 * c;
        float root = sqrtf(discriminant);
        float offset_1 = (-b - root) * G.reciprocal;
        float offset_2 = (-b + root) * G.reciprocal;
        float radius_1 = G.initial + G.change * offset_1;
        float radius_2 = G.initial + G.change * offset_2;
        offset = radius_2 >= 0.0f ? offset_2 : offset_1;
        zero = discriminant < 0.0f || !(radius_2 >= 0.0f || radius_1 >= 0.0f);
    }
    uint32_t n = G.n, index = 0;
    for (uint32_t k = 0; k < n; ++k)
        index += offset < stops[k] ? 0u : 1u;
    uint32_t i0 = index ? index - 1 : 0, i1 = index < n ? index : n - 1;
    float mix = (offset - stops[i0]) / (stops[i1] - stops[i0]);
    if (index == 0 || index == n)
        mix = 0.0f;
    float4 c0 = load4(colors + 4 * i0), c1 = load4(colors + 4 * i1);
    float4 c = make_float4(c0.x + (c1.x - c0.x) * mix, c0.y + (c1.y - c0.y) * mix,
                           c0.z + (c1.z - c0.z) * mix, c0.w + (c1.w - c0.w) * mix);
    float4 r = premultiplied(c);
    return zero ? make_float4(0.0f, 0.0f, 0.0f, 0.0f) : r;
}

// The segment's pixels under one draw, painted with a solid colour
// (SOLID), nothing (EMPTY) or the gradient G: every lane composites all its
// pixels and keeps the result where the pixel is inside the canvas and
// painted (coverage over the threshold, or an operator that repaints
// everything).
enum { FILL_SOLID, FILL_EMPTY, FILL_GRADIENT };

template <int KIND, bool MASKED, int PX>
__device__ __forceinline__ void composite_pixels(float4 solid, const gradient_setup &G,
                                                 const float *stops, const float *colors,
                                                 float4 (&back)[PX],
                                                 const int (&xs)[PX], int width, float py,
                                                 const float (&sums)[PX], bool evenodd,
                                                 bool clear, float galpha, int op,
                                                 const float (&vis)[PX],
                                                 unsigned &composites) {
#pragma clang loop unroll(full)
    for (int k = 0; k < PX; ++k) {
        float cov = coverage(sums[k], evenodd);
        bool on = xs[k] < width && (cov >= kThreshold || clear);
        if (MASKED)
            on = on && vis[k] >= kThreshold;
        float4 f = solid;
        if (KIND == FILL_EMPTY)
            f = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
        if (KIND == FILL_GRADIENT)
            f = gradient_pixel(G, stops, colors, mk((float)xs[k] + 0.5f, py));
        float4 b = back[k];
        composite(b, scaled(f, cov * galpha), op, MASKED ? vis[k] : 1.0f);
        float4 a = back[k];
        back[k] = make_float4(on ? b.x : a.x, on ? b.y : a.y, on ? b.z : a.z, on ? b.w : a.w);
        composites += on ? 1u : 0u;
    }
}

// One fine work item: the segment of PX * 64 pixels of one canvas row, every
// draw of the canvas applied in painter's order.  MASKED: the frame has draws
// under a clip mask (their visibility is looked up per pixel like coverage,
// hpp:2565-2604); the plain instantiation keeps the registers of the common
// case.  Clears L.mask before its first use.
template <bool MASKED, int PX>
__device__ __forceinline__ void fine_item_run(const frame &F, uint32_t item, seg_lds_t<PX> &L,
                                              paint_lds &P) {
    uint32_t c = find_canvas(F.items, (uint32_t)F.ncanvases, item);
    const canvas_desc &C = F.canvases[c];
    uint32_t local = item - F.items[c];
    int segs = (C.width + (64 * PX) - 1) / (64 * PX);
    int y = (int)(local / (uint32_t)segs);
    int x0 = (int)(local % (uint32_t)segs) * (64 * PX);
    int x1 = x0 + (64 * PX);
    int lane = threadIdx.x;
    float4 *bitmap = (float4 *)(uintptr_t)C.bitmap + (size_t)y * C.width;
    if (lane < 2 * PX)
        L.mask[lane] = 0;
    float4 back[PX];
    int xs[PX];
    for (int k = 0; k < PX; ++k) {
        xs[k] = x0 + lane * PX + k;
        back[k] = xs[k] < C.width ? bitmap[xs[k]] : make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    }
    __syncthreads();
    unsigned composites = 0;
    const step sentinel = {STEP_SENTINEL, 0.0f};
    int dend = C.first_draw + C.ndraws;
    for (int d0 = C.first_draw; d0 < dend; d0 += 64) {
        int dl = d0 + lane;
        bool need_shadow = false, need_main = false, has_steps = false;
        const step *st = nullptr;
        uint32_t count = 0, first = 0;
        float before = 0.0f;
        int op = 0, flags = 0, brush = 0, cmask = 0;
        float alpha = 0.0f;
        float4 color = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
        if (dl < dend) {
            const draw_desc &D = F.draws[dl];
            const draw_state &S = F.state[dl];
            op = D.op;
            cmask = D.clip_mask;
            flags = D.flags;
            brush = D.brush;
            alpha = D.global_alpha;
            if (D.shadow && S.width > 0) {
                // shadow layer (hpp:2504-2538): rows top <= y + border < bottom,
                // columns [left - border, right - border) within the canvas
                int sy = y + D.border - S.top;
                int lo = max(0, S.left - D.border);
                int hi = min(C.width, S.left + S.width - D.border);
                need_shadow = sy >= 0 && sy < S.height && lo < hi && lo < x1 && hi > x0;
            }
            if (S.rows > 0 && y >= S.row0 && y < S.row0 + S.rows) {
                const uint32_t *rowoff = F.words + S.rowoff + (y - S.row0);
                uint32_t begin = rowoff[0];
                count = rowoff[1] - begin;
                st = F.steps + S.runs + begin;
            }
            if (count) {
                uint32_t lo = 0, hi = count;
                while (lo < hi) {
                    uint32_t mid = (lo + hi) >> 1;
                    if (st[mid].x < x0)
                        lo = mid + 1;
                    else
                        hi = mid;
                }
                first = lo;
                has_steps = lo < count && st[lo].x < x1;
                before = lo > 0 ? st[lo - 1].sum : 0.0f;
            }
            // Pixels of the row get |running sum| as coverage up to the
            // canvas edge (the clip mask's last run, hpp:2636-2642, ends each
            // row); operators without bit 8 repaint every pixel (hpp:2574).
            need_main = !(op & 8) || has_steps ||
                        coverage(before, (flags & CIH_DRAW_FLAG_EVENODD) != 0) >= kThreshold;
            if (D.kind == CIH_DRAW_CLIP || D.kind == CIH_DRAW_PUT_IMAGE) // k_clip / k_put
                need_main = need_shadow = false;
            if (D.brush == CIH_BRUSH_COLOR && D.ncolors)
                color = load4((const float *)(F.payload + D.colors_at));
        }
        unsigned long long todo = __ballot(need_shadow || need_main);
        unsigned long long stepped = __ballot(has_steps);
        auto fetch = [&](unsigned long long rest) {
            step s = sentinel;
            if (rest) {
                int j = __ffsll((long long)rest) - 1;
                if ((stepped >> j) & 1) {
                    const step *sj = (const step *)(((uint64_t)(uint32_t)lane_int(
                                                          (int)((uintptr_t)st >> 32), j) << 32) |
                                                     (uint32_t)lane_int((int)(uintptr_t)st, j));
                    uint32_t at = (uint32_t)lane_int((int)first, j) + lane;
                    if (at < (uint32_t)lane_int((int)count, j))
                        s = sj[at];
                }
            }
            return s;
        };
        step next = fetch(todo);
        while (todo) {
            int j = __ffsll((long long)todo) - 1;
            todo &= todo - 1;
            step cur = next;
            next = fetch(todo);
            int d = d0 + j;
            int dop = lane_int(op, j);
            // visibility under the draw's clip mask: min(|running sum of the
            // mask's pieces at or left of the pixel|, 1)
            float vis[PX];
            for (int k = 0; k < PX; ++k)
                vis[k] = 1.0f;
            if (MASKED) {
                int m = lane_int(cmask, j);
                if (m) {
                    const mask_desc &M = ((const mask_desc *)(uintptr_t)C.masks)[m];
                    uint32_t mn = ((const uint32_t *)(uintptr_t)M.counts)[y];
                    float vb = 0.0f, vsums[PX];
                    bool vin = false;
                    if (mn)
                        vin = segment_sums((const step *)(uintptr_t)M.steps +
                                               (size_t)y * (size_t)(C.width + 2),
                                           mn, x0, L, vb, vsums);
                    bool any = false;
                    for (int k = 0; k < PX; ++k) {
                        vis[k] = smin(fabsf(vin ? vsums[k] : vb), 1.0f);
                        any = any || vis[k] >= kThreshold;
                    }
                    if (!__ballot(any))
                        continue; // nothing of this draw is visible in the segment
                }
            }
            if (lane_int((int)need_shadow, j)) {
                const draw_desc &D = F.draws[d];
                const draw_state &S = F.state[d];
                const float *row =
                    F.shadow + S.image + (size_t)(y + D.border - S.top) * S.width;
                int lo = max(0, S.left - D.border);
                int hi = min(C.width, S.left + S.width - D.border);
                float4 scolor = load4(D.shadow_color);
                for (int k = 0; k < PX; ++k) {
                    int x = xs[k];
                    if (x < lo || x >= hi)
                        continue;
                    float sv = row[x + D.border - S.left];
                    if (MASKED && !(vis[k] >= kThreshold))
                        continue;
                    composite(back[k], scaled(scolor, D.global_alpha * sv), dop,
                              MASKED ? vis[k] : 1.0f);
                    ++composites;
                }
            }
            if (!lane_int((int)need_main, j))
                continue;
            bool clear = !(dop & 8);
            bool evenodd = (lane_int(flags, j) & CIH_DRAW_FLAG_EVENODD) != 0;
            float sums[PX];
            float bj = lane_float(before, j);
            bool inside = false;
            if ((stepped >> j) & 1) {
                const step *sj = (const step *)(((uint64_t)(uint32_t)lane_int(
                                                     (int)((uintptr_t)st >> 32), j) << 32) |
                                                (uint32_t)lane_int((int)(uintptr_t)st, j));
                inside = segment_fill(sj, (uint32_t)lane_int((int)count, j),
                                      (uint32_t)lane_int((int)first, j), cur, x0, L, bj, sums);
            }
            if (!inside) {
                if (!clear && !(coverage(bj, evenodd) >= kThreshold))
                    continue;
                for (int k = 0; k < PX; ++k)
                    sums[k] = bj;
            }
            float galpha = lane_float(alpha, j);
            int dbrush = lane_int(brush, j);
            if (dbrush == CIH_BRUSH_COLOR) {
                float4 p = make_float4(lane_float(color.x, j), lane_float(color.y, j),
                                       lane_float(color.z, j), lane_float(color.w, j));
                gradient_setup none;
                composite_pixels<FILL_SOLID, MASKED>(p, none, nullptr, nullptr, back, xs,
                                                     C.width, 0.0f, sums, evenodd, clear,
                                                     galpha, dop, vis, composites);
                continue;
            }
            // Gradients and patterns: the draw's descriptor (and a gradient's
            // stops and colours) staged in LDS once for the segment's pixels.
            const draw_desc *Dg = F.draws + d;
            wave_lds_sync(); // the previous draw's reads of P are done
            const uint32_t *src = (const uint32_t *)Dg;
            for (int i = lane; i < (int)(sizeof(draw_desc) / 4); i += 64)
                P.words[i] = src[i];
            uint32_t nstops = Dg->nstops, ncolors = Dg->ncolors;
            bool staged = nstops <= STAGE_STOPS && ncolors <= STAGE_STOPS;
            const float *gstops = (const float *)(F.payload + Dg->stops_at);
            const float *gcolors = (const float *)(F.payload + Dg->colors_at);
            if (staged) {
                for (uint32_t i = lane; i < nstops; i += 64)
                    P.stops[i] = gstops[i];
                for (uint32_t i = lane; i < 4 * ncolors; i += 64)
                    P.colors[i] = gcolors[i];
            }
            wave_lds_sync();
            const draw_desc &D = *(const draw_desc *)P.words;
            float py = (float)y + 0.5f;
            if (dbrush == CIH_BRUSH_LINEAR || dbrush == CIH_BRUSH_RADIAL) {
                gradient_setup G = gradient_prepare(D);
                float4 zero4 = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
                if (G.empty)
                    composite_pixels<FILL_EMPTY, MASKED>(zero4, G, nullptr, nullptr, back, xs,
                                                         C.width, py, sums, evenodd, clear,
                                                         galpha, dop, vis, composites);
                else if (staged)
                    composite_pixels<FILL_GRADIENT, MASKED>(zero4, G, P.stops, P.colors, back,
                                                            xs, C.width, py, sums, evenodd,
                                                            clear, galpha, dop, vis, composites);
                else
                    composite_pixels<FILL_GRADIENT, MASKED>(zero4, G, gstops, gcolors, back, xs,
                                                            C.width, py, sums, evenodd, clear,
                                                            galpha, dop, vis, composites);
                continue;
            }
            // patterns: bicubic taps, only for the pixels that are painted
#pragma clang loop unroll(full)
            for (int k = 0; k < PX; ++k) {
                float cov = coverage(sums[k], evenodd);
                if (xs[k] < C.width && (cov >= kThreshold || clear) &&
                    (!MASKED || vis[k] >= kThreshold)) {
                    float4 p = paint(F, D, mk((float)xs[k] + 0.5f, py));
                    composite(back[k], scaled(p, cov * galpha), dop, MASKED ? vis[k] : 1.0f);
                    ++composites;
                }
            }
        }
    }
    unsigned long long total = composites;
    for (int w = 32; w >= 1; w >>= 1)
        total += __shfl_xor(total, w, 64);
    if (lane == 0 && total)
        atomicAdd(&F.counters[COUNTER_PIXEL_DRAWS], total);
    for (int k = 0; k < PX; ++k)
        if (xs[k] < C.width)
            bitmap[xs[k]] = back[k];
}


// The fine kernel.  One workgroup (one wave) per work item, or, PERSISTENT,
// a fixed grid whose waves take items from eight queues until all are empty:
// the items are split into eight contiguous ranges, one per XCD (a workgroup
// runs on XCD blockIdx.x % 8), so the segments of a row that read the same
// steps share an L2; a wave whose own range is empty takes items from the
// next ranges.  Each range's cursor is a frame counter cleared by k_reset.
template <bool MASKED, int PX, bool PERSISTENT>
__global__ __launch_bounds__(FINE_THREADS) void k_fine(frame F) {
    __shared__ seg_lds_t<PX> L;
    __shared__ paint_lds P;
    if (!PERSISTENT) {
        uint32_t item;
        if (!fine_item(F, item) || overflowed(F))
            return;
        fine_item_run<MASKED, PX>(F, item, L, P);
        return;
    }
    if (overflowed(F))
        return;
    uint32_t per = (F.nitems + FINE_XCDS - 1) / FINE_XCDS;
    int xcd = (int)(blockIdx.x % FINE_XCDS);
    for (int k = 0; k < FINE_XCDS; ++k) {
        int q = (xcd + k) % FINE_XCDS;
        uint32_t lo = (uint32_t)q * per;
        uint32_t hi = lo + per < F.nitems ? lo + per : F.nitems;
        if (lo >= hi)
            continue;
        for (;;) {
            uint32_t i = 0;
            if (threadIdx.x == 0)
                i = (uint32_t)atomicAdd(&F.counters[COUNTER_FINE_QUEUE + q], 1ull);
            i = __shfl(i, 0, 64);
            if (i >= hi - lo)
                break;
            fine_item_run<MASKED, PX>(F, lo + i, L, P);
        }
    }
}

// ---------------------------------------------------------------------------
// Small kernels.

__global__ void k_reset(unsigned long long *counters) {
    if (threadIdx.x < COUNTER_COUNT)
        counters[threadIdx.x] = 0;
}

__global__ void k_clear(float4 *pixels, size_t n) {
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    for (; i < n; i += (size_t)gridDim.x * blockDim.x)
        pixels[i] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
}

// RGBA8 -> premultiplied linear fp32 through the sRGB table (hpp:2852-2860,
// hpp:3402-3406).
__device__ inline float4 texel_from_rgba8(uint32_t v, const float *lut) {
    float a = (float)(v >> 24) / 255.0f;
    return make_float4(lut[v & 255] * a, lut[(v >> 8) & 255] * a, lut[(v >> 16) & 255] * a, a);
}

__global__ void k_prepare_image(const uint32_t *rgba8, float4 *texels, size_t n,
                                const float *lut) {
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    for (; i < n; i += (size_t)gridDim.x * blockDim.x)
        texels[i] = texel_from_rgba8(rgba8[i], lut);
}

__global__ void k_put_image(const uint32_t *rgba8, int w, int h, float4 *bitmap, int W, int H,
                            int x, int y, const float *lut) {
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    size_t n = (size_t)w * (size_t)h;
    for (; i < n; i += (size_t)gridDim.x * blockDim.x) {
        int ix = (int)(i % (size_t)w), iy = (int)(i / (size_t)w);
        int cx = x + ix, cy = y + iy;
        if (cx < 0 || cx >= W || cy < 0 || cy >= H)
            continue;
        bitmap[(size_t)cy * W + cx] = texel_from_rgba8(rgba8[i], lut);
    }
}

__device__ inline float clamp01(float v) { return smin(smax(v, 0.0f), 1.0f); }

// get_image_data (hpp:3363-3380) into a dense w x h RGBA8 buffer.
__global__ void k_readback(const float4 *bitmap, int W, int H, uint32_t *out, int w, int h,
                           int x, int y) {
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    size_t n = (size_t)w * (size_t)h;
    for (; i < n; i += (size_t)gridDim.x * blockDim.x) {
        int ix = (int)(i % (size_t)w), iy = (int)(i / (size_t)w);
        int cx = x + ix, cy = y + iy;
        float4 c = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
        if (0 <= cx && cx < W && 0 <= cy && cy < H)
            c = bitmap[(size_t)cy * W + cx];
        float t = c_bayer[(cy & 3) * 4 + (cx & 3)];
        if (c.w < 1.0f / 8160.0f)
            c = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
        else {
            float inv = 1.0f / c.w;
            c = make_float4(inv * c.x, inv * c.y, inv * c.z, c.w);
        }
        float r = t + delinearized(clamp01(c.x)) * 255.0f;
        float g = t + delinearized(clamp01(c.y)) * 255.0f;
        float b = t + delinearized(clamp01(c.z)) * 255.0f;
        float a = t + clamp01(c.w) * 255.0f;
        out[i] = (uint32_t)(unsigned char)(int)r | ((uint32_t)(unsigned char)(int)g << 8) |
                 ((uint32_t)(unsigned char)(int)b << 16) | ((uint32_t)(unsigned char)(int)a << 24);
    }
}

} // namespace cih

// ---------------------------------------------------------------------------
// Launchers used by the host runtime.

namespace cih {


static unsigned grid_for(size_t n, unsigned block) {
    size_t g = (n + block - 1) / block;
    return (unsigned)(g < 65536 ? (g ? g : 1) : 65536);
}

// put_image_data of draw d (hpp:3383-3408): the image's texels (already
// premultiplied linear, as put_image_data converts them) replace the canvas
// pixels they land on.  A grid-stride loop over the image.
__global__ void k_put(frame F, int d) {
    const draw_desc &D = F.draws[d];
    const canvas_desc &C = F.canvases[D.canvas];
    const image_desc &I = F.images[D.image];
    const float4 *texels = (const float4 *)(uintptr_t)I.texels;
    float4 *bitmap = (float4 *)(uintptr_t)C.bitmap;
    const int w = D.image_width, h = D.image_height;
    const long x0 = (long)D.start_x, y0 = (long)D.start_y;
    const size_t n = (size_t)w * (size_t)h;
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (size_t)gridDim.x * blockDim.x) {
        long cx = x0 + (long)(i % (size_t)w), cy = y0 + (long)(i / (size_t)w);
        if (cx < 0 || cx >= C.width || cy < 0 || cy >= C.height)
            continue;
        bitmap[(size_t)cy * (size_t)C.width + (size_t)cx] = texels[i];
    }
}

const int kFrameKernels = 8;
extern const char *const kFrameKernelNames[kFrameKernels] = {
    "k_reset", "k_geometry", "k_runs", "k_coalesce", "k_shadow_raster", "k_blur_rows",
    "k_blur_cols", "k_fine"};

// Launches a frame's kernels on `stream`.  With `events` (kFrameKernels + 1
// of them), event k is recorded before kernel k and the last after the last
// kernel, so the time of kernel k is elapsed(events[k], events[k + 1]).
hipError_t launch_frame(const frame &F, int shadow_rows, int shadow_cols, bool any_shadow,
                        const int32_t *clips, int nclips, const int32_t *puts, int nputs,
                        bool masked, hipEvent_t *events, hipStream_t stream) {
#define MARK(k)                                                                                \
    do {                                                                                       \
        if (events)                                                                            \
            (void)hipEventRecord(events[k], stream);                                           \
    } while (0)
    MARK(0);
    k_reset<<<1, 64, 0, stream>>>(F.counters);
    // puts in draw order (puts[2k] = draw, puts[2k + 1] = texels): a frame
    // holds no painting draw of a canvas before its puts
    for (int k = 0; k < nputs; ++k)
        k_put<<<grid_for((size_t)puts[2 * k + 1], 256), 256, 0, stream>>>(F, puts[2 * k]);
    MARK(1);
    if (F.ndraws > 0)
        k_geometry<<<F.ndraws, GEOM_THREADS, 0, stream>>>(F);
    MARK(2);
    if (F.ndraws > 0)
        k_runs<<<(any_shadow ? 2 : 1) * F.ndraws, GEO_THREADS, 0, stream>>>(F, any_shadow ? 2 : 1);
    MARK(3);
    if (F.ndraws > 0) {
        k_coalesce<<<COALESCE_GRID, 64, 0, stream>>>(F);
        k_coalesce_wide<<<COALESCE_GRID, 64, 0, stream>>>(F);
    }
    // clip masks in draw order (clips[2k] = draw, clips[2k + 1] = canvas height)
    for (int k = 0; k < nclips; ++k) {
        k_clip<<<(clips[2 * k + 1] + 63) / 64, 64, 0, stream>>>(F, clips[2 * k]);
        k_clip_finish<<<1, 64, 0, stream>>>(F, clips[2 * k]);
    }
    MARK(4);
    int ry = (shadow_rows + 63) / 64, cx = (shadow_cols + 63) / 64;
    if (F.ndraws > 0 && any_shadow)
        k_shadow_raster<<<dim3(F.ndraws, 64), FINE_THREADS, 0, stream>>>(F);
    MARK(5);
    if (F.ndraws > 0 && any_shadow) { // the row passes run between two transposes
        k_shadow_transpose<<<dim3(F.ndraws, 32), 256, 0, stream>>>(F, 1);
        k_blur_rows<<<dim3(F.ndraws, ry > 0 ? ry : 1), 64, 0, stream>>>(F);
        k_shadow_transpose<<<dim3(F.ndraws, 32), 256, 0, stream>>>(F, 0);
    }
    MARK(6);
    if (F.ndraws > 0 && any_shadow)
        k_blur_cols<<<dim3(F.ndraws, cx > 0 ? cx : 1), 64, 0, stream>>>(F);
    MARK(7);
    if (F.nitems) {
        // the host gives frames with masked draws the default configuration
        // (256-pixel items, one workgroup each): the masked variant has only it
        unsigned grid = FINE_XCDS * ((F.nitems + FINE_XCDS - 1) / FINE_XCDS);
        if (F.fine_persistent && grid > FINE_PERSISTENT_GRID)
            grid = FINE_PERSISTENT_GRID;
        int px = F.segment / FINE_THREADS;
        bool persistent = F.fine_persistent != 0;
        if (masked && (px != FINE_PX || persistent))
            return hipErrorInvalidValue; // cannot happen: run_kernels sees to it
        if (masked)
            k_fine<true, FINE_PX, false><<<grid, FINE_THREADS, 0, stream>>>(F);
        else if (px == 1 && persistent)
            k_fine<false, 1, true><<<grid, FINE_THREADS, 0, stream>>>(F);
        else if (px == 1)
            k_fine<false, 1, false><<<grid, FINE_THREADS, 0, stream>>>(F);
        else if (px == 2 && persistent)
            k_fine<false, 2, true><<<grid, FINE_THREADS, 0, stream>>>(F);
        else if (px == 2)
            k_fine<false, 2, false><<<grid, FINE_THREADS, 0, stream>>>(F);
        else if (persistent)
            k_fine<false, 4, true><<<grid, FINE_THREADS, 0, stream>>>(F);
        else
            k_fine<false, 4, false><<<grid, FINE_THREADS, 0, stream>>>(F);
    }
    MARK(8);
#undef MARK
    return hipGetLastError();
}

hipError_t launch_clear(void *pixels, size_t n, hipStream_t stream) {
    k_clear<<<grid_for(n, 256), 256, 0, stream>>>((float4 *)pixels, n);
    return hipGetLastError();
}

hipError_t launch_prepare_image(const void *rgba8, void *texels, size_t n, const float *lut,
                                hipStream_t stream) {
    k_prepare_image<<<grid_for(n, 256), 256, 0, stream>>>((const uint32_t *)rgba8,
                                                          (float4 *)texels, n, lut);
    return hipGetLastError();
}

hipError_t launch_put_image(const void *rgba8, int w, int h, void *bitmap, int W, int H, int x,
                            int y, const float *lut, hipStream_t stream) {
    k_put_image<<<grid_for((size_t)w * h, 256), 256, 0, stream>>>(
        (const uint32_t *)rgba8, w, h, (float4 *)bitmap, W, H, x, y, lut);
    return hipGetLastError();
}

hipError_t launch_readback(const void *bitmap, int W, int H, void *out, int w, int h, int x,
                           int y, hipStream_t stream) {
    k_readback<<<grid_for((size_t)w * h, 256), 256, 0, stream>>>(
        (const float4 *)bitmap, W, H, (uint32_t *)out, w, h, x, y);
    return hipGetLastError();
}

} // namespace cih

// ---------------------------------------------------------------------------
// Batched read-back and clear: every canvas of a batch, one launch each.

namespace cih {

__device__ inline int canvas_of(const uint64_t *prefix, int n, uint64_t i) {
    int lo = 0, hi = n; // prefix[lo] <= i < prefix[lo + 1]
    while (hi - lo > 1) {
        int mid = (lo + hi) >> 1;
        if (prefix[mid] <= i)
            lo = mid;
        else
            hi = mid;
    }
    return lo;
}

__global__ void k_readback_batch(const canvas_desc *canvases, const uint64_t *pixel_prefix,
                                 int ncanvases, uint32_t *out) {
    uint64_t total = pixel_prefix[ncanvases];
    uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    for (; i < total; i += (uint64_t)gridDim.x * blockDim.x) {
        int k = canvas_of(pixel_prefix, ncanvases, i);
        const canvas_desc &C = canvases[k];
        uint64_t local = i - pixel_prefix[k];
        int cx = (int)(local % (uint64_t)C.width), cy = (int)(local / (uint64_t)C.width);
        float4 c = ((const float4 *)(uintptr_t)C.bitmap)[local];
        float t = c_bayer[(cy & 3) * 4 + (cx & 3)];
        if (c.w < 1.0f / 8160.0f)
            c = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
        else {
            float inv = 1.0f / c.w;
            c = make_float4(inv * c.x, inv * c.y, inv * c.z, c.w);
        }
        float r = t + delinearized(clamp01(c.x)) * 255.0f;
        float g = t + delinearized(clamp01(c.y)) * 255.0f;
        float b = t + delinearized(clamp01(c.z)) * 255.0f;
        float a = t + clamp01(c.w) * 255.0f;
        out[i] = (uint32_t)(unsigned char)(int)r | ((uint32_t)(unsigned char)(int)g << 8) |
                 ((uint32_t)(unsigned char)(int)b << 16) | ((uint32_t)(unsigned char)(int)a << 24);
    }
}

__global__ void k_clear_batch(const canvas_desc *canvases, const uint64_t *pixel_prefix,
                              int ncanvases) {
    uint64_t total = pixel_prefix[ncanvases];
    uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    for (; i < total; i += (uint64_t)gridDim.x * blockDim.x) {
        int k = canvas_of(pixel_prefix, ncanvases, i);
        ((float4 *)(uintptr_t)canvases[k].bitmap)[i - pixel_prefix[k]] =
            make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    }
}

hipError_t launch_readback_batch(const void *canvases, const void *pixel_prefix, int ncanvases,
                                 uint64_t total, void *out, hipStream_t stream) {
    k_readback_batch<<<grid_for(total, 256), 256, 0, stream>>>(
        (const canvas_desc *)canvases, (const uint64_t *)pixel_prefix, ncanvases,
        (uint32_t *)out);
    return hipGetLastError();
}

hipError_t launch_clear_batch(const void *canvases, const void *pixel_prefix, int ncanvases,
                              uint64_t total, hipStream_t stream) {
    k_clear_batch<<<grid_for(total, 256), 256, 0, stream>>>(
        (const canvas_desc *)canvases, (const uint64_t *)pixel_prefix, ncanvases);
    return hipGetLastError();
}

} // namespace cih
