// cih_kernels.hip -- the MI355X (gfx950) kernels of the canvas pipeline.
//
// Per frame (a list of draws over one or more canvases):
//   k_reset        clears the arena counters (a kernel, so frames can be
//                  captured in HIP graphs without memset nodes)
//   k_geometry     one workgroup per draw: flatten (one thread per cubic),
//                  dash and stroke (one thread per subpath) -> closed polygons
//   k_runs         one workgroup per draw and pass (main, shadow): clip
//                  (hpp:2205-2238), signed-area runs (hpp:2109-2170) bucketed
//                  per pixel row; the shadow pass also sizes its image
//   k_shadow_*     shadow alpha raster, then the extended-box blur
//                  (hpp:2453-2503) with one thread per row / per column
//   k_fine         one workgroup per (canvas, row, segment): the fp32 pixels
//                  stay in registers while every draw of the canvas is
//                  applied in painter's order -- row prefix sums of the runs
//                  in 64-bit fixed point (exact, order independent), paint
//                  (hpp:2265-2377) and composite (hpp:2576-2592)
//   k_readback     get_image_data's unpremultiply/sRGB/dither (hpp:3363-3380)
// None of this is a dense contraction: the pipeline is VALU/LDS and HBM work.
#include <hip/hip_runtime.h>

#include "cih_device.hpp"
#include "cih_geometry.hpp"
#include "cih_abi.h"

namespace cih {

#define GEO_THREADS 256
#define FINE_THREADS 128
#define FINE_PX 2
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

// Bump allocation from an arena; returns ~0 (and flags overflow) when full.
__device__ inline uint64_t arena_alloc(const frame &F, int which, uint64_t n, uint64_t cap) {
    unsigned long long at = atomicAdd(&F.counters[which], (unsigned long long)n);
    if (at + n > cap) {
        atomicMax(&F.counters[COUNTER_OVERFLOW], 1ull);
        return ~0ull;
    }
    return at;
}

// Exclusive scan of one value per thread over a GEO_THREADS workgroup.
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
    // point and cubic prefixes over the subpaths
    uint64_t w = block_alloc(F, COUNTER_WORDS, 2 * (uint64_t)(nsub + 1), F.words_cap, slot);
    if (w == ~0ull)
        return out;
    uint32_t *sp_pts = F.words + w;
    uint32_t *sp_cub = sp_pts + nsub + 1;
    for (uint32_t s = threadIdx.x; s < nsub; s += blockDim.x) {
        uint32_t c = in_sub[2 * s];
        sp_pts[s] = c;
        sp_cub[s] = c >= 1 ? (c - 1) / 3 : 0;
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
        uint32_t base = sp_pts[s] + 3 * (g - sp_cub[s]);
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
        uint32_t base = sp_pts[s] + 3 * (g - sp_cub[s]);
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

// Dash (mode 0) or stroke (mode 1) every subpath of `in`, one thread each.
__device__ line_path per_subpath(const frame &F, const draw_desc &D, line_path in, int mode,
                                 uint32_t *lds, uint64_t *slot) {
    line_path out = {0, 0, 0, 0};
    uint32_t nsub = in.nsub;
    uint64_t w = block_alloc(F, COUNTER_WORDS, 3 * (uint64_t)(nsub + 1), F.words_cap, slot);
    if (w == ~0ull)
        return out;
    uint32_t *at = F.words + w; // point offset of each input subpath
    uint32_t *np = at + nsub + 1;
    uint32_t *ns = np + nsub + 1;
    const float *dash = (const float *)(F.payload + D.dash_at);
    dash_setup setup = {0, 0.0f};
    if (mode == 0)
        setup = dash_prepare(dash, (int)D.ndash, D.dash_offset);
    for (uint32_t s = threadIdx.x; s < nsub; s += blockDim.x)
        at[s] = in.sub[2 * s];
    __syncthreads();
    scan_array(at, nsub, lds);
    for (uint32_t s = threadIdx.x; s < nsub; s += blockDim.x) {
        count_points c;
        const f2 *pts = in.pts + at[s];
        uint32_t count = in.sub[2 * s];
        bool closed = in.sub[2 * s + 1] != 0;
        if (mode == 0)
            dash_subpath(c, pts, count, closed, dash, (int)D.ndash, setup, D.inverse);
        else
            stroke_subpath(c, pts, count, closed, D.line_width, D.miter_limit, D.line_join,
                           D.line_cap, D.forward, D.inverse);
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
        if (mode == 0)
            dash_subpath(wp, src, count, closed, dash, (int)D.ndash, setup, D.inverse);
        else
            stroke_subpath(wp, src, count, closed, D.line_width, D.miter_limit, D.line_join,
                           D.line_cap, D.forward, D.inverse);
        wp.finish();
    }
    __syncthreads();
    out.pts = pts;
    out.sub = sub;
    out.npts = tp;
    out.nsub = ts;
    return out;
}

__global__ __launch_bounds__(GEO_THREADS) void k_geometry(frame F) {
    __shared__ uint32_t lds[16];
    __shared__ uint64_t slot[1];
    int d = blockIdx.x;
    if (d >= F.ndraws || overflowed(F))
        return;
    const draw_desc &D = F.draws[d];
    line_path lp;
    bool stroke = D.kind == CIH_DRAW_STROKE_PATH || D.kind == CIH_DRAW_STROKE_LINES;
    draw_state &S = F.state[d];
    if (D.kind == CIH_DRAW_FILL_PATH || D.kind == CIH_DRAW_STROKE_PATH) {
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
            lp = per_subpath(F, D, lp, 0, lds, slot);
            if (!lp.pts)
                return;
            if (threadIdx.x == 0) {
                S.dash_points = (uint64_t)(uintptr_t)lp.pts;
                S.dash_counts = (uint64_t)(uintptr_t)lp.sub;
                S.dash_npoints = lp.npts;
                S.dash_nsub = lp.nsub;
            }
        }
        lp = per_subpath(F, D, lp, 1, lds, slot);
        if (!lp.pts)
            return;
    }
    if (threadIdx.x == 0) {
        S.poly_points = (uint64_t)(uintptr_t)lp.pts;
        S.poly_counts = (uint64_t)(uintptr_t)lp.sub;
        S.poly_npoints = lp.npts;
        S.poly_nsub = lp.nsub;
    }
}


// ---------------------------------------------------------------------------
// k_runs: polygons -> clipped edges -> runs bucketed by pixel row.
//
// Sinks drop runs outside the row range or the 16-bit x range, which only
// malformed (non-finite) geometry can produce; counting and writing apply the
// same test so bucket sizes always match.

struct row_counter {
    uint32_t *rows;
    int row0, nrows;
    int y;
    uint32_t n;
    __device__ void flush() {
        if (n)
            atomicAdd(&rows[y - row0], n);
        n = 0;
    }
    __device__ void run(int x, int yy, float) {
        if (yy - row0 < 0 || yy - row0 >= nrows || x < 0 || x > 0xffff)
            return;
        if (yy != y) {
            flush();
            y = yy;
        }
        ++n;
    }
};

struct row_writer {
    uint32_t *cursor;
    int row0, nrows;
    struct run *out;
    int minx, maxx;
    int nzminx, nzmaxx, nzminy, nzmaxy;
    unsigned long long minkey;
    int y, n;
    int bx[8];
    float bd[8];
    __device__ void flush() {
        if (!n)
            return;
        uint32_t at = atomicAdd(&cursor[y - row0], (uint32_t)n);
        for (int k = 0; k < n; ++k) {
            struct run r;
            r.x = bx[k];
            r.delta = bd[k];
            out[at + k] = r;
        }
        n = 0;
    }
    __device__ void run(int x, int yy, float delta) {
        if (yy - row0 < 0 || yy - row0 >= nrows || x < 0 || x > 0xffff)
            return;
        if (yy != y || n == 8) {
            flush();
            y = yy;
        }
        bx[n] = x;
        bd[n] = delta;
        ++n;
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

__global__ __launch_bounds__(GEO_THREADS) void k_runs(frame F) {
    __shared__ uint32_t lds[16];
    __shared__ uint64_t slot[1];
    __shared__ int s_ymin, s_ymax;
    __shared__ uint32_t s_rows[LDS_ROWS];
    __shared__ int s_b[6];
    __shared__ unsigned long long s_key;
    int d = blockIdx.x >> 1;
    int variant = blockIdx.x & 1;
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
    for (int r = threadIdx.x; r < rows; r += blockDim.x)
        counter[r] = 0;
    __syncthreads();
    for (uint32_t e = threadIdx.x; e < nedges; e += blockDim.x) {
        f2 from, to;
        src.get(e, from, to);
        row_counter rc = {counter, row0, rows, -1, 0};
        add_runs(rc, clamp_point(from, width, height), clamp_point(to, width, height));
        rc.flush();
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
    row_writer wr;
    wr.cursor = counter;
    wr.row0 = row0;
    wr.nrows = rows;
    wr.out = F.runs + rb;
    wr.minx = 0x7fffffff, wr.maxx = -1;
    wr.nzminx = 0x7fffffff, wr.nzmaxx = -1, wr.nzminy = 0x7fffffff, wr.nzmaxy = -1;
    wr.minkey = ~0ull;
    wr.y = -1, wr.n = 0;
    for (uint32_t e = threadIdx.x; e < nedges; e += blockDim.x) {
        f2 from, to;
        src.get(e, from, to);
        add_runs(wr, clamp_point(from, width, height), clamp_point(to, width, height));
        wr.flush();
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
        S.row0 = row0;
        S.rows = rows;
        S.rowoff = ro;
        S.runs = rb;
        S.minx = s_b[0];
        S.maxx = s_b[1];
        return;
    }
    S.srow0 = row0;
    S.srows = rows;
    S.srowoff = ro;
    S.sruns = rb;
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
    }
    S.image = img;
    S.image2 = img + size;
}

// ---------------------------------------------------------------------------
// Row-segment coverage shared by the shadow raster and the fine pass.
//
// A workgroup of FINE_THREADS owns FINE_SEGMENT pixels [x0, x0 + SEGMENT) of
// one row; thread t owns FINE_PX consecutive pixels.  Coverage at pixel x is
// |sum of run deltas at x' <= x| (hpp:2565-2575): deltas left of the segment
// form the backdrop, those inside are added into LDS in 2^-32 fixed point and
// prefix-summed.  Integer sums are exact and independent of the order in
// which threads add them.

struct seg_lds {
    long long acc[FINE_SEGMENT];
    long long wave[FINE_THREADS / 64];
    int minx, maxx;
};

__device__ inline long long quantize(float delta) {
    return __float2ll_rn(delta * 4294967296.0f);
}

__device__ inline float coverage(long long q, bool evenodd) {
    if (evenodd) {
        // extension: triangle wave of the winding area, period 2
        unsigned long long a = (unsigned long long)(q < 0 ? -q : q) & ((1ull << 33) - 1);
        if (a > (1ull << 32))
            a = (1ull << 33) - a;
        return smin((float)a * 2.3283064365386963e-10f, 1.0f);
    }
    float s = (float)q * 2.3283064365386963e-10f;
    return smin(fabsf(s), 1.0f);
}

__device__ inline long long wave_sum64(long long v) {
    for (int d = 32; d >= 1; d >>= 1)
        v += __shfl_xor(v, d, 64);
    return v;
}

// Fills q[k] with the running sum at pixel x0 + threadIdx.x*FINE_PX + k and
// returns the row's first/last run x in rmin/rmax.  Uniform call.
__device__ void segment_coverage(const run *runs, uint32_t begin, uint32_t end, int x0,
                                 seg_lds &L, long long q[FINE_PX], int &rmin, int &rmax) {
    int t = threadIdx.x;
    long long back = 0;
    int mn = 0x7fffffff, mx = -1;
    for (uint32_t i = begin + t; i < end; i += FINE_THREADS) {
        run r = runs[i];
        long long v = quantize(r.delta);
        mn = min(mn, r.x);
        mx = max(mx, r.x);
        int rel = r.x - x0;
        if (rel < 0)
            back += v;
        else if (rel < FINE_SEGMENT)
            atomicAdd((unsigned long long *)&L.acc[rel], (unsigned long long)v);
    }
    back = wave_sum64(back);
    for (int d = 32; d >= 1; d >>= 1) {
        mn = min(mn, __shfl_xor(mn, d, 64));
        mx = max(mx, __shfl_xor(mx, d, 64));
    }
    if ((t & 63) == 0) {
        L.wave[t >> 6] = back;
        atomicMin(&L.minx, mn);
        atomicMax(&L.maxx, mx);
    }
    __syncthreads();
    long long backdrop = 0;
    for (int w = 0; w < FINE_THREADS / 64; ++w)
        backdrop += L.wave[w];
    rmin = L.minx;
    rmax = L.maxx;
    long long local = 0;
    for (int k = 0; k < FINE_PX; ++k) {
        local += L.acc[t * FINE_PX + k];
        q[k] = local;
        L.acc[t * FINE_PX + k] = 0;
    }
    // exclusive scan of the per-thread totals
    int lane = t & 63, wave = t >> 6;
    long long incl = local;
    for (int d = 1; d < 64; d <<= 1) {
        long long n = __shfl_up(incl, d, 64);
        if (lane >= d)
            incl += n;
    }
    __syncthreads(); // everyone has read L.wave / minx / maxx
    if (lane == 63)
        L.wave[wave] = incl;
    if (t == 0) {
        L.minx = 0x7fffffff;
        L.maxx = -1;
    }
    __syncthreads();
    long long before = backdrop + incl - local;
    for (int w = 0; w < wave; ++w)
        before += L.wave[w];
    for (int k = 0; k < FINE_PX; ++k)
        q[k] += before;
    __syncthreads(); // L.wave free for the next call
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
    for (int py = top; py < bottom; ++py) {
        float y = fabsf(reciprocal_y * ((float)py - point.y));
        float weight_y = cubic_weight(y);
        int wy = py % bh;
        if (wy < 0)
            wy += bh;
        if (clamp)
            wy = min(max(py, 0), bh - 1);
        const float4 *row = texels + (size_t)wy * (size_t)bw;
        for (int px = left; px < right; ++px) {
            float x = fabsf(reciprocal_x * ((float)px - point.x));
            float weight_x = cubic_weight(x);
            int wx = px % bw;
            if (wx < 0)
                wx += bw;
            if (clamp)
                wx = min(max(px, 0), bw - 1);
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

__device__ float4 paint_gradient(const frame &F, const draw_desc &D, f2 point) {
    const float4 zero = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    const float *colors = (const float *)(F.payload + D.colors_at);
    const float *stops = (const float *)(F.payload + D.stops_at);
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
    uint32_t n = D.nstops;
    uint32_t index = 0; // std::upper_bound
    while (index < n && !(offset < stops[index]))
        ++index;
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

// paint_pixel at point (already offset to the pixel centre).
__device__ inline float4 paint(const frame &F, const draw_desc &D, f2 point) {
    if (D.brush == CIH_BRUSH_PATTERN || D.brush == CIH_BRUSH_IMAGE)
        return paint_pattern(F, D, apply(D.paint_inverse, point));
    if (D.ncolors == 0)
        return make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    if (D.brush == CIH_BRUSH_COLOR)
        return load4((const float *)(F.payload + D.colors_at));
    return paint_gradient(F, D, apply(D.paint_inverse, point));
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
__global__ __launch_bounds__(FINE_THREADS) void k_shadow_raster(frame F) {
    __shared__ seg_lds L;
    int d = blockIdx.x;
    if (d >= F.ndraws || overflowed(F))
        return;
    const draw_desc &D = F.draws[d];
    if (!D.shadow)
        return;
    const draw_state &S = F.state[d];
    for (int i = threadIdx.x; i < FINE_SEGMENT; i += FINE_THREADS)
        L.acc[i] = 0;
    if (threadIdx.x == 0) {
        L.minx = 0x7fffffff;
        L.maxx = -1;
    }
    __syncthreads();
    float *image = F.shadow + S.image;
    const run *runs = F.runs + S.sruns;
    const uint32_t *rowoff = F.words + S.srowoff;
    f2 offset = mk((float)D.border + D.shadow_offset_x, (float)D.border + D.shadow_offset_y);
    for (int r = blockIdx.y; r < S.height; r += gridDim.y) {
        int y = S.top + r;
        bool hit = y >= S.srow0 && y < S.srow0 + S.srows;
        uint32_t begin = hit ? rowoff[y - S.srow0] : 0, end = hit ? rowoff[y - S.srow0 + 1] : 0;
        for (int x0 = S.left; x0 < S.left + S.width; x0 += FINE_SEGMENT) {
            long long q[FINE_PX] = {0};
            int rmin = 0x7fffffff, rmax = -1;
            if (begin != end)
                segment_coverage(runs, begin, end, x0, L, q, rmin, rmax);
            for (int k = 0; k < FINE_PX; ++k) {
                int x = x0 + threadIdx.x * FINE_PX + k;
                if (x >= S.left + S.width)
                    continue;
                float alpha = 0.0f;
                if (x >= rmin && x <= rmax) {
                    float cov = coverage(q[k], (D.flags & CIH_DRAW_FLAG_EVENODD) != 0);
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

// One extended-box pass over n samples at stride `step` (hpp:2460-2481).
__device__ inline void box_pass(const float *in, float *out, int n, size_t step, int radius,
                                float weight_1, float weight_2) {
    float running = weight_1 * (radius + 1 < n ? in[(size_t)(radius + 1) * step] : 0.0f);
    for (int x = 0; x <= radius; ++x)
        running += (weight_1 + weight_2) * (x < n ? in[(size_t)x * step] : 0.0f);
    out[0] = running;
    for (int x = 1; x < n; ++x) {
        if (x >= radius + 1)
            running -= weight_2 * in[(size_t)(x - radius - 1) * step];
        if (x >= radius + 2)
            running -= weight_1 * in[(size_t)(x - radius - 2) * step];
        if (x + radius < n)
            running += weight_2 * in[(size_t)(x + radius) * step];
        if (x + radius + 1 < n)
            running += weight_1 * in[(size_t)(x + radius + 1) * step];
        out[(size_t)x * step] = running;
    }
}

// Three passes along every row (one thread per row): image -> image2 ->
// image -> image2.
__global__ __launch_bounds__(64) void k_blur_rows(frame F) {
    int d = blockIdx.x;
    if (d >= F.ndraws || overflowed(F))
        return;
    const draw_desc &D = F.draws[d];
    const draw_state &S = F.state[d];
    if (!D.shadow)
        return;
    int r = blockIdx.y * 64 + threadIdx.x;
    for (; r < S.height; r += gridDim.y * 64) {
        float *a = F.shadow + S.image + (size_t)r * S.width;
        float *b = F.shadow + S.image2 + (size_t)r * S.width;
        box_pass(a, b, S.width, 1, D.radius, D.weight_1, D.weight_2);
        box_pass(b, a, S.width, 1, D.radius, D.weight_1, D.weight_2);
        box_pass(a, b, S.width, 1, D.radius, D.weight_1, D.weight_2);
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
    int c = blockIdx.y * 64 + threadIdx.x;
    for (; c < S.width; c += gridDim.y * 64) {
        float *a = F.shadow + S.image + c;
        float *b = F.shadow + S.image2 + c;
        box_pass(b, a, S.height, (size_t)S.width, D.radius, D.weight_1, D.weight_2);
        box_pass(a, b, S.height, (size_t)S.width, D.radius, D.weight_1, D.weight_2);
        box_pass(b, a, S.height, (size_t)S.width, D.radius, D.weight_1, D.weight_2);
    }
}

// ---------------------------------------------------------------------------
// k_fine: every draw of a canvas over one row segment, pixels in registers.

__device__ inline uint32_t find_canvas(const uint32_t *prefix, uint32_t n, uint32_t item) {
    return find_prefix(prefix, n, item);
}

__global__ __launch_bounds__(FINE_THREADS) void k_fine(frame F) {
    __shared__ seg_lds L;
    uint32_t item = blockIdx.x;
    if (item >= F.nitems || overflowed(F))
        return;
    uint32_t c = find_canvas(F.items, (uint32_t)F.ncanvases, item);
    const canvas_desc &C = F.canvases[c];
    uint32_t local = item - F.items[c];
    int segs = (C.width + FINE_SEGMENT - 1) / FINE_SEGMENT;
    int y = (int)(local / (uint32_t)segs);
    int x0 = (int)(local % (uint32_t)segs) * FINE_SEGMENT;
    float4 *bitmap = (float4 *)(uintptr_t)C.bitmap + (size_t)y * C.width;
    for (int i = threadIdx.x; i < FINE_SEGMENT; i += FINE_THREADS)
        L.acc[i] = 0;
    if (threadIdx.x == 0) {
        L.minx = 0x7fffffff;
        L.maxx = -1;
    }
    float4 back[FINE_PX];
    int xs[FINE_PX];
    for (int k = 0; k < FINE_PX; ++k) {
        xs[k] = x0 + threadIdx.x * FINE_PX + k;
        back[k] = xs[k] < C.width ? bitmap[xs[k]] : make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    }
    __syncthreads();
    unsigned composites = 0;
    for (int d = C.first_draw; d < C.first_draw + C.ndraws; ++d) {
        const draw_desc &D = F.draws[d];
        const draw_state &S = F.state[d];
        int op = D.op;
        if (D.shadow && S.width > 0) {
            // shadow layer (hpp:2504-2538): rows top <= y + border < bottom,
            // columns [left - border, right - border) within the canvas
            int sy = y + D.border - S.top;
            if (sy >= 0 && sy < S.height) {
                const float *row = F.shadow + S.image + (size_t)sy * S.width;
                int lo = max(0, S.left - D.border);
                int hi = min(C.width, S.left + S.width - D.border);
                float4 color = load4(D.shadow_color);
                for (int k = 0; k < FINE_PX; ++k) {
                    int x = xs[k];
                    if (x < lo || x >= hi)
                        continue;
                    float s = row[x + D.border - S.left];
                    composite(back[k], scaled(color, D.global_alpha * s), op, 1.0f);
                    ++composites;
                }
            }
        }
        bool clear = !(op & 8); // whole-canvas operators (hpp:2574)
        bool rows_hit = S.rows > 0 && y >= S.row0 && y < S.row0 + S.rows;
        if (!rows_hit && !clear)
            continue;
        long long q[FINE_PX] = {0};
        int rmin = 0x7fffffff, rmax = -1;
        if (rows_hit) {
            const uint32_t *rowoff = F.words + S.rowoff;
            uint32_t begin = rowoff[y - S.row0], end = rowoff[y - S.row0 + 1];
            if (begin == end && !clear)
                continue;
            if (begin != end)
                segment_coverage(F.runs + S.runs, begin, end, x0, L, q, rmin, rmax);
        }
        bool evenodd = (D.flags & CIH_DRAW_FLAG_EVENODD) != 0;
        for (int k = 0; k < FINE_PX; ++k) {
            int x = xs[k];
            if (x >= C.width || (!clear && (x < rmin || x > rmax)))
                continue;
            float cov = coverage(q[k], evenodd);
            if (cov >= kThreshold || clear) {
                float4 p = paint(F, D, mk((float)x + 0.5f, (float)y + 0.5f));
                composite(back[k], scaled(p, cov * D.global_alpha), op, 1.0f);
                ++composites;
            }
        }
    }
    unsigned long long total = composites;
    for (int w = 32; w >= 1; w >>= 1)
        total += __shfl_xor(total, w, 64);
    if ((threadIdx.x & 63) == 0 && total)
        atomicAdd(&F.counters[COUNTER_PIXEL_DRAWS], total);
    for (int k = 0; k < FINE_PX; ++k)
        if (xs[k] < C.width)
            bitmap[xs[k]] = back[k];
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

// sRGB encode (hpp:1282-1284).  The power is evaluated in double and rounded
// once (correctly rounded), as close as a device can get to the host's powf.
__device__ inline float delinearized(float v) {
    return v < 0.0031308f ? 12.92f * v
                          : 1.055f * (float)pow((double)v, (double)(1.0f / 2.4f)) - 0.055f;
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

int fine_segment() { return FINE_SEGMENT; }

const int kFrameKernels = 7;
extern const char *const kFrameKernelNames[kFrameKernels] = {
    "k_reset", "k_geometry", "k_runs", "k_shadow_raster", "k_blur_rows", "k_blur_cols",
    "k_fine"};

// Launches a frame's kernels on `stream`.  With `events` (kFrameKernels + 1
// of them), event k is recorded before kernel k and the last after the last
// kernel, so the time of kernel k is elapsed(events[k], events[k + 1]).
hipError_t launch_frame(const frame &F, int shadow_rows, int shadow_cols, bool any_shadow,
                        hipEvent_t *events, hipStream_t stream) {
#define MARK(k)                                                                                \
    do {                                                                                       \
        if (events)                                                                            \
            hipEventRecord(events[k], stream);                                                 \
    } while (0)
    MARK(0);
    k_reset<<<1, 64, 0, stream>>>(F.counters);
    MARK(1);
    if (F.ndraws > 0)
        k_geometry<<<F.ndraws, GEO_THREADS, 0, stream>>>(F);
    MARK(2);
    if (F.ndraws > 0)
        k_runs<<<2 * F.ndraws, GEO_THREADS, 0, stream>>>(F);
    MARK(3);
    int ry = (shadow_rows + 63) / 64, cx = (shadow_cols + 63) / 64;
    if (F.ndraws > 0 && any_shadow)
        k_shadow_raster<<<dim3(F.ndraws, 64), FINE_THREADS, 0, stream>>>(F);
    MARK(4);
    if (F.ndraws > 0 && any_shadow)
        k_blur_rows<<<dim3(F.ndraws, ry > 0 ? ry : 1), 64, 0, stream>>>(F);
    MARK(5);
    if (F.ndraws > 0 && any_shadow)
        k_blur_cols<<<dim3(F.ndraws, cx > 0 ? cx : 1), 64, 0, stream>>>(F);
    MARK(6);
    if (F.nitems)
        k_fine<<<F.nitems, FINE_THREADS, 0, stream>>>(F);
    MARK(7);
#undef MARK
    return hipGetLastError();
}

static unsigned grid_for(size_t n, unsigned block) {
    size_t g = (n + block - 1) / block;
    return (unsigned)(g < 65536 ? (g ? g : 1) : 65536);
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
