// gfx950 kernels of the rasterization pipeline.
//
// Stage map (reference GLSL compute stage -> kernel here):
//   dice.comp                   -> k_flatten<COUNT/EMIT>: the hybrid tiler's recursive
//                                  flatten (core/d3d9/tiler.cpp:284-315), chords in DFS order
//   bin.comp                    -> k_bin_wave / k_bin_group <COUNT/EMIT>: the tile walk of
//                                  each chord (tiler.cpp:131-279), exact float order, per path
//                                  with LDS counters; also writes every tile of the path's rect
//   bound.comp                  -> (the bin count pass initialises each path's tiles)
//   propagate.comp              -> k_propagate (prepare_tiles, tiler.cpp:369-439) + z-buffer
//   sort.comp                   -> k_entries_* + stable LSD radix sort by framebuffer tile
//   fill.comp / fill.frag       -> k_fill_masks: analytic box-filtered area, fixed-point
//   tile.comp / tile.frag       -> k_composite: premultiplied src-over in fp32, one RGBA8 store
//
// This translation unit is built with -ffp-contract=off: the flatten/bin
// arithmetic must round exactly like the reference's SSE code (no FMA).
#include "pfh_internal.h"

namespace pfh {

// ---------------------------------------------------------------------------
// helpers
// ---------------------------------------------------------------------------

// Zeroes up to ZERO_SPANS arrays in one launch (the per-launch clears: one dispatch and one tail
// instead of one per array): index i of their concatenation lies in the first span whose end is
// past it.
__global__ void k_zero_spans(ZeroSpans z) {
    const uint64_t total = z.n ? z.end[z.n - 1] : 0;
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < total; i += stride) {
#pragma unroll
        for (int k = 0; k < ZERO_SPANS; k++) {
            const uint64_t start = k ? z.end[k - 1] : 0;
            if (k < z.n && i >= start && i < z.end[k]) z.p[k][i - start] = 0u;
        }
    }
}

// _mm_max_ps(a, b) / _mm_min_ps(a, b): the second operand wins on NaN or equality.
__device__ __forceinline__ float sse_max(float a, float b) { return a > b ? a : b; }
__device__ __forceinline__ float sse_min(float a, float b) { return a < b ? a : b; }

struct Seg4 {
    float2 p0, c0, c1, p1;
};

__device__ __forceinline__ float2 f2add(float2 a, float2 b) { return make_float2(a.x + b.x, a.y + b.y); }
__device__ __forceinline__ float2 f2sub(float2 a, float2 b) { return make_float2(a.x - b.x, a.y - b.y); }
__device__ __forceinline__ float2 f2muls(float2 a, float s) { return make_float2(a.x * s, a.y * s); }
__device__ __forceinline__ float2 lerp2(float2 a, float2 b, float t) { return f2add(a, f2muls(f2sub(b, a), t)); }

// core/data/segment.cpp:12-20 is_flat_cubic
__device__ __forceinline__ bool flat_cubic(const Seg4 &s) {
    float u0 = 3.0f * s.c0.x - s.p0.x - s.p0.x - s.p1.x;
    float u1 = 3.0f * s.c0.y - s.p0.y - s.p0.y - s.p1.y;
    float u2 = 3.0f * s.c1.x - s.p1.x - s.p1.x - s.p0.x;
    float u3 = 3.0f * s.c1.y - s.p1.y - s.p1.y - s.p0.y;
    u0 = u0 * u0;
    u1 = u1 * u1;
    u2 = u2 * u2;
    u3 = u3 * u3;
    return sse_max(u0, u2) + sse_max(u1, u3) <= 1.0f;  // FLATTENING_TOLERANCE, tiler.cpp:15
}

// core/data/segment.cpp:22-36 is_flat_quadratic
__device__ __forceinline__ bool flat_quad(const Seg4 &s) {
    float2 mid = f2muls(f2add(s.p0, s.p1), 0.5f);
    float2 dev = f2sub(s.c0, mid);
    return dev.x * dev.x + dev.y * dev.y <= 1.0f * 0.25f;
}

// ---------------------------------------------------------------------------
// Tile walk of one chord (hybrid semantics)
// ---------------------------------------------------------------------------

// The walk's side effects go either to global counters (atomics) or, when one wave or
// workgroup owns the whole path, to LDS arrays indexed by the tile's position in the path's
// rect (LDS = true): l_cnt holds fill counts (COUNT) or the tiles' next free fill slots
// (EMIT), l_delta the backdrop deltas and l_col the above-rect column backdrops.
// PACK: l_delta holds two tiles per word, 16 bits each biased by 0x8000, so that ±1 steps of
// either half never carry into the other while |sum| < 2^15 (a chord steps a tile at most once:
// callers use it for paths with fewer than 2^15 chords).
template <bool EMIT, bool LDS = false, bool PACK = false>
struct Walker {
    // scene / path state
    float vb_l, vb_r, vb_b;
    int32_t tl, tt, tr, tb, tw, th;
    uint32_t tbase, cbase;
    uint32_t *tile_cnt;
    int32_t *tile_delta, *col_bd;
    uint32_t *l_cnt;
    int32_t *l_delta, *l_col;
    uint4 *out;              // EMIT with dumps: fills in emission order
    uint2 *fills;            // EMIT: fills grouped by tile
    const uint32_t *tile_fill_off;
    uint32_t cap_fills;
    uint32_t pend_tile, pend_n;  // COUNT: pending increments of one tile
    uint32_t count;          // fills emitted by this chord so far
    uint32_t out_base;       // first emission slot of this chord (EMIT)
    uint32_t guard;

    __device__ __forceinline__ void count_add(uint32_t j, uint32_t n) {
        if (LDS) atomicAdd(&l_cnt[j], n);
        else atomicAdd(&tile_cnt[tbase + j], n);
    }

    // core/d3d9/object_builder.cpp:19-64 add_fill
    __device__ __forceinline__ void add_fill(float2 from, float2 to, int tx, int ty) {
        if (!(tl <= tx && tx <= tr - 1 && tt <= ty && ty <= tb - 1)) return;
        float ox = (float)tx * 16.0f, oy = (float)ty * 16.0f;
        float v0 = (from.x - ox) * 256.0f, v1 = (from.y - oy) * 256.0f;
        float v2 = (to.x - ox) * 256.0f, v3 = (to.y - oy) * 256.0f;
        uint32_t q0 = (uint32_t)rintf(sse_min(sse_max(v0, 0.0f), 4095.0f));
        uint32_t q1 = (uint32_t)rintf(sse_min(sse_max(v1, 0.0f), 4095.0f));
        uint32_t q2 = (uint32_t)rintf(sse_min(sse_max(v2, 0.0f), 4095.0f));
        uint32_t q3 = (uint32_t)rintf(sse_min(sse_max(v3, 0.0f), 4095.0f));
        if (q0 == q2) return;
        // tile within the path's rect (rect sides are below 2^24 tiles: a 24-bit multiply-add)
        uint32_t j = __umul24((uint32_t)(ty - tt), (uint32_t)tw) + (uint32_t)(tx - tl);
        if (!EMIT) {
            // consecutive fills of a chord mostly share a tile: count them locally
            if (j != pend_tile) {
                if (pend_n) count_add(pend_tile, pend_n);
                pend_tile = j;
                pend_n = 0;
            }
            pend_n++;
        } else {
            // straight into the tile's range; order inside a tile is free (coverage is
            // accumulated in fixed point)
            const uint32_t packed_x = q0 | (q1 << 16), packed_y = q2 | (q3 << 16);
            uint32_t pos;
            if (LDS) {
                pos = atomicAdd(&l_cnt[j], 1u);
            } else {
                uint32_t slot = atomicSub(&tile_cnt[tbase + j], 1u) - 1u;
                // written by this block in this launch: read from L2
                pos = __hip_atomic_load(&tile_fill_off[tbase + j], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) + slot;
            }
            if (pos < cap_fills) fills[pos] = make_uint2(packed_x, packed_y);
            if (out) {
                uint32_t d = out_base + count;
                if (d < cap_fills) out[d] = make_uint4(packed_x, packed_y, tbase + j, 0u);
            }
        }
        count++;
    }

    __device__ __forceinline__ void flush() {
        if (!EMIT && pend_n) count_add(pend_tile, pend_n);
        pend_n = 0;
    }

    // core/d3d9/object_builder.cpp:96-114 adjust_alpha_tile_backdrop
    __device__ __forceinline__ void adjust(int tx, int ty, int delta) {
        if (EMIT) return;  // backdrops are accumulated once, in the count pass
        int ox = tx - tl, oy = ty - tt;
        if (ox < 0 || ox >= tw || oy >= th) return;
        if (oy < 0) {
            if (LDS) atomicAdd(&l_col[ox], delta);
            else atomicAdd(&col_bd[cbase + ox], delta);
            return;
        }
        const uint32_t j = __umul24((uint32_t)oy, (uint32_t)tw) + (uint32_t)ox;
        if (LDS && PACK) atomicAdd(reinterpret_cast<uint32_t *>(l_delta) + (j >> 1), (uint32_t)delta << (16 * (j & 1)));
        else if (LDS) atomicAdd(&l_delta[j], delta);
        else atomicAdd(&tile_delta[tbase + j], delta);
    }

    __device__ __forceinline__ static uint32_t outcode(float2 p, float l, float t, float r, float bb) {
        uint32_t f = 0;
        if (p.x < l) f |= 1;
        else if (p.x > r) f |= 2;
        if (p.y < t) f |= 4;
        else if (p.y > bb) f |= 8;
        return f;
    }

    // common/math/basic.h:45-53 lerp with t clamped to [0, 1]
    __device__ __forceinline__ static float lerpc(float a, float bv, float t) {
        t = t < 0.0f ? 0.0f : (t > 1.0f ? 1.0f : t);
        return a + (bv - a) * t;
    }

    // core/d3d9/tiler.cpp:65-125 Cohen–Sutherland against the view box with top = -inf
    __device__ bool clip(float2 &from, float2 &to) {
        const float l = vb_l, t = -INFINITY, r = vb_r, bb = vb_b;
        uint32_t of = outcode(from, l, t, r, bb), ot = outcode(to, l, t, r, bb);
        for (int iter = 0; iter < 64; iter++) {
            if (of == 0 && ot == 0) return true;
            if ((of & ot) != 0) return false;
            bool clip_from = of > ot;
            uint32_t oc = clip_from ? of : ot;
            // the first edge of the outcode (left, right, top, bottom): a vertical edge fixes x and
            // interpolates y, a horizontal one the reverse.  The operands are selected first so
            // that one division serves every case (lanes clipping different edges run it once).
            const bool vert = (oc & 3u) != 0;
            const float edge = (oc & 1u) ? l : ((oc & 2u) ? r : ((oc & 4u) ? t : bb));
            const float a0 = vert ? from.x : from.y, a1 = vert ? to.x : to.y;
            const float b0 = vert ? from.y : from.x, b1 = vert ? to.y : to.x;
            const float other = lerpc(b0, b1, (edge - a0) / (a1 - a0));
            const float2 p = vert ? make_float2(edge, other) : make_float2(other, edge);
            if (clip_from) {
                from = p;
                of = outcode(p, l, t, r, bb);
            } else {
                to = p;
                ot = outcode(p, l, t, r, bb);
            }
        }
        guard++;
        return false;
    }

    // core/d3d9/tiler.cpp:131-279 process_line_segment
    __device__ void line(float2 from, float2 to) {
        if (!isfinite(from.x) || !isfinite(from.y) || !isfinite(to.x) || !isfinite(to.y)) return;
        if (!clip(from, to)) return;
        const float inv = 1.0f / 16.0f;
        int fx = (int)floorf(from.x * inv), fy = (int)floorf(from.y * inv);
        int ex = (int)floorf(to.x * inv), ey = (int)floorf(to.y * inv);
        float2 vec = f2sub(to, from);
        int stx = vec.x < 0 ? -1 : 1, sty = vec.y < 0 ? -1 : 1;
        float crx = ((float)fx + (vec.x >= 0 ? 1.0f : 0.0f)) * 16.0f;
        float cry = ((float)fy + (vec.y >= 0 ? 1.0f : 0.0f)) * 16.0f;
        float tmx = (crx - from.x) / vec.x, tmy = (cry - from.y) / vec.y;
        float tdx = fabsf(16.0f / vec.x), tdy = fabsf(16.0f / vec.y);
        float2 cur = from;
        int cx = fx, cy = fy;
        int last = 0;  // 0 none, 1 X, 2 Y
        int64_t dx = (int64_t)ex - fx, dy = (int64_t)ey - fy;
        int64_t lim = (dx < 0 ? -dx : dx) + (dy < 0 ? -dy : dy) + 2;
        for (int64_t it = 0;; it++) {
            if (it > lim) {
                guard++;
                break;
            }
            int next;
            if (tmx < tmy) next = 1;
            else if (tmx > tmy) next = 2;
            else next = stx > 0 ? 1 : 2;
            float nt = next == 1 ? tmx : tmy;
            nt = (1.0f < nt) ? 1.0f : nt;  // std::min(next_t, 1.0f)
            if (cx == ex && cy == ey) next = 0;
            float2 npos = f2add(from, f2muls(vec, nt));  // LineSegmentF::sample
            add_fill(cur, npos, cx, cy);
            // the auxiliary fill along the tile's left edge: (npos, corner) when stepping up out of
            // the tile, (corner, cur) when having stepped down into it; one call site for both, so
            // lanes walking up and down do not run two copies
            const float2 corner = make_float2((float)cx * 16.0f, (float)cy * 16.0f);
            const bool up = sty < 0 && next == 2, down = sty > 0 && last == 2;
            if (up || down) add_fill(up ? npos : corner, up ? corner : cur, cx, cy);
            // the backdrop step when leaving to the right or having entered from the left
            if ((stx < 0 && last == 1) || (stx > 0 && next == 1)) adjust(cx, cy, stx < 0 ? 1 : -1);
            if (next == 1) {
                tmx += tdx;
                cx += stx;
            } else if (next == 2) {
                tmy += tdy;
                cy += sty;
            } else {
                break;
            }
            cur = npos;
            last = next;
        }
    }
};

// ---------------------------------------------------------------------------
// Flatten: one thread per segment, chords out in depth-first order
// ---------------------------------------------------------------------------
//
// core/d3d9/tiler.cpp:284-315 process_segment (depth first, first half first) and
// core/data/segment.cpp:42-126 split_cubic / split_quadratic at t = 0.5.  The explicit
// stack holds pending right halves without their start point: a pending half starts
// exactly where the previously flattened piece ended (splits keep their endpoints
// bit-for-bit), so only (c0, c1, p1, depth) are stored.  The stack's first
// FLAT_LDS_LEVELS levels live in LDS, level-major with one float per lane per row so the
// lanes of a wave hit distinct banks; deeper levels (rare: curves of the benchmark
// generators flatten within depth 6) fall back to a private array.  Four levels leave room
// for 6 workgroups per CU (6 levels: 4, and both passes 12–15 % slower; 3 levels: the
// private levels' traffic costs more).  A private stack alone is placed in scratch, where a
// divergent wave's narrow stores make the
// cache write back whole lines: several times the chord output in HBM writes.
constexpr int FLAT_LDS_LEVELS = 4;

// Lines need no subdivision: k_flatten_lines, a thread per segment without the LDS stack (so
// with many more waves per CU to hide the loads), finishes every line and every segment that
// emits nothing, and its count pass marks the curves, one bit per segment in a 64-bit word per
// 64 segments, for k_flatten.  With deep_pred it lists instead the curves whose subdivision is
// estimated at deep_pred levels or more for k_flatten_deep: a halving divides the flatness
// measure of core/data/segment.cpp:12-36 by about 16 (second differences by 4, squared), so a
// measure above 16^(deep_pred - 1) takes about deep_pred levels.  The estimate only routes a
// curve; both kernels subdivide it the same way.
__device__ __forceinline__ bool deep_curve(const float2 *P, uint32_t kind, uint32_t deep_pred, DeepCurve &dc) {
    const float2 p0 = P[0], c0 = P[1], p1 = P[kind == SEG_QUAD ? 2 : 3], c1 = kind == SEG_QUAD ? c0 : P[2];
    dc.p0 = p0;
    dc.c0 = c0;
    dc.c1 = c1;
    dc.p1 = p1;
    float m;
    if (kind == SEG_QUAD) {
        const float dx = c0.x - 0.5f * (p0.x + p1.x), dy = c0.y - 0.5f * (p0.y + p1.y);
        m = 4.0f * (dx * dx + dy * dy);
    } else {
        const float u0 = 3.0f * c0.x - 2.0f * p0.x - p1.x, u1 = 3.0f * c0.y - 2.0f * p0.y - p1.y;
        const float u2 = 3.0f * c1.x - 2.0f * p1.x - p0.x, u3 = 3.0f * c1.y - 2.0f * p1.y - p0.y;
        m = fmaxf(u0 * u0, u2 * u2) + fmaxf(u1 * u1, u3 * u3);
    }
    // a non-finite curve emits nothing: k_flatten disposes of it at once
    const bool finite = isfinite(p0.x) && isfinite(p0.y) && isfinite(c0.x) && isfinite(c0.y) && isfinite(c1.x) &&
                        isfinite(c1.y) && isfinite(p1.x) && isfinite(p1.y);
    return finite && m > ldexpf(1.0f, 4 * ((int)deep_pred - 1));
}

template <bool EMIT>
__global__ __launch_bounds__(256) void k_flatten_lines(BatchDev b) {
    const uint32_t lane = threadIdx.x & 63;
    const uint32_t g0 = blockIdx.x * 256 + (threadIdx.x & ~63u);  // the wave's first segment
    if (g0 >= b.n_segs) return;  // whole wave
    const uint32_t g = g0 + lane;
    const bool in = g < b.n_segs;
    int k = find_scene_wave(b.seg_bases, b.n_scenes, g0);
    uint32_t kind = 0;
    bool brk = false, deep = false;
    float2 p0 = {}, p1 = {};
    DeepCurve dc{};
    if (in) {
        while (g >= b.seg_bases[k + 1]) k++;
        const SceneDev &sc = b.scenes[k];
        const uint2 sg = sc.segs[g - sc.seg_base];
        const uint32_t sk = sg.y & SEG_KIND_MASK;
        if (sk == SEG_LINE) {
            p0 = sc.pts[sg.x];
            p1 = sc.pts[sg.x + 1];
            // a non-finite point: process_line_segment drops the chord (tiler.cpp:133-136)
            if (isfinite(p0.x) && isfinite(p0.y) && isfinite(p1.x) && isfinite(p1.y)) {
                kind = SEG_LINE;
                brk = (sg.y & SEG_CHAIN_BREAK) != 0;
            }
        } else if (sk != 0) {
            kind = sk;  // a curve: its points are checked where it is subdivided
            if (!EMIT && b.deep_pred) {
                deep = deep_curve(sc.pts + sg.x, sk, b.deep_pred, dc);
                dc.seg = g;
                dc.scene = (uint32_t)k | (sk == SEG_QUAD ? DEEP_QUAD : 0u) | ((sg.y & SEG_CHAIN_BREAK) ? DEEP_BREAK : 0u);
            }
        }
    }
    const bool curve = kind == SEG_QUAD || kind == SEG_CUBIC;
    if (!EMIT) {
        // stream points: a chord's end (+ the chain's separator and start point)
        if (in && !curve) b.seg_cnt[g] = kind == SEG_LINE ? 1u + (brk ? 2u : 0u) : 0u;
        const uint64_t m = __ballot(curve && !deep);
        if (lane == 0) b.curve_mask[g0 >> 6] = m;
        const uint64_t dm = __ballot(deep);
        if (dm) {  // wave-uniform: one add per wave, on its region's counter
            const uint32_t slot = (g0 >> 6) & (DEEP_SLOTS - 1);
            uint32_t at = 0;
            if (lane == 0) at = atomicAdd(&b.counters[C_DEEP + slot], (uint32_t)__popcll(dm));
            at = __shfl(at, 0);
            const uint32_t below = __builtin_amdgcn_mbcnt_hi((uint32_t)(dm >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)dm, 0u));
            if (deep) b.deep_list[(uint64_t)slot * b.deep_slot_cap + at + below] = dc;
        }
    } else if (kind == SEG_LINE) {
        uint32_t o = b.seg_cnt[g];
        if (brk) {  // the chain's separator (a NaN point: its chords emit nothing) and start
            if (o < b.cap_lines) b.lines[o] = make_float2(__builtin_nanf(""), __builtin_nanf(""));
            if (o + 1 < b.cap_lines) b.lines[o + 1] = p0;
            o += 2;
        }
        if (o < b.cap_lines) b.lines[o] = p1;
    }
}

// de Casteljau at t = 0.5 (core/data/segment.cpp:42-126), one path for both kinds (lanes holding
// quads and cubics run it once).  A quad is held as (p0, c, c, p1): p01 = lerp(p0, c) and
// p23 = lerp(c, p1) are its qa and qb, and taking them in place of p012 and p123 makes
// p0123 = lerp(qa, qb), its midpoint: halves (p0, qa, qa, qc), (qc, qb, qb, p1).
__device__ __forceinline__ void split_mid(const Seg4 &s, bool quad, Seg4 &a, Seg4 &r) {
    const float t = 0.5f;
    const float2 p01 = lerp2(s.p0, s.c0, t), p12 = lerp2(s.c0, s.c1, t), p23 = lerp2(s.c1, s.p1, t);
    const float2 p012 = quad ? p01 : lerp2(p01, p12, t), p123 = quad ? p23 : lerp2(p12, p23, t);
    const float2 p0123 = lerp2(p012, p123, t);
    a = Seg4{s.p0, p01, p012, p0123};
    r = Seg4{p0123, p123, p23, s.p1};
}

// The curves k_flatten_lines marked.  Lanes are scheduled dynamically within a wave: the wave
// owns a contiguous range of segments, each iteration every lane takes one depth-first step of
// its curve, and a lane whose curve is done takes the next curve of the range (a wave-uniform
// counter, ranks by ballot).  So a wave no longer runs as long as its deepest curve while the
// lanes holding shallow curves idle.  The curves among the next 64 segments are loaded one per
// lane (one dependent load chain per 64 segments, coalesced), packed into the lowest lanes in
// segment order, and handed to the lane that takes each by lane shuffles; 64 segments without a
// curve cost one scalar load of their mask word.
template <bool EMIT>
__global__ __launch_bounds__(256) void k_flatten(BatchDev b, uint32_t segs_per_wave, uint32_t take) {
    __shared__ float fs[FLAT_LDS_LEVELS][6][256];
    __shared__ uint8_t fd[FLAT_LDS_LEVELS][256];
    const uint32_t tid = threadIdx.x, lane = tid & 63;
    const uint64_t w0 = ((uint64_t)blockIdx.x * 4 + (tid >> 6)) * segs_per_wave;
    if (w0 >= b.n_segs) return;  // whole wave; each lane uses only its own LDS column
    const uint32_t w1 = (uint32_t)min<uint64_t>(w0 + segs_per_wave, b.n_segs);
    const uint64_t lt_mask = lane == 0 ? 0ull : (~0ull >> (64 - lane));
    float stk[MAX_FLATTEN_DEPTH + 1 - FLAT_LDS_LEVELS][6];
    int sdepth[MAX_FLATTEN_DEPTH + 1 - FLAT_LDS_LEVELS];
    // scene of the next segment to load: segment indices only grow, so it is advanced, not searched
    int k = find_scene_wave(b.seg_bases, b.n_scenes, (uint32_t)w0);
    uint32_t word = (uint32_t)w0;             // first of the next 64 segments to look at (wave-uniform)
    uint32_t staged = 0, taken = 0;           // curves staged in lanes [0, staged); those handed out
    float2 q0 = {}, q1 = {}, q2 = {}, q3 = {};  // staged curve of this lane: points
    uint32_t qkind = 0, qbase = 0, qseg = 0;    // its kind (0: emits nothing), chord offset, segment
    Seg4 s;
    uint32_t g = 0, kind = 0, n = 0, base = 0, guard = 0;
    bool brk = false;  // this segment starts a chord chain (a separator and its start point first)
    int d = 0, sp = 0;
    bool active = false;
    while (true) {
        const uint64_t want = __ballot(!active);
        if (want && taken == staged && word < w1) {  // wave-uniform
            uint64_t m = 0;
            while (word < w1) {
                m = b.curve_mask[word >> 6];
                if (m) break;
                word += 64;
            }
            if (m) {
                const uint32_t gl = word + lane;
                float2 p0 = {}, p1 = {}, p2 = {}, p3 = {};
                uint32_t kd = 0, bs = 0;
                int kl = k;
                if ((m >> lane) & 1ull) {
                    while (gl >= b.seg_bases[kl + 1]) kl++;
                    const SceneDev &sc = b.scenes[kl];
                    const uint2 sg = sc.segs[gl - sc.seg_base];
                    bs = EMIT ? b.seg_cnt[gl] : 0u;
                    const float2 *P = sc.pts + sg.x;
                    p0 = P[0];
                    p1 = P[1];
                    if ((sg.y & SEG_KIND_MASK) == SEG_QUAD) {
                        p2 = p1;
                        p3 = P[2];
                    } else {
                        p2 = P[2];
                        p3 = P[3];
                    }
                    // A non-finite point makes every flattened piece non-finite, and
                    // process_line_segment drops those (tiler.cpp:133-136): nothing from it.
                    const bool finite = isfinite(p0.x) && isfinite(p0.y) && isfinite(p1.x) && isfinite(p1.y) &&
                                        isfinite(p2.x) && isfinite(p2.y) && isfinite(p3.x) && isfinite(p3.y);
                    if (finite) kd = sg.y;  // kind and chain-break flag
                }
                k = __builtin_amdgcn_readlane(kl, 63 - __clzll(m));
                // pack: a curve's lane sends its data to lane (curves below it), the other lanes to
                // the lanes above the curves, in order (a permutation of the lanes)
                const uint32_t below = __builtin_amdgcn_mbcnt_hi((uint32_t)(m >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)m, 0u));
                staged = (uint32_t)__popcll(m);
                const int to = (int)(((m >> lane) & 1ull) ? below : staged + (lane - below)) * 4;
                auto send = [&](uint32_t v) { return (uint32_t)__builtin_amdgcn_ds_permute(to, (int)v); };
                auto sendf = [&](float v) { return __uint_as_float(send(__float_as_uint(v))); };
                q0 = make_float2(sendf(p0.x), sendf(p0.y));
                q1 = make_float2(sendf(p1.x), sendf(p1.y));
                q2 = make_float2(sendf(p2.x), sendf(p2.y));
                q3 = make_float2(sendf(p3.x), sendf(p3.y));
                qkind = send(kd);
                qbase = send(bs);
                qseg = send(gl);
                taken = 0;
                word += 64;
            }
        }
        const uint32_t idle = (uint32_t)__popcll(want);
        // a hand-over (shuffles of the staged curves to the idle lanes) once `take` lanes are idle,
        // or as many as there are curves staged: fewer hand-overs, each feeding more lanes
        if (want && taken < staged && idle >= min(take, staged - taken)) {  // wave-uniform
            // idle lanes take the next staged curves, in lane order
            const uint32_t cand = taken + (uint32_t)__popcll(want & lt_mask);
            taken = min(taken + idle, staged);
            const int src = (int)(cand & 63);
            const float2 t0 = make_float2(__shfl(q0.x, src), __shfl(q0.y, src));
            const float2 t1 = make_float2(__shfl(q1.x, src), __shfl(q1.y, src));
            const float2 t2 = make_float2(__shfl(q2.x, src), __shfl(q2.y, src));
            const float2 t3 = make_float2(__shfl(q3.x, src), __shfl(q3.y, src));
            const uint32_t tk = __shfl(qkind, src);
            const uint32_t tb = EMIT ? __shfl(qbase, src) : 0u;
            const uint32_t tg = __shfl(qseg, src);
            if (!active && cand < staged) {
                g = tg;
                s = Seg4{t0, t1, t2, t3};
                kind = tk & SEG_KIND_MASK;
                brk = kind != 0 && (tk & SEG_CHAIN_BREAK) != 0;
                base = tb;
                if (EMIT && brk) {  // the chain's separator (a NaN point: its chords emit nothing) and start
                    if (base < b.cap_lines) b.lines[base] = make_float2(__builtin_nanf(""), __builtin_nanf(""));
                    if (base + 1 < b.cap_lines) b.lines[base + 1] = s.p0;
                    base += 2;
                }
                n = 0;
                d = 0;
                sp = 0;
                active = true;
            }
        }
        // every lane idle after a hand-over: nothing is staged and the range is done
        if (__ballot(active) == 0) break;
        if (!active) continue;
        bool done = false;
        if (kind == 0) {
            done = true;
        } else if (kind == SEG_LINE) {
            if (EMIT && base < b.cap_lines) b.lines[base] = s.p1;
            n = 1;
            done = true;
        } else {
            // one depth-first step: a flat piece is emitted (then the next pending half is
            // popped), otherwise it is split at t = 0.5 and the right half pushed
            const bool quad = kind == SEG_QUAD;
            const bool flat = quad ? flat_quad(s) : flat_cubic(s);
            if (flat || d >= MAX_FLATTEN_DEPTH) {
                if (!flat) guard++;
                if (EMIT && base + n < b.cap_lines) b.lines[base + n] = s.p1;
                n++;
                if (sp == 0) {
                    done = true;
                } else {
                    sp--;
                    s.p0 = s.p1;
                    if (sp < FLAT_LDS_LEVELS) {
                        s.c0 = make_float2(fs[sp][0][tid], fs[sp][1][tid]);
                        s.c1 = make_float2(fs[sp][2][tid], fs[sp][3][tid]);
                        s.p1 = make_float2(fs[sp][4][tid], fs[sp][5][tid]);
                        d = fd[sp][tid];
                    } else {
                        const int q = sp - FLAT_LDS_LEVELS;
                        s.c0 = make_float2(stk[q][0], stk[q][1]);
                        s.c1 = make_float2(stk[q][2], stk[q][3]);
                        s.p1 = make_float2(stk[q][4], stk[q][5]);
                        d = sdepth[q];
                    }
                }
            } else {
                Seg4 a, bseg;  // the halves at t = 0.5
                split_mid(s, quad, a, bseg);
                if (sp < FLAT_LDS_LEVELS) {
                    fs[sp][0][tid] = bseg.c0.x;
                    fs[sp][1][tid] = bseg.c0.y;
                    fs[sp][2][tid] = bseg.c1.x;
                    fs[sp][3][tid] = bseg.c1.y;
                    fs[sp][4][tid] = bseg.p1.x;
                    fs[sp][5][tid] = bseg.p1.y;
                    fd[sp][tid] = (uint8_t)(d + 1);
                } else {
                    const int q = sp - FLAT_LDS_LEVELS;
                    stk[q][0] = bseg.c0.x;
                    stk[q][1] = bseg.c0.y;
                    stk[q][2] = bseg.c1.x;
                    stk[q][3] = bseg.c1.y;
                    stk[q][4] = bseg.p1.x;
                    stk[q][5] = bseg.p1.y;
                    sdepth[q] = d + 1;
                }
                sp++;
                s = a;
                d++;
            }
        }
        if (done) {
            if (!EMIT) b.seg_cnt[g] = n + (brk ? 2u : 0u);  // stream points: chords (+ separator, start)
            active = false;
        }
    }
    if (!EMIT && guard) atomicAdd(&b.counters[C_GUARD], guard);
}

// The curves k_flatten_lines listed as deep, 16 lanes each (4 per wave), so that a curve of 2^d
// chords takes about 2^(d-3) serial steps instead of 2^(d+1).  Lane j of a group owns the j-th
// node, left to right, of depth DEEP_LEVELS of the curve's subdivision tree: it walks down to it
// from the root by the same splits and flatness tests as the depth-first walk of k_flatten, and
// where a node on the way is already flat (a leaf above that depth) the leftmost lane under it
// emits its chord and the others nothing.  Otherwise the lane flattens its node depth first.
// The chords of the group's lanes, in lane order, are the curve's chords in the order k_flatten
// emits them (core/d3d9/tiler.cpp:284-315: first half first).  The count pass writes the curve's
// total; the emit pass counts each lane's share again, places it by a 16-lane prefix sum and
// emits it.  Workgroups stride over the list's regions taken end to end (their lengths are only
// known on the device: each workgroup scans the DEEP_SLOTS region counts in LDS first).
constexpr int DEEP_LEVELS = 4;  // 16 lanes per curve

template <bool EMIT>
__device__ __forceinline__ uint32_t flatten_subtree(Seg4 s, int d, bool quad, float (*fs)[6][256], uint8_t (*fd)[256],
                                                    uint32_t tid, float2 *out, uint32_t room, uint32_t &guard) {
    float stk[MAX_FLATTEN_DEPTH + 1 - FLAT_LDS_LEVELS][6];
    int sdepth[MAX_FLATTEN_DEPTH + 1 - FLAT_LDS_LEVELS];
    uint32_t n = 0;
    int sp = 0;
    while (true) {
        const bool flat = quad ? flat_quad(s) : flat_cubic(s);
        if (flat || d >= MAX_FLATTEN_DEPTH) {
            if (!flat) guard++;
            if (EMIT && n < room) out[n] = s.p1;
            n++;
            if (sp == 0) break;
            sp--;
            s.p0 = s.p1;
            if (sp < FLAT_LDS_LEVELS) {
                s.c0 = make_float2(fs[sp][0][tid], fs[sp][1][tid]);
                s.c1 = make_float2(fs[sp][2][tid], fs[sp][3][tid]);
                s.p1 = make_float2(fs[sp][4][tid], fs[sp][5][tid]);
                d = fd[sp][tid];
            } else {
                const int q = sp - FLAT_LDS_LEVELS;
                s.c0 = make_float2(stk[q][0], stk[q][1]);
                s.c1 = make_float2(stk[q][2], stk[q][3]);
                s.p1 = make_float2(stk[q][4], stk[q][5]);
                d = sdepth[q];
            }
        } else {
            Seg4 a, r;
            split_mid(s, quad, a, r);
            if (sp < FLAT_LDS_LEVELS) {
                fs[sp][0][tid] = r.c0.x;
                fs[sp][1][tid] = r.c0.y;
                fs[sp][2][tid] = r.c1.x;
                fs[sp][3][tid] = r.c1.y;
                fs[sp][4][tid] = r.p1.x;
                fs[sp][5][tid] = r.p1.y;
                fd[sp][tid] = (uint8_t)(d + 1);
            } else {
                const int q = sp - FLAT_LDS_LEVELS;
                stk[q][0] = r.c0.x;
                stk[q][1] = r.c0.y;
                stk[q][2] = r.c1.x;
                stk[q][3] = r.c1.y;
                stk[q][4] = r.p1.x;
                stk[q][5] = r.p1.y;
                sdepth[q] = d + 1;
            }
            sp++;
            s = a;
            d++;
        }
    }
    return n;
}

template <bool EMIT>
__global__ __launch_bounds__(256) void k_flatten_deep(BatchDev b) {
    __shared__ float fs[FLAT_LDS_LEVELS][6][256];
    __shared__ uint8_t fd[FLAT_LDS_LEVELS][256];
    __shared__ uint32_t sstart[DEEP_SLOTS + 1];  // exclusive scan of the region counts
    __shared__ uint32_t wsum[4];
    static_assert(DEEP_SLOTS == 256, "one region count per thread");
    const uint32_t tid = threadIdx.x, j = tid & 15u;
    {
        uint32_t inc = min(b.counters[C_DEEP + tid], b.deep_slot_cap);
        for (int dl = 1; dl < 64; dl <<= 1) {
            const uint32_t o = __shfl_up(inc, dl, 64);
            if ((tid & 63u) >= (uint32_t)dl) inc += o;
        }
        if ((tid & 63u) == 63u) wsum[tid >> 6] = inc;
        __syncthreads();
        for (uint32_t w = 0; w < (tid >> 6); w++) inc += wsum[w];
        sstart[tid + 1] = inc;
        if (tid == 0) sstart[0] = 0;
        __syncthreads();
    }
    const uint32_t n = sstart[DEEP_SLOTS];
    uint32_t guard = 0;
    // the 16 lanes of a group take the same curves (uniform over the group)
    for (uint32_t i = blockIdx.x * 16 + (tid >> 4); i < n; i += gridDim.x * 16) {
        // item i: region r, the last with sstart[r] <= i (empty regions share the next one's start)
        uint32_t lo = 0, hi = DEEP_SLOTS - 1;
        while (lo < hi) {
            const uint32_t mid = (lo + hi + 1) >> 1;
            if (sstart[mid] <= i) lo = mid;
            else hi = mid - 1;
        }
        // the list is in no order: each entry carries its segment's kind, flag and points
        const DeepCurve dc = b.deep_list[(uint64_t)lo * b.deep_slot_cap + (i - sstart[lo])];
        const uint32_t g = dc.seg;
        const bool quad = (dc.scene & DEEP_QUAD) != 0;
        const bool brk = (dc.scene & DEEP_BREAK) != 0;
        Seg4 s{dc.p0, dc.c0, dc.c1, dc.p1};
        const float2 start = s.p0;
        // down to this lane's node (bits of j, most significant first: 0 the first half)
        bool leaf = false, mine = false;
        for (int lv = 0; lv < DEEP_LEVELS; lv++) {
            if (quad ? flat_quad(s) : flat_cubic(s)) {  // a leaf at depth lv < MAX_FLATTEN_DEPTH
                leaf = true;
                mine = (j & ((1u << (DEEP_LEVELS - lv)) - 1u)) == 0;
                break;
            }
            Seg4 a, r;
            split_mid(s, quad, a, r);
            s = ((j >> (DEEP_LEVELS - 1 - lv)) & 1u) ? r : a;
        }
        uint32_t cnt = leaf ? (mine ? 1u : 0u)
                            : flatten_subtree<false>(s, DEEP_LEVELS, quad, fs, fd, tid, nullptr, 0u, guard);
        const uint32_t incl = row16_inclusive_sum(cnt);
        if (!EMIT) {
            if (j == 15) b.seg_cnt[g] = incl + (brk ? 2u : 0u);  // stream points (+ separator, start)
            continue;
        }
        uint32_t base = b.seg_cnt[g];
        if (brk) {  // the chain's separator (a NaN point: its chords emit nothing) and start
            if (j == 0 && base < b.cap_lines) b.lines[base] = make_float2(__builtin_nanf(""), __builtin_nanf(""));
            if (j == 0 && base + 1 < b.cap_lines) b.lines[base + 1] = start;
            base += 2;
        }
        const uint32_t at = base + (incl - cnt);
        const uint32_t room = at < b.cap_lines ? b.cap_lines - at : 0u;
        if (leaf) {
            if (mine && room) b.lines[at] = s.p1;
        } else {
            uint32_t unused = 0;
            flatten_subtree<true>(s, DEEP_LEVELS, quad, fs, fd, tid, b.lines + at, room, unused);
        }
    }
    if (!EMIT && guard) atomicAdd(&b.counters[C_GUARD], guard);
}

// ---------------------------------------------------------------------------
// Device-wide exclusive scan (reduce / scan block sums / apply), 4096 items per block
// ---------------------------------------------------------------------------

constexpr int SCAN_THREADS = 256;
constexpr int SCAN_ITEMS = 16;
constexpr int SCAN_TILE = SCAN_THREADS * SCAN_ITEMS;

struct U2 {
    uint32_t x, y;
    __device__ __forceinline__ U2 operator+(const U2 &o) const { return U2{x + o.x, y + o.y}; }
};

__device__ __forceinline__ uint32_t shfl_up_val(uint32_t v, int d) { return __shfl_up(v, d, 64); }
__device__ __forceinline__ U2 shfl_up_val(U2 v, int d) { return U2{__shfl_up(v.x, d, 64), __shfl_up(v.y, d, 64)}; }

template <typename T>
__device__ __forceinline__ T zero_of() { return T{}; }

// Exclusive scan of one value per thread across a 256-thread block; returns the block total.
template <typename T>
__device__ T block_exclusive_scan(T v, T &total) {
    __shared__ T wave_tot[SCAN_THREADS / 64];
    int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    T inc = v;
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
        T o = shfl_up_val(inc, d);
        if (lane >= d) inc = inc + o;
    }
    if (lane == 63) wave_tot[wave] = inc;
    __syncthreads();
    T base = zero_of<T>();
    T all = zero_of<T>();
#pragma unroll
    for (int w = 0; w < SCAN_THREADS / 64; w++) {
        if (w < wave) base = base + wave_tot[w];
        all = all + wave_tot[w];
    }
    __syncthreads();
    total = all;
    T excl_in_wave = shfl_up_val(inc, 1);
    if (lane == 0) excl_in_wave = zero_of<T>();
    return base + excl_in_wave;
}

// ---------------------------------------------------------------------------
// Bin: per path, one wave (small rects) or one workgroup (large rects)
// ---------------------------------------------------------------------------
//
// The chords of a path are contiguous (flatten emits in segment order, segments are in
// path order), and every tile and column of the path's rect belongs to it alone.  So the
// owner counts fills and accumulates backdrop deltas in LDS with LDS atomics and stores
// the whole rect once (COUNT), then hands out fill slots from LDS cursors seeded with the
// tiles' scanned offsets (EMIT).  No global atomics, and the work arrays need no clearing.

// Chord range [c0, c1) of local path p of scene sc (seg_cnt holds the scanned offsets).
__device__ __forceinline__ void path_chords(const BatchDev &b, const SceneDev &sc, uint32_t p, uint32_t &c0,
                                            uint32_t &c1) {
    const uint32_t g0 = sc.seg_base + sc.path_seg_off[p], g1 = sc.seg_base + sc.path_seg_off[p + 1];
    const uint32_t total = b.counters[C_LINES];
    c0 = g0 < b.n_segs ? b.seg_cnt[g0] : total;
    c1 = g1 < b.n_segs ? b.seg_cnt[g1] : total;
    c0 = min(c0, b.cap_lines);
    c1 = min(c1, b.cap_lines);
}

__device__ __forceinline__ void wave_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// The chords [c0, c1) of a path, thread `first` of `stride` taking every stride-th.  Most
// chords lie inside one tile, where process_line_segment emits exactly one fill (from, from +
// vec) and no auxiliary fill or backdrop step: the tile crossings it would compute are all at
// t >= 1.  Those are emitted at once; the others (crossing a tile boundary) are queued per wave
// in LDS (wq, WALK_QUEUE slots) and walked 64 at a time, so the DDA's divisions and steps run
// with every lane on a chord that needs them.
template <bool EMIT, bool LDS, bool PACK = false>
__device__ __forceinline__ void walk_chords(const BatchDev &b, const SceneDev &sc, const int4 &rect, uint32_t tbase,
                                            uint32_t cbase, uint32_t *l_cnt, int32_t *l_delta, int32_t *l_col,
                                            uint32_t c0, uint32_t c1, uint32_t first, uint32_t stride,
                                            uint32_t *wq) {
    const uint32_t lane = threadIdx.x & 63;
    const uint64_t lt_mask = lane == 0 ? 0ull : (~0ull >> (64 - lane));
    uint32_t guard = 0;
    Walker<EMIT, LDS, PACK> w;
    w.vb_l = sc.vb_l;
    w.vb_r = sc.vb_r;
    w.vb_b = sc.vb_b;
    w.tl = rect.x;
    w.tt = rect.y;
    w.tr = rect.z;
    w.tb = rect.w;
    w.tw = rect.z - rect.x;
    w.th = rect.w - rect.y;
    w.tbase = tbase;
    w.cbase = cbase;
    w.tile_cnt = b.tile_cnt;
    w.tile_delta = b.tile_delta;
    w.col_bd = b.col_bd;
    w.l_cnt = l_cnt;
    w.l_delta = l_delta;
    w.l_col = l_col;
    w.out = b.dump ? b.dump_fills : nullptr;
    w.fills = b.fills;
    w.tile_fill_off = b.tile_fill_off;
    w.cap_fills = b.cap_fills;
    auto begin = [&](uint32_t i) {
        w.pend_tile = 0xffffffffu;
        w.pend_n = 0;
        w.count = 0;
        w.guard = 0;
        w.out_base = (EMIT && b.dump) ? b.line_cnt[i] : 0u;
    };
    auto end = [&](uint32_t i) {
        w.flush();
        if (!EMIT && b.dump) b.line_cnt[i] = w.count;  // fills per chord: canonical dump order only
        guard += w.guard;
    };
    uint32_t qn = 0;  // queued chords (wave-uniform)
    // chord i: stream points i and i + 1 of the path's range (the last point pairs with a NaN:
    // nothing is emitted, and the dump's per-chord count there is 0)
    auto chord = [&](uint32_t i) {
        const float2 a = b.lines[i];
        const float2 e = i + 1 < c1 ? b.lines[i + 1] : make_float2(__builtin_nanf(""), __builtin_nanf(""));
        return make_float4(a.x, a.y, e.x, e.y);
    };
    // count pass: the next chunk's chord is loaded while this one is processed (measured: the
    // count passes gain ~5 %, the emit passes, whose stores also wait, lose ~1 %)
    const float4 zero4 = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    float4 ln_next = !EMIT && c0 + first < c1 ? chord(c0 + first) : zero4;
    for (uint32_t base = c0; base < c1; base += stride) {  // uniform over the wave
        const uint32_t i = base + first;
        const float4 ln_cur = ln_next;
        if (!EMIT && base + stride + first < c1) ln_next = chord(base + stride + first);
        bool multi = false;
        if (i < c1) {
            const float4 ln = EMIT ? chord(i) : ln_cur;
            float2 from = make_float2(ln.x, ln.y), to = make_float2(ln.z, ln.w);
            begin(i);
            if (isfinite(from.x) && isfinite(from.y) && isfinite(to.x) && isfinite(to.y) && w.clip(from, to)) {
                const float inv = 1.0f / 16.0f;
                const int fx = (int)floorf(from.x * inv), fy = (int)floorf(from.y * inv);
                const int ex = (int)floorf(to.x * inv), ey = (int)floorf(to.y * inv);
                const float2 vec = f2sub(to, from);
                // a -0 component makes a crossing time -inf: left to the full walk
                const bool neg_zero = (vec.x == 0.0f && signbit(vec.x)) || (vec.y == 0.0f && signbit(vec.y));
                if (fx == ex && fy == ey && !neg_zero) w.add_fill(from, f2add(from, vec), fx, fy);
                else multi = true;
            }
            if (!multi) end(i);
        }
        const uint64_t m = __ballot(multi);
        if (multi) wq[qn + (uint32_t)__popcll(m & lt_mask)] = i;
        qn += (uint32_t)__popcll(m);
        wave_sync();
        if (qn >= 64) {  // at most 127 queued: one full batch
            qn -= 64;
            const uint32_t ci = wq[qn + lane];
            wave_sync();  // read before the next chunk's chords overwrite the slots
            const float4 ln = chord(ci);
            begin(ci);
            w.line(make_float2(ln.x, ln.y), make_float2(ln.z, ln.w));
            end(ci);
        }
    }
    if (lane < qn) {  // the rest of the queue
        const uint32_t ci = wq[lane];
        const float4 ln = chord(ci);
        begin(ci);
        w.line(make_float2(ln.x, ln.y), make_float2(ln.z, ln.w));
        end(ci);
    }
    wave_sync();  // the queue is read before the caller reuses the wave's LDS
    if (!EMIT && guard) atomicAdd(&b.counters[C_GUARD], guard);
}


// Per-path totals and offsets.  The count pass stores each path's {fills, alpha tiles}
// (path_tot, batch path order); after their scan the emit pass turns its rect's tile counts
// into the tiles' fill offsets and alpha-tile indices (a local scan, row-major), writes
// them, and uses the fill offsets as its LDS cursors.  The batch-wide order of tiles is
// path order then row-major within the rect, so this equals a scan over all tiles.

__device__ __forceinline__ U2 wave_sum_u2(U2 v) {
#pragma unroll
    for (int d = 32; d > 0; d >>= 1) {
        v.x += __shfl_xor(v.x, d, 64);
        v.y += __shfl_xor(v.y, d, 64);
    }
    return v;
}

// Inclusive prefix sums of (count, count > 0) across the wave.
__device__ __forceinline__ uint32_t alpha_rows(const BinItem &it) { return it.level != 0 ? 2u : 1u; }

__device__ __forceinline__ U2 wave_inclusive_u2(uint32_t c) {
    U2 v{c, c > 0 ? 1u : 0u};
    const int lane = threadIdx.x & 63;
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
        const U2 o = shfl_up_val(v, d);
        if (lane >= d) v = v + o;
    }
    return v;
}

// Wave: tile offsets of a rect of `area` tiles (counts in cnt_src, global) from the path's
// scanned base; offsets go to the tile arrays and, when given, to the LDS cursors.
// aw: mask rows per alpha tile (2 for a clip path's tiles, whose coverage is kept under both
// fill rules; 1 otherwise).
__device__ __forceinline__ void wave_tile_offsets(const BatchDev &b, uint32_t tbase, uint32_t area, U2 base,
                                                  uint32_t *l_cur, uint32_t aw) {
    const uint32_t lane = threadIdx.x & 63;
    for (uint32_t j0 = 0; j0 < area; j0 += 64) {
        const uint32_t j = j0 + lane;
        const uint32_t c = j < area ? b.tile_cnt[tbase + j] : 0u;
        const U2 inc = wave_inclusive_u2(c);
        if (j < area) {
            const uint32_t off = base.x + inc.x - c, ai = base.y + aw * (inc.y - (c > 0 ? 1u : 0u));
            b.tile_fill_off[tbase + j] = off;
            b.tile_alpha_idx[tbase + j] = ai;
            if (l_cur) l_cur[j] = off;
        }
        base = base + U2{__shfl(inc.x, 63, 64), aw * __shfl(inc.y, 63, 64)};
    }
}

// Block (256 threads) version, called by every thread of the block.
__device__ __forceinline__ void block_tile_offsets(const BatchDev &b, uint32_t tbase, uint32_t area, U2 base,
                                                   uint32_t *l_cur, uint32_t aw) {
    for (uint32_t j0 = 0; j0 < area; j0 += 256) {
        const uint32_t j = j0 + threadIdx.x;
        const uint32_t c = j < area ? b.tile_cnt[tbase + j] : 0u;
        U2 tot;
        const U2 ex = block_exclusive_scan<U2>(U2{c, c > 0 ? 1u : 0u}, tot);
        if (j < area) {
            b.tile_fill_off[tbase + j] = base.x + ex.x;
            b.tile_alpha_idx[tbase + j] = base.y + aw * ex.y;
            if (l_cur) l_cur[j] = base.x + ex.x;
        }
        base = base + U2{tot.x, aw * tot.y};
    }
}

// Block sum over the block's threads; `part` is 4 slots of LDS the block no longer reads.
__device__ __forceinline__ U2 block_sum_u2(U2 v, U2 *part) {
    v = wave_sum_u2(v);
    __syncthreads();  // part may alias data read just before
    if ((threadIdx.x & 63) == 0) part[threadIdx.x >> 6] = v;
    __syncthreads();
    const U2 t = part[0] + part[1] + part[2] + part[3];
    __syncthreads();
    return t;
}

// One descriptor per binned path: thread per item, class GROUP 0 (wave) or 1 (group).
template <int GROUP>
__global__ __launch_bounds__(256) void k_bin_items(BatchDev b) {
    const uint32_t n = GROUP ? b.n_bin_group : b.n_bin_wave;
    const uint32_t *bases = GROUP ? b.bin_group_bases : b.bin_wave_bases;
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    // every lane searches (lanes past the end with the last item), then those leave
    const int k = find_scene_lanes(bases, b.n_scenes, min(i, n - 1));
    if (i >= n) return;
    const SceneDev &sc = b.scenes[k];
    const uint32_t p = sc.bin_paths[(GROUP ? sc.n_bin_wave : 0u) + (i - bases[k])];
    BinItem it;
    it.rect = sc.path_rect[p];
    it.tbase = sc.tile_base + sc.path_tile_off[p];
    it.cbase = sc.col_base + sc.path_col_off[p];
    path_chords(b, sc, p, it.c0, it.c1);
    it.p = p;
    it.gp = sc.path_base + p;
    it.k = (uint32_t)k;
    it.rgba = sc.path_rgba[p];
    it.rule = sc.path_rule[p] | (sc.clip_own_rules ? RULE_CLIP_OWN : 0u);
    uint32_t cr = sc.path_clip ? sc.path_clip[p] : CLIP_NONE;
    if (cr < CLIP_EMPTY)
        cr = (cr & CLIP_GROUP) ? b.n_bin_wave + sc.bin_group_base + (cr & ~CLIP_GROUP) : sc.bin_wave_base + cr;
    it.clip = cr;
    it.level = sc.path_level ? (uint32_t)sc.path_level[p] : 0u;
    it.paint = sc.path_paint ? sc.path_paint[p] : PAINT_NONE;
    const uint32_t gi = (GROUP ? b.n_bin_wave : 0u) + i;
    b.bin_items[gi] = it;
    const uint32_t tw = (uint32_t)max(it.rect.z - it.rect.x, 0);
    for (uint32_t j = 0; j < tw; j++) b.col_item[it.cbase + j] = gi;
}

// Wave class: 4 paths per 256-thread block, one per wave, each with its own LDS slice.
template <bool EMIT>
__global__ __launch_bounds__(256) void k_bin_wave(BatchDev b) {
    __shared__ uint32_t s_cnt[4][BIN_WAVE_TILES];
    __shared__ int32_t s_delta[4][EMIT ? 1 : BIN_WAVE_TILES];
    __shared__ int32_t s_col[4][EMIT ? 1 : BIN_WAVE_COLS];
    __shared__ uint32_t s_queue[4][WALK_QUEUE];
    const uint32_t wv = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const uint32_t item = __builtin_amdgcn_readfirstlane(blockIdx.x * 4 + wv);  // uniform: scalar loads below
    if (item >= b.n_bin_wave) return;  // whole wave: no block-wide barriers below
    const BinItem it = b.bin_items[item];
    const SceneDev &sc = b.scenes[it.k];
    const int4 rect = it.rect;
    const uint32_t tw = (uint32_t)max(rect.z - rect.x, 0), th = (uint32_t)max(rect.w - rect.y, 0);
    const uint32_t area = tw * th;
    if (area > (uint32_t)BIN_WAVE_TILES || tw > (uint32_t)BIN_WAVE_COLS) return;  // not this class (host sorts)
    const uint32_t tbase = it.tbase, cbase = it.cbase;
    uint32_t *l_cnt = s_cnt[wv];
    int32_t *l_delta = s_delta[wv], *l_col = s_col[wv];
    const uint32_t gp = it.gp;
    if (EMIT) {
        const uint2 base = b.path_tot[gp];
        wave_tile_offsets(b, tbase, area, U2{base.x, base.y}, l_cnt, alpha_rows(it));
    } else {
        for (uint32_t j = lane; j < area; j += 64) {
            l_cnt[j] = 0;
            l_delta[j] = 0;
        }
        for (uint32_t j = lane; j < tw; j += 64) l_col[j] = 0;
    }
    wave_sync();
    walk_chords<EMIT, true>(b, sc, rect, tbase, cbase, l_cnt, l_delta, l_col, it.c0, it.c1, lane, 64, s_queue[wv]);
    if (!EMIT) {
        wave_sync();
        U2 tot{0, 0};
        for (uint32_t j = lane; j < area; j += 64) {
            const uint32_t c = l_cnt[j];
            b.tile_cnt[tbase + j] = c;
            b.tile_delta[tbase + j] = l_delta[j];
            tot = tot + U2{c, c > 0 ? 1u : 0u};
        }
        for (uint32_t j = lane; j < tw; j += 64) b.col_bd[cbase + j] = l_col[j];
        tot = wave_sum_u2(tot);
        if (lane == 0) b.path_tot[gp] = make_uint2(tot.x, tot.y * alpha_rows(it));
    }
}

// Group class: one path per 256-thread block; LDS when the rect fits, else global atomics
// on the path's own (first zeroed) tiles.
template <bool EMIT>
__global__ __launch_bounds__(256) void k_bin_group(BatchDev b) {
    __shared__ uint32_t s_cnt[BIN_GROUP_TILES];
    __shared__ uint32_t s_delta[EMIT ? 1 : (BIN_GROUP_TILES + 1) / 2];  // packed, two tiles per word
    __shared__ int32_t s_col[EMIT ? 1 : BIN_GROUP_COLS];
    __shared__ uint32_t s_queue[4][WALK_QUEUE];
    const uint32_t item = blockIdx.x;
    if (item >= b.n_bin_group) return;  // uniform over the block
    const BinItem it = b.bin_items[b.n_bin_wave + item];
    const SceneDev &sc = b.scenes[it.k];
    const int4 rect = it.rect;
    const uint32_t tw = (uint32_t)max(rect.z - rect.x, 0), th = (uint32_t)max(rect.w - rect.y, 0);
    const uint32_t area = tw * th;
    const uint32_t tbase = it.tbase, cbase = it.cbase;
    const uint32_t gp = it.gp;
    const uint32_t c0 = it.c0, c1 = it.c1;
    // the packed LDS deltas hold sums below 2^15 in magnitude: at most one step per chord
    const bool lds = area <= (uint32_t)BIN_GROUP_TILES && tw <= (uint32_t)BIN_GROUP_COLS &&
                     (EMIT || c1 - c0 < (1u << 15));
    if (lds) {
        if (EMIT) {
            const uint2 base = b.path_tot[gp];
            block_tile_offsets(b, tbase, area, U2{base.x, base.y}, s_cnt, alpha_rows(it));
        } else {
            for (uint32_t j = threadIdx.x; j < area; j += 256) s_cnt[j] = 0;
            for (uint32_t j = threadIdx.x; j < (area + 1) / 2; j += 256) s_delta[j] = 0x80008000u;
            for (uint32_t j = threadIdx.x; j < tw; j += 256) s_col[j] = 0;
        }
        __syncthreads();
        walk_chords<EMIT, true, true>(b, sc, rect, tbase, cbase, s_cnt, reinterpret_cast<int32_t *>(s_delta), s_col,
                                      c0, c1, threadIdx.x, 256, s_queue[threadIdx.x >> 6]);
        if (!EMIT) {
            __syncthreads();
            U2 tot{0, 0};
            for (uint32_t j = threadIdx.x; j < area; j += 256) {
                const uint32_t c = s_cnt[j];
                b.tile_cnt[tbase + j] = c;
                b.tile_delta[tbase + j] = (int32_t)((s_delta[j >> 1] >> (16 * (j & 1))) & 0xffffu) - 0x8000;
                tot = tot + U2{c, c > 0 ? 1u : 0u};
            }
            for (uint32_t j = threadIdx.x; j < tw; j += 256) b.col_bd[cbase + j] = s_col[j];
            tot = block_sum_u2(tot, reinterpret_cast<U2 *>(s_col));
            if (threadIdx.x == 0) b.path_tot[gp] = make_uint2(tot.x, tot.y * alpha_rows(it));
        }
    } else {
        if (EMIT) {
            const uint2 base = b.path_tot[gp];
            block_tile_offsets(b, tbase, area, U2{base.x, base.y}, nullptr, alpha_rows(it));
            __threadfence();  // the offsets reach L2 before any thread reads them back
            __syncthreads();
        }
        if (!EMIT) {
            for (uint32_t j = threadIdx.x; j < area; j += 256) {
                b.tile_cnt[tbase + j] = 0;
                b.tile_delta[tbase + j] = 0;
            }
            for (uint32_t j = threadIdx.x; j < tw; j += 256) b.col_bd[cbase + j] = 0;
            __threadfence();  // the zeros reach L2 before any thread's atomics
            __syncthreads();
        }
        walk_chords<EMIT, false>(b, sc, rect, tbase, cbase, nullptr, nullptr, nullptr, c0, c1, threadIdx.x, 256,
                                 s_queue[threadIdx.x >> 6]);
        if (!EMIT) {
            __threadfence();  // every thread's atomics are done before the counts are summed
            __syncthreads();
            U2 tot{0, 0};
            for (uint32_t j = threadIdx.x; j < area; j += 256) {
                const uint32_t c = __hip_atomic_load(&b.tile_cnt[tbase + j], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                tot = tot + U2{c, c > 0 ? 1u : 0u};
            }
            tot = block_sum_u2(tot, reinterpret_cast<U2 *>(s_col));  // s_col is unused on this path
            if (threadIdx.x == 0) b.path_tot[gp] = make_uint2(tot.x, tot.y * alpha_rows(it));
        }
    }
}

// ---------------------------------------------------------------------------



// Order does not matter for the block sum: each pass reads one coalesced row of 256 items.
template <typename T, typename Load>
__global__ __launch_bounds__(SCAN_THREADS) void k_scan_reduce(Load ld, uint32_t n, T *block_sums) {
    const uint64_t base = (uint64_t)blockIdx.x * SCAN_TILE + threadIdx.x;
    T s = zero_of<T>();
#pragma unroll
    for (int i = 0; i < SCAN_ITEMS; i++) {
        const uint64_t j = base + (uint64_t)i * SCAN_THREADS;
        if (j < n) s = s + ld((uint32_t)j);
    }
    T total;
    block_exclusive_scan<T>(s, total);
    if (threadIdx.x == 0) block_sums[blockIdx.x] = total;
}

// Single block: exclusive scan of nb block sums in place; block_sums[nb] = grand total.
// Each thread takes SCAN_ITEMS consecutive sums per round (the array is small and stays in
// cache), so a round covers SCAN_TILE sums.
template <typename T>
__global__ __launch_bounds__(SCAN_THREADS) void k_scan_blocks(T *block_sums, uint32_t nb) {
    T carry = zero_of<T>();
    for (uint32_t start = 0; start < nb; start += SCAN_TILE) {
        const uint32_t first = start + threadIdx.x * SCAN_ITEMS;
        T vals[SCAN_ITEMS];
        T s = zero_of<T>();
#pragma unroll
        for (int i = 0; i < SCAN_ITEMS; i++) {
            vals[i] = first + i < nb ? block_sums[first + i] : zero_of<T>();
            s = s + vals[i];
        }
        T total;
        T run = carry + block_exclusive_scan<T>(s, total);
#pragma unroll
        for (int i = 0; i < SCAN_ITEMS; i++) {
            if (first + i < nb) block_sums[first + i] = run;
            run = run + vals[i];
        }
        carry = carry + total;
    }
    if (threadIdx.x == 0) block_sums[nb] = carry;
}

// Item s of a tile sits at LDS word s + s / 16: striped rows (lane-consecutive) and blocked
// runs (16 per thread, threads 17 words apart) are both bank-conflict free.
__device__ __forceinline__ uint32_t scan_pad(uint32_t s) { return s + (s >> 4); }

// Loads and stores are striped (one coalesced row of 256 items at a time) and transposed
// through LDS to the blocked order the per-thread running sum needs.
template <typename T, typename Load, typename Store>
__global__ __launch_bounds__(SCAN_THREADS) void k_scan_apply(Load ld, Store st, uint32_t n, const T *block_sums,
                                                             uint32_t nb) {
    __shared__ T tile[SCAN_TILE + SCAN_TILE / 16];
    const uint32_t tid = threadIdx.x;
    const uint64_t base = (uint64_t)blockIdx.x * SCAN_TILE;
#pragma unroll
    for (int i = 0; i < SCAN_ITEMS; i++) {
        const uint32_t si = (uint32_t)i * SCAN_THREADS + tid;
        const uint64_t j = base + si;
        tile[scan_pad(si)] = j < n ? ld((uint32_t)j) : zero_of<T>();
    }
    __syncthreads();
    T vals[SCAN_ITEMS];
    T s = zero_of<T>();
#pragma unroll
    for (int i = 0; i < SCAN_ITEMS; i++) {
        vals[i] = tile[scan_pad(tid * SCAN_ITEMS + (uint32_t)i)];
        s = s + vals[i];
    }
    T total;
    T run = block_exclusive_scan<T>(s, total) + block_sums[blockIdx.x];  // its barriers order the reads above
#pragma unroll
    for (int i = 0; i < SCAN_ITEMS; i++) {
        tile[scan_pad(tid * SCAN_ITEMS + (uint32_t)i)] = run;
        run = run + vals[i];
    }
    __syncthreads();
#pragma unroll
    for (int i = 0; i < SCAN_ITEMS; i++) {
        const uint32_t si = (uint32_t)i * SCAN_THREADS + tid;
        const uint64_t j = base + si;
        if (j < n) st((uint32_t)j, tile[scan_pad(si)]);
    }
    if (blockIdx.x == nb - 1 && threadIdx.x == 0) st(n, block_sums[nb]);
}

// Scan functors --------------------------------------------------------------

struct LoadU32 {
    const uint32_t *p;
    __device__ uint32_t operator()(uint32_t i) const { return p[i]; }
};
struct StoreU32 {
    uint32_t *p;
    __device__ void operator()(uint32_t i, uint32_t v) const { p[i] = v; }
};
// Tiles: (fill count, has fills) -> fill offsets, alpha indices; totals -> counters.
struct LoadPathTot {
    const uint2 *tot;
    __device__ U2 operator()(uint32_t i) const {
        const uint2 v = tot[i];
        return U2{v.x, v.y};
    }
};
struct StorePathOff {
    uint2 *off;
    uint32_t *fill_end, *alpha_end, *counters;
    uint32_t n;
    __device__ void operator()(uint32_t i, U2 v) const {
        off[i] = make_uint2(v.x, v.y);
        if (i == n) {
            *fill_end = v.x;
            *alpha_end = v.y;
            counters[C_FILLS] = v.x;
            counters[C_ALPHA] = v.y;
        }
    }
};
struct LoadU32Counted {
    const uint32_t *p;
    const uint32_t *n_ptr;
    __device__ uint32_t operator()(uint32_t i) const { return i < *n_ptr ? p[i] : 0u; }
};
struct StoreCounted {
    uint32_t *p, *counters;
    int counter;
    uint32_t n;
    __device__ void operator()(uint32_t i, uint32_t v) const {
        p[i] = v;
        if (i == n && counter >= 0) counters[counter] = v;
    }
};
struct StoreEntryOff {
    uint2 *off;
    uint32_t *counters;
    uint32_t n;
    __device__ void operator()(uint32_t i, U2 v) const {
        off[i] = make_uint2(v.x, v.y);
        if (i == n) {
            counters[C_ENTRIES] = v.x;
            counters[C_ALPHA_ENTRIES] = v.y;
        }
    }
};

// ---------------------------------------------------------------------------
// Backdrop propagation down each path column (tiler.cpp:369-439) + z-buffer
// ---------------------------------------------------------------------------
//
// The column kernels walk a path column top to bottom, one thread per column.  Their
// loads do not depend on earlier rows (only the running backdrop does, in registers), so
// rows are taken COL_ROWS at a time: every load of the group is issued before the first
// is used, and a wave waits on memory once per group instead of once per row.

constexpr int COL_ROWS = 8;

// Where a tile of a path column lands in its scene's framebuffer-tile grid.
struct ColumnWalk {
    uint32_t p, rgba, rule, tbase, clip, level;
    bool own_rules;  // clip coverage under each clip path's own fill rule (PFH_CLIP_RULE_OWN)
    int rx, ry;  // the path's rect origin (tiles)
    int w, h, x, fb_row0;  // fb_row0: fb tile index (scene-local) of row 0
    uint32_t fb_base, fb_w;
    // column c of the batch, through its bin item (every column belongs to a binned path)
    __device__ ColumnWalk(const BatchDev &b, uint32_t c) {
        const BinItem it = b.bin_items[b.col_item[c]];
        const SceneDev &sc = b.scenes[it.k];
        const int4 r = it.rect;
        p = it.p;
        w = r.z - r.x;
        h = r.w - r.y;
        x = (int)(c - it.cbase);
        tbase = it.tbase;
        rgba = it.rgba;
        rule = it.rule & 1u;
        own_rules = (it.rule & RULE_CLIP_OWN) != 0;
        clip = it.clip;
        level = it.level;
        rx = r.x;
        ry = r.y;
        fb_base = sc.fb_base;
        fb_w = (uint32_t)sc.fb_w;
        fb_row0 = (r.y - sc.fb_y0) * sc.fb_w + (r.x + x - sc.fb_x0);
    }
    __device__ uint32_t tile(int row) const { return tbase + (uint32_t)(row * w + x); }
    __device__ uint32_t fb_local(int row) const { return (uint32_t)(fb_row0 + row * (int)fb_w); }
};

// A clip path's tile at given tile coordinates, as resolved by the propagate phase of its clip
// level (an earlier launch): whether it has a coverage mask (its own, or one taken from its own
// clip), whether it covers without one (solid), its backdrop, and its coverage source (the mask
// row whose record holds the chain).
struct ClipTile {
    bool inside, alpha, solid;
    int8_t bd;
    uint32_t src;
};
__device__ __forceinline__ ClipTile clip_tile_at(const BatchDev &b, const BinItem &ci, int tx, int ty) {
    ClipTile r{false, false, false, 0, 0};
    const int4 cr = ci.rect;
    if (tx < cr.x || tx >= cr.z || ty < cr.y || ty >= cr.w) return r;
    const uint32_t ct = ci.tbase + (uint32_t)((ty - cr.y) * (cr.z - cr.x) + (tx - cr.x));
    const uint8_t st = b.tile_st[ct];
    r.inside = true;
    r.alpha = st == TILE_ALPHA || st == TILE_CLIP_MASK;
    r.solid = st == TILE_SOLID;
    r.bd = b.tile_bd[ct];
    r.src = r.alpha ? b.tile_clip_ai[ct] : 0u;
    return r;
}

// Column walk down a path's tiles (core/d3d9/tiler.cpp:369-439): backdrops, tile status, alpha
// records, z-buffer, and clip resolution (tiler.cpp:385-431) against the clip path's resolved
// tile at the same coordinates: outside its rect or over a blank tile -> culled; both with masks ->
// a combine job (the tile's backdrop moves into it); a solid tile over a mask -> it takes that
// mask; otherwise unchanged (a solid clip tile passes the tile through).
// phase 0: every column (no clip paths in the batch); 1..MAX_CLIP_LEVELS: the columns of clip paths
// of that level, a clip path nested in another being resolved against it like a draw path (the
// level structure of core/d3d11/scene_builder.cpp:92-133, shaders/d3d11/propagate.comp:133-176);
// MAX_CLIP_LEVELS + 1: the draw paths' columns.  Each phase reads only tiles of earlier phases.
// CLIPS: the batch has clip paths (without them the kernel is the plain per-column walk).
// What a tile of a path column reads of its clip path: the clip's tile at the same coordinates
// and, for a clip path's tile over a clip mask, the chain record of that mask.  Both were written
// by an earlier propagate phase, so a column loads them for a group of rows before it stores
// anything (no load waits behind a store it might alias).
struct ClipRead {
    ClipTile ct;
    uint4 pr;
};
__device__ __forceinline__ ClipRead clip_read(const BatchDev &b, const ColumnWalk &cw, const BinItem &ci, int row) {
    ClipRead r{{false, false, false, 0, 0}, make_uint4(1u, 0u, 0u, 0u)};
    if (cw.clip < CLIP_EMPTY) r.ct = clip_tile_at(b, ci, cw.rx + cw.x, cw.ry + row);
    if (cw.level != 0 && r.ct.alpha && r.ct.src < b.cap_alpha) r.pr = b.clip_rec[r.ct.src];
    return r;
}

// One tile of a path column, its propagated backdrop bd and its fill range [f0, f1) known (ai: its
// mask row when it has fills), with what it reads of its clip path (CLIPS).
template <bool CLIPS>
__device__ __forceinline__ void propagate_tile(const BatchDev &b, const ColumnWalk &cw, const ClipRead &cr, int row,
                                               int8_t bd, uint32_t f0, uint32_t f1, uint32_t ai) {
    const bool is_clip = CLIPS && cw.level != 0;
    const bool clipped = CLIPS && (cw.clip < CLIP_EMPTY || cw.clip == CLIP_EMPTY);
    const bool opaque = (cw.rgba >> 24) == 255u;
    const uint32_t t = cw.tile(row);
    const bool alpha = f1 > f0;
    const ClipTile &ct = cr.ct;
    // Whether the tile covers without fills: backdrop != 0, except under the path's own even-odd
    // rule with an even backdrop where clip paths' own rules are taken (PFH_CLIP_RULE_OWN); the
    // reference's clip resolution tests backdrop != 0 alone (tiler.cpp:396-431).
    const bool covers = bd != 0 && !(CLIPS && cw.own_rules && cw.rule && (bd & 1) == 0);
    const uint32_t job_rules = CLIPS && cw.own_rules ? CLIP_JOB_OWN_RULES : 0u;
    if (CLIPS && is_clip) {
        // A clip path draws nothing itself.  Its fills make two mask rows (ai, ai + 1) that
        // keep |sum + backdrop| under the nonzero and the even-odd rule (under its own rule,
        // both, where clip paths' own rules are taken), for the tiles that take them over; its
        // coverage at the tile is that value, min'd with its own clip's where it is nested and
        // both have masks (a combine job rewrites both rows).
        const bool own = alpha && ai + 1 < b.cap_alpha;
        if (own) {
            const uint32_t r0 = cw.own_rules ? cw.rule : 0u, r1 = cw.own_rules ? cw.rule : 1u;
            b.alpha_rec[ai] = make_uint4(f0, f1, (uint32_t)(uint8_t)bd | (r0 << 8), t);
            b.alpha_rec[ai + 1] = make_uint4(f0, f1, (uint32_t)(uint8_t)bd | (r1 << 8), t);
        }
        uint8_t st = alpha ? TILE_ALPHA : (covers ? TILE_SOLID : TILE_EMPTY);
        int8_t out_bd = bd;
        uint32_t src = ai;
        uint4 rec = make_uint4(1u, 0u, 0u, 0u);  // the chain: this row alone
        if (clipped) {
            if (!ct.inside || (!ct.alpha && !ct.solid)) {
                st = TILE_EMPTY;
                out_bd = 0;
            } else if (ct.alpha && alpha) {
                out_bd = 0;
                const uint4 pr = cr.pr;
                rec = make_uint4(1u + min(pr.x, MAX_CLIP_LEVELS - 1u), ct.src, pr.y, pr.z);
                const uint32_t j = atomicAdd(&b.counters[C_CLIP_JOBS], 1u);
                if (j < b.cap_alpha) b.clip_jobs[j] = make_uint2(ai | CLIP_JOB_CLIP | job_rules, ai);
            } else if (ct.alpha && covers) {
                st = TILE_CLIP_MASK;  // its own clip's coverage, unchanged
                out_bd = ct.bd;
                src = ct.src;
            }
        }
        if (st == TILE_ALPHA && own) b.clip_rec[ai] = rec;
        if (st == TILE_ALPHA || st == TILE_CLIP_MASK) b.tile_clip_ai[t] = src;
        b.tile_st[t] = st;
        b.tile_bd[t] = out_bd;
        return;
    }
    if (CLIPS && clipped) {
        if (alpha && ai < b.cap_alpha)
            b.alpha_rec[ai] = make_uint4(f0, f1, (uint32_t)(uint8_t)bd | (cw.rule << 8), t);
        uint8_t st;
        int8_t out_bd = 0;
        if (!ct.inside || (!ct.alpha && !ct.solid)) {
            st = TILE_EMPTY;  // outside the clip's rect or over a blank clip tile: culled
        } else if (ct.alpha && alpha) {
            // both masks: min(|own + backdrop|, clip coverage) replaces the mask; backdrop 0
            st = TILE_ALPHA;
            const uint32_t j = atomicAdd(&b.counters[C_CLIP_JOBS], 1u);
            if (j < b.cap_alpha) b.clip_jobs[j] = make_uint2(ai | job_rules, ct.src);
        } else if (ct.alpha && covers) {
            st = TILE_CLIP_MASK;  // solid over a clip mask: the clip's mask and backdrop
            out_bd = ct.bd;
            b.tile_clip_ai[t] = ct.src + cw.rule;
        } else {
            // unchanged: an alpha tile, or a solid tile over a solid clip tile
            const bool even = cw.rule && (bd & 1) == 0;
            st = alpha ? TILE_ALPHA : (bd == 0 ? TILE_EMPTY : (even ? TILE_CLEAR : TILE_SOLID));
            out_bd = bd;
            if (st == TILE_SOLID && opaque) atomicMax(&b.zbuf[cw.fb_base + cw.fb_local(row)], cw.p);
        }
        b.tile_st[t] = st;
        if (b.dump) b.tile_bd[t] = out_bd;
        return;
    }
    const bool even = cw.rule && (bd & 1) == 0;
    b.tile_st[t] = alpha ? TILE_ALPHA : (bd == 0 ? TILE_EMPTY : (even ? TILE_CLEAR : TILE_SOLID));
    if (b.dump) b.tile_bd[t] = bd;
    if (alpha) {
        if (ai < b.cap_alpha) b.alpha_rec[ai] = make_uint4(f0, f1, (uint32_t)(uint8_t)bd | (cw.rule << 8), t);
    } else if (opaque && bd != 0 && !even) {
        atomicMax(&b.zbuf[cw.fb_base + cw.fb_local(row)], cw.p);
    }
}

// Column walk down a path's tiles (core/d3d9/tiler.cpp:369-439): backdrops, tile status, alpha
// records, z-buffer, and clip resolution (tiler.cpp:385-431) against the clip path's resolved
// tile at the same coordinates: outside its rect or over a blank tile -> culled; both with masks ->
// a combine job (the tile's backdrop moves into it); a solid tile over a mask -> it takes that
// mask; otherwise unchanged (a solid clip tile passes the tile through).
// phase 0: every column (no clip paths in the batch); 1..MAX_CLIP_LEVELS: the columns of clip paths
// of that level, a clip path nested in another being resolved against it like a draw path (the
// level structure of core/d3d11/scene_builder.cpp:92-133, shaders/d3d11/propagate.comp:133-176);
// MAX_CLIP_LEVELS + 1: the draw paths' columns.  Each phase reads only tiles of earlier phases.
// CLIPS: the batch has clip paths (without them the kernel is the plain per-column walk).
// A thread walks its column: rows of neighbouring columns are neighbours in memory, so the
// wave's loads coalesce.  A walk is a chain of dependent loads per group of rows, and a launch
// lasts as long as its longest walk: a column of more than `rows` tiles (a schedule choice) is
// listed instead for k_propagate_rows, which spreads its rows over 16 or 64 lanes (rows 0: none).
template <bool CLIPS>
__global__ __launch_bounds__(256) void k_propagate(BatchDev b, uint32_t phase, uint32_t rows) {
    uint32_t c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= b.n_cols) return;
    const ColumnWalk cw(b, c);
    if (CLIPS && cw.level != (phase > MAX_CLIP_LEVELS ? 0u : phase)) return;
    if (rows != 0 && (uint32_t)cw.h > rows) {
        // at most one entry per column and phase, so the two lists together hold at most n_cols
        if (cw.h > TALL_ROWS) b.tall_cols[b.n_cols - 1 - atomicAdd(&b.counters[C_TALL + phase], 1u)] = c;
        else b.tall_cols[atomicAdd(&b.counters[C_MID + phase], 1u)] = c;
        return;
    }
    BinItem ci;
    if (CLIPS && cw.clip < CLIP_EMPTY) ci = b.bin_items[cw.clip];
    int32_t running = b.col_bd[c];
    for (int row0 = 0; row0 < cw.h; row0 += COL_ROWS) {
        int32_t delta[COL_ROWS];
        uint32_t f0[COL_ROWS], f1[COL_ROWS];
#pragma unroll
        for (int i = 0; i < COL_ROWS; i++) {
            delta[i] = 0;
            f0[i] = f1[i] = 0;
            if (row0 + i < cw.h) {
                const uint32_t t = cw.tile(row0 + i);
                delta[i] = b.tile_delta[t];
                f0[i] = b.tile_fill_off[t];
                f1[i] = b.tile_fill_off[t + 1];
            }
        }
        int8_t bd[COL_ROWS];
        uint32_t ai[COL_ROWS];
        ClipRead cr[COL_ROWS];
#pragma unroll
        for (int i = 0; i < COL_ROWS; i++) {
            bd[i] = (int8_t)running;
            running += (int8_t)delta[i];
            ai[i] = 0;
            // with clip paths the group's loads go out at once (the walk is latency-bound there)
            if (row0 + i < cw.h && (CLIPS || f1[i] > f0[i])) ai[i] = b.tile_alpha_idx[cw.tile(row0 + i)];
            cr[i] = ClipRead{{false, false, false, 0, 0}, make_uint4(1u, 0u, 0u, 0u)};
            if (CLIPS && row0 + i < cw.h) cr[i] = clip_read(b, cw, ci, row0 + i);
        }
#pragma unroll
        for (int i = 0; i < COL_ROWS; i++) {
            if (row0 + i >= cw.h) break;
            propagate_tile<CLIPS>(b, cw, cr[i], row0 + i, bd[i], f0[i], f1[i], ai[i]);
        }
    }
    b.col_bd[c] = running;
}

// The columns k_propagate listed for this phase (SUB 64: the C_TALL list, SUB 16: the C_MID list),
// SUB lanes per column: a column's lanes take SUB rows at a time, and a prefix sum of the rows'
// backdrop deltas across them gives each row the backdrop the serial walk reaches there (the int8
// deltas summed in int32, as the walk does).  Every wave strides over the list (its length is
// only known on the device), 64 / SUB columns at a time.
template <bool CLIPS, int SUB>
__global__ __launch_bounds__(256) void k_propagate_rows(BatchDev b, uint32_t phase) {
    constexpr uint32_t G = 64 / SUB;  // columns per wave
    const uint32_t n = min(b.counters[(SUB == 64 ? C_TALL : C_MID) + phase], b.n_cols);
    const int lane = (int)(threadIdx.x & 63), li = lane % SUB, g = lane / SUB;
    const uint32_t waves = gridDim.x * 4;
    for (uint32_t j0 = (blockIdx.x * 4 + (threadIdx.x >> 6)) * G; j0 < n; j0 += waves * G) {  // wave-uniform
        const uint32_t j = j0 + (uint32_t)g;
        const bool has = j < n;
        const uint32_t c = has ? b.tall_cols[SUB == 64 ? b.n_cols - 1 - j : j] : 0u;
        const ColumnWalk cw(b, c);
        const int h = has ? cw.h : 0;
        BinItem ci;
        if (CLIPS && has && cw.clip < CLIP_EMPTY) ci = b.bin_items[cw.clip];
        int32_t running = has ? b.col_bd[c] : 0;
        int hmax = h;
#pragma unroll
        for (uint32_t k = 0; k < G; k++) hmax = max(hmax, __builtin_amdgcn_readlane(h, k * SUB));
        for (int row0 = 0; row0 < hmax; row0 += SUB) {
            const int row = row0 + li;
            const bool on = row < h;
            int32_t d8 = 0;
            uint32_t f0 = 0, f1 = 0, ai = 0;
            ClipRead cr{{false, false, false, 0, 0}, make_uint4(1u, 0u, 0u, 0u)};
            if (on) {
                const uint32_t t = cw.tile(row);
                d8 = (int8_t)b.tile_delta[t];
                f0 = b.tile_fill_off[t];
                f1 = b.tile_fill_off[t + 1];
                ai = b.tile_alpha_idx[t];
                if (CLIPS) cr = clip_read(b, cw, ci, row);
            }
            uint32_t incl;
            if (SUB == 16) {
                incl = row16_inclusive_sum((uint32_t)d8);
            } else {
                incl = (uint32_t)d8;
#pragma unroll
                for (int d = 1; d < 64; d <<= 1) {
                    const uint32_t o = shfl_up_val(incl, d);
                    if (lane >= d) incl += o;
                }
            }
            if (on) propagate_tile<CLIPS>(b, cw, cr, row, (int8_t)(running + (int32_t)(incl - (uint32_t)d8)), f0, f1, ai);
            running += (int32_t)__shfl(incl, g * SUB + SUB - 1, 64);
        }
        if (has && li == 0) b.col_bd[c] = running;
    }
}

// Draw entries: tiles with a mask or a nonzero backdrop that survive the z-cull
// (core/d3d9/scene_builder.cpp:57-79, shaders/d3d9/tile.vert:103-118), per path like the
// bin kernels (a wave per small rect, a block per large one) with the threads striding over
// the rect's tiles in row-major order.  The count pass stores each path's entry count;
// after their scan the write pass numbers the path's entries in row-major order by ballot
// rank, so consecutive threads write consecutive entries.  Entries of a path precede those
// of the next, which is all the sort by framebuffer tile needs to keep draw order: a path
// has at most one tile per framebuffer tile.

struct EntryTile {
    bool keep;
    uint32_t st, fbl, ai;  // ai: the alpha index of an alpha tile (write pass)
};

// e: the entry's index; ea: its index among alpha entries (for an alpha tile).  The alpha
// entries' tiles are listed in that order for the fill kernel, which rasterizes only them.
// With clip paths, the solid tiles of several draw paths can take one clip mask row: the first
// entry to list it claims it in clip_seen, the others list nothing (NO_ALPHA: the fill kernel
// skips the slot), so the row is rasterized once.
template <bool CLIPS>
__device__ __forceinline__ void write_entry(const BatchDev &b, const SceneDev &sc, uint32_t p, uint32_t rgba,
                                            uint32_t paint, uint32_t tbase, uint32_t j, const EntryTile &et,
                                            uint32_t e, uint32_t ea) {
    if (e >= b.cap_entries) return;  // past the capacity: the batch is retried
    const uint32_t ai = (et.st == TILE_ALPHA || et.st == TILE_CLIP_MASK) ? et.ai : NO_ALPHA;
    if (ai != NO_ALPHA && ea < b.cap_vis) {
        uint32_t listed = ai;
        if (CLIPS && et.st == TILE_CLIP_MASK && ai < b.cap_alpha) {
            const uint32_t bit = 1u << (ai & 31u);
            if (atomicOr(&b.clip_seen[ai >> 5], bit) & bit) {
                listed = NO_ALPHA;
                atomicAdd(&b.counters[C_CLIP_DUPS], 1u);
            }
        }
        b.alpha_vis[ea] = listed;
    }
    // even-odd solid tile with an even winding covers nothing: zero alpha
    uint32_t rx = ai, ry = et.st == TILE_CLEAR ? (rgba & 0x00ffffffu) : rgba;
    // a path with a paint record: flagged, its mask row (or none) and its paint; a tile that covers
    // nothing stays a transparent solid entry, unless its paint composites through an op (which
    // sets the alpha of every pixel of the tile, shaders/d3d9/tile.frag:585-598)
    if (paint != PAINT_NONE) {
        if (et.st != TILE_CLEAR) {
            rx = ENTRY_GRADIENT | (ai == NO_ALPHA ? ENTRY_ROW_NONE : min(ai, ENTRY_ROW_EMPTY - 1u));
            ry = paint;
        } else if (sc.paints[paint].op != 0) {
            rx = ENTRY_GRADIENT | ENTRY_ROW_EMPTY;
            ry = paint;
        } else {
            rx = NO_ALPHA;
            ry = 0u;
        }
    }
    b.keys[e] = sc.fb_base + et.fbl;  // its value, e, is generated by the sort's first pass
    b.entry_rec[e] = make_uint2(rx, ry);
    if (b.dump) b.entry_meta[e] = make_uint2(p, et.fbl);
}

// One wave over the tiles [j0, j1) of a path's rect, row-major.  Count pass: returns the
// range's {entries, alpha entries}.  Write pass: numbers them from `first` (entry, alpha entry)
// by ballot rank.
template <bool WRITE, bool CLIPS>
__device__ __forceinline__ uint2 range_entries(const BatchDev &b, const SceneDev &sc, const BinItem &it, uint32_t j0,
                                               uint32_t j1, uint2 first) {
    const uint32_t lane = threadIdx.x & 63;
    const uint64_t lt_mask = lane == 0 ? 0ull : (~0ull >> (64 - lane));
    const uint32_t p = it.p;
    const int4 r = it.rect;
    const int w = max(r.z - r.x, 0);
    const uint32_t tbase = it.tbase;
    const int fb_row0 = (r.y - sc.fb_y0) * sc.fb_w + (r.x - sc.fb_x0);
    const uint32_t rgba = it.rgba;
    const float inv_w = w > 0 ? 1.0f / (float)w : 0.0f;
    uint32_t e = first.x, n_alpha = first.y;
    // tile j's status, its framebuffer tile and that tile's z, loaded one chunk ahead; row = j / w
    // from a float estimate (off by at most one below 2^24 tiles), then corrected
    auto fetch = [&](uint32_t j, uint32_t &st, uint32_t &z, uint32_t &fbl) {
        st = TILE_EMPTY;
        z = 0;
        fbl = 0;
        if (j < j1) {
            uint32_t row = (uint32_t)((float)j * inv_w);
            if (row * (uint32_t)w > j) row--;
            else if ((row + 1) * (uint32_t)w <= j) row++;
            const uint32_t x = j - row * (uint32_t)w;
            fbl = (uint32_t)(fb_row0 + (int)row * sc.fb_w + (int)x);
            st = b.tile_st[tbase + j];
            z = b.zbuf[sc.fb_base + fbl];
        }
    };
    uint32_t n_st, n_z, n_fbl;
    fetch(j0 + lane, n_st, n_z, n_fbl);
    // write pass: a chunk's alpha indices are loaded at the end of the chunk before it
    auto alpha_of = [&](uint32_t st, uint32_t j) {
        if (!CLIPS) return st == TILE_ALPHA ? b.tile_alpha_idx[tbase + j] : NO_ALPHA;
        return st == TILE_ALPHA ? b.tile_alpha_idx[tbase + j] : (st == TILE_CLIP_MASK ? b.tile_clip_ai[tbase + j] : NO_ALPHA);
    };
    uint32_t n_ai = WRITE ? alpha_of(n_st, j0 + lane) : NO_ALPHA;
    for (uint32_t jc = j0; jc < j1; jc += 64) {  // uniform over the wave
        const uint32_t j = jc + lane;
        EntryTile et;
        et.st = n_st;
        et.fbl = n_fbl;
        et.ai = n_ai;
        et.keep = n_st != TILE_EMPTY && (int32_t)p >= (int32_t)n_z;
        fetch(jc + 64 + lane, n_st, n_z, n_fbl);
        const uint64_t m = __ballot(et.keep);
        const uint64_t ma = __ballot(et.keep && (et.st == TILE_ALPHA || (CLIPS && et.st == TILE_CLIP_MASK)));
        if (WRITE && et.keep)
            write_entry<CLIPS>(b, sc, p, rgba, it.paint, tbase, j, et, e + (uint32_t)__popcll(m & lt_mask),
                        n_alpha + (uint32_t)__popcll(ma & lt_mask));
        e += (uint32_t)__popcll(m);
        n_alpha += (uint32_t)__popcll(ma);
        if (WRITE) n_ai = alpha_of(n_st, jc + 64 + lane);
    }
    return make_uint2(e, n_alpha);
}

template <bool WRITE, bool CLIPS>
__global__ __launch_bounds__(256) void k_entries_wave(BatchDev b) {
    const uint32_t item = __builtin_amdgcn_readfirstlane(blockIdx.x * 4 + (threadIdx.x >> 6));  // uniform
    if (item >= b.n_bin_wave) return;  // whole wave
    const BinItem it = b.bin_items[item];
    const int4 r = it.rect;
    // a clip path draws nothing: no entries
    const uint32_t area = CLIPS && it.level != 0 ? 0u : (uint32_t)max(r.z - r.x, 0) * (uint32_t)max(r.w - r.y, 0);
    const uint2 first = WRITE ? b.path_ent[it.gp] : make_uint2(0u, 0u);
    const uint2 n = range_entries<WRITE, CLIPS>(b, b.scenes[it.k], it, 0, area, first);
    if (!WRITE && (threadIdx.x & 63) == 0) b.path_ent[it.gp] = n;
}

// A block per large path: wave w takes the w-th quarter of the rect's tiles (row-major, in
// whole 64-tile chunks), so the waves need no barrier.  The count pass keeps each quarter's
// counts (ent_quarter) beside the path's total; the write pass starts each wave after the
// quarters before it.
template <bool WRITE, bool CLIPS>
__global__ __launch_bounds__(256) void k_entries_group(BatchDev b) {
    const uint32_t item = blockIdx.x;
    if (item >= b.n_bin_group) return;  // uniform over the block
    const BinItem it = b.bin_items[b.n_bin_wave + item];
    const int4 r = it.rect;
    const uint32_t area = CLIPS && it.level != 0 ? 0u : (uint32_t)max(r.z - r.x, 0) * (uint32_t)max(r.w - r.y, 0);
    const uint32_t wv = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const uint32_t q = (area + 255) / 256 * 64;  // tiles per quarter, whole chunks
    const uint32_t j0 = min(wv * q, area), j1 = min(j0 + q, area);
    uint2 first = make_uint2(0u, 0u);
    if (WRITE) {
        first = b.path_ent[it.gp];
        for (uint32_t v = 0; v < wv; v++) {
            const uint2 c = b.ent_quarter[item * 4 + v];
            first.x += c.x;
            first.y += c.y;
        }
    }
    const uint2 n = range_entries<WRITE, CLIPS>(b, b.scenes[it.k], it, j0, j1, first);
    if (!WRITE) {
        if (lane == 0) b.ent_quarter[item * 4 + wv] = n;
        __shared__ uint2 part[4];
        if (lane == 0) part[wv] = n;
        __syncthreads();
        if (threadIdx.x == 0)
            b.path_ent[it.gp] = make_uint2(part[0].x + part[1].x + part[2].x + part[3].x,
                                           part[0].y + part[1].y + part[2].y + part[3].y);
    }
}

// ---------------------------------------------------------------------------
// Stable LSD radix sort of (key, value) u32 pairs, 8 bits per pass
// ---------------------------------------------------------------------------

constexpr int RS_THREADS = 256;
constexpr int RS_ITEMS = 16;
constexpr int RS_TILE = RS_THREADS * RS_ITEMS;

// The sort runs per scene: a scene's entries are contiguous and its keys lie in its own range
// of framebuffer tiles, so sorting each scene's entries by their scene-local tile index
// (16 bits for canvases up to 4096^2) sorts the whole batch.  Sort blocks never straddle
// scenes: scene k's entries [seg_e[k], seg_e[k + 1]) (clamped to the capacity) take blocks
// [seg_b[k], seg_b[k + 1]) of RS_TILE items, and the digit histogram is laid out
// [scene][digit][block of the scene], so its exclusive scan places every item inside its
// scene's range, ordered by digit, then block, then rank.
__global__ __launch_bounds__(RS_THREADS) void k_rs_segments(BatchDev b, uint32_t *seg_e, uint32_t *seg_b) {
    uint32_t carry = 0;
    for (int k0 = 0; k0 <= b.n_scenes; k0 += RS_THREADS) {
        const int k = k0 + (int)threadIdx.x;
        uint32_t e = 0, e_next = 0;
        if (k <= b.n_scenes) {
            const uint32_t total = min(b.counters[C_ENTRIES], b.cap_entries);
            e = k < b.n_scenes ? min(b.path_ent[b.scenes[k].path_base].x, total) : total;
            e_next = k + 1 < b.n_scenes ? min(b.path_ent[b.scenes[k + 1].path_base].x, total) : total;
            seg_e[k] = e;
        }
        const uint32_t nbk = k < b.n_scenes ? (e_next - e + RS_TILE - 1) / RS_TILE : 0u;
        uint32_t tot;
        const uint32_t ex = block_exclusive_scan<uint32_t>(nbk, tot);
        if (k <= b.n_scenes) seg_b[k] = carry + ex;
        carry += tot;
    }
}

// This block's scene and item range; false past the last block.
struct RsBlock {
    int k;
    uint32_t lb, nbk, hbase, e0, e1, fb_base;
};
__device__ __forceinline__ bool rs_block(const BatchDev &b, const uint32_t *seg_e, const uint32_t *seg_b, RsBlock &rb) {
    if (blockIdx.x >= seg_b[b.n_scenes]) return false;
    rb.k = find_scene_wave(seg_b, b.n_scenes, blockIdx.x);
    rb.lb = blockIdx.x - seg_b[rb.k];
    rb.nbk = seg_b[rb.k + 1] - seg_b[rb.k];
    rb.hbase = 256u * seg_b[rb.k];
    rb.e0 = seg_e[rb.k] + rb.lb * RS_TILE;
    rb.e1 = min(rb.e0 + RS_TILE, seg_e[rb.k + 1]);
    rb.fb_base = b.scenes[rb.k].fb_base;
    return true;
}

__global__ __launch_bounds__(RS_THREADS) void k_rs_hist(BatchDev b, const uint32_t *keys, const uint32_t *seg_e,
                                                        const uint32_t *seg_b, int shift, uint32_t *hist) {
    __shared__ uint32_t h[256];
    RsBlock rb;
    if (!rs_block(b, seg_e, seg_b, rb)) return;  // uniform over the block
    h[threadIdx.x] = 0;
    __syncthreads();
#pragma unroll 4
    for (int i = 0; i < RS_ITEMS; i++) {
        const uint32_t j = rb.e0 + (uint32_t)i * RS_THREADS + threadIdx.x;
        if (j < rb.e1) atomicAdd(&h[((keys[j] - rb.fb_base) >> shift) & 255u], 1u);
    }
    __syncthreads();
    hist[rb.hbase + threadIdx.x * rb.nbk + rb.lb] = h[threadIdx.x];
}

// Scatter of one 8-bit digit pass.  The block's 4096 items are first ranked stably by digit
// into LDS, then written out in digit order: runs of equal digits land at consecutive global
// positions, so stores coalesce instead of scattering one item per digit per row.  Each wave
// ranks its own contiguous 1024 items (16 rows of 64) with wave-local digit counters: a ballot
// multisplit gives an item's rank among equal digits of its row, the counter the items of
// earlier rows; only then are the waves' counts combined per digit.  offs is the exclusive
// scan of the [digit][block] histogram: the block's count of digit d is the difference of
// its entry and the next one.  With no vals, an item's value is its input index.
constexpr int RS_WAVE_ITEMS = RS_ITEMS * 64;

__global__ __launch_bounds__(RS_THREADS) void k_rs_scatter(BatchDev b, const uint32_t *keys, const uint32_t *vals,
                                                           uint32_t *keys_out, uint32_t *vals_out,
                                                           const uint32_t *seg_e, const uint32_t *seg_b, int shift,
                                                           const uint32_t *offs) {
    __shared__ uint32_t skey[RS_TILE], sval[RS_TILE];
    __shared__ uint32_t lstart[256];  // local start of each digit
    __shared__ uint32_t goff[256];    // global start of each digit's run from this block
    __shared__ uint32_t wcnt[RS_THREADS / 64][256];  // per wave: digit counts, then the wave's base per digit
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    RsBlock rb;
    if (!rs_block(b, seg_e, seg_b, rb)) return;  // uniform over the block
    const uint32_t base = rb.e0;
    const uint32_t valid_n = rb.e1 - rb.e0;
    {
        const uint32_t h = rb.hbase + threadIdx.x * rb.nbk + rb.lb;
        const uint32_t o = offs[h];
        goff[threadIdx.x] = o;
        uint32_t total;
        lstart[threadIdx.x] = block_exclusive_scan<uint32_t>(offs[h + 1] - o, total);
    }
    for (int q = threadIdx.x; q < (RS_THREADS / 64) * 256; q += RS_THREADS) (&wcnt[0][0])[q] = 0;
    __syncthreads();
    const uint64_t lt_mask = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
    uint32_t *cnt = wcnt[wave];
    uint32_t key[RS_ITEMS], val[RS_ITEMS], rank[RS_ITEMS];
    const uint32_t w0 = (uint32_t)wave * RS_WAVE_ITEMS;
#pragma unroll
    for (int r = 0; r < RS_ITEMS; r++) {
        const uint32_t li = w0 + (uint32_t)r * 64 + (uint32_t)lane;
        const bool valid = li < valid_n;
        key[r] = valid ? keys[base + li] : 0u;
        val[r] = !valid ? 0u : (vals ? vals[base + li] : (uint32_t)(base + li));
    }
#pragma unroll
    for (int r = 0; r < RS_ITEMS; r++) {
        const uint32_t li = w0 + (uint32_t)r * 64 + (uint32_t)lane;
        const bool valid = li < valid_n;
        const uint32_t d = ((key[r] - rb.fb_base) >> shift) & 255u;
        // lanes of this row holding the same digit (multisplit peer mask)
        uint64_t peers = __ballot(valid);
#pragma unroll
        for (int bit = 0; bit < 8; bit++) {
            const uint64_t m = __ballot((d >> bit) & 1u);
            peers &= ((d >> bit) & 1u) ? m : ~m;
        }
        const uint32_t rk = __popcll(peers & lt_mask);
        const uint32_t before = valid ? cnt[d] : 0u;  // items of digit d in earlier rows of this wave
        rank[r] = before + rk;
        // the last lane of each digit group advances the counter (after every lane's read: a
        // wave's LDS operations complete in order)
        if (valid && rk == (uint32_t)__popcll(peers) - 1u) cnt[d] = before + rk + 1u;
    }
    __syncthreads();
    {  // per digit: the counts of earlier waves
        const uint32_t d = threadIdx.x;
        uint32_t acc = 0;
#pragma unroll
        for (int w = 0; w < RS_THREADS / 64; w++) {
            const uint32_t c = wcnt[w][d];
            wcnt[w][d] = acc;
            acc += c;
        }
    }
    __syncthreads();
#pragma unroll
    for (int r = 0; r < RS_ITEMS; r++) {
        const uint32_t li = w0 + (uint32_t)r * 64 + (uint32_t)lane;
        if (li < valid_n) {
            const uint32_t d = ((key[r] - rb.fb_base) >> shift) & 255u;
            const uint32_t pos = lstart[d] + cnt[d] + rank[r];
            skey[pos] = key[r];
            sval[pos] = val[r];
        }
    }
    __syncthreads();
    for (uint32_t i = threadIdx.x; i < valid_n; i += RS_THREADS) {
        const uint32_t key = skey[i];
        const uint32_t d = ((key - rb.fb_base) >> shift) & 255u;
        const uint32_t g = goff[d] + (i - lstart[d]);
        keys_out[g] = key;
        vals_out[g] = sval[i];
    }
}

// ---------------------------------------------------------------------------
// Launch helpers (host side)
// ---------------------------------------------------------------------------

static inline uint32_t blocks_for(uint64_t n, uint32_t per) { return (uint32_t)((n + per - 1) / per); }

hipError_t launch_zero_spans(const ZeroSpans &z, hipStream_t st) {
    if (z.n < 0 || z.n > ZERO_SPANS) return hipErrorInvalidValue;
    const uint64_t total = z.n ? z.end[z.n - 1] : 0;
    if (total == 0) return hipSuccess;
    uint32_t nb = (uint32_t)std::min<uint64_t>(blocks_for(total, 256), 8192);
    hipLaunchKernelGGL(k_zero_spans, dim3(nb), dim3(256), 0, st, z);
    return hipGetLastError();
}

size_t scan_temp_elems(uint64_t n) { return blocks_for(n, SCAN_TILE) + 1; }

template <typename T, typename Load, typename Store>
static hipError_t scan_impl(Load ld, Store st, uint32_t n, T *tmp, hipStream_t s) {
    uint32_t nb = blocks_for(n, SCAN_TILE);
    if (nb == 0) nb = 1;
    hipLaunchKernelGGL((k_scan_reduce<T, Load>), dim3(nb), dim3(SCAN_THREADS), 0, s, ld, n, tmp);
    hipLaunchKernelGGL((k_scan_blocks<T>), dim3(1), dim3(SCAN_THREADS), 0, s, tmp, nb);
    hipLaunchKernelGGL((k_scan_apply<T, Load, Store>), dim3(nb), dim3(SCAN_THREADS), 0, s, ld, st, n, tmp, nb);
    return hipGetLastError();
}

// Exclusive scan of in[0..n) into out[0..n] (out[n] = total).  in may equal out.
hipError_t scan_u32(const uint32_t *in, uint32_t *out, uint32_t n, uint32_t *tmp, hipStream_t s) {
    return scan_impl<uint32_t>(LoadU32{in}, StoreU32{out}, n, tmp, s);
}

hipError_t scan_counted(uint32_t *inout, const uint32_t *n_ptr, uint32_t cap, uint32_t *counters, int counter,
                        uint32_t *tmp, hipStream_t s) {
    return scan_impl<uint32_t>(LoadU32Counted{inout, n_ptr}, StoreCounted{inout, counters, counter, cap}, cap, tmp, s);
}

// Exclusive scan in place with the total written to counters[counter].
hipError_t scan_total(uint32_t *inout, uint32_t n, uint32_t *counters, int counter, uint32_t *tmp, hipStream_t s) {
    return scan_impl<uint32_t>(LoadU32{inout}, StoreCounted{inout, counters, counter, n}, n, tmp, s);
}

// Exclusive scan of the per-path {entries, alpha entries} in place, totals to the counters.
hipError_t scan_entries(uint2 *inout, uint32_t n, uint32_t *counters, void *tmp, hipStream_t s) {
    return scan_impl<U2>(LoadPathTot{inout}, StoreEntryOff{inout, counters, n}, n, reinterpret_cast<U2 *>(tmp), s);
}

// Exclusive scan of the per-path {fills, alpha tiles} in place (path_tot[P] = totals), with
// the totals also in the counters and at the end of the tile offset arrays.
hipError_t scan_paths(const BatchDev &b, void *tmp, hipStream_t s) {
    return scan_impl<U2>(LoadPathTot{b.path_tot},
                         StorePathOff{b.path_tot, b.tile_fill_off + b.n_tiles, b.tile_alpha_idx + b.n_tiles, b.counters,
                                      b.n_paths},
                         b.n_paths, reinterpret_cast<U2 *>(tmp), s);
}

// segs_per_wave: segments each wave works through (a multiple of 64, at least 64); take: idle
// lanes a wave gathers before handing out segments (1..64)
hipError_t launch_flatten(const BatchDev &b, bool emit, uint32_t segs_per_wave, uint32_t take, hipStream_t s) {
    if (b.n_segs == 0) return hipSuccess;
    if (segs_per_wave < 64 || segs_per_wave % 64 || take < 1 || take > 64) return hipErrorInvalidValue;
    const dim3 lines(blocks_for(b.n_segs, 256));
    const dim3 grid(blocks_for(blocks_for(b.n_segs, segs_per_wave), 4));
    // deep curves: 16 per workgroup per round, at most 8 workgroups per CU
    const dim3 deep(std::max(1u, std::min(blocks_for(b.n_segs, 16), 2048u)));
    if (emit) {
        hipLaunchKernelGGL(k_flatten_lines<true>, lines, dim3(256), 0, s, b);
        hipLaunchKernelGGL(k_flatten<true>, grid, dim3(256), 0, s, b, segs_per_wave, take);
        if (b.deep_pred) hipLaunchKernelGGL(k_flatten_deep<true>, deep, dim3(256), 0, s, b);
    } else {
        hipLaunchKernelGGL(k_flatten_lines<false>, lines, dim3(256), 0, s, b);
        hipLaunchKernelGGL(k_flatten<false>, grid, dim3(256), 0, s, b, segs_per_wave, take);
        if (b.deep_pred) hipLaunchKernelGGL(k_flatten_deep<false>, deep, dim3(256), 0, s, b);
    }
    return hipGetLastError();
}

// grid_lines: chords the launch is sized for (exact in plain mode, capacity in batched mode)
hipError_t launch_bin_items(const BatchDev &b, hipStream_t s) {
    if (b.n_bin_wave) hipLaunchKernelGGL(k_bin_items<0>, dim3(blocks_for(b.n_bin_wave, 256)), dim3(256), 0, s, b);
    if (b.n_bin_group) hipLaunchKernelGGL(k_bin_items<1>, dim3(blocks_for(b.n_bin_group, 256)), dim3(256), 0, s, b);
    return hipGetLastError();
}

hipError_t launch_bin(const BatchDev &b, bool emit, hipStream_t s) {
    if (b.n_bin_wave) {
        const dim3 g(blocks_for(b.n_bin_wave, 4));
        if (emit) hipLaunchKernelGGL(k_bin_wave<true>, g, dim3(256), 0, s, b);
        else hipLaunchKernelGGL(k_bin_wave<false>, g, dim3(256), 0, s, b);
    }
    if (b.n_bin_group) {
        const dim3 g(b.n_bin_group);
        if (emit) hipLaunchKernelGGL(k_bin_group<true>, g, dim3(256), 0, s, b);
        else hipLaunchKernelGGL(k_bin_group<false>, g, dim3(256), 0, s, b);
    }
    return hipGetLastError();
}

// rows: columns of more tiles are walked by several lanes each (1..64; 0: none)
template <bool CLIPS>
static void propagate_phase(const BatchDev &b, uint32_t block, uint32_t rows, uint32_t phase, hipStream_t s) {
    hipLaunchKernelGGL(k_propagate<CLIPS>, dim3(blocks_for(b.n_cols, block)), dim3(block), 0, s, b, phase, rows);
    if (rows == 0) return;
    // the listed columns' waves: enough to cover the device, striding over lists of unknown length
    if (rows < (uint32_t)TALL_ROWS)
        hipLaunchKernelGGL((k_propagate_rows<CLIPS, 16>), dim3(std::min<uint32_t>(blocks_for(b.n_cols, 16), 2048)),
                           dim3(256), 0, s, b, phase);
    hipLaunchKernelGGL((k_propagate_rows<CLIPS, 64>), dim3(std::min<uint32_t>(blocks_for(b.n_cols, 4), 2048)),
                       dim3(256), 0, s, b, phase);
}

hipError_t launch_propagate(const BatchDev &b, uint32_t block, uint32_t rows, hipStream_t s) {
    if (b.n_cols == 0) return hipSuccess;
    if (block != 64 && block != 128 && block != 256) return hipErrorInvalidValue;
    if (rows > (uint32_t)TALL_ROWS) return hipErrorInvalidValue;
    if (!b.has_clips) {
        propagate_phase<false>(b, block, rows, 0u, s);
    } else {
        // clip levels from the outermost in, then the draw paths
        for (uint32_t level = 1; level <= std::min(b.clip_levels, MAX_CLIP_LEVELS); level++)
            propagate_phase<true>(b, block, rows, level, s);
        propagate_phase<true>(b, block, rows, MAX_CLIP_LEVELS + 1, s);
    }
    return hipGetLastError();
}

template <bool CLIPS>
static void launch_entries_t(const BatchDev &b, bool write, hipStream_t s) {
    if (b.n_bin_wave) {
        const dim3 g(blocks_for(b.n_bin_wave, 4));
        if (write) hipLaunchKernelGGL((k_entries_wave<true, CLIPS>), g, dim3(256), 0, s, b);
        else hipLaunchKernelGGL((k_entries_wave<false, CLIPS>), g, dim3(256), 0, s, b);
    }
    if (b.n_bin_group) {
        const dim3 g(b.n_bin_group);
        if (write) hipLaunchKernelGGL((k_entries_group<true, CLIPS>), g, dim3(256), 0, s, b);
        else hipLaunchKernelGGL((k_entries_group<false, CLIPS>), g, dim3(256), 0, s, b);
    }
}

hipError_t launch_entries(const BatchDev &b, bool write, hipStream_t s) {
    if (b.has_clips) launch_entries_t<true>(b, write, s);
    else launch_entries_t<false>(b, write, s);
    return hipGetLastError();
}

// Range [start, end) of every framebuffer tile's run in the sorted keys, marked at the run
// boundaries: item i starts its key's run when the key before differs and ends it when the
// key after differs.  Tiles with no entries keep the cleared (0, 0).
constexpr int FB_BOUNDS_ITEMS = 4;  // items per thread per step, all loads issued first

__global__ __launch_bounds__(256) void k_fb_bounds(const uint32_t *keys, const uint32_t *n_ptr, uint32_t cap,
                                                   uint32_t n_fb, uint2 *range) {
    const uint32_t n = min(*n_ptr, cap);
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t i0 = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i0 < n; i0 += FB_BOUNDS_ITEMS * stride) {
        uint32_t prev[FB_BOUNDS_ITEMS], key[FB_BOUNDS_ITEMS], next[FB_BOUNDS_ITEMS];
#pragma unroll
        for (int j = 0; j < FB_BOUNDS_ITEMS; j++) {
            const uint64_t i = i0 + j * stride;
            key[j] = n_fb;  // marks nothing
            prev[j] = next[j] = n_fb + 1;
            if (i < n) {
                key[j] = keys[i];
                if (i > 0) prev[j] = keys[i - 1];
                if (i + 1 < n) next[j] = keys[i + 1];
            }
        }
#pragma unroll
        for (int j = 0; j < FB_BOUNDS_ITEMS; j++) {
            const uint64_t i = i0 + j * stride;
            if (key[j] >= n_fb) continue;  // past the count (keys of the entry pass are below n_fb)
            if (prev[j] != key[j]) range[key[j]].x = (uint32_t)i;
            if (next[j] != key[j]) range[key[j]].y = (uint32_t)(i + 1);
        }
    }
}

hipError_t launch_fb_bounds(const uint32_t *sorted_keys, const uint32_t *n_ptr, uint32_t cap, uint32_t n_fb,
                            uint2 *range, hipStream_t s) {
    // the count is known on the device only: a grid for the capacity, strided
    const uint32_t nb = std::min<uint32_t>(blocks_for(cap, 256 * FB_BOUNDS_ITEMS), 16384);
    if (nb == 0) return hipSuccess;
    hipLaunchKernelGGL(k_fb_bounds, dim3(nb), dim3(256), 0, s, sorted_keys, n_ptr, cap, n_fb, range);
    return hipGetLastError();
}

// Sort blocks for a capacity of `cap` entries over `n_scenes` scenes (one partial block each).
static uint32_t rs_max_blocks(uint64_t cap, uint32_t n_scenes) { return blocks_for(cap, RS_TILE) + n_scenes + 1; }

size_t radix_temp_elems(uint64_t cap, uint32_t n_scenes) {
    const uint64_t nh = 256ull * rs_max_blocks(cap, n_scenes);
    return 2ull * (n_scenes + 1) + nh + 1 + scan_temp_elems(nh);
}

// Sorts the batch's entry keys (b.keys[0 .. entries)), with values starting as the item
// indices (vals is not read), scene by scene by the low `bits` bits of the scene-local
// framebuffer tile, stably; returns the buffer pair holding the result via *result_alt
// (false: keys/vals, true: keys_alt/vals_alt).
hipError_t radix_sort(const BatchDev &b, uint32_t *keys, uint32_t *vals, uint32_t *keys_alt, uint32_t *vals_alt,
                      int bits, uint32_t *tmp, bool *result_alt, hipStream_t s) {
    const uint32_t n_sc = (uint32_t)b.n_scenes;
    const uint32_t nb = rs_max_blocks(b.cap_entries, n_sc);
    const uint32_t nh = 256 * nb;
    uint32_t *seg_e = tmp, *seg_b = tmp + (n_sc + 1);
    uint32_t *hist = tmp + 2 * (n_sc + 1);
    uint32_t *scan_tmp = hist + nh + 1;
    hipLaunchKernelGGL(k_rs_segments, dim3(1), dim3(RS_THREADS), 0, s, b, seg_e, seg_b);
    bool alt = false;
    for (int shift = 0; shift < bits; shift += 8) {
        const uint32_t *ki = alt ? keys_alt : keys;
        const uint32_t *vi = shift == 0 ? nullptr : (alt ? vals_alt : vals);  // values start as the indices
        uint32_t *ko = alt ? keys : keys_alt;
        uint32_t *vo = alt ? vals : vals_alt;
        hipLaunchKernelGGL(k_rs_hist, dim3(nb), dim3(RS_THREADS), 0, s, b, ki, seg_e, seg_b, shift, hist);
        // the scan covers the capacity's blocks; entries past this launch's blocks are never read
        hipError_t e = scan_u32(hist, hist, nh, scan_tmp, s);
        if (e != hipSuccess) return e;
        hipLaunchKernelGGL(k_rs_scatter, dim3(nb), dim3(RS_THREADS), 0, s, b, ki, vi, ko, vo, seg_e, seg_b, shift, hist);
        alt = !alt;
    }
    *result_alt = alt;
    return hipGetLastError();
}

}  // namespace pfh
