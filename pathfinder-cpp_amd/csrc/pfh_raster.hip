// Per-pixel stages: coverage masks (fill) and tile composite.  Built with FMA
// contraction allowed: these results are checked against FP64 models within
// stated tolerances, not bit-for-bit.
#include "pfh_internal.h"

namespace pfh {

// ---------------------------------------------------------------------------
// Coverage masks
// ---------------------------------------------------------------------------
//
// Per pixel (px, py) of an alpha tile the reference accumulates, over the tile's
// fills, sign(from.x - to.x) * ∫_{[px,px+1] ∩ [xmin,xmax]} clamp(py + 1 - y(x), 0, 1) dx
// on top of the tile's backdrop, then applies the fill rule
// (shaders/d3d9/fill.frag:25-45 with its area LUT, shaders/d3d11/fill.comp:51-140).
// Here the integral is analytic.  For a fill crossing pixel column c over
// [lo, hi], every row below the fill's lowest point in that column receives the
// full sign * (hi - lo); only the rows the fill passes through need the clamped
// integral.  So each fill adds partial areas to the rows it crosses and one
// "full" delta to the row below them, and a prefix sum down each column at the
// end turns deltas into coverage.  Partial rows are stored telescoped (row r gets
// v_r - v_{r-1}, the row below the fill gets full - v_last), so a single
// difference array and one prefix sum give the coverage.  Accumulation is fixed
// point (2^-20): the result does not depend on the order fills arrive in.
//
// Layout: 16 lanes per alpha tile (one per pixel column), 4 tiles per wave,
// 16 tiles per 256-thread block; each lane owns its column in LDS, so no atomics.

constexpr int FILL_TILES_PER_BLOCK = 16;
constexpr int MASK_PREFETCH = 8;  // composite: coverage masks loaded ahead of their blend

__device__ __forceinline__ float avg_clamp01(float ua, float ub) {
    float a = fminf(ua, ub), bb = fmaxf(ua, ub);
    float du = bb - a;
    if (du <= 1e-6f) return fminf(fmaxf(0.5f * (a + bb), 0.0f), 1.0f);
    float full = bb > 1.0f ? bb - fmaxf(a, 1.0f) : 0.0f;
    float l2 = fmaxf(a, 0.0f), h2 = fminf(bb, 1.0f);
    float mid = h2 > l2 ? (h2 - l2) * (h2 + l2) * 0.5f : 0.0f;
    return (full + mid) * __builtin_amdgcn_rcpf(du);
}

__device__ __forceinline__ float apply_rule(float c, uint32_t even_odd) {
    if (!even_odd) return fminf(fabsf(c), 1.0f);
    float m = c - 2.0f * floorf(c * 0.5f);  // GLSL mod(c, 2)
    return fminf(fmaxf(1.0f - fabsf(1.0f - m), 0.0f), 1.0f);
}

// Adds one fill's contribution to pixel column `col` of a tile's difference array
// (column stride 1, row stride 16 words); LDS adds, since several lanes may share a column.
__device__ __forceinline__ void accumulate_fill(uint2 fl, int col, float scale, int32_t *diff) {
    const float x0 = (float)(fl.x & 0xffffu) * (1.0f / 256.0f), y0 = (float)(fl.x >> 16) * (1.0f / 256.0f);
    const float x1 = (float)(fl.y & 0xffffu) * (1.0f / 256.0f), y1 = (float)(fl.y >> 16) * (1.0f / 256.0f);
    const float xl = fminf(x0, x1), xr = fmaxf(x0, x1);
    const float lo = fmaxf((float)col, xl), hi = fminf((float)(col + 1), xr);
    if (!(hi > lo)) return;
    const float sign = x0 > x1 ? 1.0f : -1.0f;
    const float yl = x0 < x1 ? y0 : y1, yr = x0 < x1 ? y1 : y0;
    const float slope = (yr - yl) * __builtin_amdgcn_rcpf(xr - xl);
    const float ya = yl + (lo - xl) * slope, yb = yl + (hi - xl) * slope;
    const float w = hi - lo;
    const int r0 = max(0, min(15, (int)floorf(fminf(ya, yb))));
    const int r1 = max(0, min(15, (int)floorf(fmaxf(ya, yb))));
    int32_t prev = 0;
    for (int r = r0; r <= r1; r++) {
        const float rb = (float)(r + 1);
        const int32_t v = (int32_t)rintf(sign * w * avg_clamp01(rb - ya, rb - yb) * scale);
        atomicAdd(diff + r * 16 + col, v - prev);
        prev = v;
    }
    atomicAdd(diff + (r1 + 1) * 16 + col, (int32_t)rintf(sign * w * scale) - prev);
}

__device__ __forceinline__ uint32_t wave_max_of_groups(uint32_t v) {
    // v is uniform within each 16-lane group: the wave maximum is the max of lanes 0, 16, 32, 48
    uint32_t a = __builtin_amdgcn_readlane(v, 0), b = __builtin_amdgcn_readlane(v, 16);
    uint32_t c = __builtin_amdgcn_readlane(v, 32), d = __builtin_amdgcn_readlane(v, 48);
    return max(max(a, b), max(c, d));
}

// A 16-lane group per alpha tile.  Fills arrive 16 at a time (one per lane); each
// fill covers a run of pixel columns, and the group's (fill, column) pairs are
// enumerated with a prefix sum so every lane works on a pair that contributes.
__global__ __launch_bounds__(256) void k_fill_masks(BatchDev b) {
    // 17 rows per tile slot (272 words): tile slots g and g+1 land on opposite bank halves
    __shared__ int32_t diff[FILL_TILES_PER_BLOCK][17][16];
    __shared__ uint2 sfill[FILL_TILES_PER_BLOCK][16];
    __shared__ uint32_t scum[FILL_TILES_PER_BLOCK][16];
    __shared__ uint32_t out8[FILL_TILES_PER_BLOCK][64];
    const int g = threadIdx.x >> 4;   // tile slot in the block
    const int li = threadIdx.x & 15;  // lane in the tile's group (= pixel column for the output)
    const uint32_t na = min(b.counters[C_ALPHA], b.cap_alpha);
    // grid-stride over groups of 16 tiles: the grid is sized from the capacity, the count
    // is only known on the device; the loop bound is uniform over the block
    for (uint32_t first = blockIdx.x * FILL_TILES_PER_BLOCK; first < na; first += gridDim.x * FILL_TILES_PER_BLOCK) {
    const uint32_t a = first + g;
    const bool valid = a < na;
    // all lanes stay to the end: loops below have wave-uniform bounds and use shuffles
#pragma unroll
    for (int r = 0; r < 17; r++) diff[g][r][li] = 0;
    uint4 rec = make_uint4(0, 0, 0, 0);
    if (valid) rec = b.alpha_rec[a];
    // After a fill-capacity overflow a tile's range can start past the buffer (rec.x >=
    // cap_fills): clamp both ends so the count never wraps (the batch is then retried).
    const uint32_t f0 = valid ? min(rec.x, b.cap_fills) : 0u;
    const uint32_t f1 = valid ? min(rec.y, b.cap_fills) : 0u;
    const uint32_t nf = f1 > f0 ? f1 - f0 : 0u;
    const float scale = (float)(1 << FIX_SHIFT);
    const uint32_t chunks = wave_max_of_groups((nf + 15) >> 4);
    for (uint32_t ch = 0; ch < chunks; ch++) {
        const uint32_t i = ch * 16 + li;
        uint2 fl = make_uint2(0, 0);
        uint32_t cnt = 0;
        if (i < nf) {
            fl = b.fills[f0 + i];
            const float x0 = (float)(fl.x & 0xffffu) * (1.0f / 256.0f), x1 = (float)(fl.y & 0xffffu) * (1.0f / 256.0f);
            const int c0 = (int)floorf(fminf(x0, x1)), c1 = (int)ceilf(fmaxf(x0, x1)) - 1;
            cnt = (uint32_t)max(0, c1 - c0 + 1);
        }
        uint32_t incl = cnt;
#pragma unroll
        for (int d = 1; d < 16; d <<= 1) {
            const uint32_t o = (uint32_t)__shfl_up((int)incl, d, 16);
            if (li >= d) incl += o;
        }
        const uint32_t total = (uint32_t)__shfl((int)incl, 15, 16);
        sfill[g][li] = fl;
        scum[g][li] = incl;
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
        const uint32_t rounds = wave_max_of_groups((total + 15) >> 4);
        for (uint32_t rd = 0; rd < rounds; rd++) {
            const uint32_t q = rd * 16 + li;
            if (q < total) {
                // the fill owning pair q: first j with scum[j] > q
                int j = 0;
#pragma unroll
                for (int st = 8; st > 0; st >>= 1)
                    if (scum[g][j + st - 1] <= q) j += st;
                const uint2 fj = sfill[g][j];
                const uint32_t before = j ? scum[g][j - 1] : 0u;
                const float x0 = (float)(fj.x & 0xffffu) * (1.0f / 256.0f), x1 = (float)(fj.y & 0xffffu) * (1.0f / 256.0f);
                const int col = (int)floorf(fminf(x0, x1)) + (int)(q - before);
                accumulate_fill(fj, col, scale, &diff[g][0][0]);
            }
        }
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    }
    const float bd = (float)(int8_t)(rec.z & 0xffu);
    const uint32_t rule = (rec.z >> 8) & 1u;
    const float inv = 1.0f / scale;
    int32_t run = 0;
    uint8_t *bytes = reinterpret_cast<uint8_t *>(&out8[g][0]);
#pragma unroll
    for (int r = 0; r < 16; r++) {
        run += diff[g][r][li];
        const float c = bd + (float)run * inv;
        bytes[r * 16 + li] = (uint8_t)rintf(apply_rule(c, rule) * 255.0f);
    }
    // each lane stores one 16-byte row of its tile: 256 B per tile, coalesced
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    if (valid) {
        const uint4 row = reinterpret_cast<const uint4 *>(&out8[g][0])[li];
        reinterpret_cast<uint4 *>(b.masks + (uint64_t)a * 256)[li] = row;
    }
    // the next group reuses this slot's LDS
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    }
}

// ---------------------------------------------------------------------------
// Composite
// ---------------------------------------------------------------------------
//
// One wave per framebuffer tile, 4 pixels of one row per lane.  Entries arrive in
// draw order (ascending path index).  Premultiplied src-over is associative, so a
// run of consecutive solid entries (whose coverage is the same 0 or 1 for every
// pixel of the tile) is folded into one premultiplied colour by a segmented scan
// across the wave; the per-pixel loop then blends only alpha entries and folded
// runs.  The colour model follows shaders/d3d9/tile.frag:590-653 (mask alpha,
// premultiply) and BlendState::from_over (gpu/base.h:83-93); layers accumulate in
// fp32 and the framebuffer is stored once as RGBA8 (premultiplied).

__global__ __launch_bounds__(256) void k_composite(BatchDev b, const uint32_t *sorted_vals) {
    const int lane = threadIdx.x & 63;
    const uint32_t f = blockIdx.x * 4 + (threadIdx.x >> 6);
    if (f >= b.n_fb) return;
    int k = 0;
    {
        int lo = 0, hi = b.n_scenes - 1;
        while (lo < hi) {
            int mid = (lo + hi + 1) >> 1;
            if (b.fb_bases[mid] <= f) lo = mid;
            else hi = mid - 1;
        }
        k = lo;
    }
    const SceneDev &sc = b.scenes[k];
    const uint32_t fl = f - sc.fb_base;
    const int tx = sc.fb_x0 + (int)(fl % (uint32_t)sc.fb_w), ty = sc.fb_y0 + (int)(fl / (uint32_t)sc.fb_w);
    const int row = lane >> 2, col = (lane & 3) * 4;
    const int y = ty * TILE + row, x0 = tx * TILE + col;
    float r[4], g[4], bl[4], al[4];
    if (sc.load_dst && y < sc.height) {
#pragma unroll
        for (int j = 0; j < 4; j++) {
            uint32_t px = 0;
            if (x0 + j < sc.width)
                px = *reinterpret_cast<const uint32_t *>(sc.out + (int64_t)y * sc.out_pitch + (int64_t)(x0 + j) * 4);
            r[j] = (float)(px & 255u) * (1.0f / 255.0f);
            g[j] = (float)((px >> 8) & 255u) * (1.0f / 255.0f);
            bl[j] = (float)((px >> 16) & 255u) * (1.0f / 255.0f);
            al[j] = (float)(px >> 24) * (1.0f / 255.0f);
        }
    } else {
#pragma unroll
        for (int j = 0; j < 4; j++) r[j] = g[j] = bl[j] = al[j] = 0.0f;
    }
    uint32_t e0 = b.fb_cnt[f], e1 = min(b.fb_cnt[f + 1], b.cap_entries);
    for (uint32_t base = e0; base < e1; base += 64) {
        const uint32_t i = base + lane;
        const uint32_t cnt = min(64u, e1 - base);
        uint32_t ai = NO_ALPHA, c = 0;
        if (i < e1) {
            const uint4 rec = b.entry_rec[sorted_vals[i]];
            ai = rec.x;
            c = rec.y;
            if (ai != NO_ALPHA && ai >= b.cap_alpha) c = 0;  // capacity overflow: the run is retried
        }
        const float ca = (float)(c >> 24) * (1.0f / 255.0f);
        const float cr = (float)(c & 255u) * (1.0f / 255.0f);
        const float cg = (float)((c >> 8) & 255u) * (1.0f / 255.0f);
        const float cb = (float)((c >> 16) & 255u) * (1.0f / 255.0f);
        const bool valid = (uint32_t)lane < cnt;
        const bool solid = valid && ai == NO_ALPHA;
        // premultiplied colour of a solid entry; fold runs with an inclusive segmented scan
        float pr = cr * ca, pg = cg * ca, pb = cb * ca, pa = ca;
        const uint64_t solid_mask = __ballot(solid);
        // a run starts at a solid lane whose predecessor is not solid
        bool head = solid && (lane == 0 || !((solid_mask >> (lane - 1)) & 1ull));
        bool flag = head || !solid;
#pragma unroll
        for (int d = 1; d < 64; d <<= 1) {
            const float qr = __shfl_up(pr, d, 64), qg = __shfl_up(pg, d, 64), qb = __shfl_up(pb, d, 64),
                        qa = __shfl_up(pa, d, 64);
            const bool qf = __shfl_up((int)flag, d, 64);
            if (lane >= d && !flag) {
                // the entry on top is this lane's accumulated colour; the lower one is q
                const float k1 = 1.0f - pa;
                pr = pr + qr * k1;
                pg = pg + qg * k1;
                pb = pb + qb * k1;
                pa = pa + qa * k1;
            }
            if (lane >= d) flag = flag || qf;
        }
        // events: alpha entries, and the last entry of each solid run
        const uint64_t valid_mask = __ballot(valid);
        const uint64_t alpha_mask = valid_mask & ~solid_mask;
        const uint64_t run_end = solid_mask & ~(solid_mask >> 1);
        uint64_t events = alpha_mask | run_end;
        const uint32_t *mask_words = reinterpret_cast<const uint32_t *>(b.masks);
        while (events) {
            // Issue the masks of the next MASK_PREFETCH alpha events together (their loads are
            // in flight at once), then blend every event up to the last of them in order.
            uint32_t mq[MASK_PREFETCH];
            uint64_t am = events & alpha_mask;
#pragma unroll
            for (int u = 0; u < MASK_PREFETCH; u++) {
                mq[u] = 0;
                if (am) {
                    const int j = __builtin_ctzll(am);
                    am &= am - 1;
                    const uint32_t aj = __builtin_amdgcn_readlane(ai, j);
                    const uint32_t cj = __builtin_amdgcn_readlane(c, j);
                    if ((cj >> 24) != 0u) mq[u] = mask_words[(uint64_t)aj * 64 + lane];
                }
            }
            // events before the first alpha event not prefetched (all of them if none is left)
            const uint64_t upto = am ? ((1ull << __builtin_ctzll(am)) - 1ull) : ~0ull;
            uint64_t chunk = events & upto;
            events &= ~upto;
            while (chunk) {
                const int j = __builtin_ctzll(chunk);
                chunk &= chunk - 1;
                if ((alpha_mask >> j) & 1ull) {
                    const uint32_t m = mq[0];
#pragma unroll
                    for (int u = 0; u + 1 < MASK_PREFETCH; u++) mq[u] = mq[u + 1];
                    const uint32_t cj = __builtin_amdgcn_readlane(c, j);
                    const float ja = (float)(cj >> 24) * (1.0f / 255.0f);
                    if (ja == 0.0f) continue;
                    const float jr = (float)(cj & 255u) * (1.0f / 255.0f);
                    const float jg = (float)((cj >> 8) & 255u) * (1.0f / 255.0f);
                    const float jb = (float)((cj >> 16) & 255u) * (1.0f / 255.0f);
#pragma unroll
                    for (int q = 0; q < 4; q++) {
                        const float sa = ja * (float)((m >> (8 * q)) & 255u) * (1.0f / 255.0f);
                        const float k1 = 1.0f - sa;
                        r[q] = jr * sa + r[q] * k1;
                        g[q] = jg * sa + g[q] * k1;
                        bl[q] = jb * sa + bl[q] * k1;
                        al[q] = sa + al[q] * k1;
                    }
                } else {
                    const float fr = __shfl(pr, j, 64), fg = __shfl(pg, j, 64), fb = __shfl(pb, j, 64),
                                fa = __shfl(pa, j, 64);
                    const float k1 = 1.0f - fa;
#pragma unroll
                    for (int q = 0; q < 4; q++) {
                        r[q] = fr + r[q] * k1;
                        g[q] = fg + g[q] * k1;
                        bl[q] = fb + bl[q] * k1;
                        al[q] = fa + al[q] * k1;
                    }
                }
            }
        }
    }
    if (y >= sc.height) return;
    uint8_t *orow = sc.out + (int64_t)y * sc.out_pitch;
    uint32_t px[4];
#pragma unroll
    for (int q = 0; q < 4; q++) {
        const uint32_t pr = (uint32_t)rintf(fminf(fmaxf(r[q], 0.0f), 1.0f) * 255.0f);
        const uint32_t pg = (uint32_t)rintf(fminf(fmaxf(g[q], 0.0f), 1.0f) * 255.0f);
        const uint32_t pb = (uint32_t)rintf(fminf(fmaxf(bl[q], 0.0f), 1.0f) * 255.0f);
        const uint32_t pa = (uint32_t)rintf(fminf(fmaxf(al[q], 0.0f), 1.0f) * 255.0f);
        px[q] = pr | (pg << 8) | (pb << 16) | (pa << 24);
    }
    if (x0 + 3 < sc.width && ((reinterpret_cast<uintptr_t>(orow) & 15) == 0)) {
        reinterpret_cast<uint4 *>(orow)[x0 >> 2] = make_uint4(px[0], px[1], px[2], px[3]);
    } else {
#pragma unroll
        for (int q = 0; q < 4; q++)
            if (x0 + q < sc.width) reinterpret_cast<uint32_t *>(orow)[x0 + q] = px[q];
    }
}

static inline uint32_t blocks_for(uint64_t n, uint32_t per) { return (uint32_t)((n + per - 1) / per); }

hipError_t launch_fill_masks(const BatchDev &b, uint32_t grid_alpha, hipStream_t s) {
    if (grid_alpha == 0) return hipSuccess;
    // at most 8192 blocks (32 per CU): each loops over its share of the tiles
    const uint32_t nb = std::min<uint32_t>(blocks_for(grid_alpha, FILL_TILES_PER_BLOCK), 8192u);
    hipLaunchKernelGGL(k_fill_masks, dim3(nb), dim3(256), 0, s, b);
    return hipGetLastError();
}

hipError_t launch_composite(const BatchDev &b, const uint32_t *sorted_vals, hipStream_t s) {
    if (b.n_fb == 0) return hipSuccess;
    hipLaunchKernelGGL(k_composite, dim3(blocks_for(b.n_fb, 4)), dim3(256), 0, s, b, sorted_vals);
    return hipGetLastError();
}

}  // namespace pfh
