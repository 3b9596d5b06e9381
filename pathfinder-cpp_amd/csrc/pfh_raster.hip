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
// Layout: 16 lanes per alpha tile (one per pixel column), 4 tiles per wave, 16 tiles per
// 256-thread block.  A tile's difference array is column-major in LDS (each column's 17
// rows padded to 20 words, so a lane zeroes and reads its column with 16-byte accesses),
// and the coverage mask is stored column-major too: lane c writes column c's 16 bytes with
// one 16-byte store (byte 16 c + r is pixel (c, r)).

constexpr int FILL_TILES_PER_BLOCK = 16;
constexpr int FILL_COL_WORDS = 20;                          // rows 0..16 used
constexpr int FILL_GROUP_WORDS = 16 * FILL_COL_WORDS + 8;   // pad between tiles of a wave





// Adds one fill's contribution to pixel column `col` of a tile's difference array
// (FILL_COL_WORDS words per column); LDS adds, since several lanes may share a column.
// Over the column's x-window [lo, hi] (width w) the fill runs linearly between heights ymin
// and ymax.  With G(t) = ∫ max(t - y(x), 0) dx over the window (0 below ymin, w (t - ymean)
// above ymax, w (t - ymin)^2 / (2 (ymax - ymin)) between), the covered area of pixel row r is
// F(r) = ∫ clamp(r + 1 - y(x), 0, 1) dx = G(r + 1) - G(r).  The array holds F's differences
// down the column in fixed point: rows r0 = floor(ymin) .. r1 = floor(ymax) are crossed, rows
// below get the full w.  Which form of G applies at each row boundary follows from n = r1 - r0
// alone (r0 + 1 is past ymax only when n = 0, r0 + 2 only when n <= 1, r1 + 1 always), so no
// lane branches on it.  Inside the crossed span F grows by the same 2 k2 from row to row (G is
// quadratic there), so F(r0), F(r0 + 1), F(r1) and the full width are evaluated and the rows
// between add one constant: a pair costs the same whatever its height.  Values are rounded to
// the fixed-point grid once each and the differences are taken in integers, so the running sum
// is exact at r0, r0 + 1, r1 and below.
// v_cvt_rpi_i32_f32: floor(v + 0.5) as an integer, in one instruction.
__device__ __forceinline__ int32_t round_fix(float v) {
    int32_t r;
    asm("v_cvt_rpi_i32_f32 %0, %1" : "=v"(r) : "v"(v));
    return r;
}

// A fill's geometry, decoded once per fill: x range [xl, xr], the height yl at xl, the slope,
// and the sign (fill direction) carried in xl's sign bit (xl >= 0, so |x| restores it).
struct __align__(16) FillGeom {
    float xl_signed, xr, yl, slope;
};

__device__ __forceinline__ FillGeom fill_geom(uint2 fl) {
    const float x0 = (float)(fl.x & 0xffffu) * (1.0f / 256.0f), y0 = (float)(fl.x >> 16) * (1.0f / 256.0f);
    const float x1 = (float)(fl.y & 0xffffu) * (1.0f / 256.0f), y1 = (float)(fl.y >> 16) * (1.0f / 256.0f);
    const float xl = fminf(x0, x1), xr = fmaxf(x0, x1);
    const float yl = x0 < x1 ? y0 : y1, yr = x0 < x1 ? y1 : y0;
    FillGeom g;
    g.xl_signed = x0 > x1 ? xl : -xl;  // positive: the fill adds coverage
    g.xr = xr;
    g.yl = yl;
    g.slope = (yr - yl) * __builtin_amdgcn_rcpf(xr - xl);
    return g;
}

// v_max_f32 / v_min_f32 on operands that are never NaN here, without the quieting fminf and
// fmaxf apply to values loaded from memory.
__device__ __forceinline__ float max_f(float a, float b) {
    float r;
    asm("v_max_f32 %0, %1, %2" : "=v"(r) : "v"(a), "v"(b));
    return r;
}
__device__ __forceinline__ float min_f(float a, float b) {
    float r;
    asm("v_min_f32 %0, %1, %2" : "=v"(r) : "v"(a), "v"(b));
    return r;
}

__device__ __forceinline__ void accumulate_fill(const FillGeom &g, int col, float scale, int32_t *tile_diff) {
    const float xl = fabsf(g.xl_signed), xr = g.xr;
    const float lo = max_f((float)col, xl), hi = min_f((float)(col + 1), xr);
    if (!(hi > lo) || (uint32_t)col > 15u) return;
    const float ks = __builtin_copysignf(scale, g.xl_signed);  // sign of the fill, in fixed-point units
    const float ya = g.yl + (lo - xl) * g.slope, yb = g.yl + (hi - xl) * g.slope;
    const float ymin = fminf(ya, yb), ymax = fmaxf(ya, yb);
    const float ws = ks * (hi - lo);  // the full width, signed, in fixed-point units
    const float fr0 = __builtin_amdgcn_fmed3f(floorf(ymin), 0.0f, 15.0f);
    const float fr1 = __builtin_amdgcn_fmed3f(floorf(ymax), 0.0f, 15.0f);
    const int r0 = (int)fr0, r1 = (int)fr1, n = r1 - r0;
    // k2 is infinite when ymin == ymax; then n == 0 and only the linear form is selected
    const float k2 = 0.5f * ws * __builtin_amdgcn_rcpf(ymax - ymin);
    const float u0 = (fr0 - ymin) + 1.0f, u1 = fr1 - ymin;
    const float gb = ws * ((fr1 + 1.0f) - 0.5f * (ymin + ymax));  // G(r1 + 1): linear form
    const float f0 = n == 0 ? gb : k2 * u0 * u0;                    // F(r0) = G(r0 + 1)
    const float fd = gb - k2 * u1 * u1;                              // F(r1), when n >= 1
    const float f1 = n == 1 ? fd : k2 * (2.0f * u0 + 1.0f);          // F(r0 + 1), when n >= 1
    const int32_t full = round_fix(ws), v0 = round_fix(f0), v1 = round_fix(f1), vd = round_fix(fd);
    const int32_t dm = round_fix(2.0f * k2);  // F(r) - F(r - 1) for r0 + 2 <= r <= r1 - 1
    // 0 <= col <= 15 here: the column's word offset by a full-rate 24-bit multiply (left to
    // itself the compiler forms it with a quarter-rate 32-bit one)
    uint32_t col_words;
    asm("v_mul_u32_u24 %0, %1, %2" : "=v"(col_words) : "v"((uint32_t)col), "v"((uint32_t)FILL_COL_WORDS));
    int32_t *cd = tile_diff + col_words;
    atomicAdd(cd + r0, v0);
    atomicAdd(cd + r0 + 1, (n == 0 ? full : v1) - v0);
    for (int r = r0 + 2; r < r1; r++) atomicAdd(cd + r, dm);
    // n >= 2 means ymax - ymin > 1, so |dm| < 2^20: a 24-bit multiply is exact
    if (n >= 2) atomicAdd(cd + r1, vd - (v1 + __mul24(n - 2, dm)));
    if (n >= 1) atomicAdd(cd + r1 + 1, full - vd);  // n == 1: v1 == vd
}

// Inclusive prefix sum across the 64-lane wave: DPP within rows, then the row totals by row
// broadcasts (lane 15 into rows 1 and 3, then lane 31 into rows 2 and 3).
__device__ __forceinline__ uint32_t wave_inclusive_sum(uint32_t v) {
    v = row16_inclusive_sum(v);
    v += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x142, 0xa, 0xf, false);  // row_bcast:15
    v += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x143, 0xc, 0xf, false);  // row_bcast:31
    return v;
}

// v_med3_i32: the median of three.
__device__ __forceinline__ int32_t med3_i32(int32_t a, int32_t b, int32_t c) {
    int32_t r;
    asm("v_med3_i32 %0, %1, %2, %3" : "=v"(r) : "v"(a), "v"(b), "v"(c));
    return r;
}

// Coverage of one pixel from its fixed-point winding (2^-FIX_SHIFT units) under the fill rule
// (shaders/d3d11/fill.comp:134-140), rounded to nearest 8 bits and left in bits 24..31:
// (c * 255 + 2^19) >> 20 == (c * 4080 + 2^23) >> 24 for 0 <= c <= 2^20 (fits 32 bits).
// Nonzero: min(|w|, 1) = med3(w, -w, 1); even-odd: 1 - |1 - (w mod 2)| = min(m, 2 - m).
__device__ __forceinline__ uint32_t coverage_hi(int32_t w, uint32_t even_odd) {
    const int32_t one = 1 << FIX_SHIFT;
    const int32_t nz = med3_i32(w, -w, one);
    const int32_t m = w & (2 * one - 1);  // w mod 2 (two's complement: also for negative w)
    const int32_t eo = min(m, 2 * one - m);
    const uint32_t c = (uint32_t)(even_odd ? eo : nz);  // <= 2^20: a 24-bit multiply-add
    return __umul24(c, 4080u) + (1u << 23);
}

// Prefix of a column's 16 difference values on top of the backdrop, then coverage bytes packed
// 4 per word (RULE 0 nonzero, 1 even-odd, 2 per lane from `rule`).  v_perm_b32 gathers the
// top bytes: selector bytes 0-3 pick the second operand's bytes, 4-7 the first's, 0x0c gives 0.
template <int RULE>
__device__ __forceinline__ void resolve_column(int32_t run, const int32_t (&dv)[16], uint32_t rule,
                                               uint32_t (&packed)[4]) {
    const uint32_t eo = RULE == 2 ? rule : (uint32_t)RULE;
#pragma unroll
    for (int k = 0; k < 4; k++) {
        uint32_t hi[4];
#pragma unroll
        for (int r = 0; r < 4; r++) {
            run += dv[4 * k + r];
            hi[r] = coverage_hi(run, eo);
        }
        const uint32_t lo2 = __builtin_amdgcn_perm(hi[1], hi[0], 0x0c0c0703u);  // bytes 0, 1
        const uint32_t hi2 = __builtin_amdgcn_perm(hi[3], hi[2], 0x07030c0cu);  // bytes 2, 3
        packed[k] = lo2 | hi2;
    }
}

// A wave takes 4 consecutive alpha tiles (16 lanes, one pixel column each, per tile).  Their
// fills form one pooled list of up to 64 per chunk; the (fill, pixel column) pairs of a
// chunk are enumerated with a wave-wide prefix sum, so every lane works on a pair that
// contributes whatever the mix of light and heavy tiles in the wave.  The tiles are those
// some draw entry samples, in the list the entry pass wrote: a mask hidden under an opaque
// solid tile of a later path is never read, so it is not computed.  With dumps every alpha
// tile is rasterized, in index order.
__global__ __launch_bounds__(256) void k_fill_masks(BatchDev b) {
    __shared__ __align__(16) int32_t diff[FILL_TILES_PER_BLOCK * FILL_GROUP_WORDS];
    // a chunk's fills: geometry, and (first pair | tile offset in the wave's LDS << 11 | first
    // pixel column << 27); a pair window's marks: 1 where a fill's pairs start
    __shared__ FillGeom sgeom[4][64];
    __shared__ uint32_t smeta[4][64];
    __shared__ uint8_t smark[4][64];
    const uint32_t lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    const uint32_t slot = lane >> 4, li = lane & 15;  // tile slot in the wave, pixel column
    int32_t *dwave = diff + wv * 4 * FILL_GROUP_WORDS;
    int32_t *dcol = dwave + slot * FILL_GROUP_WORDS + li * FILL_COL_WORDS;
    const bool listed = !b.dump;
    const uint32_t na = min(b.counters[listed ? C_ALPHA_ENTRIES : C_ALPHA], listed ? b.cap_vis : b.cap_alpha);
    const float scale = (float)(1 << FIX_SHIFT);
    uint32_t drawn = 0;  // fills this wave read (wave-uniform)
    // row cap_alpha of the masks is all ones: the composite's full-coverage row
    if (blockIdx.x == 0 && wv == 0)
        reinterpret_cast<uint32_t *>(b.masks)[(uint64_t)b.cap_alpha * 64 + lane] = 0xffffffffu;
    // grid-stride over groups of 16 tiles: the grid is sized from the capacity, the count
    // is only known on the device; the loop bound is uniform over the block
    const uint32_t lb = (b.xcd_order & 1u) ? xcd_block(blockIdx.x, gridDim.x) : blockIdx.x;
    const uint32_t stride = gridDim.x * FILL_TILES_PER_BLOCK, first0 = lb * FILL_TILES_PER_BLOCK;
    const uint32_t mine = wv * 4 + slot;  // this lane's tile within a group
    // The tile list and the tile records are read ahead (the record one group ahead, the list
    // entry two), so a group starts with its record in registers and waits only for its fills.
    // ~0u: no tile.
    auto tile_at = [&](uint32_t idx) { return idx >= na ? ~0u : (listed ? b.alpha_vis[idx] : idx); };
    auto rec_of = [&](uint32_t a) { return a < b.cap_alpha ? b.alpha_rec[a] : make_uint4(0, 0, 0, 0); };
    uint32_t a_next = tile_at(first0 + mine), a_after = tile_at(first0 + stride + mine);
    uint4 rec_next = rec_of(a_next);
    for (uint32_t first = first0; first < na; first += stride) {
        const uint32_t a = a_next;
        const uint4 rec = rec_next;
        const bool valid = a < b.cap_alpha;
        rec_next = rec_of(a_after);
        a_next = a_after;
        a_after = tile_at(first + 2 * stride + mine);
#pragma unroll
        for (int k = 0; k < FILL_COL_WORDS / 4; k++) reinterpret_cast<uint4 *>(dcol)[k] = make_uint4(0, 0, 0, 0);
        // After a fill-capacity overflow a tile's range can start past the buffer: clamp both
        // ends so no count wraps (the batch is then retried).
        const uint32_t f0 = min(rec.x, b.cap_fills);
        const uint32_t f1 = max(f0, min(rec.y, b.cap_fills));
        // the 4 tiles' ranges (uniform over the wave), concatenated
        const uint32_t s0 = __builtin_amdgcn_readlane(f0, 0), n0 = __builtin_amdgcn_readlane(f1, 0) - s0;
        const uint32_t s1 = __builtin_amdgcn_readlane(f0, 16), n1 = __builtin_amdgcn_readlane(f1, 16) - s1;
        const uint32_t s2 = __builtin_amdgcn_readlane(f0, 32), n2 = __builtin_amdgcn_readlane(f1, 32) - s2;
        const uint32_t s3 = __builtin_amdgcn_readlane(f0, 48), n3 = __builtin_amdgcn_readlane(f1, 48) - s3;
        const uint32_t c1 = n0, c2 = c1 + n1, c3 = c2 + n2, ntot = c3 + n3;
        const uint32_t o1 = s1 - c1, o2 = s2 - c2, o3 = s3 - c3;  // wrapping differences
        drawn += ntot;
        for (uint32_t base = 0; base < ntot; base += 64) {
            const uint32_t k = base + lane;
            uint32_t cnt = 0, fs = 0, c0 = 0;
            FillGeom g{0.0f, 0.0f, 0.0f, 0.0f};
            if (k < ntot) {
                // fill k of the concatenated ranges: k minus its tile's first, plus that tile's start
                const bool ge1 = k >= c1, ge2 = k >= c2, ge3 = k >= c3;
                fs = (ge1 ? 1u : 0u) + (ge2 ? 1u : 0u) + (ge3 ? 1u : 0u);
                uint32_t off = s0;
                off = ge1 ? o1 : off;
                off = ge2 ? o2 : off;
                off = ge3 ? o3 : off;
                const uint32_t addr = k + off;
                g = fill_geom(b.fills[addr]);
                const float xl = fabsf(g.xl_signed);
                const int cl = (int)floorf(xl);
                cnt = (uint32_t)max(0, (int)ceilf(g.xr) - cl);
                c0 = (uint32_t)cl & 15u;  // 8.8 fills lie in [0, 16): cl <= 15
            }
            const uint32_t incl = wave_inclusive_sum(cnt);
            const uint32_t total = __builtin_amdgcn_readlane(incl, 63);
            const uint32_t start = incl - cnt;  // this fill's first pair (< 64 * 17)
            sgeom[wv][lane] = g;
            smeta[wv][lane] = start | ((fs * FILL_GROUP_WORDS) << 11) | (c0 << 27);
            // Pairs go in windows of 64, one per lane.  The fill owning pair q is the last one
            // starting at or before q: marks at the window's fill starts, their ballot, and the
            // count of marks up to this lane (plus the fills that started in earlier windows)
            // give it.  Every fill covers at least one column (the bin stage drops fills with
            // equal x ends), so no two fills start at the same pair.
            uint32_t carry = 0;
            for (uint32_t q0 = 0; q0 < total; q0 += 64) {
                smark[wv][lane] = 0;
                __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
                __builtin_amdgcn_wave_barrier();
                __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
                if (cnt != 0 && start >= q0 && start - q0 < 64u) smark[wv][start - q0] = 1;
                // the marks, the chunk's fills and (first chunk) the zeroed columns are visible
                __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
                __builtin_amdgcn_wave_barrier();
                __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
                const uint64_t marks = __ballot(smark[wv][lane] != 0);
                const uint32_t q = q0 + lane;
                if (q < total) {
                    const uint32_t below = __builtin_amdgcn_mbcnt_hi((uint32_t)(marks >> 32),
                                                                     __builtin_amdgcn_mbcnt_lo((uint32_t)marks, 0u));
                    const uint32_t j = min(carry + below + (uint32_t)((marks >> lane) & 1ull) - 1u, 63u);
                    const FillGeom g = sgeom[wv][j];
                    const uint32_t m = smeta[wv][j];
                    const int col = (int)(m >> 27) + (int)(q - (m & 2047u));
                    accumulate_fill(g, col, scale, dwave + ((m >> 11) & 2047u));
                }
                carry += (uint32_t)__popcll(marks);
                // the next window rewrites the marks, the next chunk the fills
                __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
                __builtin_amdgcn_wave_barrier();
                __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
            }
        }
        if (ntot == 0) {  // nothing accumulated: the zeroing still has to land first
            __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
            __builtin_amdgcn_wave_barrier();
            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
        }
        const uint32_t rule = (rec.z >> 8) & 1u;
        int32_t dv[16];
#pragma unroll
        for (int k = 0; k < 4; k++) {
            const uint4 v = reinterpret_cast<const uint4 *>(dcol)[k];
            dv[4 * k] = (int32_t)v.x;
            dv[4 * k + 1] = (int32_t)v.y;
            dv[4 * k + 2] = (int32_t)v.z;
            dv[4 * k + 3] = (int32_t)v.w;
        }
        const int32_t bd_fix = (int32_t)(int8_t)(rec.z & 0xffu) * (1 << FIX_SHIFT);  // backdrop, exact
        uint32_t packed[4];
        // the 4 tiles of a wave nearly always belong to one path: branch on a wave-uniform rule
        const uint64_t eo_mask = __ballot(rule != 0);
        if (eo_mask == 0) resolve_column<0>(bd_fix, dv, rule, packed);
        else if (eo_mask == ~0ull) resolve_column<1>(bd_fix, dv, rule, packed);
        else resolve_column<2>(bd_fix, dv, rule, packed);
        if (valid)
            reinterpret_cast<uint4 *>(b.masks + (uint64_t)a * 256)[li] = make_uint4(packed[0], packed[1], packed[2], packed[3]);
        // the next group of tiles reuses this wave's LDS
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    }
    // one add per block, spread over slots by block index (the loop above is uniform over the block)
    __shared__ uint32_t sdrawn[4];
    if (lane == 0) sdrawn[wv] = drawn;
    __syncthreads();
    if (threadIdx.x == 0)
        atomicAdd(&b.counters[C_DRAWN + (blockIdx.x & (DRAWN_SLOTS - 1))], sdrawn[0] + sdrawn[1] + sdrawn[2] + sdrawn[3]);
}

// ---------------------------------------------------------------------------
// Composite
// ---------------------------------------------------------------------------
//
// One wave per framebuffer tile, 4 pixels of one column per lane.  Entries arrive in
// draw order (ascending path index) and are composited front to back: from the top
// entry down, each adds its premultiplied colour weighted by the transmittance T left
// by the entries above it, and T shrinks by its coverage alpha.  Per pixel and entry
// that is 6 operations for a masked entry (coverage, weight, 3 colours, T) where
// back-to-front "over" needs 10.  The framebuffer's alpha is 1 - T at the end (over a
// loaded destination: the destination weighted by T is added).  Premultiplied src-over is
// associative, so a run of consecutive solid entries (whose coverage is the same 0 or 1
// for every pixel of the tile) is folded into one premultiplied colour by a segmented scan
// across the wave, and the per-pixel loop blends only masked entries and folded runs.
// The colour model follows shaders/d3d9/tile.frag:590-653 (mask alpha, premultiply) and
// BlendState::from_over (gpu/base.h:83-93); layers accumulate in fp32 and the framebuffer
// is stored once as RGBA8 (premultiplied).

// One step of the segmented inclusive scan of premultiplied colours under "over": the value
// from the DPP source lane lies below this lane's accumulated run unless a run starts (or a
// non-solid entry sits) inside this lane's span; lanes without a source are left alone.
template <int CTRL, int ROW_MASK>
__device__ __forceinline__ void over_step(bool has_src, float &pr, float &pg, float &pb, float &pa, bool &flag) {
    auto dpp = [](float v) {
        return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), CTRL, ROW_MASK, 0xf, false));
    };
    const float qr = dpp(pr), qg = dpp(pg), qb = dpp(pb), qa = dpp(pa);
    const bool qf = __builtin_amdgcn_update_dpp(0, flag ? 1 : 0, CTRL, ROW_MASK, 0xf, false) != 0;
    if (has_src && !flag) {
        const float k1 = 1.0f - pa;
        pr = pr + qr * k1;
        pg = pg + qg * k1;
        pb = pb + qb * k1;
        pa = pa + qa * k1;
    }
    if (has_src) flag = flag || qf;
}


constexpr int EV_GROUP = 4;  // blend events whose mask words are loaded together

// A gradient paint's colour at pixel centre (px, py): the tile shader's colour texture coordinate
// (shaders/d3d9/tile.vert:77, mat2(m11, m21, m12, m22) * position + (m13, m23)), for a radial
// gradient its filter's t (shaders/d3d9/tile.frag:296-323: the larger root of the two circles'
// quadratic that is not negative, nothing where the discriminant is 0), then a linear sample of the
// ramp row (RGBA8, clamp to edge or repeat along u).  Channels in byte units (0..255); false where
// the paint leaves the pixel transparent (a zero or negative discriminant, a non-finite coordinate).
__device__ __forceinline__ bool gradient_colour(const PaintRec &pr, const uint32_t *ramp, float px, float py,
                                                float &cr, float &cg, float &cb, float &ca) {
    if (pr.kind == 0) {  // a solid colour's record (a blended solid path)
        cr = (float)(pr.colour & 255u);
        cg = (float)((pr.colour >> 8) & 255u);
        cb = (float)((pr.colour >> 16) & 255u);
        ca = (float)(pr.colour >> 24);
        return true;
    }
    const float u = pr.m11 * px + pr.m12 * py + pr.m13;
    float s = u;
    if (pr.kind == 2) {
        const float v = pr.m21 * px + pr.m22 * py + pr.m23;
        const float dpx = u - pr.from_x, dpy = v - pr.from_y, dr = pr.r1 - pr.r0;
        const float a = pr.vec_x * pr.vec_x + pr.vec_y * pr.vec_y - dr * dr;
        const float b = dpx * pr.vec_x + dpy * pr.vec_y + pr.r0 * dr;
        const float c = dpx * dpx + dpy * dpy - pr.r0 * pr.r0;
        const float disc = b * b - a * c;
        if (!(disc > 0.0f)) return false;
        const float sq = sqrtf(disc);
        float t0 = (sq + b) / a, t1 = (-sq + b) / a;
        if (t0 > t1) {
            const float t = t0;
            t0 = t1;
            t1 = t;
        }
        s = pr.uv_x + (t0 >= 0.0f ? t0 : t1);
    }
    if (!isfinite(s)) return false;
    // texel space: u * 256 - 0.5 (repeat: the fraction of u first, so large u keep their precision)
    const float uu = pr.wrap ? s - floorf(s) : fminf(fmaxf(s, -1.0f), 2.0f);
    const float cx = uu * (float)RAMP_TEXELS - 0.5f;
    const float fl = floorf(cx);
    const float f = cx - fl;
    const int i0 = (int)fl;
    int k0, k1;
    if (pr.wrap) {
        k0 = i0 & (RAMP_TEXELS - 1);
        k1 = (i0 + 1) & (RAMP_TEXELS - 1);
    } else {
        k0 = min(max(i0, 0), RAMP_TEXELS - 1);
        k1 = min(max(i0 + 1, 0), RAMP_TEXELS - 1);
    }
    const uint32_t t0 = ramp[k0], t1 = ramp[k1];
    auto ch = [&](int sh) {
        const float a0 = (float)((t0 >> sh) & 255u), a1 = (float)((t1 >> sh) & 255u);
        return a0 + (a1 - a0) * f;
    };
    cr = ch(0);
    cg = ch(8);
    cb = ch(16);
    ca = ch(24);
    return true;
}

// The tile shader's composite ops (shaders/d3d9/tile.frag:444-583): B(dest, src) per channel
// for the separable ones, through HSL for hue / saturation / color / luminosity.  Colours in [0, 1].
__device__ __forceinline__ float cdiv(float num, float den) { return den != 0.0f ? num / den : 0.0f; }
__device__ __forceinline__ float3 f3(float x, float y, float z) { return make_float3(x, y, z); }
__device__ __forceinline__ float dodge1(float d, float s) { return d == 0.0f ? 0.0f : (s == 1.0f ? 1.0f : d / (1.0f - s)); }
__device__ __forceinline__ float screen1(float d, float s) { return d + s - d * s; }
__device__ __forceinline__ float hard1(float d, float s) { return s <= 0.5f ? d * 2.0f * s : screen1(d, 2.0f * s - 1.0f); }
__device__ __forceinline__ float soft1(float d, float s) {
    const float dark = d <= 0.25f ? ((16.0f * d - 12.0f) * d + 4.0f) * d : sqrtf(d);
    return d + (s * 2.0f - 1.0f) * (s <= 0.5f ? d * (1.0f - d) : dark - d);
}
// RGB -> (h, s, l) and back (the shader's HSL, h in radians)
__device__ __forceinline__ float3 to_hsl(float3 c) {
    const float v = fmaxf(fmaxf(c.x, c.y), c.z), mn = fminf(fminf(c.x, c.y), c.z);
    const float ch = v - mn, l = mn + (v - mn) * 0.5f;
    float t0, t1, t2;
    if (c.x == v) { t0 = 0.0f; t1 = c.y; t2 = c.z; }
    else if (c.y == v) { t0 = 2.0f; t1 = c.z; t2 = c.x; }
    else { t0 = 4.0f; t1 = c.x; t2 = c.y; }
    const float h = 1.0471975511965976f * cdiv(t0 * ch + t1 - t2, ch);
    return f3(h, cdiv(ch, v), l);
}
__device__ __forceinline__ float3 from_hsl(float3 hsl) {
    const float a = hsl.y * fminf(hsl.z, 1.0f - hsl.z);
    float out[3];
    const float base[3] = {0.0f, 8.0f, 4.0f};
#pragma unroll
    for (int i = 0; i < 3; i++) {
        const float t = base[i] + hsl.x * 1.9098593171027443f;
        const float k = t - 12.0f * floorf(t / 12.0f);  // GLSL mod
        out[i] = hsl.z - fminf(fmaxf(fminf(k - 3.0f, 9.0f - k), -1.0f), 1.0f) * a;
    }
    return f3(out[0], out[1], out[2]);
}
__device__ float3 composite_rgb(float3 d, float3 s, uint32_t op) {
    switch (op) {
        case 0x1: return f3(d.x * s.x, d.y * s.y, d.z * s.z);
        case 0x2: return f3(screen1(d.x, s.x), screen1(d.y, s.y), screen1(d.z, s.z));
        case 0x3: return f3(hard1(s.x, d.x), hard1(s.y, d.y), hard1(s.z, d.z));  // overlay: hard light, swapped
        case 0x4: return f3(fminf(d.x, s.x), fminf(d.y, s.y), fminf(d.z, s.z));
        case 0x5: return f3(fmaxf(d.x, s.x), fmaxf(d.y, s.y), fmaxf(d.z, s.z));
        case 0x6: return f3(dodge1(d.x, s.x), dodge1(d.y, s.y), dodge1(d.z, s.z));
        case 0x7: return f3(1.0f - dodge1(1.0f - d.x, 1.0f - s.x), 1.0f - dodge1(1.0f - d.y, 1.0f - s.y),
                            1.0f - dodge1(1.0f - d.z, 1.0f - s.z));
        case 0x8: return f3(hard1(d.x, s.x), hard1(d.y, s.y), hard1(d.z, s.z));
        case 0x9: return f3(soft1(d.x, s.x), soft1(d.y, s.y), soft1(d.z, s.z));
        case 0xa: return f3(fabsf(d.x - s.x), fabsf(d.y - s.y), fabsf(d.z - s.z));
        case 0xb: return f3(d.x + s.x - 2.0f * d.x * s.x, d.y + s.y - 2.0f * d.y * s.y, d.z + s.z - 2.0f * d.z * s.z);
        default: {  // 0xc hue, 0xd saturation, 0xe color, 0xf luminosity
            const float3 hd = to_hsl(d), hs = to_hsl(s);
            const float3 h = op == 0xc ? f3(hs.x, hd.y, hd.z)
                             : op == 0xd ? f3(hd.x, hs.y, hd.z)
                             : op == 0xe ? f3(hs.x, hs.y, hd.z)
                                         : f3(hd.x, hd.y, hs.z);
            return from_hsl(h);
        }
    }
}

// A tile with an entry that composites through an op, blended back to front (the op reads what lies
// below): per entry, from the bottom, src = its colour (a solid, or its paint at the pixel centre;
// a pixel the paint leaves out is transparent), a = src alpha x coverage, then over (op 0), or
// tile.frag's composite (:585-598): rgb = a (1 - da) src + a da B(dest, src) + (1 - a) dest,
// alpha 1.  dest starts from the target (PFH_LOAD_DST) or transparent; premultiplied, in [0, 1].
// Stores the lane's 4 pixels.
__device__ void composite_back_to_front(const BatchDev &b, const SceneDev &sc, const uint32_t *sorted_vals, uint32_t e0,
                                        uint32_t e1, int x, int y0, uint32_t lane_off, uint8_t *cptr, int64_t pitch,
                                        int height, bool load_dst) {
    float dr[4], dg[4], db[4], da[4];
#pragma unroll
    for (int q = 0; q < 4; q++) {
        uint32_t px = 0;
        if (load_dst && y0 + q >= 0 && y0 + q < height) px = *reinterpret_cast<const uint32_t *>(cptr + q * pitch);
        dr[q] = (float)(px & 255u) * (1.0f / 255.0f);
        dg[q] = (float)((px >> 8) & 255u) * (1.0f / 255.0f);
        db[q] = (float)((px >> 16) & 255u) * (1.0f / 255.0f);
        da[q] = (float)(px >> 24) * (1.0f / 255.0f);
    }
    for (uint32_t i = e0; i < e1; i++) {  // uniform over the wave
        const uint2 rec = b.entry_rec[sorted_vals[i]];
        uint32_t colour = rec.y, op = 0, row = b.cap_alpha;  // row cap_alpha: full coverage
        bool empty = false;
        const PaintRec *pr = nullptr;
        if (rec.x != NO_ALPHA && (rec.x & ENTRY_GRADIENT)) {
            const uint32_t r = rec.x & ENTRY_ROW_MASK;
            if (r == ENTRY_ROW_EMPTY) empty = true;
            else if (r != ENTRY_ROW_NONE) row = r;
            if (row > b.cap_alpha) continue;  // capacity overflow (the run is retried): blends nothing
            pr = sc.paints + rec.y;
            op = pr->op;
        } else if (rec.x != NO_ALPHA) {
            if (rec.x >= b.cap_alpha) continue;  // capacity overflow
            row = rec.x;
        }
        const uint32_t m = empty ? 0u : *reinterpret_cast<const uint32_t *>(b.masks + (uint64_t)row * 256 + lane_off);
#pragma unroll
        for (int q = 0; q < 4; q++) {
            float cr = (float)(colour & 255u), cg = (float)((colour >> 8) & 255u), cb = (float)((colour >> 16) & 255u),
                  ca = (float)(colour >> 24);
            if (pr && !gradient_colour(*pr, sc.ramps + (uint64_t)rec.y * RAMP_TEXELS, (float)x + 0.5f,
                                       (float)(y0 + q) + 0.5f, cr, cg, cb, ca))
                cr = cg = cb = ca = 0.0f;
            const float k = 1.0f / 255.0f;
            const float a = ca * k * ((float)((m >> (8 * q)) & 255u) * k);
            const float sr = cr * k, sg = cg * k, sb = cb * k;
            if (op == 0) {
                dr[q] = sr * a + dr[q] * (1.0f - a);
                dg[q] = sg * a + dg[q] * (1.0f - a);
                db[q] = sb * a + db[q] * (1.0f - a);
                da[q] = a + da[q] * (1.0f - a);
            } else {
                const float3 bl = composite_rgb(f3(dr[q], dg[q], db[q]), f3(sr, sg, sb), op);
                const float w0 = a * (1.0f - da[q]), w1 = a * da[q], w2 = 1.0f - a;
                dr[q] = w0 * sr + w1 * bl.x + w2 * dr[q];
                dg[q] = w0 * sg + w1 * bl.y + w2 * dg[q];
                db[q] = w0 * sb + w1 * bl.z + w2 * db[q];
                da[q] = 1.0f;
            }
        }
    }
#pragma unroll
    for (int q = 0; q < 4; q++) {
        uint32_t w = __builtin_amdgcn_cvt_pk_u8_f32(dr[q] * 255.0f, 0, 0u);
        w = __builtin_amdgcn_cvt_pk_u8_f32(dg[q] * 255.0f, 1, w);
        w = __builtin_amdgcn_cvt_pk_u8_f32(db[q] * 255.0f, 2, w);
        w = __builtin_amdgcn_cvt_pk_u8_f32(da[q] * 255.0f, 3, w);
        if (y0 + q >= 0 && y0 + q < height) *reinterpret_cast<uint32_t *>(cptr + q * pitch) = w;
    }
}

// PAINTS: some scene of the batch has paint records (the solid-only variant carries no code for them).
template <bool PAINTS>
__global__ __launch_bounds__(256) void k_composite(BatchDev b, const uint32_t *sorted_vals) {
    __shared__ float4 s_colour[4][64];  // per wave of the block: the chunk's premultiplied colours
    const int lane = threadIdx.x & 63;
    const uint32_t lb = (b.xcd_order & 2u) ? xcd_block(blockIdx.x, gridDim.x) : blockIdx.x;
    // the tile and its scene are uniform over the wave: held in scalar registers, so the scene
    // record and the tile's range come through scalar loads
    const uint32_t f = __builtin_amdgcn_readfirstlane(lb * (blockDim.x >> 6) + (threadIdx.x >> 6));
    if (f >= b.n_fb) return;
    const int k = __builtin_amdgcn_readfirstlane(find_scene_wave(b.fb_bases, b.n_scenes, f));
    const SceneDev &sc = b.scenes[k];
    // what the store at the end needs, read here (scalar loads) rather than after the blending
    const int width = sc.width, height = sc.height;
    uint8_t *const out = sc.out;
    const int64_t pitch = sc.out_pitch;
    const bool load_dst = sc.load_dst != 0;
    const uint32_t fl = f - sc.fb_base, fb_w = (uint32_t)sc.fb_w;
    // fl / fb_w from a float estimate (off by at most one below 2^24 tiles), then corrected
    uint32_t row = (uint32_t)((float)fl * sc.fb_inv_w);
    if (row * fb_w > fl) row--;
    else if ((row + 1) * fb_w <= fl) row++;
    const int tx = sc.fb_x0 + (int)(fl - row * fb_w), ty = sc.fb_y0 + (int)row;
    // a lane owns 4 vertically adjacent pixels: column lane / 4, rows 4 (lane % 4) .. + 3,
    // matching the column-major coverage masks (one 32-bit mask word per lane)
    const int x = tx * TILE + (lane >> 2), y0 = ty * TILE + (lane & 3) * 4;
    uint32_t lane_off = (uint32_t)lane * 4u;  // this lane's word in a mask row
    float r[4], g[4], bl[4], tr[4];  // accumulated premultiplied colour (byte units), transmittance
#pragma unroll
    for (int q = 0; q < 4; q++) {
        r[q] = g[q] = bl[q] = 0.0f;
        tr[q] = 1.0f;
    }
    const uint2 range = b.fb_range[f];
    const uint32_t e0 = range.x, e1 = min(range.y, b.cap_entries);
    if (PAINTS && sc.has_ops) {
        // a tile with an entry that composites through an op is blended back to front
        bool ops = false;
        for (uint32_t c0 = e0; c0 < e1 && !ops; c0 += 64) {  // uniform over the wave
            const uint32_t i = c0 + lane;
            bool op = false;
            if (i < e1) {
                const uint2 rec = b.entry_rec[sorted_vals[i]];
                op = rec.x != NO_ALPHA && (rec.x & ENTRY_GRADIENT) && sc.paints[rec.y].op != 0;
            }
            ops = __ballot(op) != 0;
        }
        if (ops) {
            if (x < 0 || x >= width) return;
            composite_back_to_front(b, sc, sorted_vals, e0, e1, x, y0, lane_off, out + (int64_t)y0 * pitch + (int64_t)x * 4,
                                    pitch, height, load_dst);
            return;
        }
    }
    // chunks of 64 entries from the top one down
    for (uint32_t base = e0 + (e1 > e0 ? (e1 - e0 - 1) / 64 * 64 : 0); e1 > e0; base -= 64) {
        const uint32_t i = base + lane;
        const uint32_t cnt = min(64u, e1 - base);
        // ai: the mask row this lane's entry samples; row cap_alpha is all ones (full
        // coverage), sampled by solid entries and lanes past the chunk
        uint32_t ai = b.cap_alpha, c = 0;
        bool masked = false, grad = false;
        if (i < e1) {
            const uint2 rec = b.entry_rec[sorted_vals[i]];
            c = rec.y;
            if (rec.x != NO_ALPHA) {
                masked = true;  // a gradient entry never joins a solid run, with or without a mask row
                if (PAINTS && (rec.x & ENTRY_GRADIENT)) {
                    const uint32_t row = rec.x & ENTRY_ROW_MASK;
                    grad = row == ENTRY_ROW_NONE || row < b.cap_alpha;
                    if (row != ENTRY_ROW_NONE && row < b.cap_alpha) ai = row;
                    if (!grad) c = 0;  // capacity overflow: blends nothing
                } else if (rec.x < b.cap_alpha) {
                    ai = rec.x;
                } else {
                    c = 0;  // capacity overflow (the run is retried): blends nothing
                }
            }
        }
        const uint64_t grad_mask = PAINTS ? __ballot(grad) : 0ull;
        const uint32_t paint = c;  // a gradient entry's paint index
        if (PAINTS && grad) c = 0;
        const float ca = (float)(c >> 24) * (1.0f / 255.0f);
        const bool valid = (uint32_t)lane < cnt;
        const bool solid = valid && !masked;
        // premultiplied colour, the colour channels in byte units (0..255) so that the store
        // converts them as they are; a masked entry's is also scaled by 1/255 so that it multiplies
        // the raw mask byte.  Runs of solid entries are folded by an inclusive segmented scan.
        const float sc_a = solid ? ca : ca * (1.0f / 255.0f);
        float pr = (float)(c & 255u) * sc_a, pg = (float)((c >> 8) & 255u) * sc_a, pb = (float)((c >> 16) & 255u) * sc_a;
        float pa = sc_a;
        const uint64_t solid_mask = __ballot(solid);
        // a run starts at a solid lane whose predecessor is not solid
        bool head = solid && (lane == 0 || !((solid_mask >> (lane - 1)) & 1ull));
        bool flag = head || !solid;
        // DPP steps: shifts by 1, 2, 4, 8 within each 16-lane row, then the row totals
        // (lane 15 into row 1 and lane 47 into row 3, then lane 31 into rows 2 and 3).  A step
        // changes nothing unless a run is longer than its shift (within a row) or crosses the
        // row boundary it bridges, so only those steps run (runs are mostly 1-3 entries).
        const int row = lane >> 4, rl = lane & 15;
        const uint64_t run2 = solid_mask & (solid_mask >> 1);  // bit i: lanes i, i + 1 both solid
        const uint64_t run3 = run2 & (solid_mask >> 2);
        const uint64_t run5 = run3 & (run2 >> 2) & (solid_mask >> 4);
        const uint64_t run9 = run5 & (run5 >> 4);
        if (run2) over_step<0x111, 0xf>(rl >= 1, pr, pg, pb, pa, flag);
        if (run3) over_step<0x112, 0xf>(rl >= 2, pr, pg, pb, pa, flag);
        if (run5) over_step<0x114, 0xf>(rl >= 4, pr, pg, pb, pa, flag);
        if (run9) over_step<0x118, 0xf>(rl >= 8, pr, pg, pb, pa, flag);
        if (run2 & ((1ull << 15) | (1ull << 47))) over_step<0x142, 0xa>(row == 1 || row == 3, pr, pg, pb, pa, flag);
        if (run2 & (1ull << 31)) over_step<0x143, 0xc>(row >= 2, pr, pg, pb, pa, flag);
        // events, taken from the top: masked entries, and the last (top) entry of each solid
        // run, which holds the run's folded colour.  Every event blends the same way: a solid
        // run samples an all-ones mask row, and its colour is scaled by 1/255 like a masked
        // entry's, so the loop has a single path.
        if (solid) {
            pr *= 1.0f / 255.0f;
            pg *= 1.0f / 255.0f;
            pb *= 1.0f / 255.0f;
            pa *= 1.0f / 255.0f;
        }
        const uint64_t valid_mask = __ballot(valid);
        const uint64_t alpha_mask = valid_mask & ~solid_mask;
        const uint64_t run_end = solid_mask & ~(solid_mask >> 1);
        uint64_t events = alpha_mask | run_end;
        // this lane's mask word for event j; an empty slot (j = -1) loads lane 63's row, which
        // exists, so every slot loads and no branch splits the loads
        auto mask_word = [&](int j) -> uint32_t {
            const uint32_t idx = __builtin_amdgcn_readlane(ai, j & 63);
            const uint8_t *row = b.masks + (uint64_t)idx * 256;  // uniform: a scalar base
            // the lane's byte offset, made opaque in place, so that the load keeps the
            // scalar-base form instead of a hoisted per-lane pointer plus a vector add per event
            asm volatile("" : "+v"(lane_off));
            return *reinterpret_cast<const uint32_t *>(row + lane_off);
        };
        // The chunk's event colours go through LDS: a blend reads its event's colour with one
        // broadcast LDS read instead of four lane reads, which would occupy the VALU.
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
        s_colour[threadIdx.x >> 6][lane] = make_float4(pr, pg, pb, pa);
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
        const float4 *wcol = s_colour[threadIdx.x >> 6];
        auto blend = [&](float4 cf, uint32_t m) {
            const float fr = cf.x, fg = cf.y, fb = cf.z, fa = cf.w;
#pragma unroll
            for (int q = 0; q < 4; q++) {
                const float wq = tr[q] * (float)((m >> (8 * q)) & 255u);
                r[q] = fmaf(wq, fr, r[q]);
                g[q] = fmaf(wq, fg, g[q]);
                bl[q] = fmaf(wq, fb, bl[q]);
                tr[q] = fmaf(-wq, fa, tr[q]);
            }
        };
        // Events go in groups of EV_GROUP from the top: the next group's mask words are loaded
        // (EV_GROUP loads in flight) while the current group blends, so a wave waits on memory
        // once per group.  Two register sets take turns (the loop is unrolled by two), so no
        // copy waits for a load, and scheduling barriers keep the loads ahead of the blends.
        // An event index < 0 is an empty slot: no blend.
        auto take = [&](int (&js)[EV_GROUP], uint32_t (&ms)[EV_GROUP]) {
#pragma unroll
            for (int d = 0; d < EV_GROUP; d++) {
                js[d] = 63 - __clzll(events);  // -1 once none are left
                events &= ~(1ull << (js[d] & 63));
                ms[d] = mask_word(js[d]);
            }
        };
        auto blend_group = [&](const int (&js)[EV_GROUP], const uint32_t (&ms)[EV_GROUP]) {
            float4 cf[EV_GROUP];
#pragma unroll
            for (int d = 0; d < EV_GROUP; d++) cf[d] = wcol[max(js[d], 0)];
#pragma unroll
            for (int d = 0; d < EV_GROUP; d++)
                if (js[d] >= 0) blend(cf[d], ms[d]);
        };
        if (PAINTS && grad_mask) {
            // a chunk with gradient entries: events one at a time from the top, a gradient event's
            // colour evaluated per pixel
            while (events) {
                const int j = 63 - __clzll(events);
                events &= ~(1ull << j);
                const uint32_t m = mask_word(j);
                if ((grad_mask >> j) & 1ull) {
                    const uint32_t pi = __builtin_amdgcn_readlane(paint, j);
                    const PaintRec pr = sc.paints[pi];
                    const uint32_t *ramp = sc.ramps + (uint64_t)pi * RAMP_TEXELS;
#pragma unroll
                    for (int q = 0; q < 4; q++) {
                        float cr, cg, cb, ca;
                        if (!gradient_colour(pr, ramp, (float)x + 0.5f, (float)(y0 + q) + 0.5f, cr, cg, cb, ca)) continue;
                        // premultiplied, scaled like a masked entry's colour (the mask byte is the coverage)
                        const float k = ca * (1.0f / (255.0f * 255.0f));
                        const float wq = tr[q] * (float)((m >> (8 * q)) & 255u);
                        r[q] = fmaf(wq, cr * k, r[q]);
                        g[q] = fmaf(wq, cg * k, g[q]);
                        bl[q] = fmaf(wq, cb * k, bl[q]);
                        tr[q] = fmaf(-wq, k, tr[q]);
                    }
                } else {
                    blend(wcol[j], m);
                }
            }
            if (base == e0) break;
            continue;
        }
        int ja[EV_GROUP], jb[EV_GROUP];
        uint32_t ma[EV_GROUP], mb[EV_GROUP];
        take(ja, ma);
        while (true) {
            take(jb, mb);
            __builtin_amdgcn_sched_barrier(0);
            blend_group(ja, ma);
            __builtin_amdgcn_sched_barrier(0);
            if (jb[0] < 0) break;
            take(ja, ma);
            __builtin_amdgcn_sched_barrier(0);
            blend_group(jb, mb);
            __builtin_amdgcn_sched_barrier(0);
            if (ja[0] < 0) break;
        }
        if (base == e0) break;
    }
    if (x < 0 || x >= width) return;  // a view box past the target's left or right edge
    uint8_t *const cptr = out + (int64_t)y0 * pitch + (int64_t)x * 4;  // this lane's top pixel
    float al[4];
#pragma unroll
    for (int q = 0; q < 4; q++) al[q] = fmaf(tr[q], -255.0f, 255.0f);  // byte units, like r, g, bl
    if (load_dst) {
        // the destination lies below every entry: it shows through the remaining transmittance
#pragma unroll
        for (int q = 0; q < 4; q++) {
            uint32_t px = 0;
            if (y0 + q >= 0 && y0 + q < height) px = *reinterpret_cast<const uint32_t *>(cptr + q * pitch);
            r[q] = fmaf(tr[q], (float)(px & 255u), r[q]);
            g[q] = fmaf(tr[q], (float)((px >> 8) & 255u), g[q]);
            bl[q] = fmaf(tr[q], (float)((px >> 16) & 255u), bl[q]);
            al[q] = fmaf(tr[q], (float)(px >> 24), al[q]);
        }
    }
    // v_cvt_pk_u8_f32 rounds to nearest even and saturates to [0, 255] (NaN to 0), i.e. the
    // byte is rint(clamp(v, 0, 255)), written straight into its place in the word
    uint32_t px[4];
#pragma unroll
    for (int q = 0; q < 4; q++) {
        uint32_t w = __builtin_amdgcn_cvt_pk_u8_f32(r[q], 0, 0u);
        w = __builtin_amdgcn_cvt_pk_u8_f32(g[q], 1, w);
        w = __builtin_amdgcn_cvt_pk_u8_f32(bl[q], 2, w);
        px[q] = __builtin_amdgcn_cvt_pk_u8_f32(al[q], 3, w);
    }
#pragma unroll
    for (int q = 0; q < 4; q++)
        if (y0 + q >= 0 && y0 + q < height) *reinterpret_cast<uint32_t *>(cptr + q * pitch) = px[q];
}

// The pixels of a scene's target outside its view box's framebuffer tiles, which the composite does
// not visit, cleared to transparent (the reference clears the whole destination texture,
// Canvas::draw's clear_dst); scenes drawn over their target (PFH_LOAD_DST) keep them.  Block
// (tile row, scene): its 16 pixel rows, all of a row outside the tile rows, else the columns left
// and right of them.
__global__ __launch_bounds__(256) void k_clear_outside(BatchDev b) {
    const SceneDev &sc = b.scenes[blockIdx.y];
    if (sc.load_dst) return;
    const int y_lo = sc.fb_y0 * TILE, y_hi = (sc.fb_y0 + sc.fb_h) * TILE;
    const int x_lo = sc.fb_x0 * TILE, x_hi = (sc.fb_x0 + sc.fb_w) * TILE;
    const int r0 = (int)blockIdx.x * TILE;
    if (r0 >= sc.height) return;
    // nothing outside the tiles in these rows (the whole target, for a view box that covers it)
    if (r0 >= y_lo && r0 + TILE <= y_hi && x_lo <= 0 && x_hi >= sc.width) return;
    for (int i = threadIdx.x; i < TILE * sc.width; i += blockDim.x) {
        const int y = (int)blockIdx.x * TILE + i / sc.width, x = i % sc.width;
        if (y >= sc.height) break;
        if (y >= y_lo && y < y_hi && x >= x_lo && x < x_hi) continue;  // a framebuffer tile's pixel
        *reinterpret_cast<uint32_t *>(sc.out + (int64_t)y * sc.out_pitch + (int64_t)x * 4) = 0u;
    }
}

hipError_t launch_clear_outside(const BatchDev &b, uint32_t max_rows, hipStream_t s) {
    if (b.n_scenes <= 0 || max_rows == 0) return hipSuccess;
    hipLaunchKernelGGL(k_clear_outside, dim3(max_rows, (uint32_t)b.n_scenes), dim3(256), 0, s, b);
    return hipGetLastError();
}

// ---------------------------------------------------------------------------
// Clip combine (shaders/d3d9/tile_clip_combine.frag:25-27 and the draw path's fill rule in
// shaders/d3d9/tile.frag): a draw tile with its own mask over a clip tile with a mask gets
// rule(min(|own sum + own backdrop|, |clip sum + clip backdrop|)) per pixel.
// ---------------------------------------------------------------------------

// Mean of clamp(u, 0, 1) for u linear between ua and ub.
__device__ __forceinline__ float mean_clamp01(float ua, float ub) {
    const float a = fminf(ua, ub), c = fmaxf(ua, ub);
    if (c - a <= 0.0f) return fminf(fmaxf(a, 0.0f), 1.0f);
    const float full = c > 1.0f ? c - fmaxf(a, 1.0f) : 0.0f;
    const float lo = fmaxf(a, 0.0f), hi = fminf(c, 1.0f);
    const float mid = hi > lo ? (hi - lo) * (hi + lo) * 0.5f : 0.0f;
    return (full + mid) / (c - a);
}

// Signed area of one 8.8 fill over pixel (px, py) of its tile: the box-filtered coverage
// that the area LUT of shaders/d3d9/fill.frag:25-45 tabulates, evaluated directly.
__device__ __forceinline__ float fill_area_at(uint2 fl, float px, float py) {
    const float k = 1.0f / 256.0f;
    const float x0 = (float)(fl.x & 0xffffu) * k, y0 = (float)(fl.x >> 16) * k;
    const float x1 = (float)(fl.y & 0xffffu) * k, y1 = (float)(fl.y >> 16) * k;
    if (x0 == x1) return 0.0f;
    const float xl = fminf(x0, x1), xr = fmaxf(x0, x1);
    const float yl = x0 < x1 ? y0 : y1, yr = x0 < x1 ? y1 : y0;
    const float lo = fmaxf(px, xl), hi = fminf(px + 1.0f, xr);
    if (hi <= lo) return 0.0f;
    const float slope = (yr - yl) / (xr - xl);
    const float ua = py + 1.0f - (yl + (lo - xl) * slope), ub = py + 1.0f - (yl + (hi - xl) * slope);
    return (x0 > x1 ? 1.0f : -1.0f) * (hi - lo) * mean_clamp01(ua, ub);
}

__device__ __forceinline__ float fill_rule_f(float c, uint32_t even_odd) {
    if (!even_odd) return fminf(fabsf(c), 1.0f);
    const float m = c - 2.0f * floorf(c * 0.5f);
    return fminf(fmaxf(1.0f - fabsf(1.0f - m), 0.0f), 1.0f);
}

// A workgroup per combine job (rare: only tiles where a clipped path's and its clip's masks
// meet), after k_fill_masks, whose mask for the row it replaces.  The coverage is the min over a
// chain of terms |fills + backdrop| of mask rows: the draw tile's own row (a draw job), then the
// clip's coverage source and the rows its record lists (the clip chain, up to MAX_CLIP_LEVELS).
// A draw job writes the draw path's fill rule of it to its row; a nested clip path's job writes
// the nonzero and the even-odd rule of it to its two rows.  Where clip paths' own rules are taken
// (CLIP_JOB_OWN_RULES) each term is its row's fill rule of fills + backdrop instead, and the min
// of those is the coverage, written as it is.  Each lane owns 4 pixels of one column;
// the 4 waves take every 4th fill of every term and their partial sums meet in LDS, so a long fill
// list costs a quarter of its length in serial steps.
constexpr int CLIP_TERMS = MAX_CLIP_LEVELS + 1;

__global__ __launch_bounds__(256) void k_fill_clipped(BatchDev b) {
    __shared__ float part[CLIP_TERMS][4][4][64];  // term x wave x pixel row offset x lane
    const uint32_t lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    const uint32_t n = min(b.counters[C_CLIP_JOBS], b.cap_alpha);
    const float px = (float)(lane & 15u), py0 = (float)((lane >> 4) * 4u);
    for (uint32_t j = blockIdx.x; j < n; j += gridDim.x) {  // uniform over the block
        const uint2 job = b.clip_jobs[j];
        const bool clip_job = (job.x & CLIP_JOB_CLIP) != 0, own_rules = (job.x & CLIP_JOB_OWN_RULES) != 0;
        const uint32_t out = job.x & ~(CLIP_JOB_CLIP | CLIP_JOB_OWN_RULES), src = job.y;
        bool ok = out + (clip_job ? 1u : 0u) < b.cap_alpha && src < b.cap_alpha;
        // terms: the draw tile's own row (draw jobs), then the clip chain (fixed slots, so that
        // the arrays stay in registers)
        const uint4 rec = ok ? b.clip_rec[src] : make_uint4(1u, 0u, 0u, 0u);
        const uint32_t chain = max(1u, min(rec.x, (uint32_t)MAX_CLIP_LEVELS));
        const uint32_t first = clip_job ? 1u : 0u, nt = 1u + chain;  // slots [first, nt)
        const uint32_t term[CLIP_TERMS] = {out, src, rec.y, rec.z, rec.w};
#pragma unroll
        for (int k = 0; k < CLIP_TERMS; k++) ok = ok && ((uint32_t)k < first || (uint32_t)k >= nt || term[k] < b.cap_alpha);
        int8_t bdk[CLIP_TERMS];
        uint32_t rk[CLIP_TERMS];  // each term's fill rule (its row's)
#pragma unroll
        for (int k = 0; k < CLIP_TERMS; k++) {
            bdk[k] = 0;
            rk[k] = 0;
            if ((uint32_t)k < first || (uint32_t)k >= nt) continue;
            const uint4 r = ok ? b.alpha_rec[term[k]] : make_uint4(0, 0, 0, 0);
            bdk[k] = (int8_t)(r.z & 0xffu);
            rk[k] = (r.z >> 8) & 1u;
            float acc[4] = {0.0f, 0.0f, 0.0f, 0.0f};
            const uint32_t f0 = min(r.x, b.cap_fills), f1 = max(f0, min(r.y, b.cap_fills));
            for (uint32_t f = f0 + wv; f < f1; f += 4) {
                const uint2 fl = b.fills[f];
#pragma unroll
                for (int q = 0; q < 4; q++) acc[q] += fill_area_at(fl, px, py0 + (float)q);
            }
#pragma unroll
            for (int q = 0; q < 4; q++) part[k][wv][q][lane] = acc[q];
        }
        __syncthreads();
        if (wv == 0 && ok) {
            const uint32_t rule = clip_job ? 0u : (b.alpha_rec[out].z >> 8) & 1u;
            uint32_t packed = 0, packed_eo = 0;
#pragma unroll
            for (int q = 0; q < 4; q++) {
                float m = INFINITY;
#pragma unroll
                for (int k = 0; k < CLIP_TERMS; k++) {
                    if ((uint32_t)k < first || (uint32_t)k >= nt) continue;
                    const float v = (float)bdk[k] + part[k][0][q][lane] + part[k][1][q][lane] + part[k][2][q][lane] +
                                    part[k][3][q][lane];
                    m = fminf(m, own_rules ? fill_rule_f(v, rk[k]) : fabsf(v));
                }
                const float c = own_rules ? m : fill_rule_f(m, rule), c_eo = own_rules ? m : fill_rule_f(m, 1u);
                packed |= (uint32_t)__float2int_rn(c * 255.0f) << (8 * q);
                packed_eo |= (uint32_t)__float2int_rn(c_eo * 255.0f) << (8 * q);
            }
            // column-major mask: byte 16 x + y
            const uint32_t off = (lane & 15u) * 16 + (lane >> 4) * 4;
            *reinterpret_cast<uint32_t *>(b.masks + (uint64_t)out * 256 + off) = packed;
            if (clip_job) *reinterpret_cast<uint32_t *>(b.masks + (uint64_t)(out + 1) * 256 + off) = packed_eo;
        }
        __syncthreads();  // the next job rewrites the partial sums
    }
}

static inline uint32_t blocks_for(uint64_t n, uint32_t per) { return (uint32_t)((n + per - 1) / per); }

hipError_t launch_fill_masks(const BatchDev &b, uint32_t grid_alpha, uint32_t max_blocks, hipStream_t s) {
    // at least one workgroup (it writes the all-ones row), at most max_blocks: each loops over
    // its share of the tiles
    const uint32_t nb = std::max(1u, std::min<uint32_t>(blocks_for(grid_alpha, FILL_TILES_PER_BLOCK), max_blocks));
    hipLaunchKernelGGL(k_fill_masks, dim3(nb), dim3(256), 0, s, b);
    return hipGetLastError();
}

hipError_t launch_fill_clipped(const BatchDev &b, hipStream_t s) {
    if (!b.has_clips) return hipSuccess;
    // the job count is only known on the device: the workgroups stride over it
    hipLaunchKernelGGL(k_fill_clipped, dim3(2048), dim3(256), 0, s, b);
    return hipGetLastError();
}

hipError_t launch_composite(const BatchDev &b, const uint32_t *sorted_vals, uint32_t block, hipStream_t s) {
    if (b.n_fb == 0) return hipSuccess;
    if (block != 64 && block != 128 && block != 256) return hipErrorInvalidValue;
    if (b.paints)
        hipLaunchKernelGGL(k_composite<true>, dim3(blocks_for(b.n_fb, block / 64)), dim3(block), 0, s, b, sorted_vals);
    else
        hipLaunchKernelGGL(k_composite<false>, dim3(blocks_for(b.n_fb, block / 64)), dim3(block), 0, s, b, sorted_vals);
    return hipGetLastError();
}

}  // namespace pfh
