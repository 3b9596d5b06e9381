// Internal layouts shared by the HIP kernels (pfh_kernels.hip) and the host
// runtime (pfh_runtime.hip).  Not part of the ABI.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace pfh {

constexpr int TILE = 16;                 // TILE_WIDTH / TILE_HEIGHT, core/data/data.h:15-16
// Recursion cap of process_segment: the flatness measure shrinks 16x per halving, so a
// curve with coordinates below ~2^30 px is flat long before this depth; deeper curves
// are emitted as chords and counted in C_GUARD (the oracle applies the same cap).
constexpr int MAX_FLATTEN_DEPTH = 16;
constexpr int FIX_SHIFT = 20;            // fixed-point coverage accumulation: 2^-20 px^2 units
constexpr uint32_t NO_ALPHA = 0xffffffffu;

// Paint records.  An entry of a path with a paint record (a gradient, a non-occluding solid, a
// composite op) carries ENTRY_GRADIENT in its first word (with its mask row, ENTRY_ROW_NONE in the
// low bits for a tile without a mask, or ENTRY_ROW_EMPTY for an even-odd tile of even winding,
// which covers nothing but still composites) and the scene-local paint index in its second;
// NO_ALPHA (all bits set) stays the plain solid entry.
constexpr uint32_t ENTRY_GRADIENT = 0x40000000u;
constexpr uint32_t ENTRY_ROW_MASK = 0x3fffffffu;
constexpr uint32_t ENTRY_ROW_NONE = 0x3fffffffu;
constexpr uint32_t ENTRY_ROW_EMPTY = 0x3ffffffeu;
constexpr uint32_t PAINT_NONE = 0xffffffffu;   // SceneDev::path_paint: a solid colour
constexpr int RAMP_TEXELS = 256;               // GRADIENT_TILE_LENGTH, core/paint/gradient.h
// One gradient paint as the tile shader sees it (shaders/d3d9/tile.vert:65-77, tile.frag:296-323):
// the colour texture transform and the radial filter's parameters as the RGBA16F metadata texture
// holds them (core/renderer.cpp:167-250), and its row of RAMP_TEXELS RGBA8 texels.
struct PaintRec {
    float m11, m21, m12, m22, m13, m23;  // pixel -> colour texture coordinates (radial: gradient space)
    float from_x, from_y, vec_x, vec_y;  // radial: the line between the circles' centres
    float r0, r1, uv_x, uv_y;            // radial: radii, the ramp row's uv origin
    uint32_t kind;                       // 0 solid (colour), 1 linear, 2 radial
    uint32_t wrap;                       // 0 clamp to edge, 1 repeat (REPEAT_U)
    uint32_t colour;                     // kind 0: r | g << 8 | b << 16 | a << 24
    uint32_t op;                         // composite op (COMBINER_CTRL_COMPOSITE_*): 0 normal (over)
};
// A path with a paint record: a gradient, or a solid colour whose blend mode does not occlude
// (drawn as over, but never written to the z-buffer) or composites through a non-normal op.

// Clip reference of a path (SceneDev::path_clip, then BinItem::clip).  A path clipped by a clip
// path (a draw path, or a clip path nested in another) holds that clip path's binning position: its
// index in the scene's wave-class list, or CLIP_GROUP | its index in the group-class list
// (BinItem::clip: the batch bin item index).  Whether a path is a clip path is its clip level
// (SceneDev::path_level, BinItem::level): 0 for a draw path, 1 + its nesting depth for a clip path.
constexpr uint32_t CLIP_NONE = 0xffffffffu;     // not clipped
constexpr uint32_t CLIP_EMPTY = 0xfffffffdu;    // clipped by a clip path with no tile columns: all culled
constexpr uint32_t CLIP_GROUP = 0x80000000u;
constexpr uint32_t MAX_CLIP_LEVELS = 4;         // clip paths in one nesting chain
// clip_jobs[j].x: the mask row a combine job writes; CLIP_JOB_CLIP marks a nested clip path's tile,
// whose two rows (nonzero, even-odd) both take the chain's coverage; CLIP_JOB_OWN_RULES marks a
// scene whose clip paths' coverage is taken under their own fill rules (PFH_CLIP_RULE_OWN)
constexpr uint32_t CLIP_JOB_CLIP = 0x80000000u;
constexpr uint32_t CLIP_JOB_OWN_RULES = 0x40000000u;
// BinItem::rule: the fill rule in bit 0; RULE_CLIP_OWN set for every path of a scene whose clip
// paths' coverage is taken under their own fill rules
constexpr uint32_t RULE_CLIP_OWN = 0x2u;

// Binning ownership classes (a path's tiles and columns belong to it alone, so the wave or
// workgroup that bins a path can count in LDS and store plainly).  A path goes to the wave
// class when its tile rect and column count fit one wave's LDS slice and it has at most as many
// segments as a wave has lanes, else to the group class, which uses LDS up to BIN_GROUP_TILES /
// BIN_GROUP_COLS and global atomics beyond.
constexpr int BIN_WAVE_TILES = 512;
constexpr int BIN_WAVE_COLS = 128;
constexpr int BIN_WAVE_SEGS = 64;       // more segments than lanes: the group class (its walk is 4x wider)
constexpr int BIN_GROUP_TILES = 4600;   // with the walk queues: 20 KB (emit) and 40 KB (count) of LDS
constexpr int BIN_GROUP_COLS = 512;
constexpr int WALK_QUEUE = 128;          // per wave: chords that cross tile boundaries, waiting for a full wave

enum SegKind : uint32_t { SEG_LINE = 1, SEG_QUAD = 2, SEG_CUBIC = 3 };
// Device segment word (.y): kind in the low byte (0: emits nothing), SEG_CHAIN_BREAK when the
// segment does not start where the previous emitting segment of its path ended (bit for bit):
// its chord points are then preceded by a NaN separator and its start point.
constexpr uint32_t SEG_KIND_MASK = 0xffu;
constexpr uint32_t SEG_CHAIN_BREAK = 0x100u;

// Launch schedule (chosen by pfh_autotune; every schedule gives bit-identical results).
struct Schedule {
    uint32_t fill_grid = 16384;       // max workgroups of k_fill_masks (grid-stride beyond)
    uint32_t composite_block = 256;   // threads per k_composite workgroup (a wave per fb tile)
    uint32_t column_block = 256;      // threads per workgroup of k_propagate / k_entries
    uint32_t xcd_order = 0;           // bit 0 fill, bit 1 composite: workgroups grouped per XCD
    uint32_t flatten_segs = 0;        // segments per k_flatten wave (multiple of 64), 0: from the count
    uint32_t flatten_take = 8;        // idle k_flatten lanes a wave waits for before a hand-over (1..64)
    uint32_t column_rows = 0;         // path columns of more tiles are walked by 16 or 64 lanes (1..64;
                                      // 0: every column by one lane)
    uint32_t flatten_deep = 0;        // curves whose estimated subdivision depth is at least this are
                                      // flattened by 16 lanes each (k_flatten_deep; 0: none)
};

// Inclusive prefix sum within each 16-lane row of the wave (DPP row shifts: no LDS).
__device__ __forceinline__ uint32_t row16_inclusive_sum(uint32_t v) {
    v += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x111, 0xf, 0xf, true);  // row_shr:1
    v += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x112, 0xf, 0xf, true);  // row_shr:2
    v += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x114, 0xf, 0xf, true);  // row_shr:4
    v += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x118, 0xf, 0xf, true);  // row_shr:8
    return v;
}

// MI355X dispatches consecutive workgroups round-robin over its 8 XCDs (each with its own L2).
// Logical block of hardware block b when the grid's blocks are regrouped so that each XCD
// processes one contiguous range of logical blocks (a bijection on [0, nb)).
__device__ __forceinline__ uint32_t xcd_block(uint32_t b, uint32_t nb) {
    const uint32_t q = nb >> 3, r = nb & 7u, x = b & 7u;
    return x * q + min(x, r) + (b >> 3);
}

// Scene holding batch item i: the largest k in [0, n) with bases[k] <= i, where bases holds
// the n + 1 scanned per-scene item offsets (bases[0] = 0 <= i < bases[n]).  i is the same in
// every lane and all 64 lanes are active: a 64-ary search, one vector load and one ballot per
// round (one round up to 64 scenes, two up to 4096), in place of a chain of dependent loads.
__device__ __forceinline__ int find_scene_wave(const uint32_t *bases, int n, uint32_t i) {
    const int lane = (int)(threadIdx.x & 63);
    int lo = 0, hi = n - 1;
    while (lo < hi) {
        const int step = (hi - lo + 64) >> 6;  // candidates lo, lo + step, ... cover [lo, hi]
        const int cand = lo + lane * step;
        const uint64_t ok = __ballot(cand <= hi && bases[cand] <= i);  // a prefix of the lanes, lane 0 set
        lo += (63 - __clzll(ok)) * step;
        hi = min(hi, lo + step - 1);
    }
    return lo;
}

// Scene of each lane's item, when no lane's item is below lane 0's (all 64 lanes active):
// lane 0's scene, then a short walk over the scene boundaries inside the wave's items.
__device__ __forceinline__ int find_scene_lanes(const uint32_t *bases, int n, uint32_t i) {
    int k = find_scene_wave(bases, n, __builtin_amdgcn_readfirstlane(i));
    while (bases[k + 1] <= i) k++;
    return k;
}

// What a tile draws, from its fills and propagated backdrop (written by propagate).
enum TileStatus : uint8_t {
    TILE_EMPTY = 0,  // no fills, zero backdrop: no draw entry
    TILE_ALPHA = 1,  // has fills: a coverage mask
    TILE_SOLID = 2,  // no fills, backdrop covers the whole tile
    TILE_CLEAR = 3,  // no fills, even-odd rule with an even nonzero backdrop: an entry that covers nothing
    TILE_CLIP_MASK = 4,  // no fills, nonzero backdrop, over a clip tile with fills: draws the clip's
                         // coverage (mask row tile_clip_ai)
};

// Counters written by the device pipeline (one block per batch).
enum Counter {
    C_FILLS = 0,       // total fills (count pass)
    C_ALPHA = 1,       // total alpha tiles
    C_ENTRIES = 2,     // total draw entries after the z-cull
    C_OVERFLOW = 3,    // bit 0 fills, bit 1 alpha tiles, bit 2 entries exceeded capacity
    C_GUARD = 4,       // loop guards tripped
    C_ALPHA_ENTRIES = 5,  // draw entries that sample a mask
    C_LINES = 6,       // flattened chords (lines) in the batch
    C_CLIP_JOBS = 7,   // clip combine jobs (alpha tiles over clip tiles with fills)
    C_DRAWN = 8,       // [C_DRAWN, C_DRAWN + DRAWN_SLOTS): fills the fill kernel read, spread over
                       // slots (per block, by block index) so that no single address takes every add
    C_CLIP_DUPS = 8 + 64,  // draw entries whose clip mask row an earlier entry already listed
    C_TALL = 8 + 64 + 1,   // [C_TALL, C_TALL + 6): per propagate phase, columns of more than
                           // TALL_ROWS tiles (and Schedule::column_rows) listed for a wave each
    C_MID = 8 + 64 + 1 + 6,  // [C_MID, C_MID + 6): per phase, columns of more than
                             // Schedule::column_rows tiles and at most TALL_ROWS (k_propagate_rows<16>)
    C_DEEP = 8 + 64 + 1 + 12,  // [C_DEEP, C_DEEP + DEEP_SLOTS): curves listed for k_flatten_deep, per
                               // list region
    C_COUNT = 8 + 64 + 1 + 12 + 256
};
// The deep-curve list is split into DEEP_SLOTS regions of deep_slot_cap entries, a listing wave
// taking region (its index mod DEEP_SLOTS) with that region's counter: no single address takes
// every wave's add.
constexpr uint32_t DEEP_SLOTS = 256;

// A curve listed for k_flatten_deep, with what the listing pass loaded of it: its batch segment,
// its scene | DEEP_QUAD | DEEP_BREAK, and its points (a quad as p0, c, c, p1).
constexpr uint32_t DEEP_QUAD = 0x40000000u, DEEP_BREAK = 0x80000000u;
struct __align__(8) DeepCurve {
    uint32_t seg, scene;
    float2 p0, c0, c1, p1;
};

// Arrays zeroed by one launch: p[k] has end[k] - end[k - 1] words (end[-1] = 0), k < n.
constexpr int ZERO_SPANS = 6;
struct ZeroSpans {
    uint32_t *p[ZERO_SPANS];
    uint64_t end[ZERO_SPANS];
    int n;
    void add(uint32_t *ptr, uint64_t words) {
        if (!words) return;
        p[n] = ptr;
        end[n] = (n ? end[n - 1] : 0) + words;
        n++;
    }
};
constexpr int TALL_ROWS = 64;
constexpr int DRAWN_SLOTS = 64;

// One scene of a batch as the kernels see it.  Scene-local arrays are resident
// from upload; *_base fields place the scene inside the batch's work arrays.
struct SceneDev {
    const float2 *pts;            // packed points
    const uint2 *segs;            // {first point, SegKind}
    const uint32_t *seg_path;     // local path of each segment
    const int4 *path_rect;        // tile rect (l, t, r, b) per path
    const uint32_t *path_tile_off;  // [n_paths + 1] local tile offsets
    const uint32_t *path_col_off;   // [n_paths + 1] local column offsets
    const uint32_t *col_path;     // local column -> local path
    const uint32_t *path_rgba;    // r | g << 8 | b << 16 | a << 24
    const uint32_t *path_seg_off; // [n_paths + 1] local segment offsets
    const uint32_t *bin_paths;    // local path indices: wave class, then group class
    const uint8_t *path_rule;     // 0 nonzero, 1 even-odd
    uint32_t clip_own_rules;      // 1: clip paths' coverage under their own fill rules (PFH_CLIP_RULE_OWN)
    const uint32_t *path_clip;    // [n_paths] clip reference (CLIP_*), or null: no clip paths
    const uint8_t *path_level;    // [n_paths] clip level (0 draw path, 1..4 clip path), or null
    float vb_l, vb_t, vb_r, vb_b; // view box (right edge rounded up to 16)
    int32_t fb_x0, fb_y0, fb_w, fb_h;
    int32_t width, height;
    uint32_t n_paths, n_segs, n_tiles, n_cols, n_fb;
    uint32_t n_bin_wave, n_bin_group;  // paths in each binning class
    uint32_t bin_wave_base, bin_group_base;  // first item of this scene in the batch's class lists
    uint32_t seg_base, path_base, tile_base, col_base, fb_base;
    uint8_t *out;                 // RGBA8 target
    int64_t out_pitch;
    uint32_t load_dst;            // blend over existing target contents
    float fb_inv_w;               // 1 / fb_w (0 without framebuffer tiles)
    const uint32_t *path_paint;   // [n_paths] gradient paint index or PAINT_NONE, or null: solid colours
    const PaintRec *paints;       // the scene's paint records (gradients, blended solids)
    uint32_t has_ops;             // some record composites through an op (tiles with one blend back to front)
    const uint32_t *ramps;        // [paints * RAMP_TEXELS] RGBA8 ramp texels, r | g << 8 | b << 16 | a << 24
};

// A binned path, resolved once per launch (after the chord scan) so that the per-path
// kernels start from one load instead of a chain of dependent ones.
struct BinItem {
    int4 rect;                 // tile rect (l, t, r, b)
    uint32_t tbase, cbase;     // batch index of its first tile / column
    uint32_t c0, c1;           // its chords [c0, c1)
    uint32_t p, gp, k;         // local path, batch path, scene
    uint32_t rgba, rule;       // paint, fill rule (bit 0) | RULE_CLIP_OWN
    uint32_t clip;             // CLIP_NONE / CLIP_EMPTY, or the bin item of the path's clip path
    uint32_t level;            // 0: a draw path; 1..4: a clip path at that nesting depth (never drawn)
    uint32_t paint;            // gradient paint index, or PAINT_NONE (a solid colour)
};

// Batch-wide work arrays (device pointers).  Sizes: S segments, T tiles,
// W columns, FB framebuffer tiles, F/A/E capacities for fills, alpha tiles, entries.
struct BatchDev {
    const SceneDev *scenes;
    int32_t n_scenes;
    const uint32_t *seg_bases;    // [n_scenes + 1]
    const uint32_t *tile_bases;
    const uint32_t *col_bases;
    const uint32_t *fb_bases;
    const uint32_t *bin_wave_bases;   // [n_scenes + 1] wave-class items per scene, scanned
    const uint32_t *bin_group_bases;  // [n_scenes + 1] group-class items per scene, scanned
    uint32_t n_segs, n_tiles, n_cols, n_fb, n_paths;
    uint32_t n_bin_wave, n_bin_group;
    uint32_t cap_fills, cap_alpha, cap_entries, cap_lines;
    uint32_t paints;              // some scene of the batch has gradient paints
    uint32_t cap_vis;             // alpha_vis capacity: cap_alpha, or with clip paths the larger of it and
                                  // cap_entries (solid tiles of several paths can take one clip mask)
    int32_t *tile_delta;          // [T] backdrop delta (int8 arithmetic, stored widened)
    uint32_t *tile_cnt;           // [T] fills per tile (global-atomic paths: then emit cursor)
    uint32_t *tile_fill_off;      // [T + 1] exclusive scan of fill counts
    uint32_t *tile_alpha_idx;     // [T + 1] exclusive scan of alpha flags
    int8_t *tile_bd;              // [T] propagated backdrop (written with dumps or clip paths only)
    uint32_t *tile_clip_ai;       // [T] with clip paths: mask row of a TILE_CLIP_MASK tile; for a clip
                                  // path's tile with a mask, its coverage source (below)
    uint2 *clip_jobs;             // [A] with clip paths: (mask row | CLIP_JOB_CLIP, coverage source)
    // A clip path's coverage at a tile is min over a chain of terms |fills + backdrop of a mask row|.
    // Its source s (tile_clip_ai) is the mask row whose record holds the chain: clip_rec[s] =
    // (n terms, t1, t2, t3), the first term being row s itself; rows s, s + 1 hold that coverage
    // under the nonzero and the even-odd rule.
    uint4 *clip_rec;              // [A] with clip paths
    uint32_t *clip_seen;          // [A / 32 + 1] with clip paths: mask rows already listed for the fill
    uint32_t clip_levels;         // clip levels present in the batch (0: no clip paths)
    uint8_t *tile_st;             // [T] tile status (TileStatus)
    uint2 *path_ent;              // [P + 1] {draw entries, alpha entries} per batch path, then exclusive scan
    uint2 *ent_quarter;           // [4 n_bin_group] {draw entries, alpha entries} per quarter of a group-class rect
    uint2 *path_tot;              // [P + 1] {fills, alpha tiles} per batch path, then exclusive scan
    BinItem *bin_items;           // [n_bin_wave + n_bin_group] wave class, then group class
    uint32_t *col_item;           // [W] the bin item owning each column
    int32_t *col_bd;              // [W] above-rect backdrop, then final column sum
    uint32_t *tall_cols;          // [W] columns of the current propagate phase walked by several lanes:
                                  // the C_MID list from the front, the C_TALL list from the back
    uint32_t *seg_cnt;            // [S + 1] chords per segment, then exclusive scan
    uint64_t *curve_mask;         // [S / 64 + 1] bit i of word j: segment 64 j + i is a curve (to be
                                  // subdivided), written by the flatten count pass
    DeepCurve *deep_list;         // [DEEP_SLOTS * deep_slot_cap] with deep_pred: the curves left to
                                  // k_flatten_deep, per region in no particular order
    uint32_t deep_slot_cap;       // entries per region: 64 per wave of k_flatten_lines mapped to it
    uint32_t deep_pred;           // Schedule::flatten_deep (0: every curve through k_flatten)
    // [L] flattened chords as a point stream, in emission order: per chain of chords a NaN
    // separator and the start point, then each chord's end point; chord i of a path's range
    // [c0, c1) runs from point i to point i + 1 (a pair with a NaN end emits nothing)
    float2 *lines;
    uint32_t *line_cnt;           // [L + 1] fills per chord, then exclusive scan (emission order)
    uint2 *fills;                 // [F] grouped by tile: (x0 | y0 << 16, x1 | y1 << 16)
    uint4 *dump_fills;            // [F] emission order: (x0|y0<<16, x1|y1<<16, global tile, 0)
    uint4 *alpha_rec;             // [A] (fill begin, fill end, backdrop | rule << 8, tile)
    uint32_t *alpha_vis;          // [cap_vis] mask rows of the alpha entries (tiles that survive the z-cull)
    uint8_t *masks;               // [(A + 1) * 256]: row cap_alpha is all ones (written by the fill kernel)
    uint32_t *zbuf;               // [FB]
    uint2 *fb_range;              // [FB] [start, end) of each framebuffer tile's entries in sort order
    uint32_t *keys, *vals;        // [E] sort keys / values
    uint32_t *keys_alt, *vals_alt;
    uint2 *entry_rec;             // [E] (alpha index | NO_ALPHA, rgba (alpha 0 if covers nothing)), or for a
                                  // gradient path (ENTRY_GRADIENT | mask row or ENTRY_ROW_NONE, paint index)
    uint2 *entry_meta;            // [E] with dumps only: (local path, scene-local fb tile)
    uint32_t *counters;           // [C_COUNT]
    int32_t dump;
    int32_t has_clips;            // some scene has clip paths: propagate in two phases, combine masks
    uint32_t xcd_order;           // Schedule::xcd_order
};

}  // namespace pfh
