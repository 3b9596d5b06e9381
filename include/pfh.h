/*
 * pfh.h — C ABI of the MI355X (gfx950) HIP rasterizer that replaces the
 * reference's GPU-driven rendering path for a scene:
 *
 *   Canvas::draw (core/canvas.cpp:557-567)
 *     -> SceneBuilder::build (core/scene_builder.h:16-32)
 *     -> Renderer::draw      (core/renderer.h:83)
 *
 * i.e. SceneBuilderD3D11::build (core/d3d11/scene_builder.cpp:209-223) plus
 * RendererD3D11::draw (core/d3d11/renderer.cpp:302-336) and its seven compute
 * stages (shaders/d3d11/{bound,dice,bin,propagate,sort,fill,tile}.comp), with
 * the hybrid CPU tiler's arithmetic (core/d3d9/tiler.cpp) so that tiles, fills,
 * backdrops and draw order are bit-identical to it.
 *
 * Conventions: plain pointers and sizes only; host arrays are borrowed for the
 * duration of the call; the return value is PFH_OK (0) or a negative PFH_E*
 * code, and pfh_last_error() describes the failure.  A context is externally
 * synchronised (one host thread at a time).  Nothing throws across the ABI.
 */
#ifndef PFH_H
#define PFH_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define PFH_OK 0
#define PFH_E_ARG (-1)       /* invalid argument */
#define PFH_E_HIP (-2)       /* HIP runtime error */
#define PFH_E_NOMEM (-3)     /* device allocation failed */
#define PFH_E_STATE (-4)     /* call out of order (e.g. dump before rasterize) */
#define PFH_E_LIMIT (-5)     /* a size exceeds a documented limit */
#define PFH_E_OVERFLOW (-6)  /* capacity exceeded and retries exhausted */

/* pfh_rasterize flags */
#define PFH_PLAIN 0x0u      /* per scene, stage by stage, exact sizes read back (RendererD3D11::draw structure) */
#define PFH_BATCHED 0x1u    /* all scenes in one pass, no host round trips; replayed from a captured HIP graph */
#define PFH_PREDICT 0x2u    /* size fill/alpha/entry buffers from the learned predictor (PFH_BATCHED) */
#define PFH_DUMP 0x4u       /* keep intermediates for pfh_dump (canonical fills in emission order, masks) */
#define PFH_LOAD_DST 0x8u   /* blend over the current target contents instead of clearing (clear_dst = false) */
#define PFH_NO_GRAPH 0x10u  /* PFH_BATCHED without graph capture (eager launches) */

/* Fill rules (core/data/data.h:58-63 FillRule). */
#define PFH_FILL_NONZERO 0
#define PFH_FILL_EVEN_ODD 1
#define PFH_CLIP_RULE_ABS 0    /* pfh_scene_desc::clip_rule: the reference's clip coverage */
#define PFH_CLIP_RULE_OWN 1    /* pfh_scene_desc::clip_rule: ClipPath::fill_rule honoured */

/* Blend modes (BlendMode, core/paint/effects.h:56-123, same order). */
#define PFH_BLEND_CLEAR 0
#define PFH_BLEND_COPY 1
#define PFH_BLEND_SRC_IN 2
#define PFH_BLEND_SRC_OUT 3
#define PFH_BLEND_SRC_OVER 4
#define PFH_BLEND_SRC_ATOP 5
#define PFH_BLEND_DEST_IN 6
#define PFH_BLEND_DEST_OUT 7
#define PFH_BLEND_DEST_OVER 8
#define PFH_BLEND_DEST_ATOP 9
#define PFH_BLEND_XOR 10
#define PFH_BLEND_LIGHTER 11
#define PFH_BLEND_DARKEN 12
#define PFH_BLEND_LIGHTEN 13
#define PFH_BLEND_MULTIPLY 14
#define PFH_BLEND_SCREEN 15
#define PFH_BLEND_HARD_LIGHT 16
#define PFH_BLEND_OVERLAY 17
#define PFH_BLEND_COLOR_DODGE 18
#define PFH_BLEND_COLOR_BURN 19
#define PFH_BLEND_SOFT_LIGHT 20
#define PFH_BLEND_DIFFERENCE 21
#define PFH_BLEND_EXCLUSION 22
#define PFH_BLEND_HUE 23
#define PFH_BLEND_SATURATION 24
#define PFH_BLEND_COLOR 25
#define PFH_BLEND_LUMINOSITY 26

typedef struct pfh_ctx pfh_ctx;

/* An outline: contours of points with PointFlag flags (core/data/path.h:12-57, Outline / Contour). */
typedef struct pfh_outline {
    int32_t n_contours, n_points;
    const float *points;          /* [2 * n_points] */
    const uint8_t *point_flags;   /* [n_points] PointFlag */
    const int32_t *contour_start; /* [n_contours + 1] */
    const uint8_t *contour_closed;/* [n_contours] */
} pfh_outline;

/* StrokeStyle (core/stroke.h:12-26); LineCap / LineJoin values of core/data/data.h:73-93. */
#define PFH_CAP_BUTT 0
#define PFH_CAP_SQUARE 1
#define PFH_CAP_ROUND 2
#define PFH_JOIN_MITER 0
#define PFH_JOIN_BEVEL 1
#define PFH_JOIN_ROUND 2
/* Dashes (OutlineDash, core/dash.cpp:9-126), applied before stroking as Canvas::stroke_path does
 * (core/canvas.cpp:284-289): dash_count > 0 dashes the path with the pattern
 * dashes[dash_start .. dash_start + dash_count) (on, off, on, ... lengths; finite, non-negative,
 * of positive total) starting dash_offset into it (finite, non-negative; Canvas::stroke_path
 * passes 0 whatever its line_dash_offset, the dasher's own offset argument is exposed here).
 * dash_count 0 strokes the path undashed. */
typedef struct pfh_stroke_style {
    float line_width;
    int32_t line_cap;
    int32_t line_join;
    float miter_limit;
    int32_t dash_start;
    int32_t dash_count;
    float dash_offset;
} pfh_stroke_style;


/* Gradient paints (Gradient, core/paint/gradient.h:101-176; Paint::from_gradient, core/paint/paint.h:87-96:
 * white base colour, the gradient combined "source in").  A linear gradient runs along `line`; a
 * radial one between the circles (line.from, radii[0]) and (line.to, radii[1]) of gradient space,
 * which `transform` maps to the scene.  Stops: offset and colour (r | g << 8 | b << 16 | a << 24),
 * added in the order given (Gradient::add, core/paint/gradient.cpp:7-16). */
#define PFH_GRADIENT_LINEAR 0
#define PFH_GRADIENT_RADIAL 1
#define PFH_WRAP_CLAMP 0       /* GradientWrap::Clamp */
#define PFH_WRAP_REPEAT 1      /* GradientWrap::Repeat */
typedef struct pfh_gradient {
    int32_t kind;              /* PFH_GRADIENT_* */
    int32_t wrap;              /* PFH_WRAP_* */
    float line[4];             /* from x, y, to x, y */
    float radii[2];            /* radial */
    float transform[6];        /* radial: GradientRadial::transform, Transform2 as m11, m21, m12, m22, m13, m23:
                                * (x, y) -> (m11 x + m12 y + m13, m21 x + m22 y + m23); identity 1 0 0 1 0 0 */
    int32_t stop_start, stop_count;  /* into the stop arrays; stop_count < 0: the ramp is given instead
                                      * (pfh_scene_desc::gradient_ramps) */
} pfh_gradient;

/*
 * A scene as outlines: the data of Scene::draw_paths (core/scene.h:57-61) after
 * Path2d/Canvas have applied transforms, strokes and dashes (core/canvas.cpp:180-301).
 * Point flags follow PointFlag (core/data/data.h:42-51): 0 on-curve,
 * 1 first control point (quadratic or cubic), 2 second control point (cubic).
 * Colours are solid paints, r | g << 8 | b << 16 | a << 24 (ColorU).
 */
typedef struct pfh_scene_desc {
    int32_t width, height;              /* canvas size in pixels */
    float view_box[4];                  /* left, top, right, bottom (Scene::set_view_box, core/scene.cpp:170-179) */
    int32_t n_paths, n_contours, n_points;
    const float *points;                /* [2 * n_points] */
    const uint8_t *point_flags;         /* [n_points] */
    const int32_t *contour_start;       /* [n_contours + 1] */
    const uint8_t *contour_closed;      /* [n_contours] */
    const int32_t *path_contour_start;  /* [n_paths + 1] */
    const uint8_t *path_fill_rule;      /* [n_paths] PFH_FILL_* */
    const uint32_t *path_rgba;          /* [n_paths] */
    /* Clip paths (Scene::clip_paths, core/scene.h:53-130; ClipPath::outline): the last n_clip_paths
     * paths of the arrays above are clip paths, tiled like draw paths but never drawn.  path_clip[p]
     * (NULL: no path is clipped) is the clip path of path p (DrawPath::clip_path, or for a clip path
     * ClipPath::clip_path) as an index among the clip paths, or -1.  Resolution and mask combine
     * follow core/d3d9/tiler.cpp:385-431 and shaders/d3d9/tile_clip_combine.frag:25-27: by default
     * the clip's own fill rule is not used (its coverage enters as |winding sum|; clip_rule below
     * selects ClipPath::fill_rule instead).  Nested clips (undefined in
     * the reference's hybrid tiler, core/d3d9/scene_builder.cpp:209-216) follow the GPU-driven
     * renderer's levels (core/d3d11/scene_builder.cpp:92-133): a clip path is resolved against its
     * own clip path like a draw path, and its coverage is the min over the chain; chains of at most
     * 4 clip paths (a deeper chain or a cycle fails with PFH_E_ARG). */
    int32_t n_clip_paths;
    const int32_t *path_clip;           /* [n_paths] or NULL */
    /* Strokes (Canvas::stroke_path, core/canvas.cpp:272-301): NULL, or per path a stroke style; a
     * path whose line_width is above 0 is stroked on the device at upload (as pfh_stroke_outlines)
     * and its stroke filled with the nonzero rule, whatever path_fill_rule says. */
    const pfh_stroke_style *path_stroke;
    /* The dash lengths that path_stroke's patterns index (NULL when no path is dashed). */
    const float *stroke_dashes;         /* [n_stroke_dashes] */
    int32_t n_stroke_dashes;
    /* Gradient paints: NULL, or per path an index into `gradients` (-1: the solid path_rgba).  A
     * gradient path is opaque for the z-buffer when every stop is (Paint::is_opaque,
     * core/paint/paint.cpp:9-19); its colour per pixel follows the tile shader's colour texture
     * (shaders/d3d9/tile.vert:65-77, tile.frag:296-323, 590-653) over the ramp and metadata
     * pfh_gradient_compile makes. */
    const int32_t *path_gradient;       /* [n_paths] */
    const pfh_gradient *gradients;      /* [n_gradients] */
    int32_t n_gradients;
    const float *gradient_stop_offsets; /* [n_gradient_stops] */
    const uint32_t *gradient_stop_rgba; /* [n_gradient_stops] */
    int32_t n_gradient_stops;
    /* NULL, or [n_gradients][256]: for a gradient whose stop_count is negative, its ramp as the
     * caller sampled it (texel x = Gradient::sample((x + 0.5) / 256), as GradientTileBuilder::allocate
     * does; a binding that cannot read a Gradient's stops passes this); opaque when every texel is. */
    const uint32_t *gradient_ramps;
    /* How a clip path's coverage enters its clip (ClipPath::fill_rule, core/scene.h, which neither of
     * the reference's renderers reads):
     *   PFH_CLIP_RULE_ABS (0, the reference): |winding sum| of each clip level, whatever its
     *     path_fill_rule, the draw path's rule applied to the min (core/d3d9/tiler.cpp:385-431,
     *     shaders/d3d9/tile_clip_combine.frag:25-27, tile.frag);
     *   PFH_CLIP_RULE_OWN (1): every level of the chain under its own path_fill_rule and the draw
     *     path under its own, coverage = the min of those; a tile without fills covers where its
     *     rule says (an even-odd tile with an even backdrop covers nothing). */
    int32_t clip_rule;
    /* Blend modes: NULL (every path PFH_BLEND_SRC_OVER), or per path a PFH_BLEND_* value (a clip
     * path's is not used).  As the reference's hybrid renderer treats DrawPath::blend_mode:
     *   - a destructive mode (clear, copy, src-in, dest-in, src-out, dest-atop) tiles the whole view
     *     box (core/data/built_path.cpp:21, effects.h:234-243);
     *   - only src-over and clear paths of an opaque paint occlude what lies below (built_path.cpp:40,
     *     effects.h:127-160);
     *   - the Porter-Duff modes (clear .. lighter) draw as src-over (the tile pass's blend state is
     *     always over, core/d3d9/renderer.cpp:209);
     *   - the other 15 modes draw through tile.frag's composite() (shaders/d3d9/tile.frag:514-598,
     *     the op of blend_mode_to_composite_ctrl, effects.h:184-232) against the colour composited
     *     below (the reference's hybrid pass binds no destination texture there, so this is the
     *     shader's formula with a defined destination): rgb = a (1 - da) src + a da B(dest, src) +
     *     (1 - a) dest, alpha 1, over every pixel of the path's drawn tiles. */
    const uint8_t *path_blend;          /* [n_paths] or NULL */
} pfh_scene_desc;

typedef struct pfh_stats {
    int64_t paths, segments, tiles, columns, fills, alpha_tiles, entries, fb_tiles;
    int64_t retries;          /* capacity overflows that forced a re-run */
    int64_t guard_hits;       /* loop guards tripped (inputs on which the reference would not terminate) */
    double ms_total;          /* device time of the last pfh_rasterize (HIP events) */
    double ms_stage[8];       /* plain mode: flatten, bin count, bin emit, propagate, entries, sort, fill, composite;
                                 batched mode: [6] fill and [7] composite kernels of the last launch, others 0 */
    int64_t capacity_fills, capacity_alpha, capacity_entries;
    int64_t alpha_entries;    /* draw entries that sample a coverage mask */
    int64_t lines, capacity_lines; /* flattened chords and the chord buffer's capacity */
    int64_t host_syncs;       /* device->host round trips during the last pfh_rasterize */
    int64_t fills_masked, alpha_masked; /* fills and alpha tiles the coverage pass rasterized: those of
                                           alpha tiles that draw entries sample (all of them with PFH_DUMP) */
} pfh_stats;

/* Context on HIP device `device`.  Replaces Renderer's Device/Queue pair (core/renderer.cpp:13-37). */
int pfh_create(int device, pfh_ctx **out);
void pfh_destroy(pfh_ctx *ctx);
const char *pfh_last_error(const pfh_ctx *ctx);

/*
 * Packs a scene (SegmentsIter order, core/data/contour.cpp:110-173; path tile rects,
 * core/data/built_path.cpp:8-34) and makes it resident on the device.  Replaces
 * SceneBuilderD3D11::build's CPU packing (core/d3d11/gpu_data.cpp:79-115,
 * core/d3d11/scene_builder.cpp:26-58).  *scene_id receives a handle.
 */
int pfh_scene_upload(pfh_ctx *ctx, const pfh_scene_desc *desc, int32_t *scene_id);
int pfh_scene_release(pfh_ctx *ctx, int32_t scene_id);

/* Per-path features of an uploaded scene for the sizing predictor: [n_paths * 12] floats. */
int pfh_scene_features(pfh_ctx *ctx, int32_t scene_id, float *out);

/*
 * Rasterizes n scenes into their targets (RendererD3D11::draw, core/d3d11/renderer.cpp:302-336).
 * Without a target set by pfh_scene_set_target the context owns an RGBA8 image per scene.
 */
int pfh_rasterize(pfh_ctx *ctx, const int32_t *scene_ids, int32_t n, uint32_t flags);

/*
 * pfh_rasterize for PFH_PLAIN (optionally with PFH_LOAD_DST) that allocates nothing: no hipMalloc,
 * hipFree, hipHostMalloc, event or stream creation, no allocating library call.  No counterpart in
 * the reference; it is for callers that cannot allocate device memory during a frame (code under
 * HIP graph capture, tight serving loops, harnesses that forbid allocation).  The scenes must be
 * uploaded and their targets set beforehand; the work buffers are carved, scene by scene, from one
 * static device array of PFH_NOALLOC_SCRATCH_BYTES per GPU, which the library holds until the
 * process ends.  A scene whose work does not fit fails with PFH_E_LIMIT (nothing is drawn into its
 * target).  Every call on a GPU shares that GPU's array, across contexts too, and it keeps the last
 * call's contents: calls that overlap on one GPU (two host threads, two contexts) race.
 */
#define PFH_NOALLOC_SCRATCH_BYTES (128ll << 20)
int pfh_rasterize_noalloc(pfh_ctx *ctx, const int32_t *scene_ids, int32_t n, uint32_t flags);

/* Points a scene's output at caller-owned device memory (Canvas::set_dst_texture, core/canvas.cpp:538-540). */
int pfh_scene_set_target(pfh_ctx *ctx, int32_t scene_id, void *device_rgba8, int64_t pitch_bytes);
/* Device address and pitch of a scene's RGBA8 output (premultiplied). */
int pfh_scene_target(pfh_ctx *ctx, int32_t scene_id, void **device_rgba8, int64_t *pitch_bytes);
/* Copies a scene's RGBA8 output to host memory (CommandEncoder::read_texture, gpu/command_encoder.h:228). */
int pfh_readback_rgba8(pfh_ctx *ctx, int32_t scene_id, void *dst, int64_t pitch_bytes);

/*
 * Canonical intermediates of the last PFH_DUMP rasterize for one scene, in the
 * oracle's layout.  counts: tiles, columns, fills, entries, fb tiles, paths, alpha tiles, guard hits.
 */
int pfh_dump_counts(pfh_ctx *ctx, int32_t scene_id, int64_t *counts);
int pfh_dump(pfh_ctx *ctx, int32_t scene_id, int32_t *path_rect, int64_t *tile_off, int64_t *col_off,
             int64_t *fill_off, int8_t *tile_backdrop, uint8_t *tile_alpha, int32_t *col_sum, uint16_t *fill_seg,
             uint32_t *fill_tile, uint32_t *entry_fbtile, uint32_t *entry_path, uint32_t *zbuf, int32_t *fb_rect,
             uint8_t *masks /* alpha tiles in (path, row, column) order, 256 bytes each */);

int pfh_stats_get(pfh_ctx *ctx, pfh_stats *out);

/*
 * Sizing predictor (replaces the count / read back / grow-to-2^k / retry sizing of
 * core/d3d11/renderer.cpp:537-577, 707-724, 834-845 for batched launches): per path, an
 * FP32 MLP 12 -> 64 -> 64 -> 3 predicting log1p of (flattened chords, fills, alpha tiles).
 * weights = mean[12], inv_std[12] (feature normalisation) followed by the 5187 MLP
 * parameters (layout at pfh_mlp_eval).  safety >= 1 scales the predicted capacities.
 * weights == NULL removes the predictor.
 */
int pfh_set_predictor(pfh_ctx *ctx, const float *weights, int64_t n_floats, float safety);
/* Predicted (chords, fills, alpha tiles) of an uploaded scene (-1 without a predictor). */
int pfh_scene_prediction(pfh_ctx *ctx, int32_t scene_id, float *out3);

/*
 * Exact per-path sizes of the last PFH_DUMP rasterize of a scene: flattened chords,
 * fills and alpha tiles — the training targets of the sizing predictor.
 */
int pfh_path_counts(pfh_ctx *ctx, int32_t scene_id, int32_t *lines, int32_t *fills, int32_t *alpha_tiles);

/*
 * The sizing MLP's HIP kernels on host arrays (the context stages them on the device).
 * params: 5187 floats, W1[12][64], b1[64], W2[64][64], b2[64], W3[64][3], b3[3] (W[in][out]);
 * x: n x 12 (normalised features); out / target: n x 3.
 *   pfh_mlp_eval:  out = MLP(x)
 *   pfh_mlp_grad:  gradient of the mean squared error over all n x 3 outputs, and the loss
 *   pfh_mlp_train: `epochs` of Adam over shuffled mini-batches; params updated in place;
 *                  loss_history[epochs] receives the full-data loss after each epoch.
 */
int pfh_mlp_eval(pfh_ctx *ctx, const float *params, const float *x, int64_t n, float *out);
int pfh_mlp_grad(pfh_ctx *ctx, const float *params, const float *x, const float *target, int64_t n, float *grad,
                 float *loss);
int pfh_mlp_train(pfh_ctx *ctx, const float *x, const float *target, int64_t n, int32_t epochs, int32_t batch,
                  float lr, uint64_t seed, float *params, float *loss_history);

/*
 * Launch schedule of the batched pipeline (no counterpart in the reference, whose compute stages
 * run at fixed local sizes, e.g. shaders/d3d11/fill.comp, tile.comp).  Every schedule produces bit-identical results;
 * pfh_autotune picks one for a set of scenes by timing graph replays.
 */
typedef struct pfh_schedule {
    int32_t fill_grid;        /* max workgroups of the coverage-mask kernel (it loops beyond) */
    int32_t composite_block;  /* threads per composite workgroup: 64, 128 or 256 (a wave per 16x16 tile) */
    int32_t column_block;     /* threads per workgroup of the per-column propagate kernel */
    int32_t xcd_order;        /* bit 0 fill, bit 1 composite: workgroups regrouped so that each of the
                                 8 XCDs (own L2) works on one contiguous range of tiles */
    int32_t flatten_segs;     /* segments each flatten wave works through (multiple of 64; 0: from the
                                 segment count) */
    int32_t flatten_take;     /* idle lanes a flatten wave waits for before handing them new segments
                                 (1..64; 0: 8) */
    int32_t column_rows;      /* path columns of more tiles than this are propagated by 16 or 64 lanes
                                 each instead of one (1..64; 0: every column by one lane) */
    int32_t flatten_deep;     /* curves whose subdivision depth is estimated at this or more are
                                 flattened by 16 lanes each instead of one (0..12; 0: none) */
} pfh_schedule;
int pfh_get_schedule(pfh_ctx *ctx, pfh_schedule *out);
int pfh_set_schedule(pfh_ctx *ctx, const pfh_schedule *in);
/* Coordinate descent from the current schedule: each parameter's candidates timed as the median
 * device time of `reps` replays of the batched launch of these scenes; the context keeps the
 * fastest (a change has to gain 1 %).  ms_start / ms_tuned (may be NULL): time of the starting and
 * of the chosen schedule. */
int pfh_autotune(pfh_ctx *ctx, const int32_t *scene_ids, int32_t n, uint32_t flags, int32_t reps, pfh_schedule *out,
                 double *ms_start, double *ms_tuned);

/*
 * Request loop (continuous batching).  No counterpart in the reference, which draws one scene per
 * Canvas::draw (core/canvas.cpp:557-567) with blocking submits per stage (gpu/vk/queue.cpp:29-34).
 * Requests name uploaded (resident) scenes.  Each of `slots` workspaces (own HIP stream, own work
 * buffers) takes the queued requests whose predicted device bytes fit `memory_budget_bytes` beside
 * the launches in flight (at most max_batch per launch, arrival order), launches them as one batched
 * pass and returns at once; requests arriving meanwhile join the next launch of whichever slot frees
 * first, so launches overlap on the device.  pfh_server_submit only queues; the loop advances (retires
 * finished launches, admits queued requests) in pfh_server_poll, which the caller calls in its loop.  A scene is in at most one launch at a time; a request
 * larger than the budget runs alone once nothing else is in flight.  flags: PFH_PREDICT, PFH_NO_GRAPH.
 * Latency = submit to the poll that observes completion.  Results are the scene's target, as after
 * pfh_rasterize.  pfh_scene_release refuses a scene with queued or running requests.
 */
typedef struct pfh_server_stats {
    int64_t submitted, completed, queued, in_flight;
    int64_t launches, retries;
    int64_t max_concurrent_launches;  /* slots seen busy at once */
    double mean_batch;                /* requests per launch */
    double latency_ms_p50, latency_ms_p99, latency_ms_mean, latency_ms_max;
    int64_t reserved_bytes, peak_reserved_bytes, budget_bytes;
    int64_t device_bytes;             /* bytes the slots' work buffers hold */
} pfh_server_stats;

int pfh_server_start(pfh_ctx *ctx, int32_t slots, int64_t memory_budget_bytes, int32_t max_batch, uint32_t flags);
int pfh_server_submit(pfh_ctx *ctx, int32_t scene_id, int64_t *request_id);
/* Advances the loop; waits up to wait_ms for a completion; returns up to `capacity` completed ids. */
int pfh_server_poll(pfh_ctx *ctx, double wait_ms, int64_t *done_ids, int32_t capacity, int32_t *n_done);
int pfh_server_stats_get(pfh_ctx *ctx, pfh_server_stats *out);
/* Runs every queued request to completion, then frees the slots. */
int pfh_server_stop(pfh_ctx *ctx);

/*
 * Stroke-to-fill on the device: the outlines whose nonzero fills are the strokes of `in`'s paths,
 * bit-identical to the reference's OutlineStrokeToFill::offset / into_outline
 * (core/stroke.cpp:124-195), which Canvas::stroke_path runs before a stroke reaches the scene
 * (core/canvas.cpp:272-301); here many paths are stroked at once, one style each.  A style with
 * dash_count > 0 dashes its path first with dashes[dash_start ..] (OutlineDash::dash +
 * into_outline, core/dash.cpp:48-126), bit-identical as well; dashes may be NULL when no style
 * dashes (n_dashes 0).  A pattern far finer than its path (over 2^24 dash steps) fails with
 * PFH_E_LIMIT.  Output
 * capacities in points and contours; on return n_points / n_contours hold the sizes needed and
 * PFH_E_LIMIT means a capacity was too small (call again with them; capacities of 0 ask for the
 * sizes only).  out->path_contour_start [n_paths + 1] receives each path's output contours.
 */
typedef struct pfh_outline_buf {
    float *points;               /* [2 * cap_points] */
    uint8_t *point_flags;        /* [cap_points] */
    int64_t cap_points;
    int32_t *contour_start;      /* [cap_contours + 1] */
    uint8_t *contour_closed;     /* [cap_contours] */
    int64_t cap_contours;
    int32_t *path_contour_start; /* [n_paths + 1] */
    int64_t n_points, n_contours;
} pfh_outline_buf;

/*
 * A gradient as the library draws it, on the host (no context): its ramp, the RAMP_TEXELS = 256
 * RGBA8 texels GradientTileBuilder::allocate samples (core/paint/gradient.cpp:86-123 with
 * Gradient::sample, :23-56), and its 16-float record: the colour texture transform (m11, m21, m12,
 * m22, m13, m23) and the radial filter's line from x, y, vector x, y, radii, uv origin x, y, as the
 * RGBA16F metadata texture holds them (Palette::calculate_texture_transforms,
 * core/paint/palette.cpp:398-440; PaintMetadata::filter, core/paint/paint.cpp:45-65;
 * Renderer::upload_texture_metadata, core/renderer.cpp:167-250, through third_party/umHalf's
 * truncating float-to-half), then kind (1 linear, 2 radial) and wrap as floats.  ramp_row: the
 * gradient's row in its 256-row texture page (the v coordinate).  PFH_E_ARG for a bad kind or
 * wrap, non-finite geometry or stop offsets, or a radial transform that cannot be inverted.
 */
int pfh_gradient_compile(const pfh_gradient *g, const float *stop_offsets, const uint32_t *stop_rgba,
                         int32_t ramp_row, uint32_t *ramp, float *record);

int pfh_stroke_outlines(pfh_ctx *ctx, const pfh_outline *in, const int32_t *path_contour_start, int32_t n_paths,
                        const pfh_stroke_style *styles, const float *dashes, int32_t n_dashes, pfh_outline_buf *out);

#ifdef __cplusplus
}
#endif

#endif /* PFH_H */
