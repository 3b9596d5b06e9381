"""Launch schedules and the autotuner: every schedule gives bit-identical framebuffers."""
import numpy as np
import pytest

import pfh
import predictor
import scenegen

pytestmark = pytest.mark.gpu


def test_every_schedule_gives_identical_pixels():
    ctx = pfh.Context(0)
    scenes = scenegen.batch_scenes(n_scenes=5, seed=808, side_max=1536, paths_max=2000)
    ids = [ctx.upload(s) for s in scenes]
    assert ctx.schedule() == {"fill_grid": 16384, "composite_block": 256, "column_block": 256, "xcd_order": 0,
                              "flatten_segs": 0, "flatten_take": 8, "column_rows": 0, "flatten_deep": 0}
    ctx.rasterize(ids, pfh.PFH_BATCHED)
    want = [ctx.readback(i) for i in ids]
    for sched in ({"fill_grid": 1}, {"fill_grid": 1024, "xcd_order": 1}, {"composite_block": 64, "xcd_order": 2},
                  {"composite_block": 128, "column_block": 64, "xcd_order": 3}, {"column_block": 128},
                  {"flatten_segs": 64}, {"flatten_segs": 2048}, {"flatten_take": 1},
                  {"flatten_take": 64, "flatten_segs": 128}, {"column_rows": 1}, {"column_rows": 16}, {"column_rows": 64},
                  {"column_rows": 8, "column_block": 64}, {"flatten_deep": 1}, {"flatten_deep": 4},
                  {"flatten_deep": 5, "flatten_segs": 64, "flatten_take": 1}):
        ctx.set_schedule(**sched)
        for flags in (pfh.PFH_BATCHED, pfh.PFH_BATCHED | pfh.PFH_NO_GRAPH):
            ctx.rasterize(ids, flags)
            for i, img in zip(ids, want):
                np.testing.assert_array_equal(ctx.readback(i), img, err_msg=str(sched))
        ctx.rasterize([ids[0]], pfh.PFH_PLAIN)
        np.testing.assert_array_equal(ctx.readback(ids[0]), want[0], err_msg=f"plain {sched}")
    with pytest.raises(pfh.PfhError):
        ctx.set_schedule(composite_block=96)
    with pytest.raises(pfh.PfhError):
        ctx.set_schedule(flatten_segs=100)
    with pytest.raises(pfh.PfhError):
        ctx.set_schedule(flatten_take=65)
    with pytest.raises(pfh.PfhError):
        ctx.set_schedule(column_rows=65)
    with pytest.raises(pfh.PfhError):
        ctx.set_schedule(flatten_deep=13)
    ctx.close()


def test_autotune_keeps_results_and_never_slows_down():
    ctx = pfh.Context(0)
    w = predictor.load()
    ctx.set_predictor(w, safety=1.25)
    scenes = scenegen.batch_scenes(n_scenes=6, seed=909, side_max=2048, paths_max=4000)
    ids = [ctx.upload(s) for s in scenes]
    ctx.rasterize(ids, pfh.PFH_BATCHED | pfh.PFH_PREDICT)
    want = [ctx.readback(i) for i in ids]
    sched, ms0, ms1 = ctx.autotune(ids, pfh.PFH_PREDICT, reps=3)
    print("[autotune]", sched, f"{ms0:.3f} ms -> {ms1:.3f} ms", flush=True)
    assert ms1 <= ms0 and ctx.schedule() == sched
    ctx.rasterize(ids, pfh.PFH_BATCHED | pfh.PFH_PREDICT)
    for i, img in zip(ids, want):
        np.testing.assert_array_equal(ctx.readback(i), img)
    ctx.close()
