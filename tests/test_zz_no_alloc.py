"""vt_render_frame_noalloc through the C ABI: the per-ray pass state in the
library's file-scope device array instead of context buffers, for callers
that cannot allocate.  Two calls in a row (the second frame smaller) give
the same bytes as vt_render_frame on the same scene and cache state, and
sizes beyond the array's maximum are refused before any device work."""
import ctypes

import numpy as np
import pytest

from gpu_scenes import Scene, blobs_volume, orbit_camera

pytestmark = pytest.mark.gpu


def _noalloc(vt, ctx, cam, prm, d_rgba, d_hits):
    st = vt.FrameStats()
    rc = vt.lib().vt_render_frame_noalloc(ctx.handle, ctypes.byref(cam), ctypes.byref(prm), d_rgba.ptr,
                                          d_hits.ptr if d_hits is not None else None, ctypes.byref(st))
    return rc, st


def test_noalloc_frames_equal_render_frame(vt, vto, ctx):
    vol = blobs_volume(64, 12, seed=11)
    sc = Scene(vto, vol, 0.25, 3, 2)
    sc.setup_gpu(vt, ctx, slots=6)   # 64 bricks through 6 slots: several passes, loads between them
    sizes = [(200, 150), (72, 40)]
    try:
        for k, (W, H) in enumerate(sizes):
            cam = orbit_camera(vt, W, H, (32, 32, 32), 100.0, 30.0 + 50.0 * k, 15.0)
            R = W * H
            print(f"noalloc frame {k}: {W}x{H}", flush=True)
            ref_rgba, ref_hits = ctx.alloc(R * 16), ctx.alloc(R * 32)
            ctx.cache_clear()
            st_ref = ctx.render_frame(cam, sc.prm_g, ref_rgba, ref_hits)
            d_rgba, d_hits = ctx.alloc(R * 16), ctx.alloc(R * 32)
            ctx.cache_clear()
            rc, st = _noalloc(vt, ctx, cam, sc.prm_g, d_rgba, d_hits)
            assert rc == 0, vt.last_error() if hasattr(vt, "last_error") else rc
            assert st.passes == st_ref.passes and st.passes >= 2
            assert st.bricks_uploaded == st_ref.bricks_uploaded
            got = (d_hits.download(vt.HIT_DTYPE, (R,)), d_rgba.download(np.float32, (R, 4)))
            ref = (ref_hits.download(vt.HIT_DTYPE, (R,)), ref_rgba.download(np.float32, (R, 4)))
            assert got[0].tobytes() == ref[0].tobytes()
            assert got[1].tobytes() == ref[1].tobytes()
            assert (got[0]["node"] != 0).mean() > 0.05
            for b in (ref_rgba, ref_hits, d_rgba, d_hits):
                b.free()
        # refused before any device work: too many rays, no hit buffer,
        # deferral lists longer than the array holds, trace recording on
        big = vt.make_camera(2048, 1025, (32.0, 32.0, 200.0))
        small = orbit_camera(vt, 16, 16, (32, 32, 32), 100.0, 0.0)
        d_rgba, d_hits = ctx.alloc(16 * 16 * 16), ctx.alloc(16 * 16 * 32)
        assert _noalloc(vt, ctx, big, sc.prm_g, d_rgba, d_hits)[0] == -1
        assert _noalloc(vt, ctx, small, sc.prm_g, d_rgba, None)[0] == -1
        ctx.set_march_options(skip_empty=True, max_defer=3, max_defer_late=2)
        assert _noalloc(vt, ctx, small, sc.prm_g, d_rgba, d_hits)[0] == -1
        ctx.set_march_options()
        ctx.set_frame_options(False, True)
        assert _noalloc(vt, ctx, small, sc.prm_g, d_rgba, d_hits)[0] == -5
        ctx.set_frame_options(False, False)
        assert _noalloc(vt, ctx, small, sc.prm_g, d_rgba, d_hits)[0] == 0
        d_rgba.free()
        d_hits.free()
    finally:
        ctx.set_march_options()
        ctx.set_frame_options(False, False)
        ctx.cache_destroy()
