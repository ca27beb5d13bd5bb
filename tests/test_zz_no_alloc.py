"""Allocation-free projection forward (``jaxsplat_project_forward_noalloc``) on the GPU.

The entry point runs the kernels of ``jaxsplat_project_forward`` with the
scan's scratch in a library-held device array instead of a caller workspace.
Bar: all eight outputs bit-identical to ``jaxsplat_project_forward`` on the
same inputs, over two calls in a row, the second with fewer Gaussians (the
array keeps the first call's contents); sizes above 2^26 Gaussians are
rejected with JAXSPLAT_ERR_CAPACITY before any launch.
"""

import ctypes

import numpy as np
import pytest

from jaxsplat_amd import _lib, scenes
from tests import helpers as H

pytestmark = pytest.mark.gpu

OUTS = ("cov3d", "xys", "depths", "radii", "conics", "compensation", "num_tiles_hit",
        "cum_tiles_hit")


def _outputs(t, N, fill):
    f32 = dict(device="cuda", dtype=t.float32)
    i32 = dict(device="cuda", dtype=t.int32)
    return dict(cov3d=t.full((N, 6), fill, **f32), xys=t.full((N, 2), fill, **f32),
                depths=t.full((N,), fill, **f32), radii=t.full((N,), int(fill), **i32),
                conics=t.full((N, 3), fill, **f32), compensation=t.full((N,), fill, **f32),
                num_tiles_hit=t.full((N,), int(fill), **i32),
                cum_tiles_hit=t.full((N,), int(fill), **i32))


def _run_both(n, seed):
    t = H.torch()
    L = _lib.lib()
    sc = scenes.make("hd1m", n=n, seed=seed)
    cam = sc.camera
    d = H.descriptor(n, cam.img_shape, cam.f, cam.c, sc.glob_scale, sc.clip_thresh,
                     sc.block_width)
    ins = [H.dev(np.asarray(a, np.float32)) for a in (sc.means3d, sc.scales, sc.quats,
                                                      cam.viewmat)]
    ref = _outputs(t, n, 7.0)
    wsb = L.jaxsplat_project_workspace_size(ctypes.byref(d))
    ws = t.empty(wsb, dtype=t.uint8, device="cuda")
    _lib.check(L.jaxsplat_project_forward(ctypes.byref(d), *[H.P(x) for x in ins],
                                          *[H.P(ref[k]) for k in OUTS], H.P(ws), wsb,
                                          H.stream()))
    got = _outputs(t, n, -3.0)
    _lib.check(L.jaxsplat_project_forward_noalloc(ctypes.byref(d), *[H.P(x) for x in ins],
                                                  *[H.P(got[k]) for k in OUTS], H.stream()))
    t.cuda.synchronize()
    return {k: H.host(v) for k, v in ref.items()}, {k: H.host(v) for k, v in got.items()}


def test_noalloc_projection_matches_workspace_entry_point():
    # 300k Gaussians (293 scan partitions), then 5,000 (5 partitions): the
    # second call reuses the array the first one filled
    for n, seed in ((300_000, 11), (5_000, 12)):
        print(f"noalloc projection: n={n} seed={seed}", flush=True)
        ref, got = _run_both(n, seed)
        for k in OUTS:
            assert got[k].dtype == ref[k].dtype, k
            assert np.array_equal(got[k].view(np.uint32), ref[k].view(np.uint32)), (n, k)
        assert int(got["cum_tiles_hit"][-1]) > 0


def test_noalloc_projection_rejects_oversize():
    L = _lib.lib()
    d = H.descriptor((1 << 26) + 1, (64, 64), (64.0, 64.0), (32.0, 32.0))
    dummy = ctypes.c_void_p(256)     # never dereferenced: the size check comes first
    code = L.jaxsplat_project_forward_noalloc(ctypes.byref(d), *([dummy] * 12), H.stream())
    assert code == _lib.ERR_CAPACITY
    assert b"at most" in L.jaxsplat_last_error()
