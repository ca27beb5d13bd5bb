"""Staged rasterize backward (``jaxsplat_rasterize_backward_staged``), on the GPU.

The staged backward writes each (Gaussian, tile) pair's partial sums to the
pair's intersection slot in the reference's emission order and sums them per
Gaussian, instead of adding them with float atomics.  Bars: every gradient
within |g - g_a| <= 1e-5 |g_a| + 1e-6 max|g_a| of the atomic backward on the
same inputs (the two differ only in the order of float additions); two runs
bitwise identical; every output row written (rows of Gaussians that no pixel
reaches are exactly zero, also where the output buffer held garbage).  The
float64 oracle comparison of the render path, which uses this backward
through torch autograd, is in test_gpu_modes (small scenes, every gradient,
both adjoints) and test_gpu_render_fullsize (the 1M- and 3M-Gaussian
workloads, tile subsets); test_gpu_parity and test_gpu_fullsize check the
atomic backward.
"""

import ctypes

import numpy as np
import pytest

from jaxsplat_amd import _lib, scenes
from tests import helpers as H

pytestmark = pytest.mark.gpu


def _case(nch, n=4000, seed=3, bw=16):
    sc = scenes.make("small", n=n)
    cam = sc.camera
    kw = dict(img_shape=cam.img_shape, f=cam.f, c=cam.c, glob_scale=sc.glob_scale,
              clip_thresh=sc.clip_thresh, block_width=bw)
    g = H.gpu_project(sc.means3d, sc.scales, sc.quats, cam.viewmat, **kw)
    gids, _, bins, isect = H.gpu_bin(g["xys"], g["depths"], g["radii"], g["cum_tiles_hit"],
                                     img_shape=cam.img_shape, block_width=bw)
    rng = np.random.default_rng(seed)
    N = sc.means3d.shape[0]
    colors = rng.uniform(0, 1, (N, nch)).astype(np.float32)
    bg = rng.uniform(0, 1, nch).astype(np.float32)
    return sc, cam, g, gids, bins, isect, colors, bg


def _run(kind, sc, cam, g, gids, bins, isect, colors, bg, v_img, adjoint, with_abs, bw=16):
    t = H.torch()
    N, nch = colors.shape
    H_, W = cam.img_shape
    d = H.descriptor(N, cam.img_shape, block_width=bw)
    L = _lib.lib()
    dg = H.dev(np.asarray(gids, np.int32))
    db = H.dev(np.asarray(bins, np.int32))
    xy = H.dev(np.asarray(g["xys"], np.float32).reshape(N, 2))
    cn = H.dev(np.asarray(g["conics"], np.float32).reshape(N, 3))
    col = H.dev(colors)
    op = H.dev(np.asarray(sc.opacities, np.float32))
    bgd = H.dev(bg)
    fT = t.empty((H_, W), device="cuda")
    fI = t.empty((H_, W), dtype=t.int32, device="cuda")
    img = t.empty((H_, W, nch), device="cuda")
    _lib.check(L.jaxsplat_rasterize_forward_nd(ctypes.byref(d), H.P(dg), H.P(db), isect, H.P(xy),
                                               H.P(cn), H.P(col), H.P(op), H.P(bgd), nch,
                                               H.P(fT), H.P(fI), H.P(img), H.stream()))
    vi = H.dev(v_img)
    outs = dict(v_colors=t.full((N, nch), 7.0, device="cuda"),
                v_opacity=t.full((N, 1), 7.0, device="cuda"),
                v_xy=t.full((N, 2), 7.0, device="cuda"),
                v_xy_abs=t.full((N, 2), 7.0, device="cuda") if with_abs else None,
                v_conic=t.full((N, 3), 7.0, device="cuda"))
    P = [H.P(outs[k]) for k in ("v_colors", "v_opacity", "v_xy", "v_xy_abs", "v_conic")]
    if kind == "staged":
        r = H.dev(np.asarray(g["radii"], np.int32).reshape(N))
        cum = H.dev(np.asarray(g["cum_tiles_hit"], np.int32).reshape(N))
        wsb = L.jaxsplat_rasterize_backward_staged_workspace_size(ctypes.byref(d), isect)
        ws = t.full((max(wsb, 1),), 0xA5, dtype=t.uint8, device="cuda")
        _lib.check(L.jaxsplat_rasterize_backward_staged(
            ctypes.byref(d), H.P(dg), H.P(db), isect, H.P(xy), H.P(cn), H.P(col), H.P(op),
            H.P(bgd), nch, H.P(r), H.P(cum), H.P(fT), H.P(fI), H.P(vi), None, *P, adjoint,
            H.P(ws), wsb, H.stream()))
    else:
        _lib.check(L.jaxsplat_rasterize_backward_nd(
            ctypes.byref(d), H.P(dg), H.P(db), isect, H.P(xy), H.P(cn), H.P(col), H.P(op),
            H.P(bgd), nch, H.P(fT), H.P(fI), H.P(vi), None, *P, adjoint, H.stream()))
    return {k: H.host(v) for k, v in outs.items() if v is not None}


@pytest.mark.parametrize("nch,adjoint,with_abs", [(3, 0, False), (3, 1, True), (1, 0, True),
                                                  (2, 1, False)])
def test_staged_matches_atomic_backward(nch, adjoint, with_abs):
    case = _case(nch)
    sc, cam = case[0], case[1]
    rng = np.random.default_rng(11)
    v_img = rng.normal(size=(*cam.img_shape, nch)).astype(np.float32)
    a = _run("atomic", *case, v_img, adjoint, with_abs)
    s = _run("staged", *case, v_img, adjoint, with_abs)
    for k in a:
        ref = a[k].astype(np.float64)
        tol = 1e-5 * np.abs(ref) + 1e-6 * np.abs(ref).max()
        assert np.all(np.abs(s[k] - ref) <= tol), k


@pytest.mark.parametrize("bw", [12, 13])
def test_staged_matches_atomic_backward_partial_quadrants(bw):
    # tiles of 12 and 13 pixels: the four-slot quadrant layout with lanes
    # past the tile owning no pixel, in both the staged and the atomic form
    case = _case(3, bw=bw)
    sc, cam = case[0], case[1]
    rng = np.random.default_rng(13)
    v_img = rng.normal(size=(*cam.img_shape, 3)).astype(np.float32)
    a = _run("atomic", *case, v_img, 0, False, bw=bw)
    s = _run("staged", *case, v_img, 0, False, bw=bw)
    assert np.abs(a["v_colors"]).max() > 0
    for k in a:
        ref = a[k].astype(np.float64)
        tol = 1e-5 * np.abs(ref) + 1e-6 * np.abs(ref).max()
        assert np.all(np.abs(s[k] - ref) <= tol), k


def test_staged_intersection_count_not_multiple_of_four():
    """Regression: the per-slot flags of a workspace left dirty by an earlier
    call (here 0xA5) must all be cleared, also when the intersection count is
    not a multiple of four."""
    for n in range(1500, 1540):
        case = _case(3, n=n, seed=8)
        if case[5] % 4:
            break
    assert case[5] % 4
    cam = case[1]
    rng = np.random.default_rng(13)
    v_img = rng.normal(size=(*cam.img_shape, 3)).astype(np.float32)
    a = _run("atomic", *case, v_img, 0, False)
    s = _run("staged", *case, v_img, 0, False)
    for k in a:
        ref = a[k].astype(np.float64)
        tol = 1e-5 * np.abs(ref) + 1e-6 * np.abs(ref).max()
        assert np.all(np.abs(s[k] - ref) <= tol), k


def test_staged_is_reproducible_and_writes_every_row():
    case = _case(3, n=6000, seed=4)
    sc, cam, g = case[0], case[1], case[2]
    rng = np.random.default_rng(12)
    v_img = rng.normal(size=(*cam.img_shape, 3)).astype(np.float32)
    r1 = _run("staged", *case, v_img, 0, True)
    r2 = _run("staged", *case, v_img, 0, True)
    for k in r1:
        assert np.array_equal(r1[k].view(np.uint32), r2[k].view(np.uint32)), k
    # Gaussians outside the view (radius 0) get exact zeros over the 7.0 fill
    off = np.asarray(g["radii"]).reshape(-1) == 0
    assert off.any()
    for k in r1:
        assert np.all(r1[k][off] == 0.0), k


def test_staged_inconsistent_counts_do_not_write_out_of_range():
    """cum_tiles_hit that disagrees with the binning (here shifted by one slot)
    must not be written through: the records land in other slots or none, and
    the call completes."""
    t = H.torch()
    sc, cam, g, gids, bins, isect, colors, bg = _case(3, n=2000, seed=6)
    g = dict(g)
    g["cum_tiles_hit"] = (np.asarray(g["cum_tiles_hit"]) + 1).astype(np.int32)
    v_img = np.ones((*cam.img_shape, 3), np.float32)
    out = _run("staged", sc, cam, g, gids, bins, isect, colors, bg, v_img, 0, False)
    for k, v in out.items():
        assert np.all(np.isfinite(v)), k
