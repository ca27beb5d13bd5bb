"""Opacity-limited tile rectangles (``jaxsplat_tile_rects``), on the GPU.

``rasterize`` bins each Gaussian into the tiles of its reference rectangle
where the per-pixel test o exp(-sigma) >= 1/255 can pass (bounding box of that
ellipse, with margins).  Bars:

* the rectangles lie inside the reference's (``lib/kernels/helpers.h:9-32``),
  the counts are their areas and ``cum_tiles_hit`` their inclusive sums;
  Gaussians with radius 0 or 255 o < 1 get none;
* rasterizing through these lists and through the reference's lists gives
  bit-identical images (both adjoint modes, 1 and 3 channels, the antialiased
  mode): a dropped entry never passes the per-pixel test, which leaves a
  pixel's state exactly unchanged.  The rasterizer's gradients (colors,
  opacities, xys, conics) match within |g - g_ref| <= 1e-5 |g_ref| +
  1e-7 max|g_ref|: each (Gaussian, tile) sum is the same, but a Gaussian whose
  list of tiles shrinks below 32 is summed by its own lane instead of the
  wave, in another order.  (Through the projection backward this rounding is
  amplified by ill-conditioned covariances, so the parameters' gradients are
  compared at the usual parity bars against the float64 oracle elsewhere.)  The scenes include
  strongly elongated Gaussians and opacities near 1/255, where the ellipse's
  bounding box is tightest and its margins matter.
"""

import ctypes

import numpy as np
import pytest

from jaxsplat_amd import _lib, _ops, project, rasterize, render, scenes
from tests import helpers as H

pytestmark = pytest.mark.gpu


def _stress_scene(seed, n=3000, hw=(200, 312)):
    """Elongated Gaussians (scale ratios up to ~60) at random rotations, with
    opacities spread over [1e-3, 1] on a log scale (many near 1/255)."""
    rng = np.random.default_rng(seed)
    means = np.stack([rng.uniform(-2.5, 2.5, n), rng.uniform(-1.6, 1.6, n),
                      rng.uniform(3.0, 8.0, n)], axis=1).astype(np.float32)
    base = rng.normal(np.log(0.03), 0.4, n)
    stretch = rng.uniform(0.0, np.log(60.0), n)
    scales = np.exp(np.stack([base + stretch, base, base - 0.3 * stretch], axis=1))
    q = rng.normal(size=(n, 4))
    q /= np.linalg.norm(q, axis=1, keepdims=True)
    colors = rng.uniform(0, 1, (n, 3)).astype(np.float32)
    opac = np.exp(rng.uniform(np.log(1e-3), 0.0, (n, 1))).astype(np.float32)
    H_, W = hw
    cam = dict(viewmat=scenes.IDENTITY.copy(), img_shape=hw, f=(float(W), float(W)),
               c=(W / 2.0, H_ / 2.0))
    return [means, scales.astype(np.float32), q.astype(np.float32), colors, opac], cam


def _ref_rects(xys, radii, img_shape, bw):
    H_, W = img_shape
    tx, ty = (W + bw - 1) // bw, (H_ + bw - 1) // bw
    c = xys.astype(np.float32) / np.float32(bw)
    r = radii.reshape(-1).astype(np.float32) / np.float32(bw)
    x0 = np.clip((c[:, 0] - r).astype(np.int64), 0, tx)
    x1 = np.clip((c[:, 0] + r + np.float32(1)).astype(np.int64), 0, tx)
    y0 = np.clip((c[:, 1] - r).astype(np.int64), 0, ty)
    y1 = np.clip((c[:, 1] + r + np.float32(1)).astype(np.int64), 0, ty)
    return x0, x1, y0, y1


@pytest.mark.parametrize("seed", [1, 2])
def test_tile_rects_inside_reference_rects_with_inclusive_counts(seed):
    t = H.torch()
    params, cam = _stress_scene(seed)
    bw = 16
    g = H.gpu_project(*params[:3], cam["viewmat"], img_shape=cam["img_shape"], f=cam["f"],
                      c=cam["c"], block_width=bw)
    N = params[0].shape[0]
    d = H.descriptor(N, cam["img_shape"], block_width=bw)
    L = _lib.lib()
    xys, radii = H.dev(g["xys"]), H.dev(np.asarray(g["radii"], np.int32).reshape(N))
    conics, opac = H.dev(g["conics"]), H.dev(params[4].reshape(N))
    rects = t.full((N, 2), -1, dtype=t.int32, device="cuda")
    nth = t.full((N,), -1, dtype=t.int32, device="cuda")
    cum = t.full((N,), -1, dtype=t.int32, device="cuda")
    wsb = L.jaxsplat_tile_rects_workspace_size(ctypes.byref(d))
    ws = t.empty(max(wsb, 1), dtype=t.uint8, device="cuda")
    _lib.check(L.jaxsplat_tile_rects(ctypes.byref(d), H.P(xys), H.P(radii), H.P(conics),
                                     H.P(opac), H.P(rects), H.P(nth), H.P(cum), H.P(ws), wsb,
                                     H.stream()))
    rc = H.host(rects).view(np.uint32).astype(np.int64)
    x0, y0 = rc[:, 0] & 0xFFFF, rc[:, 0] >> 16
    w, h = rc[:, 1] & 0xFFFF, rc[:, 1] >> 16
    n_h, c_h = H.host(nth), H.host(cum)
    assert np.array_equal(n_h, w * h)
    assert np.array_equal(c_h, np.cumsum(w * h))
    rx0, rx1, ry0, ry1 = _ref_rects(g["xys"], g["radii"], cam["img_shape"], bw)
    ref_n = np.asarray(g["num_tiles_hit"]).reshape(-1)
    live = (w * h) > 0
    assert np.all((x0 >= rx0) & (x0 + w <= rx1) & (y0 >= ry0) & (y0 + h <= ry1) | ~live)
    assert np.all(n_h <= ref_n)
    off = np.asarray(g["radii"]).reshape(-1) <= 0
    assert np.all(n_h[off] == 0)
    assert np.all(n_h[params[4].reshape(-1) * 255.0 < 0.999] == 0)
    # the lists are shorter on this scene
    assert n_h.sum() < 0.9 * ref_n.sum()


def _raster_both(params, cam, *, adjoint="reference", antialiased=False, nch=3, seed=0):
    """Project once, then rasterize through the opacity-limited lists and
    through the reference's lists, with xys, conics, colors and opacities as
    leaves: the image and the rasterizer's own gradients of both."""
    t = H.torch()
    rng = np.random.default_rng(seed)
    colors = params[3] if nch == 3 else rng.uniform(0, 1, (params[0].shape[0], nch)).astype(
        np.float32)
    bg = H.dev(rng.uniform(0, 1, nch).astype(np.float32))
    with t.no_grad():
        out = project(*[H.dev(x) for x in params[:3]], H.dev(cam["viewmat"]),
                      img_shape=cam["img_shape"], f=cam["f"], c=cam["c"], glob_scale=1.0,
                      clip_thresh=0.01, block_width=16, return_compensation=antialiased)
    xys, depths, radii, conics, _, cum = out[:6]
    opac = H.dev(params[4])
    if antialiased:
        opac = opac * out[6]
    w = None
    res = []
    for tight in (True, False):
        _ops.TIGHT_TILES = tight
        try:
            leaves = [x.detach().clone().requires_grad_(True)
                      for x in (H.dev(colors), opac, xys, conics)]
            img = rasterize(leaves[0], leaves[1], bg, leaves[2], depths, radii, leaves[3], cum,
                            img_shape=cam["img_shape"], block_width=16, adjoint=adjoint)
            if w is None:
                w = H.dev(rng.normal(size=img.shape).astype(np.float32))
            (img * w).sum().backward()
            res.append((H.host(img), [H.host(x.grad) for x in leaves]))
        finally:
            _ops.TIGHT_TILES = True
    return res


def _grads_match(got, ref):
    # the same per-(Gaussian, tile) sums; only their order of addition can
    # differ (lane or wave sum), so the bar is float32 rounding of the sums
    res = []
    for k, (a, b) in enumerate(zip(got, ref)):
        ok, worst = H.grad_close(a, b, rtol=1e-5, scale_frac=1e-7)
        res.append((k, ok, round(worst, 3)))
    assert all(r[1] for r in res), res


@pytest.mark.parametrize("seed,adjoint,antialiased,nch", [
    (3, "reference", False, 3), (4, "consistent", False, 3), (5, "reference", True, 3),
    (6, "reference", False, 1)])
def test_tight_lists_render_the_same(seed, adjoint, antialiased, nch):
    params, cam = _stress_scene(seed)
    (img_t, g_t), (img_r, g_r) = _raster_both(params, cam, adjoint=adjoint,
                                              antialiased=antialiased, nch=nch, seed=seed)
    assert np.array_equal(img_t.view(np.uint32), img_r.view(np.uint32))
    _grads_match(g_t, g_r)


def test_tight_lists_render_the_same_on_the_small_workload():
    sc = scenes.make("small", n=6000)
    cam = dict(viewmat=sc.camera.viewmat, img_shape=sc.camera.img_shape, f=sc.camera.f,
               c=sc.camera.c)
    (img_t, g_t), (img_r, g_r) = _raster_both(sc.params(), cam, seed=9)
    assert np.array_equal(img_t.view(np.uint32), img_r.view(np.uint32))
    _grads_match(g_t, g_r)


@pytest.mark.parametrize("case", ["behind_camera", "transparent", "one_gaussian"])
def test_tight_lists_edge_cases(case):
    """No intersection at all (every Gaussian behind the camera), none in the
    opacity-limited lists while the reference's lists are not empty (every
    opacity below 1/255), and a single Gaussian: the image is the same as
    through the reference's lists, and the gradients too (zero when nothing
    blends)."""
    params, cam = _stress_scene(11, n=400)
    if case == "behind_camera":
        params[0] = params[0].copy()
        params[0][:, 2] = -5.0
    elif case == "transparent":
        params[4] = np.full_like(params[4], 0.5 / 255.0)
    else:
        params = [p[:1].copy() for p in params]
        params[0][0] = (0.0, 0.0, 4.0)
        params[4][0] = 0.9
    (img_t, g_t), (img_r, g_r) = _raster_both(params, cam, seed=12)
    assert np.array_equal(img_t.view(np.uint32), img_r.view(np.uint32))
    _grads_match(g_t, g_r)
    if case != "one_gaussian":
        assert all(not g.any() for g in g_t)


@pytest.mark.parametrize("antialiased", [False, True])
def test_fused_render_matches_reference_lists(antialiased):
    """render() projects once with the opacity-limited rectangles
    (project_for_render) and rasterizes through them: the image is
    bit-identical to rendering through the reference's lists, and the
    colour and opacity gradients (sums of the same per-tile records) match at
    the rasterizer's bar."""
    params, cam = _stress_scene(21)
    rng = np.random.default_rng(22)
    bg = H.dev(rng.uniform(0, 1, 3).astype(np.float32))
    w = None
    res = []
    for tight in (True, False):
        _ops.TIGHT_TILES = tight
        try:
            d = [H.dev(x).requires_grad_(True) for x in params]
            img = render(*d, viewmat=H.dev(cam["viewmat"]), background=bg,
                         img_shape=cam["img_shape"], f=cam["f"], c=cam["c"], glob_scale=1.0,
                         clip_thresh=0.01, block_size=16, antialiased=antialiased)
            if w is None:
                w = H.dev(rng.normal(size=img.shape).astype(np.float32))
            (img * w).sum().backward()
            res.append((H.host(img), [H.host(d[3].grad), H.host(d[4].grad)]))
        finally:
            _ops.TIGHT_TILES = True
    assert np.array_equal(res[0][0].view(np.uint32), res[1][0].view(np.uint32))
    if not antialiased:
        _grads_match(res[0][1], res[1][1])


def test_render_without_gaussians():
    """N = 0 through the fused path and through the reference's lists: the
    background everywhere, and empty gradients."""
    t = H.torch()
    cam = dict(viewmat=scenes.IDENTITY.copy(), img_shape=(40, 56), f=(56.0, 56.0), c=(28.0, 20.0))
    bg = np.array([0.2, 0.4, 0.6], np.float32)
    for tight in (True, False):
        _ops.TIGHT_TILES = tight
        try:
            d = [H.dev(np.zeros(s, np.float32)).requires_grad_(True)
                 for s in ((0, 3), (0, 3), (0, 4), (0, 3), (0, 1))]
            img = render(*d, viewmat=H.dev(cam["viewmat"]), background=H.dev(bg),
                         img_shape=cam["img_shape"], f=cam["f"], c=cam["c"], glob_scale=1.0,
                         clip_thresh=0.01, block_size=16)
            assert np.array_equal(H.host(img), np.broadcast_to(bg, (40, 56, 3)))
            img.sum().backward()
            assert all(x.grad is not None and x.grad.shape[0] == 0 for x in d)
        finally:
            _ops.TIGHT_TILES = True
