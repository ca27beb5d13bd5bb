"""GPU parity: every stage of the HIP path through the C ABI vs the oracle.

Bars (written in each test):
* projection: bit-exact against the float32 restatement (same operation
  order, -ffp-contract=off); float64 agreement within 1e-5 relative.
* counts, cumulative counts, sorted ids/keys, tile bins: bit-exact.
* image: max-abs error <= 1e-4 vs the float64 oracle on the same float32
  inputs, excluding (and bounding the number of) pixels whose discrete
  decisions are ambiguous in float32 (oracle ``ambiguous`` mask).
* gradients: element-wise |g - g64| <= 1e-3 |g64| + 1e-4 max|g64| with the
  cotangent of ambiguous pixels set to zero for both sides.
"""

import numpy as np
import pytest

import oracle
from jaxsplat_amd import scenes
from tests import helpers as H

pytestmark = pytest.mark.gpu


def _setup(name="small", n=None, block_width=16, **over):
    sc = scenes.make(name, n=n)
    cam = sc.camera
    kw = dict(img_shape=cam.img_shape, f=cam.f, c=cam.c, glob_scale=sc.glob_scale,
              clip_thresh=sc.clip_thresh, block_width=block_width)
    kw.update(over)
    return sc, cam, kw


def _project_both(sc, cam, kw):
    g = H.gpu_project(sc.means3d, sc.scales, sc.quats, cam.viewmat, **kw)
    o32 = oracle.project.project_fwd(sc.means3d, sc.scales, sc.quats, cam.viewmat,
                                     dtype=np.float32, **kw)
    return g, o32


@pytest.mark.parametrize("bw", [16, 8, 5])
def test_project_fwd_bitexact(bw):
    sc, cam, kw = _setup(block_width=bw)
    g, o32 = _project_both(sc, cam, kw)
    for k in ("cov3d", "xys", "depths", "conics", "compensation"):
        a = g[k].reshape(o32[k].shape)
        assert np.array_equal(a.view(np.uint32), o32[k].astype(np.float32).view(np.uint32)), k
    for k in ("radii", "num_tiles_hit", "cum_tiles_hit"):
        assert np.array_equal(g[k].reshape(-1).astype(np.int64), o32[k].astype(np.int64)), k


def test_project_fwd_vs_float64():
    sc, cam, kw = _setup()
    g, _ = _project_both(sc, cam, kw)
    o64 = oracle.project.project_fwd(sc.means3d, sc.scales, sc.quats, cam.viewmat, **kw)
    vis = o64["radii"] > 0
    for k in ("xys", "conics", "cov3d"):
        a, b = g[k][vis].astype(np.float64), o64[k][vis]
        assert np.abs(a - b).max() <= 1e-5 * np.abs(b).max(), k
    # integer decisions may flip only at float32 rounding boundaries
    assert (g["radii"] != o64["radii"]).mean() < 1e-3


@pytest.mark.parametrize("split", [False, True])
@pytest.mark.parametrize("name,n,bw", [("small", None, 16), ("small", None, 8),
                                       ("small", 3000, 3), ("hd1m", 200_000, 16)])
def test_bin_sort_bitexact(name, n, bw, split):
    sc, cam, kw = _setup(name, n=n, block_width=bw)
    g = H.gpu_project(sc.means3d, sc.scales, sc.quats, cam.viewmat, **kw)
    gids, tiles, bins, isect = H.gpu_bin(g["xys"], g["depths"], g["radii"], g["cum_tiles_hit"],
                                         img_shape=cam.img_shape, block_width=bw, split=split)
    ks, gs, bs = oracle.binning.bin_and_sort(g["xys"], g["depths"], g["radii"],
                                             g["cum_tiles_hit"], img_shape=cam.img_shape,
                                             block_width=bw)
    assert isect == ks.size > 0
    assert np.array_equal(gids, gs)
    keys = (tiles.astype(np.int64) << 32) | g["depths"].view(np.int32)[gids].astype(np.int64)
    assert np.array_equal(keys, ks)
    assert np.array_equal(bins, bs)


def _raster_case(name="small", n=None, bw=16, opacity=None, bg=(0.0, 0.0, 0.0), scale=1.0):
    sc, cam, kw = _setup(name, n=n, block_width=bw)
    sc.scales = (sc.scales * scale).astype(np.float32)
    g = H.gpu_project(sc.means3d, sc.scales, sc.quats, cam.viewmat, **kw)
    gids, _, bins, isect = H.gpu_bin(g["xys"], g["depths"], g["radii"], g["cum_tiles_hit"],
                                     img_shape=cam.img_shape, block_width=bw)
    op = sc.opacities if opacity is None else opacity(sc.opacities)
    return sc, cam, g, gids, bins, isect, op, np.asarray(bg, np.float32)


@pytest.mark.parametrize("bw,bg", [(16, (0.0, 0.0, 0.0)), (8, (0.2, 0.5, 0.9)), (5, (1, 1, 1)),
                                  (12, (0.4, 0.4, 0.1)), (13, (0.0, 1.0, 0.5))])
def test_raster_fwd_vs_oracle(bw, bg):
    sc, cam, g, gids, bins, isect, op, bgv = _raster_case(bw=bw, bg=bg)
    out = H.gpu_raster_fwd(gids, bins, isect, g["xys"], g["conics"], sc.colors, op, bgv,
                           img_shape=cam.img_shape, block_width=bw)
    ref = oracle.rasterize.rasterize_fwd(gids, bins, g["xys"], g["conics"], sc.colors, op, bgv,
                                         img_shape=cam.img_shape, block_width=bw)
    ok = ~ref["ambiguous"]
    assert ref["ambiguous"].mean() < 2e-3
    assert np.abs(out["img"][ok] - ref["img"][ok]).max() <= 1e-4
    assert np.array_equal(out["final_idx"][ok], ref["final_idx"][ok])
    assert np.abs(out["final_Ts"][ok] - ref["final_Ts"][ok]).max() <= 1e-4


def test_raster_fwd_deep_termination():
    """High opacities: most pixels stop at T <= 1e-4 (termination path)."""
    sc, cam, g, gids, bins, isect, op, bgv = _raster_case(
        n=6000, opacity=lambda o: np.full_like(o, 0.97), scale=3.0)
    out = H.gpu_raster_fwd(gids, bins, isect, g["xys"], g["conics"], sc.colors, op, bgv,
                           img_shape=cam.img_shape)
    ref = oracle.rasterize.rasterize_fwd(gids, bins, g["xys"], g["conics"], sc.colors, op, bgv,
                                         img_shape=cam.img_shape, block_width=16)
    ok = ~ref["ambiguous"]
    assert (ref["final_Ts"] < 1e-3).mean() > 0.5
    assert ref["ambiguous"].mean() < 5e-3
    assert np.abs(out["img"][ok] - ref["img"][ok]).max() <= 1e-4
    assert np.array_equal(out["final_idx"][ok], ref["final_idx"][ok])


def test_raster_partial_tiles_vs_oracle():
    """Image sides that are not multiples of the 16-pixel tile (250 x 243):
    the last tile row and column are partial, so lanes outside the image
    start done.  Every other Gaussian is near opaque, so pixels also
    terminate inside those tiles.  Forward and backward vs the oracle."""
    import dataclasses
    sc = scenes.make("small", n=6000)
    sc.scales = (sc.scales * 3.0).astype(np.float32)
    cam = dataclasses.replace(sc.camera, img_shape=(250, 243), c=(121.5, 125.0))
    bw = 16
    g = H.gpu_project(sc.means3d, sc.scales, sc.quats, cam.viewmat, img_shape=cam.img_shape,
                      f=cam.f, c=cam.c, glob_scale=sc.glob_scale, clip_thresh=sc.clip_thresh,
                      block_width=bw)
    gids, _, bins, isect = H.gpu_bin(g["xys"], g["depths"], g["radii"], g["cum_tiles_hit"],
                                     img_shape=cam.img_shape, block_width=bw)
    assert isect > 0
    op = np.where(np.arange(sc.n)[:, None] % 2 == 0, 0.97, sc.opacities).astype(np.float32)
    bgv = np.asarray((0.3, 0.1, 0.6), np.float32)
    Hh, Ww = cam.img_shape
    out = H.gpu_raster_fwd(gids, bins, isect, g["xys"], g["conics"], sc.colors, op, bgv,
                           img_shape=cam.img_shape, block_width=bw)
    ref = oracle.rasterize.rasterize_fwd(gids, bins, g["xys"], g["conics"], sc.colors, op, bgv,
                                         img_shape=cam.img_shape, block_width=bw)
    ok = ~ref["ambiguous"]
    assert ref["ambiguous"].mean() < 5e-3
    assert (ref["final_Ts"] < 1e-3).mean() > 0.3           # most pixels terminate
    assert np.abs(out["img"][ok] - ref["img"][ok]).max() <= 1e-4
    assert np.array_equal(out["final_idx"][ok], ref["final_idx"][ok])
    assert np.abs(out["final_Ts"][ok] - ref["final_Ts"][ok]).max() <= 1e-4
    rng = np.random.default_rng(17)
    v_img = rng.normal(size=(Hh, Ww, 3)).astype(np.float32)
    v_alpha = rng.normal(size=(Hh, Ww)).astype(np.float32)
    v_img[ref["ambiguous"]] = 0
    v_alpha[ref["ambiguous"]] = 0
    got = H.gpu_raster_bwd(gids, bins, isect, g["xys"], g["conics"], sc.colors, op, bgv,
                           out["final_Ts"], out["final_idx"], v_img, v_alpha,
                           img_shape=cam.img_shape, block_width=bw, adjoint=0)
    exp = oracle.rasterize.rasterize_bwd(gids, bins, g["xys"], g["conics"], sc.colors, op, bgv,
                                         ref["final_Ts"], ref["final_idx"], v_img, v_alpha,
                                         img_shape=cam.img_shape, block_width=bw,
                                         mode="reference")
    for k in ("v_colors", "v_opacity", "v_xy", "v_xy_abs", "v_conic"):
        okg, worst = H.grad_close(got[k].reshape(exp[k].shape), exp[k])
        assert okg, (k, worst)


# tile widths 12..16 take the backward's four-slot quadrant layout (lanes past
# a 12- or 13-pixel tile own no pixel), up to 11 the two-slot row-major one
@pytest.mark.parametrize("adjoint,mode", [(0, "reference"), (1, "consistent")])
@pytest.mark.parametrize("bw", [16, 13, 12, 8])
def test_raster_bwd_vs_oracle(adjoint, mode, bw):
    sc, cam, g, gids, bins, isect, op, bgv = _raster_case(bw=bw, bg=(0.3, 0.1, 0.6),
                                                          opacity=lambda o: np.minimum(o * 1.3, 1.0))
    Hh, Ww = cam.img_shape
    out = H.gpu_raster_fwd(gids, bins, isect, g["xys"], g["conics"], sc.colors, op, bgv,
                           img_shape=cam.img_shape, block_width=bw)
    ref = oracle.rasterize.rasterize_fwd(gids, bins, g["xys"], g["conics"], sc.colors, op, bgv,
                                         img_shape=cam.img_shape, block_width=bw)
    rng = np.random.default_rng(7)
    v_img = rng.normal(size=(Hh, Ww, 3)).astype(np.float32)
    v_alpha = rng.normal(size=(Hh, Ww)).astype(np.float32)
    v_img[ref["ambiguous"]] = 0
    v_alpha[ref["ambiguous"]] = 0
    got = H.gpu_raster_bwd(gids, bins, isect, g["xys"], g["conics"], sc.colors, op, bgv,
                           out["final_Ts"], out["final_idx"], v_img, v_alpha,
                           img_shape=cam.img_shape, block_width=bw, adjoint=adjoint)
    exp = oracle.rasterize.rasterize_bwd(gids, bins, g["xys"], g["conics"], sc.colors, op, bgv,
                                         ref["final_Ts"], ref["final_idx"], v_img, v_alpha,
                                         img_shape=cam.img_shape, block_width=bw, mode=mode)
    for k in ("v_colors", "v_opacity", "v_xy", "v_xy_abs", "v_conic"):
        ok, worst = H.grad_close(got[k].reshape(exp[k].shape), exp[k])
        assert ok, (k, worst)


@pytest.mark.parametrize("adjoint,mode", [(0, "reference"), (1, "consistent")])
def test_project_bwd_vs_oracle(adjoint, mode):
    sc, cam, kw = _setup()
    # wide field of view clamp: put some Gaussians far off-axis
    means = sc.means3d.copy()
    means[:500, 0] *= 6.0
    g = H.gpu_project(means, sc.scales, sc.quats, cam.viewmat, **kw)
    N = sc.n
    rng = np.random.default_rng(3)
    v_xy = rng.normal(size=(N, 2)).astype(np.float32)
    v_depth = rng.normal(size=(N,)).astype(np.float32)
    v_conic = rng.normal(size=(N, 3)).astype(np.float32)
    got = H.gpu_project_bwd(means, sc.scales, sc.quats, cam.viewmat, g["radii"], g["conics"],
                            v_xy, v_depth, v_conic, adjoint=adjoint, **kw)
    exp = oracle.project.project_bwd(means, sc.scales, sc.quats, cam.viewmat, g["radii"],
                                     g["conics"], np.zeros(N), v_xy, v_depth, v_conic,
                                     np.zeros(N), img_shape=cam.img_shape, f=cam.f,
                                     glob_scale=sc.glob_scale, mode=mode)
    for k in ("v_mean3d", "v_scale", "v_quat", "v_cov2d", "v_cov3d"):
        ok, worst = H.grad_close(got[k], exp[k], rtol=1e-3, scale_frac=1e-4)
        assert ok, (k, worst)


def test_zero_intersections():
    """All Gaussians behind the camera: background image, final_T = 0, zero grads
    (reference lib/ops.cu:119-127,184-186)."""
    sc, cam, kw = _setup(n=500)
    means = sc.means3d.copy()
    means[:, 2] = -5.0
    g = H.gpu_project(means, sc.scales, sc.quats, cam.viewmat, **kw)
    assert int(g["cum_tiles_hit"][-1]) == 0
    assert not g["radii"].any() and not g["xys"].any() and not g["conics"].any()
    bg = np.array([0.25, 0.5, 0.75], np.float32)
    gids, _, bins, isect = H.gpu_bin(g["xys"], g["depths"], g["radii"], g["cum_tiles_hit"],
                                     img_shape=cam.img_shape)
    assert isect == 0 and not bins.any()
    out = H.gpu_raster_fwd(gids, bins, 0, g["xys"], g["conics"], sc.colors, sc.opacities, bg,
                           img_shape=cam.img_shape)
    assert np.all(out["img"] == bg) and not out["final_Ts"].any() and not out["final_idx"].any()
    Hh, Ww = cam.img_shape
    got = H.gpu_raster_bwd(gids, bins, 0, g["xys"], g["conics"], sc.colors, sc.opacities, bg,
                           out["final_Ts"], out["final_idx"], np.ones((Hh, Ww, 3)),
                           np.ones((Hh, Ww)), img_shape=cam.img_shape)
    assert all(not v.any() for v in got.values())


def test_forward_deterministic():
    sc, cam, g, gids, bins, isect, op, bgv = _raster_case()
    a = H.gpu_raster_fwd(gids, bins, isect, g["xys"], g["conics"], sc.colors, op, bgv,
                         img_shape=cam.img_shape)
    b = H.gpu_raster_fwd(gids, bins, isect, g["xys"], g["conics"], sc.colors, op, bgv,
                         img_shape=cam.img_shape)
    for k in a:
        assert np.array_equal(a[k], b[k])
    g2 = H.gpu_project(sc.means3d, sc.scales, sc.quats, sc.camera.viewmat,
                       img_shape=cam.img_shape, f=cam.f, c=cam.c)
    gids2, _, bins2, _ = H.gpu_bin(g2["xys"], g2["depths"], g2["radii"], g2["cum_tiles_hit"],
                                   img_shape=cam.img_shape)
    assert np.array_equal(gids, gids2) and np.array_equal(bins, bins2)
