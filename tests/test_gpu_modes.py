"""Render modes beyond the reference's default path, on the GPU.

Antialiased (compensation) mode: ``render(..., antialiased=True)`` blends with
opacities * compensation, the compensation the reference's projection
computes and drops (lib/kernels/forward.cu:537-543).  Bars: the image within
1e-4 max-abs of the float64 oracle fed the float32 projection (ambiguous
pixels excluded, < 0.2%); every gradient within
|g - g64| <= 1e-3 |g64| + 1e-4 max|g64| of the oracle chain (rasterize
backward, the product rule, projection backward with v_compensation through
helpers.h:73-87), in both adjoint modes.
"""

import numpy as np
import pytest

import oracle
from jaxsplat_amd import render, scenes
from tests import helpers as H

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("adjoint", ["reference", "consistent"])
def test_antialiased_render_vs_oracle(adjoint):
    t = H.torch()
    sc = scenes.make("small", n=3000)
    cam = sc.camera
    kw = dict(img_shape=cam.img_shape, f=cam.f, c=cam.c, glob_scale=1.0, clip_thresh=0.01,
              block_width=16)
    o = oracle.project.project_fwd(sc.means3d, sc.scales * 2, sc.quats, cam.viewmat,
                                   dtype=np.float32, **kw)
    comp = o["compensation"].astype(np.float32).reshape(-1, 1)
    op_eff = (sc.opacities * comp).astype(np.float32)
    _, gids, bins = oracle.binning.bin_and_sort(o["xys"], o["depths"], o["radii"],
                                                o["cum_tiles_hit"], img_shape=cam.img_shape,
                                                block_width=16)
    bg = np.array([0.2, 0.1, 0.3], np.float32)
    ref = oracle.rasterize.rasterize_fwd(gids, bins, o["xys"], o["conics"], sc.colors, op_eff,
                                         bg, img_shape=cam.img_shape, block_width=16)
    amb = ref["ambiguous"]
    assert amb.mean() < 2e-3

    params = [H.dev(a) for a in (sc.means3d, sc.scales * 2, sc.quats, sc.colors, sc.opacities)]
    for x in params:
        x.requires_grad_(True)
    img = render(*params, viewmat=H.dev(cam.viewmat), background=H.dev(bg),
                 img_shape=cam.img_shape, f=cam.f, c=cam.c, glob_scale=1.0, clip_thresh=0.01,
                 block_size=16, adjoint=adjoint, antialiased=True)
    got = H.host(img)
    assert np.abs(got[~amb] - ref["img"][~amb]).max() <= 1e-4
    # the mode changes the image: compensation < 1 where cov2d is small
    plain = H.host(render(*[x.detach() for x in params], viewmat=H.dev(cam.viewmat),
                          background=H.dev(bg), img_shape=cam.img_shape, f=cam.f, c=cam.c,
                          glob_scale=1.0, clip_thresh=0.01, block_size=16))
    assert np.abs(plain - got).max() > 1e-2

    w = np.random.default_rng(9).normal(size=got.shape).astype(np.float32)
    w[amb] = 0
    (img * H.dev(w)).sum().backward()
    rb = oracle.rasterize.rasterize_bwd(gids, bins, o["xys"], o["conics"], sc.colors, op_eff, bg,
                                        ref["final_Ts"], ref["final_idx"], w,
                                        np.zeros(cam.img_shape), img_shape=cam.img_shape,
                                        block_width=16, mode=adjoint)
    v_eff = rb["v_opacity"].reshape(-1, 1)
    v_op = v_eff * comp
    v_comp = (v_eff * sc.opacities).reshape(-1)
    N = sc.n
    pb = oracle.project.project_bwd(sc.means3d, sc.scales * 2, sc.quats, cam.viewmat,
                                    o["radii"], o["conics"], o["compensation"], rb["v_xy"],
                                    np.zeros(N), rb["v_conic"], v_comp, img_shape=cam.img_shape,
                                    f=cam.f, glob_scale=1.0, mode=adjoint)
    exp = [pb["v_mean3d"], pb["v_scale"], pb["v_quat"], rb["v_colors"], v_op]
    for k, (x, e) in enumerate(zip(params, exp)):
        ok, worst = H.grad_close(H.host(x.grad).reshape(e.shape), e)
        assert ok, (k, worst)


def _nd_setup(nch, seed=5):
    sc = scenes.make("small", n=3000)
    cam = sc.camera
    kw = dict(img_shape=cam.img_shape, f=cam.f, c=cam.c, glob_scale=1.0, clip_thresh=0.01,
              block_width=16)
    rng = np.random.default_rng(seed)
    col = rng.uniform(size=(sc.n, nch)).astype(np.float32)
    bg = rng.uniform(size=nch).astype(np.float32)
    return sc, cam, kw, col, bg


def _render_nd(sc, cam, col, bg, adjoint="reference", grad=True):
    params = [H.dev(a) for a in (sc.means3d, sc.scales * 2, sc.quats)] + [H.dev(col),
                                                                          H.dev(sc.opacities)]
    if grad:
        for x in params:
            x.requires_grad_(True)
    img = render(*params, viewmat=H.dev(cam.viewmat), background=H.dev(bg),
                 img_shape=cam.img_shape, f=cam.f, c=cam.c, glob_scale=1.0, clip_thresh=0.01,
                 block_size=16, adjoint=adjoint)
    return img, params


@pytest.mark.parametrize("nch", [1, 2, 4, 7])
@pytest.mark.parametrize("adjoint", ["reference", "consistent"])
def test_nchannel_render_vs_oracle(nch, adjoint):
    """N-channel rasterization: image within 1e-4 of the float64 oracle (ambiguous
    pixels excluded), every gradient within 1e-3 relative of the oracle chain."""
    sc, cam, kw, col, bg = _nd_setup(nch)
    o = oracle.project.project_fwd(sc.means3d, sc.scales * 2, sc.quats, cam.viewmat,
                                   dtype=np.float32, **kw)
    _, gids, bins = oracle.binning.bin_and_sort(o["xys"], o["depths"], o["radii"],
                                                o["cum_tiles_hit"], img_shape=cam.img_shape,
                                                block_width=16)
    ref = oracle.rasterize.rasterize_fwd(gids, bins, o["xys"], o["conics"], col, sc.opacities,
                                         bg, img_shape=cam.img_shape, block_width=16)
    amb = ref["ambiguous"]
    assert amb.mean() < 2e-3
    img, params = _render_nd(sc, cam, col, bg, adjoint)
    got = H.host(img)
    assert got.shape == (*cam.img_shape, nch)
    assert np.abs(got[~amb] - ref["img"][~amb]).max() <= 1e-4
    w = np.random.default_rng(11).normal(size=got.shape).astype(np.float32)
    w[amb] = 0
    (img * H.dev(w)).sum().backward()
    rb = oracle.rasterize.rasterize_bwd(gids, bins, o["xys"], o["conics"], col, sc.opacities, bg,
                                        ref["final_Ts"], ref["final_idx"], w,
                                        np.zeros(cam.img_shape), img_shape=cam.img_shape,
                                        block_width=16, mode=adjoint)
    N = sc.n
    pb = oracle.project.project_bwd(sc.means3d, sc.scales * 2, sc.quats, cam.viewmat,
                                    o["radii"], o["conics"], o["compensation"], rb["v_xy"],
                                    np.zeros(N), rb["v_conic"], np.zeros(N),
                                    img_shape=cam.img_shape, f=cam.f, glob_scale=1.0, mode=adjoint)
    exp = [pb["v_mean3d"], pb["v_scale"], pb["v_quat"], rb["v_colors"],
           rb["v_opacity"].reshape(-1, 1)]
    for k, (x, e) in enumerate(zip(params, exp)):
        ok, worst = H.grad_close(H.host(x.grad).reshape(e.shape), e)
        assert ok, (k, worst)


def test_nchannel_splits_into_exact_rgb_renders():
    """Six channels render as two RGB passes: each half of the image equals the
    3-channel render of those colours bit for bit, and the 3-channel
    N-channel entry point equals the typed RGB entry point bit for bit."""
    t = H.torch()
    sc, cam, kw, col, bg = _nd_setup(6, seed=8)
    with t.no_grad():
        six = H.host(_render_nd(sc, cam, col, bg, grad=False)[0])
        a = H.host(_render_nd(sc, cam, col[:, :3].copy(), bg[:3].copy(), grad=False)[0])
        b = H.host(_render_nd(sc, cam, col[:, 3:].copy(), bg[3:].copy(), grad=False)[0])
    assert np.array_equal(six[..., :3], a) and np.array_equal(six[..., 3:], b)
    pr = H.gpu_project(sc.means3d, sc.scales * 2, sc.quats, cam.viewmat, **kw)
    gids, _, bins, isect = H.gpu_bin(pr["xys"], pr["depths"], pr["radii"], pr["cum_tiles_hit"],
                                     img_shape=cam.img_shape)
    rgb = H.gpu_raster_fwd(gids, bins, isect, pr["xys"], pr["conics"], col[:, :3].copy(),
                           sc.opacities, bg[:3].copy(), img_shape=cam.img_shape)
    assert np.array_equal(rgb["img"], a)
