"""Full-size parity of the timed path: ``jaxsplat_amd.render`` with its
defaults -- projection writing the opacity-limited tile rectangles, binning
into those shorter lists, the rasterize forward, and the staged (atomic-free)
rasterize backward under torch autograd -- at ``hd1m`` (1,000,000 Gaussians,
1080x1920) and ``dense3m`` (3,000,000 Gaussians, ~107M intersections).

Bars (written in each test):
* the whole-frame image is bit-identical to the same render through the
  reference's tile lists (``TIGHT_TILES = False``) and to the rasterize
  forward over the reference lists that ``test_gpu_fullsize`` checks against
  the oracle (same float32 decisions, dropped entries never blend);
* on seeded tile subsets (plus the deepest tiles), image max-abs <= 1e-4 vs
  the float64 oracle outside float32-ambiguous pixels (at most 1% of them);
* with the image cotangent restricted to the subset (a tile's result depends
  only on its own list), all five parameter gradients that ``render``'s
  autograd returns -- means3d, scales, quats, colors, opacities -- against
  the float64 oracle chain (rasterize backward, then projection backward) on
  the reference's lists: |g - g64| <= 1e-3 |g64| + 1e-4 max|g64|, and exactly
  zero for every Gaussian outside the subset's lists.
"""

import numpy as np
import pytest

import jaxsplat_amd
import oracle
from jaxsplat_amd import _ops
from tests import helpers as H
from tests.test_gpu_fullsize import BW, _fwd_subset, _pipeline, _subset

pytestmark = pytest.mark.gpu


def _render(sc, cam, w=None, tight=True, adjoint="reference"):
    """render() through the public API; with w, also the gradients of
    sum(img * w) from torch autograd."""
    t = H.torch()
    dev = t.device("cuda")
    params = [t.from_numpy(np.ascontiguousarray(a)).to(dev).requires_grad_(True)
              for a in sc.params()]
    kw = dict(viewmat=t.from_numpy(cam.viewmat).to(dev),
              background=t.from_numpy(sc.background).to(dev), img_shape=cam.img_shape,
              f=cam.f, c=cam.c, glob_scale=sc.glob_scale, clip_thresh=sc.clip_thresh,
              block_size=BW, adjoint=adjoint)
    saved = _ops.TIGHT_TILES
    _ops.TIGHT_TILES = tight
    try:
        img = jaxsplat_amd.render(*params, **kw)
        if w is not None:
            (img * t.from_numpy(w).to(dev)).sum().backward()
    finally:
        _ops.TIGHT_TILES = saved
    grads = [H.host(p.grad) if p.grad is not None else None for p in params]
    return H.host(img), grads


@pytest.mark.parametrize("name", ["hd1m", "dense3m"])
def test_render_fullsize_image_bitexact_vs_reference_lists(name):
    sc, cam, kw, g, gids, _, bins, isect, fwd = _pipeline(name)
    print(f"[render-fullsize] {name}: render() image, tight vs reference lists", flush=True)
    img_tight, _ = _render(sc, cam, tight=True)
    img_ref, _ = _render(sc, cam, tight=False)
    assert np.array_equal(img_tight.view(np.uint32), img_ref.view(np.uint32))
    # the forward over the reference lists that test_gpu_fullsize pins to the oracle
    assert np.array_equal(img_tight.reshape(-1).view(np.uint32),
                          fwd["img"].reshape(-1).view(np.uint32))


@pytest.mark.parametrize("name,mode", [("hd1m", "reference"), ("hd1m", "consistent"),
                                       ("dense3m", "reference")])
def test_render_fullsize_gradients_vs_oracle_subset(name, mode):
    sc, cam, kw, g, gids, _, bins, isect, fwd = _pipeline(name)
    tiles = _subset(bins, n=24, deep=4, seed=31)
    ref, mask = _fwd_subset(name, tiles)
    amb = ref["ambiguous"] & mask
    assert amb.sum() <= 1e-2 * mask.sum(), amb.sum()
    rng = np.random.default_rng(37)
    w = rng.normal(size=(*cam.img_shape, 3)).astype(np.float32)
    w[~mask] = 0
    w[ref["ambiguous"]] = 0
    print(f"[render-fullsize] {name}/{mode}: render() fwd+bwd", flush=True)
    img, (gm, gs, gq, gc, go) = _render(sc, cam, w=w, adjoint=mode)
    ok = mask & ~ref["ambiguous"]
    assert np.abs(img[ok] - ref["img"][ok]).max() <= 1e-4
    # oracle chain on the reference lists of the subset
    ob = oracle.rasterize.rasterize_bwd(gids, bins, g["xys"], g["conics"], sc.colors,
                                        sc.opacities, sc.background, ref["final_Ts"],
                                        ref["final_idx"], w, np.zeros(cam.img_shape),
                                        img_shape=cam.img_shape, block_width=BW, mode=mode,
                                        tiles=tiles)
    rows = np.unique(np.concatenate([gids[bins[t, 0]:bins[t, 1]] for t in tiles]))
    zN = np.zeros(rows.size)
    op = oracle.project.project_bwd(sc.means3d[rows], sc.scales[rows], sc.quats[rows],
                                    cam.viewmat, g["radii"].reshape(-1)[rows],
                                    g["conics"][rows], g["compensation"].reshape(-1)[rows],
                                    ob["v_xy"][rows], zN, ob["v_conic"][rows], zN,
                                    img_shape=cam.img_shape, f=cam.f, glob_scale=sc.glob_scale,
                                    mode=mode)
    others = np.ones(sc.n, bool)
    others[rows] = False
    checks = [("v_colors", gc, ob["v_colors"], True), ("v_opacity", go, ob["v_opacity"], True),
              ("v_mean3d", gm, op["v_mean3d"], False), ("v_scale", gs, op["v_scale"], False),
              ("v_quat", gq, op["v_quat"], False)]
    for k, got, exp, full in checks:
        got_rows = got.reshape(sc.n, -1)[rows]
        exp_rows = exp.reshape(-1, got_rows.shape[1])
        exp_rows = exp_rows[rows] if full else exp_rows
        ok_, worst = H.grad_close(got_rows, exp_rows)
        assert ok_, (k, worst)
        assert not got.reshape(sc.n, -1)[others].any(), k
