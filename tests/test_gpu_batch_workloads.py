"""The batched workloads at their full sizes, on the GPU:

* ``orbit64``: one scene of 500,000 Gaussians seen by 64 orbit cameras at
  800x800;
* ``multi32``: 32 scenes of 10,000 to 2,000,000 Gaussians (8.0M in all), one
  camera each, image sizes 512x512, 768x1024 and 1080x1920.

The reference renders one view per call (``jaxsplat/__init__.py:6-70``), so
its result for a batch is the plain render of every view.  Bars:

* every view's batched image is bit-identical to a plain ``render()`` of that
  view (same kernels, same per-tile order of the same tile lists);
* the batched gradients equal the sum over the views of the plain per-view
  gradients of the same cotangent within |g - g_sum| <= 1e-3 |g_sum| +
  1e-4 max|g_sum| (the batch sums per item, then per Gaussian over the views;
  the plain calls add per call, so the float additions differ in order);
* two views per workload -- for multi32 a 1080x1920 scene and a 512x512 one
  -- against the float64 oracle on seeded tile subsets: image max-abs <= 1e-4
  outside float32-ambiguous pixels, and with the cotangent restricted to the
  subset of that one view, every parameter gradient of the batch against the
  oracle chain (rasterize backward, then projection backward) on the
  reference's lists of that view, at the bar above.
"""

import numpy as np
import pytest

import jaxsplat_amd
import oracle
from jaxsplat_amd import BatchRenderer, ViewSpec, scenes
from tests import helpers as H

pytestmark = pytest.mark.gpu

BW = 16
_CACHE = {}


def _workload(name):
    """(scene sizes, views, per-view (scene, camera), float32 parameter arrays
    with the scenes end to end)."""
    if name not in _CACHE:
        if name == "orbit64":
            sc = scenes.make("orbit64")
            sizes = [sc.n]
            cams = [(0, c) for c in sc.cameras]
            params = [np.ascontiguousarray(a) for a in sc.params()]
        else:
            ms = scenes.make_multi()
            sizes = [m.n for m in ms]
            cams = [(i, m.camera) for i, m in enumerate(ms)]
            params = [np.ascontiguousarray(np.concatenate(x))
                      for x in zip(*[m.params() for m in ms])]
        views = [ViewSpec(s, c.img_shape, c.f, c.c) for s, c in cams]
        _CACHE.clear()
        _CACHE[name] = (sizes, views, cams, params)
    return _CACHE[name]


def _offsets(sizes):
    return np.concatenate([[0], np.cumsum(sizes)]).astype(np.int64)


def _weights(cams, seed, only=None, mask=None):
    """Per-view cotangents: N(0, 1) images, or only view `only` (restricted to
    `mask`), zeros elsewhere."""
    rng = np.random.default_rng(seed)
    out = []
    for v, (_, c) in enumerate(cams):
        if only is None:
            out.append(rng.normal(size=(*c.img_shape, 3)).astype(np.float32))
        elif v == only:
            w = rng.normal(size=(*c.img_shape, 3)).astype(np.float32)
            w[~mask] = 0
            out.append(w)
        else:
            out.append(np.zeros((*c.img_shape, 3), np.float32))
    return out


def _batched(name, ws):
    """Batched forward + backward of sum_v <img_v, w_v>: per-view images and
    the parameter gradients."""
    t = H.torch()
    sizes, views, cams, params = _workload(name)
    dp = [H.dev(a).requires_grad_(True) for a in params]
    vms = H.dev(np.stack([c.viewmat for _, c in cams]))
    bgs = t.zeros((len(cams), 3), dtype=t.float32, device="cuda")
    r = BatchRenderer(sizes, views, glob_scale=1.0, clip_thresh=0.01, block_width=BW)
    flat = r(*dp, vms, bgs)
    W = t.cat([t.from_numpy(w.reshape(-1, 3)) for w in ws]).to("cuda")
    (flat * W).sum().backward()
    imgs = [H.host(x) for x in r.split(flat.detach())]
    return imgs, [H.host(p.grad) for p in dp]


def _plain(name, ws, views_only=None):
    """Plain render() per view (the reference's usage): images and the sum of
    the per-view parameter gradients."""
    t = H.torch()
    sizes, views, cams, params = _workload(name)
    offs = _offsets(sizes)
    dp = [H.dev(a).requires_grad_(True) for a in params]
    bg = t.zeros(3, dtype=t.float32, device="cuda")
    imgs = []
    for v, (s, c) in enumerate(cams):
        if views_only is not None and v not in views_only:
            imgs.append(None)
            continue
        a, b = int(offs[s]), int(offs[s + 1])
        img = jaxsplat_amd.render(*[x[a:b] for x in dp], viewmat=H.dev(c.viewmat), background=bg,
                                  img_shape=c.img_shape, f=c.f, c=c.c, glob_scale=1.0,
                                  clip_thresh=0.01, block_size=BW)
        (img * t.from_numpy(ws[v]).to("cuda")).sum().backward()
        imgs.append(H.host(img.detach()))
    return imgs, [H.host(p.grad) if p.grad is not None else np.zeros(p.shape, np.float32)
                  for p in dp]


@pytest.mark.parametrize("name", ["orbit64", "multi32"])
def test_batch_workload_equals_plain_renders(name):
    sizes, views, cams, params = _workload(name)
    print(f"[batch-workload] {name}: {len(cams)} views, {sum(sizes)} Gaussians", flush=True)
    ws = _weights(cams, seed=41)
    imgs_b, grads_b = _batched(name, ws)
    imgs_p, grads_p = _plain(name, ws)
    for v in range(len(cams)):
        assert np.array_equal(imgs_b[v].view(np.uint32), imgs_p[v].view(np.uint32)), v
    for k, (gb, gp) in enumerate(zip(grads_b, grads_p)):
        ok, worst = H.grad_close(gb, gp)
        assert ok, (k, worst)


def _subset(bins, n, deep, seed):
    lens = bins[:, 1] - bins[:, 0]
    nonempty = np.nonzero(lens > 0)[0]
    rng = np.random.default_rng(seed)
    pick = rng.choice(nonempty, size=min(n, nonempty.size), replace=False)
    return np.unique(np.concatenate([pick, np.argsort(lens)[-deep:]]))


def _tile_mask(tiles, img_shape):
    Hh, Ww = img_shape
    tx_n = (Ww + BW - 1) // BW
    m = np.zeros(img_shape, bool)
    for t in tiles:
        ty, tx = divmod(int(t), tx_n)
        m[ty * BW:(ty + 1) * BW, tx * BW:(tx + 1) * BW] = True
    return m


def _pick_views(name):
    sizes, views, cams, params = _workload(name)
    if name == "orbit64":
        return [5, 38]
    shapes = [tuple(c.img_shape) for _, c in cams]
    big = max((v for v in range(len(cams)) if shapes[v] == (1080, 1920)),
              key=lambda v: sizes[cams[v][0]])
    small = max((v for v in range(len(cams)) if shapes[v] == (512, 512)),
                key=lambda v: sizes[cams[v][0]])
    return [big, small]


@pytest.mark.parametrize("name", ["orbit64", "multi32"])
def test_batch_workload_views_vs_oracle(name):
    sizes, views, cams, params = _workload(name)
    offs = _offsets(sizes)
    for v in _pick_views(name):
        s, cam = cams[v]
        a, b = int(offs[s]), int(offs[s + 1])
        m, sc_, q, col, op = (x[a:b] for x in params)
        kw = dict(img_shape=cam.img_shape, f=cam.f, c=cam.c, glob_scale=1.0, clip_thresh=0.01,
                  block_width=BW)
        print(f"[batch-workload] {name}: view {v} ({cam.img_shape}, {b - a} Gaussians) vs oracle",
              flush=True)
        g = H.gpu_project(m, sc_, q, cam.viewmat, **kw)
        gids, _, bins, isect = H.gpu_bin(g["xys"], g["depths"], g["radii"], g["cum_tiles_hit"],
                                         img_shape=cam.img_shape, block_width=BW)
        tiles = _subset(bins, n=16, deep=3, seed=43 + v)
        ref = oracle.rasterize.rasterize_fwd(gids, bins, g["xys"], g["conics"], col, op,
                                             np.zeros(3, np.float32), img_shape=cam.img_shape,
                                             block_width=BW, tiles=tiles, chunk=512)
        mask = _tile_mask(tiles, cam.img_shape)
        amb = ref["ambiguous"] & mask
        assert amb.sum() <= 1e-2 * mask.sum(), amb.sum()
        wv = _weights(cams, seed=47 + v, only=v, mask=mask & ~ref["ambiguous"])
        imgs_b, (gm, gs, gq, gc, go) = _batched(name, wv)
        ok = mask & ~ref["ambiguous"]
        assert np.abs(imgs_b[v][ok] - ref["img"][ok]).max() <= 1e-4
        w = wv[v]
        ob = oracle.rasterize.rasterize_bwd(gids, bins, g["xys"], g["conics"], col, op,
                                            np.zeros(3, np.float32), ref["final_Ts"],
                                            ref["final_idx"], w, np.zeros(cam.img_shape),
                                            img_shape=cam.img_shape, block_width=BW,
                                            tiles=tiles)
        rows = np.unique(np.concatenate([gids[bins[t, 0]:bins[t, 1]] for t in tiles]))
        zN = np.zeros(rows.size)
        opb = oracle.project.project_bwd(m[rows], sc_[rows], q[rows], cam.viewmat,
                                         g["radii"].reshape(-1)[rows], g["conics"][rows],
                                         g["compensation"].reshape(-1)[rows], ob["v_xy"][rows],
                                         zN, ob["v_conic"][rows], zN, img_shape=cam.img_shape,
                                         f=cam.f, glob_scale=1.0)
        checks = [(gc, ob["v_colors"][rows]), (go, ob["v_opacity"][rows].reshape(-1, 1)),
                  (gm, opb["v_mean3d"]), (gs, opb["v_scale"]), (gq, opb["v_quat"])]
        for k, (got, exp) in enumerate(checks):
            ok_, worst = H.grad_close(got[a:b][rows], exp)
            assert ok_, (v, k, worst)
            others = np.ones(got.shape[0], bool)
            others[a + rows] = False
            assert not got[others].any(), (v, k)


def test_multi32_queue_schedule_equals_grid():
    """multi32 at full size: the work-queue rasterizer gives the flat grid's
    images and gradients bit for bit."""
    t = H.torch()
    sizes, views, cams, params = _workload("multi32")
    vms = H.dev(np.stack([c.viewmat for _, c in cams]))
    bgs = t.zeros((len(cams), 3), dtype=t.float32, device="cuda")
    ws = _weights(cams, seed=53)
    W = t.cat([t.from_numpy(w.reshape(-1, 3)) for w in ws]).to("cuda")
    res = {}
    for sched in ("grid", "queue"):
        r = BatchRenderer(sizes, views, glob_scale=1.0, clip_thresh=0.01, block_width=BW)
        dp = [H.dev(a).requires_grad_(True) for a in params]
        flat = r(*dp, vms, bgs)                    # sizes the capacity (grid schedule)
        r.plan.set_schedule(sched)
        for p in dp:
            p.grad = None
        flat = r(*dp, vms, bgs)
        (flat * W).sum().backward()
        res[sched] = [H.host(flat.detach())] + [H.host(p.grad) for p in dp]
    for a, b in zip(res["grid"], res["queue"]):
        assert np.array_equal(a.view(np.uint32), b.view(np.uint32))
