"""Batched rendering (many views of one or more scenes per call) on the GPU.

Bars (written in each test):
* per view, the batched image and final_Ts are bit-identical to the plain
  path's render of that view (same kernels, same per-tile order);
* one view against the float64 oracle at the plain path's image bar
  (max-abs <= 1e-4 outside float32-ambiguous pixels);
* gradients equal the sum over the scene's views of the plain per-view
  gradients within |g - g_sum| <= 1e-3 |g_sum| + 1e-4 max|g_sum| (the batch
  sums per item and then per Gaussian over the views, the plain path per
  Gaussian, so the float additions differ in order);
* a view that sees nothing renders the background with zero gradients;
* a capacity overflow is reported, and BatchRenderer recovers from it;
* a HIP graph of forward + backward replays to the eager results bit for bit
  (the staged backward sums in a fixed order and writes every gradient row).
"""

import numpy as np
import pytest

import oracle
from jaxsplat_amd import BatchPlan, BatchRenderer, ViewSpec, _lib, render, render_batch, scenes
from tests import helpers as H

pytestmark = pytest.mark.gpu


def _layout(seed=7):
    """Two scenes: 4000 Gaussians seen by 3 orbit cameras (one of them looking
    away from the scene) and 2500 Gaussians seen by 2 cameras; image sizes
    that are not tile multiples."""
    rng = np.random.default_rng(seed)
    a = scenes.gaussians(rng, 4000, (-1.5, -1.5, -1.5), (1.5, 1.5, 1.5), np.log(0.04))
    b = scenes.gaussians(rng, 2500, (-1, -1, 3), (1, 1, 6), np.log(0.03))
    params = [np.concatenate([x, y]) for x, y in zip(a, b)]
    cams = [
        (0, scenes.look_at((0.0, 1.0, -6.0)), (160, 200), (210.0, 200.0), (100.0, 80.0)),
        (0, scenes.look_at((5.0, -2.0, 3.0)), (144, 144), (180.0, 180.0), (72.0, 72.0)),
        (0, scenes.look_at((0.0, 0.0, -30.0), target=(0.0, 0.0, -60.0)), (64, 96),
         (100.0, 100.0), (48.0, 32.0)),
        (1, scenes.IDENTITY.copy(), (96, 128), (128.0, 128.0), (64.0, 48.0)),
        (1, scenes.look_at((0.5, 0.0, -1.0), target=(0.0, 0.0, 4.5)), (120, 200),
         (190.0, 190.0), (100.0, 60.0)),
    ]
    return params, (4000, 2500), cams


def _plan(sizes, cams, bw=16, capacity=1 << 20):
    views = [ViewSpec(s, hw, f, c) for s, _vm, hw, f, c in cams]
    return BatchPlan(sizes, views, glob_scale=1.0, clip_thresh=0.01, block_width=bw,
                     isect_capacity=capacity)


def _dev_params(params, requires_grad=False):
    t = H.torch()
    out = [H.dev(p) for p in params]
    if requires_grad:
        for x in out:
            x.requires_grad_(True)
    return out


def _views_arrays(cams):
    vm = np.stack([c[1] for c in cams]).astype(np.float32)
    bg = np.array([[0.1 * (i + 1), 0.2, 0.3] for i in range(len(cams))], np.float32)
    return vm, bg


def _scene_rows(sizes, s):
    a = sum(sizes[:s])
    return slice(a, a + sizes[s])


def _plain_view(params, sizes, cam, bg, bw):
    """Plain C-ABI path for one view: image and final_Ts."""
    s, vm, hw, f, c = cam
    rows = _scene_rows(sizes, s)
    m, sc, q, col, op = (p[rows] for p in params)
    pr = H.gpu_project(m, sc, q, vm, img_shape=hw, f=f, c=c, block_width=bw)
    gids, _, bins, isect = H.gpu_bin(pr["xys"], pr["depths"], pr["radii"], pr["cum_tiles_hit"],
                                     img_shape=hw, block_width=bw)
    out = H.gpu_raster_fwd(gids, bins, isect, pr["xys"], pr["conics"], col, op, bg,
                           img_shape=hw, block_width=bw)
    return out, pr, gids, bins


@pytest.mark.parametrize("bw", [16, 8])
def test_batch_forward_equals_plain_per_view(bw):
    t = H.torch()
    params, sizes, cams = _layout()
    plan = _plan(sizes, cams, bw)
    vm, bg = _views_arrays(cams)
    with t.no_grad():
        flat = render_batch(*_dev_params(params), H.dev(vm), H.dev(bg), plan=plan)
    n = plan.check()
    assert n > 0
    imgs = [H.host(x) for x in plan.split(flat)]
    total = 0
    for v, cam in enumerate(cams):
        out, pr, _, _ = _plain_view(params, sizes, cam, bg[v], bw)
        # the batch bins into the opacity-limited rectangles, as rasterize does
        _, tcum = H.gpu_tile_rects(pr["xys"], pr["radii"], pr["conics"],
                                   params[4][_scene_rows(sizes, cam[0])], img_shape=cam[2],
                                   block_width=bw)
        total += int(tcum[-1]) if tcum.size else 0
        assert np.array_equal(imgs[v], out["img"]), f"view {v} image"
    assert n == total                      # the exact count is the sum of the views' counts
    # the view looking away sees nothing: background only
    assert np.array_equal(imgs[2], np.broadcast_to(bg[2], imgs[2].shape))


def test_batch_final_ts_equal_plain():
    """final_Ts per view (a residual of the backward) equals the plain path's,
    except views with no intersections, which keep T = 1."""
    t = H.torch()
    params, sizes, cams = _layout()
    plan = _plan(sizes, cams)
    vm, bg = _views_arrays(cams)
    d = _dev_params(params)
    P = plan.num_pixels
    fT = t.full((P,), 7.0, device="cuda")
    fI = t.full((P,), 7, dtype=t.int32, device="cuda")
    img = t.full((P, 3), 7.0, device="cuda")
    vmd, bgd = H.dev(vm), H.dev(bg)
    _lib.check(_lib.lib().jaxsplat_batch_forward(
        plan._h, *[H.P(x) for x in d], H.P(vmd), H.P(bgd), H.P(fT), H.P(fI), H.P(img),
        H.P(plan.workspace), plan.workspace_bytes, H.stream()))
    Ts = plan.split(fT.view(-1, 1))
    for v, cam in enumerate(cams):
        out, pr, _, _ = _plain_view(params, sizes, cam, bg[v], 16)
        got = H.host(Ts[v]).reshape(out["final_Ts"].shape)
        if int(pr["cum_tiles_hit"][-1]) == 0:
            assert np.all(got == 1.0)
        else:
            assert np.array_equal(got, out["final_Ts"]), f"view {v}"


def test_batch_view_vs_float64_oracle():
    params, sizes, cams = _layout()
    t = H.torch()
    plan = _plan(sizes, cams)
    vm, bg = _views_arrays(cams)
    with t.no_grad():
        flat = render_batch(*_dev_params(params), H.dev(vm), H.dev(bg), plan=plan)
    img = H.host(plan.split(flat)[0])
    s, view, hw, f, c = cams[0]
    rows = _scene_rows(sizes, s)
    m, sc, q, col, op = (p[rows] for p in params)
    o = oracle.project.project_fwd(m, sc, q, view, img_shape=hw, f=f, c=c, glob_scale=1.0,
                                   clip_thresh=0.01, block_width=16, dtype=np.float32)
    _, gids, bins = oracle.binning.bin_and_sort(o["xys"], o["depths"], o["radii"],
                                                o["cum_tiles_hit"], img_shape=hw,
                                                block_width=16)
    ref = oracle.rasterize.rasterize_fwd(gids, bins, o["xys"], o["conics"], col, op, bg[0],
                                         img_shape=hw, block_width=16)
    ok = ~ref["ambiguous"]
    assert ref["ambiguous"].mean() < 2e-3
    assert np.abs(img[ok] - ref["img"][ok]).max() <= 1e-4


@pytest.mark.parametrize("adjoint", ["reference", "consistent"])
def test_batch_grads_equal_sum_of_plain(adjoint):
    t = H.torch()
    params, sizes, cams = _layout()
    plan = _plan(sizes, cams)
    vm, bg = _views_arrays(cams)
    rng = np.random.default_rng(3)
    w = rng.normal(size=(plan.num_pixels, 3)).astype(np.float32)
    d = _dev_params(params, requires_grad=True)
    flat = render_batch(*d, H.dev(vm), H.dev(bg), plan=plan, adjoint=adjoint)
    (flat * H.dev(w)).sum().backward()
    got = [H.host(x.grad) for x in d]

    ref = [np.zeros_like(p, dtype=np.float64) for p in params]
    ws = plan.split(H.dev(w))
    for v, (s, view, hw, f, c) in enumerate(cams):
        rows = _scene_rows(sizes, s)
        pv = [H.dev(p[rows]) for p in params]
        for x in pv:
            x.requires_grad_(True)
        img = render(*pv, viewmat=H.dev(view), background=H.dev(bg[v]), img_shape=hw, f=f, c=c,
                     glob_scale=1.0, clip_thresh=0.01, block_size=16, adjoint=adjoint)
        (img * ws[v]).sum().backward()
        for k, x in enumerate(pv):
            ref[k][rows] += H.host(x.grad)
    for k, name in enumerate(("means3d", "scales", "quats", "colors", "opacities")):
        ok, worst = H.grad_close(got[k], ref[k])
        assert ok, (name, worst)
    # Gaussians of scene 0 seen only by view 2 (which sees nothing) still get
    # the other views' gradients; all gradients are finite
    assert all(np.isfinite(g).all() for g in got)


def test_batch_capacity_overflow_and_renderer():
    t = H.torch()
    params, sizes, cams = _layout()
    vm, bg = _views_arrays(cams)
    small = _plan(sizes, cams, capacity=100)
    with t.no_grad():
        render_batch(*_dev_params(params), H.dev(vm), H.dev(bg), plan=small)
    with pytest.raises(_lib.JaxsplatError) as e:
        small.check()
    assert e.value.code == _lib.ERR_CAPACITY
    need = small.last_intersections
    assert need > 100
    full = _plan(sizes, cams, capacity=need)
    with t.no_grad():
        ref = H.host(render_batch(*_dev_params(params), H.dev(vm), H.dev(bg), plan=full))
    assert full.check() == need
    views = [ViewSpec(s, hw, f, c) for s, _, hw, f, c in cams]
    r = BatchRenderer(sizes, views, initial_capacity=100)
    with t.no_grad():
        got = H.host(r(*_dev_params(params), H.dev(vm), H.dev(bg)))
    assert r.capacity >= need
    assert np.array_equal(got, ref)
    # a later call that needs more intersections than the first one sized the
    # capacity for (most opacities below 1/255 at first) is checked and
    # rendered again, not cut off at the capacity
    faint = list(params)
    keep = np.arange(params[4].shape[0]) % 8 == 0
    faint[4] = (params[4] * np.where(keep, 1.0, 1e-3)[:, None]).astype(np.float32)
    r2 = BatchRenderer(sizes, views, initial_capacity=100)
    with t.no_grad():
        r2(*_dev_params(faint), H.dev(vm), H.dev(bg))
        assert r2.capacity < need
        got2 = H.host(r2(*_dev_params(params), H.dev(vm), H.dev(bg)))
    assert r2.capacity >= need
    assert np.array_equal(got2, ref)


def test_batch_backward_after_second_forward_raises():
    t = H.torch()
    params, sizes, cams = _layout()
    plan = _plan(sizes, cams)
    vm, bg = _views_arrays(cams)
    d = _dev_params(params, requires_grad=True)
    a = render_batch(*d, H.dev(vm), H.dev(bg), plan=plan)
    render_batch(*d, H.dev(vm), H.dev(bg), plan=plan)
    with pytest.raises(RuntimeError, match="residuals"):
        a.sum().backward()


def test_batch_hip_graph_replays_match_eager():
    """Forward + backward of the whole batch captured in one HIP graph;
    replayed three times with eager work and synchronisations in between,
    every replay equals the eager run bit for bit, image and gradients (the
    gradient buffers are refilled with garbage before each replay)."""
    t = H.torch()
    params, sizes, cams = _layout()
    plan = _plan(sizes, cams)
    vm, bg = _views_arrays(cams)
    L = _lib.lib()
    d = _dev_params(params)
    vmd, bgd = H.dev(vm), H.dev(bg)
    P = plan.num_pixels
    rng = np.random.default_rng(5)
    v_img = H.dev(rng.normal(size=(P, 3)).astype(np.float32))
    fT = t.empty((P,), device="cuda")
    fI = t.empty((P,), dtype=t.int32, device="cuda")
    img = t.empty((P, 3), device="cuda")
    grads = [t.empty_like(x) for x in d]

    def step():
        s = H.stream()
        _lib.check(L.jaxsplat_batch_forward(
            plan._h, *[H.P(x) for x in d], H.P(vmd), H.P(bgd), H.P(fT), H.P(fI), H.P(img),
            H.P(plan.workspace), plan.workspace_bytes, s))
        _lib.check(L.jaxsplat_batch_backward(
            plan._h, *[H.P(x) for x in d], H.P(vmd), H.P(bgd), H.P(fT), H.P(fI), H.P(v_img),
            *[H.P(g) for g in grads], 0, H.P(plan.workspace), plan.workspace_bytes, s))

    step()
    eager_img = H.host(img).copy()
    eager_grads = [H.host(g).copy() for g in grads]
    side = t.cuda.Stream()
    side.wait_stream(t.cuda.current_stream())
    with t.cuda.stream(side):
        step()
    t.cuda.current_stream().wait_stream(side)
    t.cuda.synchronize()
    graph = t.cuda.CUDAGraph()
    with t.cuda.graph(graph):
        step()
    for r in range(3):
        img.fill_(-1.0)
        for g in grads:
            g.fill_(123.0)
        t.cuda.synchronize()
        graph.replay()
        t.cuda.synchronize()
        assert np.array_equal(H.host(img), eager_img), f"replay {r} image"
        for k, g in enumerate(grads):
            assert np.array_equal(H.host(g).view(np.uint32), eager_grads[k].view(np.uint32)), \
                (r, k)
        # eager work between replays
        with t.no_grad():
            render_batch(*d, vmd, bgd, plan=_plan(sizes, cams))
        t.cuda.synchronize()


@pytest.mark.parametrize("bw,fd,bd", [(16, 64, 256), (16, 128, 64), (16, 256, 128),
                                      (8, 256, 64), (8, 128, 256)])
def test_batch_lds_depths_do_not_change_results(bw, fd, bd):
    """Batch depth is a pure performance knob: the image is bit-identical and
    the gradients equal within the atomics tolerance."""
    t = H.torch()
    params, sizes, cams = _layout()
    vm, bg = _views_arrays(cams)
    rng = np.random.default_rng(31)
    outs = []
    for depths in ((0, 0), (fd, bd)):
        plan = _plan(sizes, cams, bw)
        plan.set_depths(*depths)
        w = H.dev(rng.normal(size=(plan.num_pixels, 3)).astype(np.float32)) if not outs \
            else outs[0][2]
        d = _dev_params(params, requires_grad=True)
        flat = render_batch(*d, H.dev(vm), H.dev(bg), plan=plan)
        (flat * w).sum().backward()
        outs.append((H.host(flat), [H.host(x.grad) for x in d], w))
    assert np.array_equal(outs[0][0], outs[1][0])
    for a, b in zip(outs[0][1], outs[1][1]):
        ok, worst = H.grad_close(b, a, rtol=1e-4, scale_frac=1e-5)
        assert ok, worst


@pytest.mark.parametrize("bw", [16, 8])
def test_batch_queue_schedule_equals_grid(bw):
    """The work-queue schedule (tiles ordered longest list first, taken by a
    persistent grid) gives the flat grid's images, final_Ts and gradients
    bit for bit, on two calls in a row (the queue counters start over)."""
    t = H.torch()
    params, sizes, cams = _layout()
    vm, bg = _views_arrays(cams)
    out = {}
    for sched in ("grid", "queue", "queue"):
        plan = _plan(sizes, cams, bw)
        plan.set_schedule(sched)
        d = _dev_params(params, requires_grad=True)
        flat = render_batch(*d, H.dev(vm), H.dev(bg), plan=plan)
        w = t.from_numpy(np.random.default_rng(5).normal(
            size=tuple(flat.shape)).astype(np.float32)).cuda()
        (flat * w).sum().backward()
        plan.check()
        res = [H.host(flat.detach())] + [H.host(x.grad) for x in d]
        if sched in out:
            res2 = res
        else:
            out[sched] = res
    for a, b, c in zip(out["grid"], out["queue"], res2):
        assert np.array_equal(a.view(np.uint32), b.view(np.uint32))
        assert np.array_equal(a.view(np.uint32), c.view(np.uint32))
