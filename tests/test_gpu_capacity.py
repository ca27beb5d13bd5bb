"""render() with a fixed intersection capacity (``IsectCapacity``): no host
synchronisation, so a fwd+bwd step can be captured in a HIP graph.

Bars (written in each test):
* image and all five gradients bit-identical to the default render (which
  reads the exact count back to the host), on the opacity-limited and on the
  reference's tile lists; ``check()`` returns the exact intersection total;
* a capacity below the total is reported by ``check()`` (ERR_CAPACITY), the
  capacity grows to fit, and the next render equals the default again;
* a render with no intersections gives the background and zero gradients;
* a HIP graph of forward + L2 gradient + backward, replayed several times
  with eager work and synchronisations in between and with new parameter
  values copied into its inputs before each replay, equals the eager render
  of those values bit for bit, image and gradients.
"""

import numpy as np
import pytest

import jaxsplat_amd
from jaxsplat_amd import IsectCapacity, _lib, _ops, scenes
from tests import helpers as H

pytestmark = pytest.mark.gpu


def _scene(n=60_000, seed=None):
    sc = scenes.make("hd1m", n=n, seed=seed)
    sc.cameras[0].img_shape = (360, 640)
    sc.cameras[0].f = (500.0, 500.0)
    sc.cameras[0].c = (320.0, 180.0)
    return sc


def _kw(sc, t):
    cam = sc.camera
    return dict(viewmat=t.from_numpy(cam.viewmat).cuda(),
                background=t.tensor([0.1, 0.2, 0.3], dtype=t.float32, device="cuda"),
                img_shape=cam.img_shape, f=cam.f, c=cam.c, glob_scale=sc.glob_scale,
                clip_thresh=sc.clip_thresh, block_size=16)


def _fwd_bwd(params, kw, w, cap=None):
    t = H.torch()
    ps = [p.detach().clone().requires_grad_(True) for p in params]
    img = jaxsplat_amd.render(*ps, **kw, isect_capacity=cap)
    (img * w).sum().backward()
    return H.host(img), [H.host(p.grad) for p in ps]


def _same(a, b):
    return np.array_equal(np.asarray(a).view(np.uint32), np.asarray(b).view(np.uint32))


@pytest.mark.parametrize("tight", [True, False])
def test_capacity_render_equals_default(tight, monkeypatch):
    t = H.torch()
    monkeypatch.setattr(_ops, "TIGHT_TILES", tight)
    sc = _scene()
    params = [H.dev(a) for a in sc.params()]
    kw = _kw(sc, t)
    w = t.from_numpy(np.random.default_rng(3).normal(
        size=(*sc.camera.img_shape, 3)).astype(np.float32)).cuda()
    img0, g0 = _fwd_bwd(params, kw, w)
    cap = IsectCapacity(4_000_000)
    img1, g1 = _fwd_bwd(params, kw, w, cap)
    total = cap.check()
    assert total > 10_000
    assert _same(img0, img1)
    for a, b in zip(g0, g1):
        assert _same(a, b)
    # the exact total is that of the lists the render bins into
    xys, depths, radii, conics, cum, comp, rects = _ops.project_for_render(
        *params[:3], params[4], kw["viewmat"], img_shape=kw["img_shape"], f=kw["f"], c=kw["c"],
        glob_scale=1.0, clip_thresh=0.01, block_width=16)
    if tight:
        assert total == int(H.host(cum)[-1, 0])


def test_capacity_overflow_reported_and_grown():
    t = H.torch()
    sc = _scene()
    params = [H.dev(a) for a in sc.params()]
    kw = _kw(sc, t)
    w = t.ones((*sc.camera.img_shape, 3), dtype=t.float32, device="cuda")
    img0, g0 = _fwd_bwd(params, kw, w)
    cap = IsectCapacity(4_000_000)
    _fwd_bwd(params, kw, w, cap)
    total = cap.check()
    small = IsectCapacity(total // 3)
    _fwd_bwd(params, kw, w, small)
    with pytest.raises(_lib.JaxsplatError) as e:
        small.check()
    assert e.value.code == _lib.ERR_CAPACITY
    assert small.capacity >= total
    img2, g2 = _fwd_bwd(params, kw, w, small)
    assert small.check() == total
    assert _same(img0, img2)
    for a, b in zip(g0, g2):
        assert _same(a, b)


def test_capacity_zero_intersections():
    t = H.torch()
    sc = _scene(n=500)
    params = [H.dev(a) for a in sc.params()]
    params[0] = params[0] * t.tensor([1.0, 1.0, -1.0], device="cuda")   # all behind the camera
    kw = _kw(sc, t)
    w = t.ones((*sc.camera.img_shape, 3), dtype=t.float32, device="cuda")
    cap = IsectCapacity(1000)
    img, grads = _fwd_bwd(params, kw, w, cap)
    assert cap.check() == 0
    assert np.array_equal(img, np.broadcast_to(np.array([0.1, 0.2, 0.3], np.float32), img.shape))
    assert all(not g.any() for g in grads)


def test_capacity_hip_graph_replays_match_eager():
    t = H.torch()
    sc = _scene()
    kw = _kw(sc, t)
    target = t.from_numpy(scenes.target_image(sc.camera.img_shape, 5)).cuda()
    scale = 2.0 / target.numel()
    static = [H.dev(a).requires_grad_(True) for a in sc.params()]
    cap = IsectCapacity(4_000_000)

    def step():
        img = jaxsplat_amd.render(*static, **kw, isect_capacity=cap)
        img.backward((img.detach() - target) * scale)
        return img

    side = t.cuda.Stream()
    side.wait_stream(t.cuda.current_stream())
    with t.cuda.stream(side):
        for _ in range(2):
            for p in static:
                p.grad = None
            step()
    t.cuda.current_stream().wait_stream(side)
    t.cuda.synchronize()
    for p in static:
        p.grad = None
    graph = t.cuda.CUDAGraph()
    with t.cuda.graph(graph):
        g_img = step()
    g_grads = [p.grad for p in static]
    cap.check()
    for r in range(3):
        other = _scene(seed=2000 + r)               # new parameter values for this replay
        with t.no_grad():
            for p, a in zip(static, other.params()):
                p.copy_(t.from_numpy(a))
        print(f"[capacity-graph] replay {r}", flush=True)
        graph.replay()
        t.cuda.synchronize()
        got_img = H.host(g_img)
        got_grads = [H.host(g) for g in g_grads]
        assert cap.check() > 0
        # eager render of the same values (the default path, with its host
        # read of the count) between the replays; the cotangent is the
        # graph's, (img - target) * scale, formed the same way
        w = (t.from_numpy(got_img).cuda() - target) * scale
        exp_img, exp_grads = _fwd_bwd([p.detach() for p in static], kw, w)
        assert _same(got_img, exp_img), r
        for a, b in zip(got_grads, exp_grads):
            assert _same(a, b), r
