"""The fused render node (``_ops.render_fused``, ``jaxsplat_render_backward``) on the GPU.

``render`` runs the projection with the opacity-limited rectangles, the
binning and the rasterize forward as one autograd node; its backward writes
the staged rasterize records and then, in one per-Gaussian pass, sums them and
runs the projection VJP in registers.  The same sums and the same VJP code as
the two-node composition (``project_for_render`` + ``rasterize`` with
``tile_rects``, i.e. ``jaxsplat_rasterize_backward_staged_rects`` then
``jaxsplat_project_backward``), so the bar is bit-identical images and
gradients (all five), in both adjoint modes, for 1, 2 and 3 channels, with and
without a fixed intersection capacity, and at the 1M-Gaussian 1080p workload.
"""

import numpy as np
import pytest

from jaxsplat_amd import IsectCapacity, _ops, render, scenes
from tests import helpers as H

pytestmark = pytest.mark.gpu


def _run(params, cam, bg, w, fused, adjoint, capacity=None):
    _ops.FUSED_RENDER = fused
    try:
        d = [H.dev(x).requires_grad_(True) for x in params]
        img = render(*d, viewmat=H.dev(cam.viewmat), background=bg, img_shape=cam.img_shape,
                     f=cam.f, c=cam.c, glob_scale=1.0, clip_thresh=0.01, block_size=16,
                     adjoint=adjoint, isect_capacity=capacity)
        (img * w).sum().backward()
        return H.host(img), [H.host(x.grad) for x in d]
    finally:
        _ops.FUSED_RENDER = True


def _bitwise_equal(a, b):
    return np.array_equal(np.asarray(a).view(np.uint32), np.asarray(b).view(np.uint32))


@pytest.mark.parametrize("adjoint", ["reference", "consistent"])
@pytest.mark.parametrize("nch", [3, 1, 2])
def test_fused_render_bitwise_equals_composition(adjoint, nch):
    sc = scenes.make("small", n=4000)
    cam = sc.camera
    rng = np.random.default_rng(31)
    colors = rng.uniform(0, 1, (sc.n, nch)).astype(np.float32)
    params = [sc.means3d, sc.scales, sc.quats, colors, sc.opacities]
    bg = H.dev(rng.uniform(0, 1, nch).astype(np.float32))
    H_, W = cam.img_shape
    w = H.dev(rng.normal(size=(H_, W, nch)).astype(np.float32))
    img_f, g_f = _run(params, cam, bg, w, True, adjoint)
    img_c, g_c = _run(params, cam, bg, w, False, adjoint)
    assert _bitwise_equal(img_f, img_c)
    for k, (a, b) in enumerate(zip(g_f, g_c)):
        assert _bitwise_equal(a, b), f"gradient {k}"
    assert np.abs(g_f[0]).max() > 0 and np.abs(g_f[3]).max() > 0


def test_fused_render_with_capacity():
    sc = scenes.make("small", n=4000)
    cam = sc.camera
    rng = np.random.default_rng(32)
    params = sc.params()
    bg = H.dev(np.zeros(3, np.float32))
    H_, W = cam.img_shape
    w = H.dev(rng.normal(size=(H_, W, 3)).astype(np.float32))
    cap = IsectCapacity(1 << 20)
    img_f, g_f = _run(params, cam, bg, w, True, "reference", cap)
    cap.check()
    img_c, g_c = _run(params, cam, bg, w, False, "reference")
    assert _bitwise_equal(img_f, img_c)
    for k, (a, b) in enumerate(zip(g_f, g_c)):
        assert _bitwise_equal(a, b), f"gradient {k}"


def test_fused_render_hd1m_bitwise():
    sc = scenes.make("hd1m")
    cam = sc.camera
    print("fused render: hd1m, 1M Gaussians at 1080p", flush=True)
    rng = np.random.default_rng(33)
    params = sc.params()
    bg = H.dev(np.zeros(3, np.float32))
    H_, W = cam.img_shape
    w = H.dev(rng.normal(size=(H_, W, 3)).astype(np.float32))
    img_f, g_f = _run(params, cam, bg, w, True, "reference")
    img_c, g_c = _run(params, cam, bg, w, False, "reference")
    assert _bitwise_equal(img_f, img_c)
    for k, (a, b) in enumerate(zip(g_f, g_c)):
        assert _bitwise_equal(a, b), f"gradient {k}"
