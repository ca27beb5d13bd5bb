"""Batched rendering: many views of one or more scenes per call.

The reference renders one view per ``render`` call (``jaxsplat/__init__.py:6-70``),
each call sorting twice and synchronising with the host.  Here a
``BatchPlan`` fixes the layout of a batch once (scenes stored end to end in
the parameter arrays, views grouped by scene), and ``render_batch`` runs
projection, binning and rasterization of every view -- and the backward of
every view -- as kernel launches only, with no allocation and no host
synchronisation inside, so a whole step can be captured in one HIP graph.

Per view the image equals the plain ``render`` of that view with the same
``block_width`` (same kernels, same per-tile order).  Gradients of a scene's
Gaussians are summed over its views.

Intersection capacity: the workspace is sized for ``isect_capacity``
(view, tile) intersections, counted on the opacity-limited tile rectangles
that ``rasterize`` bins into (the reference's rectangles without the tiles
where every pixel fails the alpha test).  ``BatchPlan.check()`` (a host synchronisation)
reports the exact count and raises ``JaxsplatError`` with code
``ERR_CAPACITY`` when a forward overflowed; ``BatchRenderer`` sizes the
capacity on its first call and grows it when a check finds an overflow.
"""

from __future__ import annotations

import ctypes
from dataclasses import dataclass

import torch

from . import _lib


@dataclass(frozen=True)
class ViewSpec:
    """One camera of a batch: the scene it looks at, image (H, W), focal
    lengths (fx, fy) and principal point (cx, cy) in pixels."""
    scene: int
    img_shape: tuple
    f: tuple
    c: tuple


def _p(t):
    return ctypes.c_void_p(t.data_ptr()) if t is not None else None


def _stream(device):
    return ctypes.c_void_p(torch.cuda.current_stream(device).cuda_stream)


class BatchPlan:
    """Layout of a batch and its device workspace (C ABI ``jaxsplat_batch``).

    scenes: Gaussian count of each scene (rows stored consecutively).
    views: ``ViewSpec`` per view, grouped by ascending scene.
    """

    def __init__(self, scenes, views, *, glob_scale=1.0, clip_thresh=0.01, block_width=16,
                 isect_capacity, device=None):
        self.device = torch.device(device if device is not None else "cuda")
        self.scene_sizes = tuple(int(n) for n in scenes)
        self.views = tuple(views)
        self.glob_scale = float(glob_scale)
        self.clip_thresh = float(clip_thresh)
        self.block_width = int(block_width)
        self.isect_capacity = int(isect_capacity)
        sc = (_lib.Scene * len(self.scene_sizes))()
        first = 0
        for i, n in enumerate(self.scene_sizes):
            sc[i].first_gaussian, sc[i].num_gaussians = first, n
            first += n
        vw = (_lib.View * len(self.views))()
        for i, v in enumerate(self.views):
            H, W = v.img_shape
            vw[i].scene, vw[i].img_h, vw[i].img_w = int(v.scene), int(H), int(W)
            vw[i].fx, vw[i].fy = float(v.f[0]), float(v.f[1])
            vw[i].cx, vw[i].cy = float(v.c[0]), float(v.c[1])
        L = _lib.lib()
        h = ctypes.c_void_p()
        with torch.cuda.device(self.device):
            code = L.jaxsplat_batch_create(ctypes.byref(h), sc, len(self.scene_sizes), vw,
                                           len(self.views), self.glob_scale, self.clip_thresh,
                                           self.block_width, self.isect_capacity)
        if code != _lib.OK:
            raise ValueError(L.jaxsplat_last_error().decode())
        self._h = h
        self.num_gaussians = int(L.jaxsplat_batch_num_gaussians(h))
        self.num_pixels = int(L.jaxsplat_batch_num_pixels(h))
        self.pixel_offsets = [int(L.jaxsplat_batch_pixel_offset(h, v))
                              for v in range(len(self.views) + 1)]
        self.workspace_bytes = int(L.jaxsplat_batch_workspace_size(h))
        self.workspace = torch.empty(self.workspace_bytes, dtype=torch.uint8, device=self.device)
        self.generation = 0          # bumped by every forward; the backward checks it
        self.depths = (0, 0)
        self.schedule = "grid"

    def __del__(self):
        h = getattr(self, "_h", None)
        if h is not None and h.value:
            try:
                _lib.lib().jaxsplat_batch_destroy(h)
            except Exception:
                pass
            self._h = None

    @property
    def num_views(self):
        return len(self.views)

    def set_depths(self, fwd_depth=0, bwd_depth=0):
        """Gaussians staged in LDS per batch by the rasterize forward / backward
        (64, 128 or 256; 0 = workgroup size).  Results do not depend on it."""
        _lib.check(_lib.lib().jaxsplat_batch_set_depths(self._h, int(fwd_depth), int(bwd_depth)))
        self.depths = (int(fwd_depth), int(bwd_depth))

    def set_schedule(self, schedule="grid"):
        """How the rasterizer hands (view, tile) items to workgroups: "grid"
        (one workgroup per tile) or "queue" (tiles ordered longest list
        first, taken by a persistent grid from a device counter).  Results do
        not depend on it."""
        try:
            code = _lib.SCHEDULE[schedule]
        except KeyError:
            raise ValueError(f"schedule must be one of {sorted(_lib.SCHEDULE)}, "
                             f"got {schedule!r}") from None
        _lib.check(_lib.lib().jaxsplat_batch_set_schedule(self._h, code))
        self.schedule = schedule

    def split(self, flat):
        """Per-view (H, W, C) views of a (num_pixels, C) buffer."""
        out = []
        for v, spec in enumerate(self.views):
            a, b = self.pixel_offsets[v], self.pixel_offsets[v + 1]
            out.append(flat[a:b].view(spec.img_shape[0], spec.img_shape[1], -1))
        return out

    def check(self):
        """Exact intersection count of the last forward (host sync); raises
        ``JaxsplatError`` (ERR_CAPACITY) if it overflowed the capacity."""
        n = ctypes.c_int64(0)
        with torch.cuda.device(self.device):
            code = _lib.lib().jaxsplat_batch_check(self._h, _p(self.workspace), ctypes.byref(n),
                                                   _stream(self.device))
        self.last_intersections = int(n.value)
        _lib.check(code)
        return int(n.value)


def _expect(name, t, shape):
    if not isinstance(t, torch.Tensor) or tuple(t.shape) != tuple(shape) or \
            t.dtype != torch.float32 or t.device.type != "cuda":
        got = (tuple(t.shape), t.dtype, t.device) if isinstance(t, torch.Tensor) else type(t)
        raise TypeError(f"{name}: expected float32 GPU tensor of shape {tuple(shape)}, got {got}")


class _RenderBatch(torch.autograd.Function):
    @staticmethod
    def forward(ctx, plan, adjoint, means3d, scales, quats, colors, opacities, viewmats,
                backgrounds):
        dev = plan.device
        P = plan.num_pixels
        img = torch.empty((P, 3), dtype=torch.float32, device=dev)
        final_Ts = torch.empty((P,), dtype=torch.float32, device=dev)
        final_idx = torch.empty((P,), dtype=torch.int32, device=dev)
        with torch.cuda.device(dev):
            _lib.check(_lib.lib().jaxsplat_batch_forward(
                plan._h, _p(means3d), _p(scales), _p(quats), _p(colors), _p(opacities),
                _p(viewmats), _p(backgrounds), _p(final_Ts), _p(final_idx), _p(img),
                _p(plan.workspace), plan.workspace_bytes, _stream(dev)))
        plan.generation += 1
        ctx.plan = plan
        ctx.generation = plan.generation
        ctx.adjoint = adjoint
        ctx.save_for_backward(means3d, scales, quats, colors, opacities, viewmats, backgrounds,
                              final_Ts, final_idx)
        return img

    @staticmethod
    def backward(ctx, v_img):
        plan = ctx.plan
        if plan.generation != ctx.generation:
            raise RuntimeError("render_batch: another forward of this BatchPlan ran before this "
                               "backward; its residuals were overwritten")
        means3d, scales, quats, colors, opacities, viewmats, backgrounds, final_Ts, final_idx = \
            ctx.saved_tensors
        v_img = v_img.contiguous()
        v_means = torch.empty_like(means3d)
        v_scales = torch.empty_like(scales)
        v_quats = torch.empty_like(quats)
        v_colors = torch.empty_like(colors)
        v_opac = torch.empty_like(opacities)
        dev = plan.device
        with torch.cuda.device(dev):
            _lib.check(_lib.lib().jaxsplat_batch_backward(
                plan._h, _p(means3d), _p(scales), _p(quats), _p(colors), _p(opacities),
                _p(viewmats), _p(backgrounds), _p(final_Ts), _p(final_idx), _p(v_img),
                _p(v_means), _p(v_scales), _p(v_quats), _p(v_colors), _p(v_opac),
                ctx.adjoint, _p(plan.workspace), plan.workspace_bytes, _stream(dev)))
        return None, None, v_means, v_scales, v_quats, v_colors, v_opac, None, None


def render_batch(means3d, scales, quats, colors, opacities, viewmats, backgrounds, *, plan,
                 adjoint="reference"):
    """Render every view of ``plan``: returns the (num_pixels, 3) image buffer
    (``plan.split`` gives per-view (H, W, 3) views).  Differentiable in
    means3d, scales, quats, colors and opacities (all (G, k) float32, scenes
    end to end); viewmats (V, 4, 4) row-major world-to-camera; backgrounds
    (V, 3)."""
    try:
        adj = _lib.ADJOINT[adjoint]
    except KeyError:
        raise ValueError(f"adjoint must be one of {sorted(_lib.ADJOINT)}, got {adjoint!r}") \
            from None
    G, V = plan.num_gaussians, plan.num_views
    _expect("means3d", means3d, (G, 3))
    _expect("scales", scales, (G, 3))
    _expect("quats", quats, (G, 4))
    _expect("colors", colors, (G, 3))
    _expect("opacities", opacities, (G, 1))
    _expect("viewmats", viewmats, (V, 4, 4))
    _expect("backgrounds", backgrounds, (V, 3))
    return _RenderBatch.apply(plan, adj, means3d.contiguous(), scales.contiguous(),
                              quats.contiguous(), colors.contiguous(), opacities.contiguous(),
                              viewmats.contiguous(), backgrounds.contiguous())


class BatchRenderer:
    """A ``BatchPlan`` whose intersection capacity is sized from the data.

    The first call renders, reads the exact intersection count and, if it
    overflowed the initial guess, re-plans with ``headroom`` times the count
    and renders again.  Later calls check the count every ``check_every``
    calls (a host synchronisation; default every call) and re-plan and render
    again on overflow, so a returned image is never one cut off at the
    capacity.  ``check_every=0`` skips the checks after the first call: an
    overflow then leaves that call's images and gradients invalid, and the
    caller must read ``plan.check()`` (or the device status word,
    ``jaxsplat_batch_status_word``) itself -- the form for loops captured in
    a HIP graph, where no host synchronisation may run.
    """

    def __init__(self, scenes, views, *, glob_scale=1.0, clip_thresh=0.01, block_width=16,
                 headroom=1.25, check_every=1, device=None, initial_capacity=None):
        self.scenes = tuple(scenes)
        self.views = tuple(views)
        self.kw = dict(glob_scale=glob_scale, clip_thresh=clip_thresh, block_width=block_width,
                       device=device)
        self.headroom = float(headroom)
        self.check_every = int(check_every)
        n_items = sum(self.scenes[v.scene] for v in self.views)
        self.capacity = int(initial_capacity or max(1 << 16, 4 * n_items))
        self.plan = None
        self.calls = 0

    def _replan(self, capacity):
        self.capacity = int(min(capacity, 0x7FFFFF00))
        self.plan = BatchPlan(self.scenes, self.views, isect_capacity=self.capacity, **self.kw)

    def __call__(self, means3d, scales, quats, colors, opacities, viewmats, backgrounds,
                 adjoint="reference"):
        args = (means3d, scales, quats, colors, opacities, viewmats, backgrounds)
        first = self.plan is None
        if first:
            self._replan(self.capacity)
        self.calls += 1
        out = render_batch(*args, plan=self.plan, adjoint=adjoint)
        if first or (self.check_every > 0 and self.calls % self.check_every == 0):
            try:
                self.plan.check()
            except _lib.JaxsplatError as e:
                if e.code != _lib.ERR_CAPACITY:
                    raise
                n = self.plan.last_intersections
                self._replan(max(int(n * self.headroom), n + 1))
                out = render_batch(*args, plan=self.plan, adjoint=adjoint)
                self.plan.check()
        return out

    def split(self, flat):
        return self.plan.split(flat)
