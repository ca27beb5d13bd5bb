"""Autotuner of the batched render: tile size and LDS batch depths.

Knobs, timed on the caller's own batch (the cost of a batch depends on its
Gaussians, cameras and image sizes):

* ``fwd_depth`` and ``bwd_depth`` in {64, 128, 256}: Gaussians staged in LDS
  per batch by the rasterize forward and backward.  Results do not depend on
  them (every pixel walks the same entries in the same order).
* ``schedule``: the rasterizer's one workgroup per tile ("grid"), or a
  persistent grid taking the tiles longest list first from a device counter
  ("queue"), timed at the chosen depths.  Results do not depend on it.
* ``block_width`` over the candidates the caller allows (8 and/or 16).  The
  tile side is part of the reference's semantics: a Gaussian joins the lists
  of the tiles its 3-sigma rectangle touches, so a different tile size can
  add or drop contributions near tile borders (below the 1/255 alpha cut only
  where the rectangle is tight).  The tuner therefore only picks among the
  block widths passed in; the default is the reference's 16.

For each block width the intersection capacity is sized with one forward,
then the forward is timed at each forward depth and the backward at each
backward depth (HIP events on the launch stream, median of ``iters`` runs).
The choice minimising forward + backward time is cached in memory per batch
layout.
"""

from __future__ import annotations

import ctypes
from dataclasses import dataclass, field

import numpy as np
import torch

from . import _lib
from .batch import BatchPlan

DEPTHS = (64, 128, 256)
_CACHE = {}


@dataclass
class Tuning:
    block_width: int
    fwd_depth: int
    bwd_depth: int
    capacity: int
    ms: float                       # forward + backward at the chosen setting
    table: dict = field(default_factory=dict)   # (bw, "fwd"|"bwd"[_queue], depth) -> ms
    schedule: str = "grid"


def _p(t):
    return ctypes.c_void_p(t.data_ptr()) if t is not None else None


def layout_key(scenes, views, glob_scale, clip_thresh, block_widths):
    return (tuple(int(n) for n in scenes),
            tuple((v.scene, tuple(v.img_shape), tuple(v.f), tuple(v.c)) for v in views),
            float(glob_scale), float(clip_thresh), tuple(block_widths))


def _median_ms(fn, iters, stream):
    times = []
    for _ in range(iters):
        e0 = torch.cuda.Event(enable_timing=True)
        e1 = torch.cuda.Event(enable_timing=True)
        e0.record(stream)
        fn()
        e1.record(stream)
        e1.synchronize()
        times.append(e0.elapsed_time(e1))
    return float(np.median(times))


def autotune(scenes, views, means3d, scales, quats, colors, opacities, viewmats, backgrounds, *,
             block_widths=(16,), depths=DEPTHS, schedules=("grid", "queue"), iters=3,
             glob_scale=1.0, clip_thresh=0.01, headroom=1.1, device=None, cache=True):
    """Pick (block_width, fwd_depth, bwd_depth) for a batch; see the module
    docstring.  Arguments are those of ``BatchPlan`` and ``render_batch``."""
    key = layout_key(scenes, views, glob_scale, clip_thresh, block_widths) + (tuple(schedules),)
    if cache and key in _CACHE:
        return _CACHE[key]
    dev = torch.device(device) if device is not None else means3d.device
    L = _lib.lib()
    args = [x.contiguous() for x in (means3d, scales, quats, colors, opacities, viewmats,
                                     backgrounds)]
    stream = torch.cuda.current_stream(dev)
    s = ctypes.c_void_p(stream.cuda_stream)
    grads = [torch.empty_like(x) for x in args[:5]]
    table, best = {}, None
    for bw in block_widths:
        n_items = sum(int(scenes[v.scene]) for v in views)
        cap = max(1 << 16, 4 * n_items)
        while True:                       # size the capacity with one forward
            plan = BatchPlan(scenes, views, glob_scale=glob_scale, clip_thresh=clip_thresh,
                             block_width=bw, isect_capacity=cap, device=dev)
            P = plan.num_pixels
            img = torch.empty((P, 3), dtype=torch.float32, device=dev)
            fT = torch.empty((P,), dtype=torch.float32, device=dev)
            fI = torch.empty((P,), dtype=torch.int32, device=dev)

            def fwd():
                _lib.check(L.jaxsplat_batch_forward(
                    plan._h, *[_p(x) for x in args], _p(fT), _p(fI), _p(img),
                    _p(plan.workspace), plan.workspace_bytes, s))
            fwd()
            try:
                plan.check()
                break
            except _lib.JaxsplatError as e:
                if e.code != _lib.ERR_CAPACITY:
                    raise
                cap = int(plan.last_intersections * headroom) + 1024
        v_img = torch.full((P, 3), 1e-3, dtype=torch.float32, device=dev)

        def bwd():
            _lib.check(L.jaxsplat_batch_backward(
                plan._h, *[_p(x) for x in args], _p(fT), _p(fI), _p(v_img),
                *[_p(g) for g in grads], 0, _p(plan.workspace), plan.workspace_bytes, s))
        for d in depths:
            plan.set_depths(d, 0)
            fwd()                         # warm
            table[(bw, "fwd", d)] = _median_ms(fwd, iters, stream)
        fd = min(depths, key=lambda d: table[(bw, "fwd", d)])
        plan.set_depths(fd, 0)
        fwd()
        for d in depths:
            plan.set_depths(fd, d)
            bwd()                         # warm
            table[(bw, "bwd", d)] = _median_ms(bwd, iters, stream)
        bd = min(depths, key=lambda d: table[(bw, "bwd", d)])
        ms = table[(bw, "fwd", fd)] + table[(bw, "bwd", bd)]
        sched = "grid"
        if "queue" in schedules:
            # the work queue at the chosen depths
            plan.set_depths(fd, bd)
            plan.set_schedule("queue")
            fwd()
            table[(bw, "fwd_queue", fd)] = _median_ms(fwd, iters, stream)
            bwd()
            table[(bw, "bwd_queue", bd)] = _median_ms(bwd, iters, stream)
            qms = table[(bw, "fwd_queue", fd)] + table[(bw, "bwd_queue", bd)]
            if qms < ms:
                ms, sched = qms, "queue"
        if best is None or ms < best.ms:
            best = Tuning(bw, fd, bd, plan.isect_capacity, ms, schedule=sched)
        del plan
    best.table = table
    if cache:
        _CACHE[key] = best
    return best


def tuned_plan(tuning, scenes, views, *, glob_scale=1.0, clip_thresh=0.01, device=None):
    """A ``BatchPlan`` with a tuning's block width, depths and capacity."""
    plan = BatchPlan(scenes, views, glob_scale=glob_scale, clip_thresh=clip_thresh,
                     block_width=tuning.block_width, isect_capacity=tuning.capacity,
                     device=device)
    plan.set_depths(tuning.fwd_depth, tuning.bwd_depth)
    plan.set_schedule(tuning.schedule)
    return plan
