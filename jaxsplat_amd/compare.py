"""Comparator harness: the reference's ``examples/benchmark.py`` for this build.

The reference times the forward ``render`` of jaxsplat against two torch CUDA
packages (gsplat, diff_gaussian_rasterization) on 1600x900 with 50,000
Gaussians and block 16.  Every iteration draws fresh parameters, and the
clock is the wall-clock around ``block_until_ready`` (``benchmark.py:15-37,
61-67,268-289``).  Those packages do not exist for ROCm, so the experiments
here are the three ways this build renders the same inputs:

* ``plain``       -- ``jaxsplat_amd.render``, one view per call (host read of
  the intersection count, as the reference);
* ``batched``     -- the batched path with one view (kernel launches only; its
  capacity is sized on the first call);
* ``xla_targets`` -- the four reference-compatible custom-call targets driven
  as XLA drives them (the reference's own call path, with its per-call
  binning).

All three run the same kernels, so their images must be identical; the
float64 oracle comparison of the same workload is a GPU test
(``tests/test_gpu_compare.py``).

Each experiment renders the same parameters.  The harness prints the average
milliseconds per render (``torch.cuda.synchronize`` after each, like the
reference's ``block_until_ready``), writes ``test-<name>.png``, and reports each
image's max-abs difference to the first experiment's and the fraction of
pixels that differ by more than 1e-4.

Usage: ``python -m jaxsplat_amd.compare [--iters 100] [--num_points 50000]
[--out DIR]``.
"""

from __future__ import annotations

import argparse
import json
import os
import time

import numpy as np


def init_params(rng, num_points):
    """``benchmark.py:236-254``: means U(-3,3), scales U(0,0.5), normalised
    N(0,1) quaternions, colors U(0,1), opacities U(0,1)."""
    means = rng.uniform(-3, 3, size=(num_points, 3))
    scales = rng.uniform(0, 0.5, size=(num_points, 3))
    q = rng.normal(size=(num_points, 4))
    q /= np.linalg.norm(q, axis=-1, keepdims=True)
    colors = rng.uniform(size=(num_points, 3))
    opac = rng.uniform(size=(num_points, 1))
    return [a.astype(np.float32) for a in (means, scales, q, colors, opac)]


def init_coeffs(img_shape=(900, 1600)):
    """``benchmark.py:257-279``: viewmat translating z by 8, background ones,
    f = (W/2, H/2), c = (W/2, H/2), glob_scale 1, clip 0.01, block 16."""
    H, W = img_shape
    vm = np.eye(4, dtype=np.float32)
    vm[2, 3] = 8.0
    return dict(viewmat=vm, background=np.ones(3, np.float32), img_shape=(H, W),
                f=(W / 2, H / 2), c=(W / 2, H / 2), glob_scale=1.0, clip_thresh=0.01,
                block_size=16)


class Experiment:
    name = ""

    def __init__(self):
        self.total = 0.0
        self.calls = 0

    def run(self, params):
        t0 = time.perf_counter()
        img = self._render(params)
        self.total += time.perf_counter() - t0
        self.calls += 1
        return img

    def avg_ms(self):
        return 1e3 * self.total / max(self.calls, 1)


class Plain(Experiment):
    name = "plain"

    def __init__(self, coeffs):
        super().__init__()
        import torch
        self.torch = torch
        self.kw = dict(coeffs)
        self.kw["viewmat"] = torch.from_numpy(coeffs["viewmat"]).cuda()
        self.kw["background"] = torch.from_numpy(coeffs["background"]).cuda()

    def _render(self, params):
        import jaxsplat_amd
        p = [self.torch.from_numpy(a).cuda() for a in params]
        self.torch.cuda.synchronize()
        t0 = time.perf_counter()
        img = jaxsplat_amd.render(*p, **self.kw)
        self.torch.cuda.synchronize()
        self.total += time.perf_counter() - t0
        return img.cpu().numpy()

    def run(self, params):          # times the render only, not the host-to-device copy
        img = self._render(params)
        self.calls += 1
        return img


class Batched(Plain):
    name = "batched"

    def __init__(self, coeffs):
        super().__init__(coeffs)
        from jaxsplat_amd import BatchRenderer, ViewSpec
        self.coeffs = coeffs
        self.r = None
        self.mk = lambda n: BatchRenderer([n], [ViewSpec(0, coeffs["img_shape"], coeffs["f"],
                                                         coeffs["c"])],
                                          glob_scale=coeffs["glob_scale"],
                                          clip_thresh=coeffs["clip_thresh"],
                                          block_width=coeffs["block_size"], check_every=1)

    def _render(self, params):
        torch = self.torch
        p = [torch.from_numpy(a).cuda() for a in params]
        if self.r is None:
            self.r = self.mk(params[0].shape[0])
        vm = self.kw["viewmat"][None]
        bg = self.kw["background"][None]
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        with torch.no_grad():
            flat = self.r(*p, vm, bg)
        torch.cuda.synchronize()
        self.total += time.perf_counter() - t0
        return flat.view(*self.coeffs["img_shape"], 3).cpu().numpy()


class XlaTargets(Plain):
    """The four reference-compatible custom-call targets driven as XLA drives
    them: a void** array of operands then results in the reference lowering
    order and the 80-byte descriptor as the opaque string."""
    name = "xla_targets"

    def __init__(self, coeffs):
        super().__init__(coeffs)
        from . import _lib
        self.lib = _lib

    def _render(self, params):
        import ctypes
        torch, L, c = self.torch, self.lib.lib(), self.kw
        N = params[0].shape[0]
        H, W = c["img_shape"]
        d = self.lib.make_descriptor(N, (H, W), c["f"], c["c"], c["glob_scale"],
                                     c["clip_thresh"], c["block_size"])
        opaque = ctypes.string_at(ctypes.addressof(d), 80)
        f32 = dict(dtype=torch.float32, device="cuda")
        i32 = dict(dtype=torch.int32, device="cuda")
        p = [torch.from_numpy(a).cuda() for a in params]
        proj = [torch.empty((N, 6), **f32), torch.empty((N, 2), **f32), torch.empty((N, 1), **f32),
                torch.empty((N, 1), **i32), torch.empty((N, 3), **f32), torch.empty((N, 1), **f32),
                torch.empty((N, 1), **i32), torch.empty((N, 1), **i32)]
        out = [torch.empty((H, W, 1), **f32), torch.empty((H, W, 1), **i32),
               torch.empty((H, W, 3), **f32)]
        s = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)

        def call(name, bufs):
            arr = (ctypes.c_void_p * len(bufs))(*[b.data_ptr() for b in bufs])
            getattr(L, name)(s, arr, opaque, 80)
            err = L.jaxsplat_last_error()
            if err:
                raise RuntimeError(err.decode())
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        call("jaxsplat_project_fwd", p[:3] + [c["viewmat"]] + proj)
        call("jaxsplat_rasterize_fwd", [p[3], p[4], c["background"], proj[1], proj[2], proj[3],
                                        proj[4], proj[7]] + out)
        torch.cuda.synchronize()
        self.total += time.perf_counter() - t0
        return out[2].cpu().numpy()


def _save(img, path):
    try:
        from PIL import Image
        Image.fromarray((np.clip(img, 0, 1) * 255).astype(np.uint8)).save(path)
    except ImportError:
        np.save(os.path.splitext(path)[0] + ".npy", img)


def main(iters=100, num_points=50_000, out_dir=".", img_shape=(900, 1600), seed=0,
         images_out=None):
    """Runs the experiments; returns the report dict.  ``images_out`` (a dict)
    receives each experiment's float image of the first parameters."""
    coeffs = init_coeffs(img_shape)
    exps = [Plain(coeffs), Batched(coeffs), XlaTargets(coeffs)]
    rng = np.random.default_rng(seed)
    first = init_params(rng, num_points)
    images = {}
    for e in exps:
        images[e.name] = e.run(first)
        _save(images[e.name], os.path.join(out_dir, f"test-{e.name}.png"))
    for it in range(iters):
        params = init_params(rng, num_points)
        for e in exps:
            e.run(params)
    if images_out is not None:
        images_out.update(images)
    base = images[exps[0].name]
    report = {}
    for e in exps:
        d = np.abs(images[e.name].astype(np.float64) - base)
        report[e.name] = {"avg_ms": round(e.avg_ms(), 4), "calls": e.calls,
                          "max_abs_vs_plain": float(d.max()),
                          "frac_px_over_1e-4": float((d.max(axis=-1) > 1e-4).mean())}
        print(f"{e.name} avg {e.avg_ms():.4f}ms  max|diff| vs {exps[0].name} {d.max():.3g}")
    return report


if __name__ == "__main__":
    ap = argparse.ArgumentParser(prog="python -m jaxsplat_amd.compare",
                                 description="Times and compares the render implementations")
    ap.add_argument("--iters", type=int, default=100)
    ap.add_argument("--num_points", type=int, default=50_000)
    ap.add_argument("--out", default=".")
    a = ap.parse_args()
    print(json.dumps(main(a.iters, a.num_points, a.out)))
