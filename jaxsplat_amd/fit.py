"""Fitting 3D Gaussians to one image on the device: the training loop of the
reference's ``examples/single_image.py`` (``main`` :10-69, ``init`` :72-140,
``render_fn`` :143-166).

The reference steps a Python loop: ``jax.value_and_grad`` of
``mean((render(activated params) - gt)^2)`` (activations: ``quats / |quats|``,
``sigmoid(colors)``, ``sigmoid(opacities)``; means and scales raw), then
``optax.adam(lr)``.  Here one step is five kernel launches groups on one
stream -- activations, batched render forward, L2 gradient, batched render
backward, activation VJP + Adam -- with no allocation and no host
synchronisation, captured once in a HIP graph and replayed.

A render whose intersections overflow the plan's capacity never updates the
parameters: the Adam kernel reads the render's device status word and skips
the step.  ``ImageFit`` checks the status every ``check_every`` steps, grows
the capacity and re-captures the graph when needed.
"""

from __future__ import annotations

import ctypes
import math

import numpy as np
import torch

from . import _lib
from .batch import BatchPlan, ViewSpec


def _p(t):
    return ctypes.c_void_p(t.data_ptr()) if t is not None else None


def init_params(num_points, seed=0):
    """The reference initialisation (``single_image.py:72-140``) as float32
    NumPy arrays: means U([-6,-6,-1], [6,6,1]); scales U(0, 0.5); uniformly
    random unit quaternions from (u, v, w) ~ U(0,1)^3; raw colors U(0,1); raw
    opacities U(0.5, 1).  Drawn from ``numpy.random.default_rng(seed)`` (the
    reference draws from ``jax.random.key(0)``; same distributions, another
    stream)."""
    rng = np.random.default_rng(seed)
    n = int(num_points)
    means = rng.uniform([-6.0, -6.0, -1.0], [6.0, 6.0, 1.0], size=(n, 3))
    scales = rng.uniform(0.0, 0.5, size=(n, 3))
    u, v, w = rng.uniform(size=(3, n, 1))
    quats = np.hstack([np.sqrt(1 - u) * np.sin(2 * np.pi * v),
                       np.sqrt(1 - u) * np.cos(2 * np.pi * v),
                       np.sqrt(u) * np.sin(2 * np.pi * w),
                       np.sqrt(u) * np.cos(2 * np.pi * w)])
    colors = rng.uniform(size=(n, 3))
    opacities = rng.uniform(0.5, 1.0, size=(n, 1))
    return tuple(a.astype(np.float32) for a in (means, scales, quats, colors, opacities))


def camera(img_shape):
    """The reference camera (``single_image.py:103-124``): viewmat translating
    z by 8, f = (W/2, H/2), c = (W/2, H/2)."""
    H, W = img_shape
    vm = np.eye(4, dtype=np.float32)
    vm[2, 3] = 8.0
    return vm, (W / 2.0, H / 2.0), (W / 2.0, H / 2.0)


def psnr(img, target):
    mse = float(((img.float() - target.float()) ** 2).mean())
    return 10.0 * math.log10(1.0 / max(mse, 1e-20))


class ImageFit:
    """Adam fit of ``num_points`` Gaussians to ``target`` ((H, W, 3) float32 in
    [0, 1], on the GPU).  ``step(k)`` runs k training steps (graph replays
    when ``graph=True``); ``loss()`` is the loss of the last step's render."""

    def __init__(self, target, num_points=50_000, lr=0.01, seed=0, block_width=16,
                 betas=(0.9, 0.999), eps=1e-8, adjoint="reference", graph=True, check_every=25,
                 headroom=1.5, params=None):
        if not (isinstance(target, torch.Tensor) and target.dim() == 3 and target.shape[2] == 3
                and target.dtype == torch.float32 and target.device.type == "cuda"):
            raise TypeError("target must be an (H, W, 3) float32 GPU tensor")
        self.device = target.device
        self.target = target.contiguous()
        self.img_shape = tuple(target.shape[:2])
        H, W = self.img_shape
        raw = init_params(num_points, seed) if params is None else params
        dev = self.device
        self.means, self.scales, self.quats_raw, self.colors_raw, self.opac_raw = (
            torch.from_numpy(np.ascontiguousarray(a, np.float32)).to(dev) for a in raw)
        self.n = self.means.shape[0]
        vm, f, c = camera(self.img_shape)
        self.viewmats = torch.from_numpy(vm[None]).to(dev)
        self.backgrounds = torch.zeros((1, 3), dtype=torch.float32, device=dev)
        self.view = ViewSpec(0, self.img_shape, f, c)
        self.block_width = int(block_width)
        self.lr, self.b1, self.b2, self.eps = float(lr), float(betas[0]), float(betas[1]), \
            float(eps)
        self.adjoint = _lib.ADJOINT[adjoint]
        f32 = dict(dtype=torch.float32, device=dev)
        self.quats = torch.empty((self.n, 4), **f32)
        self.colors = torch.empty((self.n, 3), **f32)
        self.opac = torch.empty((self.n, 1), **f32)
        self.grads = [torch.empty((self.n, k), **f32) for k in (3, 3, 4, 3, 1)]
        self.m1 = torch.zeros((self.n, 14), **f32)
        self.m2 = torch.zeros((self.n, 14), **f32)
        self.step_dev = torch.zeros((1,), dtype=torch.int32, device=dev)
        P = H * W
        self.img = torch.empty((P, 3), **f32)
        self.final_Ts = torch.empty((P,), **f32)
        self.final_idx = torch.empty((P,), dtype=torch.int32, device=dev)
        self.v_img = torch.empty((P, 3), **f32)
        self.loss_sum = torch.zeros((1,), **f32)
        self.use_graph = bool(graph)
        self.check_every = int(check_every)
        self.headroom = float(headroom)
        self.graph = None
        self.calls = 0
        self.replans = 0
        self._plan(max(1 << 16, 16 * self.n))

    # -- plan / graph ---------------------------------------------------------
    def _plan(self, capacity):
        self.plan = BatchPlan([self.n], [self.view], glob_scale=1.0, clip_thresh=0.01,
                              block_width=self.block_width,
                              isect_capacity=int(min(capacity, 0x7FFFFF00)), device=self.device)
        self.status = ctypes.c_void_p(_lib.lib().jaxsplat_batch_status_word(
            self.plan._h, _p(self.plan.workspace)))
        self.graph = None

    def _launch(self):
        L = _lib.lib()
        s = ctypes.c_void_p(torch.cuda.current_stream(self.device).cuda_stream)
        pl = self.plan
        n = self.n
        _lib.check(L.jaxsplat_fit_activate(n, _p(self.quats_raw), _p(self.colors_raw),
                                           _p(self.opac_raw), _p(self.quats), _p(self.colors),
                                           _p(self.opac), s))
        act = (_p(self.means), _p(self.scales), _p(self.quats), _p(self.colors), _p(self.opac))
        _lib.check(L.jaxsplat_batch_forward(pl._h, *act, _p(self.viewmats), _p(self.backgrounds),
                                            _p(self.final_Ts), _p(self.final_idx), _p(self.img),
                                            _p(pl.workspace), pl.workspace_bytes, s))
        _lib.check(L.jaxsplat_fit_l2(self.img.numel(), _p(self.img), _p(self.target),
                                     _p(self.v_img), _p(self.loss_sum), s))
        _lib.check(L.jaxsplat_batch_backward(pl._h, *act, _p(self.viewmats),
                                             _p(self.backgrounds), _p(self.final_Ts),
                                             _p(self.final_idx), _p(self.v_img),
                                             *[_p(g) for g in self.grads], self.adjoint,
                                             _p(pl.workspace), pl.workspace_bytes, s))
        _lib.check(L.jaxsplat_fit_adam(n, _p(self.means), _p(self.scales), _p(self.quats_raw),
                                       _p(self.colors_raw), _p(self.opac_raw),
                                       *[_p(g) for g in self.grads], _p(self.m1), _p(self.m2),
                                       _p(self.step_dev), self.status, self.lr, self.b1, self.b2,
                                       self.eps, s))

    def _capture(self):
        side = torch.cuda.Stream(self.device)
        side.wait_stream(torch.cuda.current_stream(self.device))
        torch.cuda.synchronize(self.device)
        g = torch.cuda.CUDAGraph()
        with torch.cuda.graph(g, stream=side):
            self._launch()
        self.graph = g

    def _checked(self):
        """True if the last render fitted the capacity; otherwise re-plan."""
        try:
            self.plan.check()
            return True
        except _lib.JaxsplatError as e:
            if e.code != _lib.ERR_CAPACITY:
                raise
            need = self.plan.last_intersections
            self.replans += 1
            self._plan(max(int(need * self.headroom), need + 1))
            return False

    # -- public ---------------------------------------------------------------
    def step(self, k=1):
        for _ in range(int(k)):
            first = self.calls == 0
            self.calls += 1
            if first:
                # size the capacity eagerly: a skipped (overflowing) step is redone
                self._launch()
                while not self._checked():
                    self._launch()
                continue
            if self.use_graph:
                if self.graph is None:
                    self._capture()
                self.graph.replay()
            else:
                self._launch()
            if self.check_every > 0 and self.calls % self.check_every == 0:
                self._checked()

    @property
    def steps_applied(self):
        return int(self.step_dev.item())

    def loss(self):
        """mean((img - target)^2) of the last step's render (host sync)."""
        return float(self.loss_sum.item()) / self.img.numel()

    def render(self):
        """Image of the current parameters, (H, W, 3) (plain render path)."""
        from . import render
        vm, f, c = camera(self.img_shape)
        with torch.no_grad():
            q = self.quats_raw / self.quats_raw.norm(dim=1, keepdim=True)
            return render(self.means, self.scales, q, torch.sigmoid(self.colors_raw),
                          torch.sigmoid(self.opac_raw), viewmat=self.viewmats[0],
                          background=self.backgrounds[0], img_shape=self.img_shape, f=f, c=c,
                          glob_scale=1.0, clip_thresh=0.01, block_size=self.block_width)
