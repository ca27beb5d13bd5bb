"""Seeded synthetic scenes for the benchmark workloads and the parity tests.

Every scene is plain NumPy float32 in the reference API layout
(``jaxsplat/__init__.py:27-40``): means3d (N,3), scales (N,3) (linear, the
exp of the drawn log-scales), quats (N,4) w-first and normalised, colors
(N,3), opacities (N,1), a row-major (4,4) viewmat, background (3,).

Draw order inside one scene (one ``numpy.random.default_rng(seed)``):
means, log-scales, quats, colors, opacity logits.  The L2 target image is
drawn from its own ``default_rng(target_seed)`` as U(0,1) of shape (H,W,3).

Workloads (N Gaussians, image H x W, focal f, principal point c):

* ``small``   N=10,000, 256x256, f=256, c=128; x,y~U(-2,2), z~U(2,6);
              log-scale~N(ln 0.02, 0.3); opacity=sigmoid(N(0,1)).
* ``hd1m``    N=1,000,000, 1080x1920, f=1500, c=(960,540);
              x~U(-4,4), y~U(-2.5,2.5), z~U(2,10); otherwise as ``small``.
* ``dense3m`` N=3,000,000 as ``hd1m`` but log-scale~N(ln 0.05, 0.3) and
              opacity=sigmoid(N(4, 0.5)) (deep per-tile lists).
* ``orbit64`` N=500,000, means~U(-2,2)^3, 64 look-at cameras on a radius-8
              orbit, 800x800, f=800.
* ``multi32`` 32 scenes, N log-uniform in [1e4, 2e6], one camera each.
"""

from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np

IDENTITY = np.eye(4, dtype=np.float32)


@dataclass
class Camera:
    viewmat: np.ndarray
    img_shape: tuple
    f: tuple
    c: tuple


@dataclass
class Scene:
    name: str
    means3d: np.ndarray
    scales: np.ndarray
    quats: np.ndarray
    colors: np.ndarray
    opacities: np.ndarray
    cameras: list
    background: np.ndarray = field(default_factory=lambda: np.zeros(3, np.float32))
    glob_scale: float = 1.0
    clip_thresh: float = 0.01
    block_width: int = 16
    seed: int = 0

    @property
    def n(self):
        return self.means3d.shape[0]

    @property
    def camera(self):
        return self.cameras[0]

    def params(self):
        return (self.means3d, self.scales, self.quats, self.colors, self.opacities)


def _sigmoid(x):
    return (1.0 / (1.0 + np.exp(-x))).astype(np.float32)


def gaussians(rng, n, lo, hi, log_scale_mean, log_scale_std=0.3,
              opacity_logit_mean=0.0, opacity_logit_std=1.0):
    means = rng.uniform(lo, hi, size=(n, 3)).astype(np.float32)
    scales = np.exp(rng.normal(log_scale_mean, log_scale_std, size=(n, 3))).astype(np.float32)
    q = rng.normal(size=(n, 4))
    q = (q / np.linalg.norm(q, axis=1, keepdims=True)).astype(np.float32)
    colors = rng.uniform(0.0, 1.0, size=(n, 3)).astype(np.float32)
    opac = _sigmoid(rng.normal(opacity_logit_mean, opacity_logit_std, size=(n, 1)))
    return means, scales, q, colors, opac


def look_at(eye, target=(0.0, 0.0, 0.0), up=(0.0, 1.0, 0.0)):
    """World-to-camera viewmat [R | -R C] with the camera looking down +z."""
    eye = np.asarray(eye, np.float64)
    fwd = np.asarray(target, np.float64) - eye
    fwd /= np.linalg.norm(fwd)
    right = np.cross(np.asarray(up, np.float64), fwd)
    if np.linalg.norm(right) < 1e-9:
        right = np.cross(np.array([1.0, 0.0, 0.0]), fwd)
    right /= np.linalg.norm(right)
    down = np.cross(fwd, right)
    R = np.stack([right, down, fwd], axis=0)
    vm = np.eye(4)
    vm[:3, :3] = R
    vm[:3, 3] = -R @ eye
    return vm.astype(np.float32)


def make(name, seed=None, n=None):
    """Build one of the named workloads (optionally with a smaller ``n``)."""
    if name == "small":
        seed = 1000 if seed is None else seed
        rng = np.random.default_rng(seed)
        n = 10_000 if n is None else n
        g = gaussians(rng, n, (-2, -2, 2), (2, 2, 6), np.log(0.02))
        cam = Camera(IDENTITY.copy(), (256, 256), (256.0, 256.0), (128.0, 128.0))
    elif name in ("hd1m", "dense3m"):
        dense = name == "dense3m"
        seed = (1002 if dense else 1001) if seed is None else seed
        rng = np.random.default_rng(seed)
        n = (3_000_000 if dense else 1_000_000) if n is None else n
        g = gaussians(rng, n, (-4, -2.5, 2), (4, 2.5, 10),
                      np.log(0.05 if dense else 0.02),
                      opacity_logit_mean=4.0 if dense else 0.0,
                      opacity_logit_std=0.5 if dense else 1.0)
        cam = Camera(IDENTITY.copy(), (1080, 1920), (1500.0, 1500.0), (960.0, 540.0))
    elif name == "orbit64":
        seed = 1003 if seed is None else seed
        rng = np.random.default_rng(seed)
        n = 500_000 if n is None else n
        g = gaussians(rng, n, (-2, -2, -2), (2, 2, 2), np.log(0.02))
        az = rng.uniform(0, 2 * np.pi, 64)
        el = np.deg2rad(rng.uniform(-30, 30, 64))
        cams = []
        for a, e in zip(az, el):
            eye = 8.0 * np.array([np.cos(e) * np.sin(a), np.sin(e), np.cos(e) * np.cos(a)])
            cams.append(Camera(look_at(eye), (800, 800), (800.0, 800.0), (400.0, 400.0)))
        return Scene(name, *g, cameras=cams, seed=seed)
    else:
        raise ValueError(f"unknown workload {name!r}")
    return Scene(name, *g, cameras=[cam], seed=seed)


def make_multi(seed=1004, count=32):
    """``multi32``: ``count`` scenes with log-uniform N and mixed image sizes."""
    rng = np.random.default_rng(seed)
    sizes = [(512, 512), (768, 1024), (1080, 1920)]
    out = []
    for s in range(count):
        n = int(np.exp(rng.uniform(np.log(1e4), np.log(2e6))))
        H, W = sizes[int(rng.integers(0, 3))]
        g = gaussians(rng, n, (-4, -2.5, 2), (4, 2.5, 10), np.log(0.02))
        cam = Camera(IDENTITY.copy(), (H, W), (float(W), float(W)), (W / 2.0, H / 2.0))
        out.append(Scene(f"multi32[{s}]", *g, cameras=[cam], seed=seed))
    return out


def target_image(img_shape, seed):
    rng = np.random.default_rng(seed)
    return rng.uniform(0.0, 1.0, size=(*img_shape, 3)).astype(np.float32)
