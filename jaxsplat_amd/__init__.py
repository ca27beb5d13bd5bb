"""jaxsplat_amd: MI355X-native differentiable 3D Gaussian splatting.

Drop-in for the reference API of yklcs/jaxsplat (``jaxsplat/__init__.py:6-73``):
``render``, ``project`` and ``rasterize`` with the same keyword-only arguments,
outputs and cotangent routing, on torch GPU tensors, backed by hand-written
HIP kernels for gfx950 reached through the C ABI of ``libjaxsplat_amd.so``.

``adjoint="reference"`` (default) follows the reference backward's formulas,
including its fwd/bwd alpha-clamp mismatch and its unclamped EWA Jacobian;
``adjoint="consistent"`` gives the exact adjoint of the forward.  Either
matches its float64 restatement up to float32 rounding, which includes the
borderline discrete decisions: the forward updates T as fma(-T, alpha, T)
and accepts a pair by comparing sigma * log2(e) with log2(255 o) on a conic
pre-scaled to base 2, so a pair with alpha within rounding of 1/255, or a T
within rounding of 1e-4, may be decided the other way than the reference's
o * exp(-sigma) >= 1/255 and T (1 - alpha) <= 1e-4 would decide it.  The
parity tests exclude (and count) the pixels where such a decision is
ambiguous in float32.
"""

from __future__ import annotations

from . import _ops
from ._ops import IsectCapacity, ShapeDtypeError, bin_and_sort, project, rasterize
from .batch import BatchPlan, BatchRenderer, ViewSpec, render_batch

__all__ = ["render", "project", "rasterize", "bin_and_sort", "ShapeDtypeError", "IsectCapacity",
           "BatchPlan", "BatchRenderer", "ViewSpec", "render_batch"]


def render(means3d, scales, quats, colors, opacities, *, viewmat, background, img_shape, f, c,
           glob_scale, clip_thresh, block_size=16, adjoint="reference", antialiased=False,
           isect_capacity=None):
    """Render 3D Gaussians to an (H, W, 3) image, differentiably.

    Same contract as the reference ``render`` (``jaxsplat/__init__.py:6-70``):
    means3d (N,3), scales (N,3), quats (N,4) (w first, normalised), colors
    (N,3), opacities (N,1); viewmat (4,4) row-major world-to-camera;
    background (3,); img_shape (H, W); f (fx, fy); c (cx, cy); glob_scale;
    clip_thresh; block_size in [1, 16] (tile side).  Differentiable in means3d,
    scales, quats, colors and opacities.  colors (N, C) with background (C,)
    renders C channels (features, depth, ...) with the same blend.

    ``antialiased=True`` (not in the reference API) blends with
    ``opacities * compensation``, the compensation that the reference's
    projection computes and drops (``forward.cu:537-543``); gradients flow
    through both factors.

    ``isect_capacity`` (an ``IsectCapacity``, not in the reference API): no
    host synchronisation -- the binning uses buffers of that fixed capacity
    and counts on the device, so a whole fwd+bwd step is kernel launches and
    may be captured in a HIP graph; ``isect_capacity.check()`` reports an
    overflow afterwards.
    """
    fusable = (isinstance(colors, _ops.torch.Tensor) and colors.dim() == 2 and
               1 <= colors.shape[1] <= 3)
    if _ops.TIGHT_TILES and _ops.FUSED_RENDER and not antialiased and fusable:
        # one autograd node (up to 3 channels): projection with the
        # opacity-limited rectangles and splat records, binning,
        # rasterization; its backward sums the rasterize records and runs the
        # projection adjoint in one per-Gaussian pass
        return _ops.render_fused(means3d, scales, quats, colors, opacities, viewmat=viewmat,
                                 background=background, img_shape=img_shape, f=f, c=c,
                                 glob_scale=glob_scale, clip_thresh=clip_thresh,
                                 block_width=block_size, adjoint=adjoint,
                                 isect_capacity=isect_capacity)
    if _ops.TIGHT_TILES:
        # one projection pass also yields the opacity-limited tile rectangles
        # (of the raw opacities: compensation <= 1 only shrinks the region)
        xys, depths, radii, conics, cum, comp, rects = _ops.project_for_render(
            means3d, scales, quats, opacities, viewmat, img_shape=img_shape, f=f, c=c,
            glob_scale=glob_scale, clip_thresh=clip_thresh, block_width=block_size,
            adjoint=adjoint)
        if antialiased:
            opacities = opacities * comp
        return rasterize(colors, opacities, background, xys, depths, radii, conics, cum,
                         img_shape=img_shape, block_width=block_size, adjoint=adjoint,
                         tile_rects=rects, isect_capacity=isect_capacity)
    out = project(means3d, scales, quats, viewmat, img_shape=img_shape, f=f, c=c,
                  glob_scale=glob_scale, clip_thresh=clip_thresh, block_width=block_size,
                  adjoint=adjoint, return_compensation=antialiased)
    xys, depths, radii, conics, _num_tiles_hit, cum_tiles_hit = out[:6]
    if antialiased:
        opacities = opacities * out[6]
    return rasterize(colors, opacities, background, xys, depths, radii, conics, cum_tiles_hit,
                     img_shape=img_shape, block_width=block_size, adjoint=adjoint,
                     isect_capacity=isect_capacity)
