"""Differentiable ``project`` / ``rasterize`` on torch device memory.

Mirrors the reference's ``jax.custom_vjp`` pair for each op
(``jaxsplat/_project/__init__.py:56-208`` and
``jaxsplat/_rasterize/__init__.py:40-190``) with ``torch.autograd.Function``:
the forward calls the HIP kernels through the C ABI, saves residuals, and the
backward calls the hand-written adjoint kernels.  Cotangent routing follows the
reference: ``project`` is differentiable in means3d, scales, quats (``None``
for viewmat, ``_project/__init__.py:205``); ``rasterize`` in colors,
opacities, xys, conics (``None`` for background, depths, radii,
cum_tiles_hit, ``_rasterize/__init__.py:187``).  The dropped outputs
(compensation, img_alpha) carry zero cotangents, as in the reference.

Difference in residuals (same results): the sorted Gaussian ids and tile bins
of the forward are kept for the backward instead of sorting twice
(``lib/ops.cu:188-198``), and cov3d is recomputed in the backward instead of
being stored.

The rasterize backward of up to three channels sums each Gaussian's
per-tile partial gradients in a fixed order (``jaxsplat_rasterize_backward_
staged``) instead of with float atomics, so it is bitwise reproducible.
``STAGED_BACKWARD = False`` (or ``JAXSPLAT_AMD_STAGED_BACKWARD=0`` in the
environment) selects the atomic backward, which differs only in rounding.

``rasterize`` bins each Gaussian into its opacity-limited tile rectangle
(``jaxsplat_tile_rects``): the reference's rectangle without the tiles where
every pixel fails the per-pixel test o exp(-sigma) >= 1/255.  Those entries
never blend, so the image and gradients are the reference's, from shorter tile
lists; only the internal sorted indices (the final_idx residual) differ.
``TIGHT_TILES = False`` (or ``JAXSPLAT_AMD_TIGHT_TILES=0``) bins into the
reference's rectangles given by ``cum_tiles_hit``.
"""

from __future__ import annotations

import ctypes
import os

import torch

from . import _lib

STAGED_BACKWARD = os.environ.get("JAXSPLAT_AMD_STAGED_BACKWARD", "1") != "0"
TIGHT_TILES = os.environ.get("JAXSPLAT_AMD_TIGHT_TILES", "1") != "0"
FUSED_RENDER = os.environ.get("JAXSPLAT_AMD_FUSED_RENDER", "1") != "0"
# the staged backward's record slots, and the capacity of one binning, are 31-bit
STAGED_MAX_INTERSECTIONS = (1 << 30) - 1


class ShapeDtypeError(AssertionError, TypeError):
    """Raised where the reference's abstract evaluation asserts (``_types.py:30-33``)."""


def _expect(name, t, shape, dtype):
    if not isinstance(t, torch.Tensor):
        raise ShapeDtypeError(f"{name}: expected a torch.Tensor, got {type(t).__name__}")
    if tuple(t.shape) != tuple(shape) or t.dtype != dtype:
        raise ShapeDtypeError(
            f"{name}: expected shape {tuple(shape)} {dtype}, got {tuple(t.shape)} {t.dtype}")
    if t.device.type != "cuda":
        raise ShapeDtypeError(f"{name}: expected a GPU tensor, got device {t.device}")


def _p(t):
    return ctypes.c_void_p(t.data_ptr()) if t is not None else None


def _stream(device):
    return ctypes.c_void_p(torch.cuda.current_stream(device).cuda_stream)


def _adjoint(mode):
    try:
        return _lib.ADJOINT[mode]
    except KeyError:
        raise ValueError(f"adjoint must be one of {sorted(_lib.ADJOINT)}, got {mode!r}") from None


def _project_forward(d, means3d, scales, quats, viewmat, opacities=None):
    """Projection through the C ABI.  With opacities (the fused render) the
    tile counts are those of the opacity-limited rectangles, also returned.
    Returns (xys, depths, radii, conics, num_tiles_hit, cum_tiles_hit,
    compensation, rects)."""
    N = means3d.shape[0]
    dev = means3d.device
    L = _lib.lib()
    xys = torch.empty((N, 2), dtype=torch.float32, device=dev)
    depths = torch.empty((N, 1), dtype=torch.float32, device=dev)
    radii = torch.empty((N, 1), dtype=torch.int32, device=dev)
    conics = torch.empty((N, 3), dtype=torch.float32, device=dev)
    nth = torch.empty((N, 1), dtype=torch.int32, device=dev)
    cum = torch.empty((N, 1), dtype=torch.int32, device=dev)
    comp = torch.empty((N, 1), dtype=torch.float32, device=dev)
    rects = torch.empty((N if opacities is not None else 0, 2), dtype=torch.int32, device=dev)
    ws_bytes = L.jaxsplat_project_workspace_size(ctypes.byref(d))
    ws = torch.empty(max(ws_bytes, 1), dtype=torch.uint8, device=dev)
    with torch.cuda.device(dev):
        if opacities is not None:
            _lib.check(L.jaxsplat_project_forward_rects(
                ctypes.byref(d), _p(means3d), _p(scales), _p(quats), _p(viewmat),
                _p(opacities), None, _p(xys), _p(depths), _p(radii), _p(conics), _p(comp),
                _p(nth), _p(cum), _p(rects), _p(ws), ws_bytes, _stream(dev)))
        else:
            _lib.check(L.jaxsplat_project_forward(
                ctypes.byref(d), _p(means3d), _p(scales), _p(quats), _p(viewmat), None,
                _p(xys), _p(depths), _p(radii), _p(conics), _p(comp), _p(nth), _p(cum),
                _p(ws), ws_bytes, _stream(dev)))
    return xys, depths, radii, conics, nth, cum, comp, rects


def _project_backward(d, adjoint, means3d, scales, quats, viewmat, radii, conics, comp, v_xys,
                      v_depths, v_conics, v_comp):
    """Projection adjoint through the C ABI; None cotangents are zero.
    Returns (v_mean3d, v_scale, v_quat)."""
    N = means3d.shape[0]
    dev = means3d.device
    v_xys = torch.zeros((N, 2), dtype=torch.float32, device=dev) if v_xys is None \
        else v_xys.contiguous()
    v_conics = torch.zeros((N, 3), dtype=torch.float32, device=dev) if v_conics is None \
        else v_conics.contiguous()
    v_depths = v_depths.contiguous() if v_depths is not None else None
    v_comp = v_comp.contiguous() if v_comp is not None else None
    v_mean3d = torch.empty_like(means3d)
    v_scale = torch.empty_like(scales)
    v_quat = torch.empty_like(quats)
    L = _lib.lib()
    with torch.cuda.device(dev):
        _lib.check(L.jaxsplat_project_backward(
            ctypes.byref(d), _p(means3d), _p(scales), _p(quats), _p(viewmat),
            _p(radii), _p(conics), _p(comp), _p(v_xys), _p(v_depths), _p(v_conics),
            _p(v_comp), _p(v_mean3d), _p(v_scale), _p(v_quat), None, None, adjoint,
            _stream(dev)))
    return v_mean3d, v_scale, v_quat


class _Project(torch.autograd.Function):
    @staticmethod
    def forward(ctx, cfg, means3d, scales, quats, viewmat, opacities=None):
        # with opacities (the fused render): the counts are those of the
        # opacity-limited tile rectangles, returned as an eighth output
        img_shape, f, c, glob_scale, clip_thresh, block_width, adjoint = cfg
        N = means3d.shape[0]
        d = _lib.make_descriptor(N, img_shape, f, c, glob_scale, clip_thresh, block_width)
        xys, depths, radii, conics, nth, cum, comp, rects = _project_forward(
            d, means3d, scales, quats, viewmat, opacities)
        ctx.d = d
        ctx.adjoint = adjoint
        ctx.save_for_backward(means3d, scales, quats, viewmat, radii, conics, comp)
        ctx.mark_non_differentiable(radii, nth, cum, rects)
        ctx.set_materialize_grads(False)
        return xys, depths, radii, conics, nth, cum, comp, rects

    @staticmethod
    def backward(ctx, v_xys, v_depths, _v_radii, v_conics, _v_nth, _v_cum, v_comp, _v_rects):
        means3d, scales, quats, viewmat, radii, conics, comp = ctx.saved_tensors
        v_mean3d, v_scale, v_quat = _project_backward(
            ctx.d, ctx.adjoint, means3d, scales, quats, viewmat, radii, conics, comp, v_xys,
            v_depths, v_conics, v_comp)
        return None, v_mean3d, v_scale, v_quat, None, None


def project(means3d, scales, quats, viewmat, *, img_shape, f, c, glob_scale, clip_thresh,
            block_width, adjoint="reference", return_compensation=False):
    """Project 3D Gaussians to the image plane (reference ``_project/__init__.py:56-87``).

    Returns ``(xys (N,2), depths (N,1), radii (N,1) int32, conics (N,3),
    num_tiles_hit (N,1) int32, cum_tiles_hit (N,1) int32)``.  The reference
    types the two counts uint32; torch's uint32 lacks arithmetic, so they are
    int32 here (the values are identical, all < 2^31).

    ``return_compensation=True`` appends the differentiable compensation
    (N,1) = sqrt(det(cov2d) / det(cov2d + 0.3 I)) that the reference computes
    and drops (``forward.cu:537-543``, ``_project/__init__.py:78-87``); its
    cotangent reaches the projection backward (``helpers.h:73-87``).
    """
    adj = _adjoint(adjoint)
    N = means3d.shape[0] if isinstance(means3d, torch.Tensor) else -1
    _expect("means3d", means3d, (N, 3), torch.float32)
    _expect("scales", scales, (N, 3), torch.float32)
    _expect("quats", quats, (N, 4), torch.float32)
    _expect("viewmat", viewmat, (4, 4), torch.float32)
    cfg = (tuple(img_shape), tuple(f), tuple(c), float(glob_scale), float(clip_thresh),
           int(block_width), adj)
    out = _Project.apply(cfg, means3d.contiguous(), scales.contiguous(), quats.contiguous(),
                         viewmat.contiguous(), None)
    return out[:7] if return_compensation else out[:6]


def project_for_render(means3d, scales, quats, opacities, viewmat, *, img_shape, f, c,
                       glob_scale, clip_thresh, block_width, adjoint="reference"):
    """The projection of the fused render: ``project`` whose tile counts are
    those of the opacity-limited rectangles of ``opacities`` (no gradient
    flows to them through this op).  Returns ``(xys, depths, radii, conics,
    cum_tiles_hit, compensation, tile_rects)`` for ``rasterize(...,
    tile_rects=...)``."""
    adj = _adjoint(adjoint)
    N = means3d.shape[0] if isinstance(means3d, torch.Tensor) else -1
    _expect("means3d", means3d, (N, 3), torch.float32)
    _expect("scales", scales, (N, 3), torch.float32)
    _expect("quats", quats, (N, 4), torch.float32)
    _expect("opacities", opacities, (N, 1), torch.float32)
    _expect("viewmat", viewmat, (4, 4), torch.float32)
    cfg = (tuple(img_shape), tuple(f), tuple(c), float(glob_scale), float(clip_thresh),
           int(block_width), adj)
    xys, depths, radii, conics, _nth, cum, comp, rects = _Project.apply(
        cfg, means3d.contiguous(), scales.contiguous(), quats.contiguous(), viewmat.contiguous(),
        opacities.detach().contiguous())
    return xys, depths, radii, conics, cum, comp, rects


def bin_and_sort(xys, depths, radii, cum_tiles_hit, *, img_shape, block_width,
                 return_tiles=False):
    """Run the binning stage alone (tests, tools).  Returns
    ``(gaussian_ids_sorted, tile_bins (T,2)[, tile_ids_sorted], num_intersects)``."""
    N = xys.shape[0]
    dev = xys.device
    d = _lib.make_descriptor(N, img_shape, (1.0, 1.0), (0.0, 0.0), 1.0, 0.0, block_width)
    isect = int(cum_tiles_hit.reshape(-1)[-1].item()) if N > 0 else 0
    return _bin(d, xys, depths, radii, isect, dev, return_tiles)


def _bin(d, xys, depths, radii, isect, dev, return_tiles=False):
    L = _lib.lib()
    ntiles = d.grid_dim_2d.x * d.grid_dim_2d.y
    gids = torch.empty(max(isect, 1), dtype=torch.int32, device=dev)
    tiles = torch.empty(max(isect, 1), dtype=torch.int32, device=dev) if return_tiles else None
    bins = torch.empty((ntiles, 2), dtype=torch.int32, device=dev)
    ws_bytes = L.jaxsplat_bin_workspace_size(ctypes.byref(d), isect)
    ws = torch.empty(max(ws_bytes, 1), dtype=torch.uint8, device=dev)
    with torch.cuda.device(dev):
        _lib.check(L.jaxsplat_bin_and_sort(
            ctypes.byref(d), _p(xys), _p(depths), _p(radii), isect, _p(gids), _p(tiles),
            _p(bins), _p(ws), ws_bytes, _stream(dev)))
    out = (gids[:isect], bins)
    if return_tiles:
        out = out + (tiles[:isect],)
    return out + (isect,)


class IsectCapacity:
    """A fixed intersection capacity for ``render`` / ``rasterize`` calls
    that must not synchronise with the host (not in the reference API).

    By default the binning reads the intersection count back to the host to
    size its buffers, as the reference does (``lib/ops.cu:425-431``); that
    host read keeps a fwd+bwd step out of a HIP graph.  With
    ``isect_capacity=cap`` the buffers are sized for ``cap.capacity``
    intersections instead and the kernels count on the device, so a step is
    kernel launches only.  Each binning records on the device whether it
    overflowed the capacity and its exact total; ``check()`` reads them (a
    host synchronisation) and raises ``JaxsplatError`` (``ERR_CAPACITY``)
    after an overflow -- whose images and gradients miss the intersections
    beyond the capacity -- having grown ``capacity`` to ``headroom`` times the
    largest total seen (re-run, and re-capture a graph, after that).

    Results equal the default path bit for bit, except that a render with no
    intersections at all keeps final_Ts = 1 internally (the reference writes 0
    there); its image (the background) and gradients (zero) are the same.
    """

    def __init__(self, capacity, device=None, headroom=1.25):
        capacity = int(capacity)
        if not 1 <= capacity <= STAGED_MAX_INTERSECTIONS:
            raise ValueError(f"capacity must be in [1, 2^30 - 1], got {capacity}")
        self.capacity = capacity
        self.headroom = float(headroom)
        self.device = torch.device(device if device is not None else "cuda")
        self._err = torch.zeros(1, dtype=torch.int32, device=self.device)
        self._total = torch.zeros(1, dtype=torch.int64, device=self.device)
        self.last_total = 0

    def _record(self, ws):
        # emission workspace header: error word at byte 0, exact total (uint64) at byte 8
        torch.bitwise_or(self._err, ws[0:4].view(torch.int32), out=self._err)
        torch.maximum(self._total, ws[8:16].view(torch.int64), out=self._total)

    def check(self):
        """Largest intersection total since the last check (host sync);
        raises ``JaxsplatError`` after an overflow or a binning error."""
        err, total = int(self._err.item()), int(self._total.item())
        self._err.zero_()
        self._total.zero_()
        self.last_total = total
        if err & 8:
            old = self.capacity
            self.capacity = min(max(int(total * self.headroom), total + 1),
                                STAGED_MAX_INTERSECTIONS)
            raise _lib.JaxsplatError(
                _lib.ERR_CAPACITY, f"{total} intersections exceeded the capacity {old}; "
                f"capacity grown to {self.capacity}")
        if err:
            raise _lib.JaxsplatError(_lib.ERR_CAPACITY, f"binning error (code {err})")
        return total


def _emit(d, dws, isect, dev, capacity):
    """Emission, tile sort and bin edges over a depth-stage workspace, for
    isect intersections (the exact count, or capacity.capacity).  Returns
    (gaussian_ids_sorted, tile_bins, isect)."""
    L = _lib.lib()
    ntiles = d.grid_dim_2d.x * d.grid_dim_2d.y
    gids = torch.empty(max(isect, 1), dtype=torch.int32, device=dev)
    bins = torch.empty((ntiles, 2), dtype=torch.int32, device=dev)
    ws_bytes = L.jaxsplat_bin_emit_workspace_size(ctypes.byref(d), isect)
    ws = torch.empty(max(ws_bytes, 16), dtype=torch.uint8, device=dev)
    _lib.check(L.jaxsplat_bin_emit_sort(ctypes.byref(d), isect, _p(dws), _p(gids), None,
                                        _p(bins), _p(ws), ws_bytes, _stream(dev)))
    if capacity is not None:
        capacity._record(ws)
    return gids[:isect], bins, isect


def _read_count(cum, dev):
    """Queue a copy of the last inclusive count to pinned memory; returns
    (host tensor, event) -- the caller waits on the event after queueing the
    depth stage, so that the read overlaps it."""
    count = torch.empty(1, dtype=torch.int32, pin_memory=True)
    count.copy_(cum.reshape(-1)[-1:], non_blocking=True)
    ready = torch.cuda.Event()
    ready.record()
    return count, ready


def _bin_overlapped(d, xys, depths, radii, cum_tiles_hit, dev, capacity=None):
    """Binning into the reference's tile rectangles.  By default the host read
    of num_intersects = cum_tiles_hit[N-1] (which sizes the outputs, as in the
    reference) overlaps the depth stage: the count is copied to pinned memory
    and the depth sort queued behind it before the host waits.  With a
    capacity (IsectCapacity) nothing is read back.  Returns
    (gaussian_ids_sorted, tile_bins, isect)."""
    N = xys.shape[0]
    if N == 0:
        return _bin(d, xys, depths, radii, 0, dev)
    L = _lib.lib()
    with torch.cuda.device(dev):
        if capacity is None:
            count, ready = _read_count(cum_tiles_hit, dev)
        dws_bytes = L.jaxsplat_bin_depth_workspace_size(ctypes.byref(d))
        dws = torch.empty(max(dws_bytes, 1), dtype=torch.uint8, device=dev)
        _lib.check(L.jaxsplat_bin_depth_sort(ctypes.byref(d), _p(xys), _p(depths), _p(radii),
                                             _p(dws), dws_bytes, _stream(dev)))
        if capacity is None:
            ready.synchronize()
            isect = int(count[0])
        else:
            isect = capacity.capacity
        return _emit(d, dws, isect, dev, capacity)


def _bin_tight(d, xys, depths, radii, conics, opacities, dev, capacity=None):
    """Binning into the opacity-limited tile rectangles: the rectangles and
    their inclusive counts (jaxsplat_tile_rects), then _bin_rects.  Returns
    (gaussian_ids_sorted, tile_bins, isect, rects, cum)."""
    N = xys.shape[0]
    L = _lib.lib()
    rects = torch.empty((max(N, 1), 2), dtype=torch.int32, device=dev)
    cum = torch.empty((max(N, 1), 1), dtype=torch.int32, device=dev)
    if N == 0:
        gids, bins, isect = _bin(d, xys, depths, radii, 0, dev)
        return gids, bins, isect, rects, cum
    with torch.cuda.device(dev):
        rws_bytes = L.jaxsplat_tile_rects_workspace_size(ctypes.byref(d))
        rws = torch.empty(max(rws_bytes, 1), dtype=torch.uint8, device=dev)
        _lib.check(L.jaxsplat_tile_rects(ctypes.byref(d), _p(xys), _p(radii), _p(conics),
                                         _p(opacities), _p(rects), None, _p(cum), _p(rws),
                                         rws_bytes, _stream(dev)))
    gids, bins, isect = _bin_rects(d, depths, rects, cum, dev, capacity)
    return gids, bins, isect, rects, cum


def _bin_rects(d, depths, rects, cum, dev, capacity=None):
    """Binning into given tile rectangles with their inclusive counts cum
    (jaxsplat_bin_rects_sort: emission straight from the rectangles and
    counts, no depth stage), after a host read of the total, or (with a
    capacity) no read.  Returns (gaussian_ids_sorted, tile_bins, isect)."""
    L = _lib.lib()
    ntiles = d.grid_dim_2d.x * d.grid_dim_2d.y
    with torch.cuda.device(dev):
        if capacity is None:
            count, ready = _read_count(cum, dev)
            ready.synchronize()
            isect = int(count[0])
        else:
            isect = capacity.capacity
        gids = torch.empty(max(isect, 1), dtype=torch.int32, device=dev)
        bins = torch.empty((ntiles, 2), dtype=torch.int32, device=dev)
        ws_bytes = L.jaxsplat_bin_emit_workspace_size(ctypes.byref(d), isect)
        ws = torch.empty(max(ws_bytes, 16), dtype=torch.uint8, device=dev)
        _lib.check(L.jaxsplat_bin_rects_sort(ctypes.byref(d), _p(depths), _p(rects), _p(cum),
                                             isect, _p(gids), None, _p(bins), _p(ws), ws_bytes,
                                             _stream(dev)))
    if capacity is not None:
        capacity._record(ws)
    return gids[:isect], bins, isect


def staged_backward_applies(nch, isect):
    """Whether the rasterize backward takes the staged entry point: up to
    three channels and fewer than 2^30 intersections (its record slots are
    31-bit); otherwise the atomic N-channel backward, which has neither limit
    and differs only in the order of the float additions."""
    return STAGED_BACKWARD and 1 <= nch <= 3 and 0 <= isect <= STAGED_MAX_INTERSECTIONS


def _rasterize_forward(d, gids, bins, isect, xys, conics, colors, opacities, background,
                       img_shape):
    """Rasterize forward of binned lists.  Returns (img (H,W,C), final_Ts
    (H,W,1), final_idx (H,W,1))."""
    N, nch = colors.shape
    dev = colors.device
    H, W = img_shape
    final_Ts = torch.empty((H, W, 1), dtype=torch.float32, device=dev)
    final_idx = torch.empty((H, W, 1), dtype=torch.int32, device=dev)
    img = torch.empty((H, W, nch), dtype=torch.float32, device=dev)
    L = _lib.lib()
    with torch.cuda.device(dev):
        _lib.check(L.jaxsplat_rasterize_forward_nd(
            ctypes.byref(d), _p(gids), _p(bins), isect, _p(xys), _p(conics), _p(colors),
            _p(opacities), _p(background), nch, _p(final_Ts), _p(final_idx), _p(img),
            _stream(dev)))
    return img, final_Ts, final_idx


def _staged_workspace(d, isect, nch, dev):
    """Workspace of the staged rasterize backward (records and bitmap), or
    None where that backward does not apply or its 52 bytes per intersection
    do not fit (the atomic backward is then used).  Returns (ws, bytes)."""
    if not staged_backward_applies(nch, isect):
        return None, 0
    L = _lib.lib()
    wsb = L.jaxsplat_rasterize_backward_staged_workspace_size(ctypes.byref(d), isect)
    try:
        return torch.empty(max(wsb, 1), dtype=torch.uint8, device=dev), wsb
    except torch.OutOfMemoryError:
        return None, 0


def _rasterize_backward(d, isect, adjoint, colors, opacities, background, xys, conics, gids,
                        bins, final_Ts, final_idx, radii, cum_tiles_hit, rects, v_img):
    """Rasterize adjoint: the staged backward (per-intersection records summed
    per Gaussian: no float atomics, bitwise reproducible) when it applies,
    else the atomic N-channel backward.  rects: the opacity-limited tile
    rectangles the binning used, or None for the reference's.  Returns
    (v_colors, v_opacity, v_xy, v_conic)."""
    dev = colors.device
    v_img = v_img.contiguous()
    v_colors = torch.empty_like(colors)
    v_opacity = torch.empty_like(opacities)
    v_xy = torch.empty_like(xys)
    v_conic = torch.empty_like(conics)
    L = _lib.lib()
    nch = colors.shape[1]
    ws, wsb = _staged_workspace(d, isect, nch, dev)
    with torch.cuda.device(dev):
        if ws is not None and rects is not None:
            _lib.check(L.jaxsplat_rasterize_backward_staged_rects(
                ctypes.byref(d), _p(gids), _p(bins), isect, _p(xys), _p(conics), _p(colors),
                _p(opacities), _p(background), nch, _p(radii), _p(rects), _p(cum_tiles_hit),
                _p(final_Ts), _p(final_idx), _p(v_img), None, _p(v_colors), _p(v_opacity),
                _p(v_xy), None, _p(v_conic), adjoint, _p(ws), wsb, _stream(dev)))
        elif ws is not None:
            _lib.check(L.jaxsplat_rasterize_backward_staged(
                ctypes.byref(d), _p(gids), _p(bins), isect, _p(xys), _p(conics), _p(colors),
                _p(opacities), _p(background), nch, _p(radii), _p(cum_tiles_hit),
                _p(final_Ts), _p(final_idx), _p(v_img), None, _p(v_colors), _p(v_opacity),
                _p(v_xy), None, _p(v_conic), adjoint, _p(ws), wsb, _stream(dev)))
        else:
            _lib.check(L.jaxsplat_rasterize_backward_nd(
                ctypes.byref(d), _p(gids), _p(bins), isect, _p(xys), _p(conics), _p(colors),
                _p(opacities), _p(background), nch, _p(final_Ts), _p(final_idx), _p(v_img),
                None, _p(v_colors), _p(v_opacity), _p(v_xy), None, _p(v_conic), adjoint,
                _stream(dev)))
    return v_colors, v_opacity, v_xy, v_conic


class _Rasterize(torch.autograd.Function):
    @staticmethod
    def forward(ctx, cfg, colors, opacities, background, xys, depths, radii, conics,
                cum_tiles_hit, tile_rects=None, capacity=None):
        img_shape, block_width, adjoint = cfg
        N, nch = colors.shape
        dev = colors.device
        d = _lib.make_descriptor(N, img_shape, (1.0, 1.0), (0.0, 0.0), 1.0, 0.0, block_width)
        rects = None
        if tile_rects is not None:
            # rectangles and counts from project_for_render
            rects = tile_rects
            if N == 0:
                gids, bins, isect = _bin(d, xys, depths, radii, 0, dev)
            else:
                gids, bins, isect = _bin_rects(d, depths, rects, cum_tiles_hit, dev, capacity)
        elif TIGHT_TILES:
            gids, bins, isect, rects, cum_tiles_hit = _bin_tight(d, xys, depths, radii, conics,
                                                                 opacities, dev, capacity)
        else:
            gids, bins, isect = _bin_overlapped(d, xys, depths, radii, cum_tiles_hit, dev,
                                                capacity)
        img, final_Ts, final_idx = _rasterize_forward(d, gids, bins, isect, xys, conics, colors,
                                                      opacities, background, img_shape)
        ctx.d = d
        ctx.isect = isect
        ctx.adjoint = adjoint
        ctx.tight = rects is not None
        ctx.save_for_backward(colors, opacities, background, xys, conics, gids, bins, final_Ts,
                              final_idx, radii, cum_tiles_hit,
                              rects if rects is not None else radii)
        return img

    @staticmethod
    def backward(ctx, v_img):
        colors, opacities, background, xys, conics, gids, bins, final_Ts, final_idx, radii, \
            cum_tiles_hit, rects = ctx.saved_tensors
        v_colors, v_opacity, v_xy, v_conic = _rasterize_backward(
            ctx.d, ctx.isect, ctx.adjoint, colors, opacities, background, xys, conics, gids, bins,
            final_Ts, final_idx, radii, cum_tiles_hit, rects if ctx.tight else None, v_img)
        return None, v_colors, v_opacity, None, v_xy, None, None, v_conic, None, None, None


class _Render(torch.autograd.Function):
    """The fused render as one autograd node: the projection with the
    opacity-limited rectangles, which also writes one 48-byte splat record per
    Gaussian (position, opacity and its alpha limit, conic, colour,
    rectangle), the binning into those rectangles and the rasterize forward,
    staging its entries from the records; the backward is
    ``jaxsplat_render_backward`` (the staged rasterize records, staged from
    the splat records, then one per-Gaussian pass that sums them and runs the
    projection VJP in registers), or, where the staged backward does not
    apply, the rasterize adjoint followed by the projection adjoint.  Results
    are those of ``project_for_render`` + ``rasterize`` bit for bit."""

    @staticmethod
    def forward(ctx, cfg, means3d, scales, quats, colors, opacities, viewmat, background,
                capacity):
        img_shape, f, c, glob_scale, clip_thresh, block_width, adjoint = cfg
        N, nch = colors.shape
        dev = means3d.device
        H, W = img_shape
        d = _lib.make_descriptor(N, img_shape, f, c, glob_scale, clip_thresh, block_width)
        L = _lib.lib()
        # no xys or per-Gaussian counts: the splat records hold the positions
        # and cum the counts
        depths = torch.empty((N, 1), dtype=torch.float32, device=dev)
        radii = torch.empty((N, 1), dtype=torch.int32, device=dev)
        conics = torch.empty((N, 3), dtype=torch.float32, device=dev)
        cum = torch.empty((N, 1), dtype=torch.int32, device=dev)
        rects = torch.empty((N, 2), dtype=torch.int32, device=dev)
        splats = torch.empty((N, 12), dtype=torch.float32, device=dev)
        ws_bytes = L.jaxsplat_project_workspace_size(ctypes.byref(d))
        ws = torch.empty(max(ws_bytes, 1), dtype=torch.uint8, device=dev)
        with torch.cuda.device(dev):
            _lib.check(L.jaxsplat_project_forward_splats(
                ctypes.byref(d), _p(means3d), _p(scales), _p(quats), _p(viewmat),
                _p(opacities), _p(colors), nch, None, _p(depths), _p(radii), _p(conics),
                None, None, _p(cum), _p(rects), _p(splats), _p(ws), ws_bytes,
                _stream(dev)))
        if N == 0:
            gids, bins, isect = _bin(d, splats[:, :2].contiguous(), depths, radii, 0, dev)
        else:
            gids, bins, isect = _bin_rects(d, depths, rects, cum, dev, capacity)
        final_Ts = torch.empty((H, W, 1), dtype=torch.float32, device=dev)
        final_idx = torch.empty((H, W, 1), dtype=torch.int32, device=dev)
        img = torch.empty((H, W, nch), dtype=torch.float32, device=dev)
        with torch.cuda.device(dev):
            _lib.check(L.jaxsplat_rasterize_forward_splats(
                ctypes.byref(d), _p(gids), _p(bins), isect, _p(splats), _p(background), nch,
                _p(final_Ts), _p(final_idx), _p(img), _stream(dev)))
        ctx.d = d
        ctx.isect = isect
        ctx.adjoint = adjoint
        ctx.save_for_backward(means3d, scales, quats, viewmat, colors, opacities, background,
                              conics, radii, cum, rects, splats, gids, bins, final_Ts,
                              final_idx)
        return img

    @staticmethod
    def backward(ctx, v_img):
        means3d, scales, quats, viewmat, colors, opacities, background, conics, radii, \
            cum, rects, splats, gids, bins, final_Ts, final_idx = ctx.saved_tensors
        dev = means3d.device
        nch = colors.shape[1]
        ws, wsb = _staged_workspace(ctx.d, ctx.isect, nch, dev)
        if ws is None:
            xys = splats[:, :2].contiguous()
            v_colors, v_opacity, v_xy, v_conic = _rasterize_backward(
                ctx.d, ctx.isect, ctx.adjoint, colors, opacities, background, xys, conics, gids,
                bins, final_Ts, final_idx, radii, cum, rects, v_img)
            v_mean3d, v_scale, v_quat = _project_backward(
                ctx.d, ctx.adjoint, means3d, scales, quats, viewmat, radii, conics, None, v_xy,
                None, v_conic, None)
            return None, v_mean3d, v_scale, v_quat, v_colors, v_opacity, None, None, None
        v_img = v_img.contiguous()
        v_colors = torch.empty_like(colors)
        v_opacity = torch.empty_like(opacities)
        v_mean3d = torch.empty_like(means3d)
        v_scale = torch.empty_like(scales)
        v_quat = torch.empty_like(quats)
        L = _lib.lib()
        with torch.cuda.device(dev):
            _lib.check(L.jaxsplat_render_backward(
                ctypes.byref(ctx.d), _p(gids), _p(bins), ctx.isect, _p(means3d), _p(scales),
                _p(quats), _p(viewmat), _p(splats), _p(conics), _p(background), nch, _p(radii),
                _p(rects), _p(cum), _p(final_Ts), _p(final_idx), _p(v_img), _p(v_colors),
                _p(v_opacity), _p(v_mean3d), _p(v_scale), _p(v_quat), ctx.adjoint, _p(ws), wsb,
                _stream(dev)))
        return None, v_mean3d, v_scale, v_quat, v_colors, v_opacity, None, None, None


def render_fused(means3d, scales, quats, colors, opacities, *, viewmat, background, img_shape, f,
                 c, glob_scale, clip_thresh, block_width, adjoint="reference",
                 isect_capacity=None):
    """``render`` as one autograd node (``_Render``): the opacity-limited
    binning and the fused backward tail.  Same arguments and result as
    ``render`` without the antialiased mode."""
    adj = _adjoint(adjoint)
    N = means3d.shape[0] if isinstance(means3d, torch.Tensor) else -1
    nch = colors.shape[1] if isinstance(colors, torch.Tensor) and colors.dim() == 2 else 3
    if nch < 1:
        raise ShapeDtypeError("colors: need at least one channel")
    _expect("means3d", means3d, (N, 3), torch.float32)
    _expect("scales", scales, (N, 3), torch.float32)
    _expect("quats", quats, (N, 4), torch.float32)
    _expect("colors", colors, (N, nch), torch.float32)
    _expect("opacities", opacities, (N, 1), torch.float32)
    _expect("viewmat", viewmat, (4, 4), torch.float32)
    _expect("background", background, (nch,), torch.float32)
    if isect_capacity is not None and not isinstance(isect_capacity, IsectCapacity):
        raise TypeError("isect_capacity must be an IsectCapacity")
    cfg = (tuple(img_shape), tuple(f), tuple(c), float(glob_scale), float(clip_thresh),
           int(block_width), adj)
    return _Render.apply(cfg, means3d.contiguous(), scales.contiguous(), quats.contiguous(),
                         colors.contiguous(), opacities.contiguous(), viewmat.contiguous(),
                         background.contiguous(), isect_capacity)


def rasterize(colors, opacities, background, xys, depths, radii, conics, cum_tiles_hit, *,
              img_shape, block_width, adjoint="reference", tile_rects=None,
              isect_capacity=None):
    """Rasterize projected Gaussians (reference ``_rasterize/__init__.py:40-68``).

    Returns the (H, W, C) image for colors (N, C) and background (C,).  The
    reference API has C = 3; other channel counts use the N-channel
    rasterizer (``lib/kernels/forward.cu:213-337``, never called by the
    reference), with the same blend per channel in fp32.

    ``tile_rects`` (N, 2) int32 (not in the reference API): the
    opacity-limited tile rectangles of ``project_for_render``, whose counts
    ``cum_tiles_hit`` then are.

    ``isect_capacity`` (an ``IsectCapacity``, not in the reference API): bin
    into buffers of that fixed capacity, counting on the device, instead of
    reading the intersection count back to the host -- no host
    synchronisation, so the call may be captured in a HIP graph.
    """
    adj = _adjoint(adjoint)
    N = colors.shape[0] if isinstance(colors, torch.Tensor) else -1
    nch = colors.shape[1] if isinstance(colors, torch.Tensor) and colors.dim() == 2 else 3
    if nch < 1:
        raise ShapeDtypeError("colors: need at least one channel")
    _expect("colors", colors, (N, nch), torch.float32)
    _expect("opacities", opacities, (N, 1), torch.float32)
    _expect("background", background, (nch,), torch.float32)
    _expect("xys", xys, (N, 2), torch.float32)
    _expect("depths", depths, (N, 1), torch.float32)
    _expect("radii", radii, (N, 1), torch.int32)
    _expect("conics", conics, (N, 3), torch.float32)
    _expect("cum_tiles_hit", cum_tiles_hit, (N, 1), torch.int32)
    if tile_rects is not None:
        _expect("tile_rects", tile_rects, (N, 2), torch.int32)
        tile_rects = tile_rects.contiguous()
    if isect_capacity is not None and not isinstance(isect_capacity, IsectCapacity):
        raise TypeError("isect_capacity must be an IsectCapacity")
    cfg = (tuple(img_shape), int(block_width), adj)
    return _Rasterize.apply(cfg, colors.contiguous(), opacities.contiguous(),
                            background.contiguous(), xys.contiguous(), depths.contiguous(),
                            radii.contiguous(), conics.contiguous(), cum_tiles_hit.contiguous(),
                            tile_rects, isect_capacity)
