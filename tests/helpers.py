"""Shared test helpers: run each stage on the GPU through the C ABI (ctypes +
torch device buffers) and compare against the oracle."""

from __future__ import annotations

import ctypes

import numpy as np

from jaxsplat_amd import _lib


def torch():
    import torch as _t
    return _t


def dev(a, dtype=None):
    t = torch()
    a = np.ascontiguousarray(a)
    out = t.from_numpy(a).to("cuda")
    return out if dtype is None else out.to(dtype)


def host(t):
    torch().cuda.synchronize()
    return t.detach().cpu().numpy()


def P(t):
    return ctypes.c_void_p(t.data_ptr()) if t is not None else None


def stream():
    t = torch()
    return ctypes.c_void_p(t.cuda.current_stream().cuda_stream)


def descriptor(n, img_shape, f=(1.0, 1.0), c=(0.0, 0.0), glob_scale=1.0, clip_thresh=0.01,
               block_width=16):
    return _lib.make_descriptor(n, img_shape, f, c, glob_scale, clip_thresh, block_width)


def gpu_project(means3d, scales, quats, viewmat, *, img_shape, f, c, glob_scale=1.0,
                clip_thresh=0.01, block_width=16):
    t = torch()
    N = means3d.shape[0]
    d = descriptor(N, img_shape, f, c, glob_scale, clip_thresh, block_width)
    L = _lib.lib()
    ins = [dev(np.asarray(a, np.float32)) for a in (means3d, scales, quats, viewmat)]
    outs = dict(
        cov3d=t.full((N, 6), 7.0, device="cuda"), xys=t.full((N, 2), 7.0, device="cuda"),
        depths=t.full((N,), 7.0, device="cuda"), radii=t.full((N,), 7, dtype=t.int32, device="cuda"),
        conics=t.full((N, 3), 7.0, device="cuda"), compensation=t.full((N,), 7.0, device="cuda"),
        num_tiles_hit=t.full((N,), 7, dtype=t.int32, device="cuda"),
        cum_tiles_hit=t.full((N,), 7, dtype=t.int32, device="cuda"))
    wsb = L.jaxsplat_project_workspace_size(ctypes.byref(d))
    ws = t.empty(max(wsb, 1), dtype=t.uint8, device="cuda")
    _lib.check(L.jaxsplat_project_forward(
        ctypes.byref(d), *[P(x) for x in ins], *[P(outs[k]) for k in (
            "cov3d", "xys", "depths", "radii", "conics", "compensation", "num_tiles_hit",
            "cum_tiles_hit")], P(ws), wsb, stream()))
    return {k: host(v) for k, v in outs.items()}


def packed_rects(xys, radii, img_shape, block_width):
    """The reference tile rectangles packed as the library packs them
    (x0 | y0 << 16, w | h << 16), zero for radius 0, as (N, 2) int32."""
    import oracle
    H_, W_ = img_shape
    tx = (W_ + block_width - 1) // block_width
    ty = (H_ + block_width - 1) // block_width
    xys = np.asarray(xys, np.float32).reshape(-1, 2)
    radii = np.asarray(radii).reshape(-1)
    x0, x1, y0, y1 = oracle.project.tile_rect(xys[:, 0], xys[:, 1], radii.astype(np.float32),
                                              tx, ty, block_width)
    on = radii > 0
    w = np.where(on, x1 - x0, 0).astype(np.uint32)
    h = np.where(on, y1 - y0, 0).astype(np.uint32)
    lo = np.where(on, x0.astype(np.uint32) | (y0.astype(np.uint32) << 16), 0).astype(np.uint32)
    return np.stack([lo, w | (h << 16)], 1).view(np.int32)


def gpu_bin(xys, depths, radii, cum_tiles_hit, *, img_shape, block_width=16, split=False,
            rects=False):
    """Binning through jaxsplat_bin_and_sort, or (split) through the depth
    stage and the emission stage with separate workspaces, or (rects) through
    jaxsplat_bin_rects_sort with the reference rectangles packed on the host
    and cum_tiles_hit."""
    t = torch()
    N = xys.shape[0]
    d = descriptor(N, img_shape, block_width=block_width)
    L = _lib.lib()
    isect = int(np.asarray(cum_tiles_hit).reshape(-1)[-1]) if N else 0
    ntiles = d.grid_dim_2d.x * d.grid_dim_2d.y
    gids = t.full((max(isect, 1),), -5, dtype=t.int32, device="cuda")
    tiles = t.full((max(isect, 1),), -5, dtype=t.int32, device="cuda")
    bins = t.full((ntiles, 2), -5, dtype=t.int32, device="cuda")
    x = dev(np.asarray(xys, np.float32).reshape(N, 2))
    dp = dev(np.asarray(depths, np.float32).reshape(N))
    r = dev(np.asarray(radii, np.int32).reshape(N))
    if rects:
        rc = dev(packed_rects(xys, radii, img_shape, block_width))
        cm = dev(np.asarray(cum_tiles_hit, np.int32).reshape(N))
        wsb = L.jaxsplat_bin_emit_workspace_size(ctypes.byref(d), isect)
        ws = t.full((max(wsb, 1),), 0xA5, dtype=t.uint8, device="cuda")
        _lib.check(L.jaxsplat_bin_rects_sort(ctypes.byref(d), P(dp), P(rc), P(cm), isect,
                                             P(gids), P(tiles), P(bins), P(ws), wsb, stream()))
    elif split:
        dwb = L.jaxsplat_bin_depth_workspace_size(ctypes.byref(d))
        dws = t.full((max(dwb, 1),), 0xA5, dtype=t.uint8, device="cuda")
        _lib.check(L.jaxsplat_bin_depth_sort(ctypes.byref(d), P(x), P(dp), P(r), P(dws), dwb,
                                             stream()))
        wsb = L.jaxsplat_bin_emit_workspace_size(ctypes.byref(d), isect)
        ws = t.full((max(wsb, 1),), 0xA5, dtype=t.uint8, device="cuda")
        _lib.check(L.jaxsplat_bin_emit_sort(ctypes.byref(d), isect, P(dws), P(gids), P(tiles),
                                            P(bins), P(ws), wsb, stream()))
    else:
        wsb = L.jaxsplat_bin_workspace_size(ctypes.byref(d), isect)
        ws = t.empty(max(wsb, 1), dtype=t.uint8, device="cuda")
        _lib.check(L.jaxsplat_bin_and_sort(ctypes.byref(d), P(x), P(dp), P(r), isect, P(gids),
                                           P(tiles), P(bins), P(ws), wsb, stream()))
    _lib.check(L.jaxsplat_bin_check(ctypes.byref(d), isect, P(ws), stream()))
    return host(gids)[:isect], host(tiles)[:isect], host(bins), isect


def gpu_tile_rects(xys, radii, conics, opacities, *, img_shape, block_width=16):
    """jaxsplat_tile_rects: opacity-limited rectangles (N, 2) uint32 and the
    inclusive sums of their tile counts."""
    t = torch()
    N = xys.shape[0]
    d = descriptor(N, img_shape, block_width=block_width)
    L = _lib.lib()
    rects = t.empty((max(N, 1), 2), dtype=t.int32, device="cuda")
    cum = t.empty((max(N, 1),), dtype=t.int32, device="cuda")
    wsb = L.jaxsplat_tile_rects_workspace_size(ctypes.byref(d))
    ws = t.empty(max(wsb, 1), dtype=t.uint8, device="cuda")
    ins = [dev(np.asarray(xys, np.float32).reshape(N, 2)),
           dev(np.asarray(radii, np.int32).reshape(N)),
           dev(np.asarray(conics, np.float32).reshape(N, 3)),
           dev(np.asarray(opacities, np.float32).reshape(N))]
    _lib.check(L.jaxsplat_tile_rects(ctypes.byref(d), *[P(x) for x in ins], P(rects), None,
                                     P(cum), P(ws), wsb, stream()))
    return host(rects)[:N].view(np.uint32), host(cum)[:N]


def gpu_raster_fwd(gids, bins, isect, xys, conics, colors, opacities, background, *, img_shape,
                   block_width=16):
    t = torch()
    N = xys.shape[0]
    H, W = img_shape
    d = descriptor(N, img_shape, block_width=block_width)
    L = _lib.lib()
    fT = t.full((H, W), 7.0, device="cuda")
    fI = t.full((H, W), 7, dtype=t.int32, device="cuda")
    img = t.full((H, W, 3), 7.0, device="cuda")
    g = dev(np.asarray(gids, np.int32).reshape(-1) if isect else np.zeros(1, np.int32))
    b = dev(np.asarray(bins, np.int32))
    ins = [dev(np.asarray(a, np.float32)) for a in (xys, conics, colors, opacities, background)]
    _lib.check(L.jaxsplat_rasterize_forward_binned(
        ctypes.byref(d), P(g), P(b), isect, *[P(x) for x in ins], P(fT), P(fI), P(img),
        stream()))
    return dict(final_Ts=host(fT), final_idx=host(fI), img=host(img))


def gpu_raster_bwd(gids, bins, isect, xys, conics, colors, opacities, background, final_Ts,
                   final_idx, v_img, v_alpha, *, img_shape, block_width=16, adjoint=0,
                   with_abs=True):
    t = torch()
    N = xys.shape[0]
    d = descriptor(N, img_shape, block_width=block_width)
    L = _lib.lib()
    outs = dict(v_colors=t.full((N, 3), 7.0, device="cuda"), v_opacity=t.full((N,), 7.0, device="cuda"),
                v_xy=t.full((N, 2), 7.0, device="cuda"),
                v_xy_abs=t.full((N, 2), 7.0, device="cuda") if with_abs else None,
                v_conic=t.full((N, 3), 7.0, device="cuda"))
    g = dev(np.asarray(gids, np.int32).reshape(-1) if isect else np.zeros(1, np.int32))
    b = dev(np.asarray(bins, np.int32))
    ins = [dev(np.asarray(a, np.float32)) for a in (xys, conics, colors, opacities, background,
                                                    final_Ts)]
    fI = dev(np.asarray(final_idx, np.int32))
    vi = dev(np.asarray(v_img, np.float32))
    va = dev(np.asarray(v_alpha, np.float32)) if v_alpha is not None else None
    _lib.check(L.jaxsplat_rasterize_backward_binned(
        ctypes.byref(d), P(g), P(b), isect, *[P(x) for x in ins], P(fI), P(vi), P(va),
        P(outs["v_colors"]), P(outs["v_opacity"]), P(outs["v_xy"]), P(outs["v_xy_abs"]),
        P(outs["v_conic"]), adjoint, stream()))
    return {k: host(v) for k, v in outs.items() if v is not None}


def gpu_project_bwd(means3d, scales, quats, viewmat, radii, conics, v_xy, v_depth, v_conic, *,
                    img_shape, f, c, glob_scale=1.0, clip_thresh=0.01, block_width=16,
                    adjoint=0, compensation=None, v_compensation=None):
    t = torch()
    N = means3d.shape[0]
    d = descriptor(N, img_shape, f, c, glob_scale, clip_thresh, block_width)
    L = _lib.lib()
    ins = [dev(np.asarray(a, np.float32)) for a in (means3d, scales, quats, viewmat)]
    r = dev(np.asarray(radii, np.int32).reshape(N))
    cn = dev(np.asarray(conics, np.float32))
    comp = dev(np.asarray(compensation, np.float32)) if compensation is not None else None
    vxy = dev(np.asarray(v_xy, np.float32))
    vd = dev(np.asarray(v_depth, np.float32)) if v_depth is not None else None
    vc = dev(np.asarray(v_conic, np.float32))
    vcomp = dev(np.asarray(v_compensation, np.float32)) if v_compensation is not None else None
    outs = dict(v_mean3d=t.full((N, 3), 7.0, device="cuda"), v_scale=t.full((N, 3), 7.0, device="cuda"),
                v_quat=t.full((N, 4), 7.0, device="cuda"), v_cov2d=t.full((N, 3), 7.0, device="cuda"),
                v_cov3d=t.full((N, 6), 7.0, device="cuda"))
    _lib.check(L.jaxsplat_project_backward(
        ctypes.byref(d), *[P(x) for x in ins], P(r), P(cn), P(comp), P(vxy), P(vd), P(vc),
        P(vcomp), *[P(outs[k]) for k in ("v_mean3d", "v_scale", "v_quat", "v_cov2d", "v_cov3d")],
        adjoint, stream()))
    return {k: host(v) for k, v in outs.items()}


def grad_close(got, ref, rtol=1e-3, scale_frac=1e-4):
    """Element-wise |got-ref| <= rtol*|ref| + scale_frac*max|ref|; returns (ok, worst)."""
    got = np.asarray(got, np.float64)
    ref = np.asarray(ref, np.float64)
    scale = float(np.abs(ref).max()) if ref.size else 0.0
    err = np.abs(got - ref)
    tol = rtol * np.abs(ref) + scale_frac * scale + 1e-30
    worst = float((err / tol).max()) if ref.size else 0.0
    return worst <= 1.0, worst
