"""Rasterize forward/backward oracle in float64 (test infrastructure only).

Restates ``rasterize_fwd`` (reference ``lib/kernels/forward.cu:339-462``) and
``rasterize_bwd`` (``lib/kernels/backward.cu:153-343``), plus the
zero-intersection paths of ``lib/ops.cu:119-127,184-186`` (image = background,
``final_Ts = 0``, ``final_idx = 0``, zero gradients).

Per-pixel semantics (forward, front to back over the tile's sorted list):
``sigma = 0.5*(a dx^2 + c dy^2) + b dx dy`` with ``d = xy - (j+0.5, i+0.5)``;
``alpha = min(0.999, o*exp(-sigma))``; skip if ``sigma < 0`` or
``alpha < 1/255``; stop (without adding) once ``T*(1-alpha) <= 1e-4``;
``C += c*alpha*T``; ``final_idx`` = global sorted index of the last added
Gaussian (0 if none); ``out = C + T*bg``.

Backward (reference mode, ``backward.cu:242-313``): back to front over the
sorted indices ``<= final_idx`` with ``alpha = min(0.99, o*vis)``, T recovered
by division, and the clamp's zero slope ignored.  ``mode="consistent"`` uses
0.999 and zeroes the alpha slope where the clamp is active, which makes it the
exact adjoint of the forward.

The vectorised functions work per tile with cumulative products; the
``*_bruteforce`` functions are plain sequential loops for tiny inputs.

``ambiguous`` marks pixels where a float32 evaluation may take a different
discrete decision than float64 (alpha within 1e-4 relative of 1/255, the
transmittance within its float32 error estimate of 1e-4, or sigma within
float32 rounding of 0).  Parity tests exclude (and count) those pixels.
"""

from __future__ import annotations

import numpy as np

ALPHA_MIN = 1.0 / 255.0
T_EPS = 1e-4


def _tile_pixels(tile, tiles_x, bw, H, W):
    ty, tx = divmod(tile, tiles_x)
    ii, jj = np.meshgrid(np.arange(ty * bw, ty * bw + bw),
                         np.arange(tx * bw, tx * bw + bw), indexing="ij")
    ii, jj = ii.reshape(-1), jj.reshape(-1)
    keep = (ii < H) & (jj < W)
    return ii[keep], jj[keep]


def _pair_terms(xy, conic, px, py):
    dx = xy[None, :, 0] - px[:, None]
    dy = xy[None, :, 1] - py[:, None]
    a, b, c = conic[None, :, 0], conic[None, :, 1], conic[None, :, 2]
    t1, t2, t3 = 0.5 * (a * dx * dx), 0.5 * (c * dy * dy), b * dx * dy
    sigma = (t1 + t2) + t3
    return dx, dy, sigma, np.abs(t1) + np.abs(t2) + np.abs(t3)


def rasterize_fwd(gids_sorted, tile_bins, xys, conics, colors, opacities,
                  background, *, img_shape, block_width, num_intersects=None,
                  tiles=None, chunk=2048):
    """Vectorised forward.  Returns dict img (H,W,C), final_Ts (H,W),
    final_idx (H,W) int32, ambiguous (H,W) bool, for colors (N,C) and
    background (C,) (C = 3 in the reference API; any C for the N-channel
    rasterizer, whose blend is the same per channel).  ``tiles`` restricts
    the work to a subset of tile ids (other pixels stay zero / unflagged)."""
    H, W = img_shape
    bw = block_width
    tiles_x = (W + bw - 1) // bw
    tiles_y = (H + bw - 1) // bw
    xys = np.asarray(xys, np.float64).reshape(-1, 2)
    conics = np.asarray(conics, np.float64).reshape(-1, 3)
    colors = np.asarray(colors, np.float64)
    colors = colors.reshape(colors.shape[0], -1)
    nch = colors.shape[1]
    opac = np.asarray(opacities, np.float64).reshape(-1)
    bg = np.asarray(background, np.float64).reshape(nch)
    gids_sorted = np.asarray(gids_sorted).reshape(-1)
    tile_bins = np.asarray(tile_bins).reshape(-1, 2)
    if num_intersects is None:
        num_intersects = gids_sorted.size
    img = np.zeros((H, W, nch))
    final_T = np.zeros((H, W))
    final_idx = np.zeros((H, W), np.int32)
    amb = np.zeros((H, W), bool)
    if num_intersects == 0:                      # ops.cu:119-127
        img[:] = bg
        return dict(img=img, final_Ts=final_T, final_idx=final_idx, ambiguous=amb)
    tile_list = range(tiles_x * tiles_y) if tiles is None else tiles
    for tile in tile_list:
        ii, jj = _tile_pixels(int(tile), tiles_x, bw, H, W)
        px, py = jj + 0.5, ii + 0.5
        P = ii.size
        T = np.ones(P)
        C = np.zeros((P, nch))
        last = np.zeros(P, np.int64)
        done = np.zeros(P, bool)
        ab = np.zeros(P, bool)
        Terr = np.zeros(P)                        # relative float32 error estimate of T
        s, e = int(tile_bins[tile, 0]), int(tile_bins[tile, 1])
        for k0 in range(s, e, chunk):
            if done.all():
                break
            k1 = min(e, k0 + chunk)
            ids = gids_sorted[k0:k1]
            dx, dy, sigma, tsum = _pair_terms(xys[ids], conics[ids], px, py)
            araw = opac[None, ids] * np.exp(-sigma)
            alpha = np.minimum(0.999, araw)
            live = ~done[:, None]
            valid = (sigma >= 0) & (alpha >= ALPHA_MIN) & live
            om = np.where(valid, 1.0 - alpha, 1.0)
            cp = np.cumprod(om, axis=1)
            Tafter = T[:, None] * cp
            Tbefore = T[:, None] * np.concatenate([np.ones((P, 1)), cp[:, :-1]], axis=1)
            stop = valid & (Tafter <= T_EPS)
            anystop = stop.any(axis=1)
            kstop = np.where(anystop, stop.argmax(axis=1), k1 - k0)
            kk = np.arange(k1 - k0)[None, :]
            contrib = valid & (kk < kstop[:, None])
            considered = live & (kk <= kstop[:, None])
            # float32 decision ambiguity
            ea = 4e-7 * tsum + 3e-7 * (1.0 + sigma)
            erel = np.cumsum(np.where(valid, alpha / (1.0 - alpha) * ea, 0.0), axis=1) \
                + Terr[:, None] + 1e-7 * (kk + 1)
            near_a = np.abs(araw * 255.0 - 1.0) < 4.0 * ea + 1e-6
            near_t = valid & (np.abs(Tafter / T_EPS - 1.0) < 8.0 * erel + 1e-6)
            near_s = (sigma < 4e-7 * tsum) & (opac[None, ids] >= ALPHA_MIN)
            ab |= (considered & (near_a | near_t | near_s)).any(axis=1)
            w = np.where(contrib, alpha * Tbefore, 0.0)
            C += w @ colors[ids]
            has = contrib.any(axis=1)
            lastk = (k1 - k0 - 1) - np.argmax(contrib[:, ::-1], axis=1)
            last = np.where(has, k0 + lastk, last)
            idx_end = np.minimum(kstop, k1 - k0 - 1)
            T = np.where(anystop, Tbefore[np.arange(P), idx_end], Tafter[:, -1])
            Terr = np.take_along_axis(erel, idx_end[:, None], axis=1)[:, 0]
            done |= anystop
        img[ii, jj] = C + T[:, None] * bg[None, :]
        final_T[ii, jj] = T
        final_idx[ii, jj] = last
        amb[ii, jj] = ab
    return dict(img=img, final_Ts=final_T, final_idx=final_idx, ambiguous=amb)


def rasterize_bwd(gids_sorted, tile_bins, xys, conics, colors, opacities,
                  background, final_Ts, final_idx, v_img, v_img_alpha, *,
                  img_shape, block_width, mode="reference", num_intersects=None,
                  tiles=None, pix_chunk=64):
    """Vectorised backward.  Returns dict v_colors (N,C), v_opacity (N,),
    v_xy (N,2), v_xy_abs (N,2), v_conic (N,3) for colors (N,C)."""
    if mode not in ("reference", "consistent"):
        raise ValueError(mode)
    clamp = 0.99 if mode == "reference" else 0.999
    H, W = img_shape
    bw = block_width
    tiles_x = (W + bw - 1) // bw
    tiles_y = (H + bw - 1) // bw
    xys = np.asarray(xys, np.float64).reshape(-1, 2)
    N = xys.shape[0]
    conics = np.asarray(conics, np.float64).reshape(-1, 3)
    colors = np.asarray(colors, np.float64)
    colors = colors.reshape(colors.shape[0], -1)
    nch = colors.shape[1]
    opac = np.asarray(opacities, np.float64).reshape(-1)
    bg = np.asarray(background, np.float64).reshape(nch)
    fT = np.asarray(final_Ts, np.float64).reshape(H, W)
    fI = np.asarray(final_idx, np.int64).reshape(H, W)
    vimg = np.asarray(v_img, np.float64).reshape(H, W, nch)
    valpha = np.asarray(v_img_alpha, np.float64).reshape(H, W)
    gids_sorted = np.asarray(gids_sorted).reshape(-1)
    tile_bins = np.asarray(tile_bins).reshape(-1, 2)
    if num_intersects is None:
        num_intersects = gids_sorted.size
    out = dict(v_colors=np.zeros((N, nch)), v_opacity=np.zeros(N), v_xy=np.zeros((N, 2)),
               v_xy_abs=np.zeros((N, 2)), v_conic=np.zeros((N, 3)))
    if num_intersects == 0:                      # ops.cu:184-186
        return out
    acc = np.zeros((N, nch + 8))
    tile_list = range(tiles_x * tiles_y) if tiles is None else tiles
    for tile in tile_list:
        s, e = int(tile_bins[tile, 0]), int(tile_bins[tile, 1])
        if e <= s:
            continue
        ii_all, jj_all = _tile_pixels(int(tile), tiles_x, bw, H, W)
        ids = gids_sorted[s:e]
        xy, cn, rgb, op = xys[ids], conics[ids], colors[ids], opac[ids]
        for p0 in range(0, ii_all.size, pix_chunk):
            ii, jj = ii_all[p0:p0 + pix_chunk], jj_all[p0:p0 + pix_chunk]
            px, py = jj + 0.5, ii + 0.5
            Tf = fT[ii, jj]
            fin = fI[ii, jj] - s
            vo = vimg[ii, jj]
            voa = valpha[ii, jj]
            dx, dy, sigma, _ = _pair_terms(xy, cn, px, py)
            vis = np.exp(-sigma)
            raw = op[None, :] * vis
            alpha = np.minimum(clamp, raw)
            kk = np.arange(e - s)[None, :]
            m = (sigma >= 0) & (alpha >= ALPHA_MIN) & (kk <= fin[:, None])
            om = np.where(m, 1.0 - alpha, 1.0)
            Prev = np.cumprod(om[:, ::-1], axis=1)[:, ::-1]     # prod_{j>=k}
            Tk = Tf[:, None] / Prev
            fac = np.where(m, alpha * Tk, 0.0)
            contrib = rgb[None, :, :] * fac[:, :, None]           # (P,K,C)
            csum = np.cumsum(contrib[:, ::-1, :], axis=1)[:, ::-1, :]
            Sbuf = csum - contrib                                 # sum over j > k
            ra = 1.0 / (1.0 - alpha)
            va = ((rgb[None] * Tk[:, :, None] - Sbuf * ra[:, :, None]) * vo[:, None, :]).sum(-1)
            va = va + Tf[:, None] * ra * voa[:, None]
            va = va - Tf[:, None] * ra * (bg[None, :] * vo).sum(-1)[:, None]
            va = np.where(m, va, 0.0)
            slope = np.ones_like(raw) if mode == "reference" else (raw < clamp).astype(float)
            vsig = -op[None, :] * vis * va * slope
            vcon = np.stack([0.5 * vsig * dx * dx, vsig * dx * dy, 0.5 * vsig * dy * dy], -1)
            vxy = np.stack([vsig * (cn[None, :, 0] * dx + cn[None, :, 1] * dy),
                            vsig * (cn[None, :, 1] * dx + cn[None, :, 2] * dy)], -1)
            vop = vis * va * slope
            vrgb = fac[:, :, None] * vo[:, None, :]
            per = np.concatenate([vrgb, vcon, vxy, np.abs(vxy), vop[:, :, None]], -1)
            per = np.where(m[:, :, None], per, 0.0)
            np.add.at(acc, ids, per.sum(axis=0))
    out["v_colors"] = acc[:, 0:nch]
    out["v_conic"] = acc[:, nch:nch + 3]
    out["v_xy"] = acc[:, nch + 3:nch + 5]
    out["v_xy_abs"] = acc[:, nch + 5:nch + 7]
    out["v_opacity"] = acc[:, nch + 7]
    return out


# ----------------------------------------------------------------------------
# Brute-force sequential restatements (tiny inputs only)
# ----------------------------------------------------------------------------

def rasterize_fwd_bruteforce(gids_sorted, tile_bins, xys, conics, colors,
                             opacities, background, *, img_shape, block_width):
    import math
    H, W = img_shape
    bw = block_width
    tiles_x = (W + bw - 1) // bw
    opacities = np.asarray(opacities, np.float64).reshape(-1)
    nch = np.asarray(colors).reshape(len(opacities), -1).shape[1]
    img = np.zeros((H, W, nch))
    fT = np.zeros((H, W))
    fI = np.zeros((H, W), np.int32)
    for i in range(H):
        for j in range(W):
            tile = (i // bw) * tiles_x + (j // bw)
            s, e = int(tile_bins[tile][0]), int(tile_bins[tile][1])
            T, C, cur = 1.0, [0.0] * nch, 0
            for k in range(s, e):
                g = int(gids_sorted[k])
                dx = float(xys[g][0]) - (j + 0.5)
                dy = float(xys[g][1]) - (i + 0.5)
                a, b, c = (float(v) for v in conics[g])
                sigma = 0.5 * (a * dx * dx + c * dy * dy) + b * dx * dy
                alpha = min(0.999, float(opacities[g]) * math.exp(-sigma))
                if sigma < 0 or alpha < ALPHA_MIN:
                    continue
                nT = T * (1.0 - alpha)
                if nT <= T_EPS:
                    break
                for ch in range(nch):
                    C[ch] += float(colors[g][ch]) * alpha * T
                T = nT
                cur = k
            for ch in range(nch):
                img[i, j, ch] = C[ch] + T * float(background[ch])
            fT[i, j] = T
            fI[i, j] = cur
    return dict(img=img, final_Ts=fT, final_idx=fI)


def rasterize_bwd_bruteforce(gids_sorted, tile_bins, xys, conics, colors,
                             opacities, background, final_Ts, final_idx, v_img,
                             v_img_alpha, *, img_shape, block_width,
                             mode="reference"):
    import math
    clamp = 0.99 if mode == "reference" else 0.999
    H, W = img_shape
    bw = block_width
    tiles_x = (W + bw - 1) // bw
    N = len(xys)
    opacities = np.asarray(opacities, np.float64).reshape(-1)
    nch = np.asarray(colors).reshape(N, -1).shape[1]
    vc = np.zeros((N, nch))
    vo = np.zeros(N)
    vxy = np.zeros((N, 2))
    vxya = np.zeros((N, 2))
    vcn = np.zeros((N, 3))
    for i in range(H):
        for j in range(W):
            tile = (i // bw) * tiles_x + (j // bw)
            s, e = int(tile_bins[tile][0]), int(tile_bins[tile][1])
            Tf = float(final_Ts[i][j])
            T = Tf
            buf = [0.0] * nch
            fin = int(final_idx[i][j])
            v = [float(x) for x in v_img[i][j]]
            va_out = float(v_img_alpha[i][j])
            for k in range(e - 1, s - 1, -1):
                if k > fin:
                    continue
                g = int(gids_sorted[k])
                dx = float(xys[g][0]) - (j + 0.5)
                dy = float(xys[g][1]) - (i + 0.5)
                a, b, c = (float(x) for x in conics[g])
                sigma = 0.5 * (a * dx * dx + c * dy * dy) + b * dx * dy
                vis = math.exp(-sigma)
                raw = float(opacities[g]) * vis
                alpha = min(clamp, raw)
                if sigma < 0 or alpha < ALPHA_MIN:
                    continue
                ra = 1.0 / (1.0 - alpha)
                T *= ra
                fac = alpha * T
                rgb = [float(x) for x in colors[g]]
                v_alpha = 0.0
                for ch in range(nch):
                    vc[g, ch] += fac * v[ch]
                    v_alpha += (rgb[ch] * T - buf[ch] * ra) * v[ch]
                v_alpha += Tf * ra * va_out
                for ch in range(nch):
                    v_alpha += -Tf * ra * float(background[ch]) * v[ch]
                for ch in range(nch):
                    buf[ch] += rgb[ch] * fac
                slope = 1.0 if (mode == "reference" or raw < clamp) else 0.0
                v_sigma = -float(opacities[g]) * vis * v_alpha * slope
                vcn[g] += [0.5 * v_sigma * dx * dx, v_sigma * dx * dy, 0.5 * v_sigma * dy * dy]
                lx = v_sigma * (a * dx + b * dy)
                ly = v_sigma * (b * dx + c * dy)
                vxy[g] += [lx, ly]
                vxya[g] += [abs(lx), abs(ly)]
                vo[g] += vis * v_alpha * slope
    return dict(v_colors=vc, v_opacity=vo, v_xy=vxy, v_xy_abs=vxya, v_conic=vcn)
