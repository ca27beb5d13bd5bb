"""Independent float64 torch forward of the render path, used only to pin the
oracle's adjoints with autograd.

Written from the mathematics (EWA splatting, front-to-back alpha blending),
not from the oracle: a quaternion rotation, Sigma = R S S^T R^T, the pinhole
Jacobian, the +0.3 blur, the 2x2 inverse, the per-pixel blend.  The discrete
decisions of the blend (which pairs are valid, where a pixel stops) are
passed in as masks, so autograd differentiates exactly the smooth function
that the adjoint kernels differentiate.
"""

from __future__ import annotations

import numpy as np
import torch


def rotmat(q):
    w, x, y, z = q.unbind(-1)
    return torch.stack([
        torch.stack([1 - 2 * (y * y + z * z), 2 * (x * y - w * z), 2 * (x * z + w * y)], -1),
        torch.stack([2 * (x * y + w * z), 1 - 2 * (x * x + z * z), 2 * (y * z - w * x)], -1),
        torch.stack([2 * (x * z - w * y), 2 * (y * z + w * x), 1 - 2 * (x * x + y * y)], -1),
    ], -2)


def project(means, scales, quats, viewmat, fx, fy, cx, cy, img_shape, glob_scale=1.0,
            clamp=True, jac_unclamped=False):
    """Returns xys (N,2), conic (N,3) = (a, b, c) of Sigma2^-1, depth (N,).

    clamp: clamp the camera-space mean to 1.3 x the field of view inside the
    Jacobian (the forward); jac_unclamped: build the Jacobian from the
    unclamped mean (the value used by the reference backward)."""
    H, W = img_shape
    Rv = viewmat[:3, :3]
    tv = viewmat[:3, 3]
    t = means @ Rv.T + tv
    tx, ty, tz = t.unbind(-1)
    if clamp and not jac_unclamped:
        lx, ly = 1.3 * (0.5 * W / fx), 1.3 * (0.5 * H / fy)
        tx = tz * torch.clamp(tx / tz, -lx, lx)
        ty = tz * torch.clamp(ty / tz, -ly, ly)
    R = rotmat(quats)
    S = torch.diag_embed(glob_scale * scales)
    M = R @ S
    Sig = M @ M.transpose(-1, -2)
    zero = torch.zeros_like(tz)
    J = torch.stack([
        torch.stack([fx / tz, zero, -fx * tx / tz ** 2], -1),
        torch.stack([zero, fy / tz, -fy * ty / tz ** 2], -1),
    ], -2)
    T = J @ Rv
    C2 = T @ Sig @ T.transpose(-1, -2) + 0.3 * torch.eye(2, dtype=means.dtype)
    det = C2[:, 0, 0] * C2[:, 1, 1] - C2[:, 0, 1] * C2[:, 1, 0]
    conic = torch.stack([C2[:, 1, 1] / det, -C2[:, 0, 1] / det, C2[:, 0, 0] / det], -1)
    x, y, z = t.unbind(-1)
    xys = torch.stack([fx * x / (z + 1e-6) + cx, fy * y / (z + 1e-6) + cy], -1)
    return xys, conic, z


def blend(xys, conics, colors, opac, bg, pix, order, valid, clamp_val=0.999):
    """Per-pixel front-to-back blend.

    pix: (P, 2) pixel centres; order: list (per pixel) of Gaussian ids in
    depth order that contribute (valid and before termination); valid pairs
    are given by `valid` (list of bool lists, same shape as order).  Returns
    colours (P, 3) and final transmittance (P,).
    """
    outs, Ts = [], []
    for p in range(pix.shape[0]):
        T = torch.ones((), dtype=xys.dtype)
        C = torch.zeros(3, dtype=xys.dtype)
        for g, ok in zip(order[p], valid[p]):
            if not ok:
                continue
            d = xys[g] - pix[p]
            a, b, c = conics[g]
            sig = 0.5 * (a * d[0] ** 2 + c * d[1] ** 2) + b * d[0] * d[1]
            al = opac[g] * torch.exp(-sig)
            al = torch.minimum(al, torch.tensor(clamp_val, dtype=xys.dtype))
            C = C + colors[g] * al * T
            T = T * (1 - al)
        outs.append(C + T * bg)
        Ts.append(T)
    return torch.stack(outs), torch.stack(Ts)


def decisions(xys, conics, opac, pix, order):
    """Discrete decisions of the reference forward for each pixel (float64)."""
    xys = np.asarray(xys, np.float64)
    conics = np.asarray(conics, np.float64)
    opac = np.asarray(opac, np.float64).reshape(-1)
    keep, flags = [], []
    for p in range(pix.shape[0]):
        T = 1.0
        ids, ok = [], []
        for g in order[p]:
            d = xys[g] - pix[p]
            a, b, c = conics[g]
            sig = 0.5 * (a * d[0] ** 2 + c * d[1] ** 2) + b * d[0] * d[1]
            al = min(0.999, opac[g] * np.exp(-sig))
            if sig < 0 or al < 1.0 / 255.0:
                continue
            if T * (1 - al) <= 1e-4:
                break
            ids.append(int(g))
            ok.append(True)
            T *= 1 - al
        keep.append(ids)
        flags.append(ok)
    return keep, flags
