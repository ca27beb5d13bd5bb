"""Projection forward/backward oracle (test infrastructure only).

Restates ``project_gaussians_fwd`` (reference ``lib/kernels/forward.cu:13-131``)
with ``scale_rot_to_cov3d`` (``forward.cu:547-568``), ``project_cov3d_ewa``
(``forward.cu:465-544``) and the helpers of ``lib/kernels/helpers.h``
(``clip_near_plane`` :204-211, ``quat_to_rotmat`` :140-153, ``scale_to_mat``
:194-201, ``compute_cov2d_bounds`` :34-57, ``project_pix`` :122-128,
``get_tile_bbox``/``get_bbox`` :9-32), and ``project_gaussians_bwd``
(``lib/kernels/backward.cu:345-418``) with ``project_cov3d_ewa_vjp``
(:421-524), ``scale_rot_to_cov3d_vjp`` (:528-570) and the VJP helpers
(``helpers.h:60-87,90-98,131-138,155-192``).

Matrices are held as nested 3x3 lists of ``(N,)`` arrays indexed ``[row][col]``
in mathematical order.  The reference stores glm-style column-major
``m[col][row]`` (``lib/common.h:25-137``); every product below sums over the
inner index in the order 0, 1, 2, which is the order of the reference's
``mat3::operator*`` (``common.h:59-71``).  With ``dtype=np.float32`` every
intermediate is rounded to float32 in the same operation order as the HIP
projection kernel (compiled with ``-ffp-contract=off``), so the two agree
bit for bit.

Outputs of culled Gaussians follow the reference's zero-initialised buffers
(``lib/ops.cu:260-322``): near-plane culled -> all zero; ``det == 0`` -> only
cov3d written; empty tile rectangle -> cov3d and conic written, the rest zero.
"""

from __future__ import annotations

import numpy as np


def _mm(a, b):
    return [[a[r][0] * b[0][c] + a[r][1] * b[1][c] + a[r][2] * b[2][c]
             for c in range(3)] for r in range(3)]


def _tr(a):
    return [[a[c][r] for c in range(3)] for r in range(3)]


def quat_to_rotmat(q, dt):
    """``helpers.h:140-153``; quaternion is (w, x, y, z), not normalised."""
    w, x, y, z = (q[:, i].astype(dt) for i in range(4))
    two, one = dt(2.0), dt(1.0)
    return [
        [one - two * (y * y + z * z), two * (x * y - w * z), two * (x * z + w * y)],
        [two * (x * y + w * z), one - two * (x * x + z * z), two * (y * z - w * x)],
        [two * (x * z - w * y), two * (y * z + w * x), one - two * (x * x + y * y)],
    ]


def scale_to_mat(s, glob_scale, dt):
    """``helpers.h:194-201``: diag(glob_scale * s)."""
    n = s.shape[0]
    z = np.zeros(n, dt)
    g = dt(glob_scale)
    return [[g * s[:, 0].astype(dt), z, z], [z, g * s[:, 1].astype(dt), z],
            [z, z, g * s[:, 2].astype(dt)]]


def scale_rot_to_cov3d(scales, glob_scale, quats, dt):
    """``forward.cu:547-568``: upper triangle of (R S)(R S)^T."""
    R = quat_to_rotmat(quats, dt)
    S = scale_to_mat(scales, glob_scale, dt)
    M = _mm(R, S)
    tmp = _mm(M, _tr(M))
    return np.stack([tmp[0][0], tmp[0][1], tmp[0][2], tmp[1][1], tmp[1][2],
                     tmp[2][2]], axis=1)


def transform_4x3(vm, p, dt):
    """``helpers.h:101-108``: R p + t with a row-major 4x4 viewmat."""
    v = vm.reshape(-1).astype(dt)
    x, y, z = (p[:, i].astype(dt) for i in range(3))
    return [v[4 * r + 0] * x + v[4 * r + 1] * y + v[4 * r + 2] * z + v[4 * r + 3]
            for r in range(3)]


def _view_rot(vm, dt):
    v = vm.reshape(-1).astype(dt)
    return [[v[4 * r + c] for c in range(3)] for r in range(3)]


def tan_fov(img_shape, f):
    """``forward.cu:70-71`` evaluates ``0.5 * W / fx`` in double, then rounds."""
    H, W = img_shape
    return float(np.float32(0.5 * W / np.float64(np.float32(f[0])))), \
        float(np.float32(0.5 * H / np.float64(np.float32(f[1]))))


def project_cov3d_ewa(means3d, cov3d, vm, fx, fy, tfx, tfy, dt):
    """``forward.cu:465-544``: EWA cov2d (+0.3 blur) and compensation."""
    W = _view_rot(vm, dt)
    t = transform_4x3(vm, means3d, dt)
    tx, ty, tz = t
    lim_x = dt(1.3) * dt(tfx)
    lim_y = dt(1.3) * dt(tfy)
    tx = tz * np.minimum(lim_x, np.maximum(-lim_x, tx / tz))
    ty = tz * np.minimum(lim_y, np.maximum(-lim_y, ty / tz))
    one = dt(1.0)
    rz = one / tz
    rz2 = rz * rz
    fx, fy = dt(fx), dt(fy)
    zero = np.zeros_like(tz)
    J = [[fx * rz, zero, -fx * tx * rz2],
         [zero, fy * rz, -fy * ty * rz2],
         [zero, zero, zero]]
    T = _mm(J, W)
    c = [cov3d[:, i] for i in range(6)]
    V = [[c[0], c[1], c[2]], [c[1], c[3], c[4]], [c[2], c[4], c[5]]]
    cov = _mm(_mm(T, V), _tr(T))
    c00, c11, c01 = cov[0][0], cov[1][1], cov[1][0]
    det_orig = c00 * c11 - c01 * c01
    blur = dt(0.3)
    a = c00 + blur
    b = c01
    cc = c11 + blur
    det_blur = a * cc - b * b
    with np.errstate(divide="ignore", invalid="ignore"):
        comp = np.sqrt(np.maximum(dt(0.0), det_orig / det_blur))
    return np.stack([a, b, cc], axis=1), comp


def compute_cov2d_bounds(cov2d, dt):
    """``helpers.h:34-57``: conic = inverse cov2d, 3-sigma radius (ceil)."""
    x, y, z = cov2d[:, 0], cov2d[:, 1], cov2d[:, 2]
    det = x * z - y * y
    ok = det != dt(0.0)
    with np.errstate(divide="ignore", invalid="ignore"):
        inv_det = dt(1.0) / det
        conic = np.stack([z * inv_det, -y * inv_det, x * inv_det], axis=1)
        b = dt(0.5) * (x + z)
        v1 = b + np.sqrt(np.maximum(dt(0.1), b * b - det))
        v2 = b - np.sqrt(np.maximum(dt(0.1), b * b - det))
        radius = np.ceil(dt(3.0) * np.sqrt(np.maximum(v1, v2)))
    return ok, conic, radius


def project_pix(fx, fy, cx, cy, pv, dt):
    """``helpers.h:122-128``: pinhole centre with ``1/(z + 1e-6)``."""
    rw = dt(1.0) / (pv[2] + dt(1e-6))
    return (pv[0] * rw * dt(fx) + dt(cx), pv[1] * rw * dt(fy) + dt(cy))


def tile_rect(cx, cy, radius, tiles_x, tiles_y, block_width):
    """``helpers.h:9-32``: tile-space rectangle [min, max) in float32.

    ``radius`` is the float radius (``ceil`` result); the emission kernel of
    the reference recomputes it from the int radius, which is the same value.
    Truncation ``(int)`` is toward zero, then clamped to [0, tiles].
    """
    f32 = np.float32
    bw = f32(block_width)
    tcx = np.asarray(cx, f32) / bw
    tcy = np.asarray(cy, f32) / bw
    tr = np.asarray(radius, f32) / bw
    with np.errstate(invalid="ignore", over="ignore"):
        xmin = np.clip(np.trunc(tcx - tr), 0, tiles_x).astype(np.int64)
        xmax = np.clip(np.trunc(tcx + tr + f32(1.0)), 0, tiles_x).astype(np.int64)
        ymin = np.clip(np.trunc(tcy - tr), 0, tiles_y).astype(np.int64)
        ymax = np.clip(np.trunc(tcy + tr + f32(1.0)), 0, tiles_y).astype(np.int64)
    return xmin, xmax, ymin, ymax


def project_fwd(means3d, scales, quats, viewmat, *, img_shape, f, c, glob_scale,
                clip_thresh, block_width, dtype=np.float64):
    """Restatement of ``project_gaussians_fwd`` (``forward.cu:13-131``).

    Returns a dict with the 8 reference outputs (``lib/ops.cu:251-258``):
    cov3d (N,6), xys (N,2), depths (N,), radii (N,) int32, conics (N,3),
    compensation (N,), num_tiles_hit (N,) int64, cum_tiles_hit (N,) int64.
    The tile rectangle is always evaluated in float32 (it is an integer
    decision made by the reference on float32 values).
    """
    dt = np.dtype(dtype).type
    H, W = img_shape
    N = means3d.shape[0]
    tiles_x = (W + block_width - 1) // block_width
    tiles_y = (H + block_width - 1) // block_width
    means3d = np.asarray(means3d).astype(dt)
    scales = np.asarray(scales).astype(dt)
    quats = np.asarray(quats).astype(dt)
    viewmat = np.asarray(viewmat).astype(dt)

    out = dict(
        cov3d=np.zeros((N, 6), dt), xys=np.zeros((N, 2), dt),
        depths=np.zeros(N, dt), radii=np.zeros(N, np.int32),
        conics=np.zeros((N, 3), dt), compensation=np.zeros(N, dt),
        num_tiles_hit=np.zeros(N, np.int64), cum_tiles_hit=np.zeros(N, np.int64),
    )
    pv = transform_4x3(viewmat, means3d, dt)
    alive = ~(pv[2] <= dt(clip_thresh))                       # forward.cu:44-47
    idx = np.nonzero(alive)[0]
    if idx.size:
        m = means3d[idx]
        cov3d = scale_rot_to_cov3d(scales[idx], glob_scale, quats[idx], dt)
        out["cov3d"][idx] = cov3d                             # forward.cu:62-63
        tfx, tfy = tan_fov(img_shape, f)
        cov2d, comp = project_cov3d_ewa(m, cov3d, viewmat, f[0], f[1], tfx, tfy, dt)
        ok, conic, radius = compute_cov2d_bounds(cov2d, dt)
        idx2 = idx[ok]
        out["conics"][idx2] = conic[ok]                       # forward.cu:93
        pvi = [p[idx2] for p in pv]
        cx_, cy_ = project_pix(f[0], f[1], c[0], c[1], pvi, dt)
        rad = radius[ok]
        xmin, xmax, ymin, ymax = tile_rect(cx_, cy_, rad, tiles_x, tiles_y, block_width)
        area = (xmax - xmin) * (ymax - ymin)
        hit = area > 0                                        # forward.cu:106-110
        g = idx2[hit]
        out["num_tiles_hit"][g] = area[hit]
        out["depths"][g] = pvi[2][hit]
        out["radii"][g] = rad[hit].astype(np.int32)
        out["xys"][g, 0] = cx_[hit]
        out["xys"][g, 1] = cy_[hit]
        out["compensation"][g] = comp[ok][hit]
    out["cum_tiles_hit"] = np.cumsum(out["num_tiles_hit"])    # ops.cu:607-643
    return out


# ----------------------------------------------------------------------------
# Backward
# ----------------------------------------------------------------------------

def quat_to_rotmat_vjp(q, vR):
    """``helpers.h:155-192``: derivative of the unnormalised polynomial R(q)."""
    w, x, y, z = (q[:, i] for i in range(4))
    # vR is [row][col]; the reference indexes v_R[col][row].
    R = lambda col, row: vR[row][col]  # noqa: E731
    vw = 2.0 * (x * (R(1, 2) - R(2, 1)) + y * (R(2, 0) - R(0, 2)) + z * (R(0, 1) - R(1, 0)))
    vx = 2.0 * (-2.0 * x * (R(1, 1) + R(2, 2)) + y * (R(0, 1) + R(1, 0))
                + z * (R(0, 2) + R(2, 0)) + w * (R(1, 2) - R(2, 1)))
    vy = 2.0 * (x * (R(0, 1) + R(1, 0)) - 2.0 * y * (R(0, 0) + R(2, 2))
                + z * (R(1, 2) + R(2, 1)) + w * (R(2, 0) - R(0, 2)))
    vz = 2.0 * (x * (R(0, 2) + R(2, 0)) + y * (R(1, 2) + R(2, 1))
                - 2.0 * z * (R(0, 0) + R(1, 1)) + w * (R(0, 1) - R(1, 0)))
    return np.stack([vw, vx, vy, vz], axis=1)


def cov2d_to_conic_vjp(conics, v_conic):
    """``helpers.h:60-71``: v_cov2d = -X G X packed as (xx, xy + yx, yy)."""
    a, b, c_ = conics[:, 0], conics[:, 1], conics[:, 2]
    X = [[a, b], [b, c_]]
    G = [[v_conic[:, 0], v_conic[:, 1] / 2.0], [v_conic[:, 1] / 2.0, v_conic[:, 2]]]
    XG = [[X[r][0] * G[0][cc] + X[r][1] * G[1][cc] for cc in range(2)] for r in range(2)]
    VS = [[-(XG[r][0] * X[0][cc] + XG[r][1] * X[1][cc]) for cc in range(2)] for r in range(2)]
    return np.stack([VS[0][0], VS[0][1] + VS[1][0], VS[1][1]], axis=1)


def cov2d_to_compensation_vjp(conics, compensation, v_compensation):
    """``helpers.h:73-87``: d comp^2 / d cov2d = (1 - comp^2) conic - 0.3 I det(conic)."""
    a, b, c_ = conics[:, 0], conics[:, 1], conics[:, 2]
    inv_det = a * c_ - b * b
    omc = 1.0 - compensation * compensation
    vsq = v_compensation * 0.5 / (compensation + 1e-6)
    return np.stack([vsq * (omc * a - 0.3 * inv_det), 2.0 * vsq * (omc * b),
                     vsq * (omc * c_ - 0.3 * inv_det)], axis=1)


def project_bwd(means3d, scales, quats, viewmat, radii, conics, compensation,
                v_xy, v_depth, v_conic, v_compensation, *, img_shape, f,
                glob_scale, mode="reference"):
    """Restatement of ``project_gaussians_bwd`` (``backward.cu:345-418``), float64.

    ``mode="reference"`` follows the reference exactly: the EWA VJP builds J
    from the *unclamped* camera-space mean (``backward.cu:445-466``) although
    the forward clamps it (``forward.cu:498-501``).
    ``mode="consistent"`` is the true adjoint of the forward: J from the
    clamped mean and the clamp differentiated (zero slope where clamped).
    cov3d is recomputed from scales/quats (the reference reads the saved one;
    the values are identical).

    Returns dict v_mean3d (N,3), v_scale (N,3), v_quat (N,4), v_cov2d (N,3),
    v_cov3d (N,6); rows with ``radii <= 0`` are zero (``backward.cu:369-371``).
    """
    if mode not in ("reference", "consistent"):
        raise ValueError(mode)
    dt = np.float64
    N = means3d.shape[0]
    out = dict(v_mean3d=np.zeros((N, 3)), v_scale=np.zeros((N, 3)),
               v_quat=np.zeros((N, 4)), v_cov2d=np.zeros((N, 3)),
               v_cov3d=np.zeros((N, 6)))
    idx = np.nonzero(np.asarray(radii).reshape(-1) > 0)[0]
    if idx.size == 0:
        return out
    m = np.asarray(means3d, dt)[idx]
    s = np.asarray(scales, dt)[idx]
    q = np.asarray(quats, dt)[idx]
    vm = np.asarray(viewmat, dt)
    cn = np.asarray(conics, dt).reshape(N, 3)[idx]
    comp = np.asarray(compensation, dt).reshape(N)[idx]
    vxy = np.asarray(v_xy, dt).reshape(N, 2)[idx]
    vdep = np.asarray(v_depth, dt).reshape(N)[idx]
    vcon = np.asarray(v_conic, dt).reshape(N, 3)[idx]
    vcomp = np.asarray(v_compensation, dt).reshape(N)[idx]
    fx, fy = float(f[0]), float(f[1])
    Wr = _view_rot(vm, dt)
    v = vm.reshape(-1)

    # v_mean from v_xy: transform_4x3_rot_only_transposed(project_pix_vjp)
    pv = transform_4x3(vm, m, dt)
    rw = 1.0 / (pv[2] + 1e-6)                                   # helpers.h:131-138
    vpx, vpy = fx * vxy[:, 0], fy * vxy[:, 1]
    vview = [vpx * rw, vpy * rw, -(vpx * pv[0] + vpy * pv[1]) * rw * rw]
    vmean = [v[0 + c] * vview[0] + v[4 + c] * vview[1] + v[8 + c] * vview[2]
             for c in range(3)]                                 # helpers.h:90-98
    vmean = [vmean[c] + v[8 + c] * vdep for c in range(3)]     # backward.cu:385-388

    # v_cov2d from v_conic (helpers.h:60-71) and v_compensation (:73-87)
    v2 = cov2d_to_conic_vjp(cn, vcon) + cov2d_to_compensation_vjp(cn, comp, vcomp)
    vc2 = [v2[:, 0], v2[:, 1], v2[:, 2]]

    # EWA VJP (backward.cu:421-524)
    cov3d = scale_rot_to_cov3d(s, glob_scale, q, dt)
    t = transform_4x3(vm, m, dt)
    tx, ty, tz = t
    if mode == "consistent":
        H, W = img_shape
        lim_x = 1.3 * (0.5 * W / fx)
        lim_y = 1.3 * (0.5 * H / fy)
        ux, uy = tx / tz, ty / tz
        inx = (ux > -lim_x) & (ux < lim_x)
        iny = (uy > -lim_y) & (uy < lim_y)
        cxv = np.clip(ux, -lim_x, lim_x)
        cyv = np.clip(uy, -lim_y, lim_y)
        txc, tyc = tz * cxv, tz * cyv
    else:
        txc, tyc = tx, ty
    rz = 1.0 / tz
    rz2 = rz * rz
    zero = np.zeros_like(tz)
    J = [[fx * rz, zero, -fx * txc * rz2], [zero, fy * rz, -fy * tyc * rz2],
         [zero, zero, zero]]
    cc3 = [cov3d[:, i] for i in range(6)]
    V = [[cc3[0], cc3[1], cc3[2]], [cc3[1], cc3[3], cc3[4]], [cc3[2], cc3[4], cc3[5]]]
    vcov = [[vc2[0], 0.5 * vc2[1], zero], [0.5 * vc2[1], vc2[2], zero], [zero, zero, zero]]
    T = _mm(J, Wr)
    vV = _mm(_mm(_tr(T), vcov), T)
    vT = [[x1 + x2 for x1, x2 in zip(r1, r2)] for r1, r2 in
          zip(_mm(_mm(vcov, T), _tr(V)), _mm(_mm(_tr(vcov), T), V))]
    vcov3d = np.stack([vV[0][0], vV[0][1] + vV[1][0], vV[0][2] + vV[2][0], vV[1][1],
                       vV[1][2] + vV[2][1], vV[2][2]], axis=1)
    vJ = _mm(vT, _tr(Wr))
    rz3 = rz2 * rz
    # reference indexes v_J[col][row]: v_J[2][0] -> vJ[0][2] etc.
    vtx = -fx * rz2 * vJ[0][2]
    vty = -fy * rz2 * vJ[1][2]
    vtz = (-fx * rz2 * vJ[0][0] + 2.0 * fx * txc * rz3 * vJ[0][2]
           - fy * rz2 * vJ[1][1] + 2.0 * fy * tyc * rz3 * vJ[1][2])
    if mode == "consistent":
        # tx' = tz * clamp(tx/tz): d/dtx = 1 inside, 0 clamped; d/dtz = 0 inside, +-lim clamped
        vtz = vtz + np.where(inx, 0.0, cxv) * vtx + np.where(iny, 0.0, cyv) * vty
        vtx = np.where(inx, vtx, 0.0)
        vty = np.where(iny, vty, 0.0)
    vt = [vtx, vty, vtz]
    vmean = [vmean[cc] + vt[0] * Wr[0][cc] + vt[1] * Wr[1][cc] + vt[2] * Wr[2][cc]
             for cc in range(3)]

    # scale/rot VJP (backward.cu:528-570)
    R = quat_to_rotmat(q, dt)
    S = scale_to_mat(s, glob_scale, dt)
    M = _mm(R, S)
    h = 0.5
    vVs = [[vcov3d[:, 0], h * vcov3d[:, 1], h * vcov3d[:, 2]],
           [h * vcov3d[:, 1], vcov3d[:, 3], h * vcov3d[:, 4]],
           [h * vcov3d[:, 2], h * vcov3d[:, 4], vcov3d[:, 5]]]
    vM = [[2.0 * e for e in row] for row in _mm(vVs, M)]
    vscale = [(R[0][i] * vM[0][i] + R[1][i] * vM[1][i] + R[2][i] * vM[2][i]) * glob_scale
              for i in range(3)]
    vR = _mm(vM, S)
    vq = quat_to_rotmat_vjp(q, vR)

    out["v_mean3d"][idx] = np.stack(vmean, axis=1)
    out["v_scale"][idx] = np.stack(vscale, axis=1)
    out["v_quat"][idx] = vq
    out["v_cov2d"][idx] = np.stack(vc2, axis=1)
    out["v_cov3d"][idx] = vcov3d
    return out
