"""CPU tests that pin the projection oracle (closed forms, scipy rotations,
float32 vs float64, torch float64 autograd for the adjoint)."""

import numpy as np
import pytest
import torch

import oracle
from jaxsplat_amd import scenes
from tests import torch_ref

KW = dict(img_shape=(64, 96), f=(80.0, 70.0), c=(48.0, 32.0), glob_scale=1.0,
          clip_thresh=0.01, block_width=16)


def test_axis_aligned_closed_form():
    """Identity rotation on the optical axis: cov2d = diag(f^2 s^2 / z^2) + 0.3."""
    z = np.array([2.0, 5.0, 9.0])
    s = np.array([[0.05, 0.02, 0.3], [0.1, 0.1, 0.1], [0.01, 0.2, 0.04]])
    means = np.stack([np.zeros(3), np.zeros(3), z], 1)
    quats = np.tile([1.0, 0, 0, 0], (3, 1))
    out = oracle.project.project_fwd(means, s, quats, np.eye(4), **KW)
    fx, fy = KW["f"]
    a = fx ** 2 * s[:, 0] ** 2 / z ** 2 + 0.3
    c = fy ** 2 * s[:, 1] ** 2 / z ** 2 + 0.3
    np.testing.assert_allclose(out["cov3d"], np.stack([s[:, 0] ** 2, 0 * z, 0 * z, s[:, 1] ** 2,
                                                       0 * z, s[:, 2] ** 2], 1), rtol=1e-12)
    np.testing.assert_allclose(out["conics"], np.stack([1 / a, 0 * z, 1 / c], 1), rtol=1e-12)
    np.testing.assert_allclose(out["xys"], np.tile(KW["c"], (3, 1)), atol=1e-9)
    np.testing.assert_allclose(out["depths"], z)
    expect_r = np.ceil(3 * np.sqrt(np.maximum(a, c)))
    np.testing.assert_array_equal(out["radii"], expect_r.astype(np.int32))
    comp = np.sqrt((a - 0.3) * (c - 0.3) / (a * c))
    np.testing.assert_allclose(out["compensation"], comp, rtol=1e-12)


def test_rotation_matches_scipy():
    from scipy.spatial.transform import Rotation
    rng = np.random.default_rng(0)
    q = rng.normal(size=(50, 4))
    q /= np.linalg.norm(q, axis=1, keepdims=True)
    R = np.array(oracle.project.quat_to_rotmat(q, np.float64)).transpose(2, 0, 1)
    Rs = Rotation.from_quat(q[:, [1, 2, 3, 0]]).as_matrix()
    np.testing.assert_allclose(R, Rs, atol=1e-12)
    s = np.exp(rng.normal(size=(50, 3)))
    cov = oracle.project.scale_rot_to_cov3d(s, 1.0, q, np.float64)
    full = Rs @ (s[:, :, None] ** 2 * np.eye(3)) @ Rs.transpose(0, 2, 1)
    np.testing.assert_allclose(cov, full[:, [0, 0, 0, 1, 1, 2], [0, 1, 2, 1, 2, 2]], rtol=1e-10,
                               atol=1e-14)


def test_culling_semantics():
    """z <= clip: everything zero; off-screen: conic written, radius/count zero
    (reference forward.cu:44-47,93,106-116)."""
    means = np.array([[0.0, 0.0, -1.0], [0.0, 0.0, 0.005], [50.0, 0.0, 2.0], [0.0, 0.0, 3.0]])
    s = np.full((4, 3), 0.05)
    q = np.tile([1.0, 0, 0, 0], (4, 1))
    out = oracle.project.project_fwd(means, s, q, np.eye(4), **KW)
    for g in (0, 1):
        assert not out["cov3d"][g].any() and not out["conics"][g].any()
        assert out["radii"][g] == 0 and out["num_tiles_hit"][g] == 0
    assert out["cov3d"][2].any() and out["conics"][2].any()
    assert out["radii"][2] == 0 and out["num_tiles_hit"][2] == 0 and not out["xys"][2].any()
    assert out["radii"][3] > 0 and out["num_tiles_hit"][3] > 0
    assert out["cum_tiles_hit"][-1] == out["num_tiles_hit"].sum()


def test_float32_restatement_close_to_float64():
    sc = scenes.make("small", n=3000)
    cam = sc.camera
    kw = dict(img_shape=cam.img_shape, f=cam.f, c=cam.c, glob_scale=1.0, clip_thresh=0.01,
              block_width=16)
    a = oracle.project.project_fwd(sc.means3d, sc.scales, sc.quats, cam.viewmat, **kw)
    b = oracle.project.project_fwd(sc.means3d, sc.scales, sc.quats, cam.viewmat,
                                   dtype=np.float32, **kw)
    vis = a["radii"] > 0
    for k in ("xys", "conics", "cov3d", "depths"):
        np.testing.assert_allclose(b[k][vis], a[k][vis], rtol=2e-5, atol=1e-6 * np.abs(a[k]).max())
    assert (a["radii"] != b["radii"]).mean() < 2e-3


def _grads(mode, jac_unclamped, means, scales, quats, vm, vxy, vdep, vcon, fx, fy, cx, cy, shape):
    t = lambda a: torch.tensor(a, dtype=torch.float64, requires_grad=True)  # noqa: E731
    m, s, q = t(means), t(scales), t(quats)
    xys, conic, z = torch_ref.project(m, s, q, torch.tensor(vm, dtype=torch.float64), fx, fy, cx,
                                      cy, shape, clamp=True, jac_unclamped=jac_unclamped)
    L = (xys * torch.tensor(vxy)).sum() + (z * torch.tensor(vdep)).sum() + \
        (conic * torch.tensor(vcon)).sum()
    L.backward()
    return m.grad.numpy(), s.grad.numpy(), q.grad.numpy()


@pytest.mark.parametrize("mode,jac_unclamped", [("consistent", False), ("reference", True)])
def test_project_bwd_matches_autograd(mode, jac_unclamped):
    rng = np.random.default_rng(1)
    n = 40
    means = np.stack([rng.uniform(-3, 3, n), rng.uniform(-2, 2, n), rng.uniform(2, 6, n)], 1)
    means[:8, 0] *= 4.0                                    # force the fov clamp
    scales = np.exp(rng.normal(np.log(0.1), 0.3, (n, 3)))
    quats = rng.normal(size=(n, 4))                         # unnormalised on purpose
    vm = scenes.look_at([0.3, -0.2, -0.5], [0.0, 0.0, 4.0]).astype(np.float64)
    fx, fy, cx, cy = 60.0, 55.0, 40.0, 30.0
    shape = (60, 80)
    fwd = oracle.project.project_fwd(means, scales, quats, vm, img_shape=shape, f=(fx, fy),
                                     c=(cx, cy), glob_scale=1.0, clip_thresh=0.01, block_width=8)
    keep = fwd["radii"] > 0
    assert keep.sum() > 20
    means, scales, quats = means[keep], scales[keep], quats[keep]
    n = means.shape[0]
    vxy, vdep, vcon = rng.normal(size=(n, 2)), rng.normal(size=n), rng.normal(size=(n, 3))
    fwd = oracle.project.project_fwd(means, scales, quats, vm, img_shape=shape, f=(fx, fy),
                                     c=(cx, cy), glob_scale=1.0, clip_thresh=0.01, block_width=8)
    got = oracle.project.project_bwd(means, scales, quats, vm, fwd["radii"], fwd["conics"],
                                     fwd["compensation"], vxy, vdep, vcon, np.zeros(n),
                                     img_shape=shape, f=(fx, fy), glob_scale=1.0, mode=mode)
    gm, gs, gq = _grads(mode, jac_unclamped, means, scales, quats, vm, vxy, vdep, vcon, fx, fy,
                        cx, cy, shape)
    np.testing.assert_allclose(got["v_mean3d"], gm, rtol=1e-8, atol=1e-10)
    np.testing.assert_allclose(got["v_scale"], gs, rtol=1e-8, atol=1e-10)
    np.testing.assert_allclose(got["v_quat"], gq, rtol=1e-8, atol=1e-10)


def test_modes_agree_without_clamp():
    sc = scenes.make("small", n=300)
    cam = sc.camera
    kw = dict(img_shape=cam.img_shape, f=cam.f, c=cam.c, glob_scale=1.0, clip_thresh=0.01,
              block_width=16)
    fwd = oracle.project.project_fwd(sc.means3d, sc.scales, sc.quats, cam.viewmat, **kw)
    n = sc.n
    rng = np.random.default_rng(2)
    args = (rng.normal(size=(n, 2)), rng.normal(size=n), rng.normal(size=(n, 3)), np.zeros(n))
    a, b = (oracle.project.project_bwd(sc.means3d, sc.scales, sc.quats, cam.viewmat, fwd["radii"],
                                       fwd["conics"], fwd["compensation"], *args,
                                       img_shape=cam.img_shape, f=cam.f, glob_scale=1.0, mode=m)
            for m in ("reference", "consistent"))
    for k in a:
        np.testing.assert_allclose(a[k], b[k], rtol=1e-12, atol=1e-14)


def _sym_params(rng, n):
    A = rng.normal(size=(n, 2, 2))
    cov = A @ A.transpose(0, 2, 1) + 0.5 * np.eye(2)     # blurred cov2d (PD)
    return cov[:, [0, 0, 1], [0, 1, 1]]


def test_conic_vjp_matches_autograd():
    """cov2d_to_conic_vjp (helpers.h:60-71) against autograd of the 2x2 inverse;
    parameters are the upper triangle (a, b, c), b shared by both off-diagonals."""
    rng = np.random.default_rng(4)
    c2 = torch.tensor(_sym_params(rng, 10), requires_grad=True)
    a, b, c = c2.unbind(-1)
    det = a * c - b * b
    conic = torch.stack([c / det, -b / det, a / det], -1)
    v = torch.tensor(rng.normal(size=(10, 3)))
    (conic * v).sum().backward()
    got = oracle.project.cov2d_to_conic_vjp(conic.detach().numpy(), v.numpy())
    np.testing.assert_allclose(got, c2.grad.numpy(), rtol=1e-10, atol=1e-12)


def test_compensation_vjp_matches_autograd():
    """cov2d_to_compensation_vjp (helpers.h:73-87) against autograd of
    comp = sqrt(det(cov2d - 0.3 I) / det(cov2d)); the +1e-6 regulariser of the
    reference sets the tolerance."""
    rng = np.random.default_rng(5)
    c2 = torch.tensor(_sym_params(rng, 10), requires_grad=True)
    a, b, c = c2.unbind(-1)
    det = a * c - b * b
    comp = torch.sqrt(((a - 0.3) * (c - 0.3) - b * b) / det)
    v = torch.tensor(rng.normal(size=10))
    (comp * v).sum().backward()
    conic = torch.stack([c / det, -b / det, a / det], -1).detach().numpy()
    got = oracle.project.cov2d_to_compensation_vjp(conic, comp.detach().numpy(), v.numpy())
    np.testing.assert_allclose(got, c2.grad.numpy(), rtol=1e-4)
