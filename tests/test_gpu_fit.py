"""The image-fitting loop (the training step of the reference's
examples/single_image.py) on the GPU.

Bars (written in each test):
* fused activations vs torch float64: normalised quats and sigmoids within
  2e-6 relative;
* fused L2 gradient: v_img within 1e-6 relative of 2 (img - target) / n, the
  loss sum within 1e-5 relative;
* fused activation VJP + Adam: one step from a random state against a float64
  restatement (torch autograd through the activations, optax's update rule):
  each parameter's update within 1e-5 of its size plus float32 rounding of
  the parameter, each moment within the float32 rounding of its operands plus
  the error bound of the pulled-back gradient (the quaternion normalisation
  VJP may cancel), the step counter advanced;
* an overflowing render skips the update (parameters, moments and counter
  unchanged) and ImageFit grows the capacity and continues;
* graph replays and eager launches give the same loss trajectory (1e-3
  relative; float atomics differ in order);
* the trained Gaussians improve the PSNR of a seeded synthetic target by more
  than 8 dB over the untrained ones (the method without training).
"""

import numpy as np
import pytest

from jaxsplat_amd import _lib
from jaxsplat_amd.fit import ImageFit, init_params, psnr
from tests import helpers as H

pytestmark = pytest.mark.gpu


def _target(shape=(96, 128), seed=0):
    Hh, Ww = shape
    yy, xx = np.mgrid[0:Hh, 0:Ww]
    x, y = xx / Ww, yy / Hh
    r = 0.5 + 0.4 * np.sin(7 * x + 2 * y)
    g = np.where((x - 0.55) ** 2 + (y - 0.45) ** 2 < 0.06, 0.9, 0.2)
    b = 0.2 + 0.6 * y
    img = np.stack([r, g, b], -1)
    img += np.random.default_rng(seed).normal(0, 0.02, img.shape)
    return np.clip(img, 0, 1).astype(np.float32)


def test_fit_activate_matches_torch():
    t = H.torch()
    rng = np.random.default_rng(1)
    n = 5000
    q, c, o = (rng.normal(size=(n, k)).astype(np.float32) for k in (4, 3, 1))
    out = [t.empty((n, k), device="cuda") for k in (4, 3, 1)]
    dq, dc, do = H.dev(q), H.dev(c), H.dev(o)
    _lib.check(_lib.lib().jaxsplat_fit_activate(n, H.P(dq), H.P(dc), H.P(do),
                                                *[H.P(x) for x in out], H.stream()))
    q64 = q.astype(np.float64)
    exp = [q64 / np.linalg.norm(q64, axis=1, keepdims=True), 1 / (1 + np.exp(-c.astype(float))),
           1 / (1 + np.exp(-o.astype(float)))]
    for got, e in zip(out, exp):
        assert np.abs(H.host(got) - e).max() <= 2e-6 * np.abs(e).max()


# scalar path (n odd), float4 path, and a 1080p RGB image (grid-stride loop
# over 2048 workgroups); two calls in a row each write the loss afresh, with
# the same bits (partial sums added in workgroup order)
@pytest.mark.parametrize("n", [3 * 97 * 131, 4 * 97 * 131, 1080 * 1920 * 3])
def test_fit_l2_matches_torch(n):
    t = H.torch()
    rng = np.random.default_rng(2)
    img, tgt = rng.uniform(size=(2, n)).astype(np.float32)
    v = t.empty(n, device="cuda")
    loss = t.full((1,), 5.0, device="cuda")
    d_img, d_tgt = H.dev(img), H.dev(tgt)     # keep both alive until the kernel ran
    got = []
    for _ in range(2):
        _lib.check(_lib.lib().jaxsplat_fit_l2(n, H.P(d_img), H.P(d_tgt), H.P(v), H.P(loss),
                                              H.stream()))
        got.append(H.host(loss).copy())
    d = img.astype(np.float64) - tgt
    assert np.abs(H.host(v) - 2 * d / n).max() <= 1e-6 * np.abs(2 * d / n).max()
    assert abs(float(got[0][0]) - (d * d).sum()) <= 1e-5 * (d * d).sum()
    assert got[0].view(np.uint32)[0] == got[1].view(np.uint32)[0]


def _adam_ref(raw, g_act, m1, m2, t_, lr, b1, b2, eps):
    torch = H.torch()
    q = torch.tensor(raw[2], dtype=torch.float64, requires_grad=True)
    c = torch.tensor(raw[3], dtype=torch.float64, requires_grad=True)
    o = torch.tensor(raw[4], dtype=torch.float64, requires_grad=True)
    obj = ((q / q.norm(dim=1, keepdim=True)) * torch.tensor(g_act[2], dtype=torch.float64)).sum() \
        + (torch.sigmoid(c) * torch.tensor(g_act[3], dtype=torch.float64)).sum() \
        + (torch.sigmoid(o) * torch.tensor(g_act[4], dtype=torch.float64)).sum()
    obj.backward()
    G = np.concatenate([g_act[0], g_act[1], q.grad.numpy(), c.grad.numpy(), o.grad.numpy()], 1)
    # float32 error bound of the pulled-back gradient: the quaternion VJP
    # (g - q^ (q^.g)) / |q| may cancel, so bound it by its operands
    qn = np.linalg.norm(raw[2].astype(np.float64), axis=1, keepdims=True)
    gq = np.abs(g_act[2]).astype(np.float64)
    dG = 4e-7 * np.abs(G)
    dG[:, 6:10] = 1e-6 * (gq + gq.sum(1, keepdims=True)) / qn
    # sigmoid VJP g s (1 - s): float32 (1 - s) carries ~ulp(1) / (1 - s) relative error
    sg = 1.0 / (1.0 + np.exp(-np.concatenate([raw[3], raw[4]], 1).astype(np.float64)))
    dG[:, 10:14] = 1e-6 * np.abs(G[:, 10:14]) * (1.0 + sg / (1.0 - sg))
    P = np.concatenate([a.astype(np.float64) for a in raw], 1)
    m = b1 * m1 + (1 - b1) * G
    v = b2 * m2 + (1 - b2) * G * G
    mh = m / (1 - b1 ** (t_ + 1))
    vh = v / (1 - b2 ** (t_ + 1))
    tol_m = 3e-7 * (b1 * np.abs(m1) + (1 - b1) * np.abs(G)) + (1 - b1) * dG
    tol_v = 3e-7 * (b2 * np.abs(m2) + (1 - b2) * G * G) + (1 - b2) * 2 * np.abs(G) * dG
    return P - lr * mh / (np.sqrt(vh) + eps), m, v, tol_m, tol_v


def test_fit_adam_matches_restatement():
    t = H.torch()
    rng = np.random.default_rng(3)
    n = 4000
    raw = [rng.normal(size=(n, k)).astype(np.float32) for k in (3, 3, 4, 3, 1)]
    g_act = [rng.normal(size=(n, k)).astype(np.float32) for k in (3, 3, 4, 3, 1)]
    m1 = rng.normal(0, 0.1, size=(n, 14)).astype(np.float32)
    m2 = rng.uniform(0, 0.01, size=(n, 14)).astype(np.float32)
    step = 3
    d_raw = [H.dev(a) for a in raw]
    d_g = [H.dev(a) for a in g_act]
    d_m1, d_m2 = H.dev(m1), H.dev(m2)
    d_step = t.full((1,), step, dtype=t.int32, device="cuda")
    lr, b1, b2, eps = 0.01, 0.9, 0.999, 1e-8
    _lib.check(_lib.lib().jaxsplat_fit_adam(n, *[H.P(x) for x in d_raw], *[H.P(x) for x in d_g],
                                            H.P(d_m1), H.P(d_m2), H.P(d_step), None, lr, b1, b2,
                                            eps, H.stream()))
    # the kernel takes float32 hyperparameters: 1 - 0.999f is 1.3e-5 relative away
    # from 1e-3, so the restatement uses the same float32-rounded values
    h32 = [float(np.float32(x)) for x in (lr, b1, b2, eps)]
    P, m, v, tol_m, tol_v = _adam_ref(raw, g_act, m1.astype(np.float64), m2.astype(np.float64),
                                      step, *h32)
    got = np.concatenate([H.host(x) for x in d_raw], 1)
    p0 = np.concatenate(raw, 1).astype(np.float64)
    # the update (up to lr * m^ / eps where v^ is tiny) to 1e-5 of its size,
    # plus float32 rounding of the stored parameter
    tol = 1e-5 * np.abs(P - p0) + 2.5e-7 * np.abs(P) + 1e-9
    assert (np.abs(got - P) <= tol).all(), (np.abs(got - P) / tol).max()
    # per element, float32 rounding of the operands plus the gradient's bound
    assert (np.abs(H.host(d_m1) - m) <= tol_m).all()
    assert (np.abs(H.host(d_m2) - v) <= tol_v).all()
    assert int(H.host(d_step)[0]) == step + 1


def test_fit_overflow_skips_update_and_recovers():
    t = H.torch()
    tgt = H.dev(_target())
    fit = ImageFit(tgt, num_points=2000, seed=4, graph=False, check_every=1)
    fit._plan(100)                          # far too small for this scene
    before = [x.clone() for x in (fit.means, fit.scales, fit.quats_raw, fit.colors_raw,
                                  fit.opac_raw, fit.m1, fit.m2, fit.step_dev)]
    fit._launch()
    t.cuda.synchronize()
    for a, b in zip(before, (fit.means, fit.scales, fit.quats_raw, fit.colors_raw, fit.opac_raw,
                             fit.m1, fit.m2, fit.step_dev)):
        assert t.equal(a, b)
    with pytest.raises(_lib.JaxsplatError):
        fit.plan.check()
    assert not fit._checked()               # re-plans with room for the render
    fit._launch()
    assert fit._checked()
    assert fit.steps_applied == 1


def test_fit_graph_matches_eager():
    tgt = H.dev(_target())
    a = ImageFit(tgt, num_points=2000, seed=5, graph=True)
    b = ImageFit(tgt, num_points=2000, seed=5, graph=False)
    la, lb = [], []
    for _ in range(8):
        a.step()
        b.step()
        la.append(a.loss())
        lb.append(b.loss())
    assert a.steps_applied == b.steps_applied == 8
    np.testing.assert_allclose(la, lb, rtol=1e-3)


def test_fit_gain_over_untrained():
    t = H.torch()
    tgt_np = _target()
    tgt = H.dev(tgt_np)
    fit = ImageFit(tgt, num_points=3000, lr=0.01, seed=6)
    p0 = psnr(fit.render(), tgt)
    losses = []
    for _ in range(10):
        fit.step(30)
        losses.append(fit.loss())
    p1 = psnr(fit.render(), tgt)
    print(f"[fit] PSNR untrained {p0:.2f} dB -> trained {p1:.2f} dB after {fit.steps_applied} "
          f"steps; loss {losses[0]:.4f} -> {losses[-1]:.4f}; replans {fit.replans}", flush=True)
    assert fit.steps_applied == 300
    assert losses[-1] < losses[0]
    assert p1 - p0 > 8.0
