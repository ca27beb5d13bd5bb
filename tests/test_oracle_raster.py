"""CPU tests that pin the binning and rasterization oracle: brute force,
closed forms, torch float64 autograd of an independent blend, and central
finite differences of the float64 forward."""

import numpy as np
import pytest
import torch

import oracle
from jaxsplat_amd import scenes
from tests import torch_ref


def _tiny(n=60, shape=(40, 48), bw=8, scale=8.0, seed=1000, opac=None):
    sc = scenes.make("small", n=n, seed=seed)
    kw = dict(img_shape=shape, f=(40.0, 40.0), c=(shape[1] / 2, shape[0] / 2), glob_scale=1.0,
              clip_thresh=0.01, block_width=bw)
    pr = oracle.project.project_fwd(sc.means3d, sc.scales * scale, sc.quats, np.eye(4),
                                    dtype=np.float32, **kw)
    ks, gs, bins = oracle.binning.bin_and_sort(pr["xys"], pr["depths"], pr["radii"],
                                               pr["cum_tiles_hit"], img_shape=shape,
                                               block_width=bw)
    op = sc.opacities if opac is None else opac(sc.opacities)
    return sc, kw, pr, ks, gs, bins, op


@pytest.mark.parametrize("bw", [3, 8, 16])
def test_binning_matches_bruteforce(bw):
    sc, kw, pr, ks, gs, bins, _ = _tiny(n=120, shape=(50, 70), bw=bw)
    kb, gb, bb = oracle.binning.bin_and_sort_bruteforce(pr["xys"], pr["depths"], pr["radii"],
                                                        img_shape=(50, 70), block_width=bw)
    assert ks.size == pr["cum_tiles_hit"][-1] > 0
    np.testing.assert_array_equal(ks, kb)
    np.testing.assert_array_equal(gs, gb)
    np.testing.assert_array_equal(bins, bb)


def test_binning_invariants():
    sc = scenes.make("small")
    cam = sc.camera
    pr = oracle.project.project_fwd(sc.means3d, sc.scales, sc.quats, cam.viewmat,
                                    dtype=np.float32, img_shape=cam.img_shape, f=cam.f, c=cam.c,
                                    glob_scale=1.0, clip_thresh=0.01, block_width=16)
    ks, gs, bins = oracle.binning.bin_and_sort(pr["xys"], pr["depths"], pr["radii"],
                                               pr["cum_tiles_hit"], img_shape=cam.img_shape,
                                               block_width=16)
    assert np.all(np.diff(ks) >= 0)
    np.testing.assert_array_equal(np.bincount(gs, minlength=sc.n), pr["num_tiles_hit"])
    tiles = ks >> 32
    for t in np.unique(tiles)[:50]:
        s, e = bins[t]
        assert np.all(tiles[s:e] == t) and (s == 0 or tiles[s - 1] != t)
    empty = np.setdiff1d(np.arange(bins.shape[0]), np.unique(tiles))
    assert not bins[empty].any()


@pytest.mark.parametrize("bw", [8, 16])
def test_tile_lists_match_sorted_runs(bw):
    """Per-tile lists (used where the full sort is too large) equal the runs of
    the full emission + stable sort, for every tile, empty ones included."""
    sc = scenes.make("small")
    cam = sc.camera
    pr = oracle.project.project_fwd(sc.means3d, sc.scales * 3, sc.quats, cam.viewmat,
                                    dtype=np.float32, img_shape=cam.img_shape, f=cam.f, c=cam.c,
                                    glob_scale=1.0, clip_thresh=0.01, block_width=bw)
    _, gs, bins = oracle.binning.bin_and_sort(pr["xys"], pr["depths"], pr["radii"],
                                              pr["cum_tiles_hit"], img_shape=cam.img_shape,
                                              block_width=bw)
    tiles = np.arange(bins.shape[0])
    lists = oracle.binning.tile_lists(pr["xys"], pr["depths"], pr["radii"], tiles,
                                      img_shape=cam.img_shape, block_width=bw)
    assert sum(x.size for x in lists) == gs.size > 0
    for t, lst in zip(tiles, lists):
        np.testing.assert_array_equal(lst, gs[bins[t, 0]:bins[t, 1]])


@pytest.mark.parametrize("opac", [None, lambda o: np.full_like(o, 0.995)])
def test_raster_fwd_matches_bruteforce(opac):
    sc, kw, pr, ks, gs, bins, op = _tiny(opac=opac, scale=12.0)
    bg = np.array([0.1, 0.2, 0.3])
    fv = oracle.rasterize.rasterize_fwd(gs, bins, pr["xys"], pr["conics"], sc.colors, op, bg,
                                        img_shape=(40, 48), block_width=8, chunk=7)
    fb = oracle.rasterize.rasterize_fwd_bruteforce(gs, bins, pr["xys"], pr["conics"], sc.colors,
                                                   op, bg, img_shape=(40, 48), block_width=8)
    np.testing.assert_allclose(fv["img"], fb["img"], atol=1e-13)
    np.testing.assert_allclose(fv["final_Ts"], fb["final_Ts"], atol=1e-13)
    np.testing.assert_array_equal(fv["final_idx"], fb["final_idx"])
    if opac is not None:
        assert fb["final_Ts"].min() < 1e-3       # termination exercised


@pytest.mark.parametrize("mode", ["reference", "consistent"])
def test_raster_bwd_matches_bruteforce(mode):
    sc, kw, pr, ks, gs, bins, op = _tiny(scale=12.0, opac=lambda o: np.minimum(o * 1.9, 0.999))
    bg = np.array([0.1, 0.2, 0.3])
    fb = oracle.rasterize.rasterize_fwd_bruteforce(gs, bins, pr["xys"], pr["conics"], sc.colors,
                                                   op, bg, img_shape=(40, 48), block_width=8)
    rng = np.random.default_rng(0)
    vi, va = rng.normal(size=(40, 48, 3)), rng.normal(size=(40, 48))
    bv = oracle.rasterize.rasterize_bwd(gs, bins, pr["xys"], pr["conics"], sc.colors, op, bg,
                                        fb["final_Ts"], fb["final_idx"], vi, va, img_shape=(40, 48),
                                        block_width=8, mode=mode, pix_chunk=5)
    bb = oracle.rasterize.rasterize_bwd_bruteforce(gs, bins, pr["xys"], pr["conics"], sc.colors,
                                                   op, bg, fb["final_Ts"], fb["final_idx"], vi, va,
                                                   img_shape=(40, 48), block_width=8, mode=mode)
    for k in bv:
        np.testing.assert_allclose(bv[k], bb[k], rtol=1e-10, atol=1e-12)


@pytest.mark.parametrize("nch", [1, 5])
@pytest.mark.parametrize("mode", ["reference", "consistent"])
def test_nchannel_matches_bruteforce(nch, mode):
    """N-channel blend: vectorised forward and backward vs the sequential loops."""
    sc, kw, pr, ks, gs, bins, op = _tiny(scale=12.0, opac=lambda o: np.minimum(o * 1.9, 0.999))
    rng = np.random.default_rng(nch)
    col = rng.uniform(size=(sc.n, nch))
    bg = rng.uniform(size=nch)
    fv = oracle.rasterize.rasterize_fwd(gs, bins, pr["xys"], pr["conics"], col, op, bg,
                                        img_shape=(40, 48), block_width=8, chunk=7)
    fb = oracle.rasterize.rasterize_fwd_bruteforce(gs, bins, pr["xys"], pr["conics"], col, op, bg,
                                                   img_shape=(40, 48), block_width=8)
    assert fv["img"].shape == (40, 48, nch)
    np.testing.assert_allclose(fv["img"], fb["img"], atol=1e-13)
    vi, va = rng.normal(size=(40, 48, nch)), rng.normal(size=(40, 48))
    bv = oracle.rasterize.rasterize_bwd(gs, bins, pr["xys"], pr["conics"], col, op, bg,
                                        fb["final_Ts"], fb["final_idx"], vi, va, img_shape=(40, 48),
                                        block_width=8, mode=mode, pix_chunk=5)
    bb = oracle.rasterize.rasterize_bwd_bruteforce(gs, bins, pr["xys"], pr["conics"], col, op, bg,
                                                   fb["final_Ts"], fb["final_idx"], vi, va,
                                                   img_shape=(40, 48), block_width=8, mode=mode)
    for k in bv:
        np.testing.assert_allclose(bv[k], bb[k], rtol=1e-10, atol=1e-12)


def test_nchannel_separable_and_alpha_closed_form():
    """Colours do not take part in the blend's decisions, so a 5-channel render
    equals its 2 + 3 channel split, and the geometry gradients add up over the
    split (the alpha cotangent counted once).  One channel of colour 1 on a
    zero background renders the accumulated alpha 1 - T."""
    sc, kw, pr, ks, gs, bins, op = _tiny(scale=12.0, opac=lambda o: np.minimum(o * 1.9, 0.999))
    rng = np.random.default_rng(7)
    col = rng.uniform(size=(sc.n, 5))
    bg = rng.uniform(size=5)
    args = (gs, bins, pr["xys"], pr["conics"])
    kw2 = dict(img_shape=(40, 48), block_width=8)
    f5 = oracle.rasterize.rasterize_fwd(*args, col, op, bg, **kw2)
    fa = oracle.rasterize.rasterize_fwd(*args, col[:, :2], op, bg[:2], **kw2)
    fb = oracle.rasterize.rasterize_fwd(*args, col[:, 2:], op, bg[2:], **kw2)
    np.testing.assert_allclose(f5["img"], np.concatenate([fa["img"], fb["img"]], -1), atol=1e-14)
    vi, va = rng.normal(size=(40, 48, 5)), rng.normal(size=(40, 48))
    g5 = oracle.rasterize.rasterize_bwd(*args, col, op, bg, f5["final_Ts"], f5["final_idx"], vi,
                                        va, mode="consistent", **kw2)
    ga = oracle.rasterize.rasterize_bwd(*args, col[:, :2], op, bg[:2], f5["final_Ts"],
                                        f5["final_idx"], vi[..., :2], va, mode="consistent", **kw2)
    gb = oracle.rasterize.rasterize_bwd(*args, col[:, 2:], op, bg[2:], f5["final_Ts"],
                                        f5["final_idx"], vi[..., 2:], np.zeros_like(va),
                                        mode="consistent", **kw2)
    np.testing.assert_allclose(g5["v_colors"], np.concatenate([ga["v_colors"], gb["v_colors"]], 1),
                               rtol=1e-12, atol=1e-14)
    for k in ("v_opacity", "v_xy", "v_conic"):
        np.testing.assert_allclose(g5[k], ga[k] + gb[k], rtol=1e-9, atol=1e-12)
    f1 = oracle.rasterize.rasterize_fwd(*args, np.ones((sc.n, 1)), op, np.zeros(1), **kw2)
    np.testing.assert_allclose(f1["img"][..., 0], 1.0 - f1["final_Ts"], atol=1e-14)


def test_single_gaussian_closed_form():
    """One Gaussian: pixel = c*alpha + bg*(1-alpha) where alpha >= 1/255, else bg."""
    xys = np.array([[10.3, 7.7]])
    conics = np.array([[0.05, 0.01, 0.08]])
    colors = np.array([[0.9, 0.4, 0.2]])
    op = np.array([[0.8]])
    bg = np.array([0.05, 0.1, 0.15])
    gids = np.array([0])
    bins = np.array([[0, 1]])
    f = oracle.rasterize.rasterize_fwd(gids, bins, xys, conics, colors, op, bg, img_shape=(16, 16),
                                       block_width=16)
    jj, ii = np.meshgrid(np.arange(16) + 0.5, np.arange(16) + 0.5)
    dx, dy = 10.3 - jj, 7.7 - ii
    sig = 0.5 * (0.05 * dx * dx + 0.08 * dy * dy) + 0.01 * dx * dy
    al = np.minimum(0.999, 0.8 * np.exp(-sig))
    al = np.where(al >= 1 / 255, al, 0.0)
    exp = colors[0] * al[..., None] + bg * (1 - al)[..., None]
    np.testing.assert_allclose(f["img"], exp, atol=1e-14)
    np.testing.assert_allclose(f["final_Ts"], 1 - al, atol=1e-14)


def test_termination_closed_form():
    """Three coincident opaque Gaussians (alpha = 0.999): the second would drop
    T to 1e-6 <= 1e-4, so only the first is added (forward.cu:432-438)."""
    xys = np.tile([[8.5, 8.5]], (3, 1))
    conics = np.tile([[1e-4, 0.0, 1e-4]], (3, 1))
    colors = np.array([[1.0, 0, 0], [0, 1.0, 0], [0, 0, 1.0]])
    op = np.ones((3, 1))
    f = oracle.rasterize.rasterize_fwd(np.arange(3), np.array([[0, 3]]), xys, conics, colors, op,
                                       np.zeros(3), img_shape=(16, 16), block_width=16)
    al = np.minimum(0.999, np.exp(-0.5 * 1e-4 * ((np.arange(16) + 0.5 - 8.5) ** 2)[:, None]
                                  - 0.5 * 1e-4 * ((np.arange(16) + 0.5 - 8.5) ** 2)[None, :]))
    np.testing.assert_allclose(f["img"][..., 0], al, rtol=1e-12)
    assert not f["img"][..., 1:].any()
    np.testing.assert_array_equal(f["final_idx"], 0)
    np.testing.assert_allclose(f["final_Ts"], 1 - al, rtol=1e-12)


def _autograd_blend(sc, pr, gs, bins, op, bg, shape, bw, vi, va):
    """Gradients of sum(img*vi) + sum(alpha_img*va) through torch_ref.blend."""
    H, W = shape
    tiles_x = (W + bw - 1) // bw
    pix, order = [], []
    for i in range(H):
        for j in range(W):
            t = (i // bw) * tiles_x + j // bw
            pix.append([j + 0.5, i + 0.5])
            order.append(list(gs[bins[t, 0]:bins[t, 1]]))
    pix = np.array(pix)
    keep, flags = torch_ref.decisions(pr["xys"], pr["conics"], op, pix, order)
    xy = torch.tensor(np.asarray(pr["xys"], np.float64), requires_grad=True)
    cn = torch.tensor(np.asarray(pr["conics"], np.float64), requires_grad=True)
    col = torch.tensor(np.asarray(sc.colors, np.float64), requires_grad=True)
    o = torch.tensor(np.asarray(op, np.float64).reshape(-1), requires_grad=True)
    img, T = torch_ref.blend(xy, cn, col, o, torch.tensor(bg), torch.tensor(pix), keep, flags)
    L = (img * torch.tensor(vi.reshape(-1, 3))).sum() + ((1 - T) * torch.tensor(va.reshape(-1))).sum()
    L.backward()
    return dict(v_xy=xy.grad.numpy(), v_conic=cn.grad.numpy(), v_colors=col.grad.numpy(),
                v_opacity=o.grad.numpy())


def test_raster_bwd_consistent_matches_autograd():
    """The consistent adjoint is the exact derivative of the forward blend."""
    shape, bw = (24, 24), 8
    sc, kw, pr, ks, gs, bins, op = _tiny(n=40, shape=shape, bw=bw, scale=10.0,
                                         opac=lambda o: np.minimum(o * 1.9, 0.9995))
    bg = np.array([0.2, 0.1, 0.3])
    f = oracle.rasterize.rasterize_fwd(gs, bins, pr["xys"], pr["conics"], sc.colors, op, bg,
                                       img_shape=shape, block_width=bw)
    assert not f["ambiguous"].any()
    rng = np.random.default_rng(1)
    vi, va = rng.normal(size=(*shape, 3)), rng.normal(size=shape)
    got = oracle.rasterize.rasterize_bwd(gs, bins, pr["xys"], pr["conics"], sc.colors, op, bg,
                                         f["final_Ts"], f["final_idx"], vi, va, img_shape=shape,
                                         block_width=bw, mode="consistent")
    exp = _autograd_blend(sc, pr, gs, bins, op, bg, shape, bw, vi, va)
    for k in exp:
        np.testing.assert_allclose(got[k].reshape(exp[k].shape), exp[k], rtol=1e-7, atol=1e-10)


def test_raster_bwd_reference_equals_consistent_below_clamp():
    """With every o*exp(-sigma) < 0.99 the reference adjoint (0.99 clamp,
    backward.cu:261) equals the exact one."""
    shape, bw = (24, 24), 8
    sc, kw, pr, ks, gs, bins, op = _tiny(n=40, shape=shape, bw=bw, scale=10.0,
                                         opac=lambda o: np.minimum(o, 0.98))
    bg = np.array([0.2, 0.1, 0.3])
    f = oracle.rasterize.rasterize_fwd(gs, bins, pr["xys"], pr["conics"], sc.colors, op, bg,
                                       img_shape=shape, block_width=bw)
    rng = np.random.default_rng(2)
    vi, va = rng.normal(size=(*shape, 3)), rng.normal(size=shape)
    a, b = (oracle.rasterize.rasterize_bwd(gs, bins, pr["xys"], pr["conics"], sc.colors, op, bg,
                                           f["final_Ts"], f["final_idx"], vi, va, img_shape=shape,
                                           block_width=bw, mode=m)
            for m in ("reference", "consistent"))
    for k in a:
        np.testing.assert_allclose(a[k], b[k], rtol=1e-12, atol=1e-14)


def _render64(params, sc, kw, bg, target):
    means, scales, quats, colors, op = params
    pr = oracle.project.project_fwd(means, scales, quats, np.eye(4), **kw)
    ks, gs, bins = oracle.binning.bin_and_sort(pr["xys"].astype(np.float32),
                                               pr["depths"].astype(np.float32), pr["radii"],
                                               pr["cum_tiles_hit"], img_shape=kw["img_shape"],
                                               block_width=kw["block_width"])
    f = oracle.rasterize.rasterize_fwd(gs, bins, pr["xys"], pr["conics"], colors, op, bg,
                                       img_shape=kw["img_shape"], block_width=kw["block_width"])
    return float(((f["img"] - target) ** 2).sum()), pr, gs, bins, f


def test_full_pipeline_finite_differences():
    """Consistent adjoint of the whole path vs central differences (eps 1e-6,
    float64) on every parameter of a small scene; pairs whose perturbation
    changes a discrete decision are rare and tolerated (<= 2% of entries)."""
    shape, bw = (24, 32), 8
    sc = scenes.make("small", n=30, seed=5)
    kw = dict(img_shape=shape, f=(30.0, 30.0), c=(16.0, 12.0), glob_scale=1.0, clip_thresh=0.01,
              block_width=bw)
    means = sc.means3d.astype(np.float64)
    scales = (sc.scales * 8).astype(np.float64)
    quats = sc.quats.astype(np.float64)
    colors = sc.colors.astype(np.float64)
    op = np.minimum(sc.opacities.astype(np.float64) * 1.5, 0.95)
    bg = np.array([0.1, 0.2, 0.3])
    target = np.random.default_rng(3).uniform(size=(*shape, 3))
    params = [means, scales, quats, colors, op]
    L0, pr, gs, bins, f = _render64(params, sc, kw, bg, target)
    v_img = 2 * (f["img"] - target)
    rb = oracle.rasterize.rasterize_bwd(gs, bins, pr["xys"], pr["conics"], colors, op, bg,
                                        f["final_Ts"], f["final_idx"], v_img, np.zeros(shape),
                                        img_shape=shape, block_width=bw, mode="consistent")
    pb = oracle.project.project_bwd(means, scales, quats, np.eye(4), pr["radii"], pr["conics"],
                                    pr["compensation"], rb["v_xy"], np.zeros(sc.n), rb["v_conic"],
                                    np.zeros(sc.n), img_shape=shape, f=kw["f"], glob_scale=1.0,
                                    mode="consistent")
    grads = [pb["v_mean3d"], pb["v_scale"], pb["v_quat"], rb["v_colors"], rb["v_opacity"]]
    eps = 1e-6
    bad = total = 0
    for pi, (p, g) in enumerate(zip(params, grads)):
        g = g.reshape(p.shape)
        for idx in np.ndindex(p.shape):
            if pr["radii"][idx[0]] <= 0:
                continue
            old = p[idx]
            p[idx] = old + eps
            Lp = _render64(params, sc, kw, bg, target)[0]
            p[idx] = old - eps
            Lm = _render64(params, sc, kw, bg, target)[0]
            p[idx] = old
            fd = (Lp - Lm) / (2 * eps)
            total += 1
            if abs(fd - g[idx]) > 1e-4 * max(1.0, abs(fd)) + 1e-6:
                bad += 1
    assert total > 300
    assert bad <= 0.02 * total, (bad, total)
