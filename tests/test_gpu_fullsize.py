"""Full-size parity at the benchmark workloads: ``hd1m`` (1,000,000 Gaussians
at 1080x1920, ~8.4M intersections) and ``dense3m`` (3,000,000 Gaussians with
deep tile lists, ~107M intersections), where the whole oracle is too slow.

Bars (written in each test):
* projection: bit-exact against the float32 restatement for every Gaussian;
* binning: exact, through properties that together fix the reference's sorted
  intersections -- (tile, depth bits, id) strictly increasing, every
  (id, tile) pair inside that id's tile rectangle, each id's count equal to its
  num_tiles_hit (so the pairs are exactly the emitted set), bin edges equal to
  the runs of the sorted tile ids; hd1m is also compared with the full oracle
  sort, and seeded tiles with the oracle's per-tile lists;
* rasterize forward on a seeded tile subset (plus the deepest tiles): image
  max-abs <= 1e-4 vs the float64 oracle outside float32-ambiguous pixels
  (at most 1% of the subset), final_idx equal there;
* adjoint chain (rasterize backward then projection backward) with the image
  cotangent restricted to the subset (a tile's result depends only on its own
  list): every gradient within |g - g64| <= 1e-3 |g64| + 1e-4 max|g64|;
* dense3m: central finite differences (eps = 1e-4, float64) of 64 seeded
  contributing Gaussians' directional derivatives against the consistent
  adjoint, within 1e-3 |fd| + 1e-4 max|fd|, on a 4x4-tile region.  The
  forward is differentiated with the discrete decisions of the unperturbed
  point held fixed (which pairs blend, where each pixel stops): with
  opacities near 1 almost every perturbation flips some pixel's stopping
  decision, and the adjoint is the derivative of the smooth piece.
  Directions whose difference quotient changes between eps and eps/2 cross
  the alpha clamp's kink and are excluded (at most 6).
"""

import numpy as np
import pytest

import oracle
from jaxsplat_amd import scenes
from tests import helpers as H
from tests import torch_ref

pytestmark = pytest.mark.gpu

BW = 16
_CACHE = {}


def _pipeline(name):
    """Projection, binning and rasterize forward of the whole workload on the
    GPU (once per workload)."""
    if name not in _CACHE:
        print(f"[fullsize] {name}: GPU pipeline", flush=True)
        sc = scenes.make(name)
        cam = sc.camera
        kw = dict(img_shape=cam.img_shape, f=cam.f, c=cam.c, glob_scale=sc.glob_scale,
                  clip_thresh=sc.clip_thresh, block_width=BW)
        g = H.gpu_project(sc.means3d, sc.scales, sc.quats, cam.viewmat, **kw)
        gids, tiles, bins, isect = H.gpu_bin(g["xys"], g["depths"], g["radii"],
                                             g["cum_tiles_hit"], img_shape=cam.img_shape,
                                             block_width=BW)
        fwd = H.gpu_raster_fwd(gids, bins, isect, g["xys"], g["conics"], sc.colors,
                               sc.opacities, sc.background, img_shape=cam.img_shape,
                               block_width=BW)
        _CACHE.clear()                       # keep one workload's arrays at a time
        _CACHE[name] = (sc, cam, kw, g, gids, tiles, bins, isect, fwd)
    return _CACHE[name]


def _tiles_xy(img_shape):
    H_, W_ = img_shape
    return (W_ + BW - 1) // BW, (H_ + BW - 1) // BW


def _subset(bins, n=40, deep=8, seed=11):
    lens = bins[:, 1] - bins[:, 0]
    nonempty = np.nonzero(lens > 0)[0]
    rng = np.random.default_rng(seed)
    pick = rng.choice(nonempty, size=min(n, nonempty.size), replace=False)
    return np.unique(np.concatenate([pick, np.argsort(lens)[-deep:]]))


def _tile_mask(tiles, img_shape):
    tx_n, _ = _tiles_xy(img_shape)
    m = np.zeros(img_shape, bool)
    for t in tiles:
        ty, tx = divmod(int(t), tx_n)
        m[ty * BW:(ty + 1) * BW, tx * BW:(tx + 1) * BW] = True
    return m


@pytest.mark.parametrize("name", ["hd1m", "dense3m"])
def test_fullsize_projection_bitexact(name):
    sc, cam, kw, g, *_ = _pipeline(name)
    o32 = oracle.project.project_fwd(sc.means3d, sc.scales, sc.quats, cam.viewmat,
                                     dtype=np.float32, **kw)
    for k in ("cov3d", "xys", "depths", "conics", "compensation"):
        a = g[k].reshape(o32[k].shape).astype(np.float32)
        assert np.array_equal(a.view(np.uint32), o32[k].astype(np.float32).view(np.uint32)), k
    for k in ("radii", "num_tiles_hit", "cum_tiles_hit"):
        assert np.array_equal(g[k].reshape(-1).astype(np.int64), o32[k].astype(np.int64)), k


@pytest.mark.parametrize("name", ["hd1m", "dense3m"])
def test_fullsize_binning_exact(name):
    sc, cam, kw, g, gids, tiles, bins, isect, _ = _pipeline(name)
    N = sc.n
    assert isect == int(g["cum_tiles_hit"][-1]) > 0
    t = tiles.astype(np.int64)
    gi = gids.astype(np.int64)
    dbits = g["depths"].reshape(-1).astype(np.float32).view(np.int32)[gi].astype(np.int64)
    # (tile, depth bits, id) strictly increasing
    d_t, d_d, d_g = np.diff(t), np.diff(dbits), np.diff(gi)
    assert ((d_t > 0) | ((d_t == 0) & ((d_d > 0) | ((d_d == 0) & (d_g > 0))))).all()
    del d_t, d_d, d_g, dbits
    # each id appears exactly num_tiles_hit times ...
    assert np.array_equal(np.bincount(gi, minlength=N), g["num_tiles_hit"].reshape(-1))
    # ... and only in tiles of its rectangle
    tx_n, ty_n = _tiles_xy(cam.img_shape)
    xmin, xmax, ymin, ymax = oracle.project.tile_rect(
        g["xys"][:, 0], g["xys"][:, 1], g["radii"].reshape(-1).astype(np.float32), tx_n, ty_n,
        BW)
    ty, tx = np.divmod(t, tx_n)
    assert ((xmin[gi] <= tx) & (tx < xmax[gi]) & (ymin[gi] <= ty) & (ty < ymax[gi])).all()
    del ty, tx
    # bin edges are the runs of the sorted tile ids; empty tiles (0, 0)
    T = tx_n * ty_n
    lo = np.searchsorted(t, np.arange(T), "left")
    hi = np.searchsorted(t, np.arange(T), "right")
    exp = np.stack([np.where(hi > lo, lo, 0), np.where(hi > lo, hi, 0)], axis=1)
    assert np.array_equal(bins.astype(np.int64), exp)
    # seeded tiles against the oracle's per-tile lists
    sub = _subset(bins)
    lists = oracle.binning.tile_lists(g["xys"], g["depths"], g["radii"], sub,
                                      img_shape=cam.img_shape, block_width=BW)
    for tt, lst in zip(sub, lists):
        assert np.array_equal(gids[bins[tt, 0]:bins[tt, 1]], lst), tt
    if name == "hd1m":
        ks, gs, ob = oracle.binning.bin_and_sort(g["xys"], g["depths"], g["radii"],
                                                 g["cum_tiles_hit"], img_shape=cam.img_shape,
                                                 block_width=BW)
        assert np.array_equal(gids, gs)
        assert np.array_equal(t, ks >> 32)
        assert np.array_equal(bins, ob)


def _fwd_subset(name, tiles):
    sc, cam, kw, g, gids, _, bins, isect, fwd = _pipeline(name)
    ref = oracle.rasterize.rasterize_fwd(gids, bins, g["xys"], g["conics"], sc.colors,
                                         sc.opacities, sc.background, img_shape=cam.img_shape,
                                         block_width=BW, tiles=tiles, chunk=512)
    mask = _tile_mask(tiles, cam.img_shape)
    return ref, mask


@pytest.mark.parametrize("name", ["hd1m", "dense3m"])
def test_fullsize_raster_fwd_subset(name):
    sc, cam, kw, g, gids, _, bins, isect, fwd = _pipeline(name)
    tiles = _subset(bins)
    ref, mask = _fwd_subset(name, tiles)
    amb = ref["ambiguous"] & mask
    # the subset holds the deepest tiles, where the float32 error bound of T
    # grows with the list length: at most 1% of the pixels are excluded
    assert amb.sum() <= 1e-2 * mask.sum(), amb.sum()
    ok = mask & ~ref["ambiguous"]
    assert np.abs(fwd["img"][ok] - ref["img"][ok]).max() <= 1e-4
    assert np.array_equal(fwd["final_idx"].reshape(mask.shape)[ok], ref["final_idx"][ok])


@pytest.mark.parametrize("name,mode", [("hd1m", "reference"), ("dense3m", "reference"),
                                       ("dense3m", "consistent")])
def test_fullsize_adjoint_chain_subset(name, mode):
    sc, cam, kw, g, gids, _, bins, isect, fwd = _pipeline(name)
    tiles = _subset(bins, n=24, deep=4, seed=13)
    ref, mask = _fwd_subset(name, tiles)
    rng = np.random.default_rng(17)
    w = rng.normal(size=(*cam.img_shape, 3)).astype(np.float32)
    w[~mask] = 0
    w[ref["ambiguous"]] = 0
    adj = 0 if mode == "reference" else 1
    N = sc.n
    rb = H.gpu_raster_bwd(gids, bins, isect, g["xys"], g["conics"], sc.colors, sc.opacities,
                          sc.background, fwd["final_Ts"], fwd["final_idx"], w, None,
                          img_shape=cam.img_shape, block_width=BW, adjoint=adj, with_abs=False)
    pb = H.gpu_project_bwd(sc.means3d, sc.scales, sc.quats, cam.viewmat, g["radii"],
                           g["conics"], rb["v_xy"], None, rb["v_conic"], adjoint=adj, **kw)
    ob = oracle.rasterize.rasterize_bwd(gids, bins, g["xys"], g["conics"], sc.colors,
                                        sc.opacities, sc.background, ref["final_Ts"],
                                        ref["final_idx"], w, np.zeros(cam.img_shape),
                                        img_shape=cam.img_shape, block_width=BW, mode=mode,
                                        tiles=tiles)
    for k in ("v_colors", "v_opacity", "v_xy", "v_conic"):
        ok, worst = H.grad_close(rb[k].reshape(ob[k].shape), ob[k])
        assert ok, (k, worst)
    rows = np.unique(np.concatenate([gids[bins[t, 0]:bins[t, 1]] for t in tiles]))
    zN = np.zeros(rows.size)
    op = oracle.project.project_bwd(sc.means3d[rows], sc.scales[rows], sc.quats[rows],
                                    cam.viewmat, g["radii"].reshape(-1)[rows],
                                    g["conics"][rows], g["compensation"].reshape(-1)[rows],
                                    ob["v_xy"][rows], zN, ob["v_conic"][rows], zN,
                                    img_shape=cam.img_shape, f=cam.f, glob_scale=sc.glob_scale,
                                    mode=mode)
    others = np.ones(N, bool)
    others[rows] = False
    for k in ("v_mean3d", "v_scale", "v_quat"):
        ok, worst = H.grad_close(pb[k][rows], op[k])
        assert ok, (k, worst)
        assert not pb[k][others].any(), k


def test_dense3m_finite_differences():
    sc, cam, kw, g, gids, _, bins, isect, fwd = _pipeline("dense3m")
    tx_n, ty_n = _tiles_xy(cam.img_shape)
    tx0, ty0 = tx_n // 2 - 2, ty_n // 2 - 2
    region = np.array([(ty0 + dy) * tx_n + tx0 + dx for dy in range(4) for dx in range(4)])
    ref, mask = _fwd_subset("dense3m", region)
    rng = np.random.default_rng(23)
    w = rng.normal(size=(*cam.img_shape, 3)).astype(np.float32)
    w[~mask] = 0
    w[ref["ambiguous"]] = 0
    # consistent adjoint on the GPU
    rb = H.gpu_raster_bwd(gids, bins, isect, g["xys"], g["conics"], sc.colors, sc.opacities,
                          sc.background, fwd["final_Ts"], fwd["final_idx"], w, None,
                          img_shape=cam.img_shape, block_width=BW, adjoint=1, with_abs=False)
    pb = H.gpu_project_bwd(sc.means3d, sc.scales, sc.quats, cam.viewmat, g["radii"],
                           g["conics"], rb["v_xy"], None, rb["v_conic"], adjoint=1, **kw)
    grad = np.concatenate([pb["v_mean3d"], pb["v_scale"], pb["v_quat"], rb["v_colors"],
                           rb["v_opacity"].reshape(-1, 1)], axis=1).astype(np.float64)

    # float64 forward of the region's pixels with the discrete decisions of
    # theta held fixed (which pairs blend, where each pixel stops): the smooth
    # function the adjoint differentiates
    lists = [gids[bins[t, 0]:bins[t, 1]] for t in region]
    members = np.unique(np.concatenate(lists))
    loc = np.full(sc.n, -1, np.int64)
    loc[members] = np.arange(members.size)
    p64 = oracle.project.project_fwd(sc.means3d[members].astype(np.float64),
                                     sc.scales[members].astype(np.float64),
                                     sc.quats[members].astype(np.float64), cam.viewmat, **kw)
    xy0, cn0 = p64["xys"], p64["conics"]
    col0 = sc.colors[members].astype(np.float64)
    op0 = sc.opacities[members].reshape(-1).astype(np.float64)
    pix, order, wpix = [], [], []
    for ti, t in enumerate(region):
        ty, tx = divmod(int(t), tx_n)
        for r in range(BW):
            for cc in range(BW):
                i, j = ty * BW + r, tx * BW + cc
                pix.append((j + 0.5, i + 0.5))
                order.append(loc[lists[ti]])
                wpix.append(w[i, j])
    pix, wpix = np.array(pix), np.array(wpix, np.float64)
    keep, _ = torch_ref.decisions(xy0, cn0, op0, pix, order)
    K = max(len(k) for k in keep)
    ids = np.full((len(keep), K), -1, np.int64)
    for p_, k in enumerate(keep):
        ids[p_, :len(k)] = k
    bg = np.asarray(sc.background, np.float64)

    def frozen_loss(rows, lj, xy_j, cn_j, col_j, op_j):
        gi = ids[rows]
        m = gi >= 0
        gg = np.where(m, gi, 0)
        xy, cn, col, op = xy0[gg], cn0[gg], col0[gg], op0[gg]
        sel = gi == lj
        xy[sel], cn[sel], col[sel], op[sel] = xy_j, cn_j, col_j, op_j
        d = xy - pix[rows, None, :]
        sig = 0.5 * (cn[..., 0] * d[..., 0] ** 2 + cn[..., 2] * d[..., 1] ** 2) + \
            cn[..., 1] * d[..., 0] * d[..., 1]
        al = np.where(m, np.minimum(op * np.exp(-sig), 0.999), 0.0)
        Tb = np.cumprod(np.concatenate([np.ones((al.shape[0], 1)), 1.0 - al[:, :-1]], axis=1),
                        axis=1)
        img = (col * (al * Tb)[..., None]).sum(axis=1) + np.prod(1.0 - al, axis=1)[:, None] * bg
        return float((img * wpix[rows]).sum())

    contrib = members[np.abs(grad[members]).max(axis=1) > 0]
    contrib = contrib[[(ids == loc[x]).any() for x in contrib]]
    assert contrib.size >= 64
    pick = np.random.default_rng(29).choice(contrib, size=64, replace=False)
    eps = 1e-4
    gd, fd, flagged = [], [], 0
    for j in pick:
        j = int(j)
        lj = int(loc[j])
        rows = np.nonzero((ids == lj).any(axis=1))[0]
        theta = [sc.means3d[j].astype(np.float64), sc.scales[j].astype(np.float64),
                 sc.quats[j].astype(np.float64), sc.colors[j].astype(np.float64),
                 float(sc.opacities[j, 0])]
        d = rng.normal(size=14)
        d[3:6] *= theta[1]                   # scales perturbed relative to their size

        def at(e):
            m, s_, q = theta[0] + e * d[0:3], theta[1] + e * d[3:6], theta[2] + e * d[6:10]
            pj = oracle.project.project_fwd(m[None], s_[None], q[None], cam.viewmat, **kw)
            return frozen_loss(rows, lj, pj["xys"][0], pj["conics"][0],
                               theta[3] + e * d[10:13], theta[4] + e * d[13])
        vals = [at(e) for e in (eps, -eps, eps / 2, -eps / 2)]
        f1 = (vals[0] - vals[1]) / (2 * eps)
        f2 = (vals[2] - vals[3]) / eps
        if abs(f1 - f2) > 1e-4 * max(abs(f1), abs(f2)) + 1e-9:
            flagged += 1                     # a kink (alpha clamp) inside [-eps, eps]
            continue
        gd.append(float(grad[j] @ d))
        fd.append(f1)
    gd, fd = np.array(gd), np.array(fd)
    print(f"[fullsize] FD: {gd.size} checked, {flagged} excluded", flush=True)
    assert flagged <= 6
    scale = np.abs(fd).max()
    err = np.abs(gd - fd)
    assert (err <= 1e-3 * np.abs(fd) + 1e-4 * scale).all(), (err / (np.abs(fd) + 1e-30)).max()
