"""Golden fixtures (tests/golden/*.npz, written by tests/golden/make_golden.py):
the oracle must keep reproducing them (CPU), and the HIP path must match
them through the C ABI (GPU)."""

import os

import numpy as np
import pytest

import oracle
from tests import helpers as H

G = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def load(name):
    with np.load(os.path.join(G, name), allow_pickle=False) as z:
        return {k: z[k] for k in z.files}


def _tiny_kw():
    return dict(img_shape=(40, 48), f=(40.0, 40.0), c=(24.0, 20.0), glob_scale=1.0,
                clip_thresh=0.01, block_width=8)


def _small_kw():
    return dict(img_shape=(128, 128), f=(128.0, 128.0), c=(64.0, 64.0), glob_scale=1.0,
                clip_thresh=0.01, block_width=16)


@pytest.mark.parametrize("name,kwf", [("tiny.npz", _tiny_kw), ("small2k.npz", _small_kw)])
def test_oracle_reproduces_fixture(name, kwf):
    z = load(name)
    kw = kwf()
    pr = oracle.project.project_fwd(z["means3d"], z["scales"], z["quats"],
                                    np.eye(4, dtype=np.float32), dtype=np.float32, **kw)
    for k, v in pr.items():
        np.testing.assert_array_equal(v, z[f"proj_{k}"], err_msg=k)
    ks, gs, bins = oracle.binning.bin_and_sort(pr["xys"], pr["depths"], pr["radii"],
                                               pr["cum_tiles_hit"], img_shape=kw["img_shape"],
                                               block_width=kw["block_width"])
    np.testing.assert_array_equal(ks, z["keys"])
    np.testing.assert_array_equal(gs, z["gids"])
    np.testing.assert_array_equal(bins, z["bins"])
    bg = z["background"] if "background" in z else np.zeros(3, np.float32)
    f = oracle.rasterize.rasterize_fwd(gs, bins, pr["xys"], pr["conics"], z["colors"],
                                       z["opacities"], bg, img_shape=kw["img_shape"],
                                       block_width=kw["block_width"])
    np.testing.assert_allclose(f["img"], z["img"], rtol=0, atol=1e-6 if name != "tiny.npz" else 1e-13)
    np.testing.assert_array_equal(f["final_idx"], z["final_idx"])
    if name == "tiny.npz":
        for mode in ("reference", "consistent"):
            b = oracle.rasterize.rasterize_bwd(gs, bins, pr["xys"], pr["conics"], z["colors"],
                                               z["opacities"], bg, f["final_Ts"], f["final_idx"],
                                               z["v_img"], z["v_alpha"], img_shape=(40, 48),
                                               block_width=8, mode=mode)
            for k, v in b.items():
                np.testing.assert_allclose(v, z[f"{mode}_{k}"], rtol=1e-12, atol=1e-14)


@pytest.mark.gpu
@pytest.mark.parametrize("name,kwf", [("tiny.npz", _tiny_kw), ("small2k.npz", _small_kw)])
def test_gpu_matches_fixture(name, kwf):
    z = load(name)
    kw = kwf()
    g = H.gpu_project(z["means3d"], z["scales"], z["quats"], np.eye(4, dtype=np.float32), **kw)
    for k in ("xys", "depths", "conics", "cov3d", "compensation"):
        assert np.array_equal(g[k].reshape(z[f"proj_{k}"].shape).view(np.uint32),
                              z[f"proj_{k}"].astype(np.float32).view(np.uint32)), k
    for k in ("radii", "num_tiles_hit", "cum_tiles_hit"):
        assert np.array_equal(g[k].reshape(-1), z[f"proj_{k}"].astype(np.int32)), k
    gids, tiles, bins, isect = H.gpu_bin(g["xys"], g["depths"], g["radii"], g["cum_tiles_hit"],
                                         img_shape=kw["img_shape"], block_width=kw["block_width"])
    assert np.array_equal(gids, z["gids"]) and np.array_equal(bins, z["bins"])
    keys = (tiles.astype(np.int64) << 32) | z["proj_depths"].view(np.int32)[gids].astype(np.int64)
    assert np.array_equal(keys, z["keys"])
    bg = z["background"] if "background" in z else np.zeros(3, np.float32)
    out = H.gpu_raster_fwd(gids, bins, isect, g["xys"], g["conics"], z["colors"], z["opacities"],
                           bg, img_shape=kw["img_shape"], block_width=kw["block_width"])
    ok = ~z["ambiguous"]
    assert np.abs(out["img"][ok] - z["img"][ok]).max() <= 1e-4
    assert np.array_equal(out["final_idx"][ok], z["final_idx"][ok])
    if name == "tiny.npz":
        vi, va = z["v_img"].copy(), z["v_alpha"].copy()
        assert not z["ambiguous"].any()
        for adj, mode in ((0, "reference"), (1, "consistent")):
            got = H.gpu_raster_bwd(gids, bins, isect, g["xys"], g["conics"], z["colors"],
                                   z["opacities"], bg, out["final_Ts"], out["final_idx"], vi, va,
                                   img_shape=(40, 48), block_width=8, adjoint=adj)
            for k in ("v_colors", "v_opacity", "v_xy", "v_xy_abs", "v_conic"):
                ok_, worst = H.grad_close(got[k].reshape(z[f"{mode}_{k}"].shape), z[f"{mode}_{k}"])
                assert ok_, (mode, k, worst)
