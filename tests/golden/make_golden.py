"""Writes the golden fixtures of tests/golden/ from the oracle.

Run from the repository root: ``python tests/golden/make_golden.py``.  The
reference publishes no test vectors, so these fixtures freeze the oracle's
outputs on seeded inputs (jaxsplat_amd.scenes) to catch regressions of both
the oracle and the HIP path:

* ``tiny.npz``: 64 Gaussians, 40x48 image, 8x8 tiles, background (0.1, 0.2,
  0.3), opacities raised so that pixels terminate; float32 projection outputs,
  sorted ids, tile bins, float64 image / final_T / final_idx, and float64
  gradients of both adjoints for seeded cotangents.
* ``small2k.npz``: 2000 Gaussians of the small workload at 128x128 (f=128,
  c=64), 16x16 tiles: float32 projection outputs, sorted keys/ids, tile bins
  and the float32-rounded image.
"""

import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

import oracle  # noqa: E402
from jaxsplat_amd import scenes  # noqa: E402

OUT = os.path.dirname(os.path.abspath(__file__))


def tiny():
    sc = scenes.make("small", n=64, seed=77)
    shape, bw = (40, 48), 8
    kw = dict(img_shape=shape, f=(40.0, 40.0), c=(24.0, 20.0), glob_scale=1.0, clip_thresh=0.01,
              block_width=bw)
    scales = (sc.scales * 12).astype(np.float32)
    opac = np.minimum(sc.opacities * 1.8, 0.999).astype(np.float32)
    pr = oracle.project.project_fwd(sc.means3d, scales, sc.quats, np.eye(4, dtype=np.float32),
                                    dtype=np.float32, **kw)
    ks, gs, bins = oracle.binning.bin_and_sort(pr["xys"], pr["depths"], pr["radii"],
                                               pr["cum_tiles_hit"], img_shape=shape,
                                               block_width=bw)
    bg = np.array([0.1, 0.2, 0.3], np.float32)
    f = oracle.rasterize.rasterize_fwd(gs, bins, pr["xys"], pr["conics"], sc.colors, opac, bg,
                                       img_shape=shape, block_width=bw)
    rng = np.random.default_rng(99)
    v_img = rng.normal(size=(*shape, 3)).astype(np.float32)
    v_alpha = rng.normal(size=shape).astype(np.float32)
    grads = {}
    for mode in ("reference", "consistent"):
        b = oracle.rasterize.rasterize_bwd(gs, bins, pr["xys"], pr["conics"], sc.colors, opac, bg,
                                           f["final_Ts"], f["final_idx"], v_img, v_alpha,
                                           img_shape=shape, block_width=bw, mode=mode)
        for k, v in b.items():
            grads[f"{mode}_{k}"] = v
    np.savez_compressed(
        os.path.join(OUT, "tiny.npz"), means3d=sc.means3d, scales=scales, quats=sc.quats,
        colors=sc.colors, opacities=opac, background=bg, v_img=v_img, v_alpha=v_alpha,
        **{f"proj_{k}": v for k, v in pr.items()}, keys=ks, gids=gs, bins=bins,
        img=f["img"], final_Ts=f["final_Ts"], final_idx=f["final_idx"],
        ambiguous=f["ambiguous"], **grads)


def small2k():
    sc = scenes.make("small", n=2000)
    shape, bw = (128, 128), 16
    kw = dict(img_shape=shape, f=(128.0, 128.0), c=(64.0, 64.0), glob_scale=1.0,
              clip_thresh=0.01, block_width=bw)
    pr = oracle.project.project_fwd(sc.means3d, sc.scales, sc.quats, np.eye(4, dtype=np.float32),
                                    dtype=np.float32, **kw)
    ks, gs, bins = oracle.binning.bin_and_sort(pr["xys"], pr["depths"], pr["radii"],
                                               pr["cum_tiles_hit"], img_shape=shape,
                                               block_width=bw)
    f = oracle.rasterize.rasterize_fwd(gs, bins, pr["xys"], pr["conics"], sc.colors,
                                       sc.opacities, sc.background, img_shape=shape,
                                       block_width=bw)
    np.savez_compressed(
        os.path.join(OUT, "small2k.npz"), means3d=sc.means3d, scales=sc.scales, quats=sc.quats,
        colors=sc.colors, opacities=sc.opacities,
        **{f"proj_{k}": v for k, v in pr.items()}, keys=ks, gids=gs, bins=bins,
        img=f["img"].astype(np.float32), final_idx=f["final_idx"], ambiguous=f["ambiguous"])


if __name__ == "__main__":
    tiny()
    small2k()
    for n in ("tiny.npz", "small2k.npz"):
        print(n, os.path.getsize(os.path.join(OUT, n)))
