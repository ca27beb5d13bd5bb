"""Depth sort of the binning on the GPU: key ranges that need 1, 2, 3 and 4 passes.

The depth stage sorts only the key bits that vary among the Gaussians that
emit, in digits of at most 9 bits, with the number of passes decided on the
device from the keys themselves.  These cases craft depth keys whose varying
bits span each pass count (and the degenerate cases: all keys equal, nothing
emits, heavy ties), bin them through ``jaxsplat_bin_and_sort`` and the split
depth + emission entry points, and require the sorted ids, the reconstructed
64-bit keys and the tile bins to be bit-identical to the oracle's stable
(tile, depth bits, id) order (``oracle.binning``, reference
``lib/ops.cu:722-760``).
"""

import numpy as np
import pytest

import oracle
from tests import helpers as H

pytestmark = pytest.mark.gpu

IMG = (96, 128)
BW = 16


def _case(kind, n=20_000, seed=7):
    rng = np.random.default_rng(seed)
    H_, W_ = IMG
    xys = np.stack([rng.uniform(-10, W_ + 10, n), rng.uniform(-10, H_ + 10, n)], 1)
    xys = xys.astype(np.float32)
    radii = rng.integers(1, 20, n).astype(np.int32)
    base = np.array(3.0, np.float32).view(np.uint32)
    if kind == "equal":                 # no varying bit: the first pass alone
        bits = np.full(n, base, np.uint32)
    elif kind == "ties":                # few distinct keys, many ties broken by id
        bits = base + rng.integers(0, 4, n).astype(np.uint32)
    elif kind == "bits9":               # varies in the low 9 bits: one pass
        bits = base + rng.integers(0, 1 << 9, n).astype(np.uint32)
    elif kind == "bits17":              # 17 bits: two passes
        bits = base + rng.integers(0, 1 << 17, n).astype(np.uint32)
    elif kind == "octaves":             # depths in [2, 10): 25 bits, three passes
        bits = rng.uniform(2.0, 10.0, n).astype(np.float32).view(np.uint32)
    elif kind == "wide":                # [0.02, 5000): 31 bits, four passes
        bits = np.exp(rng.uniform(np.log(0.02), np.log(5000.0), n)).astype(
            np.float32).view(np.uint32)
    elif kind == "none":                # nothing emits
        bits = rng.uniform(2.0, 10.0, n).astype(np.float32).view(np.uint32)
        radii[:] = 0
    else:
        raise ValueError(kind)
    depths = bits.astype(np.uint32).view(np.float32)
    # a share of culled Gaussians (radius 0) among them
    radii[rng.random(n) < 0.2] = 0
    tiles_x = (W_ + BW - 1) // BW
    tiles_y = (H_ + BW - 1) // BW
    x0, x1, y0, y1 = oracle.project.tile_rect(xys[:, 0], xys[:, 1], radii.astype(np.float32),
                                              tiles_x, tiles_y, BW)
    cnt = np.where(radii > 0, (x1 - x0) * (y1 - y0), 0)
    cum = np.cumsum(cnt).astype(np.int32)
    return xys, depths, radii, cum


@pytest.mark.parametrize("split", [False, True])
@pytest.mark.parametrize("kind", ["equal", "ties", "bits9", "bits17", "octaves", "wide"])
def test_ranged_depth_sort_bitexact(kind, split):
    xys, depths, radii, cum = _case(kind)
    gids, tiles, bins, isect = H.gpu_bin(xys, depths, radii, cum, img_shape=IMG, block_width=BW,
                                         split=split)
    ks, gs, bs = oracle.binning.bin_and_sort(xys, depths, radii, cum, img_shape=IMG,
                                             block_width=BW)
    assert isect == ks.size > 0
    assert np.array_equal(gids, gs), kind
    keys = (tiles.astype(np.int64) << 32) | depths.view(np.int32)[gids].astype(np.int64)
    assert np.array_equal(keys, ks), kind
    assert np.array_equal(bins, bs), kind


def test_ranged_depth_sort_nothing_emits():
    # the split entry points run the depth stage whatever the count: an empty
    # set of kept keys
    xys, depths, radii, cum = _case("none")
    assert int(cum[-1]) == 0
    gids, tiles, bins, isect = H.gpu_bin(xys, depths, radii, cum, img_shape=IMG, block_width=BW,
                                         split=True)
    assert isect == 0
    assert not bins.any()
