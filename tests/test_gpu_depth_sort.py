"""Binning on the GPU over depth-key distributions of different spreads.

The binning emits the pairs in Gaussian id order, sorts them stably by tile
and then orders every tile's run by (depth bits, position): runs of up to
1536 entries by buckets in LDS, up to 16384 by buckets with a global scratch,
longer ones by a chunked radix sort through global memory.  These cases craft
depth keys that vary in 0, 2, 9, 17, 25 and 31 bits (all keys equal, heavy
ties broken by id -- four buckets holding every entry --, a few octaves of
depth, depths over five decades) plus the case where nothing emits, with
short runs only, with longer runs, and with short, long and longest runs at
once (a cluster of Gaussians over a few tiles), bin them through
``jaxsplat_bin_and_sort``, the split depth + emission entry points and
``jaxsplat_bin_rects_sort`` (packed rectangles with their inclusive counts), and
require the sorted ids, the reconstructed 64-bit keys and the tile bins to be
bit-identical to the oracle's stable (tile, depth bits, id) order
(``oracle.binning``, reference ``lib/ops.cu:722-760``).
"""

import numpy as np
import pytest

import oracle
from tests import helpers as H

pytestmark = pytest.mark.gpu

IMG = (96, 128)
BW = 16


def _case(kind, n=20_000, seed=7, cluster=False):
    rng = np.random.default_rng(seed)
    H_, W_ = IMG
    xys = np.stack([rng.uniform(-10, W_ + 10, n), rng.uniform(-10, H_ + 10, n)], 1)
    if cluster:                         # half of them over a few tiles: long and short runs
        m = n // 2
        xys[:m] = rng.uniform(20, 44, (m, 2))
    xys = xys.astype(np.float32)
    radii = rng.integers(1, 20, n).astype(np.int32)
    base = np.array(3.0, np.float32).view(np.uint32)
    if kind == "equal":                 # no varying bit
        bits = np.full(n, base, np.uint32)
    elif kind == "ties":                # few distinct keys, many ties broken by id
        bits = base + rng.integers(0, 4, n).astype(np.uint32)
    elif kind == "bits9":               # varies in the low 9 bits
        bits = base + rng.integers(0, 1 << 9, n).astype(np.uint32)
    elif kind == "bits17":              # 17 bits
        bits = base + rng.integers(0, 1 << 17, n).astype(np.uint32)
    elif kind == "octaves":             # depths in [2, 10): 25 bits
        bits = rng.uniform(2.0, 10.0, n).astype(np.float32).view(np.uint32)
    elif kind == "wide":                # [0.02, 5000): 31 bits
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


PATHS = {"one": {}, "split": {"split": True}, "rects": {"rects": True}}


@pytest.mark.parametrize("path", list(PATHS))
@pytest.mark.parametrize("kind", ["equal", "ties", "bits9", "bits17", "octaves", "wide"])
def test_depth_key_spreads_bitexact(kind, path):
    xys, depths, radii, cum = _case(kind)
    gids, tiles, bins, isect = H.gpu_bin(xys, depths, radii, cum, img_shape=IMG, block_width=BW,
                                         **PATHS[path])
    ks, gs, bs = oracle.binning.bin_and_sort(xys, depths, radii, cum, img_shape=IMG,
                                             block_width=BW)
    assert isect == ks.size > 0
    assert np.array_equal(gids, gs), kind
    keys = (tiles.astype(np.int64) << 32) | depths.view(np.int32)[gids].astype(np.int64)
    assert np.array_equal(keys, ks), kind
    assert np.array_equal(bins, bs), kind


@pytest.mark.parametrize("path", ["one", "rects"])
@pytest.mark.parametrize("n,cluster", [(40_000, False), (90_000, False), (30_000, True),
                                       (70_000, True)])
@pytest.mark.parametrize("kind", ["equal", "ties", "bits9", "octaves", "wide"])
def test_long_tile_runs_bitexact(kind, n, cluster, path):
    # runs longer than 1536 entries (all three kernels with the largest
    # cluster: runs of up to 23.5k entries)
    xys, depths, radii, cum = _case(kind, n=n, seed=11, cluster=cluster)
    gids, tiles, bins, isect = H.gpu_bin(xys, depths, radii, cum, img_shape=IMG, block_width=BW,
                                         **PATHS[path])
    ks, gs, bs = oracle.binning.bin_and_sort(xys, depths, radii, cum, img_shape=IMG,
                                             block_width=BW)
    lengths = bs[:, 1] - bs[:, 0]
    assert lengths.max() > 1536
    if cluster and n <= 30_000:                     # short and long runs together
        assert 0 < lengths.min() <= 1536
    if n >= 70_000 and cluster:                     # runs beyond 16384 entries too
        assert lengths.max() > 16384
    assert isect == ks.size
    assert np.array_equal(gids, gs), kind
    keys = (tiles.astype(np.int64) << 32) | depths.view(np.int32)[gids].astype(np.int64)
    assert np.array_equal(keys, ks), kind
    assert np.array_equal(bins, bs), kind


def test_depth_stage_nothing_emits():
    # the split entry points run the depth stage whatever the count: nothing
    # is emitted or sorted
    xys, depths, radii, cum = _case("none")
    assert int(cum[-1]) == 0
    for path in ("split", "rects"):
        gids, tiles, bins, isect = H.gpu_bin(xys, depths, radii, cum, img_shape=IMG,
                                             block_width=BW, **PATHS[path])
        assert isect == 0
        assert not bins.any()
