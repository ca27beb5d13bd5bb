"""Intersection emission, sort order and tile bin edges (test infrastructure only).

Restates ``map_gaussian_to_intersects`` (reference
``lib/kernels/forward.cu:135-179``), the stable radix sort of
``ops::sort_and_bin`` (``lib/ops.cu:722-760``; CUB ``DeviceRadixSort::SortPairs``
on bits ``[0, 32 + msb)``, a stable LSD sort, third-party and not vendored) and
``get_tile_bin_edges`` (``forward.cu:184-208``).

Key of one intersection: ``(tile_id << 32) | (int64)(int32 bits of depth)``
(``forward.cu:168-173``).  Emission order is ascending Gaussian id (offsets
from the inclusive scan, ``forward.cu:166``), so a stable sort yields the
total order (tile_id, depth bits, gaussian id).  Tile bins are ``[start, end)``
and tiles with no intersection stay ``(0, 0)`` (``ops.cu:482-488``).
"""

from __future__ import annotations

import numpy as np

from .project import tile_rect


def emit(xys, depths, radii, cum_tiles_hit, *, img_shape, block_width):
    """Unsorted (keys, gaussian_ids) in the reference emission order.

    ``xys``/``depths`` must be the float32 projection outputs; the tile
    rectangle is recomputed from them and the integer radius exactly as
    ``forward.cu:152-161`` does.
    """
    H, W = img_shape
    tiles_x = (W + block_width - 1) // block_width
    tiles_y = (H + block_width - 1) // block_width
    xys = np.asarray(xys, np.float32).reshape(-1, 2)
    depths = np.asarray(depths, np.float32).reshape(-1)
    radii = np.asarray(radii, np.int32).reshape(-1)
    cum = np.asarray(cum_tiles_hit, np.int64).reshape(-1)
    n_isect = int(cum[-1]) if cum.size else 0
    keys = np.zeros(n_isect, np.int64)
    gids = np.zeros(n_isect, np.int32)
    vis = np.nonzero(radii > 0)[0]
    if vis.size == 0 or n_isect == 0:
        return keys, gids
    xmin, xmax, ymin, ymax = tile_rect(xys[vis, 0], xys[vis, 1],
                                       radii[vis].astype(np.float32),
                                       tiles_x, tiles_y, block_width)
    w = xmax - xmin
    area = w * (ymax - ymin)
    start = np.where(vis == 0, 0, cum[np.maximum(vis - 1, 0)])
    depth_bits = depths[vis].view(np.int32).astype(np.int64)  # sign-extending
    # expand each Gaussian's rectangle in row-major tile order
    rep = np.repeat(np.arange(vis.size), area)
    k = np.arange(rep.size) - np.repeat(np.cumsum(area) - area, area)
    ty = ymin[rep] + k // np.maximum(w[rep], 1)
    tx = xmin[rep] + k % np.maximum(w[rep], 1)
    tile = ty * tiles_x + tx
    pos = start[rep] + k
    keys[pos] = (tile << 32) | depth_bits[rep]
    gids[pos] = vis[rep].astype(np.int32)
    return keys, gids


def sort_pairs(keys, gids, num_tiles):
    """Stable LSD order on bits ``[0, 32 + msb)``, ``msb = bitlen(num_tiles) + 1``."""
    msb = int(num_tiles).bit_length() + 1
    nbits = 32 + msb
    mask = (np.int64(1) << nbits) - 1 if nbits < 64 else np.int64(-1)
    sk = keys & mask
    order = np.argsort(sk.view(np.uint64) if nbits >= 64 else sk, kind="stable")
    return keys[order], gids[order]


def tile_bins(keys_sorted, num_tiles):
    """``get_tile_bin_edges`` (``forward.cu:184-208``): (num_tiles, 2) int32."""
    bins = np.zeros((num_tiles, 2), np.int32)
    n = keys_sorted.size
    if n == 0:
        return bins
    tiles = (keys_sorted >> 32).astype(np.int64)
    change = np.nonzero(tiles[1:] != tiles[:-1])[0] + 1
    starts = np.concatenate([[0], change])
    ends = np.concatenate([change, [n]])
    bins[tiles[starts], 0] = starts
    bins[tiles[starts], 1] = ends
    return bins


def bin_and_sort(xys, depths, radii, cum_tiles_hit, *, img_shape, block_width):
    """Full integer stage: returns (keys_sorted, gids_sorted, tile_bins)."""
    H, W = img_shape
    num_tiles = ((W + block_width - 1) // block_width) * \
        ((H + block_width - 1) // block_width)
    keys, gids = emit(xys, depths, radii, cum_tiles_hit, img_shape=img_shape,
                      block_width=block_width)
    ks, gs = sort_pairs(keys, gids, num_tiles)
    return ks, gs, tile_bins(ks, num_tiles)


def bin_and_sort_bruteforce(xys, depths, radii, *, img_shape, block_width):
    """Pure-Python restatement for tiny inputs: nested loops + ``sorted``."""
    H, W = img_shape
    tiles_x = (W + block_width - 1) // block_width
    tiles_y = (H + block_width - 1) // block_width
    items = []
    xys = np.asarray(xys, np.float32).reshape(-1, 2)
    depths = np.asarray(depths, np.float32).reshape(-1)
    radii = np.asarray(radii, np.int32).reshape(-1)
    for g in range(xys.shape[0]):
        if radii[g] <= 0:
            continue
        x0, x1, y0, y1 = (int(v[0]) for v in tile_rect(
            xys[g:g + 1, 0], xys[g:g + 1, 1], np.float32([radii[g]]),
            tiles_x, tiles_y, block_width))
        dbits = int(depths[g:g + 1].view(np.int32)[0])
        for i in range(y0, y1):
            for j in range(x0, x1):
                items.append((i * tiles_x + j, dbits, g))
    items.sort()
    keys = np.array([(t << 32) | d for t, d, _ in items], np.int64)
    gids = np.array([g for _, _, g in items], np.int32)
    bins = np.zeros((tiles_x * tiles_y, 2), np.int32)
    for p, (t, _, _) in enumerate(items):
        if p == 0 or items[p - 1][0] != t:
            bins[t, 0] = p
        bins[t, 1] = p + 1
    return keys, gids, bins


def tile_lists(xys, depths, radii, tiles, *, img_shape, block_width):
    """Sorted Gaussian ids of each tile in ``tiles``: the run of that tile in
    the reference's sorted intersections (emission of ``forward.cu:135-179``
    restricted to one tile, in the order (depth bits, id) of the stable sort
    of ``ops.cu:722-760``), without emitting or sorting the other tiles."""
    H, W = img_shape
    tiles_x = (W + block_width - 1) // block_width
    tiles_y = (H + block_width - 1) // block_width
    xys = np.asarray(xys, np.float32).reshape(-1, 2)
    depths = np.asarray(depths, np.float32).reshape(-1)
    radii = np.asarray(radii, np.int32).reshape(-1)
    vis = np.nonzero(radii > 0)[0]
    xmin, xmax, ymin, ymax = tile_rect(xys[vis, 0], xys[vis, 1],
                                       radii[vis].astype(np.float32),
                                       tiles_x, tiles_y, block_width)
    out = []
    for t in tiles:
        ty, tx = divmod(int(t), tiles_x)
        g = vis[(xmin <= tx) & (tx < xmax) & (ymin <= ty) & (ty < ymax)]
        dbits = depths[g].view(np.int32)
        out.append(g[np.lexsort((g, dbits))].astype(np.int32))
    return out
