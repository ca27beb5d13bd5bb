"""C ABI: the shared library loads and exports every symbol that
include/jaxsplat_amd.h declares; host-only entry points (descriptor,
workspace sizes, error reporting) behave as documented.  No GPU calls."""

import ctypes
import re
from pathlib import Path

import pytest

from jaxsplat_amd import _lib

HEADER = Path(__file__).resolve().parents[1] / "include" / "jaxsplat_amd.h"


def header_functions():
    text = re.sub(r"/\*.*?\*/", "", HEADER.read_text(), flags=re.S)
    return sorted(set(re.findall(r"\b(jaxsplat_[a-z0-9_]+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    lib = ctypes.CDLL(str(_lib.LIB_PATH))
    names = header_functions()
    assert len(names) >= 15
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing


def test_python_binding_covers_header():
    assert set(header_functions()) == set(_lib.EXPORTS)


def test_descriptor_layout_matches_reference():
    D = _lib.Descriptor
    assert ctypes.sizeof(D) == 80
    assert D.intrins.offset == 16 and D.glob_scale.offset == 32
    assert D.block_width.offset == 40 and D.grid_dim_2d.offset == 52
    assert D.block_dim_2d.offset == 64


def test_make_descriptor_fields():
    d = _lib.make_descriptor(1000, (1080, 1920), (1500.0, 1400.0), (960.0, 540.0), 1.5, 0.01, 16)
    assert (d.img_shape.x, d.img_shape.y, d.img_shape.z) == (1920, 1080, 1)
    assert list(d.intrins) == [1500.0, 1400.0, 960.0, 540.0]
    assert (d.grid_dim_2d.x, d.grid_dim_2d.y) == (120, 68)
    assert d.block_dim_1d == 256 and d.grid_dim_1d == 4
    assert (d.block_dim_2d.x, d.block_dim_2d.y, d.block_dim_2d.z) == (16, 16, 1)


@pytest.mark.parametrize("kw", [dict(block_width=0), dict(block_width=17),
                                dict(clip_thresh=-0.1), dict(img_shape=(0, 10))])
def test_make_descriptor_rejects(kw):
    args = dict(num_points=10, img_shape=(32, 32), f=(10.0, 10.0), c=(16.0, 16.0),
                glob_scale=1.0, clip_thresh=0.01, block_width=16)
    args.update(kw)
    with pytest.raises(ValueError) as e:
        _lib.make_descriptor(**args)
    assert str(e.value)
    assert _lib.lib().jaxsplat_last_error().decode() == str(e.value)


def test_workspace_sizes_scale():
    L = _lib.lib()
    small = _lib.make_descriptor(1000, (256, 256), (256, 256), (128, 128), 1.0, 0.01, 16)
    big = _lib.make_descriptor(1_000_000, (1080, 1920), (1500, 1500), (960, 540), 1.0, 0.01, 16)
    assert 0 < L.jaxsplat_project_workspace_size(ctypes.byref(small)) < \
        L.jaxsplat_project_workspace_size(ctypes.byref(big))
    a = L.jaxsplat_bin_workspace_size(ctypes.byref(big), 1_000_000)
    b = L.jaxsplat_bin_workspace_size(ctypes.byref(big), 8_000_000)
    assert 0 < a < b
    # the two stream-sized buffers per intersection dominate: <= 32 B / intersection + O(N)
    assert b - a < 32 * 7_000_000


def test_version_string():
    assert b"gfx950" in _lib.lib().jaxsplat_version()


def _batch_args(scenes=((0, 10), (10, 5)), views=((0, 32, 32), (1, 16, 24)), bw=16, cap=1000):
    sc = (_lib.Scene * len(scenes))()
    for i, (a, n) in enumerate(scenes):
        sc[i].first_gaussian, sc[i].num_gaussians = a, n
    vw = (_lib.View * len(views))()
    for i, (s, h, w) in enumerate(views):
        vw[i].scene, vw[i].img_h, vw[i].img_w = s, h, w
        vw[i].fx = vw[i].fy = 50.0
        vw[i].cx, vw[i].cy = w / 2, h / 2
    return sc, len(scenes), vw, len(views), 1.0, 0.01, bw, cap


@pytest.mark.parametrize("kw,msg", [
    (dict(scenes=((0, 10), (12, 5))), b"consecutive"),
    (dict(views=((1, 32, 32), (0, 16, 24))), b"grouped"),
    (dict(views=((0, 0, 32),)), b"image size"),
    (dict(views=((2, 8, 8),)), b"grouped"),
    (dict(bw=17), b"block_width"),
    (dict(cap=0), b"isect_capacity"),
    (dict(views=()), b"num_views"),
])
def test_batch_create_rejects(kw, msg):
    """Batch layout validation runs before any device call, so it is
    checked here without a GPU."""
    L = _lib.lib()
    h = ctypes.c_void_p()
    code = L.jaxsplat_batch_create(ctypes.byref(h), *_batch_args(**kw))
    assert code == _lib.ERR_INVALID
    assert msg in L.jaxsplat_last_error()
    assert not h.value


@pytest.mark.parametrize("field,value,msg", [("block_width", 0, b"block_width"),
                                             ("block_width", 17, b"block_width"),
                                             ("clip_thresh", -0.5, b"clip_thresh")])
@pytest.mark.parametrize("target", ["jaxsplat_project_fwd", "jaxsplat_project_bwd",
                                    "jaxsplat_rasterize_fwd", "jaxsplat_rasterize_bwd"])
def test_legacy_targets_reject_invalid_descriptor(target, field, value, msg):
    """An 80-byte descriptor with an invalid field (one make_descriptor would
    refuse) is rejected by every custom-call target before any device work;
    the message names the field.  (With NULL result buffers there is nothing
    to fill; the GPU test of the fill is in test_gpu_legacy.)"""
    L = _lib.lib()
    d = _lib.make_descriptor(64, (32, 32), (32.0, 32.0), (16.0, 16.0), 1.0, 0.01, 16)
    setattr(d, field, value)
    arr = (ctypes.c_void_p * 18)()
    getattr(L, target)(None, arr, ctypes.string_at(ctypes.addressof(d), 80), 80)
    assert msg in L.jaxsplat_last_error()


def _fake(n):
    """Host addresses that are never dereferenced: the calls below fail their
    argument checks before any GPU work."""
    return ctypes.c_void_p(0x1000 * n)


def test_opacity_limited_entry_points_reject_bad_arguments():
    L = _lib.lib()
    d = _lib.make_descriptor(1000, (256, 256), (256, 256), (128, 128), 1.0, 0.01, 16)
    D = ctypes.byref(d)
    P = _fake
    wsb = L.jaxsplat_tile_rects_workspace_size(D)
    assert wsb == L.jaxsplat_project_workspace_size(D) > 0
    # NULL required pointers
    assert L.jaxsplat_tile_rects(D, None, P(1), P(2), P(3), P(4), None, P(5), P(6), wsb,
                                 None) == _lib.ERR_INVALID
    assert L.jaxsplat_tile_rects(D, P(1), P(2), P(3), P(4), None, None, P(5), P(6), wsb,
                                 None) == _lib.ERR_INVALID
    # short workspace
    assert L.jaxsplat_tile_rects(D, P(1), P(2), P(3), P(4), P(5), None, P(6), P(7), wsb - 1,
                                 None) == _lib.ERR_WORKSPACE
    assert L.jaxsplat_project_forward_rects(D, P(1), P(2), P(3), P(4), None, None, P(6), P(7),
                                            P(8), P(9), None, P(10), P(11), P(12), P(13), wsb,
                                            None) == _lib.ERR_INVALID
    assert L.jaxsplat_project_forward_rects(D, P(1), P(2), P(3), P(4), P(5), None, P(6), P(7),
                                            P(8), P(9), None, P(10), P(11), P(12), P(13),
                                            wsb - 1, None) == _lib.ERR_WORKSPACE
    dwsb = L.jaxsplat_bin_depth_workspace_size(D)
    assert L.jaxsplat_bin_depth_sort_rects(D, P(1), P(2), None, P(3), dwsb,
                                           None) == _lib.ERR_INVALID
    assert L.jaxsplat_bin_depth_sort_rects(D, P(1), P(2), P(3), P(4), dwsb - 1,
                                           None) == _lib.ERR_WORKSPACE
    # staged backward with rectangles: channels, adjoint, NULL rectangles, workspace
    bwsb = L.jaxsplat_rasterize_backward_staged_workspace_size(D, 100)

    def staged(ch=3, rects=P(20), adj=0, ws=bwsb):
        return L.jaxsplat_rasterize_backward_staged_rects(
            D, P(1), P(2), 100, P(3), P(4), P(5), P(6), P(7), ch, P(8), rects, P(9), P(10),
            P(11), P(12), None, P(13), P(14), P(15), None, P(16), adj, P(17), ws, None)
    assert staged(ch=4) == _lib.ERR_INVALID
    assert staged(adj=7) == _lib.ERR_INVALID
    assert staged(rects=None) == _lib.ERR_INVALID
    assert staged(ws=bwsb - 1) == _lib.ERR_WORKSPACE
    assert b"workspace too small" in L.jaxsplat_last_error()


def test_batch_set_schedule_rejects():
    L = _lib.lib()
    assert L.jaxsplat_batch_set_schedule(None, 0) == _lib.ERR_INVALID
    assert set(_lib.SCHEDULE.values()) == {0, 1}
