"""The four reference-compatible custom-call targets (jaxsplat_project_fwd,
jaxsplat_project_bwd, jaxsplat_rasterize_fwd, jaxsplat_rasterize_bwd), driven
exactly as XLA would drive them: a void** array of operands then results in the
reference lowering order (lib/ops.cu:246-258, 336-354, 408-419, 502-519) and
the 80-byte descriptor as the opaque string.  Results must equal the typed
entry points (same kernels) and the oracle."""

import ctypes

import numpy as np
import pytest

import oracle
from jaxsplat_amd import _lib, scenes
from tests import helpers as H

pytestmark = pytest.mark.gpu


def _call(name, bufs, d):
    L = _lib.lib()
    arr = (ctypes.c_void_p * len(bufs))(*[b.data_ptr() for b in bufs])
    fn = getattr(L, name)
    fn(H.stream(), arr, ctypes.string_at(ctypes.addressof(d), 80), 80)
    H.torch().cuda.synchronize()
    assert L.jaxsplat_last_error() == b"", L.jaxsplat_last_error()


def test_legacy_targets_match_typed_path_and_oracle():
    t = H.torch()
    sc = scenes.make("small", n=3000)
    cam = sc.camera
    N = sc.n
    Hh, Ww = cam.img_shape
    d = _lib.make_descriptor(N, cam.img_shape, cam.f, cam.c, 1.0, 0.01, 16)
    f32 = dict(dtype=t.float32, device="cuda")
    i32 = dict(dtype=t.int32, device="cuda")
    means, scales, quats = (H.dev(a) for a in (sc.means3d, sc.scales, sc.quats))
    vm = H.dev(cam.viewmat)
    outs = [t.zeros((N, 6), **f32), t.zeros((N, 2), **f32), t.zeros((N, 1), **f32),
            t.zeros((N, 1), **i32), t.zeros((N, 3), **f32), t.zeros((N, 1), **f32),
            t.zeros((N, 1), **i32), t.zeros((N, 1), **i32)]
    _call("jaxsplat_project_fwd", [means, scales, quats, vm] + outs, d)
    cov3d, xys, depths, radii, conics, comp, nth, cum = outs
    typed = H.gpu_project(sc.means3d, sc.scales, sc.quats, cam.viewmat, img_shape=cam.img_shape,
                          f=cam.f, c=cam.c)
    for k, v in zip(("cov3d", "xys", "depths", "radii", "conics", "compensation",
                     "num_tiles_hit", "cum_tiles_hit"), outs):
        assert np.array_equal(H.host(v).reshape(typed[k].shape), typed[k]), k

    colors, opac = H.dev(sc.colors), H.dev(sc.opacities)
    bg = H.dev(np.array([0.2, 0.3, 0.4], np.float32))
    fT, fI, img = t.zeros((Hh, Ww, 1), **f32), t.zeros((Hh, Ww, 1), **i32), t.zeros((Hh, Ww, 3), **f32)
    _call("jaxsplat_rasterize_fwd", [colors, opac, bg, xys, depths, radii, conics, cum, fT, fI, img],
          d)
    gids, _, bins, isect = H.gpu_bin(typed["xys"], typed["depths"], typed["radii"],
                                     typed["cum_tiles_hit"], img_shape=cam.img_shape)
    ref = oracle.rasterize.rasterize_fwd(gids, bins, typed["xys"], typed["conics"], sc.colors,
                                         sc.opacities, np.array([0.2, 0.3, 0.4]),
                                         img_shape=cam.img_shape, block_width=16)
    ok = ~ref["ambiguous"]
    assert np.abs(H.host(img)[ok] - ref["img"][ok]).max() <= 1e-4

    rng = np.random.default_rng(0)
    v_img = rng.normal(size=(Hh, Ww, 3)).astype(np.float32)
    v_img[ref["ambiguous"]] = 0
    vi, va = H.dev(v_img), t.zeros((Hh, Ww, 1), **f32)
    grads = [t.full((N, 3), 9.0, **f32), t.full((N, 1), 9.0, **f32), t.full((N, 2), 9.0, **f32),
             t.full((N, 2), 9.0, **f32), t.full((N, 3), 9.0, **f32)]
    _call("jaxsplat_rasterize_bwd",
          [colors, opac, bg, xys, depths, radii, conics, cum, fT, fI, vi, va] + grads, d)
    exp = oracle.rasterize.rasterize_bwd(gids, bins, typed["xys"], typed["conics"], sc.colors,
                                         sc.opacities, np.array([0.2, 0.3, 0.4]),
                                         ref["final_Ts"], ref["final_idx"], v_img,
                                         np.zeros((Hh, Ww)), img_shape=cam.img_shape,
                                         block_width=16, mode="reference")
    for k, g in zip(("v_colors", "v_opacity", "v_xy", "v_xy_abs", "v_conic"), grads):
        ok_, worst = H.grad_close(H.host(g).reshape(exp[k].shape), exp[k])
        assert ok_, (k, worst)

    v_xy = grads[2]
    v_depth = t.zeros((N, 1), **f32)
    v_conic = grads[4]
    v_comp = t.zeros((N, 1), **f32)
    pg = [t.full((N, 3), 9.0, **f32), t.full((N, 3), 9.0, **f32), t.full((N, 4), 9.0, **f32),
          t.full((N, 3), 9.0, **f32), t.full((N, 6), 9.0, **f32)]
    _call("jaxsplat_project_bwd", [means, scales, quats, vm, cov3d, xys, radii, conics, comp,
                                   v_comp, v_xy, v_depth, v_conic] + pg, d)
    pexp = oracle.project.project_bwd(sc.means3d, sc.scales, sc.quats, cam.viewmat,
                                      typed["radii"], typed["conics"], typed["compensation"],
                                      H.host(v_xy), np.zeros(N), H.host(v_conic), np.zeros(N),
                                      img_shape=cam.img_shape, f=cam.f, glob_scale=1.0)
    for k, g in zip(("v_mean3d", "v_scale", "v_quat", "v_cov2d", "v_cov3d"), pg):
        ok_, worst = H.grad_close(H.host(g), pexp[k])
        assert ok_, (k, worst)


POISON = 0x7FC00000


def _call_failing(name, bufs, d, msg):
    L = _lib.lib()
    arr = (ctypes.c_void_p * len(bufs))(*[b.data_ptr() for b in bufs])
    getattr(L, name)(H.stream(), arr, ctypes.string_at(ctypes.addressof(d), 80), 80)
    H.torch().cuda.synchronize()
    assert msg in L.jaxsplat_last_error(), L.jaxsplat_last_error()


def _poisoned(t):
    a = H.host(t).reshape(-1)
    return a.size > 0 and bool((a.view(np.uint32) == POISON).all())


@pytest.mark.parametrize("case", ["count_too_small", "count_negative"])
def test_legacy_rasterize_failure_poisons_results(case):
    """An inconsistent cum_tiles_hit (its last entry below the intersections
    that xys and radii give, or negative) is an error of the rasterize
    targets: every result buffer is filled with NaN bits (0x7FC00000 in the
    integer final_idx) instead of keeping its previous contents."""
    t = H.torch()
    sc = scenes.make("small", n=3000)
    cam = sc.camera
    N = sc.n
    Hh, Ww = cam.img_shape
    d = _lib.make_descriptor(N, cam.img_shape, cam.f, cam.c, 1.0, 0.01, 16)
    f32 = dict(dtype=t.float32, device="cuda")
    i32 = dict(dtype=t.int32, device="cuda")
    g = H.gpu_project(sc.means3d, sc.scales, sc.quats, cam.viewmat, img_shape=cam.img_shape,
                      f=cam.f, c=cam.c)
    cum = g["cum_tiles_hit"].reshape(N, 1).copy()
    assert cum[-1, 0] > 100
    cum[-1, 0] = cum[-1, 0] // 2 if case == "count_too_small" else -5
    msg = b"binning error" if case == "count_too_small" else b"negative"
    colors, opac = H.dev(sc.colors), H.dev(sc.opacities)
    bg = H.dev(np.zeros(3, np.float32))
    xys, depths = H.dev(g["xys"]), H.dev(g["depths"].reshape(N, 1))
    radii, conics = H.dev(g["radii"].reshape(N, 1)), H.dev(g["conics"])
    dcum = H.dev(cum)
    outs = [t.zeros((Hh, Ww, 1), **f32), t.zeros((Hh, Ww, 1), **i32),
            t.zeros((Hh, Ww, 3), **f32)]
    _call_failing("jaxsplat_rasterize_fwd",
                  [colors, opac, bg, xys, depths, radii, conics, dcum] + outs, d, msg)
    assert all(_poisoned(o) for o in outs)
    fT, fI = t.ones((Hh, Ww, 1), **f32), t.zeros((Hh, Ww, 1), **i32)
    vi, va = t.ones((Hh, Ww, 3), **f32), t.zeros((Hh, Ww, 1), **f32)
    grads = [t.zeros((N, 3), **f32), t.zeros((N, 1), **f32), t.zeros((N, 2), **f32),
             t.zeros((N, 2), **f32), t.zeros((N, 3), **f32)]
    _call_failing("jaxsplat_rasterize_bwd",
                  [colors, opac, bg, xys, depths, radii, conics, dcum, fT, fI, vi, va] + grads,
                  d, msg)
    assert all(_poisoned(o) for o in grads)


def test_legacy_project_failure_poisons_results():
    """A descriptor the typed entry points reject (block_width 0 written into
    the 80 bytes directly) leaves every projection result NaN-filled."""
    t = H.torch()
    N = 500
    d = _lib.make_descriptor(N, (64, 64), (64.0, 64.0), (32.0, 32.0), 1.0, 0.01, 16)
    d.block_width = 0
    f32 = dict(dtype=t.float32, device="cuda")
    i32 = dict(dtype=t.int32, device="cuda")
    ins = [t.zeros((N, 3), **f32), t.ones((N, 3), **f32), t.ones((N, 4), **f32),
           t.eye(4, **f32)]
    outs = [t.zeros((N, 6), **f32), t.zeros((N, 2), **f32), t.zeros((N, 1), **f32),
            t.zeros((N, 1), **i32), t.zeros((N, 3), **f32), t.zeros((N, 1), **f32),
            t.zeros((N, 1), **i32), t.zeros((N, 1), **i32)]
    _call_failing("jaxsplat_project_fwd", ins + outs, d, b"block_width")
    assert all(_poisoned(o) for o in outs)


def test_legacy_rejects_bad_descriptor():
    L = _lib.lib()
    arr = (ctypes.c_void_p * 12)()
    L.jaxsplat_project_fwd(None, arr, b"x" * 10, 10)
    assert b"descriptor" in L.jaxsplat_last_error()
