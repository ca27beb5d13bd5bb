"""Benchmark: fwd+bwd Mpix/s of the Gaussian-splatting render path on one MI355X.

Workload (default ``hd1m``): N=1,000,000 Gaussians, 1080x1920, fx=fy=1500,
c=(960,540), block 16, seeded synthetic scene (``jaxsplat_amd.scenes``), L2
loss against a seeded U(0,1) target.  One step = render forward (projection,
binning/sort, rasterize) + loss + backward (rasterize adjoint, projection
adjoint) through the public API ``jaxsplat_amd.render`` and torch autograd.
Inputs are resident in HBM before the timed region.

Prints ONE JSON line.  Besides the contract fields it carries:
* ``roofline``: the dominant kernel (rasterize backward) -- algorithmic bytes
  per launch / its average duration measured with HIP events on the launch
  stream, against the 8 TB/s HBM peak;
* ``cpu_baseline``: the float64 NumPy oracle timed on a bounded sample of the
  same workload on this host (1 core);
* ``stages``: per-stage milliseconds of the same pipeline;
* ``batched``: the multi-view workload ``orbit64`` (64 cameras around one
  500,000-Gaussian scene, 800x800 each) and the uneven multi-scene workload
  ``multi32`` (32 scenes of 10k-2M Gaussians, one camera each, mixed image
  sizes): one HIP-graph replay of the autotuned batched forward + L2 gradient
  + backward of all views (``jaxsplat_amd.batch``, ``jaxsplat_amd.tune``)
  against one plain ``render`` fwd+bwd call per view, both in Mpix/s over
  all views.
"""

from __future__ import annotations

import argparse
import ctypes
import json
import os
import sys
import time

import numpy as np

HBM_PEAK_GBS = 8000.0


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--workload", default="hd1m", choices=["small", "hd1m", "dense3m"])
    ap.add_argument("--cpu-baseline", dest="cpu_baseline", action="store_true", default=True)
    ap.add_argument("--no-cpu-baseline", dest="cpu_baseline", action="store_false")
    ap.add_argument("--cpu-tiles", type=int, default=200)
    ap.add_argument("--stage-iters", type=int, default=10)
    ap.add_argument("--batched", dest="batched", action="store_true", default=True)
    ap.add_argument("--no-batched", dest="batched", action="store_false")
    ap.add_argument("--batched-steps", type=int, default=5)
    ap.add_argument("--batched-workloads", default="orbit64,multi32")
    ap.add_argument("--copy-calibration-mb", type=int, default=0,
                    help="run a device-to-device copy of this many MiB first (HBM counter "
                         "calibration under rocprofv3 --pmc)")
    return ap.parse_args()


def main():
    args = parse()
    if args.gpus != 1:
        print(f"bench.py measures one GPU only; --gpus {args.gpus} is not supported "
              "(the path would shard by view or scene, each GPU holding all Gaussians, with one "
              "all-reduce of the per-Gaussian gradients per step; only the one-GPU share is built)",
              file=sys.stderr)
        sys.exit(2)
    import torch
    import jaxsplat_amd
    from jaxsplat_amd import _lib, scenes
    from jaxsplat_amd._ops import _bin, _p, _stream

    if not torch.cuda.is_available():
        print("bench.py needs a GPU", file=sys.stderr)
        sys.exit(1)
    _lib.lib()  # fail loudly if the HIP library is missing
    dev = torch.device("cuda:0")
    torch.cuda.set_device(dev)

    if args.copy_calibration_mb:
        n = args.copy_calibration_mb * (1 << 20) // 4
        src = torch.ones(n, dtype=torch.float32, device=dev)
        dst = torch.empty_like(src)
        for _ in range(3):
            dst.copy_(src)
        torch.cuda.synchronize()
        del src, dst

    sc = scenes.make(args.workload)
    cam = sc.camera
    Himg, Wimg = cam.img_shape
    tgt_seed = {"small": 2000, "hd1m": 2001, "dense3m": 2002}[args.workload]
    target = torch.from_numpy(scenes.target_image(cam.img_shape, tgt_seed)).to(dev)
    params = [torch.from_numpy(a).to(dev).requires_grad_(True) for a in sc.params()]
    viewmat = torch.from_numpy(cam.viewmat).to(dev)
    background = torch.from_numpy(sc.background).to(dev)
    kw = dict(viewmat=viewmat, background=background, img_shape=cam.img_shape, f=cam.f, c=cam.c,
              glob_scale=sc.glob_scale, clip_thresh=sc.clip_thresh, block_size=sc.block_width)

    def step():
        for p in params:
            p.grad = None
        img = jaxsplat_amd.render(*params, **kw)
        # mean((img - target)^2) as in the reference example's loss, as
        # torch's MSE (one reduction kernel forward, one elementwise backward)
        loss = torch.nn.functional.mse_loss(img, target)
        loss.backward()
        return loss

    for _ in range(args.warmup):
        step()
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    t0 = time.perf_counter()
    e0.record()
    for _ in range(args.steps):
        step()
    e1.record()
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    ms = e0.elapsed_time(e1) / args.steps
    mpix = Himg * Wimg / (ms * 1e-3) / 1e6

    stages, roof, extra = stage_breakdown(torch, sc, cam, params, target, args.stage_iters)

    batched = [batched_vs_plain(torch, w, args.batched_steps)
               for w in args.batched_workloads.split(",") if w] if args.batched else None

    cpu = cpu_baseline(sc, cam, tgt_seed, args.cpu_tiles) if args.cpu_baseline else None

    rec = {
        "metric": "fwd+bwd Mpix/s at 1M Gaussians 1080p, 1 GPU; % HBM roofline; batched vs plain",
        "value": round(mpix, 2),
        "unit": "Mpix/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded scene and U(0,1) target)",
        "config": {"workload": args.workload, "num_gaussians": sc.n, "img_shape": [Himg, Wimg],
                   "block_width": sc.block_width, "f": list(cam.f), "c": list(cam.c),
                   "path": "plain single view (jaxsplat_amd.render + torch autograd)",
                   "parallelism": "none (1 GPU)"},
        "roofline": roof,
        "cpu_baseline": cpu,
        "stages_ms": stages,
        "stats": extra,
        "batched": batched,
        "wall_s": round(wall, 3),
    }
    print(json.dumps(rec), flush=True)


def stage_breakdown(torch, sc, cam, params, target, iters):
    """Run the same pipeline through the C ABI with HIP events between stages."""
    from jaxsplat_amd import _lib
    from jaxsplat_amd._ops import TIGHT_TILES, _bin, _bin_rects, _p, _stream
    L = _lib.lib()
    dev = params[0].device
    N = sc.n
    Himg, Wimg = cam.img_shape
    means3d, scales, quats, colors, opac = [p.detach() for p in params]
    viewmat = torch.from_numpy(cam.viewmat).to(dev)
    bg = torch.from_numpy(sc.background).to(dev)
    d = _lib.make_descriptor(N, cam.img_shape, cam.f, cam.c, sc.glob_scale, sc.clip_thresh,
                             sc.block_width)
    f32 = dict(dtype=torch.float32, device=dev)
    i32 = dict(dtype=torch.int32, device=dev)
    xys, depths, conics = torch.empty((N, 2), **f32), torch.empty((N, 1), **f32), \
        torch.empty((N, 3), **f32)
    radii, nth, cum = torch.empty((N, 1), **i32), torch.empty((N, 1), **i32), \
        torch.empty((N, 1), **i32)
    # the render path's projection also writes the opacity-limited tile
    # rectangles and counts them (reference rectangles and counts below)
    rects, tnth, tcum = torch.empty((N, 2), **i32), torch.empty((N, 1), **i32), \
        torch.empty((N, 1), **i32)
    wsb = L.jaxsplat_project_workspace_size(ctypes.byref(d))
    ws = torch.empty(wsb, dtype=torch.uint8, device=dev)
    fT, fI = torch.empty((Himg, Wimg), **f32), torch.empty((Himg, Wimg), **i32)
    img = torch.empty((Himg, Wimg, 3), **f32)
    vc, vo, vxy, vcn = (torch.empty((N, 3), **f32), torch.empty((N, 1), **f32),
                        torch.empty((N, 2), **f32), torch.empty((N, 3), **f32))
    vm, vs, vq = torch.empty((N, 3), **f32), torch.empty((N, 3), **f32), torch.empty((N, 4), **f32)
    names = ["project_fwd", "bin_sort", "raster_fwd", "loss_grad", "raster_bwd", "project_bwd"]
    acc = {k: 0.0 for k in names}
    s = _stream(dev)
    bws = None
    for it in range(iters + 2):
        ev = [torch.cuda.Event(enable_timing=True) for _ in range(len(names) + 1)]
        ev[0].record()
        if TIGHT_TILES:
            _lib.check(L.jaxsplat_project_forward_rects(
                ctypes.byref(d), _p(means3d), _p(scales), _p(quats), _p(viewmat), _p(opac),
                None, _p(xys), _p(depths), _p(radii), _p(conics), None, _p(tnth), _p(tcum),
                _p(rects), _p(ws), wsb, s))
        else:
            _lib.check(L.jaxsplat_project_forward(
                ctypes.byref(d), _p(means3d), _p(scales), _p(quats), _p(viewmat), None, _p(xys),
                _p(depths), _p(radii), _p(conics), None, _p(nth), _p(cum), _p(ws), wsb, s))
        ev[1].record()
        # the render path's binning: into the opacity-limited rectangles, or
        # into the reference's
        if TIGHT_TILES:
            gids, bins, isect = _bin_rects(d, depths, radii, rects, tcum, dev)
            brects, bcum = rects, tcum
        else:
            isect = int(cum[-1, 0].item())
            gids, bins, _ = _bin(d, xys, depths, radii, isect, dev)
            brects, bcum = None, cum
        ev[2].record()
        _lib.check(L.jaxsplat_rasterize_forward_binned(
            ctypes.byref(d), _p(gids), _p(bins), isect, _p(xys), _p(conics), _p(colors), _p(opac),
            _p(bg), _p(fT), _p(fI), _p(img), s))
        ev[3].record()
        v_img = (2.0 / img.numel()) * (img - target)
        bwsb = L.jaxsplat_rasterize_backward_staged_workspace_size(ctypes.byref(d), isect)
        if bws is None or bws.numel() < bwsb:
            bws = torch.empty(bwsb, dtype=torch.uint8, device=dev)
        ev[4].record()
        # the render path's backward: staged records + per-Gaussian reduction
        if brects is not None:
            _lib.check(L.jaxsplat_rasterize_backward_staged_rects(
                ctypes.byref(d), _p(gids), _p(bins), isect, _p(xys), _p(conics), _p(colors),
                _p(opac), _p(bg), 3, _p(radii), _p(brects), _p(bcum), _p(fT), _p(fI), _p(v_img),
                None, _p(vc), _p(vo), _p(vxy), None, _p(vcn), 0, _p(bws), bwsb, s))
        else:
            _lib.check(L.jaxsplat_rasterize_backward_staged(
                ctypes.byref(d), _p(gids), _p(bins), isect, _p(xys), _p(conics), _p(colors),
                _p(opac), _p(bg), 3, _p(radii), _p(bcum), _p(fT), _p(fI), _p(v_img), None,
                _p(vc), _p(vo), _p(vxy), None, _p(vcn), 0, _p(bws), bwsb, s))
        ev[5].record()
        _lib.check(L.jaxsplat_project_backward(
            ctypes.byref(d), _p(means3d), _p(scales), _p(quats), _p(viewmat), _p(radii),
            _p(conics), None, _p(vxy), None, _p(vcn), None, _p(vm), _p(vs), _p(vq), None, None,
            0, s))
        ev[6].record()
        torch.cuda.synchronize()
        if it >= 2:
            for k, n in enumerate(names):
                acc[n] += ev[k].elapsed_time(ev[k + 1]) / iters
    # algorithmic bytes of the dominant kernel (rasterize backward), per launch:
    #   40 B per consumed intersection (id 4 + xy 8 + opacity 4 + conic 12 + color 12)
    # + 20 B per pixel (final T 4 + final idx 4 + v_img 12)
    # + 36 B per visible Gaussian (v_rgb 12 + v_opacity 4 + v_xy 8 + v_conic 12 written)
    # + 8 B per tile (bin read)
    # The intersections are counted on the reference's tile lists (exact
    # binning of the same projection, outside the timed region), also when the
    # timed path bins into the shorter opacity-limited lists.
    bw = sc.block_width
    tiles_x, tiles_y = (Wimg + bw - 1) // bw, (Himg + bw - 1) // bw

    def consumed(bins_t, fI_t):
        bins_np = bins_t.cpu().numpy().astype(np.int64)
        pad = np.full((tiles_y * bw, tiles_x * bw), -1, np.int64)
        pad[:Himg, :Wimg] = fI_t.cpu().numpy().astype(np.int64)
        tmax = pad.reshape(tiles_y, bw, tiles_x, bw).max(axis=(1, 3)).reshape(-1)
        return int(np.clip(tmax - bins_np[:, 0] + 1, 0, bins_np[:, 1] - bins_np[:, 0]).sum())

    isect_binned, L_b_binned = isect, consumed(bins, fI)
    if brects is not None:
        # the reference's counts of the same projection
        _lib.check(L.jaxsplat_project_forward(
            ctypes.byref(d), _p(means3d), _p(scales), _p(quats), _p(viewmat), None, _p(xys),
            _p(depths), _p(radii), _p(conics), None, _p(nth), _p(cum), _p(ws), wsb, s))
    isect = int(cum[-1, 0].item())
    if brects is not None:
        gids, bins, _ = _bin(d, xys, depths, radii, isect, dev)
        fT2, fI2, img2 = torch.empty_like(fT), torch.empty_like(fI), torch.empty_like(img)
        _lib.check(L.jaxsplat_rasterize_forward_binned(
            ctypes.byref(d), _p(gids), _p(bins), isect, _p(xys), _p(conics), _p(colors), _p(opac),
            _p(bg), _p(fT2), _p(fI2), _p(img2), s))
        L_b = consumed(bins, fI2)
    else:
        L_b = L_b_binned
    V = int((radii > 0).sum().item())
    T = tiles_x * tiles_y
    bytes_bwd = 40 * L_b + 20 * Himg * Wimg + 36 * V + 8 * T
    t_bwd = acc["raster_bwd"] * 1e-3
    achieved = bytes_bwd / t_bwd / 1e9
    roof = {"kernel": STAGED_BWD,
            "bound": "hbm", "achieved": round(achieved, 2),
            "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 5),
            "traffic": measured_traffic(sc.name), "avg_ms": round(acc["raster_bwd"], 4),
            "bytes_per_launch": bytes_bwd}
    # whole-step algorithmic bytes (B_alg) for the combined fraction
    P = 2 if T > 256 else 1
    B_alg = 232 * N + 20 * V + (20 + 24 * P) * isect + 40 * 2 * L_b + 24 * T + 40 * Himg * Wimg
    extra = {"num_intersects": isect, "visible": V, "L_b": L_b,
             "num_intersects_binned": isect_binned, "L_b_binned": L_b_binned,
             "B_alg_bytes": int(B_alg), "stage_sum_ms": round(sum(acc.values()), 4),
             "step_frac_hbm": round(B_alg / (sum(acc.values()) * 1e-3) / 1e9 / HBM_PEAK_GBS, 5)}
    return {k: round(v, 4) for k, v in acc.items()}, roof, extra


def _batched_workload(torch, name, dev):
    """Scenes, views, parameters and targets of a batched workload."""
    from jaxsplat_amd import ViewSpec, scenes
    if name == "orbit64":
        sc = scenes.make("orbit64")
        sizes = [sc.n]
        cams = [(0, c) for c in sc.cameras]
        params = sc.params()
        meta = dict(num_gaussians=sc.n, img_shape=list(sc.cameras[0].img_shape))
        glob, clip = sc.glob_scale, sc.clip_thresh
    elif name == "multi32":
        ms = scenes.make_multi()
        sizes = [m.n for m in ms]
        cams = [(i, m.camera) for i, m in enumerate(ms)]
        params = [np.concatenate(x) for x in zip(*[m.params() for m in ms])]
        meta = dict(num_gaussians=int(sum(sizes)), scene_sizes=sizes,
                    img_shapes=[list(m.camera.img_shape) for m in ms])
        glob, clip = ms[0].glob_scale, ms[0].clip_thresh
    else:
        raise ValueError(name)
    views = [ViewSpec(s_, c.img_shape, c.f, c.c) for s_, c in cams]
    params = [torch.from_numpy(np.ascontiguousarray(a)).to(dev) for a in params]
    vms = torch.from_numpy(np.stack([c.viewmat for _, c in cams])).to(dev)
    bgs = torch.zeros((len(cams), 3), dtype=torch.float32, device=dev)
    seed0 = {"orbit64": 2003, "multi32": 2004}[name]
    targets = [torch.from_numpy(scenes.target_image(c.img_shape, seed0 + v)).to(dev)
               for v, (_, c) in enumerate(cams)]
    return sizes, views, cams, params, vms, bgs, targets, meta, glob, clip


def batched_vs_plain(torch, name, steps, warmup=2):
    """All views of a batched workload per step: the batched path (autotuned,
    one HIP graph replay of forward + L2 gradient + backward) against one
    plain render fwd+bwd call per view (the reference's usage)."""
    import jaxsplat_amd
    from jaxsplat_amd import _lib
    from jaxsplat_amd._ops import _p, _stream
    from jaxsplat_amd.tune import autotune, tuned_plan
    dev = torch.device("cuda:0")
    sizes, views, cams, params, vms, bgs, targets, meta, glob, clip = \
        _batched_workload(torch, name, dev)
    V = len(views)
    tuning = autotune(sizes, views, *params, vms, bgs, glob_scale=glob, clip_thresh=clip)
    plan = tuned_plan(tuning, sizes, views, glob_scale=glob, clip_thresh=clip, device=dev)
    P = plan.num_pixels
    L = _lib.lib()
    target = torch.cat([t.reshape(-1, 3) for t in targets])
    # per-pixel weight 2 / (3 H W) of its view: d mean((img - target)^2)
    wpix = torch.cat([torch.full((t.shape[0] * t.shape[1], 1), 2.0 / t.numel(),
                                 dtype=torch.float32, device=dev) for t in targets])
    fT = torch.empty((P,), dtype=torch.float32, device=dev)
    fI = torch.empty((P,), dtype=torch.int32, device=dev)
    img = torch.empty((P, 3), dtype=torch.float32, device=dev)
    v_img = torch.empty((P, 3), dtype=torch.float32, device=dev)
    grads = [torch.empty_like(x) for x in params]

    def step():
        s = _stream(dev)
        _lib.check(L.jaxsplat_batch_forward(plan._h, *[_p(x) for x in params], _p(vms),
                                            _p(bgs), _p(fT), _p(fI), _p(img),
                                            _p(plan.workspace), plan.workspace_bytes, s))
        torch.mul(torch.sub(img, target), wpix, out=v_img)
        _lib.check(L.jaxsplat_batch_backward(plan._h, *[_p(x) for x in params], _p(vms),
                                             _p(bgs), _p(fT), _p(fI), _p(v_img),
                                             *[_p(g) for g in grads], 0, _p(plan.workspace),
                                             plan.workspace_bytes, s))

    side = torch.cuda.Stream()
    side.wait_stream(torch.cuda.current_stream())
    with torch.cuda.stream(side):
        for _ in range(warmup):
            step()
    torch.cuda.current_stream().wait_stream(side)
    torch.cuda.synchronize()
    isect = plan.check()
    graph = torch.cuda.CUDAGraph()
    with torch.cuda.graph(graph):
        step()
    for _ in range(warmup):
        graph.replay()
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(steps):
        graph.replay()
    e1.record()
    torch.cuda.synchronize()
    ms_b = e0.elapsed_time(e1) / steps
    plan.check()

    pp = [x.clone().requires_grad_(True) for x in params]
    offs = np.concatenate([[0], np.cumsum(sizes)])

    def plain_step():
        for x in pp:
            x.grad = None
        for v, (sidx, c) in enumerate(cams):
            a, b = int(offs[sidx]), int(offs[sidx + 1])
            out = jaxsplat_amd.render(*[x[a:b] for x in pp], viewmat=vms[v], background=bgs[v],
                                      img_shape=c.img_shape, f=c.f, c=c.c, glob_scale=glob,
                                      clip_thresh=clip, block_size=tuning.block_width)
            torch.nn.functional.mse_loss(out, targets[v]).backward()

    for _ in range(warmup):
        plain_step()
    torch.cuda.synchronize()
    e0.record()
    for _ in range(steps):
        plain_step()
    e1.record()
    torch.cuda.synchronize()
    ms_p = e0.elapsed_time(e1) / steps
    gb = torch.cat([g.reshape(-1) for g in grads])
    gp = torch.cat([x.grad.reshape(-1) for x in pp])
    rel = float((gb - gp).abs().max() / gp.abs().max().clamp_min(1e-30))
    return dict(workload=name, views=V, **meta, num_intersects=isect,
                tuning={"block_width": tuning.block_width, "fwd_depth": tuning.fwd_depth,
                        "bwd_depth": tuning.bwd_depth,
                        "candidates_ms": {f"bw{k[0]}_{k[1]}{k[2]}": round(v, 3)
                                          for k, v in tuning.table.items()}},
                ms_batched_graph=round(ms_b, 3), ms_plain_per_view_calls=round(ms_p, 3),
                mpix_s_batched=round(P / (ms_b * 1e-3) / 1e6, 1),
                mpix_s_plain=round(P / (ms_p * 1e-3) / 1e6, 1),
                speedup=round(ms_p / ms_b, 3), grad_max_rel_diff=rel)


# the rasterize backward of the render path: its record kernel and per-Gaussian
# reduction (the bitmap fill is a 1 MB memset at 1080p)
STAGED_BWD = "raster_bwd_kernel<*, *, *, true, *> + raster_bwd_reduce_kernel"


def measured_traffic(workload):
    """HBM bytes per launch of the dominant kernel from the committed
    rocprofv3 FETCH_SIZE/WRITE_SIZE runs (scaled by a copy-kernel
    calibration), or None when no measurement for this workload exists."""
    p = os.path.join(os.path.dirname(os.path.abspath(__file__)), "profiles",
                     f"{workload}_traffic.json")
    try:
        with open(p) as fh:
            rec = json.load(fh)
        return rec["bytes_per_launch"] if rec.get("kernel") == STAGED_BWD else None
    except (OSError, ValueError, KeyError):
        return None


def cpu_baseline(sc, cam, tgt_seed, n_tiles):
    """Oracle (float64 NumPy, 1 core) on a bounded sample of the same workload."""
    import oracle
    from jaxsplat_amd import scenes
    Himg, Wimg = cam.img_shape
    bw = sc.block_width
    kw = dict(img_shape=cam.img_shape, f=cam.f, c=cam.c, glob_scale=sc.glob_scale,
              clip_thresh=sc.clip_thresh, block_width=bw)
    t0 = time.perf_counter()
    pr = oracle.project.project_fwd(sc.means3d, sc.scales, sc.quats, cam.viewmat, **kw)
    ks, gs, bins = oracle.binning.bin_and_sort(pr["xys"].astype(np.float32),
                                               pr["depths"].astype(np.float32), pr["radii"],
                                               pr["cum_tiles_hit"], img_shape=cam.img_shape,
                                               block_width=bw)
    t_pre = time.perf_counter() - t0
    T = bins.shape[0]
    rng = np.random.default_rng(0)
    nonempty = np.nonzero(bins[:, 1] > bins[:, 0])[0]
    tiles = np.sort(rng.choice(nonempty, size=min(n_tiles, nonempty.size), replace=False))
    target = scenes.target_image(cam.img_shape, tgt_seed)
    t1 = time.perf_counter()
    f = oracle.rasterize.rasterize_fwd(gs, bins, pr["xys"], pr["conics"], sc.colors,
                                       sc.opacities, sc.background, img_shape=cam.img_shape,
                                       block_width=bw, tiles=tiles)
    v_img = 2.0 * (f["img"] - target) / target.size
    b = oracle.rasterize.rasterize_bwd(gs, bins, pr["xys"], pr["conics"], sc.colors, sc.opacities,
                                       sc.background, f["final_Ts"], f["final_idx"], v_img,
                                       np.zeros(cam.img_shape), img_shape=cam.img_shape,
                                       block_width=bw, tiles=tiles)
    vproj = oracle.project.project_bwd(sc.means3d, sc.scales, sc.quats, cam.viewmat, pr["radii"],
                                       pr["conics"], pr["compensation"], b["v_xy"],
                                       np.zeros(sc.n), b["v_conic"], np.zeros(sc.n),
                                       img_shape=cam.img_shape, f=cam.f, glob_scale=sc.glob_scale)
    t_tiles = time.perf_counter() - t1
    px = tiles.size * bw * bw
    t_eff = t_tiles + t_pre * tiles.size / T
    return {"value": round(px / t_eff / 1e6, 5), "unit": "Mpix/s", "cores": 1, "kind": "port",
            "sample": (f"{args_workload(sc)}: float64 oracle; projection + binning of all "
                       f"{sc.n} Gaussians ({t_pre:.1f} s, prorated {tiles.size}/{T} tiles) + "
                       f"rasterize fwd+bwd of {tiles.size} seeded random non-empty tiles "
                       f"({t_tiles:.1f} s) + projection bwd"),
            "host_cpu": cpu_model(), "host_logical_cpus": os.cpu_count()}


def args_workload(sc):
    return sc.name


def cpu_model():
    try:
        with open("/proc/cpuinfo") as fh:
            for ln in fh:
                if ln.startswith("model name"):
                    return ln.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


if __name__ == "__main__":
    main()
