"""Benchmark: fwd+bwd Mpix/s of the Gaussian-splatting render path on one MI355X.

Workload (default ``hd1m``): N=1,000,000 Gaussians, 1080x1920, fx=fy=1500,
c=(960,540), block 16, seeded synthetic scene (``jaxsplat_amd.scenes``), L2
loss against a seeded U(0,1) target.  One step = render forward (projection,
binning/sort, rasterize) + loss + backward (rasterize adjoint, projection
adjoint) through the public API ``jaxsplat_amd.render`` and torch autograd.
The headline (``value``, ``ms_per_step``) times that step captured once in a
HIP graph and replayed (``render`` with an ``IsectCapacity``, so nothing
synchronises with the host, and the fused L2 loss kernel); ``eager`` times the
same step launched eagerly with torch's ``mse_loss``.  Inputs are resident in
HBM before the timed region.

Prints ONE JSON line.  Besides the contract fields it carries:
* ``roofline``: the dominant kernels (the render backward: the rasterize
  adjoint's record kernel and the per-Gaussian tail that sums the records and
  runs the projection adjoint) -- algorithmic bytes per launch / their average
  duration measured with HIP events on the launch stream, against the 8 TB/s
  HBM peak -- plus the VALU-issue figures of both rasterize kernels from the
  committed rocprofv3 counters (``profiles/hd1m_counters.json``, used only
  while its source hash matches the kernels built here);
* ``stages``: per-stage milliseconds of the same pipeline (HIP events, and the
  library's stage marks inside the binning and the backward), each stage's
  algorithmic bytes and its fraction of the HBM peak;
* ``stats``: the whole-step fraction of the HBM peak with the algorithmic byte
  count B_alg of the reference algorithm (reference tile lists, 64-bit key
  sort), with L_f and L_b measured, and with this implementation's own
  per-stage byte model;
* ``graph``: the headline step's details (capacity, intersections, loss);
* ``cpu_baseline``: the float64 NumPy oracle on a bounded sample of the same
  workload, on a process pool over the host's cores;
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
import hashlib
import json
import math
import os
import sys
import time
from pathlib import Path

import numpy as np

HBM_PEAK_GBS = 8000.0
ROOT = Path(__file__).resolve().parent


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--workload", default="hd1m", choices=["small", "hd1m", "dense3m"])
    ap.add_argument("--cpu-baseline", dest="cpu_baseline", action="store_true", default=True)
    ap.add_argument("--no-cpu-baseline", dest="cpu_baseline", action="store_false")
    ap.add_argument("--cpu-tiles-per-worker", type=int, default=60)
    ap.add_argument("--cpu-workers", type=int, default=0,
                    help="CPU-baseline processes (0: the host's cores, at most 16)")
    ap.add_argument("--stage-iters", type=int, default=10)
    ap.add_argument("--graph", dest="graph", action="store_true", default=True)
    ap.add_argument("--no-graph", dest="graph", action="store_false")
    ap.add_argument("--batched", dest="batched", action="store_true", default=True)
    ap.add_argument("--no-batched", dest="batched", action="store_false")
    ap.add_argument("--batched-steps", type=int, default=5)
    ap.add_argument("--batched-workloads", default="orbit64,multi32")
    ap.add_argument("--copy-calibration-mb", type=int, default=0,
                    help="run a device-to-device copy of this many MiB first (HBM counter "
                         "calibration under rocprofv3 --pmc)")
    ap.add_argument("--summarize-profiles", metavar="DIR", default=None,
                    help="no GPU work: summarise the rocprofv3 outputs under DIR (trace/, fetch/, "
                         "write/, sq/, lds/ of the profiling commands) into profiles/")
    ap.add_argument("--profile-command", default="",
                    help="with --summarize-profiles: the profiled command, quoted in the summaries")
    return ap.parse_args()


def main():
    args = parse()
    if args.summarize_profiles:
        summarize_profiles(Path(args.summarize_profiles), args.workload, args.copy_calibration_mb,
                           args.profile_command)
        return
    if args.gpus != 1:
        print(f"bench.py measures one GPU only; --gpus {args.gpus} is not supported "
              "(the path would shard by view or scene, each GPU holding all Gaussians, with one "
              "all-reduce of the per-Gaussian gradients per step; only the one-GPU share is built)",
              file=sys.stderr)
        sys.exit(2)
    from jaxsplat_amd import scenes
    sc = scenes.make(args.workload)
    cam = sc.camera
    tgt_seed = {"small": 2000, "hd1m": 2001, "dense3m": 2002}[args.workload]
    # the CPU baseline runs first, in forked worker processes, before this
    # process touches the GPU
    cpu = cpu_baseline(sc, cam, tgt_seed, args) if args.cpu_baseline else None

    import torch
    import jaxsplat_amd
    from jaxsplat_amd import _lib
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

    Himg, Wimg = cam.img_shape
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
    ms_eager = e0.elapsed_time(e1) / args.steps
    mpix_eager = Himg * Wimg / (ms_eager * 1e-3) / 1e6

    stages, roof, extra = stage_breakdown(torch, sc, cam, params, target, args.stage_iters)
    graph = graph_leg(torch, sc, cam, params, kw, target, args.steps, args.warmup,
                      extra["num_intersects_binned"]) if args.graph else None
    # the headline step is the HIP-graph replay of the same step (the way a
    # training loop runs it: no host synchronisation, no per-launch host
    # work), with the eager figure beside it
    if graph is not None:
        ms, mpix, path = graph["ms_per_step"], graph["mpix_s"], graph["path"]
    else:
        ms, mpix = ms_eager, mpix_eager
        path = "jaxsplat_amd.render + torch mse_loss + autograd backward, eager"
    step_fractions(extra, ms, ms_eager)

    batched = [batched_vs_plain(torch, w, args.batched_steps)
               for w in args.batched_workloads.split(",") if w] if args.batched else None

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
                   "path": "plain single view: " + path,
                   "parallelism": "none (1 GPU)"},
        "eager": {"path": "jaxsplat_amd.render + torch mse_loss + autograd backward, eager "
                          "(the binning reads the intersection count back to the host)",
                  "ms_per_step": round(ms_eager, 4), "mpix_s": round(mpix_eager, 1)},
        "roofline": roof,
        "cpu_baseline": cpu,
        "stages": stages,
        "stats": extra,
        "graph": graph,
        "batched": batched,
        "wall_s": round(wall, 3),
    }
    print(json.dumps(rec), flush=True)


# ---------------------------------------------------------------------------
# Committed counter summaries (rocprofv3) and their validity
# ---------------------------------------------------------------------------
RENDER_BWD = "raster_bwd_kernel<*, *, *, true, *> + render_bwd_tail_kernel"


def source_hash():
    """sha256 (16 hex digits) of the HIP sources the library is built from."""
    h = hashlib.sha256()
    for p in sorted((ROOT / "jaxsplat_amd" / "csrc").glob("*")):
        if p.suffix in (".hip", ".hpp", ".h"):
            h.update(p.name.encode())
            h.update(p.read_bytes())
    h.update((ROOT / "include" / "jaxsplat_amd.h").read_bytes())
    return h.hexdigest()[:16]


def committed_counters(workload):
    """The committed rocprofv3 counter summary of this workload, if it was
    measured on the kernels built from the current sources (else None and a
    reason)."""
    p = ROOT / "profiles" / f"{workload}_counters.json"
    try:
        rec = json.loads(p.read_text())
    except (OSError, ValueError):
        return None, f"no {p.name}"
    if rec.get("source_sha16") != source_hash():
        return None, f"{p.name} was measured on other kernel sources (stale)"
    return rec, None


def _plain_name(name):
    """A kernel name without its return type and anonymous namespaces."""
    name = name.replace("(anonymous namespace)::", "")
    return name[5:] if name.startswith("void ") else name


def _kernel_entry(rec, prefix):
    for name, v in rec.get("kernels", {}).items():
        if _plain_name(name).startswith(prefix):
            return v
    return None


# ---------------------------------------------------------------------------
# Per-stage timing and byte accounting
# ---------------------------------------------------------------------------
def consumed_lengths(torch, gids, bins, xys, conics, opac, fT, fI, img_shape, bw, chunk=1 << 17):
    """(L_f, L_b) of one render on the given tile lists: L_b = sum over tiles
    of (largest final index - start + 1); L_f = sum over tiles of the
    consumed prefix of the forward: (largest termination index - start + 1)
    when every pixel of the tile terminated, else the whole list.  A pixel
    terminated iff some pair after its last contributor passes the alpha test
    (the first such pair is where it stopped); a pixel with T = 1 had no
    contributor and never terminated.  Pairs are evaluated on the GPU in
    float32 with the reference's formula; borderline pairs may be decided
    differently from the kernel, which moves L_f by a few entries at most."""
    H, W = img_shape
    tx_n, ty_n = (W + bw - 1) // bw, (H + bw - 1) // bw
    T = tx_n * ty_n
    P = bw * bw
    dev = gids.device
    b = bins.to(torch.int64)
    lens = (b[:, 1] - b[:, 0]).clamp_min(0)
    fT = fT.reshape(-1)
    fI = fI.reshape(-1).to(torch.int64)
    # per tile pixel: image index and inside mask
    py = torch.arange(P, device=dev) // bw
    px = torch.arange(P, device=dev) % bw
    tiles = torch.arange(T, device=dev)
    gy = (tiles // tx_n)[:, None] * bw + py[None, :]
    gx = (tiles % tx_n)[:, None] * bw + px[None, :]
    inside = (gy < H) & (gx < W)
    lin = torch.where(inside, gy * W + gx, torch.zeros_like(gy))
    fidx_t = torch.where(inside, fI[lin], torch.full_like(lin, -1))
    # L_b
    tmax = fidx_t.max(dim=1).values
    L_b = int(torch.clamp(tmax - b[:, 0] + 1, min=torch.zeros_like(lens), max=lens).sum())
    # L_f
    big = torch.iinfo(torch.int64).max
    contrib = inside & (fT[lin] < 1.0)
    term = torch.full((T * P,), big, dtype=torch.int64, device=dev)
    n = int(lens.sum())
    tile_of = torch.repeat_interleave(tiles, lens)
    ks = torch.repeat_interleave(b[:, 0], lens) + \
        (torch.arange(n, device=dev) - torch.repeat_interleave(torch.cumsum(lens, 0) - lens, lens))
    xy = xys.reshape(-1, 2)
    cn = conics.reshape(-1, 3)
    op = opac.reshape(-1)
    for k0 in range(0, n, chunk):
        t = tile_of[k0:k0 + chunk]
        k = ks[k0:k0 + chunk]
        g = gids[k].to(torch.int64)
        dx = xy[g, 0][:, None] - (gx[t].to(torch.float32) + 0.5)
        dy = xy[g, 1][:, None] - (gy[t].to(torch.float32) + 0.5)
        sig = 0.5 * (cn[g, 0][:, None] * dx * dx + cn[g, 2][:, None] * dy * dy) + \
            cn[g, 1][:, None] * dx * dy
        alpha = torch.clamp(op[g][:, None] * torch.exp(-sig), max=0.999)
        ok = (sig >= 0) & (alpha >= 1.0 / 255.0) & contrib[t] & (k[:, None] > fidx_t[t])
        cand = torch.where(ok, k[:, None].expand_as(ok), torch.full_like(ok, big, dtype=torch.int64))
        dest = t[:, None] * P + torch.arange(P, device=dev)[None, :]
        term.scatter_reduce_(0, dest.reshape(-1), cand.reshape(-1), reduce="amin")
    term = term.reshape(T, P)
    done = (term < big) | ~inside
    all_done = done.all(dim=1) & (lens > 0)
    tmax_term = torch.where(inside, term, torch.full_like(term, -1)).max(dim=1).values
    lf_tile = torch.where(all_done, tmax_term - b[:, 0] + 1, lens)
    L_f = int(lf_tile.sum())
    return L_f, L_b, int(all_done.sum())


def _popcount32(torch, words):
    w = words.to(torch.int64) & 0xFFFFFFFF
    c = torch.zeros_like(w)
    for s in range(32):
        c += (w >> s) & 1
    return int(c.sum())


def stage_breakdown(torch, sc, cam, params, target, iters):
    """The render path through the C ABI with HIP events between stages and
    the library's stage marks inside the binning and the staged backward."""
    from jaxsplat_amd import _lib
    from jaxsplat_amd._ops import _bin, _p, _stream
    L = _lib.lib()
    dev = params[0].device
    N = sc.n
    Himg, Wimg = cam.img_shape
    bw = sc.block_width
    means3d, scales, quats, colors, opac = [p.detach() for p in params]
    viewmat = torch.from_numpy(cam.viewmat).to(dev)
    bg = torch.from_numpy(sc.background).to(dev)
    d = _lib.make_descriptor(N, cam.img_shape, cam.f, cam.c, sc.glob_scale, sc.clip_thresh, bw)
    D = ctypes.byref(d)
    f32 = dict(dtype=torch.float32, device=dev)
    i32 = dict(dtype=torch.int32, device=dev)
    xys, depths, conics = torch.empty((N, 2), **f32), torch.empty((N, 1), **f32), \
        torch.empty((N, 3), **f32)
    comp = torch.empty((N, 1), **f32)
    radii, nth, cum = torch.empty((N, 1), **i32), torch.empty((N, 1), **i32), \
        torch.empty((N, 1), **i32)
    # the render path's projection writes the opacity-limited tile rectangles
    # and counts them (reference rectangles and counts further below)
    rects, tnth, tcum = torch.empty((N, 2), **i32), torch.empty((N, 1), **i32), \
        torch.empty((N, 1), **i32)
    wsb = L.jaxsplat_project_workspace_size(D)
    ws = torch.empty(wsb, dtype=torch.uint8, device=dev)
    s = _stream(dev)

    splats = torch.empty((N, 12), **f32)

    def project(opacity_limited, analysis=False):
        if opacity_limited:
            # the render path: opacity-limited rectangles and splat records
            # (xys and the per-Gaussian counts only for the analysis below:
            # the render path writes neither)
            _lib.check(L.jaxsplat_project_forward_splats(
                D, _p(means3d), _p(scales), _p(quats), _p(viewmat), _p(opac), _p(colors), 3,
                _p(xys) if analysis else None, _p(depths), _p(radii), _p(conics), None,
                _p(tnth) if analysis else None, _p(tcum), _p(rects), _p(splats), _p(ws), wsb,
                s))
        else:
            _lib.check(L.jaxsplat_project_forward(
                D, _p(means3d), _p(scales), _p(quats), _p(viewmat), None, _p(xys), _p(depths),
                _p(radii), _p(conics), _p(comp), _p(nth), _p(cum), _p(ws), wsb, s))

    project(True, analysis=True)
    isect = int(tcum[-1, 0].item())                 # read once, outside the timed loop
    ntiles = d.grid_dim_2d.x * d.grid_dim_2d.y
    ewsb = L.jaxsplat_bin_emit_workspace_size(D, isect)
    ews = torch.empty(ewsb, dtype=torch.uint8, device=dev)
    gids = torch.empty(max(isect, 1), **i32)
    bins = torch.empty((ntiles, 2), **i32)
    fT, fI = torch.empty((Himg, Wimg), **f32), torch.empty((Himg, Wimg), **i32)
    img = torch.empty((Himg, Wimg, 3), **f32)
    v_img = torch.empty_like(img)
    vc, vo, vxy, vcn = (torch.empty((N, 3), **f32), torch.empty((N, 1), **f32),
                        torch.empty((N, 2), **f32), torch.empty((N, 3), **f32))
    vm, vs, vq = torch.empty((N, 3), **f32), torch.empty((N, 3), **f32), torch.empty((N, 4), **f32)
    bwsb = L.jaxsplat_rasterize_backward_staged_workspace_size(D, isect)
    bws = torch.empty(bwsb, dtype=torch.uint8, device=dev)
    names = ["project_fwd", "emission", "tile_sort", "bin_edges", "tile_depth_sort",
             "raster_fwd", "loss_grad", "raster_bwd_records", "bwd_tail"]
    acc = {k: 0.0 for k in names}
    marks = [torch.cuda.Event(enable_timing=True) for _ in range(4)]
    for m in marks:
        m.record()                                  # creates the event
    torch.cuda.synchronize()
    handles = (ctypes.c_void_p * 4)(*[m.cuda_event for m in marks])
    L.jaxsplat_set_stage_marks(handles, 4)
    try:
        for it in range(iters + 2):
            ev = [torch.cuda.Event(enable_timing=True) for _ in range(7)]
            ev[0].record()
            project(True)
            ev[1].record()
            ev[2].record()
            _lib.check(L.jaxsplat_bin_rects_sort(D, _p(depths), _p(rects), _p(tcum), isect,
                                                 _p(gids), None, _p(bins), _p(ews), ewsb, s))
            ev[3].record()
            _lib.check(L.jaxsplat_rasterize_forward_splats(
                D, _p(gids), _p(bins), isect, _p(splats), _p(bg), 3, _p(fT), _p(fI), _p(img),
                s))
            ev[4].record()
            torch.mul(torch.sub(img, target), 2.0 / img.numel(), out=v_img)
            ev[5].record()
            # the render backward: the staged records, then the per-Gaussian
            # tail (record sums + projection adjoint)
            _lib.check(L.jaxsplat_render_backward(
                D, _p(gids), _p(bins), isect, _p(means3d), _p(scales), _p(quats), _p(viewmat),
                _p(splats), _p(conics), _p(bg), 3, _p(radii), _p(rects), _p(tcum), _p(fT),
                _p(fI), _p(v_img), _p(vc), _p(vo), _p(vm), _p(vs), _p(vq), 0, _p(bws), bwsb,
                s))
            ev[6].record()
            torch.cuda.synchronize()
            if it >= 2:
                t = [ev[0].elapsed_time(ev[1]),
                     ev[2].elapsed_time(marks[0]), marks[0].elapsed_time(marks[1]),
                     marks[1].elapsed_time(marks[3]), marks[3].elapsed_time(ev[3]),
                     ev[3].elapsed_time(ev[4]),
                     ev[4].elapsed_time(ev[5]), ev[5].elapsed_time(marks[2]),
                     marks[2].elapsed_time(ev[6])]
                for k, n in enumerate(names):
                    acc[n] += t[k] / iters
    finally:
        L.jaxsplat_set_stage_marks(None, 0)
    _lib.check(L.jaxsplat_bin_check(D, isect, _p(ews), s))

    # ---- counts on the render path's (opacity-limited) lists ----
    V = int((radii > 0).sum().item())
    E = int((tnth > 0).sum().item())
    R = _popcount32(torch, bws[((48 * max(isect, 1) + 255) // 256) * 256:].view(torch.int32)
                    [:(isect + 31) // 32])
    L_ft, L_bt, _ = consumed_lengths(torch, gids[:isect], bins, xys, conics, opac, fT, fI,
                                     cam.img_shape, bw)
    # ---- the reference algorithm's counts: its tile lists of the same projection ----
    project(False)
    I_ref = int(cum[-1, 0].item())
    rgids, rbins, _ = _bin(d, xys, depths, radii, I_ref, dev)
    fT2, fI2, img2 = torch.empty_like(fT), torch.empty_like(fI), torch.empty_like(img)
    _lib.check(L.jaxsplat_rasterize_forward_binned(
        D, _p(rgids), _p(rbins), I_ref, _p(xys), _p(conics), _p(colors), _p(opac), _p(bg),
        _p(fT2), _p(fI2), _p(img2), s))
    L_f, L_b, tiles_done = consumed_lengths(torch, rgids, rbins, xys, conics, opac, fT2, fI2,
                                            cam.img_shape, bw)
    same_img = bool(torch.equal(img2.view(torch.int32), img.view(torch.int32)))
    T = ntiles
    Px = Himg * Wimg

    # ---- algorithmic bytes of each stage of this implementation ----
    # (documented term by term in the design notes; N Gaussians, V visible,
    # E with a non-empty opacity-limited rectangle, I = isect pairs binned,
    # L_ft / L_bt consumed by the forward / backward, R records flushed).
    # The projection reads 56 B per Gaussian (means, scales, quats, opacity,
    # colour) and writes 80 (depth, radius, conic, inclusive sum, rectangle,
    # 48-byte splat record) + 8 of scan; a staged list entry costs
    # its id and its splat record (52 B), in the backward also its inclusive
    # sum; a flushed record is 36 B.
    stage_bytes = {
        "project_fwd": 144 * N,
        "emission": 24 * N + 8 * isect,
        "tile_sort": 40 * isect,
        "bin_edges": 4 * isect + 8 * T,
        "tile_depth_sort": 12 * isect + 8 * T,
        "raster_fwd": 52 * L_ft + 8 * T + 20 * Px,
        "loss_grad": 36 * Px,
        "raster_bwd_records": 56 * L_bt + 8 * T + 20 * Px + 36 * R + (isect + 7) // 8,
        "bwd_tail": 36 * R + (isect + 7) // 8 + 16 * N + 56 * V + 4 * (N - V) + 40 * N,
    }
    # the raster kernels are bound by VALU issue, the per-tile depth sort by
    # the latency of its LDS ranking rounds and barriers, the rest by HBM
    bound_by = {"raster_fwd": "valu", "raster_bwd_records": "valu", "tile_depth_sort": "lds"}
    stages = {}
    for k in names:
        ms_k = acc[k]
        gbs = stage_bytes[k] / (ms_k * 1e-3) / 1e9 if ms_k > 0 else 0.0
        stages[k] = {"ms": round(ms_k, 4), "alg_bytes": int(stage_bytes[k]),
                     "GBs": round(gbs, 1), "frac_hbm": round(gbs / HBM_PEAK_GBS, 4),
                     "bound": bound_by.get(k, "hbm")}

    # ---- roofline of the dominant kernels: the render backward ----
    # (the rasterize adjoint's record kernel and the per-Gaussian tail that
    # sums the records and runs the projection adjoint).  Algorithmic bytes
    # per launch, counted on the reference's lists of the same projection
    # (L_b): 40 B per consumed intersection (id 4 + xy 8 + opacity 4 + conic
    # 12 + colour 12) + 20 B per pixel (final T 4 + final index 4 + v_img 12)
    # + 8 B per tile (bin read) + 16 B per Gaussian (v_rgb 12 + v_opacity 4
    # written) + the projection adjoint's 56 B per visible Gaussian read
    # (means, scales, quats 40, radius 4, conic 12) + 4 B per culled one
    # (radius) + 40 B per Gaussian written (v_mean3d, v_scale, v_quat)
    bytes_bwd = 40 * L_b + 20 * Px + 8 * T + 16 * N + 56 * V + 4 * (N - V) + 40 * N
    ms_bwd = acc["raster_bwd_records"] + acc["bwd_tail"]
    achieved = bytes_bwd / (ms_bwd * 1e-3) / 1e9
    counters, why = committed_counters(sc.name)
    roof = {"kernel": RENDER_BWD, "bound": "hbm", "achieved": round(achieved, 2),
            "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 5),
            "traffic": None, "avg_ms": round(ms_bwd, 4), "bytes_per_launch": int(bytes_bwd)}
    if counters is not None:
        kr = _kernel_entry(counters, "jsa::raster_bwd_kernel<false, false, 64, true, 4>")
        kd = _kernel_entry(counters, "jsa::render_bwd_tail_kernel")
        kf = _kernel_entry(counters, "jsa::raster_fwd_kernel<256>")
        if kr and kd and kr.get("hbm_bytes") and kd.get("hbm_bytes"):
            roof["traffic"] = int(kr["hbm_bytes"] + kd["hbm_bytes"])
            roof["traffic_over_alg"] = round(roof["traffic"] / bytes_bwd, 3)
        valu = {}
        for key, k_ in (("raster_bwd_records", kr), ("raster_fwd", kf)):
            if k_ and k_.get("valu_busy") is not None:
                ent = L_bt if key == "raster_bwd_records" else L_ft
                valu[key] = {"valu_busy": k_["valu_busy"],
                             "valu_insts_per_launch": k_.get("valu_insts"),
                             "valu_insts_per_entry": round(k_["valu_insts"] / ent, 1)
                             if k_.get("valu_insts") else None,
                             "lds_bank_conflict_frac": k_.get("lds_bank_conflict_frac"),
                             "l2_atomics_per_launch": k_.get("l2_atomics")}
        roof["valu"] = valu
        roof["counters_source"] = f"profiles/{sc.name}_counters.json"
    else:
        roof["counters_note"] = why

    # ---- whole-step fractions of the HBM peak ----
    # B_alg of the reference algorithm (one sort per step, 64-bit keys on
    # 32 + ceil(log2 T) bits in 8-bit LSD passes)
    P = math.ceil((32 + math.ceil(math.log2(T))) / 8)
    B_alg = 232 * N + 20 * V + (20 + 24 * P) * I_ref + 40 * (L_f + L_b) + 24 * T + 40 * Px
    step_ms = sum(acc.values())
    B_impl = sum(stage_bytes.values())
    extra = {"num_intersects": I_ref, "visible": V, "emitting": E, "L_f": L_f, "L_b": L_b,
             "tiles_all_terminated": tiles_done, "sort_passes_P": P,
             "num_intersects_binned": isect, "L_f_binned": L_ft, "L_b_binned": L_bt,
             "records_flushed": R, "images_equal_on_both_lists": same_img,
             "B_alg_bytes": int(B_alg), "stage_sum_ms": round(step_ms, 4),
             "impl_bytes_per_step": int(B_impl)}
    return stages, roof, extra


def step_fractions(extra, ms, ms_eager):
    """Whole-step fractions of the HBM peak: B_alg (the reference algorithm's
    bytes) and this implementation's per-stage byte model, over the headline
    step time and over the eager step's."""
    def frac(b, t_ms):
        return round(b / (t_ms * 1e-3) / 1e9 / HBM_PEAK_GBS, 5)
    extra["step_frac_hbm"] = frac(extra["B_alg_bytes"], ms)
    extra["impl_step_frac_hbm"] = frac(extra["impl_bytes_per_step"], ms)
    extra["step_frac_hbm_eager"] = frac(extra["B_alg_bytes"], ms_eager)


# ---------------------------------------------------------------------------
# The same step as one HIP graph replay
# ---------------------------------------------------------------------------
def graph_leg(torch, sc, cam, params, kw, target, steps, warmup, isect_hint):
    """render() with an IsectCapacity (no host synchronisation) + the fused
    L2 loss and gradient (jaxsplat_fit_l2: loss sum and v_img in one kernel) +
    autograd backward, captured once in a HIP graph and replayed."""
    import jaxsplat_amd
    from jaxsplat_amd import _lib
    from jaxsplat_amd._ops import _p, _stream
    L = _lib.lib()
    dev = params[0].device
    cap = jaxsplat_amd.IsectCapacity(int(isect_hint * 1.1) + 4096, device=dev)
    static = [p.detach().clone().requires_grad_(True) for p in params]
    v_img = torch.empty_like(target)
    loss_sum = torch.zeros(1, dtype=torch.float32, device=dev)

    def gstep():
        img = jaxsplat_amd.render(*static, **kw, isect_capacity=cap)
        _lib.check(L.jaxsplat_fit_l2(img.numel(), _p(img), _p(target), _p(v_img), _p(loss_sum),
                                     _stream(dev)))
        img.backward(v_img)

    side = torch.cuda.Stream()
    side.wait_stream(torch.cuda.current_stream())
    with torch.cuda.stream(side):
        for _ in range(2):
            for p in static:
                p.grad = None
            gstep()
    torch.cuda.current_stream().wait_stream(side)
    torch.cuda.synchronize()
    for p in static:
        p.grad = None
    graph = torch.cuda.CUDAGraph()
    with torch.cuda.graph(graph):
        gstep()
    for _ in range(warmup):
        graph.replay()
    torch.cuda.synchronize()
    cap.check()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(steps):
        graph.replay()
    e1.record()
    torch.cuda.synchronize()
    total = cap.check()                    # raises if a replay overflowed the capacity
    ms = e0.elapsed_time(e1) / steps
    H, W = cam.img_shape
    return {"path": "jaxsplat_amd.render(isect_capacity=...) + fused L2 loss and gradient "
                    "(jaxsplat_fit_l2) + autograd backward, captured once, one HIP graph "
                    "replay per step",
            "ms_per_step": round(ms, 4), "mpix_s": round(H * W / (ms * 1e-3) / 1e6, 1),
            "capacity": cap.capacity, "num_intersects": total,
            "loss": float(loss_sum.item()) / target.numel()}


# ---------------------------------------------------------------------------
# Batched workloads
# ---------------------------------------------------------------------------
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


# tile sides the autotuner may pick per workload: the tile side is part of
# the reference's semantics (a Gaussian joins the lists of the tiles its
# 3-sigma rectangle touches), so the plain calls it is compared with use the
# same side; the uneven multi-scene workload allows 8 and 16
BLOCK_WIDTHS = {"orbit64": (16,), "multi32": (8, 16)}


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
    tuning = autotune(sizes, views, *params, vms, bgs, glob_scale=glob, clip_thresh=clip,
                      block_widths=BLOCK_WIDTHS[name])
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
                tuning={"block_widths_allowed": list(BLOCK_WIDTHS[name]),
                        "block_width": tuning.block_width, "fwd_depth": tuning.fwd_depth,
                        "bwd_depth": tuning.bwd_depth, "schedule": tuning.schedule,
                        "candidates_ms": {f"bw{k[0]}_{k[1]}{k[2]}": round(v, 3)
                                          for k, v in tuning.table.items()}},
                ms_batched_graph=round(ms_b, 3), ms_plain_per_view_calls=round(ms_p, 3),
                mpix_s_batched=round(P / (ms_b * 1e-3) / 1e6, 1),
                mpix_s_plain=round(P / (ms_p * 1e-3) / 1e6, 1),
                speedup=round(ms_p / ms_b, 3), grad_max_rel_diff=rel)


# ---------------------------------------------------------------------------
# CPU baseline: the float64 oracle on a process pool
# ---------------------------------------------------------------------------
_CPU = {}


def _cpu_worker_init():
    try:
        from threadpoolctl import threadpool_limits
        threadpool_limits(1)
    except ImportError:
        pass


def _cpu_tiles(tiles):
    """Rasterize forward + backward of the given tiles (one worker)."""
    import oracle
    c = _CPU
    t0 = time.perf_counter()
    f = oracle.rasterize.rasterize_fwd(c["gs"], c["bins"], c["xys"], c["conics"], c["colors"],
                                       c["opac"], c["bg"], img_shape=c["img_shape"],
                                       block_width=c["bw"], tiles=tiles)
    v_img = 2.0 * (f["img"] - c["target"]) / c["target"].size
    oracle.rasterize.rasterize_bwd(c["gs"], c["bins"], c["xys"], c["conics"], c["colors"],
                                   c["opac"], c["bg"], f["final_Ts"], f["final_idx"], v_img,
                                   np.zeros(c["img_shape"]), img_shape=c["img_shape"],
                                   block_width=c["bw"], tiles=tiles)
    return time.perf_counter() - t0


def cpu_baseline(sc, cam, tgt_seed, args):
    """The float64 NumPy oracle (``kind: port``) on a bounded sample of the
    same workload: projection and binning of all N Gaussians and the
    projection backward of all N (one process each, prorated to the sampled
    tiles' share of the frame and spread over the pool's cores), plus the
    rasterize forward + backward of seeded random non-empty tiles, split
    over a pool of forked worker processes (BLAS threads pinned to 1)."""
    import multiprocessing as mp
    import oracle
    from jaxsplat_amd import scenes
    bw = sc.block_width
    kw = dict(img_shape=cam.img_shape, f=cam.f, c=cam.c, glob_scale=sc.glob_scale,
              clip_thresh=sc.clip_thresh, block_width=bw)
    workers = args.cpu_workers or min(16, os.cpu_count() or 1)
    t0 = time.perf_counter()
    pr = oracle.project.project_fwd(sc.means3d, sc.scales, sc.quats, cam.viewmat, **kw)
    ks, gs, bins = oracle.binning.bin_and_sort(pr["xys"].astype(np.float32),
                                               pr["depths"].astype(np.float32), pr["radii"],
                                               pr["cum_tiles_hit"], img_shape=cam.img_shape,
                                               block_width=bw)
    t_pre = time.perf_counter() - t0
    rng = np.random.default_rng(0)
    vxy = rng.normal(size=(sc.n, 2)) * 1e-6
    vcn = rng.normal(size=(sc.n, 3)) * 1e-6
    t1 = time.perf_counter()
    oracle.project.project_bwd(sc.means3d, sc.scales, sc.quats, cam.viewmat, pr["radii"],
                               pr["conics"], pr["compensation"], vxy, np.zeros(sc.n), vcn,
                               np.zeros(sc.n), img_shape=cam.img_shape, f=cam.f,
                               glob_scale=sc.glob_scale)
    t_pbwd = time.perf_counter() - t1
    T = bins.shape[0]
    nonempty = np.nonzero(bins[:, 1] > bins[:, 0])[0]
    n_tiles = min(workers * args.cpu_tiles_per_worker, nonempty.size)
    tiles = np.sort(rng.choice(nonempty, size=n_tiles, replace=False))
    _CPU.update(gs=gs, bins=bins, xys=pr["xys"], conics=pr["conics"], colors=sc.colors,
                opac=sc.opacities, bg=sc.background, img_shape=cam.img_shape, bw=bw,
                target=scenes.target_image(cam.img_shape, tgt_seed))
    chunks = [c for c in np.array_split(tiles, workers) if c.size]
    t2 = time.perf_counter()
    with mp.get_context("fork").Pool(len(chunks), initializer=_cpu_worker_init) as pool:
        per = pool.map(_cpu_tiles, chunks)
    t_tiles = time.perf_counter() - t2
    _CPU.clear()
    px = n_tiles * bw * bw
    # the serial stages' share of the sampled tiles, spread over the cores
    t_eff = t_tiles + (t_pre + t_pbwd) * n_tiles / T / workers
    return {"value": round(px / t_eff / 1e6, 5), "unit": "Mpix/s", "cores": workers,
            "kind": "port",
            "sample": (f"{sc.name}: float64 oracle; rasterize fwd+bwd of {n_tiles} seeded random "
                       f"non-empty tiles of {T} on {len(chunks)} worker processes "
                       f"({t_tiles:.1f} s wall, {sum(per):.1f} s CPU); projection + binning "
                       f"({t_pre:.1f} s) and projection bwd ({t_pbwd:.1f} s) of all {sc.n} "
                       f"Gaussians on one core, prorated to {n_tiles}/{T} tiles and divided "
                       f"over the {workers} cores"),
            "host_cpu": cpu_model(), "host_logical_cpus": os.cpu_count()}


# ---------------------------------------------------------------------------
# rocprofv3 output -> committed summaries (no GPU work)
# ---------------------------------------------------------------------------
# the profiling runs, each into DIR/<sub> with -o run --output-format csv
PROFILE_PASSES = {
    "trace": "--kernel-trace --stats",
    "fetch": "--pmc FETCH_SIZE",
    "write": "--pmc WRITE_SIZE",
    "sq": "--pmc SQ_WAVES SQ_INSTS_VALU SQ_INSTS_SALU SQ_INSTS_LDS SQ_ACTIVE_INST_VALU "
          "SQ_WAVE_CYCLES SQ_BUSY_CYCLES GRBM_GUI_ACTIVE",
    "lds": "--pmc SQ_LDS_BANK_CONFLICT SQ_LDS_ADDR_CONFLICT SQ_LDS_IDX_ACTIVE SQ_ACTIVE_INST_LDS "
           "SQ_WAIT_INST_LDS GRBM_GUI_ACTIVE TCC_ATOMIC_sum",
}
NUM_CU = 256
NUM_XCD = 8


def _read_csv_rows(paths):
    import csv
    rows = []
    for p in paths:
        with open(p, newline="") as fh:
            rows.extend(csv.DictReader(fh))
    return rows


def _counter_table(sub_dir):
    """{kernel name: {counter: [value per dispatch]}} from a --pmc run."""
    out = {}
    rows = _read_csv_rows(sorted(sub_dir.rglob("*counter_collection.csv")))
    per = {}
    for r in rows:
        key = (r["Dispatch_Id"], r["Counter_Name"])
        per[key] = per.get(key, 0.0) + float(r["Counter_Value"])
        out.setdefault(r["Kernel_Name"], {}).setdefault(r["Counter_Name"], {})
        out[r["Kernel_Name"]][r["Counter_Name"]][r["Dispatch_Id"]] = per[key]
    return {k: {c: list(d.values()) for c, d in v.items()} for k, v in out.items()}


def _short(name, n=72):
    return name if len(name) <= n else name[:n]


def summarize_profiles(prof, workload, calib_mb, command):
    """Kernel statistics, HBM traffic (FETCH_SIZE / WRITE_SIZE scaled by the
    copy calibration), VALU issue and LDS / atomic counters per kernel, written
    as profiles/<workload>_counters.json (read back by the bench line while
    its source hash matches the kernels built here) and two Markdown tables."""
    prof = Path(prof)
    rec = {"workload": workload, "source_sha16": source_hash(),
           "command": command or "(not given)", "passes": {}, "kernels": {}}
    kern = rec["kernels"]
    stats = _read_csv_rows(sorted((prof / "trace").rglob("*kernel_stats.csv")))
    for r in stats:
        kern.setdefault(r["Name"], {}).update(
            calls=int(r["Calls"]), avg_us=float(r["AverageNs"]) / 1e3,
            min_us=float(r["MinNs"]) / 1e3, max_us=float(r["MaxNs"]) / 1e3,
            pct=float(r["Percentage"]))
    rec["passes"]["trace"] = PROFILE_PASSES["trace"]
    # copy calibration: the calib_mb MiB device copies are the copy kernel's
    # dispatches with the largest counts; bytes moved each way = calib_mb MiB
    scale = {}
    for sub, ctr in (("fetch", "FETCH_SIZE"), ("write", "WRITE_SIZE")):
        if not (prof / sub).is_dir():
            continue
        rec["passes"][sub] = PROFILE_PASSES[sub]
        tab = _counter_table(prof / sub)
        if calib_mb:
            copies = [v for k, c in tab.items() if "copyBuffer" in k for v in c.get(ctr, [])]
            if copies:
                top = sorted(copies)[-3:]
                scale[sub] = calib_mb * 1024.0 / (sum(top) / len(top))   # KiB per KiB counted
        for k, c in tab.items():
            if ctr in c:
                vals = c[ctr]
                kern.setdefault(k, {})[ctr.lower() + "_kib"] = sum(vals) / len(vals)
    rec["fetch_scale"] = round(scale.get("fetch", 1.0), 4)
    rec["write_scale"] = round(scale.get("write", 1.0), 4)
    for k, v in kern.items():
        if "fetch_size_kib" in v and "write_size_kib" in v:
            v["hbm_bytes"] = int(1024 * (v["fetch_size_kib"] * rec["fetch_scale"] +
                                         v["write_size_kib"] * rec["write_scale"]))
            if v.get("avg_us"):
                v["hbm_GBs"] = round(v["hbm_bytes"] / (v["avg_us"] * 1e-6) / 1e9, 1)
    for sub in ("sq", "lds"):
        if not (prof / sub).is_dir():
            continue
        rec["passes"][sub] = PROFILE_PASSES[sub]
        for k, c in _counter_table(prof / sub).items():
            v = kern.setdefault(k, {})
            mean = {n: sum(x) / len(x) for n, x in c.items()}
            cyc = mean.get("GRBM_GUI_ACTIVE", 0.0) / NUM_XCD     # per XCD
            if sub == "sq":
                v["waves"] = mean.get("SQ_WAVES")
                v["valu_insts"] = mean.get("SQ_INSTS_VALU")
                v["salu_insts"] = mean.get("SQ_INSTS_SALU")
                v["lds_insts"] = mean.get("SQ_INSTS_LDS")
                v["cycles"] = cyc
                if cyc and "SQ_ACTIVE_INST_VALU" in mean:
                    v["valu_busy"] = round(100.0 * mean["SQ_ACTIVE_INST_VALU"] / NUM_CU / cyc, 1)
            else:
                v["lds_bank_conflict_cycles"] = mean.get("SQ_LDS_BANK_CONFLICT")
                v["lds_addr_conflict_cycles"] = mean.get("SQ_LDS_ADDR_CONFLICT")
                v["lds_active_cycles"] = mean.get("SQ_LDS_IDX_ACTIVE")
                v["lds_active_inst_cycles"] = mean.get("SQ_ACTIVE_INST_LDS")
                v["lds_wait_inst_cycles"] = mean.get("SQ_WAIT_INST_LDS")
                v["l2_atomics"] = mean.get("TCC_ATOMIC_sum", mean.get("TCC_ATOMIC"))
                if cyc and mean.get("SQ_LDS_BANK_CONFLICT") is not None:
                    v["lds_bank_conflict_frac"] = round(
                        mean["SQ_LDS_BANK_CONFLICT"] / NUM_CU / cyc, 4)
                if mean.get("SQ_LDS_IDX_ACTIVE"):
                    v["lds_bank_conflict_per_active"] = round(
                        mean.get("SQ_LDS_BANK_CONFLICT", 0.0) / mean["SQ_LDS_IDX_ACTIVE"], 4)
    out_dir = ROOT / "profiles"
    (out_dir / f"{workload}_counters.json").write_text(json.dumps(rec, indent=1) + "\n")

    # the library's kernels and the calibration copy; the bench's own torch
    # analysis kernels (consumed lengths, outside the timed regions) are left out
    order = sorted((k for k in kern if "avg_us" in kern[k] and
                    (_plain_name(k).startswith("jsa::") or "copyBuffer" in k)),
                   key=lambda k: -kern[k]["avg_us"] * kern[k]["calls"])
    cmds = "\n".join(f"rocprofv3 {PROFILE_PASSES[s]} -d DIR/{s} -o run --output-format csv -- "
                     f"{command}" for s in rec["passes"])
    lines = [f"# rocprofv3 kernel statistics and HBM traffic: {workload} fwd+bwd step", "",
             "Commands (repository root, MI355X, ROCm 7.2; DIR = an output directory), then "
             f"`python bench.py --summarize-profiles DIR --workload {workload} "
             f"--copy-calibration-mb {calib_mb}`:", "", "```", cmds, "```", "",
             f"Source hash of the profiled kernels: `{rec['source_sha16']}`.  HBM bytes = "
             f"FETCH_SIZE x {rec['fetch_scale']} + WRITE_SIZE x {rec['write_scale']} (KiB per "
             f"dispatch; scales from the {calib_mb} MiB copy calibration).  Library kernels "
             "only: the torch kernels of the bench's consumed-length analysis (outside the "
             "timed regions) are left out, so % GPU time is of the whole profiled run.", "",
             "| kernel | calls | avg us | min us | max us | % GPU time | HBM MB / dispatch | GB/s |",
             "|---|---:|---:|---:|---:|---:|---:|---:|"]
    for k in order:
        v = kern[k]
        hb = v.get("hbm_bytes")
        lines.append(f"| `{_short(_plain_name(k))}` | {v['calls']} | {v['avg_us']:.1f} | {v['min_us']:.1f} | "
                     f"{v['max_us']:.1f} | {v['pct']:.2f} | "
                     f"{hb / 1e6:.1f} | {v.get('hbm_GBs', 0):.0f} |" if hb else
                     f"| `{_short(_plain_name(k))}` | {v['calls']} | {v['avg_us']:.1f} | {v['min_us']:.1f} | "
                     f"{v['max_us']:.1f} | {v['pct']:.2f} | - | - |")
    (out_dir / f"{workload}_step_kernels.md").write_text("\n".join(lines) + "\n")

    lines = [f"# SQ, LDS and L2-atomic counters: {workload} fwd+bwd step", "",
             "Same commands as the kernel statistics.  Counters are summed over the 8 XCDs "
             "by rocprofv3 and averaged over dispatches.  cycles = GRBM_GUI_ACTIVE / 8; "
             "VALUBusy % = 100 SQ_ACTIVE_INST_VALU / 256 CUs / cycles; LDS bank-conflict "
             "fraction = SQ_LDS_BANK_CONFLICT / 256 CUs / cycles (share of CU cycles the LDS "
             "is stalled by bank conflicts); per active = SQ_LDS_BANK_CONFLICT / "
             "SQ_LDS_IDX_ACTIVE.", "",
             "| kernel | VALU insts | VALU / wave | VALUBusy % | LDS insts | LDS bank-conflict frac "
             "| conflict / LDS active | L2 atomics |",
             "|---|---:|---:|---:|---:|---:|---:|---:|"]

    def f(x, fmt):
        return format(x, fmt) if x is not None else "-"
    for k in order:
        v = kern[k]
        if "valu_insts" not in v and "lds_bank_conflict_cycles" not in v:
            continue
        vpw = v["valu_insts"] / v["waves"] if v.get("valu_insts") and v.get("waves") else None
        lines.append(f"| `{_short(_plain_name(k))}` | {f(v.get('valu_insts'), '.3g')} | {f(vpw, '.0f')} | "
                     f"{f(v.get('valu_busy'), '.1f')} | {f(v.get('lds_insts'), '.3g')} | "
                     f"{f(v.get('lds_bank_conflict_frac'), '.4f')} | "
                     f"{f(v.get('lds_bank_conflict_per_active'), '.4f')} | "
                     f"{f(v.get('l2_atomics'), '.3g')} |")
    (out_dir / f"{workload}_sq_counters.md").write_text("\n".join(lines) + "\n")
    print(f"wrote profiles/{workload}_counters.json, _step_kernels.md, _sq_counters.md")


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
