#!/usr/bin/env python3
"""Benchmark: 512^2 frames/s from 4 RGBD views (the real-time SAILOR path).

One step = one frame of the whole hot path on synthetic inputs already
resident in HBM: TSDF fusion of 4 x 1024^2 RGBD views into a 256^3 grid,
multi-resolution merge, two-layer tree (32^3 coarse cells of 8^3 voxels),
512^2 target rays, DDA + 48 adaptive samples per ray, per-ray sort, fused
per-view MLP (hidden 256 x 4, 64-channel features from [4,64,256,256] maps)
+ occupancy compositing + Hydra attention over the 4 views + RGB head.
The frame is captured once into a hipGraph and replayed: `warmup` untimed
replays, then `steps` timed ones with a HIP event pair around each;
frames/s = 1 / median (p99 and mean reported beside it).

`value` is the FP32 form's frames/s.  The BF16 form (MFMA bf16, fp32
accumulate) is timed the same way and its PSNR against the FP32 render is
reported (cal_psnr definition).  `roofline` is computed for the dominant
kernel (the fused radiance kernel) from HIP events on its stream;
`cpu_baseline` times the CPU restatement on a bounded sample.

  python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-cpu] [--workload c2]
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = ("512² frames/s from 4 RGBD views; fwd+bwd Mrays/s, BF16 vs FP32 at "
          "ΔPSNR≤0.05")
PEAKS = {"fp32": 157.3, "bf16": 2500.0, "fp16": 2500.0}  # dense MFMA TFLOP/s (MI355X)
HBM_GBS = 8000.0  # MI355X HBM3E peak, GB/s


def progress(msg):
    """One line per finished step on stderr (the JSON result goes to stdout)."""
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def mlp_flops_per_sample_view(F, H):
    """Algorithmic FLOPs of the per-(sample, view) MLP: 2 * sum(in*out)."""
    return 2 * ((63 + F) * H + 3 * H * H)


def time_frames(pipe, steps, warmup, stats=None):
    """Seconds per frame: the frame's hipGraph is replayed `warmup` times
    untimed, then exactly `steps` times between two device synchronisations,
    with a HIP event pair around every replay on the replay stream.  Returns
    the median event time; `stats` (a dict) receives median / p99 / mean of
    the per-frame times and the wall-clock mean of the bracketed loop."""
    import numpy as np
    import torch
    pipe.capture()
    for _ in range(warmup):
        pipe.replay()
    st = torch.cuda.current_stream()
    evs = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True))
           for _ in range(steps)]
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for e0, e1 in evs:
        e0.record(st)
        pipe.replay()
        e1.record(st)
    torch.cuda.synchronize()
    wall = (time.perf_counter() - t0) / steps
    ms = np.array([e0.elapsed_time(e1) for e0, e1 in evs])
    if stats is not None:
        stats.update(median_ms=float(np.median(ms)), p99_ms=float(np.percentile(ms, 99)),
                     mean_ms=float(ms.mean()), wall_mean_ms=wall * 1e3)
    return float(np.median(ms)) / 1e3


def stage_times(pipe, reps=5):
    """Per-stage GPU time (ms) with HIP events on the launch stream."""
    import torch
    st = torch.cuda.current_stream()
    res = {}
    for name, fn in (("build", pipe.build), ("trace", pipe.trace), ("shade", pipe.shade)):
        fn()
        ev = [torch.cuda.Event(enable_timing=True) for _ in range(2)]
        ev[0].record(st)
        for _ in range(reps):
            fn()
        ev[1].record(st)
        torch.cuda.synchronize()
        res[name] = ev[0].elapsed_time(ev[1]) / reps
    return res


def hydra_flops_per_ray(F, V, inner=32):
    """Algorithmic FLOPs of the Hydra encoder (2 stacks: q, k, v, fc, w1, w2
    per view token) and the RGB head (F -> 64 -> 32 -> 3) for one ray."""
    stack = 2 * V * (3 * F * inner + inner * F + F * inner + inner * F)
    return 2 * stack + 2 * (F * 64 + 64 * 32 + 32 * 3)


def frame_roofline(pipe, stages, prec, n_valid):
    """Frame-level roofline: each stage's fraction of its own roofline,
    weighted by its share of the frame's GPU time.
      build  (TSDF + merge + tree)  HBM: 16 B written per voxel + the V maps
             (depth, colour) read once + 4 B per fine voxel read by the merge
             and again by the tree + 32 B per node written
      trace  (rays + DDA + samples, then the per-ray sort)  HBM: 24 B rays +
             4 B valid length + 8 S B samples written per ray, 8 S B read and
             8 S B written by the sort
      sample kernel  MFMA: 2 sum(in * out) per (valid sample, view)
      Hydra + RGB head  vector FP32 (157.3 TFLOP/s): hydra_flops_per_ray per
             valid ray
      the rest of the shade stage (ray listing) counts as time at 0."""
    cfg, mc = pipe.cfg, pipe.mc
    N = cfg.dim ** 3
    R = cfg.target * cfg.target
    nodes = sum(pipe.check()[k] for k in ("n_coarse", "n_fine"))
    b_build = N * 16 + cfg.n_views * cfg.img * cfg.img * 8 + 2 * N * 4 + nodes * 32
    b_trace = R * (24 + 4 + 8 * cfg.S) + R * 16 * cfg.S
    f_mlp = n_valid * cfg.S * cfg.n_views * mlp_flops_per_sample_view(mc.F, mc.H)
    f_hyd = n_valid * hydra_flops_per_ray(mc.F, cfg.n_views)
    rows = {
        "build": (stages["build"], b_build / (stages["build"] / 1e3) / 1e9 / HBM_GBS),
        "trace": (stages["trace"], b_trace / (stages["trace"] / 1e3) / 1e9 / HBM_GBS),
        "sample_kernel": (stages["sample_kernel"],
                          f_mlp / (stages["sample_kernel"] / 1e3) / 1e12 / PEAKS[prec]),
        "hydra_kernel": (stages["hydra_kernel"],
                         f_hyd / (stages["hydra_kernel"] / 1e3) / 1e12 / PEAKS["fp32"]),
        "shade_rest": (max(stages["shade"] - stages["sample_kernel"] - stages["hydra_kernel"],
                           0.0), 0.0),
    }
    total = sum(t for t, _ in rows.values())
    frac = sum(t / total * f for t, f in rows.values())
    return {"frac": frac, "stage_ms_and_frac": {k: [round(t, 4), round(f, 4)]
                                                 for k, (t, f) in rows.items()},
            "algorithmic": {"build_bytes": b_build, "trace_bytes": b_trace,
                            "sample_flops": f_mlp, "hydra_flops": f_hyd}}


def kernel_ms(pipe, stage, reps=10):
    """Average duration (ms) of one launch of a sailor_render_rays stage
    (2 = the fused sample kernel, 4 = Hydra/RGB), from HIP events recorded on
    the stream the kernel is launched on (torch's current stream)."""
    import torch
    st = torch.cuda.current_stream()
    # stage 1 lists the rays in an arbitrary order, stage 2 stores per-view
    # features by list position and stage 4 reads them: time a stage only on
    # top of the earlier stages of the same list
    pipe.shade(1)
    pipe.shade(2)
    ev = [torch.cuda.Event(enable_timing=True) for _ in range(2)]
    ev[0].record(st)
    for _ in range(reps):
        pipe.shade(stage)
    ev[1].record(st)
    torch.cuda.synchronize()
    pipe.shade(7)
    torch.cuda.synchronize()
    return ev[0].elapsed_time(ev[1]) / reps


TRAFFIC_FILE = os.path.join(ROOT, "profiles", "render_kernel_traffic_c2.md")


def measured_traffic(prec):
    """HBM bytes per launch of the sample kernel from the committed rocprofv3
    FETCH_SIZE / WRITE_SIZE summary (lines `traffic_bytes <prec>: <bytes>`),
    or None when no summary is present."""
    import re
    try:
        text = open(TRAFFIC_FILE).read()
    except OSError:
        return None
    m = re.search(r"traffic_bytes %s:\s*([0-9.eE+]+)" % prec, text)
    return float(m.group(1)) if m else None


def profile_kernels(cfg):
    """Short fixed sequence for rocprofv3 counter passes: a 256 MiB device copy
    (the known-size calibration kernel for FETCH_SIZE / WRITE_SIZE), then
    three launches of the sample kernel per precision."""
    import torch
    from sailor_amd.pipeline import FramePipeline
    n = 64 * 1024 * 1024
    src = torch.ones(n, dtype=torch.float32, device="cuda")
    dst = torch.empty_like(src)
    dst.copy_(src)
    torch.cuda.synchronize()
    for prec in ("fp32", "bf16"):
        p = FramePipeline.synthetic(cfg, prec=prec)
        p.run()
        torch.cuda.synchronize()
        print(f"profile {prec}: valid rays {p.check()['valid_rays']}", flush=True)
        p.shade(1)
        for _ in range(3):
            p.shade(2)
        torch.cuda.synchronize()
        del p


def precision_form(cfg, steps, warmup, n_err_rays=512, n_err_voxels=200000):
    """Config-3 block: the C2 scene rendered at 1024^2 (1.05 M rays, 48
    samples).  Forms: FP32 plain; FP32 with Kahan-compensated TSDF and
    compositing sums; BF16 and FP16 (feature maps quantised per channel with
    power-of-two scales from the channel's max |x|, folded into the first
    layer; MLP on MFMA with FP32 accumulation; compensated sums).  Per form:
    frames/s, PSNR against the FP32 render, and max-abs / RMS error against
    FP64 of the per-sample occupancy (sigma), RGB and depth on n_err_rays
    seeded valid rays (float64 restatement oracle/nerf_ref.py) and of the
    TSDF on n_err_voxels seeded observed voxels (float64 restatement of the
    fusion)."""
    import numpy as np
    import torch
    from oracle import nerf_ref
    from oracle import oracle as orc
    from sailor_amd.pipeline import FramePipeline
    from sailor_amd.render import render_rays
    out = {"workload": f"{cfg.name}: C2 scene, {cfg.target}^2 render, {cfg.S} samples/ray",
           "error_sample": f"{n_err_rays} seeded valid rays (sigma, rgb, depth), "
                           f"{n_err_voxels} seeded observed voxels (tsdf), against float64"}
    forms = (("fp32", dict(prec="fp32")),
             ("fp32_compensated", dict(prec="fp32", compensated=True)),
             ("bf16", dict(prec="bf16", compensated=True)),
             ("fp16", dict(prec="fp16", compensated=True)))
    ref = mask = sub = ref64 = vox = t64 = None
    errs = {}
    for name, kw in forms:
        p = FramePipeline.synthetic(cfg, **kw)
        p.run()
        torch.cuda.synchronize()
        sec = time_frames(p, steps, warmup)
        rgb = p.rgb[0].cpu().numpy()
        if ref is None:
            ref = rgb
            mask = (p.sorted[0, :, 0] >= 0).cpu().numpy()
            out["valid_rays"] = int(mask.sum())
            valid = np.nonzero(mask)[0]
            sub = np.sort(np.random.default_rng(33).choice(valid, n_err_rays, replace=False))
            ref64 = nerf_ref.render_rays(p.weights, p.sorted[:, sub].cpu().numpy(),
                                         p.rays[:, sub].cpu().numpy(), p.K.cpu().numpy(),
                                         p.RT.cpu().numpy(), p.fmaps_np[None],
                                         p.bbox.cpu().numpy(), p.zdir.cpu().numpy(),
                                         (cfg.img, cfg.img))
            seen = np.nonzero(p.weight[0].cpu().numpy().reshape(-1) > 0)[0]
            vox = np.sort(np.random.default_rng(34).choice(seen, min(n_err_voxels, seen.size),
                                                           replace=False))
            sc = p.scene
            t64, _ = orc.integrate_f64(sc.K()[None], sc.RT()[None],
                                       None, sc.depth[None], p.bbox.cpu().numpy(),
                                       p.vs.cpu().numpy().reshape(-1), (cfg.dim,) * 3,
                                       voxels=vox)
        out[name] = {"frames_per_s": 1.0 / sec, "ms_per_step": sec * 1e3}
        progress(f"c3 {name}: {sec * 1e3:.2f} ms per frame")
        if name != "fp32":
            out[name]["psnr_vs_fp32_db"] = psnr(ref, rgb, mask)
            out[name]["speedup_vs_fp32"] = out["fp32"]["ms_per_step"] / (sec * 1e3)
        idx = torch.from_numpy(sub).cuda()
        maps = p.fmap_lo if p.half_maps else p.fmap
        o = render_rays(p.sorted[:, idx].contiguous(), p.rays[:, idx].contiguous(), p.K, p.RT,
                        (cfg.img, cfg.img), maps, p.bbox, p.zdir, p.model, want_occ=True,
                        compensated=p.compensated,
                        feature_scales=p.fscale if p.quantize else None)
        torch.cuda.synchronize()
        tsdf = p.tsdf[0].cpu().numpy().reshape(-1)[vox].astype(np.float64)
        row = {}
        for key, got, want in (("sigma", o["occ"][0], ref64[4][0]), ("rgb", o["rgb"][0], ref64[0][0]),
                               ("depth", o["depth"][0], ref64[1][0]),
                               ("tsdf", tsdf, t64[0])):
            g = got.double().cpu().numpy() if torch.is_tensor(got) else got
            d = np.abs(g - want.numpy() if torch.is_tensor(want) else g - want)
            row[key] = {"max_abs": float(d.max()), "rms": float(np.sqrt((d ** 2).mean()))}
        errs[name] = row
        del p
        torch.cuda.empty_cache()
    out["error_vs_fp64"] = errs
    return out


def _median_ms(fn, steps):
    """Median over `steps` calls of fn's GPU time (HIP events on the current
    stream, which the library's launches use)."""
    import numpy as np
    import torch
    st = torch.cuda.current_stream()
    ts = []
    for _ in range(steps):
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record(st)
        fn()
        e1.record(st)
        e1.synchronize()
        ts.append(e0.elapsed_time(e1))
    return float(np.median(ts))


def measured_peaks(steps=20):
    """What this box reaches against the datasheet peaks: HBM by a 2 GiB
    device-to-device copy (read + write bytes over the copy's time), MFMA by
    square 8192 GEMMs in bf16, fp16 and fp32 (torch.matmul on contiguous 2-D
    operands, i.e. the vendor GEMM library).  Reported beside the datasheet
    fractions, never in their place."""
    import torch
    n = (2 << 30) // 4
    src = torch.empty(n, dtype=torch.float32, device="cuda").uniform_()
    dst = torch.empty_like(src)
    for _ in range(3):
        dst.copy_(src)
    ms = _median_ms(lambda: dst.copy_(src), steps)
    hbm = 2.0 * n * 4 / (ms / 1e3) / 1e9
    del src, dst
    out = {"hbm_copy_GB_per_s": hbm, "hbm_datasheet_GB_per_s": HBM_GBS,
           "hbm_copy_frac_of_datasheet": hbm / HBM_GBS}
    m = 8192
    for name, dt in (("bf16", torch.bfloat16), ("fp16", torch.float16), ("fp32", torch.float32)):
        a = torch.randn(m, m, device="cuda").to(dt).contiguous()
        b = torch.randn(m, m, device="cuda").to(dt).contiguous()
        for _ in range(3):
            torch.matmul(a, b)
        ms = _median_ms(lambda: torch.matmul(a, b), max(5, steps // 2))
        tf = 2.0 * m ** 3 / (ms / 1e3) / 1e12
        out[f"{name}_gemm_TFLOP_per_s"] = tf
        out[f"{name}_gemm_frac_of_datasheet"] = tf / PEAKS[name]
        del a, b
    torch.cuda.synchronize()
    return out


def training_form(steps, warmup, with_cpu=True):
    """Config-4 block (differentiation form): 8 target views x 256^2 rays of
    the config-1 scene and model (4 x 1024^2 source views, 64 samples/ray).
    One step = fused forward (occupancies kept) + MSE gradient against an
    FP32 render made with perturbed weights + backward (Hydra/RGB head,
    compositing reverse scan, per-view MLP with per-ray recompute, weight-
    gradient GEMMs).  Mrays/s = all 524,288 rays / median(fwd + bwd)."""
    import numpy as np
    import torch
    from sailor_amd.model import NerfWeights
    from sailor_amd.pipeline import CONFIGS
    from sailor_amd.train import RenderTrainer, TrainingBatch, mse_grad
    cfg = CONFIGS["c4"]
    tb = TrainingBatch(cfg, 8)
    mc = tb.pipe.mc
    w = NerfWeights.init(mc.F, mc.H)
    w_tgt = NerfWeights.init(mc.F, mc.H, perturb_std=0.01)
    target = RenderTrainer(w_tgt, "fp32").forward(*tb.inputs())["rgb"].clone()
    out = {"workload": (f"{cfg.name}: 8 target views x {cfg.target}^2 rays, {cfg.n_views} x "
                        f"{cfg.img}^2 source views, {cfg.dim}^3 grid, {cfg.S} samples/ray, "
                        f"MLP {mc.H}x4, F={mc.F}; MSE to an FP32 render with weights + N(0,0.01)"),
           "rays": tb.n_rays, "valid_rays": tb.n_valid}
    for prec in ("fp32", "bf16"):
        tr = RenderTrainer(w, prec)

        def fwd():
            return tr.forward(*tb.inputs())

        def step():
            o = fwd()
            return tr.backward(mse_grad(o["rgb"], target, tb.n_valid))
        for _ in range(warmup):
            step()
        torch.cuda.synchronize()
        f_ms = _median_ms(fwd, steps)
        s_ms = _median_ms(step, steps)
        out[prec] = {"mrays_per_s": tb.n_rays / (s_ms / 1e3) / 1e6, "fwd_ms": f_ms,
                     "fwd_bwd_ms": s_ms, "fwd_bwd_over_fwd": s_ms / f_ms}
        progress(f"c4 {prec}: fwd {f_ms:.2f} ms, fwd+bwd {s_ms:.2f} ms")
        del tr
    if with_cpu:
        from oracle import nerf_ref
        valid = torch.nonzero(tb.t[0, :, 0] >= 0).flatten().cpu().numpy()
        sub = valid[:: max(1, valid.size // 2048)][:2048]
        p = tb.pipe
        P = nerf_ref.params_of(w, dtype=torch.float32, requires_grad=True)
        args = (tb.t[:, sub].cpu().numpy(), tb.rays[:, sub].cpu().numpy(), p.K.cpu().numpy(),
                p.RT.cpu().numpy(), p.fmaps_np[None], p.bbox.cpu().numpy(),
                p.zdir.cpu().numpy(), tb.img_hw)
        tgt = target[:, sub].cpu()
        t0 = time.perf_counter()
        rgb = nerf_ref.render_rays(w, *args, dtype=torch.float32, params=P)[0]
        ((rgb - tgt) ** 2).mean().backward()
        dt = time.perf_counter() - t0
        t_all = dt * tb.n_valid / len(sub)
        out["cpu_baseline"] = {
            "value": tb.n_rays / t_all / 1e6, "unit": "Mrays/s (fwd+bwd)",
            "cores": torch.get_num_threads(), "kind": "port",
            "sample": (f"PyTorch-CPU fp32 autograd of the restatement on {len(sub)} of "
                       f"{tb.n_valid} valid rays ({dt:.2f} s), extrapolated per valid ray")}
    return out


def fastnerf_infer(steps):
    """InferNerf.infer_nerf (FastNerf inference surface, packed half
    parameters) on one GUI-sized call: 65,536 rays x 48 samples x 4 geometry
    and 4 texture views, synthetic inputs.  Algorithmic FLOPs: per (view,
    sample) 2*(18*64 + 3*64^2) feature MLP + 2*(18*16 + 16*8) geometry MLP;
    per sample 2*(2*64^2 + 64) occupancy MLP; per (ray, view) two Hydra stacks
    2*2*(3*30*32 + 32*30 + 30*32 + 32*30); per ray 2*(30*64 + 64*32 + 32*3)."""
    import torch
    from sailor_amd import InferNerf
    from sailor_amd.model import fastnerf_params
    B, R, S, Vg, Vt = 1, 65536, 48, 4, 4
    pd, pc, _ = fastnerf_params()
    dev = "cuda"
    t = torch.sort(torch.rand((B, R, S), device=dev) * 2 + 1, -1)[0]
    n_g, n_t = B * Vg * R * S, B * Vt * R * S
    z = torch.rand((n_g, 1), device=dev) * 2 - 1
    ps = torch.rand((n_g, 1), device=dev) * 2 - 1
    geo = torch.randn((n_g, 16), device=dev)
    rgb = torch.rand((n_t, 3), device=dev)
    frgb = torch.randn((n_t, 16), device=dev)
    dirs = torch.randn((n_t, 3), device=dev)
    pdt = torch.from_numpy(pd).to(dev)
    pct = torch.from_numpy(pc).to(dev)

    flops = (Vg * R * S * 2 * (18 * 64 + 3 * 64 * 64) + Vt * R * S * 2 * (18 * 16 + 16 * 8)
             + R * S * 2 * (2 * 64 * 64 + 64) + Vt * R * 2 * 2 * (3 * 30 * 32 + 3 * 32 * 30)
             + R * 2 * (30 * 64 + 64 * 32 + 32 * 3))
    out = {"workload": f"infer_nerf: {R} rays x {S} samples, {Vg} geometry / {Vt} texture views",
           "flops": flops}
    for prec in ("fp16", "fp32"):
        def go():
            return InferNerf.infer_nerf(t, z, ps, geo, rgb, frgb, dirs, pdt, pct, B, R, S, Vg,
                                        Vt, 0, precision=prec)
        go()
        torch.cuda.synchronize()
        ms = _median_ms(go, steps)
        out[prec] = {"ms": ms, "rays_per_s": R / (ms / 1e3),
                     "tflops": flops / (ms / 1e3) / 1e12}
    # density pass alone (dominant: ~97 % of the FLOPs), fp16 MFMA form
    pd16 = pdt.half().contiguous()

    def dens():
        InferNerf.infer_density(z, ps, geo, pd16, B, R * S, Vg, 0)
    dens()
    torch.cuda.synchronize()
    dms = _median_ms(dens, steps)
    dflops = Vg * R * S * 2 * (18 * 64 + 3 * 64 * 64) + R * S * 2 * (2 * 64 * 64 + 64)
    out["density_fp16"] = {"ms": dms, "flops": dflops,
                           "tflops": dflops / (dms / 1e3) / 1e12,
                           "frac_of_fp16_mfma_peak": dflops / (dms / 1e3) / 1e12 / PEAKS["fp16"]}
    return out


def fastnerf_upsampling(steps):
    """The GUI's sub-pixel upsampling (fastnerf.forward_upsampling) on one
    512^2 target: 262,144 rays, ~60% on the surface (the rest outside the
    depth range), 4 sub-pixel rays each blended from 2 of 4 views with
    1024^2 x 16 colour features, 1024^2 colours and 512^2 depths; synthetic
    inputs.  Algorithmic FLOPs per upsampled sub-pixel ray: the blending MLP
    2*(40*64 + 64*64 + 64*32 + 62*64 + 64*32 + 32*3) = 29,632."""
    import numpy as np
    import torch
    from sailor_amd import scenes
    from sailor_amd.fastnerf import forward_upsampling
    from sailor_amd.model import fastnerf_params
    dev = "cuda"
    side, V = 512, 4
    R = side * side
    g = torch.Generator(device=dev).manual_seed(21)
    cams = scenes.ring_cameras(V, 600.0, 512, 512)
    tgt = scenes.target_camera(512, 512, 600.0)
    f = lambda a: torch.from_numpy(np.ascontiguousarray(a, np.float32)).to(dev)
    ks, rs = f(np.stack([c.K for c in cams])), f(np.stack([c.R for c in cams]))
    ts = f(np.stack([c.T for c in cams]))
    yy, xx = np.meshgrid(np.arange(side) + 0.5, np.arange(side) + 0.5, indexing="ij")
    h = np.stack([xx.ravel(), yy.ravel(), np.ones(R)], 1)
    d = (np.linalg.inv(tgt.K.astype(np.float64)) @ h.T).T @ tgt.R.astype(np.float64)
    d /= np.linalg.norm(d, axis=-1, keepdims=True)
    o = -tgt.R.T.astype(np.float64) @ tgt.T.astype(np.float64)
    ray_ori = f(np.broadcast_to(o, (R, 3)))[None]
    ray_dir = f(d)[None]
    on = torch.rand((1, R, 1), device=dev, generator=g) < 0.6
    o_d = torch.where(on, 2.0 + torch.rand((1, R, 1), device=dev, generator=g),
                      torch.zeros((1, R, 1), device=dev))
    o_c = torch.rand((1, R, 3), device=dev, generator=g)
    o_f = torch.randn((1, R, 30), device=dev, generator=g)
    rgbf = torch.randn((V, 1024, 1024, 16), device=dev, generator=g)
    rgbs = torch.rand((V, 1024, 1024, 3), device=dev, generator=g)
    deps = 2.0 + torch.rand((V, 512, 512), device=dev, generator=g)
    pb = torch.from_numpy(fastnerf_params()[2]).to(dev)

    n_up = int(on.sum()) * 4
    flops = n_up * 2 * (40 * 64 + 64 * 64 + 64 * 32 + 62 * 64 + 64 * 32 + 32 * 3)
    out = {"workload": f"forward_upsampling: {side}^2 rays ({n_up // 4} on the surface) x 4 "
                       "sub-pixels, 2 of 4 views with 1024^2 features", "flops": flops}
    tK, tR = f(tgt.K), f(tgt.R)
    for prec in ("fp16", "fp32"):
        def go():
            return forward_upsampling(ray_ori, ray_dir, o_d, o_c, o_f, tK, tR, ks, rs, ts, rgbf,
                                      rgbs, deps, (0, 1), pb, precision=prec)
        go()
        torch.cuda.synchronize()
        ms = _median_ms(go, steps)
        out[prec] = {"ms": ms, "subpixel_rays_per_s": n_up / (ms / 1e3),
                     "tflops": flops / (ms / 1e3) / 1e12}
    return out


def training_kernels(steps, C=8):
    """Training-only octree kernels on the C2 scene (256^3 tree, 512^2 target
    rays = 262,144): recursive octree_traversal, adaptive
    ray_voxels_points_sampling (48 samples), trilinear_aggregate_features
    forward and backward with C=8 channels.  Trilinear roofline: algorithmic
    bytes per (valid sample, channel) 32 B corner values + 4 B output (fwd),
    4 B gradient + 32 B corner gradients (bwd), plus 8 B (index, depth) per
    sample; HBM-bound."""
    import torch
    from sailor_amd._lib import call, ptr, stream_of
    from sailor_amd.pipeline import CONFIGS, FramePipeline
    cfg = CONFIGS["c2"]
    p = FramePipeline.synthetic(cfg, prec="fp32")
    p.build()
    p.trace()
    torch.cuda.synchronize()
    dev = p.dev
    st = stream_of(dev)
    R, S = cfg.target * cfg.target, cfg.S
    ori, dirs = p.rays[..., :3].contiguous(), p.rays[..., 3:].contiguous()
    nc, nf = 3 * p.dc, 512
    o = dict(device=dev)
    idx_c = torch.empty((1, R, nc), dtype=torch.int32, **o)
    t0c, t1c = torch.empty((1, R, nc), **o), torch.empty((1, R, nc), **o)
    idx = torch.empty((1, R, nf), dtype=torch.int32, **o)
    t0, t1 = torch.empty((1, R, nf), **o), torch.empty((1, R, nf), **o)
    vl = torch.empty((1, R), **o)
    status = torch.zeros(3, dtype=torch.int32, **o)

    def trav():
        call("sailor_octree_traversal", ptr(p.occ1), ptr(p.node_pos), ptr(p.node_info), p.cap,
             ptr(p.bbox), ptr(p.vs), 1, cfg.dim, p.dc, cfg.rate, ptr(ori), ptr(dirs), R, nc, nf,
             -1, -1, ptr(idx_c), ptr(t0c), ptr(t1c), ptr(idx), ptr(t0), ptr(t1), ptr(vl),
             ptr(status), st)
    trav()
    torch.cuda.synchronize()
    stat = status.tolist()
    if stat[2]:
        raise RuntimeError("training_kernels: traversal lists overflowed")
    noise = torch.rand((1, R, S), **o)
    sidx = torch.empty((1, R, S), dtype=torch.int32, **o)
    stt = torch.empty((1, R, S), **o)

    def samp():
        call("sailor_octree_sampling", ptr(idx), ptr(t0), ptr(t1), ptr(noise), ptr(vl),
             ptr(p.node_pos), p.cap, ptr(p.vs), 1, R, nf, stat[1], S, 2, ptr(sidx), ptr(stt),
             st)
    samp()
    feats = torch.randn((1, C, R, S, 8), **o)
    out = torch.empty((1, C, R, S), **o)
    grad = torch.randn((1, C, R, S), **o)
    gfeats = torch.empty_like(feats)

    def fwd():
        call("sailor_trilinear_aggregate", ptr(p.node_pos), p.cap, ptr(ori), ptr(dirs),
             ptr(sidx), ptr(stt), ptr(feats), 1, R, S, C, ptr(out), st)

    def bwd():
        call("sailor_trilinear_aggregate_backward", ptr(p.node_pos), p.cap, ptr(ori),
             ptr(dirs), ptr(sidx), ptr(stt), ptr(grad), 1, R, S, C, ptr(gfeats), st)
    res = {"workload": f"c2 tree, {R} rays, {S} samples, C={C}"}
    res["octree_traversal_ms"] = _median_ms(trav, steps)
    res["max_hits_coarse_fine"] = stat[:2]
    res["adaptive_sampling_ms"] = _median_ms(samp, steps)
    n_valid = int((sidx >= 0).sum())
    res["valid_samples"] = n_valid
    for name, fn in (("trilinear_fwd", fwd), ("trilinear_bwd", bwd)):
        ms = _median_ms(fn, steps)
        # valid samples read their C x 8 corner values (fwd) or C gradients
        # (bwd); every slot is written (zeros for dropped samples) and reads
        # its index and depth
        if name == "trilinear_fwd":
            nbytes = n_valid * C * 32 + R * S * C * 4 + R * S * 8
        else:
            nbytes = n_valid * C * 4 + R * S * C * 32 + R * S * 8
        gbs = nbytes / (ms / 1e3) / 1e9
        res[name] = {"ms": ms, "algorithmic_bytes": nbytes, "GB_per_s": gbs,
                     "frac_of_hbm_peak": gbs / HBM_GBS}
    return res


def gui_frame(steps):
    """The GUI's per-frame render on the FastNerf path (gui/gui_render.py:
    921-934) at its own settings (4 x 512^2 source views, 128^3 grid / 4,
    512^2 target, 32 samples, upsampled 4 sub-pixels per ray): TSDF + tree,
    trace + sort, per-view features + infer_nerf, upsampling, as one frame;
    the valid-ray compaction syncs the host as in the reference."""
    import torch
    from sailor_amd.gui import GuiFrame
    out = {"workload": "gui frame: 4 x 512^2 views, 128^3 grid / 4, 512^2 target, 32 samples, "
                       "x4 sub-pixel upsampling, FastNerf (16-ch 128^2 features, 16-ch 1024^2 "
                       "high-resolution features)"}
    for prec in ("fp16", "fp32"):
        g = GuiFrame(precision=prec)
        g.render()
        torch.cuda.synchronize()
        ms = _median_ms(g.render, steps)
        out[prec] = {"ms_per_frame": ms, "frames_per_s": 1e3 / ms}
        out["valid_rays"] = int(torch.all(g.pipe.sorted[0] != -1, dim=-1).sum())
        del g
    return out


def standalone_kernels(steps):
    """The stand-alone HBM-bound kernels of the path at their configuration
    sizes, each against the 8 TB/s HBM roofline with the per-unit
    algorithmic bytes of the path's measurement table:
      volume_rendering_occ_forward   C2: 262,144 rays x 48 samples, RGB;
                                     20 B per sample read, 24 B per ray written
      volume_rendering_training_forward / _backward  C4: 524,288 rays x 48,
                                     RGB; fwd 24 B per sample (t, sort index,
                                     sigma, ...) + 24 B per ray; bwd fwd bytes
                                     + 16 B per sample written + 24 B per ray
      freq_encode_forward            4,194,304 points, degree 10: 12 B read +
                                     252 B written per point
      depth_normalize                C2: 4 x 1024^2 depths: 4 B read (twice:
                                     min/max, then normalise) + 8 B written
                                     per pixel
    Inputs are seeded synthetic tensors resident in HBM."""
    import torch
    from sailor_amd import VoxelEncoding as VE
    dev = "cuda"
    g = torch.Generator(device=dev).manual_seed(9)
    out = {}

    def rec(name, fn, nbytes, unit):
        fn()
        torch.cuda.synchronize()
        ms = _median_ms(fn, steps)
        gbs = nbytes / (ms / 1e3) / 1e9
        out[name] = {"ms": ms, "algorithmic_bytes": nbytes, "GB_per_s": gbs,
                     "frac_of_hbm_peak": gbs / HBM_GBS, "unit": unit}

    R, S = 262144, 48
    t = torch.sort(2.0 + torch.rand((1, R, S), device=dev, generator=g), -1)[0]
    occ = torch.rand((1, R, S, 1), device=dev, generator=g)
    rgb = torch.rand((1, R, S, 3), device=dev, generator=g)
    rec("volume_rendering_occ_forward", lambda: VE.volume_rendering_occ_forward(t, occ, rgb, 0, 0.0),
        R * S * 20 + R * 24, "C2 rays x 48 samples")
    R4 = 524288
    t4 = torch.sort(2.0 + torch.rand((1, R4, S), device=dev, generator=g), -1)[0]
    si = torch.arange(S, device=dev, dtype=torch.int32).expand(1, R4, S).contiguous()
    sig = torch.rand((1, R4, S, 1), device=dev, generator=g)
    f4 = torch.rand((1, R4, S, 3), device=dev, generator=g)
    fo = VE.volume_rendering_training_forward(t4, si, sig, f4, 0, 0.0, True)
    grads = [torch.rand_like(x) for x in fo]
    rec("volume_rendering_training_forward",
        lambda: VE.volume_rendering_training_forward(t4, si, sig, f4, 0, 0.0, True),
        R4 * S * 24 + R4 * 24, "C4 rays x 48 samples")
    rec("volume_rendering_training_backward",
        lambda: VE.volume_rendering_training_backward(*grads, *fo, t4, si, sig, f4, 0, 0.0, True),
        R4 * S * 40 + R4 * 48, "C4 rays x 48 samples")
    N = 4194304
    x = torch.rand((N, 3), device=dev, generator=g) * 2 - 1
    rec("freq_encode_forward", lambda: VE.freq_encode_forward(x, 10, 0), N * (12 + 252),
        "points")
    d = 1.5 + torch.rand((1, 4, 1024, 1024), device=dev, generator=g) * 2
    rec("depth_normalize", lambda: VE.depth_normalize(d, 0.2, 4.0, -1, False, True, 0),
        4 * 1024 * 1024 * (4 + 4 + 8), "pixels")
    return out


def batching_form(steps, n_frames=16, precs=("fp32", "bf16")):
    """Config-5 block: 16 frames of uneven size (own seeded scene each,
    performer scale U(0.6, 1.2), 2-6 source views of 1024^2, 256^3 grid,
    512^2 render, 48 samples) replayed as one hipGraph (frames on 1 or 4
    streams), each frame's chunk size picked by the autotuner; against the
    same frames as 16 sequential single-frame graph launches (per-frame
    latency p50/p99).  fp32 (plain form) and bf16."""
    import numpy as np
    import torch
    from sailor_amd.pipeline import CONFIGS, BatchPipeline
    base = CONFIGS["c2"]
    out = {"workload": (f"c5: {n_frames} frames, 2-6 views of {base.img}^2, scale U(0.6,1.2), "
                        f"{base.dim}^3 grid, {base.target}^2 render, {base.S} samples")}
    for prec in precs:
        t0 = time.perf_counter()
        bp = BatchPipeline(base, n_frames=n_frames, prec=prec)
        setup_s = time.perf_counter() - t0
        progress(f"c5 {prec}: {n_frames} frames set up in {setup_s:.1f} s")
        tuned = bp.autotune()
        # sequential baseline: every frame alone, with its tuned chunk and the
        # whole chip for its sample kernel
        lat = []
        for p, ns in zip(bp.pipes, tuned["chunk_samples"]):
            keep = p.chunk_samples
            p.chunk_samples = ns
            p.capture()
            lat.append(_median_ms(p.replay, steps))
            p.graph = None
            p.chunk_samples = keep
        progress(f"c5 {prec}: tuned {tuned['grid_share']}% of the chip per sample kernel, "
                 f"{tuned['n_streams']} streams")
        batched = {}
        for ns in sorted({1, 4, tuned["n_streams"]}):
            bp.capture(n_streams=ns)
            bp.replay()
            torch.cuda.synchronize()
            ms = _median_ms(bp.replay, steps)
            batched[f"{ns}_streams"] = {"frames_per_s": n_frames / (ms / 1e3), "ms_per_batch": ms}
            bp.graph = None
        chunks = tuned
        progress(f"c5 {prec}: batched {batched}")
        info = [p.check() for p in bp.pipes]
        if "views" not in out:
            out["views"] = [c.n_views for c in bp.cfgs]
            out["scale"] = [round(c.scale, 3) for c in bp.cfgs]
            out["occupied_voxels"] = [i["occupied"] for i in info]
            out["valid_rays"] = [i["valid_rays"] for i in info]
        best = max(batched.values(), key=lambda r: r["frames_per_s"])
        out[prec] = {
            "chunk_samples": chunks, "batched": batched,
            "batched_frames_per_s": best["frames_per_s"],
            "sequential": {"frames_per_s": n_frames / (sum(lat) / 1e3),
                           "latency_ms_p50": float(np.percentile(lat, 50)),
                           "latency_ms_p99": float(np.percentile(lat, 99)),
                           "latency_ms": [round(x, 3) for x in lat]},
            "setup_s": setup_s}
        del bp
        torch.cuda.empty_cache()
    return out


def psnr(a, b, mask):
    import numpy as np
    m = mask[..., None].astype(np.float64)
    mse = ((np.abs(a * 255 * m - b * 255 * m)) ** 2).sum() / max(m.sum(), 1)
    return float(10 * np.log10(255 * 255 / max(mse, 1e-30)))


def cpu_baseline(pipe, n_rays_sample=1024):
    """CPU restatement timed on the host: oracle TSDF/tree/traversal/sampling
    in C (OpenMP) on the full frame, PyTorch-CPU MLP + compositing on a fixed
    subset of valid rays, extrapolated to all valid rays."""
    import numpy as np
    import torch
    from oracle import nerf_ref
    from oracle import oracle as orc
    from sailor_amd import scenes
    cfg = pipe.cfg
    sc = pipe.scene
    K = sc.K()[None]
    RT = sc.RT()[None]
    R = np.ascontiguousarray(RT[..., :3])
    T = np.ascontiguousarray(RT[..., 3])
    colors = scenes.pack_colors(sc.rgb)[None]
    depths = sc.depth[None]
    t0 = time.perf_counter()
    o = orc.integrate(K, RT, R, T, colors, depths, (cfg.dim,) * 3,
                      fixed_bbox=scenes.fixed_cube(cfg.cube)[None])
    t_int = time.perf_counter() - t0
    t0 = time.perf_counter()
    oc, ocnt = orc.merge(o["occ"], cfg.rate)
    nf = int(o["num_occ"].sum())
    tr = orc.tree_build(o["occ"], oc, o["voxel_size"], o["vol_bbox"], nf, nf + int(ocnt.sum()))
    t_tree = time.perf_counter() - t0
    cam = pipe.target_cam
    t0 = time.perf_counter()
    r6 = orc.rays(cam.K[None, None], cam.R[None, None], cam.T[None, None], cfg.target,
                  cfg.target)[0, 0].reshape(1, -1, 6)
    trv = orc.traverse(tr["vol_coarse"], tr["node_pos"], tr["node_info"], o["vol_bbox"],
                       o["voxel_size"], cfg.dim, cfg.rate, r6[..., :3], r6[..., 3:], 1024)
    noise = np.full((1, r6.shape[1], cfg.S), 0.5, np.float32)
    _, st = orc.sampling(trv["idx"], trv["tmin"], trv["tmax"], noise, trv["valid_len"],
                         tr["node_pos"], o["voxel_size"], cfg.S, cfg.rate, trv["gmax"])
    srt, _ = orc.sort_samplings(st)
    t_trace = time.perf_counter() - t0
    valid = np.nonzero(srt[0, :, 0] >= 0)[0]
    sub = valid[:n_rays_sample]
    t0 = time.perf_counter()
    with torch.no_grad():
        nerf_ref.render_rays(pipe.weights, srt[:, sub], r6[:, sub], K, RT, pipe.fmaps_np[None],
                             o["vol_bbox"], cam.R[2][None], (cfg.img, cfg.img),
                             dtype=torch.float32)
    t_mlp_sub = time.perf_counter() - t0
    t_mlp = t_mlp_sub * len(valid) / max(len(sub), 1)
    total = t_int + t_tree + t_trace + t_mlp
    return {
        "value": 1.0 / total, "unit": "frames/s", "cores": torch.get_num_threads(),
        "kind": "port",
        "sample": (f"full frame for TSDF/tree/traversal/sampling (C oracle, OpenMP); "
                   f"PyTorch-CPU fp32 MLP+compositing on {len(sub)} of {len(valid)} valid rays, "
                   f"extrapolated per ray; stage s: integrate {t_int:.2f}, tree {t_tree:.2f}, "
                   f"trace {t_trace:.2f}, mlp {t_mlp:.2f} (sample {t_mlp_sub:.2f})"),
    }


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=100)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--workload", default="c2")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-bf16", action="store_true")
    ap.add_argument("--no-early-exit", action="store_true")
    ap.add_argument("--no-c3", action="store_true",
                    help="skip the 1024^2 precision-form block")
    ap.add_argument("--no-c4", action="store_true",
                    help="skip the differentiation-form (fwd+bwd) block")
    ap.add_argument("--no-c5", action="store_true",
                    help="skip the 16-frame batching block")
    ap.add_argument("--batching-only", action="store_true",
                    help="run only the 16-frame batching block and print it")
    ap.add_argument("--no-infer", action="store_true",
                    help="skip the FastNerf inference-surface block")
    ap.add_argument("--infer-only", action="store_true",
                    help="run only the FastNerf inference-surface block and print it")
    ap.add_argument("--training-only", action="store_true",
                    help="run only the differentiation-form block and print it")
    ap.add_argument("--kernels-only", action="store_true",
                    help="run only the training-only octree kernel block and print it")
    ap.add_argument("--profile-kernels", action="store_true",
                    help="run the fixed launch sequence used for rocprofv3 counter passes")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py: this build runs on one GPU only; multi-GPU figures are not produced "
              "(--gpus must be 1)", file=sys.stderr)
        sys.exit(2)
    import numpy as np
    import torch
    from sailor_amd.pipeline import CONFIGS, FramePipeline

    cfg = CONFIGS[args.workload]
    torch.cuda.set_device(0)
    if args.profile_kernels:
        profile_kernels(cfg)
        return
    if args.batching_only:
        print(json.dumps(batching_form(args.steps)))
        return
    if args.infer_only:
        print(json.dumps(fastnerf_infer(args.steps)))
        print(json.dumps(fastnerf_upsampling(args.steps)))
        print(json.dumps(gui_frame(args.steps)))
        return
    if args.training_only:
        print(json.dumps(training_form(args.steps, args.warmup, with_cpu=not args.no_cpu)))
        return
    if args.kernels_only:
        print(json.dumps(training_kernels(args.steps)))
        print(json.dumps(standalone_kernels(args.steps)))
        return
    p32 = FramePipeline.synthetic(cfg, prec="fp32")
    p32.run()
    torch.cuda.synchronize()
    info = p32.check()
    timing = {}
    sec = time_frames(p32, args.steps, args.warmup, timing)
    progress(f"fp32 frame: median {timing['median_ms']:.2f} ms, p99 {timing['p99_ms']:.2f} ms")
    stages = stage_times(p32)
    # roofline of the dominant kernel, the fused sample kernel (stage 2 of
    # sailor_render_rays): algorithmic FLOPs count evaluated (valid) samples
    # of the per-view MLP only
    mc = p32.mc
    n_valid = info["valid_rays"]
    flops = n_valid * cfg.S * cfg.n_views * mlp_flops_per_sample_view(mc.F, mc.H)
    k32 = kernel_ms(p32, 2)
    stages["sample_kernel"] = k32
    stages["hydra_kernel"] = kernel_ms(p32, 4)
    achieved = flops / (k32 / 1e3) / 1e12
    frame_rl = frame_roofline(p32, stages, "fp32", n_valid)
    result = {
        "metric": METRIC, "value": 1.0 / sec, "unit": "frames/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": sec * 1e3,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "fp32",
        "data": "synthetic (seeded capsule-human RGBD scene, N(0,1) feature maps, Xavier weights)",
        "config": {
            "workload": (f"{cfg.name}: {cfg.n_views} views {cfg.img}^2 RGBD -> {cfg.dim}^3 TSDF "
                         f"({cfg.cube} m cube), tree {cfg.dim // cfg.rate}^3 x {cfg.rate}^3, "
                         f"{cfg.target}^2 render, {cfg.S} samples/ray, MLP {mc.H}x4, "
                         f"Hydra over {cfg.n_views} views, F={mc.F}, feature maps "
                         f"[{cfg.n_views},{mc.F},{mc.Hf},{mc.Wf}]"),
            "occupied_voxels": info["occupied"], "tree_nodes": [info["n_coarse"], info["n_fine"]],
            "valid_rays": n_valid, "stage_ms": stages, "graph": True,
            "frame_timing": timing,
        },
        "roofline": {
            "bound": "mfma", "achieved": achieved, "peak": PEAKS["fp32"], "unit": "TFLOP/s",
            "frac": achieved / PEAKS["fp32"], "traffic": measured_traffic("fp32"),
            "kernel": "render_kernel<64,256,fp32,24> (sailor_render_rays stage 2)",
            "flops_per_launch": flops, "avg_launch_ms": k32,
            "frame": frame_rl,
        },
    }
    rgb32 = p32.rgb[0].cpu().numpy()
    valid_mask = (p32.sorted[0, :, 0] >= 0).cpu().numpy()
    if not args.no_bf16:
        for prec in ("bf16", "fp16"):
            plo = FramePipeline.synthetic(cfg, prec=prec)
            plo.run()
            torch.cuda.synchronize()
            t_lo = {}
            s_lo = time_frames(plo, args.steps, args.warmup, t_lo)
            st_lo = stage_times(plo)
            k_lo = kernel_ms(plo, 2)
            st_lo["sample_kernel"] = k_lo
            st_lo["hydra_kernel"] = kernel_ms(plo, 4)
            rgb_lo = plo.rgb[0].cpu().numpy()
            progress(f"{prec} frame: median {t_lo['median_ms']:.2f} ms")
            result[prec] = {
                "frames_per_s": 1.0 / s_lo, "ms_per_step": s_lo * 1e3,
                "speedup_vs_fp32": sec / s_lo,
                "psnr_vs_fp32_db": psnr(rgb32, rgb_lo, valid_mask),
                "stage_ms": st_lo,
                "sample_kernel_tflops": flops / (k_lo / 1e3) / 1e12,
                "sample_kernel_frac_of_peak": flops / (k_lo / 1e3) / 1e12 / PEAKS[prec],
                "traffic": measured_traffic(prec),
                "frame_timing": t_lo,
                "frame_roofline": frame_roofline(plo, st_lo, prec, n_valid),
            }
            del plo
    if not args.no_early_exit:
        pe = FramePipeline.synthetic(cfg, prec="fp32", early_exit=True)
        pe.run()
        torch.cuda.synchronize()
        result["fp32_early_exit"] = {
            "frames_per_s": 1.0 / time_frames(pe, args.steps, args.warmup),
            "identical_image": bool(np.array_equal(pe.rgb[0].cpu().numpy(), rgb32)),
        }
        del pe
    progress("early-exit frame done")
    if not args.no_c3 and args.workload == "c2":
        result["c3_precision_form"] = precision_form(CONFIGS["c3"], max(3, args.steps // 4),
                                                     args.warmup)
    if not args.no_c4 and args.workload == "c2":
        result["c4_training_form"] = training_form(max(3, args.steps // 4), args.warmup,
                                                   with_cpu=not args.no_cpu)
    if not args.no_c5 and args.workload == "c2":
        result["c5_batching_form"] = batching_form(max(3, args.steps // 4))
    if not args.no_infer and args.workload == "c2":
        result["fastnerf_infer"] = fastnerf_infer(max(3, args.steps // 4))
        result["fastnerf_upsampling"] = fastnerf_upsampling(max(3, args.steps // 4))
        result["gui_frame"] = gui_frame(max(3, args.steps // 4))
        result["standalone_kernels"] = standalone_kernels(max(3, args.steps // 4))
    mp = measured_peaks()
    result["measured_peaks"] = mp
    result["roofline"]["frac_of_measured_peak"] = (
        result["roofline"]["achieved"] / mp["fp32_gemm_TFLOP_per_s"])
    for prec in ("bf16", "fp16"):
        if prec in result:
            result[prec]["sample_kernel_frac_of_measured_gemm"] = (
                result[prec]["sample_kernel_tflops"] / mp[f"{prec}_gemm_TFLOP_per_s"])
    progress("peak probes done")
    if not args.no_cpu:
        progress("cpu baseline")
        result["cpu_baseline"] = cpu_baseline(p32)
    else:
        result["cpu_baseline"] = None
    print(json.dumps(result))


if __name__ == "__main__":
    main()
