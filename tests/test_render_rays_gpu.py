"""GPU parity of the fused render path:
  * sailor_trace_sample (fused rays + DDA + sampling) vs the oracle chain,
    bit for bit;
  * sailor_render_rays (fused gather + PE + MLP on MFMA + compositing + Hydra +
    RGB head) vs the float64 PyTorch restatement oracle/nerf_ref.py;
  * bf16 / fp16 forms vs the fp32 form (PSNR);
  * the whole frame captured in a hipGraph replays to the eager result.
"""
import numpy as np
import pytest
import torch

from oracle import nerf_ref
from oracle import oracle as orc

pytestmark = pytest.mark.gpu


def small_cfg():
    from sailor_amd.pipeline import FrameConfig
    # config-1 geometry and model widths with a reduced target image
    return FrameConfig("c1_small", 4, 512, 600.0, 128, 8, 48, 56.25, 64, "c1")


_pipes = {}


def pipe(prec="fp32", cfg=None):
    from sailor_amd.pipeline import FramePipeline
    cfg = cfg or small_cfg()
    key = (cfg.name, prec)
    if key not in _pipes:
        p = FramePipeline.synthetic(cfg, prec=prec)
        p.run()
        torch.cuda.synchronize()
        _pipes[key] = p
    return _pipes[key]


def test_fused_trace_matches_oracle_chain():
    _check_trace(pipe())


def test_trace_with_coarse_nodes_after_a_gap():
    """Tree built with the coarse nodes numbered from 70,000 on (the build's
    coarse_start; the fine nodes stay at [0, n_fine)): the fused trace on
    that layout equals the oracle chain on the same arrays."""
    from sailor_amd.pipeline import FramePipeline
    p = FramePipeline.synthetic(small_cfg(), prec="fp32", coarse_start=70000)
    p.run()
    torch.cuda.synchronize()
    info = p.node_info.cpu().numpy()
    assert (info[70000:, 0] >= 0).any() and not (info[:70000, 0] >= 0).all()
    _check_trace(p)


def _check_trace(p):
    cfg = p.cfg
    info = p.check()
    assert info["valid_rays"] > 50, info
    # oracle chain on the GPU-built tree (the tree itself is checked bit-exactly
    # in test_geometry_gpu.py)
    vc = p.occ1.cpu().numpy()
    pos = p.node_pos.cpu().numpy()
    nfo = p.node_info.cpu().numpy()
    bbox = p.bbox.cpu().numpy()
    vs = p.vs.cpu().numpy().reshape(-1)
    cam = p.target_cam
    r6 = orc.rays(cam.K[None, None], cam.R[None, None], cam.T[None, None], cfg.target,
                  cfg.target)[0, 0].reshape(1, -1, 6)
    np.testing.assert_array_equal(p.rays.cpu().numpy(), r6)
    trv = orc.traverse(vc, pos, nfo, bbox, vs, cfg.dim, cfg.rate, r6[..., :3], r6[..., 3:], 1024)
    # the trace kernel replays up to 44 hits per ray from LDS and walks the
    # DDA again for longer rays: the scene must exercise both
    hits = (trv["idx"] >= 0).sum(-1)
    assert ((hits > 0) & (hits <= 44)).any() and (hits > 44).any(), np.bincount(hits.ravel())
    noise = np.full((1, r6.shape[1], cfg.S), 0.5, np.float32)
    sidx, st = orc.sampling(trv["idx"], trv["tmin"], trv["tmax"], noise, trv["valid_len"], pos,
                            vs, cfg.S, cfg.rate, trv["gmax"])
    np.testing.assert_array_equal(p.sidx.cpu().numpy(), sidx)
    np.testing.assert_array_equal(p.st.cpu().numpy(), st)
    np.testing.assert_array_equal(p.vl.cpu().numpy(), trv["valid_len"])
    srt, order = orc.sort_samplings(st)
    np.testing.assert_array_equal(p.sorted.cpu().numpy(), srt)
    np.testing.assert_array_equal(p.order.cpu().numpy(), order)


def _reference(p):
    cfg = p.cfg
    t = p.sorted.cpu().numpy()
    rays = p.rays.cpu().numpy()
    fm = p.fmaps_np[None]  # [1,V,F,Hf,Wf]
    return nerf_ref.render_rays(p.weights, t, rays, p.K.cpu().numpy(), p.RT.cpu().numpy(), fm,
                                p.bbox.cpu().numpy(), p.zdir.cpu().numpy(), (cfg.img, cfg.img))


_ref_cache = {}


def test_render_fp32_matches_float64_reference():
    from sailor_amd.render import render_rays
    p = pipe()
    out = render_rays(p.sorted, p.rays, p.K, p.RT, (p.cfg.img, p.cfg.img), p.fmap, p.bbox,
                      p.zdir, p.model, want_occ=True)
    torch.cuda.synchronize()
    ref = _ref_cache.setdefault("fp32", _reference(p))
    rgb, dep, ws, asum, occ = [r.numpy() for r in ref]
    valid = p.sorted[0, :, 0].cpu().numpy() >= 0
    assert valid.sum() > 50
    g_occ = out["occ"].cpu().numpy()
    print("max|occ| err", np.abs(g_occ - occ).max(), "max|rgb| err",
          np.abs(out["rgb"].cpu().numpy() - rgb).max())
    np.testing.assert_allclose(g_occ, occ, atol=2e-5)
    np.testing.assert_allclose(out["aux"].cpu().numpy()[..., 0], ws, atol=1e-4)
    np.testing.assert_allclose(out["aux"].cpu().numpy()[..., 1], asum, atol=1e-4)
    np.testing.assert_allclose(out["depth"].cpu().numpy(), dep, atol=5e-4)
    np.testing.assert_allclose(out["rgb"].cpu().numpy(), rgb, atol=1e-4)
    # the pipeline's own shade() call wrote the same image
    np.testing.assert_allclose(p.rgb.cpu().numpy(), out["rgb"].cpu().numpy(), atol=0)


def psnr(a, b, mask):
    """cal_psnr (utils_render/utils_render.py:25-35) on [0,1] images -> 255 scale."""
    m = mask[..., None].astype(np.float64)
    mse = ((np.abs(a * 255 * m - b * 255 * m)) ** 2).sum() / max(m.sum(), 1)
    return 10 * np.log10(255 * 255 / max(mse, 1e-20))


@pytest.mark.parametrize("prec", ["bf16", "fp16"])
def test_low_precision_forms(prec):
    p32 = pipe("fp32")
    plo = pipe(prec)
    valid = p32.sorted[0, :, 0].cpu().numpy() >= 0
    a = p32.rgb[0].cpu().numpy()
    b = plo.rgb[0].cpu().numpy()
    ps = psnr(a, b, valid)
    print(f"{prec}: PSNR vs fp32 = {ps:.2f} dB")
    assert ps > 40.0
    np.testing.assert_allclose(plo.depth.cpu().numpy(), p32.depth.cpu().numpy(), atol=5e-3)


def test_graph_replay_matches_eager():
    from sailor_amd.pipeline import FramePipeline
    p = FramePipeline.synthetic(small_cfg(), prec="fp32")
    p.run()
    torch.cuda.synchronize()
    eager = (p.rgb.clone(), p.depth.clone(), p.sorted.clone(), p.node_info.clone())
    p.capture()
    for _ in range(3):
        p.rgb.zero_()
        p.depth.zero_()
        p.replay()
        torch.cuda.synchronize()
        assert torch.equal(p.rgb, eager[0])
        assert torch.equal(p.depth, eager[1])
        assert torch.equal(p.sorted, eager[2])
        assert torch.equal(p.node_info, eager[3])


def test_early_exit_gives_same_image():
    from sailor_amd.pipeline import FramePipeline
    p = FramePipeline.synthetic(small_cfg(), prec="fp32", early_exit=True)
    p.run()
    torch.cuda.synchronize()
    ref = pipe("fp32")
    np.testing.assert_allclose(p.rgb.cpu().numpy(), ref.rgb.cpu().numpy(), atol=0)
    np.testing.assert_allclose(p.depth.cpu().numpy(), ref.depth.cpu().numpy(), atol=0)


@pytest.mark.parametrize("prec,N,chunk", [("fp32", 1000, 0), ("fp32", 37, 8),
                                          ("bf16", 1000, 0), ("fp16", 531, 24)])
def test_density_query_matches_float64(prec, N, chunk):
    """sailor_density_points (infer_density path) on two batches of points
    in and around the scene box, including a partial last group of points,
    vs the float64 restatement.  fp32: max |d alpha| <= 1e-5; 16-bit forms
    (16-bit maps and MLP): <= 2e-2."""
    from sailor_amd.model import NerfWeights
    from sailor_amd.render import PackedModel, density_points
    p = pipe()
    rng = np.random.default_rng(7 + N)
    B = 2
    bbox = p.bbox.cpu().numpy()
    lo, hi = bbox[0, :, 0], bbox[0, :, 1]
    pts = rng.uniform(lo, hi, (B, N, 3)).astype(np.float32)
    K = np.repeat(p.K.cpu().numpy(), B, 0)
    RT = np.repeat(p.RT.cpu().numpy(), B, 0)
    fm_cl = np.repeat(p.fmap.cpu().numpy(), B, 0)
    bb = np.repeat(bbox, B, 0)
    model = PackedModel(p.weights, prec, "cuda")
    dt = {"fp32": torch.float32, "bf16": torch.bfloat16, "fp16": torch.float16}[prec]
    cu = lambda a: torch.from_numpy(np.ascontiguousarray(a)).cuda()
    got = density_points(cu(pts), cu(K), cu(RT), (p.cfg.img, p.cfg.img), cu(fm_cl).to(dt),
                         cu(bb), model, chunk_samples=chunk)
    torch.cuda.synchronize()
    ref = nerf_ref.density(p.weights, pts, K, RT, np.transpose(fm_cl, (0, 1, 4, 2, 3)), bb,
                           (p.cfg.img, p.cfg.img)).numpy()
    err = np.max(np.abs(got.cpu().numpy() - ref))
    assert err <= (1e-5 if prec == "fp32" else 2e-2), err


@pytest.mark.parametrize("views", [5, 6, 8])
def test_render_more_views_matches_float64_reference(views):
    """5 to 8 source views (8 view slots per sample): fp32 render vs the
    float64 restatement (same tolerances as the 4-view test), and the bf16
    form against it by PSNR."""
    from sailor_amd.pipeline import FrameConfig
    from sailor_amd.render import render_rays
    cfg = FrameConfig(f"c1_v{views}", views, 512, 600.0, 128, 8, 40, 47.0, 48, "c1")
    p = pipe("fp32", cfg)
    out = render_rays(p.sorted, p.rays, p.K, p.RT, (cfg.img, cfg.img), p.fmap, p.bbox, p.zdir,
                      p.model, want_occ=True)
    torch.cuda.synchronize()
    rgb, dep, ws, asum, occ = [r.numpy() for r in _reference(p)]
    valid = p.sorted[0, :, 0].cpu().numpy() >= 0
    assert valid.sum() > 50
    np.testing.assert_allclose(out["occ"].cpu().numpy(), occ, atol=2e-5)
    np.testing.assert_allclose(out["depth"].cpu().numpy(), dep, atol=5e-4)
    np.testing.assert_allclose(out["rgb"].cpu().numpy(), rgb, atol=1e-4)
    plo = pipe("bf16", cfg)
    ps = psnr(p.rgb[0].cpu().numpy(), plo.rgb[0].cpu().numpy(), valid)
    assert ps > 40.0, ps


def test_batch_graph_matches_eager_frames():
    """Frames of uneven size (2-6 views, own scale) captured into one graph
    over 4 streams: each replay writes exactly each frame's eager image, and
    the replayed images match the float64 restatement."""
    from sailor_amd.pipeline import BatchPipeline, FrameConfig
    base = FrameConfig("c5_small", 4, 256, 300.0, 128, 8, 40, 47.0, 48, "c1")
    bp = BatchPipeline(base, n_frames=5, prec="fp32", seed=77, workers=1)
    assert len({c.n_views for c in bp.cfgs}) > 1
    bp.autotune(candidates=(8, 16, 24))
    bp.run()
    torch.cuda.synchronize()
    eager = [(p.rgb.clone(), p.depth.clone()) for p in bp.pipes]
    assert all(int((p.sorted[0, :, 0] >= 0).sum()) > 20 for p in bp.pipes)
    bp.capture(n_streams=4)
    for _ in range(2):
        for p in bp.pipes:
            p.rgb.zero_()
            p.depth.zero_()
        bp.replay()
        torch.cuda.synchronize()
        for p, (r, d) in zip(bp.pipes, eager):
            assert torch.equal(p.rgb, r) and torch.equal(p.depth, d)
    # the replayed images against the float64 restatement, on the frames
    # with the fewest and the most views (4- and 8-view-slot kernels)
    views = [c.n_views for c in bp.cfgs]
    for i in sorted({int(np.argmin(views)), int(np.argmax(views))}):
        p = bp.pipes[i]
        valid = torch.nonzero(p.sorted[0, :, 0] >= 0).flatten().cpu().numpy()
        sub = valid[np.linspace(0, len(valid) - 1, min(120, len(valid))).astype(int)]
        rgb, dep = nerf_ref.render_rays(p.weights, p.sorted[:, sub].cpu().numpy(),
                                        p.rays[:, sub].cpu().numpy(), p.K.cpu().numpy(),
                                        p.RT.cpu().numpy(), p.fmaps_np[None],
                                        p.bbox.cpu().numpy(), p.zdir.cpu().numpy(),
                                        (p.cfg.img, p.cfg.img))[:2]
        err_rgb = np.abs(p.rgb[0, sub].cpu().numpy() - rgb[0].numpy()).max()
        err_dep = np.abs(p.depth[0, sub].cpu().numpy() - dep[0].numpy()).max()
        print(f"batched frame {i} ({views[i]} views): max|rgb| err {err_rgb:.2e}, "
              f"max|depth| err {err_dep:.2e}")
        assert err_rgb < 1e-4 and err_dep < 5e-4


@pytest.mark.parametrize("views,prec", [(4, "fp32"), (4, "bf16"), (6, "fp32"), (6, "bf16")])
def test_every_launch_shape_renders_the_same_image(views, prec):
    """Columns of the MLP are independent and the compositing and token sums
    run in sample order, so every instantiated chunk size (4 view slots:
    8/16/24/48; 8 slots: fp32 8/12/16, 16-bit 4/8/12/16) and a sample kernel
    launched on part of the chip (grid share 40 %) give the default launch's
    image bit for bit; chunks whose LDS does not fit are rejected."""
    from sailor_amd import _lib
    from sailor_amd.pipeline import FrameConfig
    cfg = FrameConfig(f"shapes_v{views}", views, 512, 600.0, 128, 8, 40, 47.0, 48, "c1")
    p = pipe(prec, cfg)
    ref = (p.rgb.clone(), p.depth.clone(), p.aux.clone())
    cands = (8, 16, 24, 48) if views <= 4 else ((8, 12, 16) if prec == "fp32" else (4, 8, 12, 16))
    ran = []
    for ns in cands:
        for share in (0, 40):
            p.chunk_samples = ns | (share << 8)
            p.rgb.zero_()
            try:
                p.shade()
            except _lib.SailorError as e:
                assert "LDS" in str(e), e
                continue
            torch.cuda.synchronize()
            assert torch.equal(p.rgb, ref[0]) and torch.equal(p.depth, ref[1]), (ns, share)
            assert torch.equal(p.aux, ref[2]), (ns, share)
            ran.append((ns, share))
    p.chunk_samples = 0
    assert len(ran) >= 4, ran
    print(views, prec, "launch shapes checked:", ran)


@pytest.mark.parametrize("views,prec", [(4, "bf16"), (4, "fp16"), (6, "bf16")])
def test_mfma_hydra_matches_rounded_float64_on_the_same_tokens(views, prec):
    """The 16-bit forms' Hydra encoder + RGB head on MFMA, fed the tokens
    the sample kernel left in the workspace, against the float64
    restatement with every product's operands rounded to the same 16-bit
    type (nerf_ref.hydra_rgb(mm_dtype)): the remaining difference is fp32
    against float64 accumulation (|d rgb| <= 2e-4)."""
    from sailor_amd.pipeline import FrameConfig
    from sailor_amd.train import workspace_views
    cfg = FrameConfig(f"hyd_v{views}", views, 512, 600.0, 128, 8, 48, 56.25, 48, "c1")
    p = pipe(prec, cfg)
    p.shade()
    torch.cuda.synchronize()
    n_rays = cfg.target ** 2
    vp = 4 if views <= 4 else 8
    n_valid = int(p.render_ws[:4].view(torch.int32)[0].item())
    lst = p.render_ws[256:256 + 4 * n_rays].view(torch.int32)[:n_valid].long()
    toff = 256 + (n_rays * 4 + 255) // 256 * 256
    tok = p.render_ws[toff:toff + n_valid * vp * p.mc.F * 4].view(torch.float32)
    tok = tok.view(n_valid, vp, p.mc.F)[:, :views].double().cpu()
    assert n_valid > 50
    w = p.weights
    D = lambda a: torch.tensor(np.asarray(a), dtype=torch.float64)
    hyd = [{k: D(v) for k, v in st.items()} for st in w.hydra]
    rgbw = [(D(a), D(b)) for a, b in w.rgb]
    dt = torch.bfloat16 if prec == "bf16" else torch.float16
    ref = nerf_ref.hydra_rgb(hyd, rgbw, tok, mm_dtype=dt).numpy()
    got = p.rgb[0][lst.cuda()].double().cpu().numpy()
    err = np.abs(got - ref)
    print(f"{views} views {prec}: MFMA Hydra vs rounded float64 max {err.max():.2e} "
          f"(99.9% {np.quantile(err, 0.999):.2e})")
    assert np.quantile(err, 0.999) <= 2e-4 and err.max() <= 2e-3, err.max()
