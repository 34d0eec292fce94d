"""The precision form's pieces on the GPU:
  * sailor_quantize_channels: power-of-two per-channel scales from max |x|
    and the scaled 16-bit maps, bit for bit against numpy / torch rounding;
  * sailor_scale_packed_layer: the first layer's feature inputs times the
    scales, bit for bit;
  * a frame rendered from quantised maps equals the unquantised bf16 frame
    bit for bit (power-of-two scales are exact; fp16 differs only below its
    normal range), and stays accurate where the unquantised fp16 form
    overflows (a channel of magnitude 1e5);
  * Kahan-compensated TSDF sums and compositing sums: no less accurate
    against float64 than the plain float32 recurrences.
"""
import numpy as np
import pytest
import torch

from oracle import nerf_ref
from oracle import oracle as orc

pytestmark = pytest.mark.gpu


def small_cfg(name="prec_small"):
    from sailor_amd.pipeline import FrameConfig
    return FrameConfig(name, 4, 512, 600.0, 128, 8, 48, 56.25, 64, "c1")


@pytest.mark.parametrize("prec,C", [("bf16", 64), ("fp16", 64), ("bf16", 32), ("fp16", 2),
                                     ("bf16", 512)])
def test_quantize_channels_bitexact(prec, C):
    """C a multiple of 4 up to 256 takes the 16-byte-load reduction, C = 2
    and C = 512 the scalar one; 5000 rows leave a remainder for both, and a
    channel's maximum sits in the last row."""
    from sailor_amd._lib import call, ptr, stream_of
    from sailor_amd.model import PREC_IDS
    rng = np.random.default_rng(1)
    rows = 5000
    x = rng.standard_normal((rows, C)).astype(np.float32)
    x *= np.float32(10.0) ** rng.uniform(-6, 6, C).astype(np.float32)   # wide per-channel ranges
    x[:, C - 1] *= np.float32(1e-3)
    x[rows - 1, C - 1] = 3.0e4 if prec == "bf16" else 3.0e3              # maximum in the tail
    if C > 7:
        x[:, 5] = 0.0                                                    # all-zero channel
        x[:, 7] = 4.0                                                    # exact power of two
    xt = torch.from_numpy(x).cuda()
    dt = torch.bfloat16 if prec == "bf16" else torch.float16
    out = torch.empty((rows, C), dtype=dt, device="cuda")
    sc = torch.empty(C, dtype=torch.float32, device="cuda")
    ws = torch.empty(4 * C, dtype=torch.uint8, device="cuda")
    call("sailor_quantize_channels", ptr(xt), rows, C, PREC_IDS[prec], ptr(out), ptr(sc), ptr(ws),
         ws.numel(), stream_of())
    m = np.abs(x).max(0)
    _, e = np.frexp(m)
    want_s = np.where(m > 0, np.ldexp(1.0, e), 1.0).astype(np.float32)
    np.testing.assert_array_equal(sc.cpu().numpy(), want_s)
    if C > 7:
        assert sc[7].item() == 8.0 and sc[5].item() == 1.0
    want_q = torch.from_numpy(x / want_s).to(dt)
    assert torch.equal(out.cpu(), want_q)
    assert float(out.float().abs().max()) <= 1.0


def test_scale_packed_layer_bitexact():
    from sailor_amd._lib import call, ptr, stream_of
    from sailor_amd.model import NerfWeights
    from sailor_amd.render import PackedModel
    m = PackedModel(NerfWeights.init(32, 128), "bf16", "cuda")
    w0, kidx = m.first_layer()
    sc = torch.from_numpy(np.ldexp(1.0, np.arange(32) % 7 - 3).astype(np.float32)).cuda()
    out = torch.empty(w0.numel(), dtype=torch.bfloat16, device="cuda")
    call("sailor_scale_packed_layer", ptr(w0), ptr(kidx), w0.numel(), 32, ptr(sc), 1, ptr(out),
         stream_of())
    k = kidx.long().cpu()
    s_full = torch.ones(w0.numel())
    feat = k < 32
    s_full[feat] = sc.cpu()[k[feat]]
    assert torch.equal(out.cpu(), (w0.cpu() * s_full).to(torch.bfloat16))
    # unscaled inputs (positional encoding) are the packed 16-bit weights as they were
    assert torch.equal(out.cpu()[~feat], m.w_dense[:w0.numel()].cpu()[~feat])


@pytest.mark.parametrize("prec", ["bf16", "fp16"])
def test_quantised_frame_equals_unquantised(prec):
    """bf16 (8 exponent bits): bit-identical.  fp16: dividing by s_c moves
    inputs below 2^-14 s_c into fp16's subnormal range (fewer significant
    bits), so inputs within a few binades of zero round differently: the
    frames agree to > 90 dB PSNR."""
    from sailor_amd.pipeline import FramePipeline
    pq = FramePipeline.synthetic(small_cfg(), prec=prec, quantize=True)
    pu = FramePipeline.synthetic(small_cfg(), prec=prec, quantize=False)
    for p in (pq, pu):
        p.run()
    torch.cuda.synchronize()
    assert float(pq.fscale.max()) > 1.0               # the maps really were scaled
    if prec == "bf16":
        assert torch.equal(pq.rgb, pu.rgb) and torch.equal(pq.depth, pu.depth)
        assert torch.equal(pq.aux, pu.aux)
    else:
        mse = float(((pq.rgb.double() - pu.rgb.double()) ** 2).mean())
        ps = 10 * np.log10(1.0 / max(mse, 1e-30))
        print(f"fp16 quantised vs unquantised: PSNR {ps:.1f} dB")
        assert ps > 90.0


def test_quantisation_guards_the_fp16_range():
    """Channel 0 of every map times 1e5 and the first layer's weights of
    that channel divided by 1e5: the same function in exact arithmetic.
    Unquantised fp16 maps overflow (|x| > 65504); quantised maps keep the
    render within 40 dB PSNR of fp32."""
    from sailor_amd.model import NerfWeights
    from sailor_amd.pipeline import FramePipeline
    w = NerfWeights.init(32, 128)
    w.dense_w[0] = w.dense_w[0].copy()
    w.dense_w[0][:, 63] /= np.float32(1e5)             # feature channel 0 (after the 63 PE dims)
    res = {}
    for name, prec, q in (("fp32", "fp32", None), ("q", "fp16", True), ("plain", "fp16", False)):
        p = FramePipeline.synthetic(small_cfg("prec_range"), prec=prec, weights=w, quantize=q)
        p.fmap[..., 0] *= 1e5
        p.run()
        torch.cuda.synchronize()
        res[name] = p
    valid = res["fp32"].sorted[0, :, 0] >= 0
    ref = res["fp32"].rgb[0, valid].double()
    mse = lambda p: float(((p.rgb[0, valid].double() - ref) ** 2).mean())
    psnr_q = 10 * np.log10(1.0 / max(mse(res["q"]), 1e-30))
    plain = res["plain"].rgb[0, valid]
    plain_bad = (not bool(torch.isfinite(plain).all())) or \
        10 * np.log10(1.0 / max(mse(res["plain"]), 1e-30)) < 30
    print(f"fp16 with a 1e5-scale channel: quantised PSNR {psnr_q:.1f} dB, "
          f"unquantised broken: {plain_bad}")
    assert psnr_q > 40.0
    assert plain_bad


def test_compensated_tsdf_not_less_accurate():
    """Kahan-compensated TSDF vs the plain running average, both against the
    float64 restatement (errors at the float32 level either way)."""
    from sailor_amd import VoxelEncoding as VE
    from sailor_amd import scenes
    from tests.helpers import scene_inputs
    inp = scene_inputs(0, 6, 256, 256, 300.0)
    cu = lambda a: torch.from_numpy(np.ascontiguousarray(a)).cuda()
    fb = scenes.fixed_cube()[None]
    outs = {}
    for comp in (False, True):
        occ = torch.full((1, 96, 96, 96), -1, dtype=torch.int32, device="cuda")
        num = torch.zeros((1, 1), dtype=torch.int32, device="cuda")
        outs[comp] = VE.integrate(occ, num, cu(inp["K"]), cu(inp["RT"]), cu(inp["R"]),
                                  cu(inp["T"]), cu(inp["colors"]), cu(inp["depths"]),
                                  cu(inp["masks"]), 1.0, 0.025, 3.0, 13.5, 13.5, 0,
                                  fixed_bbox=cu(fb), compensated=comp) + (occ,)
    bbox = outs[False][4].cpu().numpy()
    vs = outs[False][5].cpu().numpy().reshape(-1)
    t64, w64 = orc.integrate_f64(inp["K"], inp["RT"], inp["colors"], inp["depths"], bbox, vs,
                                 (96, 96, 96))
    seen = w64 > 1                                   # voxels averaged over several views
    errs = {}
    for comp in (False, True):
        d = np.abs(outs[comp][0].cpu().numpy().astype(np.float64) - t64)[seen]
        errs[comp] = (float(d.max()), float(np.sqrt((d ** 2).mean())))
    print("tsdf vs float64 (max, rms): plain", errs[False], "compensated", errs[True])
    assert errs[True][1] <= errs[False][1] * 1.05
    assert errs[True][0] <= max(errs[False][0], 2.5e-7)
    np.testing.assert_array_equal(outs[True][1].cpu().numpy(), outs[False][1].cpu().numpy())


def test_compensated_compositing_not_less_accurate():
    """fp32 render with Kahan-compensated compositing sums vs the plain
    sums, both against the float64 restatement on the listed rays."""
    from sailor_amd.pipeline import FramePipeline
    from sailor_amd.render import render_rays
    p = FramePipeline.synthetic(small_cfg(), prec="fp32")
    p.run()
    torch.cuda.synchronize()
    valid = torch.nonzero(p.sorted[0, :, 0] >= 0).flatten().cpu().numpy()
    sub = valid[:300]
    ref = nerf_ref.render_rays(p.weights, p.sorted[:, sub].cpu().numpy(),
                               p.rays[:, sub].cpu().numpy(), p.K.cpu().numpy(),
                               p.RT.cpu().numpy(), p.fmaps_np[None], p.bbox.cpu().numpy(),
                               p.zdir.cpu().numpy(), (p.cfg.img, p.cfg.img))
    rows = {}
    for comp in (False, True):
        out = render_rays(p.sorted, p.rays, p.K, p.RT, (p.cfg.img, p.cfg.img), p.fmap, p.bbox,
                          p.zdir, p.model, compensated=comp)
        torch.cuda.synchronize()
        rows[comp] = {k: float(np.sqrt(((g.cpu().numpy()[0, sub].astype(np.float64)
                                         - r[0].numpy()) ** 2).mean()))
                      for k, g, r in (("rgb", out["rgb"], ref[0]), ("depth", out["depth"], ref[1]),
                                      ("ws", out["aux"][..., 0], ref[2]))}
    print("rms vs float64: plain", rows[False], "compensated", rows[True])
    for k in rows[True]:
        assert rows[True][k] <= rows[False][k] * 1.1 + 1e-9, (k, rows)
