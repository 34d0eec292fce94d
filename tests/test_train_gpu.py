"""GPU checks of the differentiation form against torch autograd of the
float64 restatement (oracle/nerf_ref.py):
  * Hydra + RGB head backward: dL/d token and the small-weight gradient.
"""
import numpy as np
import pytest
import torch

from oracle import nerf_ref

pytestmark = pytest.mark.gpu

_pipes = {}


def pipe(prec="fp32"):
    from sailor_amd.pipeline import FrameConfig, FramePipeline
    cfg = FrameConfig("c1_small", 4, 512, 600.0, 128, 8, 48, 56.25, 64, "c1")
    if prec not in _pipes:
        p = FramePipeline.synthetic(cfg, prec=prec)
        p.run()
        torch.cuda.synchronize()
        _pipes[prec] = p
    return _pipes[prec]


def _small_layout(w, tensors):
    """Flatten tensors named like NerfWeights.small_buffer's parts."""
    parts = [tensors["sigma_w"], torch.cat([tensors["sigma_b"].reshape(1),
                                            torch.zeros(3, dtype=torch.float64)])]
    for st in tensors["hydra"]:
        parts += [st["q"], st["k"], st["v"], st["fc"], st["w1"], st["b1"], st["w2"], st["b2"]]
    (w1, b1), (w2, b2), (w3, b3) = tensors["rgb"]
    parts += [w1, b1, w2, b2, w3, torch.cat([b3, torch.zeros(1, dtype=torch.float64)])]
    return torch.cat([p.reshape(-1) for p in parts])


def test_hydra_backward_matches_autograd():
    from sailor_amd.train import BackwardScratch, hydra_backward, workspace_views
    p = pipe()
    cfg, w = p.cfg, p.weights
    n_rays = cfg.target * cfg.target
    p.shade()
    torch.cuda.synchronize()
    n_valid, lst, tok = workspace_views(p.render_ws, n_rays, p.mc.F)
    assert n_valid > 50
    g = torch.Generator(device="cpu").manual_seed(5)
    grad_rgb = torch.randn((n_rays, 3), generator=g).cuda()
    scr = BackwardScratch(1, n_rays, p.mc.F, p.mc.H, "cuda")
    dtok, gsmall = hydra_backward(grad_rgb, 1, n_rays, cfg.n_views, p.model, p.render_ws, scr)
    torch.cuda.synchronize()

    D = lambda a: torch.tensor(np.asarray(a), dtype=torch.float64, requires_grad=True)
    hyd = [{k: D(v) for k, v in st.items()} for st in w.hydra]
    rgbw = [(D(a), D(b)) for a, b in w.rgb]
    t64 = tok.double().cpu().requires_grad_()
    rgb = nerf_ref.hydra_rgb(hyd, rgbw, t64)
    # the forward's colour of the listed rays is this function of the tokens
    np.testing.assert_allclose(rgb.detach().numpy(),
                               p.rgb[0][lst.long()].double().cpu().numpy(), atol=2e-5)
    loss = (rgb * grad_rgb[lst.long()].double().cpu()).sum()
    loss.backward()
    np.testing.assert_allclose(dtok[:n_valid].double().cpu().numpy(), t64.grad.numpy(),
                               rtol=1e-4, atol=1e-6)
    zero = lambda n: torch.zeros(n, dtype=torch.float64)
    ref = _small_layout(w, {
        "sigma_w": zero(p.mc.H), "sigma_b": zero(1),
        "hydra": [{k: v.grad for k, v in st.items()} for st in hyd],
        "rgb": [(a.grad, b.grad) for a, b in rgbw]}).numpy()
    got = gsmall.double().cpu().numpy()
    scale = np.max(np.abs(ref))
    assert np.max(np.abs(got - ref)) <= 1e-4 * scale, (np.max(np.abs(got - ref)), scale)
