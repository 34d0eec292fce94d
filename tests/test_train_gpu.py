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


def _grad_errors(got, ref):
    """(max |got - ref| / max |ref|, ||got - ref|| / ||ref||) of one tensor."""
    got = np.asarray(got, np.float64).reshape(-1)
    ref = np.asarray(ref, np.float64).reshape(-1)
    scale = max(np.max(np.abs(ref)), 1e-30)
    return (np.max(np.abs(got - ref)) / scale,
            np.linalg.norm(got - ref) / max(np.linalg.norm(ref), 1e-30))


def _flat_grads(gr):
    """Named flat list of a gradient dict (RenderTrainer.backward layout)."""
    out = [(f"dense_w{l}", gr["dense_w"][l]) for l in range(4)]
    out += [(f"dense_b{l}", gr["dense_b"][l]) for l in range(4)]
    out += [("sigma_w", gr["sigma_w"]), ("sigma_b", gr["sigma_b"])]
    for s, st in enumerate(gr["hydra"]):
        out += [(f"hydra{s}.{k}", v) for k, v in st.items()]
    for i, (a, b) in enumerate(gr["rgb"]):
        out += [(f"rgb{i}.w", a), (f"rgb{i}.b", b)]
    return out


@pytest.mark.parametrize("prec,batch_items", [("fp32", 128), ("bf16", 4096)])
def test_render_backward_matches_autograd(prec, batch_items):
    """All parameter gradients of L = sum(g . rgb) from the fused backward vs
    torch autograd through the float64 restatement; g is non-zero on a random
    subset of the listed rays (the others must contribute exactly nothing).
    Tolerance: fp32 max error <= 1e-4 of the tensor's largest entry; bf16
    relative L2 error <= 1e-2."""
    from sailor_amd.model import NerfWeights
    from sailor_amd.train import RenderTrainer
    p = pipe()
    cfg = p.cfg
    w = NerfWeights.init(p.mc.F, p.mc.H, perturb_std=0.01)
    w.dense_b = [np.random.default_rng(40 + l).normal(0, 0.01, p.mc.H).astype(np.float32)
                 for l in range(4)]
    tr = RenderTrainer(w, prec, "cuda", batch_items=batch_items)
    n_rays = cfg.target * cfg.target
    img_hw = (cfg.img, cfg.img)
    out = tr.forward(p.sorted, p.rays, p.K, p.RT, img_hw, p.fmap, p.bbox, p.zdir)
    torch.cuda.synchronize()
    valid = torch.nonzero(p.sorted[0, :, 0] >= 0).flatten().cpu().numpy()
    assert valid.size > 100
    rng = np.random.default_rng(9)
    sub = np.sort(rng.choice(valid, size=min(160, valid.size), replace=False))
    g = np.zeros((1, n_rays, 3), np.float32)
    g[0, sub] = rng.normal(size=(sub.size, 3)).astype(np.float32)
    grads = tr.backward(torch.from_numpy(g).cuda())
    torch.cuda.synchronize()

    P = nerf_ref.params_of(w, requires_grad=True)
    rgb, _, _, _, occ = nerf_ref.render_rays(
        w, p.sorted[:, sub].cpu().numpy(), p.rays[:, sub].cpu().numpy(), p.K.cpu().numpy(),
        p.RT.cpu().numpy(), p.fmaps_np[None], p.bbox.cpu().numpy(), p.zdir.cpu().numpy(),
        img_hw, params=P)
    if prec == "fp32":
        np.testing.assert_allclose(out["rgb"][0, sub].double().cpu().numpy(),
                                   rgb[0].detach().numpy(), atol=2e-5)
    (rgb * torch.from_numpy(g[:, sub]).double()).sum().backward()
    ref = dict(dense_w=[x.grad for x in P["dense_w"]], dense_b=[x.grad for x in P["dense_b"]],
               sigma_w=P["sigma_w"].grad, sigma_b=P["sigma_b"].grad,
               hydra=[{k: v.grad for k, v in st.items()} for st in P["hydra"]],
               rgb=[(a.grad, b.grad) for a, b in P["rgb"]])
    bad = []
    for (name, got), (_, want) in zip(_flat_grads(grads), _flat_grads(ref)):
        got = got.detach().double().cpu().numpy() if torch.is_tensor(got) else got
        mx, l2 = _grad_errors(got, want.numpy())
        if (prec == "fp32" and mx > 1e-4) or (prec != "fp32" and l2 > 1e-2):
            bad.append((name, mx, l2))
    assert not bad, bad
