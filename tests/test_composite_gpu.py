"""GPU parity of the compositing the reference trains with
(_volume_render_torch, voxel_encoding_helper.py:132-186), now on the HIP
kernels behind sailor_amd.voxel_encoding_helper.volume_render_torch, against
the float64 tensor restatement oracle/composite_ref.py:
  * forward of both forms, with and without the valid-ray mask and with
    Kahan-compensated sums, at 48 samples per ray;
  * gradients through torch.autograd (the HIP backward) vs float64 torch
    autograd of the restatement: relative error <= 1e-4 per tensor;
  * central finite differences (h = 1e-3, float64) of an MSE loss with
    respect to 256 random sigma inputs against the HIP gradient.
"""
import numpy as np
import pytest
import torch

from oracle import composite_ref as cr

pytestmark = pytest.mark.gpu


def _case(occ_mode, B=2, R=3000, S=48, F=3, seed=0):
    rng = np.random.default_rng(seed)
    t = np.sort(rng.uniform(0.5, 4.0, (B, R, S)), -1).astype(np.float32)
    if occ_mode:
        s = rng.uniform(0.0, 0.35, (B, R, S, 1)).astype(np.float32)
    else:
        s = rng.uniform(0.0, 30.0, (B, R, S, 1)).astype(np.float32)
    f = rng.uniform(0.0, 1.0, (B, R, S, F)).astype(np.float32)
    valid = rng.random((B, R)) > 0.1
    return t, s, f, valid


def _rel(a, b):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-30))


@pytest.mark.parametrize("occ_mode", [False, True])
@pytest.mark.parametrize("masked", [False, True])
def test_forward_matches_float64(occ_mode, masked):
    from sailor_amd.voxel_encoding_helper import volume_render_torch
    t, s, f, valid = _case(occ_mode, seed=1 + 2 * occ_mode + masked)
    vr = torch.from_numpy(valid).cuda() if masked else None
    fn = volume_render_torch.forward_occ if occ_mode else volume_render_torch.forward
    outs = fn(torch.from_numpy(t).cuda(), torch.from_numpy(s).cuda(),
              torch.from_numpy(f).cuda(), 4.0, vr)
    ref = cr.render(torch.from_numpy(t).double(), torch.from_numpy(s).double(),
                    torch.from_numpy(f).double(), occ_mode,
                    torch.from_numpy(valid) if masked else None)
    for name, o, r in zip(("feats", "depth", "ws", "alpha"), outs, ref):
        o, r = o.cpu().numpy(), r.numpy()
        assert o.shape == r.shape, (name, o.shape, r.shape)
        err = np.abs(o - r).max()
        assert err <= 2e-5 * max(1.0, np.abs(r).max()), (name, err)
        if masked:
            assert np.all(o[~valid] == 0), name


@pytest.mark.parametrize("occ_mode", [False, True])
def test_compensated_sums_not_worse(occ_mode):
    """Kahan-compensated sums (precision form) against float64: no worse
    than the plain float32 sums on every output."""
    import sailor_amd.VoxelEncoding as VE
    t, s, f, _ = _case(occ_mode, B=1, R=4000, S=48, seed=9)
    ref = cr.render(torch.from_numpy(t).double(), torch.from_numpy(s).double(),
                    torch.from_numpy(f).double(), occ_mode)
    args = [torch.from_numpy(a).cuda().contiguous() for a in (t, s[..., 0], f)]
    plain = VE.volume_render_torch_forward(*args, None, occ_mode, compensated=False)
    comp = VE.volume_render_torch_forward(*args, None, occ_mode, compensated=True)
    for name, p, c, r in zip(("feats", "depth", "ws", "alpha"), plain, comp, ref):
        ep = np.abs(p.cpu().numpy() - r.numpy()).max()
        ec = np.abs(c.cpu().numpy() - r.numpy()).max()
        print(f"{'occ' if occ_mode else 'density'} {name}: plain {ep:.2e} compensated {ec:.2e}")
        assert ec <= ep * 1.05 + 1e-7, (name, ep, ec)


@pytest.mark.parametrize("occ_mode", [False, True])
def test_gradients_match_float64_autograd(occ_mode):
    from sailor_amd.voxel_encoding_helper import volume_render_torch
    t, s, f, valid = _case(occ_mode, R=2000, seed=20 + occ_mode)
    rng = np.random.default_rng(30)
    tgt = rng.uniform(0, 1, (2, 2000, 3)).astype(np.float32)
    fn = volume_render_torch.forward_occ if occ_mode else volume_render_torch.forward
    sc = torch.from_numpy(s).cuda().requires_grad_()
    fc = torch.from_numpy(f).cuda().requires_grad_()
    vr = torch.from_numpy(valid).cuda()
    o = fn(torch.from_numpy(t).cuda(), sc, fc, 4.0, vr)
    loss = ((o[0] - torch.from_numpy(tgt).cuda()) ** 2).mean() + 0.1 * o[1].mean() \
        + 0.05 * o[2].square().mean() + 0.2 * o[3].mean()
    gs, gf = torch.autograd.grad(loss, [sc, fc])
    s64 = torch.from_numpy(s).double().requires_grad_()
    f64 = torch.from_numpy(f).double().requires_grad_()
    r = cr.render(torch.from_numpy(t).double(), s64, f64, occ_mode, torch.from_numpy(valid))
    loss64 = ((r[0] - torch.from_numpy(tgt).double()) ** 2).mean() + 0.1 * r[1].mean() \
        + 0.05 * r[2].square().mean() + 0.2 * r[3].mean()
    rs, rf = torch.autograd.grad(loss64, [s64, f64])
    es, ef = _rel(gs.cpu().numpy(), rs.numpy()), _rel(gf.cpu().numpy(), rf.numpy())
    print(f"{'occ' if occ_mode else 'density'}: rel err sigma {es:.2e} feats {ef:.2e}")
    assert es <= 1e-4 and ef <= 1e-4, (es, ef)
    assert np.all(gs.cpu().numpy()[~valid] == 0)


@pytest.mark.parametrize("occ_mode", [False, True])
def test_sigma_gradient_central_differences(occ_mode):
    """256 random sigma inputs: (L(s + h) - L(s - h)) / 2h of the float64
    restatement, h = 1e-3, vs the HIP gradient; |fd - g| <= 1e-4 max|g|."""
    from sailor_amd.voxel_encoding_helper import volume_render_torch
    t, s, f, _ = _case(occ_mode, B=1, R=512, seed=40 + occ_mode)
    rng = np.random.default_rng(41)
    tgt = torch.from_numpy(rng.uniform(0, 1, (1, 512, 3)))
    fn = volume_render_torch.forward_occ if occ_mode else volume_render_torch.forward
    sc = torch.from_numpy(s).cuda().requires_grad_()
    o = fn(torch.from_numpy(t).cuda(), sc, torch.from_numpy(f).cuda(), 4.0)
    loss = ((o[0] - tgt.float().cuda()) ** 2).sum() + o[1].sum()
    g, = torch.autograd.grad(loss, [sc])
    g = g.double().cpu().numpy().reshape(-1)
    t64, f64 = torch.from_numpy(t).double(), torch.from_numpy(f).double()
    s64 = torch.from_numpy(s).double().reshape(-1)

    def L(sv):
        r = cr.render(t64, sv.reshape(s.shape), f64, occ_mode)
        return float(((r[0] - tgt) ** 2).sum() + r[1].sum())
    picks = rng.choice(s64.numel(), size=256, replace=False)
    h = 1e-3
    scale = np.abs(g).max()
    worst = 0.0
    for i in picks:
        sp, sm = s64.clone(), s64.clone()
        sp[i] += h
        sm[i] -= h
        fd = (L(sp) - L(sm)) / (2 * h)
        worst = max(worst, abs(fd - g[i]) / scale)
    print(f"{'occ' if occ_mode else 'density'}: worst |fd - g| / max|g| = {worst:.2e}")
    assert worst <= 1e-4, worst
