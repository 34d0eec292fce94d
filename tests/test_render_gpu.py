"""GPU parity of volume compositing (forward and backward) and the frequency
encoding against the CPU oracle, through the C ABI."""
import numpy as np
import pytest
import torch

from oracle import oracle as orc

pytestmark = pytest.mark.gpu


def _rand_rays(B=2, R=513, S=48, F=3, seed=0, invalid_frac=0.1):
    rng = np.random.default_rng(seed)
    t = np.sort(rng.uniform(0.5, 4.0, (B, R, S)).astype(np.float32), -1)
    # some rays fully invalid, some with -1 holes (sorted to the front)
    inv = rng.random((B, R)) < invalid_frac
    t[inv] = -1.0
    holes = rng.random((B, R, S)) < 0.05
    t = np.where(holes, np.float32(-1.0), t)
    t = np.sort(t, -1)
    occ = rng.uniform(0, 1, (B, R, S)).astype(np.float32) ** 3
    feats = rng.standard_normal((B, R, S, F)).astype(np.float32)
    return t, occ, feats


# streamed LDS tiles (S <= 48 at 16 rays, F = 8 at 4 rays with 8 lanes per
# ray, F = 5 in two passes of 8-lane groups), the chunked LDS form for rows
# longer than a tile (S = 200), the per-lane form for F > 8, a ray count
# that gives each persistent workgroup many tiles, and tiles holding only
# two rays (B = 2, R = 1)
@pytest.mark.parametrize("F,S,R", [(3, 48, 513), (8, 48, 513), (33, 48, 513), (3, 37, 513),
                                   (5, 21, 513), (1, 7, 513), (3, 200, 513), (3, 48, 70001),
                                   (3, 48, 1), (8, 48, 1)])
def test_occ_forward_bitexact(F, S, R):
    from sailor_amd import VoxelEncoding as VE
    t, occ, feats = _rand_rays(R=R, F=F, S=S)
    of, od, ow, oa = VE.volume_rendering_occ_forward(
        torch.from_numpy(t).cuda(), torch.from_numpy(occ).cuda(),
        torch.from_numpy(feats).cuda(), 0, 1e-4)
    rf, rd, rw, ra = orc.render_occ(t, occ, feats, 1e-4)
    np.testing.assert_array_equal(of.cpu().numpy(), rf)
    np.testing.assert_array_equal(od.cpu().numpy()[..., 0], rd)
    np.testing.assert_array_equal(ow.cpu().numpy()[..., 0], rw)
    np.testing.assert_array_equal(oa.cpu().numpy()[..., 0], ra)


# (F, S) cover each launch form of the compositing: LDS tiles of 16 rays
# (S <= 48, S*F <= 144), 8 rays (S = 52) and 4 rays (F = 8), the per-lane global
# form for rows too long for a tile (S = 200) and for more than 8 channels
# (F = 12).  602 rays leave a partial last tile for every tile size; the
# R = 70001 case gives each persistent workgroup many tiles; R = 1 leaves a
# single tile of two rays.
@pytest.mark.parametrize("F,S,R", [(3, 32, 301), (3, 48, 301), (3, 52, 301), (8, 48, 301),
                                   (5, 21, 301), (1, 7, 301), (3, 200, 301), (12, 48, 301),
                                   (3, 48, 70001), (3, 48, 1), (5, 21, 1)])
@pytest.mark.parametrize("occ_mode", [True, False])
def test_training_forward_backward(occ_mode, F, S, R):
    from sailor_amd import VoxelEncoding as VE
    rng = np.random.default_rng(3 + F + S)
    B = 2
    t = rng.uniform(0.5, 4.0, (B, R, S)).astype(np.float32)
    t[rng.random((B, R, S)) < 0.05] = -1.0
    t[0, :7] = -1.0
    order = np.argsort(t, -1, kind="stable").astype(np.int32)
    if S > 2 and occ_mode:
        # rows whose sort index repeats an entry: both walks read the
        # original inputs and the later gradient write wins.  Occupancy form
        # only: in the density form a repeat makes delta negative, exp then
        # sees large positive arguments and the gradient sums cancel, so the
        # 1-ulp expf difference between device and C library grows without
        # bound.
        dup = rng.random((B, R)) < 0.1
        order[dup, S // 2] = order[dup, 1]
    sigma = (rng.uniform(0, 1, (B, R, S)) if occ_mode else rng.uniform(0, 30, (B, R, S))
             ).astype(np.float32)
    feats = rng.standard_normal((B, R, S, F)).astype(np.float32)
    g = [rng.standard_normal(s).astype(np.float32) for s in ((B, R, F), (B, R, 1), (B, R, 1),
                                                             (B, R, 1))]
    cu = lambda a: torch.from_numpy(np.ascontiguousarray(a)).cuda()
    outs = VE.volume_rendering_training_forward(cu(t), cu(order), cu(sigma), cu(feats), 0,
                                                1e-4, occ_mode)
    ref = orc.render_train_fwd(t, order, sigma, feats, 1e-4, occ_mode)
    tol = 0 if occ_mode else 2e-5
    # density form: the device expf and the C library's differ by up to 1 ulp,
    # which can move a ray's transmittance across the early-termination
    # threshold; such a ray then legitimately differs in every output.  At
    # many rays a budget of 1e-4 of the rays (at least 2) may do so; all other
    # rays meet the tolerance.
    budget = 0 if (occ_mode or R <= 1000) else max(2, int(1e-4 * B * R))
    bad = np.zeros(B * R, bool)

    def check(got, want, tol_):
        nonlocal bad
        got = got.reshape(want.shape)
        off = ~np.isclose(got, want, rtol=tol_, atol=tol_)
        bad |= off.reshape(B * R, -1).any(1)

    for a, b in zip(outs, ref):
        check(a.cpu().numpy(), b, tol)
    fwd_bad = int(bad.sum())
    gf, gs = VE.volume_rendering_training_backward(*[cu(x) for x in g], *outs, cu(t), cu(order),
                                                   cu(sigma), cu(feats), 0, 1e-4, occ_mode)
    rgf, rgs = orc.render_train_bwd(*g, *[o.cpu().numpy() for o in outs], t, order, sigma,
                                    feats, 1e-4, occ_mode)
    check(gf.cpu().numpy(), rgf, tol)
    check(gs.cpu().numpy()[..., 0], rgs, 0 if occ_mode else 1e-4)
    if budget:
        print(f"density form, {B * R} rays: {fwd_bad} forward / {int(bad.sum())} total rays "
              f"outside the tolerance (budget {budget})")
    assert int(bad.sum()) <= budget, (int(bad.sum()), budget, np.nonzero(bad)[0][:10])


def test_freq_encoding():
    from sailor_amd import VoxelEncoding as VE
    rng = np.random.default_rng(1)
    x = rng.uniform(-1, 1, (4097, 3)).astype(np.float32)
    y = VE.freq_encode_forward(torch.from_numpy(x).cuda(), 10, 0).cpu().numpy()
    y64 = orc.freq_f64(x, 10)
    # column order of the reference: [x, sin f0, cos f0, sin f1, ...]; the
    # float32 argument 2^f*x (f <= 9, |x| <= 1) carries up to 2^9 * 2^-24 * 2 of
    # rounding error, so the float64 closed form agrees to ~6e-5
    np.testing.assert_allclose(y, y64, atol=1e-4)
    np.testing.assert_allclose(y, orc.freq_fwd(x, 10), atol=2e-6)
    gy = rng.standard_normal(y.shape).astype(np.float32)
    gx = VE.freq_encode_backward(torch.from_numpy(gy).cuda(), torch.from_numpy(x).cuda(),
                                 torch.from_numpy(y).cuda(), 10, 0).cpu().numpy()
    np.testing.assert_allclose(gx, orc.freq_bwd(gy, y, 3, 10), rtol=1e-5, atol=1e-4)


def test_autograd_volume_render_matches_torch_formulation():
    """volume_render (HIP fwd+bwd of the training compositing kernel) vs
    float64 autograd through the tensor restatement of the reference's
    PyTorch occupancy compositing (oracle/composite_ref.py) on rays with no
    early termination."""
    from oracle import composite_ref as cr
    from sailor_amd.voxel_encoding_helper import volume_render
    rng = np.random.default_rng(5)
    B, R, S, F = 1, 256, 24, 3
    t = np.sort(rng.uniform(0.5, 4.0, (B, R, S)).astype(np.float32), -1)
    occ = rng.uniform(0.0, 0.3, (B, R, S, 1)).astype(np.float32)
    feats = rng.standard_normal((B, R, S, F)).astype(np.float32)
    order = np.tile(np.arange(S, dtype=np.int32), (B, R, 1))
    tc = torch.from_numpy(t).cuda()
    oc = torch.from_numpy(occ).cuda().requires_grad_(True)
    fc = torch.from_numpy(feats).cuda().requires_grad_(True)
    outs = volume_render(tc, torch.from_numpy(order).cuda(), oc, fc, 0.0, True)
    loss = outs[0].square().sum() + outs[1].sum()
    g_o, g_f = torch.autograd.grad(loss, [oc, fc])
    oc2 = torch.from_numpy(occ).double().requires_grad_(True)
    fc2 = torch.from_numpy(feats).double().requires_grad_(True)
    ref = cr.render(torch.from_numpy(t).double(), oc2, fc2, occ_mode=True)
    loss2 = ref[0].square().sum() + ref[1].sum()
    r_o, r_f = torch.autograd.grad(loss2, [oc2, fc2])
    # the training kernel uses 1 - alpha, the tensor form 1 - alpha + 1e-6:
    # relative differences ~1e-5
    np.testing.assert_allclose(outs[0].detach().cpu().numpy(), ref[0].detach().numpy(),
                               rtol=1e-4, atol=1e-4)
    np.testing.assert_allclose(g_f.cpu().numpy(), r_f.numpy(), rtol=1e-3, atol=1e-4)
    np.testing.assert_allclose(g_o.cpu().numpy(), r_o.numpy(), rtol=1e-3, atol=2e-3)
