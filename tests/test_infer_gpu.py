"""GPU parity of the FastNerf inference surface (sailor_amd.InferNerf, the
drop-in for the reference's InferNerf module) against the float64
restatement oracle/fastnerf_ref.py on seeded inputs with the reference's
packed half parameters.

precision="fp32" (vector ALUs): fp32 arithmetic vs float64 on the same half
parameters, |d| <= 1e-4 + 1e-4 |ref| (NaN colour features where the
reference's epsilon-free normalisation gives 0/0: rays without samples).

precision="fp16" (default; density MLPs on fp16 MFMA): against the
restatement with the density layers' inputs rounded to half
(act_dtype=float16), |d| <= 1e-3 + 1e-3 |ref|: the fp32 and float64 sums
can round an activation to neighbouring halves (2^-11 relative), which moves
the occupancies by ~1e-4.  Against the unrounded float64 values the colour error is
that of half inputs (checked <= 2e-2), and the colour features are nearer
the rounded restatement than the float64 one (it is the one matched).  The
colour path of this form (Hydra stacks and RGB head on fp16 MFMA) rounds
the inputs of those products to half as well."""
import numpy as np
import pytest
import torch

from oracle import fastnerf_ref as fr
from tests import fastnerf_cases as fc

pytestmark = pytest.mark.gpu


def _cu(a, dtype=torch.float32):
    return torch.from_numpy(np.ascontiguousarray(a)).to("cuda", dtype)


@pytest.mark.parametrize("B,R,S,Vg,Vt", [(2, 64, 16, 4, 4), (1, 45, 24, 6, 3), (1, 8, 8, 1, 1)])
def test_infer_nerf_matches_float64(B, R, S, Vg, Vt):
    from sailor_amd import InferNerf
    pd, pc, _ = fc.params()
    t, z, ps, geo, rgb, frgb, dirs = fc.nerf_inputs(B, R, S, Vg, Vt)
    d, c, f = InferNerf.infer_nerf(_cu(t), _cu(z).view(-1, 1), _cu(ps).view(-1, 1), _cu(geo),
                                   _cu(rgb), _cu(frgb), _cu(dirs), _cu(pd, torch.float16),
                                   _cu(pc, torch.float16), B, R, S, Vg, Vt, 0,
                                   precision="fp32")
    torch.cuda.synchronize()
    rd, rc, rf = fr.infer_nerf(t, z, ps, geo, rgb, frgb, dirs, pd, pc, B, R, S, Vg, Vt)
    for got, ref in ((d, rd), (c, rc), (f, rf)):
        np.testing.assert_allclose(got.double().cpu().numpy(), ref.numpy(), rtol=1e-4,
                                   atol=1e-4, equal_nan=True)
    # rays without samples: 0/0 in the normalisation gives NaN features; the
    # colour head's max(x, 0) turns them into a finite colour, as in the
    # reference
    dropped = t[..., -1] < 0
    assert np.isnan(f.cpu().numpy()[dropped]).all()
    assert not np.isnan(f.cpu().numpy()[~dropped]).any() and np.isfinite(c.cpu().numpy()).all()


def test_infer_density_matches_float64():
    from sailor_amd import InferNerf
    pd, _, _ = fc.params()
    B, N, V = 2, 777, 4
    _, z, ps, geo, _, _, _ = fc.nerf_inputs(B, N, 1, V, 1, seed=9)
    out = InferNerf.infer_density(_cu(z).view(-1, 1), _cu(ps).view(-1, 1), _cu(geo),
                                  _cu(pd, torch.float16), B, N, V, 0, precision="fp32")
    torch.cuda.synchronize()
    ref = fr.infer_density(z, ps, geo, pd, B, N, V)
    assert out.shape == (B, 1, N)
    np.testing.assert_allclose(out.double().cpu().numpy(), ref.numpy(), rtol=1e-4, atol=1e-5)


@pytest.mark.parametrize("B,R,S,Vg,Vt", [(2, 64, 16, 4, 4), (1, 45, 24, 6, 3), (1, 8, 8, 1, 1),
                                         (1, 30, 8, 2, 2), (1, 12, 8, 9, 2)])
def test_infer_nerf_fp16_mfma(B, R, S, Vg, Vt):
    """Default precision: 1, 2, 4 and 8 view slots on MFMA, 9 views on the
    fp32 fallback."""
    from sailor_amd import InferNerf
    pd, pc, _ = fc.params()
    t, z, ps, geo, rgb, frgb, dirs = fc.nerf_inputs(B, R, S, Vg, Vt)
    d, c, f = InferNerf.infer_nerf(_cu(t), _cu(z).view(-1, 1), _cu(ps).view(-1, 1), _cu(geo),
                                   _cu(rgb), _cu(frgb), _cu(dirs), _cu(pd, torch.float16),
                                   _cu(pc, torch.float16), B, R, S, Vg, Vt, 0)
    torch.cuda.synchronize()
    act = torch.float16 if Vg <= 8 else None
    rd, rc, rf = fr.infer_nerf(t, z, ps, geo, rgb, frgb, dirs, pd, pc, B, R, S, Vg, Vt,
                               act_dtype=act)
    tol = 1e-3 if act is not None else 1e-4
    for got, ref in ((d, rd), (c, rc), (f, rf)):
        np.testing.assert_allclose(got.double().cpu().numpy(), ref.numpy(), rtol=tol, atol=tol,
                                   equal_nan=True)
    if act is not None:
        _, rc64, rf64 = fr.infer_nerf(t, z, ps, geo, rgb, frgb, dirs, pd, pc, B, R, S, Vg, Vt)
        e64 = np.nanmax(np.abs(c.double().cpu().numpy() - rc64.numpy()))
        f64 = np.nanmax(np.abs(f.double().cpu().numpy() - rf64.numpy()))
        f16 = np.nanmax(np.abs(f.double().cpu().numpy() - rf.numpy()))
        print(f"colour vs float64 {e64:.2e}; features vs float64 {f64:.2e}, vs the rounded "
              f"restatement {f16:.2e}")
        assert e64 <= 2e-2 and f16 <= f64


def test_infer_density_fp16_mfma():
    from sailor_amd import InferNerf
    pd, _, _ = fc.params()
    for B, N, V in ((2, 777, 4), (1, 301, 3), (1, 64, 8)):
        _, z, ps, geo, _, _, _ = fc.nerf_inputs(B, N, 1, V, 1, seed=9)
        out = InferNerf.infer_density(_cu(z).view(-1, 1), _cu(ps).view(-1, 1), _cu(geo),
                                      _cu(pd, torch.float16), B, N, V, 0)
        torch.cuda.synchronize()
        ref = fr.infer_density(z, ps, geo, pd, B, N, V, act_dtype=torch.float16)
        np.testing.assert_allclose(out.double().cpu().numpy(), ref.numpy(), rtol=1e-3,
                                   atol=1e-4)


def test_infer_blending_matches_float64():
    from sailor_amd import InferNerf
    _, _, pb = fc.params()
    rng = np.random.default_rng(4)
    nf = rng.normal(0, 1, (1000, 40)).astype(np.float32)
    rf = rng.normal(0, 1, (1000, 30)).astype(np.float32)
    out = InferNerf.infer_blending(_cu(nf), _cu(rf), _cu(pb, torch.float16), 0)
    torch.cuda.synchronize()
    ref = fr.infer_blending(nf, rf, pb)
    np.testing.assert_allclose(out.double().cpu().numpy(), ref.numpy(), rtol=1e-4, atol=1e-5)
