"""Pins oracle/composite_ref.py, the tensor restatement of the reference's
PyTorch compositing (voxel_encoding_helper.py:132-186), on the CPU:
closed forms (constant sigma on equal spacing, constant occupancy), the
zero weight of the last density-form sample, the valid mask, an explicit
per-sample loop, and central finite differences of its autograd."""
import numpy as np
import pytest
import torch

from oracle import composite_ref as cr

D = torch.float64


def _rows(occ_mode, B=2, R=5, S=12, F=3, seed=0):
    g = torch.Generator().manual_seed(seed)
    t = torch.sort(torch.rand((B, R, S), generator=g, dtype=D) * 3 + 0.5, -1)[0]
    if occ_mode:
        s = torch.rand((B, R, S), generator=g, dtype=D) * 0.6
    else:
        s = torch.rand((B, R, S), generator=g, dtype=D) * 8
    f = torch.randn((B, R, S, F), generator=g, dtype=D)
    return t, s, f


def test_density_constant_sigma_closed_form():
    S, delta, sig = 10, 0.05, 7.0
    t = (1.0 + delta * torch.arange(S, dtype=D)).reshape(1, 1, S)
    s = torch.full((1, 1, S), sig, dtype=D)
    f = torch.ones((1, 1, S, 1), dtype=D)
    of, od, ow, oa = cr.render(t, s, f, occ_mode=False)
    q = np.exp(-delta * sig)
    w = (1 - q) * q ** np.arange(S)
    w[-1] = 0.0                                  # the last interval is 0
    assert abs(ow.item() - (1 - q ** (S - 1))) < 1e-14
    assert abs(od.item() - (w * t.numpy().ravel()).sum()) < 1e-13
    assert abs(oa.item() - (w * np.where(np.arange(S) < S - 1, 1 - q, 0.0)).sum()) < 1e-14
    assert abs(of.item() - ow.item()) < 1e-14


def test_occupancy_constant_alpha_geometric_series():
    S, a = 16, 0.3
    t = torch.linspace(1, 2, S, dtype=D).reshape(1, 1, S)
    s = torch.full((1, 1, S), a, dtype=D)
    f = torch.full((1, 1, S, 2), 2.0, dtype=D)
    of, od, ow, oa = cr.render(t, s, f, occ_mode=True)
    c = 1 - a + 1e-6
    assert abs(ow.item() - a * (1 - c ** S) / (1 - c)) < 1e-13
    assert abs(oa.item() - a * ow.item()) < 1e-13
    np.testing.assert_allclose(of.numpy().ravel(), [2 * ow.item()] * 2, rtol=1e-14)


@pytest.mark.parametrize("occ_mode", [False, True])
def test_matches_per_sample_loop_and_valid_mask(occ_mode):
    t, s, f = _rows(occ_mode)
    valid = torch.tensor([[True, False, True, True, False], [False, True, True, True, True]])
    of, od, ow, oa = cr.render(t, s, f, occ_mode, valid)
    for b in range(2):
        for r in range(5):
            if not valid[b, r]:
                assert of[b, r].abs().max() == 0 and od[b, r] == 0 and ow[b, r] == 0
                continue
            lf, ld, lw, la = cr.render_loop(t[b, r], s[b, r], f[b, r], occ_mode)
            np.testing.assert_allclose(of[b, r].numpy(), lf, rtol=1e-12, atol=1e-14)
            np.testing.assert_allclose([od[b, r, 0], ow[b, r, 0], oa[b, r, 0]], [ld, lw, la],
                                       rtol=1e-12, atol=1e-14)


@pytest.mark.parametrize("occ_mode", [False, True])
def test_autograd_matches_central_differences(occ_mode):
    t, s, f = _rows(occ_mode, B=1, R=3, S=9, F=2, seed=3)
    g = torch.Generator().manual_seed(4)
    wts = [torch.randn(x, generator=g, dtype=D) for x in ((1, 3, 2), (1, 3, 1), (1, 3, 1),
                                                          (1, 3, 1))]

    def loss(s_, f_):
        return sum((o * w).sum() for o, w in zip(cr.render(t, s_, f_, occ_mode), wts))
    s1, f1 = s.clone().requires_grad_(), f.clone().requires_grad_()
    gs, gf = torch.autograd.grad(loss(s1, f1), [s1, f1])
    h = 1e-6
    for idx in [(0, 0, 0), (0, 1, 4), (0, 2, 7), (0, 2, 8)]:
        sp, sm = s.clone(), s.clone()
        sp[idx] += h
        sm[idx] -= h
        fd = (loss(sp, f) - loss(sm, f)) / (2 * h)
        assert abs(fd - gs[idx]) < 1e-7 * max(1.0, abs(float(gs[idx]))), (idx, fd, gs[idx])
    for idx in [(0, 0, 0, 1), (0, 2, 5, 0)]:
        fp, fm = f.clone(), f.clone()
        fp[idx] += h
        fm[idx] -= h
        fd = (loss(s, fp) - loss(s, fm)) / (2 * h)
        assert abs(fd - gf[idx]) < 1e-7 * max(1.0, abs(float(gf[idx])))
