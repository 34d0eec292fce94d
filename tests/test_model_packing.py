"""CPU checks of the MFMA weight packing (sailor_amd/model.py) against the
fragment rules of the fused kernel (sailor_amd/csrc/mlp_layout.h):
emulating the MFMA lane layouts, packed-A x kernel-read-B must equal W @ x."""
import numpy as np
import pytest

from sailor_amd import model as M


def unit_of(H, w, g, mt, i):
    """Output unit held by accumulator element i of m-tile mt, lane group g,
    wave w (mlp.hip unit_of)."""
    return (H // 4) * w + 4 * (H // 64) * g + 4 * mt + i


def test_unit_map_gives_each_lane_group_consecutive_units():
    for H in (128, 256):
        MTW = H // 64
        seen = []
        for w in range(4):
            for g in range(4):
                units = [unit_of(H, w, g, mt, i) for mt in range(MTW) for i in range(4)]
                assert units == list(range(units[0], units[0] + 4 * MTW))
                seen += units
        assert sorted(seen) == list(range(H))


@pytest.mark.parametrize("H,CE", [(128, 4), (256, 4), (128, 8), (256, 8)])
def test_swizzle_is_a_permutation_of_each_column(H, CE):
    """act_chunk(col, c) = col*H + (c ^ (col & 15)) * CE keeps every chunk of
    a column inside that column and hits each slot once."""
    nchunk = H // CE
    for col in range(64):
        phys = [col * H + (c ^ (col & 15)) * CE for c in range(nchunk)]
        assert sorted(phys) == [col * H + CE * c for c in range(nchunk)]


def unit_of32(H, w, mt, h, r):
    """Output unit held by accumulator r of 32x32 m-tile mt for lanes with
    l/32 = h, wave w (mlp.hip unit_of32)."""
    return (H // 4) * w + 32 * mt + 16 * h + r


@pytest.mark.parametrize("H,K", [(128, 96), (128, 128), (256, 128), (256, 256)])
def test_packed_fp32_layer_times_fragments_equals_matvec(H, K):
    """16x16x4 MFMA: lane l of step s multiplies A[l%16][k] by B[k][l%16]
    with k = 16(s/4) + 4(l/16) + s%4; D row 4g+i of m-tile mt is unit_of."""
    rng = np.random.default_rng(0)
    W = rng.standard_normal((H, K)).astype(np.float32)
    X = rng.standard_normal((K, 16))                     # 16 columns = one N-tile
    A = M.pack_layer(W, "fp32").astype(np.float64)       # [waves, steps, 64, MTW, 1]
    waves, KS, _, MTW, _ = A.shape
    D = np.zeros((H, 16))
    count = np.zeros(K, int)
    for w in range(waves):
        for s in range(KS):
            for lane in range(64):
                g, m = lane // 16, lane % 16
                kk = 16 * (s >> 2) + 4 * g + (s & 3)
                if w == 0 and m == 0:
                    count[kk] += 1
                for mt in range(MTW):
                    D[unit_of(H, w, m // 4, mt, m % 4)] += A[w, s, lane, mt, 0] * X[kk]
    np.testing.assert_allclose(D, W.astype(np.float64) @ X, rtol=1e-5, atol=1e-4)
    assert (count == 1).all()  # each input is reduced exactly once per row


@pytest.mark.parametrize("H,K", [(128, 96), (128, 128), (256, 128), (256, 256)])
def test_packed_16bit_layer_times_fragments_equals_matvec(H, K):
    """32x32x16 MFMA (lane layout checked on gfx950): lane l of step u holds
    A[l%32][16u + 8(l/32) + j]; D row m = 8q + 4h + i of m-tile mt lands in
    accumulator 4q + i of lanes with l/32 = h, which the kernel writes as
    unit_of32(w, mt, h, 4q + i)."""
    rng = np.random.default_rng(0)
    W = rng.standard_normal((H, K)).astype(np.float32)
    X = rng.standard_normal((K, 32))                     # 32 columns = one N-tile
    A = M.pack_layer(W, "bf16").astype(np.float64)       # [waves, steps, 64, M32, 8]
    waves, U, _, M32, _ = A.shape
    assert M32 == H // 128
    D = np.zeros((H, 32))
    count = np.zeros(K, int)
    for w in range(waves):
        for u in range(U):
            for lane in range(64):
                m, h = lane % 32, lane // 32
                for j in range(8):
                    kk = 16 * u + 8 * h + j
                    if w == 0 and m == 0:
                        count[kk] += 1
                    for mt in range(M32):
                        unit = unit_of32(H, w, mt, (m // 4) % 2, 4 * (m // 8) + m % 4)
                        D[unit] += A[w, u, lane, mt, j] * X[kk]
    np.testing.assert_allclose(D, W.astype(np.float64) @ X, rtol=1e-5, atol=1e-4)
    assert (count == 1).all()


def test_first_layer_natural_order():
    rng = np.random.default_rng(1)
    F, H = 64, 256
    W0 = rng.standard_normal((H, M.PE_DIMS + F)).astype(np.float32)
    Wn = M.natural_first_layer(W0, F)
    pe = rng.standard_normal(M.PE_DIMS)
    feat = rng.standard_normal(F)
    x_ref = np.concatenate([pe, feat])            # model input order
    x_ker = np.concatenate([feat, pe, [0.0]])     # kernel LDS order
    np.testing.assert_allclose(Wn @ x_ker, W0 @ x_ref, rtol=1e-5, atol=1e-4)


def test_small_buffer_matches_kernel_layout():
    from sailor_amd import _lib
    for F, H in [(32, 128), (64, 256)]:
        w = M.NerfWeights.init(F, H)
        n = w.small_buffer().size
        assert n == _lib.lib().sailor_render_small_weights_floats(F, H)


def test_reference_model_runs_on_cpu():
    """The float64 restatement produces sane outputs on a tiny synthetic ray set."""
    import torch
    from oracle import nerf_ref
    w = M.NerfWeights.init(32, 128)
    rng = np.random.default_rng(1)
    B, R, S, V = 1, 5, 8, 4
    t = np.sort(rng.uniform(2.0, 3.0, (B, R, S)), -1)
    t[0, 0] = -1.0  # dropped ray
    rays = np.zeros((B, R, 6))
    rays[..., 2] = 2.5
    rays[..., 5] = -1.0
    K = np.tile(np.array([[100., 0, 64], [0, 100., 64], [0, 0, 1]]), (B, V, 1, 1))
    RT = np.tile(np.concatenate([np.eye(3), np.array([[0], [0], [3.0]])], 1), (B, V, 1, 1))
    fm = rng.standard_normal((B, V, 32, 16, 16))
    bbox = np.array([[[-1, 1], [-1, 1], [-1, 1]]], np.float64)
    zdir = np.array([[0, 0, 1.0]])
    rgb, dep, ws, asum, occ = nerf_ref.render_rays(w, t, rays, K, RT, fm, bbox, zdir, (128, 128))
    assert torch.all(rgb[0, 0] == 0) and torch.all(rgb[0, 1:] > 0)
    assert torch.all((ws[0, 1:] > 0) & (ws[0, 1:] <= 1 + 1e-9))


def test_split_small_inverts_small_buffer():
    from sailor_amd.model import NerfWeights
    w = NerfWeights.init(32, 128, perturb_std=0.01)
    buf = w.small_buffer()
    sp = w.split_small(buf)
    np.testing.assert_array_equal(sp["sigma_w"], w.sigma_w)
    assert sp["sigma_b"] == np.float32(w.sigma_b)
    for a, b in zip(sp["hydra"], w.hydra):
        for k in b:
            np.testing.assert_array_equal(a[k], b[k])
    for (a, b), (c, d) in zip(sp["rgb"], w.rgb):
        np.testing.assert_array_equal(a, c)
        np.testing.assert_array_equal(b, d)


def test_device_gather_packing_matches_pack_layer():
    """RadianceField repacks its parameters with one gather per layer
    (train._pack_index): identical to model.pack_layer for every precision
    and layer shape of the benchmark models."""
    import numpy as np
    import torch
    from sailor_amd.model import NerfWeights, natural_first_layer, pack_layer
    from sailor_amd.train import _pack_index
    for F, H in ((32, 128), (64, 256)):
        w = NerfWeights.init(F, H)
        for prec in ("fp32", "bf16"):
            first = natural_first_layer(w.dense_w[0], F)
            got = torch.from_numpy(first).reshape(-1)[_pack_index(prec, H, F + 64)].numpy()
            np.testing.assert_array_equal(got, pack_layer(first, prec).reshape(-1))
            Wt = np.ascontiguousarray(w.dense_w[2].T)
            got = torch.from_numpy(Wt).reshape(-1)[_pack_index(prec, H, H)].numpy()
            np.testing.assert_array_equal(got, pack_layer(Wt, prec).reshape(-1))
