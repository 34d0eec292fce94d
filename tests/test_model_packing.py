"""CPU checks of the MFMA weight packing (sailor_amd/model.py) against the
fragment rules the fused kernel uses (sailor_amd/csrc/mlp_layout.h):
emulating the MFMA lane layouts, packed-A x kernel-built-B must equal W @ x."""
import numpy as np
import pytest

from sailor_amd import model as M


def kernel_b_index(prec, F, H, first, s, g, j):
    """Input index the kernel places in B-fragment slot (step s, lane group g,
    slot j) -- mlp.hip render_kernel layer lambdas (-1 = padding)."""
    cpl = F // 4
    if prec == "fp32":
        if first:
            fs = cpl
            if s < fs:
                return M.PE_DIMS + cpl * g + s
            q = 16 * g + (s - fs)
            return q if q < M.PE_DIMS else -1
        return 16 * (s >> 2) + 4 * g + (s & 3)
    if first:
        fs = F // 32
        if s < fs:
            return M.PE_DIMS + cpl * g + 8 * s + j
        q = 16 * g + 8 * (s - fs) + j
        return q if q < M.PE_DIMS else -1
    return 32 * s + 4 * g + j if j < 4 else 32 * s + 16 + 4 * g + (j - 4)


@pytest.mark.parametrize("prec", ["fp32", "bf16"])
@pytest.mark.parametrize("F,H", [(32, 128), (64, 256)])
@pytest.mark.parametrize("first", [True, False])
def test_packed_layer_times_fragments_equals_matvec(prec, F, H, first):
    rng = np.random.default_rng(0)
    K = M.PE_DIMS + F if first else H
    W = rng.standard_normal((H, K)).astype(np.float64)
    X = rng.standard_normal((K, 16)).astype(np.float64)  # 16 columns = one N-tile
    A = M.pack_layer(W.astype(np.float32), F, H, prec, first).astype(np.float64)
    KS, _, MT, EPL = A.shape
    D = np.zeros((H, 16))
    seen = []
    for s in range(KS):
        for lane in range(64):
            g, m = lane // 16, lane % 16
            for j in range(EPL):
                kk = kernel_b_index(prec, F, H, first, s, g, j)
                seen.append(kk)
                xval = X[kk] if kk >= 0 else np.zeros(16)
                for mt in range(MT):
                    # A[m][k-row] * B[k-row][n] summed over the k-rows of the step
                    D[16 * mt + m] += A[s, lane, mt, j] * xval
    np.testing.assert_allclose(D, W @ X, rtol=1e-5, atol=1e-4)
    # every input appears exactly once per output row
    vals = [k for k in seen[::1] if k >= 0]
    counts = np.bincount(vals, minlength=K) // 16  # 16 m-lanes per k-row
    assert (counts == 1).all()


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
