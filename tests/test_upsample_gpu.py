"""GPU parity of the GUI's sub-pixel upsampling (sailor_rays_interpolate,
sailor_forward_upsampling through sailor_amd.fastnerf) against the float64
restatement oracle/upsample_ref.py.

Tolerances: directions 1e-5 (fp32 3x3 inverses of K with entries ~600);
colours, precision "fp32": 2e-4 absolute: the bilinear features move with
the ~1e-5 pixel error of the fp32 projection times map gradients of order 1
per pixel, and the blending MLP passes that on; precision "fp16" (MFMA, the
default): 2e-3 absolute against the restatement with half-rounded layer
inputs: a feature that moves by 1e-5 can round to the neighbouring half
(2^-11 relative), which moves the weights by ~1e-4 -- 1e-3 at worst.  Sub-pixels whose nearest-mode tap lies within
1e-3 pixel of a rounding tie are left out of the colour comparison (fp32
and float64 may round those to different pixels); rays outside the depth
range must be exactly zero."""
import numpy as np
import pytest
import torch

from oracle import upsample_ref as ur
from tests import fastnerf_cases as fc
from tests import upsample_cases as uc

pytestmark = pytest.mark.gpu


def _cu(a):
    return torch.from_numpy(np.ascontiguousarray(a)).cuda()


def test_rays_interpolate_matches_float64():
    from sailor_amd.fastnerf import rays_interpolate
    cams = []
    for seed in (1, 2):
        c = uc.case(R=200, seed=seed)
        cams.append(c)
    d = np.stack([c["ray_dir"] for c in cams])
    K = np.stack([c["tar_K"] for c in cams])
    R = np.stack([c["tar_R"] for c in cams])
    K[1, 0, 0] = 700.0                                   # two different cameras
    got = rays_interpolate(_cu(d), 200, _cu(K)[:, None], _cu(R)[:, None])
    torch.cuda.synchronize()
    ref = ur.rays_interpolate(d, K, R).reshape(1, -1, 3)
    np.testing.assert_allclose(got.double().cpu().numpy(), ref.numpy(), atol=1e-5)


def _ties(tr, eps=1e-3):
    pix = tr["pixels"].numpy()                           # [taps, n, 4, 2]
    frac = pix - np.floor(pix)
    return np.any(np.abs(frac - 0.5) < eps, axis=(0, 3))  # [n, 4]


@pytest.mark.parametrize("seed,nei,prec", [(7, (0, 1), "fp32"), (8, (1, 0), "fp32"),
                                           (9, (3, 0), "fp32"), (7, (0, 1), "fp16"),
                                           (9, (3, 0), "fp16")])
def test_forward_upsampling_matches_float64(seed, nei, prec):
    from sailor_amd.fastnerf import forward_upsampling, to_channels_last
    c = uc.case(R=300, seed=seed)
    _, _, pb = fc.params()
    got = forward_upsampling(_cu(c["ray_ori"])[None], _cu(c["ray_dir"])[None],
                             _cu(c["o_d"]).reshape(1, -1, 1), _cu(c["o_c"])[None],
                             _cu(c["o_c_feat"])[None], _cu(c["tar_K"]), _cu(c["tar_R"]),
                             _cu(c["ks"]), _cu(c["rs"]), _cu(c["ts"]),
                             to_channels_last(_cu(c["rgb_feats"])),
                             to_channels_last(_cu(c["rgbs"])), _cu(c["depths"]),
                             torch.tensor(nei, device="cuda"), _cu(pb).half(), precision=prec)
    torch.cuda.synchronize()
    assert got.shape == (1, 300, 4, 3)
    got = got[0].double().cpu().numpy()
    tr = {}
    ref = ur.forward_upsampling(c["ray_ori"], c["ray_dir"], c["o_d"], c["o_c"], c["o_c_feat"],
                                c["tar_K"], c["tar_R"], c["ks"], c["rs"], c["ts"],
                                c["rgb_feats"], c["rgbs"], c["depths"], nei, pb, trace=tr,
                                act_dtype=torch.float16 if prec == "fp16" else None).numpy()
    valid = tr["valid"].numpy()
    other = np.setdiff1d(np.arange(300), valid)
    assert other.size >= 3 and np.all(got[other] == 0)
    keep = ~_ties(tr)
    assert keep.mean() > 0.95
    np.testing.assert_allclose(got[valid][keep], ref[valid][keep],
                               atol=2e-4 if prec == "fp32" else 2e-3)
    # the blend is not degenerate: neighbour colours and the coarse colour both matter
    w = tr["weights"].numpy()
    assert w.std() > 0.01 and (got[valid] > 0).mean() > 0.5
