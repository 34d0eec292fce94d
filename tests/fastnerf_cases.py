"""Seeded synthetic inputs for the FastNerf inference surface (packed half
parameters with Xavier-like scales, per-view sample features, sorted depths
with dropped rays and skipped samples)."""
import numpy as np

from oracle import fastnerf_ref as fr


def params(seed=11):
    from sailor_amd.model import fastnerf_params
    pd, pc, pb = fastnerf_params(seed)
    assert pd.size == fr.DENSITY_PARAMS and pc.size == fr.COLOR_PARAMS
    assert pb.size == fr.BLEND_PARAMS
    return pd, pc, pb


def nerf_inputs(B, R, S, Vg, Vt, seed=5):
    rng = np.random.default_rng(seed)
    t = np.sort(rng.uniform(1.0, 3.0, (B, R, S)), -1).astype(np.float32)
    t[:, ::7] = -1.0                                 # rays without samples
    holes = rng.random((B, R, S)) < 0.1              # skipped samples
    holes[..., -1] = False
    t[holes & (t[..., -1:] > 0)] = -1.0
    n_g, n_t = B * Vg * R * S, B * Vt * R * S
    z = rng.uniform(-1, 1, n_g).astype(np.float32)
    ps = rng.uniform(-1, 1, n_g).astype(np.float32)
    geo = rng.normal(0, 1, (n_g, 16)).astype(np.float32)
    rgb = rng.uniform(0, 1, (n_t, 3)).astype(np.float32)
    frgb = rng.normal(0, 1, (n_t, 16)).astype(np.float32)
    dirs = rng.normal(0, 1, (n_t, 3)).astype(np.float32)
    return t, z, ps, geo, rgb, frgb, dirs
