"""Seeded synthetic inputs for the FastNerf inference surface (packed half
parameters with Xavier-like scales, per-view sample features, sorted depths
with dropped rays and skipped samples)."""
import numpy as np

from oracle import fastnerf_ref as fr


def _layers_params(rng, layers):
    """Packed [W (out x in) | b (out)] blocks in order, as float16."""
    parts = []
    for n_in, n_out, bias in layers:
        a = np.sqrt(6.0 / (n_in + n_out))
        parts.append(rng.uniform(-a, a, n_out * n_in))
        if bias:
            parts.append(rng.normal(0, 0.05, n_out))
    return np.concatenate(parts).astype(np.float16)


def params(seed=11):
    rng = np.random.default_rng(seed)
    pd = _layers_params(rng, [(18, 64, 1)] + [(64, 64, 1)] * 3 + [(64, 64, 1), (64, 64, 1),
                                                                   (64, 1, 1), (18, 16, 1),
                                                                   (16, 8, 1)])
    stack = [(30, 32, 0)] * 3 + [(32, 30, 0), (30, 32, 1), (32, 30, 1)]
    pc = _layers_params(rng, stack + stack + [(30, 64, 1), (64, 32, 1), (32, 3, 1)])
    pb = _layers_params(rng, [(40, 64, 1), (64, 64, 1), (64, 32, 1), (62, 64, 1), (64, 32, 1),
                              (32, 3, 1)])
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
