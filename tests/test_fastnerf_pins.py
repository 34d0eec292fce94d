"""Pins oracle/fastnerf_ref.py against checks that do not share its code
(CPU only):

* Every MLP of c_lib/FastNerf/src/infer_mlp.cu as a torch.nn.Sequential of
  nn.Linear / nn.ReLU / nn.Sigmoid modules whose weights are loaded from the
  packed half vector at offsets written out here from the kernels'
  constants (not through the restatement's unpackers):
    - feature MLP 18->64->64->64->64, ReLU after every layer, input order
      [z, tpsdf, geo (16)]                         infer_mlp.cu:150-186
    - geometry-feature MLP 18->16->8 at 22081, ReLU after both  :182-185
    - occupancy MLP 64->64->64->1 at 13696, ReLU, ReLU, sigmoid  :205-239
    - RGB head 30->64->32->3 at 11644 of the colour vector, ReLU, ReLU,
      sigmoid                                      :427-432
* Hydra stacks (:242-445, the semantics of SRONet/aggregation.py:101-171)
  against two closed forms per stack:
    - q = k = v = identity (30 of 32 rows), fc = its transpose, zero
      feed-forward: x'_v = x_v + x^_v * sum_u x^_u * x_u   (x^ = x / |x|)
    - v = 0 (attention output 0), w1 = identity, b1 = -0.1, w2 = 2 identity,
      b2 = 0.05: x' = x + 2 relu(x - 0.1) + 0.05
* Planted mistakes: copies of the restatement with one ReLU dropped, the
  occupancy sigmoid dropped, the attention residual dropped, the
  feed-forward residual dropped, or z and tpsdf swapped each fail one of
  the checks above.
"""
import inspect

import numpy as np
import pytest
import torch
from torch import nn

from oracle import fastnerf_ref as fr
from tests import fastnerf_cases as fc

D = torch.float64


def _seq(vec, offset, layers, acts):
    """nn.Sequential of Linear layers (n_in, n_out) read from a packed
    vector starting at `offset` (W [out][in] row-major, then b [out]), each
    followed by the module in `acts`."""
    mods, o = [], offset
    for (n_in, n_out), act in zip(layers, acts):
        lin = nn.Linear(n_in, n_out).to(D)
        with torch.no_grad():
            lin.weight.copy_(vec[o:o + n_in * n_out].reshape(n_out, n_in))
            o += n_in * n_out
            lin.bias.copy_(vec[o:o + n_out])
            o += n_out
        mods += [lin, act]
    return nn.Sequential(*mods), o


def _vec(p):
    return torch.from_numpy(np.asarray(p, np.float16).astype(np.float64))


def _density_inputs(n=64, seed=2):
    rng = np.random.default_rng(seed)
    z = rng.uniform(-1, 1, n).astype(np.float32)
    ps = rng.uniform(-1, 1, n).astype(np.float32)
    geo = rng.normal(0, 1, (n, 16)).astype(np.float32)
    x = torch.cat([torch.tensor(z, dtype=D)[:, None], torch.tensor(ps, dtype=D)[:, None],
                   torch.tensor(geo, dtype=D)], 1)
    return z, ps, geo, x


def check_density_layers(ref):
    """Feature, geometry and occupancy MLPs of `ref` vs nn modules."""
    pd, _, _ = fc.params()
    v = _vec(pd)
    feat, end = _seq(v, 0, [(18, 64), (64, 64), (64, 64), (64, 64)], [nn.ReLU()] * 4)
    assert end == 13696
    occ, end = _seq(v, 13696, [(64, 64), (64, 64), (64, 1)], [nn.ReLU(), nn.ReLU(), nn.Sigmoid()])
    assert end == 22081
    geo_m, end = _seq(v, 22081, [(18, 16), (16, 8)], [nn.ReLU(), nn.ReLU()])
    assert end == 22521 == pd.size
    z, ps, geo, x = _density_inputs()
    with torch.no_grad():
        h, g = ref.density_feats(z, ps, geo, pd)
        np.testing.assert_allclose(h.numpy(), feat(x).numpy(), rtol=1e-12, atol=1e-12)
        np.testing.assert_allclose(g.numpy(), geo_m(x).numpy(), rtol=1e-12, atol=1e-12)
        m = torch.randn(32, 64, dtype=D, generator=torch.Generator().manual_seed(4))
        np.testing.assert_allclose(ref.occupancy(m, pd).numpy(), occ(m)[:, 0].numpy(),
                                   rtol=1e-12, atol=1e-12)


def _stack_params(q, k, v, fc_, w1, b1, w2, b2):
    """Packed half vector of one Hydra stack (5822 values) in the kernel's
    order: q, k, v [32][30], fc [30][32], w1 [32][30], b1 [32], w2 [30][32],
    b2 [30]."""
    parts = [q, k, v, fc_, w1, b1, w2, b2]
    out = np.concatenate([np.asarray(p, np.float64).reshape(-1) for p in parts])
    assert out.size == 5822
    return out


def _color_vec(stack, rgb_tail):
    return np.concatenate([stack, stack, rgb_tail]).astype(np.float16)


def check_hydra_closed_forms(ref):
    """Two stacks with the closed-form parameters above; colour features of
    token 0 compared with numpy."""
    rng = np.random.default_rng(21)
    tok = rng.uniform(-1, 1, (5, 4, 30))
    tok[np.abs(tok) < 0.05] = 0.3           # keep x - 0.1 away from the kink
    E = np.zeros((32, 30))
    E[:30, :30] = np.eye(30)
    Z30, Z32 = np.zeros((30, 32)), np.zeros((32, 30))
    _, pc, _ = fc.params()
    rgb_tail = np.asarray(pc, np.float64)[11644:]
    # identity projections, zero feed-forward
    st = _stack_params(E, E, E, E.T, Z32, np.zeros(32), Z30, np.zeros(30))
    f, _ = ref.hydra_rgb(torch.tensor(tok), _color_vec(st, rgb_tail))
    x = tok.copy()
    for _ in range(2):
        xh = x / np.linalg.norm(x, axis=-1, keepdims=True)
        x = x + xh * (xh * x).sum(1, keepdims=True)
    np.testing.assert_allclose(f.numpy(), x[:, 0], rtol=1e-10, atol=1e-10)
    # v = 0, feed-forward x + 2 relu(x - 0.1) + 0.05
    rq = rng.normal(0, 0.3, (32, 30))
    st = _stack_params(rq, rq[::-1].copy(), np.zeros((32, 30)), rng.normal(0, 0.3, (30, 32)),
                       E, np.full(32, -0.1), 2 * E.T, np.full(30, 0.05))
    f, _ = ref.hydra_rgb(torch.tensor(tok), _color_vec(st, rgb_tail))
    x = tok.copy()
    for _ in range(2):
        # b1 = -0.1 and b2 = 0.05 are rounded to half in the packed vector
        x = x + 2 * np.maximum(x + float(np.float16(-0.1)), 0) + float(np.float16(0.05))
    np.testing.assert_allclose(f.numpy(), x[:, 0], rtol=1e-10, atol=1e-10)


def check_rgb_head(ref):
    """RGB head on token 0 after the stacks (zero stacks: tokens pass
    through unchanged) vs nn.Sequential."""
    _, pc, _ = fc.params()
    v = _vec(pc)
    head, end = _seq(v, 11644, [(30, 64), (64, 32), (32, 3)], [nn.ReLU(), nn.ReLU(), nn.Sigmoid()])
    assert end == 15807 == pc.size
    zero = _stack_params(*[np.zeros(s) for s in ((32, 30),) * 3 + ((30, 32), (32, 30), (32,),
                                                                  (30, 32), (30,))])
    # v = 0 and a zero feed-forward leave every token as it is
    rng = np.random.default_rng(22)
    q = rng.normal(0, 0.3, (32, 30))
    st = zero.copy()
    st[:960] = q.reshape(-1)
    st[960:1920] = q.reshape(-1)
    tok = torch.tensor(rng.normal(0, 1, (7, 4, 30)))
    f, rgb = ref.hydra_rgb(tok, _color_vec(st, np.asarray(pc, np.float64)[11644:]))
    np.testing.assert_allclose(f.numpy(), tok[:, 0].numpy(), rtol=1e-12, atol=1e-12)
    with torch.no_grad():
        np.testing.assert_allclose(rgb.numpy(), head(tok[:, 0]).numpy(), rtol=1e-12, atol=1e-12)


def test_density_mlps_match_nn_modules():
    check_density_layers(fr)


def test_hydra_stacks_match_closed_forms():
    check_hydra_closed_forms(fr)


def test_rgb_head_matches_nn_sequential():
    check_rgb_head(fr)


def _mutant(func_name, old, new):
    """A copy of the restatement module whose function `func_name` has the
    source text `old` replaced by `new` (planted on the CPU, in the oracle)."""
    import types
    src = inspect.getsource(getattr(fr, func_name))
    if old not in src:
        raise LookupError(f"{func_name} has no {old!r}")   # not an AssertionError
    mod = types.ModuleType("fastnerf_mutant")
    mod.__dict__.update(fr.__dict__)
    exec(compile(src.replace(old, new, 1), "<mutant>", "exec"), mod.__dict__)
    return mod


MUTANTS = {
    "feature_relu_dropped": ("density_feats", "h = _relu(_lin(_round(h, act_dtype), wb))",
                             "h = _lin(_round(h, act_dtype), wb)", check_density_layers),
    "geometry_relu_dropped": ("density_feats", "g = _relu(_lin(_round(_relu(",
                              "g = (_lin(_round(_relu(", check_density_layers),
    "z_tpsdf_swapped": ("density_feats", "[_t(z, dtype)[:, None], _t(tpsdf, dtype)[:, None]",
                        "[_t(tpsdf, dtype)[:, None], _t(z, dtype)[:, None]",
                        check_density_layers),
    "occupancy_sigmoid_dropped": ("occupancy", "torch.sigmoid(_lin(h, sigma[2]))",
                                  "(_lin(h, sigma[2]))", check_density_layers),
    "attention_residual_dropped": ("hydra_rgb", "res = x + out", "res = out",
                                   check_hydra_closed_forms),
    "feedforward_residual_dropped": ("hydra_rgb", "+ st[\"b2\"] + res", "+ st[\"b2\"]",
                                     check_hydra_closed_forms),
    "rgb_relu_dropped": ("hydra_rgb", "h = _relu(_lin(r(h), rgb[1]))", "h = _lin(r(h), rgb[1])",
                         check_rgb_head),
}


@pytest.mark.parametrize("name", sorted(MUTANTS))
def test_planted_mistakes_fail_the_pins(name):
    func, old, new, check = MUTANTS[name]
    mutant = _mutant(func, old, new)
    with pytest.raises(AssertionError):
        check(mutant)
