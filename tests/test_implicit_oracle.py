"""The implicit-grid oracle (oracle/implicit.py) pinned without retyping it.

- Skip wiring: with every layer zero except the skip layer's encoding block, the MLP collapses to
  feature_linear(relu(W_pe PE(x) + b)), a closed form that fixes the concatenation order
  ([encoding, hidden], model/mlps.py:76-77).
- The MLP against torch: the float64 torch.nn.Linear chain of diver_amd.model.mlps.ImplicitMLP
  with the same weights (the library routine for each layer).
- The integral: per-corner evaluation (integrate_mlp) equals fp64.integrate_einsum over the dense
  grid of MLP values (dense_grid), two independent forms of integrator.py's algebra.
"""
import numpy as np
import torch

from oracle import fp64
from oracle.implicit import dense_grid, implicit_mlp, integrate_mlp
from diver_amd.model.mlps import ImplicitMLP


def _state(m):
    return {k: v.detach().double().numpy() for k, v in m.state_dict().items()}


def test_skip_wiring_closed_form():
    torch.manual_seed(0)
    m = ImplicitMLP(4, 16, [2], 8, 3).double()
    sd = _state(m)
    for k in sd:
        if k.startswith('pts_linears.'):
            sd[k] = np.zeros_like(sd[k])
    rng = np.random.default_rng(1)
    w_pe = rng.normal(size=(16, 21))
    b = rng.normal(size=16)
    sd['pts_linears.2.weight'][:, :21] = w_pe
    sd['pts_linears.2.bias'] = b
    for i in (3,):
        sd[f'pts_linears.{i}.weight'] = np.eye(16)
    x = rng.uniform(-1, 1, (50, 3))
    got = implicit_mlp(sd, [2], 3, x)
    h = np.maximum(np.maximum(fp64.positional_encoding(x, 3) @ w_pe.T + b, 0) @ np.eye(16).T, 0)
    expect = h @ sd['feature_linear.weight'].T + sd['feature_linear.bias']
    np.testing.assert_allclose(got, expect, rtol=1e-12, atol=1e-12)


def test_mlp_matches_torch_linear_chain():
    torch.manual_seed(2)
    m = ImplicitMLP(5, 32, [3], 12, 4).double()
    x = np.random.default_rng(3).uniform(-1, 1, (40, 3))
    ref = m(torch.from_numpy(x)).detach().numpy()
    np.testing.assert_allclose(implicit_mlp(_state(m), [3], 4, x), ref, rtol=1e-12, atol=1e-12)


def test_integral_per_corner_equals_dense_einsum():
    torch.manual_seed(4)
    m = ImplicitMLP(3, 16, [1], 8, 2).double()
    sd = _state(m)
    N = 6  # corners per axis (voxel_num 5)
    rng = np.random.default_rng(5)
    cell = rng.integers(0, N - 1, (60, 3))
    ci = (cell + rng.uniform(0, 1, (60, 3))).astype(np.float32)
    co = (cell + rng.uniform(0, 1, (60, 3))).astype(np.float32)
    grid = dense_grid(sd, [1], 2, N)
    np.testing.assert_allclose(integrate_mlp(sd, [1], 2, N, ci, co),
                               fp64.integrate_einsum(grid, ci, co), rtol=1e-10, atol=1e-12)
