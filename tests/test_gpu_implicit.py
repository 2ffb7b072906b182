"""Implicit voxel grid (integrate_mlp, init_voxels) on libdiver against the FP64 oracle
(oracle/implicit.py) and the reference algorithm's float64 autograd.

Stated tolerances (features are O(1)):
  integrate_mlp forward: |GPU - FP64| <= 1e-4 (FP32 MLP GEMMs + FP32 table integral);
  parameter gradients: within 1e-5 of the largest |gradient| of torch autograd through the
  reference algorithm in FP32 on the GPU (unique (z, y, x) corners, MLP, gather, corner-weight
  sum), and within 2e-2 of its float64 CPU autograd (ReLU gates near zero flip between FP32 and
  FP64 evaluations of the 2^5 pi encoding band, moving a few first-layer gradients);
  init_voxels: the baked grid within 1e-4 of the FP64 MLP at every corner, and decoding through it
  matches decoding through integrate_mlp within 1e-4.
"""
import numpy as np
import pytest
import torch

from oracle import fp64
from oracle.implicit import dense_grid, integrate_mlp as oracle_integrate_mlp
from diver_amd.model.mlps import ImplicitMLP
from diver_amd.utils.integrator import integrate, integrate_mlp

pytestmark = pytest.mark.gpu

N = 17  # corners per axis (voxel_num 16)
SKIPS, BANDS = [2], 6


def _mlp(dev):
    torch.manual_seed(11)
    return ImplicitMLP(4, 64, SKIPS, 32, BANDS).to(dev)


def _segments(K, seed):
    rng = np.random.default_rng(seed)
    cell = rng.integers(0, N - 1, (K, 3))
    ci = (cell + rng.uniform(0, 1, (K, 3))).astype(np.float32)
    co = (cell + rng.uniform(0, 1, (K, 3))).astype(np.float32)
    return ci, co


def _sd(m):
    return {k: v.detach().double().cpu().numpy() for k, v in m.state_dict().items()}


def test_integrate_mlp_matches_oracle(dev):
    m = _mlp(dev)
    ci, co = _segments(3000, 1)
    got = integrate_mlp(m, N, 32, torch.from_numpy(ci).to(dev), torch.from_numpy(co).to(dev))
    ref = oracle_integrate_mlp(_sd(m), SKIPS, BANDS, N, ci, co)
    assert np.abs(got.detach().cpu().numpy() - ref).max() <= 1e-4


def _reference_algorithm(m, ci, co, dtype, device):
    """integrator.py:3-60 in torch (unique corners, MLP, gather, corner weights) at dtype."""
    ci_t, co_t = torch.from_numpy(ci).to(device), torch.from_numpy(co).to(device)
    pmin = torch.min((ci_t + 1e-4).long(), (co_t + 1e-4).long())
    xmin, ymin, zmin = pmin.split(1, -1)
    xmax, ymax, zmax = (pmin + 1).clamp_max(N - 1).split(1, -1)
    corners = torch.stack([zmin, ymin, xmin, zmin, ymin, xmax, zmin, ymax, xmin, zmin, ymax, xmax,
                           zmax, ymin, xmin, zmax, ymin, xmax, zmax, ymax, xmin, zmax, ymax, xmax],
                          -1).reshape(-1, 3)
    cu, rev = corners.unique(dim=0, return_inverse=True)
    Fu = m((cu.float() / (N * 0.5) - 1.0).to(dtype))
    F = Fu[rev].reshape(-1, 8, Fu.shape[-1])
    w, _ = fp64.corner_weights(ci, co)
    return (torch.from_numpy(w).to(device=device, dtype=dtype)[:, :, None] * F).sum(1)


def test_integrate_mlp_gradients(dev):
    m = _mlp(dev)
    ci, co = _segments(2000, 2)
    g_out = torch.from_numpy(np.random.default_rng(3).normal(size=(2000, 32)).astype(np.float32))

    def grads(feat_fn):
        m.zero_grad(set_to_none=True)
        (feat_fn() * g_out.to(dev)).sum().backward()
        return [p.grad.double().cpu().clone() for p in m.parameters()]

    ours = grads(lambda: integrate_mlp(m, N, 32, torch.from_numpy(ci).to(dev),
                                       torch.from_numpy(co).to(dev)))
    ref32 = grads(lambda: _reference_algorithm(m, ci, co, torch.float32, dev))
    m64 = ImplicitMLP(4, 64, SKIPS, 32, BANDS).double()
    m64.load_state_dict({k: v.double().cpu() for k, v in m.state_dict().items()})
    (_reference_algorithm(m64, ci, co, torch.float64, 'cpu') * g_out.double()).sum().backward()
    for (name, _), a, r, q in zip(m.named_parameters(), ours, ref32, m64.parameters()):
        assert (a - r).abs().max() <= 1e-5 * r.abs().max() + 1e-9, name
        assert (a - q.grad).abs().max() <= 2e-2 * q.grad.abs().max() + 1e-9, name


def test_init_voxels_and_baked_decode(dev):
    from types import SimpleNamespace
    from diver_amd.model.diver import DIVeR
    hp = SimpleNamespace(voxel_num=N - 1, voxel_dim=32, grid_size=2.0, mask_scale=1.0,
                         white_back=True, fine=0, mlp_point=[32, 1, 33], mlp_view=[32, 1],
                         dir_encode=4, implicit=True, implicit_network_depth=4,
                         implicit_channels=64, implicit_skips=SKIPS, implicit_point_encode=BANDS)
    model = DIVeR(hp).to(dev)
    model.mlp_voxel.load_state_dict(_mlp(dev).state_dict())
    ci, co = _segments(1000, 4)
    ci_t, co_t = torch.from_numpy(ci).to(dev), torch.from_numpy(co).to(dev)
    with torch.no_grad():
        implicit = integrate_mlp(model.mlp_voxel, N, 32, ci_t, co_t)
    model.init_voxels(True)
    grid = model.voxels.detach()
    ref = dense_grid(_sd(model.mlp_voxel), SKIPS, BANDS, N)
    assert tuple(grid.shape) == (N, N, N, 32)
    assert np.abs(grid.cpu().numpy() - ref).max() <= 1e-4
    with torch.no_grad():
        baked = integrate(grid, ci_t, co_t)
    assert (baked - implicit).abs().max().item() <= 1e-4
