"""diver_adjoint_weight_grads (the decoder adjoint's weight gradients) against FP64 reductions of the
same operands, for the three decoder shapes the adjoint supports, on row counts that leave a partial
LDS stage, fill one workgroup, and spread over the most workgroups with a ragged last slab.

The reference forms these sums through torch autograd of its nn.Linear layers (model/mlps.py:21-31
through model/diver.py:159-171); here the FP64 reference is the textbook definition of each weight
gradient of y = W x + b, summed over rows:
  dW = sum_k dy_k x_k^T, db = sum_k dy_k,
with x = F for the density row and W_c, PE(d) of the row's ray for W_v, r1 for W_2, r2 for W_o.

Tolerance: |g - g_64| <= 1e-5 * sum_k |dy_k| |x_k| + 1e-30 per element (FP32 partial sums of at most
a few hundred rows per workgroup, FP64 across workgroups).
"""
import numpy as np
import pytest
import torch

from diver_amd import calls
from diver_amd.decoder import CanonicalDecoder, Decoder

pytestmark = pytest.mark.gpu

PE = 27


def _pe64(d):
    parts = [d]
    for l in range(4):
        parts.append(np.sin(np.pi * 2.0 ** l * d))
        parts.append(np.cos(np.pi * 2.0 ** l * d))
    return np.concatenate(parts, 1)


def _operands(n, R, C, H, seed):
    g = np.random.default_rng(seed)
    f = lambda *s: g.standard_normal(s).astype(np.float32)
    rows = f(n, 9 + C)
    ray = np.sort(g.integers(0, R, n)).astype(np.int64)
    d = f(R, 3)
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    r1 = np.maximum(f(n, H), 0)
    r2 = np.maximum(f(n, H), 0)
    do, ds, dr1, dr2 = f(n, 3), f(n), f(n, H), f(n, H)
    return rows, ray, d, r1, r2, do, ds, dr1, dr2


def _reference(ops, C, H):
    rows, ray, d, r1, r2, do, ds, dr1, dr2 = (x.astype(np.float64) if x.dtype != np.int64 else x
                                               for x in ops)
    F = rows[:, 9:]
    pe = _pe64(d)[ray]
    blocks, bounds = [], []
    for dy, x in ((ds[:, None], F), (ds[:, None], None), (dr1, F), (dr1, pe), (dr1, None),
                  (dr2, r1), (dr2, None), (do, r2), (do, None)):
        if x is None:
            blocks.append(dy.sum(0).reshape(-1))
            bounds.append(np.abs(dy).sum(0).reshape(-1))
        else:
            blocks.append((dy.T @ x).reshape(-1))
            bounds.append((np.abs(dy).T @ np.abs(x)).reshape(-1))
    return np.concatenate(blocks), np.concatenate(bounds)


@pytest.mark.parametrize('C,H', [(32, 32), (32, 64), (64, 64)])
@pytest.mark.parametrize('n', [0, 1, 45, 1000, 70001])
def test_weight_grads_match_fp64(dev, C, H, n):
    canon = CanonicalDecoder(C, H, 0, False, 1, 1,
                             np.random.default_rng(1).standard_normal(
                                 C + 1 + H * C + H * PE + H + H * H + H + 3 * H + 3).astype(np.float32))
    dec = Decoder(canon)
    R = max(1, n // 3)
    ops = _operands(n, R, C, H, seed=n + C + H)
    t = [torch.from_numpy(x).to(dev) for x in ops]
    print(f'weight grads C={C} H={H} n={n}', flush=True)
    got = calls.adjoint_weight_grads(dec, n, *t).cpu().numpy().astype(np.float64)
    ref, bound = _reference(ops, C, H)
    assert got.shape == ref.shape
    err = np.abs(got - ref)
    assert np.all(err <= 1e-5 * bound + 1e-30), float(np.max(err / (bound + 1e-30)))
    # a fixed summation order: the same inputs give the same bits
    again = calls.adjoint_weight_grads(dec, n, *t).cpu().numpy()
    assert np.array_equal(again, got.astype(np.float32))
