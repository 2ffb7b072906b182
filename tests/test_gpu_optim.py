"""One-pass Adam (diver_adam_step) against torch.optim.Adam, the optimizer the reference trains with
(train.py:39-47).

Both run the same update rule in FP32 with a different operation order (FMA contraction, where the
products are grouped), so they differ by rounding only.  After 6 steps at lr 1e-2 the bound is
|ours - torch| <= 1e-7 + 1e-6 |torch| for the parameters (about 8 ulp at |p| = 1; an update off by
one part in 10^4, e.g. a bias correction or a (1 - beta) taken from the wrong power, misses it by
orders of magnitude), and 1e-5 relative for the moments.  Sizes cover the float4 body and the 1-3
element tail.
"""
import pytest
import torch

from diver_amd.optim import Adam

pytestmark = pytest.mark.gpu


def _close(a, b):
    a, b = a.detach(), b.detach()
    return float(((a - b).abs() - 1e-6 * b.abs()).max()) <= 1e-7


def _pair(n, seed, dev):
    g = torch.Generator(device='cpu').manual_seed(seed)
    p0 = torch.randn(n, generator=g)
    grads = [torch.randn(n, generator=g) * 1e-2 for _ in range(6)]
    a = torch.nn.Parameter(p0.clone().to(dev))
    b = torch.nn.Parameter(p0.clone().to(dev))
    return a, b, [x.to(dev) for x in grads]


@pytest.mark.parametrize('n', [1, 3, 4, 1027, 1 << 20])
@pytest.mark.parametrize('wd', [0.0, 1e-2])
def test_adam_matches_torch(dev, n, wd):
    a, b, grads = _pair(n, n + 7, dev)
    ours = Adam([a], lr=1e-2, weight_decay=wd)
    ref = torch.optim.Adam([b], lr=1e-2, weight_decay=wd)
    for gr in grads:
        a.grad = gr.clone()
        b.grad = gr.clone()
        ours.step()
        ref.step()
    torch.cuda.synchronize()
    assert _close(a, b)
    st, rs = ours.state[a], ref.state[b]
    assert float((st['exp_avg'] - rs['exp_avg']).abs().max()) <= 1e-7
    assert float(((st['exp_avg_sq'] - rs['exp_avg_sq']).abs() / rs['exp_avg_sq'].clamp_min(1e-12)).max()) <= 1e-5
    assert st['step'] == len(grads)


def test_adam_clears_gradient_in_the_same_pass(dev):
    a, b, grads = _pair(4099, 3, dev)
    ours = Adam([a], lr=1e-2, clear_grad=True)
    ref = torch.optim.Adam([b], lr=1e-2)
    for gr in grads:
        a.grad.copy_(gr) if a.grad is not None else setattr(a, 'grad', gr.clone())
        b.grad = gr.clone()
        ours.step()
        ref.step()
        torch.cuda.synchronize()
        assert int(torch.count_nonzero(a.grad)) == 0
    assert _close(a, b)


def test_adam_refuses_unsupported_parameters(dev):
    x = torch.nn.Parameter(torch.zeros(8, 8, device=dev).t())  # not contiguous
    x.grad = torch.zeros(8, 8, device=dev).t()
    with pytest.raises(TypeError):
        Adam([x]).step()


def test_adam_unaligned_views(dev):
    # parameter and gradient 4 bytes into their storage (autograd can hand over such gradient
    # views): the element-by-element form, same update
    a, b, grads = _pair(1031, 5, dev)
    store = torch.zeros(1032, device=dev)
    store[1:] = a.detach()
    pa = torch.nn.Parameter(store[1:])
    ours = Adam([pa], lr=1e-2)
    ref = torch.optim.Adam([b], lr=1e-2)
    gstore = torch.zeros(1032, device=dev)
    for gr in grads:
        gstore[1:] = gr
        pa.grad = gstore[1:]
        b.grad = gr.clone()
        ours.step()
        ref.step()
    torch.cuda.synchronize()
    assert pa.data_ptr() % 16 != 0 and pa.grad.data_ptr() % 16 != 0
    assert _close(pa, b)
