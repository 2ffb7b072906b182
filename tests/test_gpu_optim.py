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


@pytest.mark.parametrize('clear', [False, True])
def test_adam_rows_equals_full_step(dev, clear):
    """The row-restricted step (diver_adam_step_rows) over every row that receives a gradient
    leaves parameters and moments bit for bit as the full one-pass step over all elements, and
    within rounding of torch.optim.Adam; rows never listed keep their values exactly."""
    R, L = 3001, 32
    g = torch.Generator(device='cpu').manual_seed(5)
    p0 = torch.randn(R * L, generator=g)
    rows = torch.randperm(R, generator=g)[:700].sort().values
    touched = torch.zeros(R, dtype=torch.bool)
    touched[rows] = True
    grads = []
    for _ in range(6):
        gr = torch.randn(R, L, generator=g) * 1e-2
        gr[~touched] = 0.0
        grads.append(gr.reshape(-1).to(dev))
    a = torch.nn.Parameter(p0.clone().to(dev))
    b = torch.nn.Parameter(p0.clone().to(dev))
    c = torch.nn.Parameter(p0.clone().to(dev))
    ours = Adam([{'params': [a], 'rows': rows.to(dev), 'row_len': L}], lr=1e-2, clear_grad=clear)
    full = Adam([b], lr=1e-2, clear_grad=clear)
    ref = torch.optim.Adam([c], lr=1e-2)
    for gr in grads:
        a.grad = gr.clone()
        b.grad = gr.clone()
        c.grad = gr.clone()
        ours.step()
        full.step()
        ref.step()
        if clear:
            assert float(a.grad.abs().max()) == 0.0
    torch.cuda.synchronize()
    assert torch.equal(a, b)
    assert torch.equal(ours.state[a]['exp_avg'], full.state[b]['exp_avg'])
    assert torch.equal(ours.state[a]['exp_avg_sq'], full.state[b]['exp_avg_sq'])
    assert _close(a, c)
    untouched = (~touched).repeat_interleave(L).to(dev)
    assert torch.equal(a.detach()[untouched], p0.to(dev)[untouched])


def test_adam_rows_refuses_weight_decay(dev):
    from diver_amd import _lib
    a = torch.nn.Parameter(torch.zeros(64, device=dev))
    a.grad = torch.zeros_like(a)
    opt = Adam([{'params': [a], 'rows': torch.arange(2, device=dev), 'row_len': 32}], lr=1e-2,
               weight_decay=1e-2)
    with pytest.raises(_lib.DiverError):
        opt.step()
