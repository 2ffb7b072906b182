"""diver_adjoint_weight_grads_noalloc (the allocation-free form, partials in the library's own
device array) against diver_adjoint_weight_grads (partials in a caller workspace), through the
C-ABI: bit-identical gradients on two calls in a row, the second with fewer rows, so the second
call runs over what the first left in the shared array.  Both forms run the same kernels with the
same workgroup split, so the bar is equality of every bit."""
import ctypes

import numpy as np
import pytest
import torch

from diver_amd import _lib
from diver_amd.decoder import CanonicalDecoder, Decoder

pytestmark = pytest.mark.gpu

PE = 27


def _call(dec, n, ops, grad, ws=None):
    ptrs = [t.data_ptr() for t in ops]
    stream = torch.cuda.current_stream().cuda_stream
    if ws is None:
        _lib.call('diver_adjoint_weight_grads_noalloc', dec.handle, n, *ptrs, grad.data_ptr(), stream)
    else:
        _lib.call('diver_adjoint_weight_grads', dec.handle, n, *ptrs, grad.data_ptr(),
                  ws.data_ptr(), ws.numel() * 4, stream)


@pytest.mark.parametrize('C,H', [(32, 64), (64, 64)])
def test_no_alloc_weight_grads_match_workspace_form(dev, C, H):
    nparam = C + 1 + H * C + H * PE + H + H * H + H + 3 * H + 3
    dec = Decoder(CanonicalDecoder(C, H, 0, False, 1, 1,
                                   np.random.default_rng(2).standard_normal(nparam).astype(np.float32)))
    L = _lib.lib()
    for n in (90001, 3001):
        print(f'no-alloc weight grads C={C} H={H} n={n}', flush=True)
        g = torch.Generator(device='cpu').manual_seed(n)
        R = n // 5 + 1
        rows = torch.randn(n, 9 + C, generator=g)
        ray = torch.sort(torch.randint(0, R, (n,), generator=g)).values
        d = torch.nn.functional.normalize(torch.randn(R, 3, generator=g), dim=1)
        r1, r2 = torch.randn(n, H, generator=g).relu(), torch.randn(n, H, generator=g).relu()
        ops = [t.to(dev).contiguous() for t in (rows, ray, d, r1, r2, torch.randn(n, 3, generator=g),
                                                torch.randn(n, generator=g),
                                                torch.randn(n, H, generator=g),
                                                torch.randn(n, H, generator=g))]
        need = L.diver_adjoint_weight_grads_workspace(ctypes.c_void_p(dec.handle), n)
        assert need > 0
        ws = torch.empty(need // 4, dtype=torch.float32, device=dev)
        want = torch.empty(nparam, dtype=torch.float32, device=dev)
        got = torch.full((nparam,), float('nan'), dtype=torch.float32, device=dev)
        _call(dec, n, ops, want, ws)
        _call(dec, n, ops, got)
        torch.cuda.synchronize()
        assert torch.equal(got, want)
        assert bool(torch.isfinite(got).all())
