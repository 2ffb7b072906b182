"""Decoder parameters in the library's canonical form, and the device handle that holds them.

Three sources map onto the one canonical form (see include/diver.h, "Canonical decoder"):
- the reference's offline decoder state dict (``mlp1.*``/``mlp2.*`` of model/diver.py:45-54);
  the last mlp1 layer and the first mlp2 layer have no activation between them, so they are
  fused exactly as convert.py:70-77 does for its real-time blob;
- the reference's real-time blob (convert.py:63-140, consumed by upload_weight); its PE columns
  are stored per axis as [raw, sin x4, cos x4] and are mapped back to the torch order of
  model/mlps.py:12-18; the DIVeR32 blob expects features pre-multiplied by layer 0
  (convert.py:149-150), which becomes an identity first layer;
- the benchmark decoder: density softplus(f[0]) and a colour MLP on [f[1:], PE(d)].
"""
from __future__ import annotations

import ctypes
import dataclasses

import numpy as np

from . import _lib

PE_DIM = 27
# torch PE index of kernel position (axis i, slot e): e=0 raw, 1..4 sin bands, 5..8 cos bands
_PE_KERNEL_ORDER = np.array([0, 3, 9, 15, 21, 6, 12, 18, 24,
                             1, 4, 10, 16, 22, 7, 13, 19, 25,
                             2, 5, 11, 17, 23, 8, 14, 20, 26])


@dataclasses.dataclass
class CanonicalDecoder:
    channels: int
    hidden: int
    n_pre: int
    pre_identity: bool
    density_act: int          # 0 relu, 1 softplus
    n_post: int
    params: np.ndarray        # float32, layout of include/diver.h

    def desc(self) -> _lib.DecoderDesc:
        return _lib.DecoderDesc(self.channels, self.hidden, self.n_pre, int(self.pre_identity),
                                self.density_act, self.n_post)

    # ---- unpacked views (host, float32) ---------------------------------------------------
    def layers(self) -> dict:
        """The canonical matrices as a dict (used by tests and by the adjoint)."""
        C, H = self.channels, self.hidden
        Z = H if self.n_pre else C
        p = self.params
        o = 0
        out = {'pre': []}
        for i in range(self.n_pre):
            if i == 0 and self.pre_identity:
                out['pre'].append((None, p[o:o + H])); o += H
            else:
                k = C if i == 0 else H
                w = p[o:o + H * k].reshape(H, k); o += H * k
                b = p[o:o + H]; o += H
                out['pre'].append((w, b))
        out['w_s'] = p[o:o + Z]; o += Z
        out['b_s'] = p[o]; o += 1
        out['W_c'] = p[o:o + H * Z].reshape(H, Z); o += H * Z
        out['W_v'] = p[o:o + H * PE_DIM].reshape(H, PE_DIM); o += H * PE_DIM
        out['b_c'] = p[o:o + H]; o += H
        out['post'] = []
        for _ in range(self.n_post):
            w = p[o:o + H * H].reshape(H, H); o += H * H
            b = p[o:o + H]; o += H
            out['post'].append((w, b))
        out['W_o'] = p[o:o + 3 * H].reshape(3, H); o += 3 * H
        out['b_o'] = p[o:o + 3]; o += 3
        assert o == p.size
        return out


def _cat(*xs) -> np.ndarray:
    return np.concatenate([np.asarray(x, np.float64).reshape(-1) for x in xs]).astype(np.float32)


def from_baseline(layers, channels: int) -> CanonicalDecoder:
    """Benchmark decoder: ``layers`` = [(W1 [H, C-1+27], b1), (W2 [H,H], b2), (W3 [3,H], b3)]."""
    (w1, b1), (w2, b2), (w3, b3) = layers
    H = w1.shape[0]
    C = channels
    assert w1.shape[1] == C - 1 + PE_DIM and w2.shape == (H, H) and w3.shape == (3, H)
    ws = np.zeros(C); ws[0] = 1.0
    wc = np.zeros((H, C)); wc[:, 1:] = w1[:, :C - 1]
    wv = w1[:, C - 1:]
    params = _cat(ws, [0.0], wc, wv, b1, w2, b2, w3, b3)
    return CanonicalDecoder(C, H, 0, False, 1, 1, params)


def from_state_dict(sd: dict, channels: int = 32) -> CanonicalDecoder:
    """Reference offline decoder (model/diver.py:45-54 state dict, float arrays or tensors)."""
    sd = {k: np.asarray(v.detach().cpu().numpy() if hasattr(v, 'detach') else v, np.float64)
          for k, v in sd.items() if k.startswith('mlp1.') or k.startswith('mlp2.')}
    pre = []
    i = 0
    while f'mlp1.{i + 2}.weight' in sd:
        pre.append((sd[f'mlp1.{i}.weight'], sd[f'mlp1.{i}.bias']))
        i += 2
    wl, bl = sd[f'mlp1.{i}.weight'], sd[f'mlp1.{i}.bias']
    H = pre[-1][0].shape[0] if pre else channels
    w2, b2 = sd['mlp2.0.weight'], sd['mlp2.0.bias']
    if 'mlp2.4.weight' in sd:
        raise ValueError('mlp_view depth > 1 is not supported by the canonical decoder')
    w3, b3 = sd['mlp2.2.weight'], sd['mlp2.2.bias']
    cx = wl.shape[0] - 1
    if w2.shape[0] != H:
        raise ValueError(f'mlp_view width {w2.shape[0]} must equal mlp_point width {H}')
    w2x, w2v = w2[:, :cx], w2[:, cx:]
    assert w2v.shape[1] == PE_DIM
    wc = w2x @ wl[1:]
    bc = w2x @ bl[1:] + b2
    chunks = []
    for (w, b) in pre:
        chunks += [w, b]
    params = _cat(*chunks, wl[0], [bl[0]], wc, w2v, bc, w3, b3)
    return CanonicalDecoder(channels, H, len(pre), False, 0, 0, params)


def from_rt_blob(params, hidden: int) -> CanonicalDecoder:
    """Reference real-time blob (convert.py:63-140; 3168 floats for DIVeR32, 8288 for DIVeR64)."""
    p = np.asarray(params.detach().cpu().numpy() if hasattr(params, 'detach') else params,
                   np.float64).reshape(-1)
    H = hidden
    inv = np.argsort(_PE_KERNEL_ORDER)  # kernel slot of each torch PE index
    if H == 32:
        if p.size != 3168:
            raise ValueError(f'DIVeR32 blob has 3168 floats, got {p.size}')
        o = 0
        b0 = p[o:o + 32]; o += 32
        b1 = p[o:o + 32]; o += 32
        w1 = p[o:o + 1024].reshape(32, 32).T; o += 1024
        pre = [b0, w1, b1]
    elif H == 64:
        if p.size != 8288:
            raise ValueError(f'DIVeR64 blob has 8288 floats, got {p.size}')
        o = 0
        b0 = p[o:o + 64]; o += 64
        w0 = p[o:o + 32 * 64].reshape(32, 64).T; o += 32 * 64
        pre = [w0, b0]
    else:
        raise ValueError('hidden must be 32 or 64')
    bs = p[o]; o += 1
    ws = p[o:o + H]; o += H
    bc = p[o:o + H]; o += H
    wc = p[o:o + H * H].reshape(H, H).T; o += H * H
    wv_k = p[o:o + PE_DIM * H].reshape(PE_DIM, H).T; o += PE_DIM * H  # [H, kernel slot]
    wv = wv_k[:, inv]
    bo = p[o:o + 3]; o += 3
    wo = p[o:o + H * 3].reshape(H, 3).T; o += H * 3
    params_c = _cat(*pre, ws, [bs], wc, wv, bc, wo, bo)
    if H == 32:
        return CanonicalDecoder(32, 32, 2, True, 0, 0, params_c)
    return CanonicalDecoder(32, 64, 1, False, 0, 0, params_c)


class Decoder:
    """Device handle of a canonical decoder (created on the current HIP device)."""

    def __init__(self, canon: CanonicalDecoder):
        self.canon = canon
        desc = canon.desc()
        params = np.ascontiguousarray(canon.params, np.float32)
        handle = ctypes.c_void_p()
        _lib.call('diver_decoder_create', ctypes.byref(desc),
                  params.ctypes.data_as(ctypes.POINTER(ctypes.c_float)), params.size,
                  ctypes.byref(handle))
        self.handle = handle.value

    def close(self):
        if getattr(self, 'handle', None):
            try:
                _lib.lib().diver_decoder_destroy(self.handle)
            finally:
                self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @property
    def nbytes(self) -> int:
        """Device bytes the handle holds: FP32 parameters plus the BF16 and FP16 matrix-core
        operand packs (at most the same element count each, at 2 bytes)."""
        return int(self.canon.params.size) * (4 + 2 + 2)

    @property
    def channels(self):
        return self.canon.channels

    @property
    def hidden(self):
        return self.canon.hidden


def param_count(canon: CanonicalDecoder) -> int:
    d = canon.desc()
    return int(_lib.lib().diver_decoder_param_count(ctypes.byref(d)))
