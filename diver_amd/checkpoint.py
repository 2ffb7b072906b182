"""The reference's deployment formats on the MI355X path: ``weight.pth`` and ``serialize.pth``.

- ``weight.pth`` (built by convert.py:23-55, read by eval.py:44-52): the model's state dict without
  the implicit MLP, with ``voxels`` stored as a sparse COO tensor over the corners of the occupied
  voxels and ``voxel_mask`` the bool [N]^3 occupancy (alpha map > 1e-2).
- ``serialize.pth`` (built by convert.py:57-161, read by eval_rt.py:52-82):
  ``{'f': features [K, C], 'i': int32 corner index z (N+1)^2 + y (N+1) + x, 'm': bool [K] (the
  corner is an occupied voxel's own index), 'p': the real-time decoder blob}``; DIVeR32 features
  are pre-multiplied by the first decoder layer (convert.py:148-150).

``weight_from_checkpoint``, ``rt_blob`` and ``serialize`` build these formats from a state dict
(the product-side counterpart of convert.py); ``load_weight`` and ``load_serialized`` read them with
a weights-only loader into device tensors and decoders that the drop-in ops, the real-time
renderer and the fused renderer take.  The blob's branch is chosen by the decoder's shape (the
width of ``mlp1``'s hidden layers), not by the feature width as convert.py:61-63,102 does, which
would send DIVeR64 checkpoints down the DIVeR32 branch.
"""
from __future__ import annotations

import dataclasses
import os
from typing import Optional, Union

import numpy as np
import torch

from .decoder import PE_DIM, CanonicalDecoder, from_rt_blob, from_state_dict

# kernel slot order of the colour layer's PE columns (per axis: raw, sin x4, cos x4) as torch
# indices of model/mlps.py:12-18 (convert.py:89-92)
PE_KERNEL_ORDER = np.array([0, 3, 9, 15, 21, 6, 12, 18, 24,
                            1, 4, 10, 16, 22, 7, 13, 19, 25,
                            2, 5, 11, 17, 23, 8, 14, 20, 26])
BLOB_PAD = 28  # zeros at the end of the blob (convert.py:100,140)


def _np(v) -> np.ndarray:
    if isinstance(v, torch.Tensor):
        v = v.detach().cpu()
        if v.is_sparse:
            v = v.to_dense()
        return v.numpy()
    return np.asarray(v)


def _load(src) -> dict:
    if isinstance(src, (str, os.PathLike)):
        return torch.load(src, map_location='cpu', weights_only=True)
    return src


def _corner_set(mask: np.ndarray) -> np.ndarray:
    """[K, 3] (z, y, x) corners of the occupied voxels, sorted and unique (convert.py:35-43)."""
    z, y, x = np.nonzero(mask)
    d = np.array([[i >> 2 & 1, i >> 1 & 1, i & 1] for i in range(8)])
    idx = np.stack([z, y, x], 1)[None] + d[:, None]
    return np.unique(idx.reshape(-1, 3), axis=0)


def weight_from_checkpoint(state_dict: dict, alpha_map, threshold: float = 1e-2) -> dict:
    """convert.py:23-55: the training checkpoint's state dict (keys ``model.*``; the implicit
    ``mlp_voxel`` is dropped) and the pruning alpha map -> the ``weight.pth`` dict, voxels as a
    sparse COO tensor over the corners of the voxels with alpha above ``threshold``."""
    weight = {}
    for k, v in state_dict.items():
        if 'mlp_voxel' in k:
            continue
        if 'model.' in k:
            weight[k.replace('model.', '')] = v
    mask = torch.as_tensor(_np(alpha_map)) > threshold
    weight['voxel_mask'] = mask
    voxels = torch.as_tensor(_np(weight['voxels']))
    idx = torch.from_numpy(_corner_set(mask.numpy()))
    weight['voxels'] = torch.sparse_coo_tensor(idx.T, voxels[idx[:, 0], idx[:, 1], idx[:, 2]],
                                               tuple(voxels.shape))
    return weight


def rt_blob(state_dict: dict) -> np.ndarray:
    """convert.py:63-140: the real-time kernels' decoder blob (float32) from the offline decoder's
    ``mlp1``/``mlp2`` layers.  mlp1's last layer and mlp2's first are fused (no activation lies
    between them), biases are folded in as leading columns, the PE columns go to the kernel order
    and 28 zeros pad the end.  DIVeR32 (two hidden mlp1 layers of width 32): 3168 floats, its
    features to be pre-multiplied by mlp1.0 (``serialize``); DIVeR64 (one of width 64): 8288."""
    sd = {k: _np(v).astype(np.float64) for k, v in state_dict.items()
          if k.startswith('mlp1.') or k.startswith('mlp2.')}
    n_hidden = sum(1 for k in sd if k.startswith('mlp1.') and k.endswith('.weight')) - 1
    H = sd['mlp1.0.weight'].shape[0]
    if (n_hidden, H) not in ((2, 32), (1, 64)):
        raise ValueError(f'no real-time kernel for {n_hidden} hidden mlp1 layers of width {H} '
                         '(DIVeR32: 2 x 32, DIVeR64: 1 x 64)')
    last = f'mlp1.{2 * n_hidden}'
    wl, bl = sd[last + '.weight'], sd[last + '.bias']        # [1 + H, H]
    w2, b2 = sd['mlp2.0.weight'], sd['mlp2.0.bias']          # [H, H + 27]
    w3, b3 = sd['mlp2.2.weight'], sd['mlp2.2.bias']          # [3, H]
    w_sig = np.concatenate([[bl[0]], wl[0]])                                  # [1 + H]
    w_col = np.concatenate([(b2 + w2[:, :H] @ bl[1:])[:, None], w2[:, :H] @ wl[1:],
                            w2[:, H:][:, PE_KERNEL_ORDER]], 1)               # [H, 1 + H + 27]
    w_out = np.concatenate([b3[:, None], w3], 1)                             # [3, 1 + H]
    if H == 64:
        w0 = np.concatenate([sd['mlp1.0.bias'][:, None], sd['mlp1.0.weight']], 1)  # [64, 33]
        head = [w0.T.reshape(-1)]
    else:
        w1 = np.concatenate([sd['mlp1.2.bias'][:, None], sd['mlp1.2.weight']], 1)  # [32, 33]
        head = [sd['mlp1.0.bias'], w1.T.reshape(-1)]
    parts = head + [w_sig, w_col.T.reshape(-1), w_out.T.reshape(-1), np.zeros(BLOB_PAD)]
    return np.concatenate(parts).astype(np.float32)


def serialize(weight: dict) -> dict:
    """convert.py:57-161: the ``weight.pth`` dict -> the ``serialize.pth`` dict {f, i, m, p}."""
    vox = weight['voxels']
    vox = vox.coalesce() if vox.is_sparse else vox.to_sparse(3).coalesce()
    n1 = vox.shape[0]
    N = n1 - 1
    idxs = vox.indices()                       # [3, K] (z, y, x)
    feats = vox.values().to(torch.float64)     # [K, C]
    blob = rt_blob(weight)
    if weight['mlp1.0.weight'].shape[0] == 32:
        feats = feats @ torch.as_tensor(_np(weight['mlp1.0.weight']), dtype=torch.float64).T
    mask = torch.as_tensor(_np(weight['voxel_mask'])).bool()
    valid = (idxs < N).all(0)
    is_corner = torch.zeros(idxs.shape[1], dtype=torch.bool)
    is_corner[valid] = mask[idxs[0][valid], idxs[1][valid], idxs[2][valid]]
    inds = (idxs[0] * n1 + idxs[1]) * n1 + idxs[2]
    return {'f': feats.to(torch.float32).contiguous(), 'i': inds.to(torch.int32),
            'm': is_corner, 'p': torch.from_numpy(blob)}


@dataclasses.dataclass
class WeightCheckpoint:
    """A loaded ``weight.pth``: dense corner features, occupancy and the decoder."""
    voxels: torch.Tensor          # [(N+1)^3 grid, C] float32 (dense; zero off the stored corners)
    mask: torch.Tensor            # [N, N, N] bool
    decoder_state: dict           # mlp1.* / mlp2.* tensors
    state: dict                   # the whole state dict (dense voxels), for DIVeR.load_state_dict

    def decoder(self) -> CanonicalDecoder:
        return from_state_dict(self.decoder_state, self.voxels.shape[-1])


def load_weight(src: Union[str, os.PathLike, dict], device=None) -> WeightCheckpoint:
    """eval.py:44-52: read ``weight.pth`` (weights-only), densify the sparse voxels."""
    w = dict(_load(src))
    vox = w['voxels']
    vox = vox.to_dense() if vox.is_sparse else vox
    w['voxels'] = vox.to(torch.float32)
    dev = torch.device(device) if device is not None else torch.device('cpu')
    dec = {k: v for k, v in w.items() if k.startswith('mlp1.') or k.startswith('mlp2.')}
    return WeightCheckpoint(w['voxels'].to(dev).contiguous(),
                            torch.as_tensor(w['voxel_mask']).bool().to(dev).contiguous(), dec, w)


@dataclasses.dataclass
class SerializedScene:
    """A loaded ``serialize.pth`` (eval_rt.py:52-82) on the device."""
    n: int
    features: torch.Tensor        # [K, C] rows (pre-multiplied by mlp1.0 for DIVeR32)
    voxel_map: torch.Tensor       # [(N+1)^3] int32 row of each corner (0 where none is stored)
    mask: torch.Tensor            # [N, N, N] bool
    blob: np.ndarray              # the real-time decoder blob
    hidden: int

    def rt_decoder(self) -> CanonicalDecoder:
        """The blob decoder as mlp_eval consumes it (DIVeR32: features pre-multiplied)."""
        return from_rt_blob(self.blob, self.hidden)

    def fused_decoder(self) -> CanonicalDecoder:
        """The same decoder for the fused renderer over ``dense_voxels()``: DIVeR32's
        pre-multiplied first layer becomes an explicit identity layer (the integral is linear, so
        integrating pre-multiplied features equals multiplying the integral)."""
        c = self.rt_decoder()
        if not c.pre_identity:
            return c
        H = c.hidden
        b0, rest = c.params[:H], c.params[H:]
        params = np.concatenate([np.eye(H, dtype=np.float32).reshape(-1), b0, rest])
        return CanonicalDecoder(c.channels, H, c.n_pre, False, c.density_act, c.n_post,
                                params.astype(np.float32))

    def dense_voxels(self) -> torch.Tensor:
        """[(N+1)^3 grid, C]: the stored rows at their corners, zero elsewhere."""
        n1 = self.n + 1
        C = self.features.shape[-1]
        out = torch.zeros((n1 ** 3, C), dtype=torch.float32, device=self.features.device)
        out[self._inds] = self.features
        return out.reshape(n1, n1, n1, C)

    _inds: Optional[torch.Tensor] = None


def load_serialized(src: Union[str, os.PathLike, dict], voxel_num: int,
                    device=None) -> SerializedScene:
    """eval_rt.py:52-82: read ``serialize.pth`` (weights-only) and rebuild the occupancy mask
    (from the rows flagged ``m``) and the corner map (row of each stored corner)."""
    w = _load(src)
    dev = torch.device(device) if device is not None else torch.device('cpu')
    n1 = voxel_num + 1
    inds = torch.as_tensor(w['i']).long()
    m = torch.as_tensor(w['m']).bool()
    zz, rem = inds // (n1 * n1), inds % (n1 * n1)
    yy, xx = rem // n1, rem % n1
    mask = torch.zeros((voxel_num,) * 3, dtype=torch.bool)
    mask[zz[m], yy[m], xx[m]] = True
    vmap = torch.zeros(n1 ** 3, dtype=torch.int32)
    vmap[inds] = torch.arange(inds.numel(), dtype=torch.int32)
    blob = _np(w['p']).astype(np.float32).reshape(-1)
    hidden = {3168: 32, 8288: 64}.get(blob.size)
    if hidden is None:
        raise ValueError(f'a real-time blob has 3168 (DIVeR32) or 8288 (DIVeR64) floats, '
                         f'got {blob.size}')
    s = SerializedScene(voxel_num, torch.as_tensor(w['f']).to(torch.float32).to(dev).contiguous(),
                        vmap.to(dev), mask.to(dev).contiguous(), blob, hidden)
    s._inds = inds.to(dev)
    return s
