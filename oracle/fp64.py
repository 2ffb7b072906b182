"""CPU ORACLE -- test infrastructure only (tests/, the smoke check and bench.py's CPU-baseline leg
use it as the checker; the product path never imports it).

FP64 restatements of the reference's arithmetic after traversal:
- ``integrate_einsum``      <- utils/integrator.py:62-113 (explicit grid)
- ``integrate_weights``     <- utils/mlp_evaluation/src/mlp_eval.cu:72-101 (corner-weight form)
- ``integrate_gauss``       <- the mathematics: mean of the trilinear interpolant over the segment,
                               2-point Gauss-Legendre (exact for the cubic integrand)
- ``positional_encoding``   <- model/mlps.py:6-18
- ``mlp_chain``             <- model/mlps.py:21-31 + model/diver.py:159-171 (offline decoder)
- ``rt_blob_decoder``       <- utils/mlp_evaluation{,64}/src/mlp_eval.cu:103-209 (+ convert.py
                               blob layout :63-140)
- ``baseline_decoder``      <- the benchmark decoder (softplus density, colour MLP on
                               [f[1:], PE(d)])
- ``composite``             <- model/diver.py:183-207
- ``render_front_to_back``  <- utils/mlp_evaluation/src/mlp_eval.cu:58-224 with configurable
                               thresholds (front-to-back, early termination)

The voxel corner used by every integral is ``min(int(p_in + 1e-4), int(p_out + 1e-4))`` evaluated in
float32 exactly as the reference does (utils/integrator.py:75, mlp_eval.cu:72-74), so corner indices
are bit-exact with the GPU path; everything after that is float64.
"""
from __future__ import annotations

import math

import numpy as np

F32_EPS = np.float32(1e-4)


def voxel_corner(coord_in: np.ndarray, coord_out: np.ndarray) -> np.ndarray:
    ci = (np.asarray(coord_in, np.float32) + F32_EPS).astype(np.int64)
    co = (np.asarray(coord_out, np.float32) + F32_EPS).astype(np.int64)
    return np.minimum(ci, co)


def _gather_corners(voxels: np.ndarray, pmin: np.ndarray):
    """F[dz, dy, dx, K, C] with the upper corner clamped to the grid (integrator.py:79)."""
    n1 = voxels.shape[0]
    lo = pmin
    hi = np.minimum(pmin + 1, n1 - 1)
    xs = (lo[:, 0], hi[:, 0])
    ys = (lo[:, 1], hi[:, 1])
    zs = (lo[:, 2], hi[:, 2])
    out = np.empty((2, 2, 2, pmin.shape[0], voxels.shape[-1]), np.float64)
    for dz in (0, 1):
        for dy in (0, 1):
            for dx in (0, 1):
                out[dz, dy, dx] = voxels[zs[dz], ys[dy], xs[dx]]
    return out


def integrate_einsum(voxels: np.ndarray, coord_in: np.ndarray, coord_out: np.ndarray) -> np.ndarray:
    """utils/integrator.py:62-113 in float64 (A/B/C term form)."""
    pmin = voxel_corner(coord_in, coord_out)
    F = _gather_corners(voxels, pmin)
    P = np.stack([np.asarray(coord_in, np.float64) - pmin,
                  np.asarray(coord_out, np.float64) - pmin], 0).transpose(0, 2, 1)  # [2,3,K]
    A = P.sum(0).prod(0) / 12 + P.prod(1).sum(0) / 6
    X = F[1] - F[0]
    X = X[0] - X[1]
    X = X[0] - X[1]
    feature = A[:, None] * X
    P_roll = P[:, [1, 2, 0]]
    B = (P * P_roll).sum(0) / 3 + (P * P_roll[[1, 0]]).sum(0) / 6
    Y = (F[0, 0, 0][None] - F[[0, 0, 1], [1, 1, 0], 0]
         - F[[0, 1, 0], 0, [1, 0, 1]] + F[[0, 1, 1], [1, 1, 0], [1, 0, 1]])
    feature = feature + np.einsum('dk,dkc->kc', B, Y)
    Cc = P.sum(0) / 2
    Z = F[[0, 0, 1], [0, 1, 0], [1, 0, 0]] - F[0, 0, 0][None]
    feature = feature + np.einsum('dk,dkc->kc', Cc, Z)
    return feature + F[0, 0, 0]


def corner_weights(coord_in: np.ndarray, coord_out: np.ndarray) -> tuple[np.ndarray, np.ndarray]:
    """mlp_eval.cu:72-101: the 8 corner weights (float64) and the voxel corner."""
    pmin = voxel_corner(coord_in, coord_out)
    a = np.asarray(coord_in, np.float64) - pmin
    b = np.asarray(coord_out, np.float64) - pmin
    x0, y0, z0 = a[:, 0], a[:, 1], a[:, 2]
    x1, y1, z1 = b[:, 0], b[:, 1], b[:, 2]
    z01, y01, x01 = z0 + z1, y0 + y1, x0 + x1
    w = np.empty((a.shape[0], 8))
    w[:, 7] = (2 * x0 * y0 * z0 + 2 * x1 * y1 * z1 + x01 * y01 * z01) / 12
    w2 = (y01 * z01 + y0 * z0 + y1 * z1) / 6
    w[:, 5] = (x01 * z01 + x0 * z0 + x1 * z1) / 6 - w[:, 7]
    w4 = (x01 * y01 + x0 * y0 + x1 * y1) / 6
    x01, y01, z01 = x01 * 0.5, y01 * 0.5, z01 * 0.5
    w[:, 4] = z01 - w[:, 5] - w2
    w[:, 0] = 1 - w[:, 4] - y01 - x01 + w4
    w[:, 6] = w2 - w[:, 7]
    w[:, 2] = y01 - w[:, 6] - w4
    w[:, 1] = x01 - w[:, 5] - w4
    w[:, 3] = -w[:, 7] + w4
    return w, pmin


def integrate_weights(voxels: np.ndarray, coord_in, coord_out) -> np.ndarray:
    w, pmin = corner_weights(coord_in, coord_out)
    n1 = voxels.shape[0]
    out = np.zeros((pmin.shape[0], voxels.shape[-1]))
    for j in range(8):
        dx, dy, dz = j & 1, (j >> 1) & 1, j >> 2
        z = np.minimum(pmin[:, 2] + dz, n1 - 1)
        y = np.minimum(pmin[:, 1] + dy, n1 - 1)
        x = np.minimum(pmin[:, 0] + dx, n1 - 1)
        out += w[:, j:j + 1] * voxels[z, y, x]
    return out


def trilerp(voxels: np.ndarray, pmin: np.ndarray, p: np.ndarray) -> np.ndarray:
    """Trilinear interpolation inside the voxel at ``pmin`` (local coordinates may extrapolate)."""
    n1 = voxels.shape[0]
    loc = p - pmin
    out = np.zeros((p.shape[0], voxels.shape[-1]))
    for j in range(8):
        dx, dy, dz = j & 1, (j >> 1) & 1, j >> 2
        wx = loc[:, 0] if dx else 1 - loc[:, 0]
        wy = loc[:, 1] if dy else 1 - loc[:, 1]
        wz = loc[:, 2] if dz else 1 - loc[:, 2]
        z = np.minimum(pmin[:, 2] + dz, n1 - 1)
        y = np.minimum(pmin[:, 1] + dy, n1 - 1)
        x = np.minimum(pmin[:, 0] + dx, n1 - 1)
        out += (wx * wy * wz)[:, None] * voxels[z, y, x]
    return out


def integrate_gauss(voxels: np.ndarray, coord_in, coord_out) -> np.ndarray:
    """Mean of the trilinear interpolant along [p_in, p_out] by 2-point Gauss-Legendre."""
    pmin = voxel_corner(coord_in, coord_out)
    a = np.asarray(coord_in, np.float64)
    b = np.asarray(coord_out, np.float64)
    s = 0.5 / math.sqrt(3.0)
    p1 = a + (0.5 - s) * (b - a)
    p2 = a + (0.5 + s) * (b - a)
    return 0.5 * (trilerp(voxels, pmin, p1) + trilerp(voxels, pmin, p2))


# ---------------------------------------------------------------------------------------------
# decoders
# ---------------------------------------------------------------------------------------------

def positional_encoding(x: np.ndarray, bands: int = 4) -> np.ndarray:
    """model/mlps.py:12-18: [x, sin(2^0 pi x), cos(2^0 pi x), ..., cos(2^(L-1) pi x)]."""
    x = np.asarray(x, np.float64)
    enc = [x]
    for l in range(bands):
        enc.append(np.sin((2 ** l * math.pi) * x))
        enc.append(np.cos((2 ** l * math.pi) * x))
    return np.concatenate(enc, -1)


def _linear(x, w, b):
    return x @ np.asarray(w, np.float64).T + np.asarray(b, np.float64)


def _relu(x):
    return np.maximum(x, 0.0)


def _sigmoid(x):
    return 1.0 / (1.0 + np.exp(-x))


def softplus(x):
    return np.logaddexp(0.0, x)


def mlp_chain(sd: dict, feature: np.ndarray, d: np.ndarray, bands: int = 4):
    """Offline decoder: (sigma_raw, rgb) per row (model/diver.py:159-171)."""
    x = np.asarray(feature, np.float64)
    i = 0
    while f'mlp1.{i + 2}.weight' in sd:
        x = _relu(_linear(x, sd[f'mlp1.{i}.weight'], sd[f'mlp1.{i}.bias']))
        i += 2
    x = _linear(x, sd[f'mlp1.{i}.weight'], sd[f'mlp1.{i}.bias'])
    sigma, x = x[:, 0], x[:, 1:]
    x = np.concatenate([x, positional_encoding(d, bands)], -1)
    j = 0
    while f'mlp2.{j + 2}.weight' in sd:
        x = _relu(_linear(x, sd[f'mlp2.{j}.weight'], sd[f'mlp2.{j}.bias']))
        j += 2
    x = _linear(x, sd[f'mlp2.{j}.weight'], sd[f'mlp2.{j}.bias'])
    return sigma, _sigmoid(x)


def baseline_decoder(layers, feature: np.ndarray, d: np.ndarray, bands: int = 4):
    """Benchmark decoder: sigma_raw = f[0] (softplus applied by the caller); colour MLP on
    [f[1:], PE(d)] with ReLU between layers and a final sigmoid."""
    f = np.asarray(feature, np.float64)
    x = np.concatenate([f[:, 1:], positional_encoding(d, bands)], -1)
    for k, (w, b) in enumerate(layers):
        x = _linear(x, w, b)
        if k + 1 < len(layers):
            x = _relu(x)
    return f[:, 0], _sigmoid(x)


_PE_KERNEL_ORDER = np.array([0, 3, 9, 15, 21, 6, 12, 18, 24,
                             1, 4, 10, 16, 22, 7, 13, 19, 25,
                             2, 5, 11, 17, 23, 8, 14, 20, 26])


def rt_blob_decoder(params: np.ndarray, hidden: int, feature: np.ndarray, d: np.ndarray):
    """The real-time kernels' decoder from the serialized blob (mlp_eval.cu:103-209).

    For hidden=32 ``feature`` is the *pre-multiplied* integral (features @ W0^T, convert.py:149-150)
    and the blob starts with the layer-0 bias; for hidden=64 ``feature`` is the raw integral.
    Returns (sigma_raw, rgb); alpha = 1 - exp(-relu(sigma_raw)).
    """
    p = np.asarray(params, np.float64)
    f = np.asarray(feature, np.float64)
    H = hidden
    if H == 32:
        b0 = p[0:32]
        h1 = _relu(f + b0)
        off = 32
        b1 = p[off:off + 32]; off += 32
        w1t = p[off:off + 32 * 32].reshape(32, 32); off += 32 * 32
        h2 = h1 @ w1t + b1
    else:
        b0 = p[0:64]
        off = 64
        w0t = p[off:off + 32 * 64].reshape(32, 64); off += 32 * 64
        h2 = f @ w0t + b0
    bs = p[off]; off += 1
    ws = p[off:off + H]; off += H
    sigma = _relu(h2) @ ws + bs
    b3 = p[off:off + H]; off += H
    w2t = p[off:off + H * H].reshape(H, H); off += H * H
    wv = p[off:off + 27 * H].reshape(27, H); off += 27 * H
    pe = positional_encoding(d, 4)[:, _PE_KERNEL_ORDER]
    h3 = _relu(h2) @ w2t + b3 + pe @ wv
    bo = p[off:off + 3]; off += 3
    wot = p[off:off + H * 3].reshape(H, 3)
    out = _relu(h3) @ wot + bo
    return sigma, _sigmoid(out)


# ---------------------------------------------------------------------------------------------
# compositing
# ---------------------------------------------------------------------------------------------

def composite(sigma: np.ndarray, color: np.ndarray, mask: np.ndarray, white_back: bool = True):
    """model/diver.py:183-207 in float64: returns (rgb [B,3], weight [B,M])."""
    s = np.maximum(np.asarray(sigma, np.float64), 0.0)
    alpha = 1.0 - np.exp(-s * mask)
    shifted = np.concatenate([np.ones((alpha.shape[0], 1)), 1.0 - alpha[:, :-1]], 1)
    weight = alpha * np.cumprod(shifted, -1)
    rgb = (weight[:, :, None] * color).sum(1)
    if white_back:
        rgb = rgb + (1.0 - weight.sum(1, keepdims=True))
    return rgb, weight


def render_front_to_back(ray_index: np.ndarray, alpha: np.ndarray, color: np.ndarray,
                         n_rays: int, t_stop: float = 1e-4, alpha_skip: float = 0.0,
                         rel: float = 1e-3, return_near: bool = False):
    """Front-to-back compositing of rows sorted by (ray, depth) with early termination.

    A row whose alpha is below ``alpha_skip`` is skipped entirely (transmittance unchanged,
    mlp_eval.cu:153-155); a ray stops after the row that takes its transmittance below ``t_stop``
    (mlp_eval.cu:218-222).  Returns (rgb_accum [R,3], T [R], rows_used [R]), and with
    ``return_near`` also near [R]: the ray met a row whose alpha lies within ``rel`` (relative) of
    alpha_skip or whose transmittance after the row lies within ``rel`` of t_stop, so an FP32
    evaluation may take the other branch there (skip or not, stop or go on).
    """
    rgb = np.zeros((n_rays, 3))
    T = np.ones(n_rays)
    used = np.zeros(n_rays, np.int64)
    done = np.zeros(n_rays, bool)
    near = np.zeros(n_rays, bool)
    for k in range(ray_index.shape[0]):
        r = ray_index[k]
        if done[r]:
            continue
        used[r] += 1
        a = alpha[k]
        if alpha_skip > 0 and abs(a - alpha_skip) <= rel * alpha_skip:
            near[r] = True
        if a < alpha_skip:
            continue
        rgb[r] += T[r] * a * color[k]
        T[r] = T[r] * (1.0 - a)
        if t_stop > 0 and abs(T[r] - t_stop) <= rel * t_stop:
            near[r] = True
        if T[r] < t_stop:
            done[r] = True
    if return_near:
        return rgb, T, used, near
    return rgb, T, used


def rt_blob_from_state(sd: dict, hidden: int) -> np.ndarray:
    """The real-time blob convert.py:63-140 builds from the offline decoder's state dict
    (layer fusion of mlp1's last layer into mlp2's first, PE columns reordered per axis, 28 zeros
    of padding).  For hidden=32 the features must be pre-multiplied by mlp1.0 (convert.py:149-150)."""
    sd = {k: np.asarray(v, np.float64) for k, v in sd.items()}
    H = hidden
    last = 'mlp1.4' if H == 32 else 'mlp1.2'
    l21, b21 = sd[last + '.weight'], sd[last + '.bias']
    l22, b22 = sd['mlp2.0.weight'], sd['mlp2.0.bias']
    l4 = np.concatenate([l21[:1], l22[:, :H] @ l21[1:]], 0)
    b4 = np.concatenate([b21[:1], l22[:, :H] @ b21[1:] + b22], 0)
    w1 = np.concatenate([b4[0:1], l4[0]], 0)
    w2 = np.concatenate([b4[1:, None], l4[1:], l22[:, H:]], 1)
    w2 = np.concatenate([w2[:, :-27], w2[:, -27:][:, _PE_KERNEL_ORDER]], -1)
    w3 = np.concatenate([sd['mlp2.2.bias'][:, None], sd['mlp2.2.weight']], -1)
    if H == 64:
        w0 = np.concatenate([sd['mlp1.0.bias'][:, None], sd['mlp1.0.weight']], 1)
        return np.concatenate([w0.T.reshape(-1), w1, w2.T.reshape(-1), w3.T.reshape(-1),
                               np.zeros(28)])
    win = np.concatenate([sd['mlp1.0.bias'][:, None], sd['mlp1.0.weight']], 1)
    w0 = np.concatenate([sd['mlp1.2.bias'][:, None], sd['mlp1.2.weight']], 1)
    return np.concatenate([win[:, 0], w0.T.reshape(-1), w1, w2.T.reshape(-1), w3.T.reshape(-1),
                           np.zeros(28)])


def mlp_eval_round(coords, voxels_dense, blob, hidden, d, rgba, finish, alpha_skip, t_stop,
                   near=None, rel=1e-3):
    """One mlp_eval call (mlp_eval.cu:50-230) in float64, in place on rgba / finish (numpy).

    coords [8, B, 6]; voxels_dense [(N+1)^3 grid, C] already in the blob's feature space
    (pre-multiplied for hidden=32).  ``near`` (bool [B], optional) is set for rays that meet an
    alpha within ``rel`` of alpha_skip or a transmittance within ``rel`` of t_stop."""
    B = coords.shape[1]
    for b in range(B):
        if finish[b]:
            continue
        for h in range(8):
            c = coords[h, b]
            if c[0] < 0:
                finish[b] = 1
                break
            F = integrate_weights(voxels_dense, c[None, :3], c[None, 3:])
            s, rgb = rt_blob_decoder(blob, hidden, F, np.asarray(d[b:b + 1], np.float64))
            alpha = 1.0 - math.exp(-max(s[0], 0.0))
            if near is not None and alpha_skip > 0 and abs(alpha - alpha_skip) <= rel * alpha_skip:
                near[b] = True
            if alpha < alpha_skip:
                continue
            T = rgba[b, 3]
            rgba[b, :3] += T * alpha * rgb[0]
            rgba[b, 3] = T * (1.0 - alpha)
            if near is not None and t_stop > 0 and abs(rgba[b, 3] - t_stop) <= rel * t_stop:
                near[b] = True
            if rgba[b, 3] < t_stop:
                finish[b] = 1
                break
