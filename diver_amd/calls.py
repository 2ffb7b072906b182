"""Tensor-level calls into libdiver (device tensors in, device tensors out).

Each function validates device, dtype and contiguity, then enqueues the library call on torch's
current stream.  There is no CPU fallback: CPU tensors raise.
"""
from __future__ import annotations

import ctypes

import torch

from . import _lib


def _dev(*ts):
    dev = None
    for t in ts:
        if t is None:
            continue
        if not t.is_cuda:
            raise ValueError('libdiver operates on device tensors; got a CPU tensor')
        if dev is None:
            dev = t.device
        elif t.device != dev:
            raise ValueError('tensors are on different devices')
    return dev


def _f32(t, name):
    if t.dtype != torch.float32:
        raise TypeError(f'{name} must be float32, got {t.dtype}')
    return t.contiguous()


def _u8(t, name):
    if t.dtype not in (torch.bool, torch.uint8):
        raise TypeError(f'{name} must be bool/uint8, got {t.dtype}')
    return t.contiguous()


def _i32(t, name):
    if t.dtype != torch.int32:
        raise TypeError(f'{name} must be int32, got {t.dtype}')
    return t.contiguous()


def _s(t):
    return _lib.stream_of(t)


def masked_intersect_raw(o, v, mask, xyzmin, voxel_num, voxel_size, mask_scale, max_n,
                         counts_only=False):
    """One launch of diver_masked_intersect; returns (isect, num, tns) (isect/tns None when
    counts_only)."""
    dev = _dev(o, v, mask)
    o, v, mask = _f32(o, 'o'), _f32(v, 'v'), _u8(mask, 'mask')
    B = o.shape[0]
    num = torch.empty(B, dtype=torch.int32, device=dev)
    isect = tns = None
    if not counts_only:
        isect = torch.empty((B, max_n, 6), dtype=torch.float32, device=dev)
        tns = torch.empty((B, max_n), dtype=torch.float32, device=dev)
    _lib.call('diver_masked_intersect', B, int(max_n), float(xyzmin), 0.0, float(voxel_num),
              float(voxel_size), float(mask_scale), o.data_ptr(), v.data_ptr(), mask.data_ptr(),
              _lib.ptr(isect), num.data_ptr(), _lib.ptr(tns), _s(o))
    return isect, num, tns


def ray_voxel_intersect_raw(o, v, xyzmin, voxel_num, voxel_size, max_n, counts_only=False):
    dev = _dev(o, v)
    o, v = _f32(o, 'o'), _f32(v, 'v')
    B = o.shape[0]
    num = torch.empty(B, dtype=torch.int32, device=dev)
    isect = tns = None
    if not counts_only:
        isect = torch.empty((B, max_n, 3), dtype=torch.float32, device=dev)
        tns = torch.empty((B, max_n), dtype=torch.float32, device=dev)
    _lib.call('diver_ray_voxel_intersect', B, int(max_n), float(xyzmin), 0.0, float(voxel_num),
              float(voxel_size), o.data_ptr(), v.data_ptr(), _lib.ptr(isect), num.data_ptr(),
              _lib.ptr(tns), _s(o))
    return isect, num, tns


def aabb_intersect(coords, v, center, finish, rgba, octrees, voxel_num):
    _dev(coords, v, center, finish, rgba, octrees)
    for t, n in ((coords, 'coords'), (v, 'v'), (center, 'center'), (rgba, 'rgba')):
        if t.dtype != torch.float32 or not t.is_contiguous():
            raise TypeError(f'{n} must be a contiguous float32 tensor (modified in place)')
    if finish.dtype not in (torch.bool, torch.uint8) or not finish.is_contiguous():
        raise TypeError('finish must be a contiguous bool tensor (modified in place)')
    octrees = _u8(octrees, 'octrees')
    B = v.numel() // 3
    _lib.call('diver_aabb_intersect', B, coords.data_ptr(), v.data_ptr(), center.data_ptr(),
              finish.data_ptr(), rgba.data_ptr(), octrees.data_ptr(), int(voxel_num), _s(v))


REF_THREAD_STRIDE = 256 * 1024  # grid-stride width of the reference's ray_march launch


def ray_march(coords, v, mask, finish, voxel_num, thread_stride=REF_THREAD_STRIDE):
    _dev(coords, v, mask, finish)
    if coords.dtype != torch.float32 or not coords.is_contiguous():
        raise TypeError('coords must be a contiguous float32 tensor (modified in place)')
    v, mask, finish = _f32(v, 'v'), _u8(mask, 'mask'), _u8(finish, 'finish')
    B = v.numel() // 3
    _lib.call('diver_ray_march', B, coords.data_ptr(), v.data_ptr(), mask.data_ptr(),
              finish.data_ptr(), int(voxel_num), int(thread_stride), _s(v))


def integrate(voxels, coord_in, coord_out):
    dev = _dev(voxels, coord_in, coord_out)
    voxels = _f32(voxels, 'voxels')
    ci, co = _f32(coord_in, 'coord_in'), _f32(coord_out, 'coord_out')
    n1, C = voxels.shape[0], voxels.shape[-1]
    K = ci.shape[0]
    out = torch.empty((K, C), dtype=torch.float32, device=dev)
    _lib.call('diver_integrate', voxels.data_ptr(), n1, C, K, ci.data_ptr(), co.data_ptr(),
              out.data_ptr(), _s(voxels))
    return out


def integrate_backward(voxels_shape, coord_in, coord_out, grad_feature, grad_voxels=None):
    dev = _dev(coord_in, coord_out, grad_feature)
    ci, co = _f32(coord_in, 'coord_in'), _f32(coord_out, 'coord_out')
    g = _f32(grad_feature, 'grad_feature')
    if grad_voxels is None:
        grad_voxels = torch.zeros(voxels_shape, dtype=torch.float32, device=dev)
    n1, C = voxels_shape[0], voxels_shape[-1]
    _lib.call('diver_integrate_backward', n1, C, ci.shape[0], ci.data_ptr(), co.data_ptr(),
              g.data_ptr(), grad_voxels.data_ptr(), _s(g))
    return grad_voxels


def composite(sigma, color, mask, white_back):
    dev = _dev(sigma, color, mask)
    sigma, color, mask = _f32(sigma, 'sigma'), _f32(color, 'color'), _u8(mask, 'mask')
    B, M = sigma.shape
    rgb = torch.empty((B, 3), dtype=torch.float32, device=dev)
    weight = torch.empty((B, M), dtype=torch.float32, device=dev)
    _lib.call('diver_composite', B, M, sigma.data_ptr(), color.data_ptr(), mask.data_ptr(),
              int(bool(white_back)), rgb.data_ptr(), weight.data_ptr(), _s(sigma))
    return rgb, weight


def composite_backward(sigma, color, mask, white_back, grad_rgb, grad_weight=None):
    dev = _dev(sigma, color, mask, grad_rgb)
    sigma, color, mask = _f32(sigma, 'sigma'), _f32(color, 'color'), _u8(mask, 'mask')
    grad_rgb = _f32(grad_rgb, 'grad_rgb')
    if grad_weight is not None:
        grad_weight = _f32(grad_weight, 'grad_weight')
    B, M = sigma.shape
    gs = torch.empty((B, M), dtype=torch.float32, device=dev)
    gc = torch.empty((B, M, 3), dtype=torch.float32, device=dev)
    _lib.call('diver_composite_backward', B, M, sigma.data_ptr(), color.data_ptr(),
              mask.data_ptr(), int(bool(white_back)), grad_rgb.data_ptr(),
              _lib.ptr(grad_weight), gs.data_ptr(), gc.data_ptr(), _s(sigma))
    return gs, gc


def mlp_eval(decoder, rgba, coord, features, voxel_map, v, finish, alpha_skip, t_stop):
    _dev(rgba, coord, features, voxel_map, v, finish)
    if rgba.dtype != torch.float32 or not rgba.is_contiguous():
        raise TypeError('rgba must be a contiguous float32 tensor (modified in place)')
    if finish.dtype not in (torch.bool, torch.uint8) or not finish.is_contiguous():
        raise TypeError('finish must be a contiguous bool tensor (modified in place)')
    coord, features, v = _f32(coord, 'coord'), _f32(features, 'features'), _f32(v, 'v')
    voxel_map = _i32(voxel_map, 'voxel_map')
    B = v.numel() // 3
    _lib.call('diver_mlp_eval', decoder.handle, B, coord.data_ptr(), features.data_ptr(),
              voxel_map.data_ptr(), int(voxel_map.shape[0]), v.data_ptr(), rgba.data_ptr(),
              finish.data_ptr(), float(alpha_skip), float(t_stop), _s(v))


PRECISIONS = {'fp32': 0, 'bf16': 1, 'fp16': 2}
TRAVERSALS = {'offline': 0, 'realtime': 1}


def grid_struct(voxels, mask, octrees, xyzmin, voxel_size, blocks=None, bricks=None, tail=None):
    n = mask.shape[0]
    p = lambda t: t.data_ptr() if t is not None else None
    return _lib.Grid(voxels.data_ptr(), mask.data_ptr(), p(octrees), p(blocks), p(bricks), int(n),
                     int(voxels.shape[-1]), float(xyzmin), float(voxel_size), p(tail))


def brick_bits(mask):
    """Bit-packed mask, one int64 word per 4^3 brick (see include/diver.h)."""
    _dev(mask)
    mask = _u8(mask, 'mask')
    n = mask.shape[0]
    bricks = torch.empty((n // 4) ** 3, dtype=torch.int64, device=mask.device)
    _lib.call('diver_brick_bits', mask.data_ptr(), int(n), bricks.data_ptr(), _s(mask))
    return bricks


def tail_bits(blocks, n):
    """Tail bits of a mask from its 8^3-block bits (n a multiple of 8): the block bits dilated by
    one block (int32 words, same numbering).  A grid carrying them ends each walk once nothing
    occupied can lie ahead (see include/diver.h)."""
    _dev(blocks)
    if blocks.dtype != torch.int32:
        raise TypeError('blocks must be int32 words')
    nb = int(n) // 8
    tail = torch.empty((nb ** 3 + 31) // 32, dtype=torch.int32, device=blocks.device)
    _lib.call('diver_tail_bits', blocks.data_ptr(), int(n), tail.data_ptr(), _s(blocks))
    return tail


def block_bits(mask):
    """8^3-block occupancy bits of an [n^3] mask (int32 words, bit b of word b // 32)."""
    _dev(mask)
    mask = _u8(mask, 'mask')
    n = mask.shape[0]
    nb = n // 8
    words = (nb ** 3 + 31) // 32
    bits = torch.empty(words, dtype=torch.int32, device=mask.device)
    _lib.call('diver_block_bits', mask.data_ptr(), int(n), bits.data_ptr(), _s(mask))
    return bits


def _grid_of(grid):
    return grid_struct(grid.voxels, grid.mask, grid.octrees, grid.xyzmin, grid.voxel_size,
                       getattr(grid, 'blocks', None), getattr(grid, 'bricks', None),
                       getattr(grid, 'tail', None))


def render_opts(precision='fp32', traversal='realtime', t_stop=1e-4, alpha_skip=0.0):
    return _lib.RenderOpts(PRECISIONS[precision], TRAVERSALS[traversal], float(t_stop),
                           float(alpha_skip))


def walk_rays(grid, o, d, traversal='realtime', max_n=None):
    """The occupied intervals the fused renderer composites: (isect [R,max_n,6], num [R])."""
    dev = _dev(grid.voxels, grid.mask, o, d)
    o, d = _f32(o, 'o'), _f32(d, 'd')
    R = o.shape[0]
    max_n = 3 * grid.n if max_n is None else max_n
    isect = torch.empty((R, max_n, 6), dtype=torch.float32, device=dev)
    num = torch.empty(R, dtype=torch.int32, device=dev)
    g = _grid_of(grid)
    _lib.call('diver_walk_rays', ctypes.byref(g), TRAVERSALS[traversal], R, o.data_ptr(),
              d.data_ptr(), int(max_n), isect.data_ptr(), num.data_ptr(), _s(o))
    return isect, num


def ray_suffix(row_offset, count, weight, color, background=None):
    """Per-ray suffix sums over the adjoint's rows (diver_ray_suffix): out[k] = sum over the rows
    j > k of the same ray of weight[j] * color[j], plus background[ray]; [K, 3] float32."""
    tensors = [row_offset, count, weight, color] + ([background] if background is not None else [])
    dev = _dev(*tensors)
    count = _i32(count, 'count')
    if row_offset.dtype != torch.int64:
        raise TypeError('row_offset must be int64')
    row_offset = row_offset.contiguous()
    w, c = _f32(weight, 'weight'), _f32(color, 'color')
    if c.dim() != 2 or c.shape[1] != 3 or w.shape[0] != c.shape[0]:
        raise ValueError('color must be [K, 3] with K = weight.shape[0]')
    if count.numel() != row_offset.numel():
        raise ValueError('count and row_offset sizes differ')
    bg = _f32(background, 'background') if background is not None else None
    if bg is not None and bg.numel() != count.numel():
        raise ValueError('background must have one value per ray')
    out = torch.empty((w.shape[0], 3), dtype=torch.float32, device=dev)
    _lib.call('diver_ray_suffix', count.numel(), row_offset.data_ptr(), count.data_ptr(),
              w.data_ptr(), c.data_ptr(), bg.data_ptr() if bg is not None else None,
              out.data_ptr(), _s(w))
    return out


def adjoint_decoder_forward(decoder, rows, ray, d):
    """Forward half of the decoder adjoint over the rows of render_rows (diver_adjoint_decoder_
    forward): (r1 [K, H], r2 [K, H], color [K, 3], weight [K]) for K = rows.shape[0]."""
    dev = _dev(rows, ray, d)
    rows, d = _f32(rows, 'rows'), _f32(d, 'd')
    if ray.dtype != torch.int64 or ray.numel() != rows.shape[0]:
        raise TypeError('ray must be int64 with one entry per row')
    ray = ray.contiguous()
    K, H = rows.shape[0], decoder.hidden
    r1 = torch.empty((K, H), dtype=torch.float32, device=dev)
    r2 = torch.empty((K, H), dtype=torch.float32, device=dev)
    color = torch.empty((K, 3), dtype=torch.float32, device=dev)
    weight = torch.empty((K,), dtype=torch.float32, device=dev)
    _lib.call('diver_adjoint_decoder_forward', decoder.handle, K, rows.data_ptr(), ray.data_ptr(),
              d.data_ptr(), r1.data_ptr(), r2.data_ptr(), color.data_ptr(), weight.data_ptr(),
              _s(rows))
    return r1, r2, color, weight


def adjoint_decoder_backward(decoder, rows, ray, n_valid, grad_rgb, suffix, r1, r2, color, weight,
                             sparsity_scale=None):
    """Backward half of the decoder adjoint (diver_adjoint_decoder_backward): (d_feature
    [n_valid, C], d_out [K, 3], d_s [K], d_r1 [K, H], d_r2 [K, H]).  sparsity_scale: a 0-dim
    device tensor (dL/d(sigma_loss) / (B M)) or None."""
    tensors = [rows, ray, grad_rgb, suffix, r1, r2, color, weight]
    dev = _dev(*tensors)
    rows, g, suffix = _f32(rows, 'rows'), _f32(grad_rgb, 'grad_rgb'), _f32(suffix, 'suffix')
    r1, r2, color, weight = (_f32(x, n) for x, n in ((r1, 'r1'), (r2, 'r2'), (color, 'color'),
                                                       (weight, 'weight')))
    ray = ray.contiguous()
    K, H, C = rows.shape[0], decoder.hidden, decoder.channels
    n_valid = int(n_valid)
    if not 0 <= n_valid <= K or suffix.shape[0] < n_valid:
        raise ValueError('need 0 <= n_valid <= rows and suffix rows >= n_valid')
    sc = None
    if sparsity_scale is not None:
        sc = _f32(sparsity_scale.reshape(1), 'sparsity_scale')
    d_feature = torch.empty((n_valid, C), dtype=torch.float32, device=dev)
    d_out = torch.empty((K, 3), dtype=torch.float32, device=dev)
    d_s = torch.empty((K,), dtype=torch.float32, device=dev)
    d_r1 = torch.empty((K, H), dtype=torch.float32, device=dev)
    d_r2 = torch.empty((K, H), dtype=torch.float32, device=dev)
    _lib.call('diver_adjoint_decoder_backward', decoder.handle, K, n_valid, rows.data_ptr(),
              ray.data_ptr(), g.data_ptr(), suffix.data_ptr(), _lib.ptr(sc), r1.data_ptr(),
              r2.data_ptr(), color.data_ptr(), weight.data_ptr(), d_feature.data_ptr(),
              d_out.data_ptr(), d_s.data_ptr(), d_r1.data_ptr(), d_r2.data_ptr(), _s(rows))
    return d_feature, d_out, d_s, d_r1, d_r2


def adjoint_weight_grads(decoder, n_rows, rows, ray, d, r1, r2, d_out, d_s, d_r1, d_r2,
                         no_alloc=False):
    """Weight gradients of the decoder adjoint over the first n_rows rows
    (diver_adjoint_weight_grads): the canonical parameter gradient (as many elements as the
    decoder's parameters) of dW_c, dW_v, dW_2, dW_o, dw_s and the biases, summed in a fixed order.
    With no_alloc the library's own partial-sum array is used (diver_adjoint_weight_grads_noalloc;
    calls on one device must not overlap) instead of a workspace allocated here."""
    tensors = [rows, ray, d, r1, r2, d_out, d_s, d_r1, d_r2]
    dev = _dev(*tensors)
    n = int(n_rows)
    for t, name in ((rows, 'rows'), (r1, 'r1'), (r2, 'r2'), (d_out, 'd_out'), (d_s, 'd_s'),
                    (d_r1, 'd_r1'), (d_r2, 'd_r2')):
        if t.dtype != torch.float32 or not t.is_contiguous():
            raise TypeError(f'{name} must be a contiguous float32 tensor')
        if t.shape[0] < n:
            raise ValueError(f'{name} has fewer than n_rows={n} rows')
    if ray.dtype != torch.int64 or not ray.is_contiguous() or ray.shape[0] < n:
        raise TypeError('ray must be a contiguous int64 tensor of at least n_rows entries')
    d = _f32(d, 'd')
    grad = torch.empty((int(decoder.canon.params.size),), dtype=torch.float32, device=dev)
    if no_alloc:
        _lib.call('diver_adjoint_weight_grads_noalloc', decoder.handle, n, rows.data_ptr(),
                  ray.data_ptr(), d.data_ptr(), r1.data_ptr(), r2.data_ptr(), d_out.data_ptr(),
                  d_s.data_ptr(), d_r1.data_ptr(), d_r2.data_ptr(), grad.data_ptr(), _s(rows))
        return grad
    L = _lib.lib()
    need = L.diver_adjoint_weight_grads_workspace(decoder.handle, n)
    if need < 0:
        raise ValueError('the weight gradients need a decoder without pre-layers and '
                         'channels/hidden 32/32, 32/64 or 64/64')
    ws = torch.empty((max(need, 4) // 4,), dtype=torch.float32, device=dev)
    _lib.call('diver_adjoint_weight_grads', decoder.handle, n, rows.data_ptr(), ray.data_ptr(),
              d.data_ptr(), r1.data_ptr(), r2.data_ptr(), d_out.data_ptr(), d_s.data_ptr(),
              d_r1.data_ptr(), d_r2.data_ptr(), grad.data_ptr(), ws.data_ptr(), ws.numel() * 4,
              _s(rows))
    return grad


def positional_encoding(d):
    """PE(d) [R, 27] as the fused kernels evaluate it (diver_positional_encoding)."""
    dev = _dev(d)
    d = _f32(d, 'd')
    pe = torch.empty((d.shape[0], 27), dtype=torch.float32, device=dev)
    _lib.call('diver_positional_encoding', d.shape[0], d.data_ptr(), pe.data_ptr(), _s(d))
    return pe


def render_rows(grid, decoder, o, d, opts, count, row_offset, n_rows, capacity=None):
    """Recompute pass of the adjoint: rows [n_rows, 9 + C] (see include/diver.h).  With capacity
    > n_rows the tensor has capacity rows and rows n_rows.. are zero."""
    dev = _dev(grid.voxels, grid.mask, o, d, count, row_offset)
    o, d = _f32(o, 'o'), _f32(d, 'd')
    count = _i32(count, 'count')
    if row_offset.dtype != torch.int64:
        raise TypeError('row_offset must be int64')
    row_offset = row_offset.contiguous()
    cap = int(n_rows) if capacity is None else max(int(capacity), int(n_rows))
    rows = torch.empty((cap, 9 + grid.channels), dtype=torch.float32, device=dev)
    if cap > n_rows:
        rows[int(n_rows):].zero_()
    g = _grid_of(grid)
    _lib.call('diver_render_rows', ctypes.byref(g), decoder.handle, ctypes.byref(opts),
              o.shape[0], o.data_ptr(), d.data_ptr(), count.data_ptr(), row_offset.data_ptr(),
              rows.data_ptr(), _s(o))
    return rows


def decoder_update(decoder, params, first: int = 0):
    """Copy canonical parameters (device float32 tensor, the whole canonical vector) into the
    decoder handle.  With first > 0 only params[first:] is copied (diver_decoder_update_range);
    first >= channels + 1 leaves the density row, and with it the channel-0 density path, alone."""
    _dev(params)
    p = _f32(params.detach(), 'params')
    if first == 0:
        _lib.call('diver_decoder_update', decoder.handle, p.data_ptr(), p.numel(), _s(p))
    else:
        tail = p[first:].contiguous()
        _lib.call('diver_decoder_update_range', decoder.handle, tail.data_ptr(), first,
                  tail.numel(), _s(p))


def render_rays(grid, decoder, o, d, opts, with_count=True):
    """grid: an object with voxels, mask, octrees, xyzmin, voxel_size (renderer.SceneGrid)."""
    dev = _dev(grid.voxels, grid.mask, o, d)
    o, d = _f32(o, 'o'), _f32(d, 'd')
    R = o.shape[0]
    rgba = torch.empty((R, 4), dtype=torch.float32, device=dev)
    count = torch.empty(R, dtype=torch.int32, device=dev) if with_count else None
    g = _grid_of(grid)
    _lib.call('diver_render_rays', ctypes.byref(g), decoder.handle, ctypes.byref(opts), R,
              o.data_ptr(), d.data_ptr(), rgba.data_ptr(), _lib.ptr(count), _s(o))
    return rgba, count


def render_rays_sparsity(grid, decoder, o, d, opts):
    """render_rays plus, per ray, the sum of log(1 + s^2 / 0.5) over its evaluated intervals
    (s = density pre-activation): (rgba [R, 4], count [R], sparsity [R])."""
    dev = _dev(grid.voxels, grid.mask, o, d)
    o, d = _f32(o, 'o'), _f32(d, 'd')
    R = o.shape[0]
    rgba = torch.empty((R, 4), dtype=torch.float32, device=dev)
    count = torch.empty(R, dtype=torch.int32, device=dev)
    sparsity = torch.empty(R, dtype=torch.float32, device=dev)
    g = _grid_of(grid)
    _lib.call('diver_render_rays_sparsity', ctypes.byref(g), decoder.handle, ctypes.byref(opts),
              R, o.data_ptr(), d.data_ptr(), rgba.data_ptr(), count.data_ptr(),
              sparsity.data_ptr(), _s(o))
    return rgba, count, sparsity


def render_rays_segments(grid, decoder, o, d, opts, seg_cap, sparsity=True):
    """The training forward with each evaluated interval's (entry, exit) recorded
    (diver_render_rays_segments): (rgba [R, 4], count [R], sparsity [R] or None,
    segments [R, seg_cap, 6])."""
    dev = _dev(grid.voxels, grid.mask, o, d)
    o, d = _f32(o, 'o'), _f32(d, 'd')
    R = o.shape[0]
    rgba = torch.empty((R, 4), dtype=torch.float32, device=dev)
    count = torch.empty(R, dtype=torch.int32, device=dev)
    sreg = torch.empty(R, dtype=torch.float32, device=dev) if sparsity else None
    segments = torch.empty((R, int(seg_cap), 6), dtype=torch.float32, device=dev)
    g = _grid_of(grid)
    _lib.call('diver_render_rays_segments', ctypes.byref(g), decoder.handle, ctypes.byref(opts),
              R, o.data_ptr(), d.data_ptr(), rgba.data_ptr(), count.data_ptr(), _lib.ptr(sreg),
              segments.data_ptr(), int(seg_cap), _s(o))
    return rgba, count, sreg, segments


def render_rows_segments(grid, decoder, opts, count, row_offset, n_rows, segments, with_ray=False):
    """diver_render_rows from recorded segments [R, seg_cap, 6] (every count <= seg_cap):
    rows [n_rows, 9 + C], the same as render_rows' re-walk; with_ray: (rows, ray [n_rows] int64,
    the ray of each row)."""
    dev = _dev(grid.voxels, grid.mask, count, row_offset, segments)
    count = _i32(count, 'count')
    if row_offset.dtype != torch.int64:
        raise TypeError('row_offset must be int64')
    row_offset = row_offset.contiguous()
    segments = _f32(segments, 'segments')
    R, cap = segments.shape[0], segments.shape[1]
    if count.shape[0] != R:
        raise ValueError('count and segments disagree on the ray count')
    rows = torch.empty((int(n_rows), 9 + grid.channels), dtype=torch.float32, device=dev)
    ray = torch.empty((int(n_rows),), dtype=torch.int64, device=dev) if with_ray else None
    g = _grid_of(grid)
    _lib.call('diver_render_rows_segments', ctypes.byref(g), decoder.handle, ctypes.byref(opts), R,
              count.data_ptr(), row_offset.data_ptr(), segments.data_ptr(), int(cap),
              rows.data_ptr(), _lib.ptr(ray), _s(count))
    return (rows, ray) if with_ray else rows


def corner_keys(n1, coord_in, coord_out):
    """int64 [K, 8]: keys z n1^2 + y n1 + x of the 8 corners of each row's integral."""
    dev = _dev(coord_in, coord_out)
    ci, co = _f32(coord_in, 'coord_in'), _f32(coord_out, 'coord_out')
    K = ci.shape[0]
    keys = torch.empty((K, 8), dtype=torch.int64, device=dev)
    _lib.call('diver_corner_keys', int(n1), K, ci.data_ptr(), co.data_ptr(), keys.data_ptr(), _s(ci))
    return keys


def integrate_table(table, corner_row, coord_in, coord_out):
    """[K, C]: the segment integral with corner j of row r taken from table[corner_row[r, j]]."""
    dev = _dev(table, corner_row, coord_in, coord_out)
    table = _f32(table, 'table')
    ci, co = _f32(coord_in, 'coord_in'), _f32(coord_out, 'coord_out')
    if corner_row.dtype != torch.int64:
        raise TypeError('corner_row must be int64')
    crow = corner_row.contiguous()
    K, C = ci.shape[0], table.shape[-1]
    if table.dim() != 2 or tuple(crow.shape) != (K, 8) or tuple(co.shape) != (K, 3):
        raise ValueError('table must be [U, C], corner_row [K, 8] and coordinates [K, 3]')
    out = torch.empty((K, C), dtype=torch.float32, device=dev)
    _lib.call('diver_integrate_table', table.data_ptr(), int(C), K, crow.data_ptr(), ci.data_ptr(),
              co.data_ptr(), out.data_ptr(), _s(ci))
    return out


def integrate_table_backward(n_rows_table, corner_row, coord_in, coord_out, grad_feature):
    """[U, C]: d feature / d table ^T grad_feature."""
    dev = _dev(corner_row, coord_in, coord_out, grad_feature)
    ci, co = _f32(coord_in, 'coord_in'), _f32(coord_out, 'coord_out')
    g = _f32(grad_feature, 'grad_feature')
    K, C = g.shape
    if corner_row.dtype != torch.int64 or tuple(corner_row.shape) != (K, 8) \
            or tuple(ci.shape) != (K, 3) or tuple(co.shape) != (K, 3):
        raise ValueError('corner_row must be int64 [K, 8] and coordinates [K, 3]')
    gt = torch.zeros((int(n_rows_table), C), dtype=torch.float32, device=dev)
    _lib.call('diver_integrate_table_backward', int(C), K, corner_row.contiguous().data_ptr(),
              ci.data_ptr(), co.data_ptr(), g.data_ptr(), gt.data_ptr(), _s(ci))
    return gt


def alpha_map(grid, decoder, o, d, pad=0, out=None):
    """Max rendering weight per voxel over rays (o, d): float32 [N^3] at x + y N + z N^2, walked
    over grid.mask.  Accumulates into `out` when given (it must start at zero)."""
    dev = _dev(grid.voxels, grid.mask, o, d)
    o, d = _f32(o, 'o'), _f32(d, 'd')
    n = grid.mask.shape[0]
    if out is None:
        out = torch.zeros(n ** 3, dtype=torch.float32, device=dev)
    elif out.dtype != torch.float32 or out.numel() != n ** 3 or not out.is_contiguous():
        raise ValueError('out must be a contiguous float32 tensor of N^3 elements')
    g = _grid_of(grid)
    _lib.call('diver_alpha_map', ctypes.byref(g), decoder.handle, o.shape[0], o.data_ptr(),
              d.data_ptr(), int(pad), out.data_ptr(), _s(o))
    return out


def render_camera(grid, decoder, c2w, width, height, focal, opts, with_count=True, out=None):
    dev = _dev(grid.voxels, grid.mask)
    rgba = out if out is not None else torch.empty((height, width, 4), dtype=torch.float32,
                                                   device=dev)
    count = torch.empty((height, width), dtype=torch.int32, device=dev) if with_count else None
    g = _grid_of(grid)
    m = (ctypes.c_float * 12)(*[float(x) for x in list(c2w.reshape(-1))])
    _lib.call('diver_render_camera', ctypes.byref(g), decoder.handle, ctypes.byref(opts), m,
              int(width), int(height), float(focal), rgba.data_ptr(), _lib.ptr(count),
              _s(grid.voxels))
    return rgba, count


def render_batch(grid, decoder, requests, opts, with_count=True):
    """Render several requests in one launch.

    requests: list of dicts, either {'c2w', 'width', 'height', 'focal'} (camera; with 'fy',
    'cx', 'cy' as well a general pinhole camera with fx = focal, see camera_rays_pinhole) or
    {'o', 'd'} (explicit rays [R,3]).  Returns a list of (rgba, count) in request order, rgba
    [H,W,4] / [R,4] as render_camera / render_rays return them.
    """
    dev = _dev(grid.voxels, grid.mask)
    if len(requests) > _lib.MAX_REQUESTS:
        raise ValueError(f'at most {_lib.MAX_REQUESTS} requests per launch')
    table = (_lib.Request * max(1, len(requests)))()
    outs, keep = [], []
    for i, rq in enumerate(requests):
        r = table[i]
        if 'c2w' in rq:
            W, H = int(rq['width']), int(rq['height'])
            rgba = torch.empty((H, W, 4), dtype=torch.float32, device=dev)
            count = torch.empty((H, W), dtype=torch.int32, device=dev) if with_count else None
            for k, x in enumerate(list(rq['c2w'].reshape(-1))[:12]):
                r.c2w[k] = float(x)
            r.width, r.height, r.focal = W, H, float(rq['focal'])
            if 'fy' in rq:
                r.fy, r.cx, r.cy = float(rq['fy']), float(rq['cx']), float(rq['cy'])
        else:
            _dev(grid.voxels, rq['o'], rq['d'])
            o, d = _f32(rq['o'], 'o'), _f32(rq['d'], 'd')
            if o.shape != d.shape or o.dim() != 2 or o.shape[1] != 3:
                raise ValueError('o and d must both be [R, 3]')
            R = o.shape[0]
            rgba = torch.empty((R, 4), dtype=torch.float32, device=dev)
            count = torch.empty((R,), dtype=torch.int32, device=dev) if with_count else None
            r.width, r.n_rays, r.o, r.d = 0, R, o.data_ptr(), d.data_ptr()
            keep += [o, d]
        r.rgba = rgba.data_ptr()
        r.count = _lib.ptr(count)
        outs.append((rgba, count))
    g = _grid_of(grid)
    _lib.call('diver_render_batch', ctypes.byref(g), decoder.handle, ctypes.byref(opts), table,
              len(requests), _s(grid.voxels))
    return outs


def camera_rays(c2w, width, height, focal, device):
    o = torch.empty((height * width, 3), dtype=torch.float32, device=device)
    d = torch.empty((height * width, 3), dtype=torch.float32, device=device)
    m = (ctypes.c_float * 12)(*[float(x) for x in list(c2w.reshape(-1))])
    _lib.call('diver_camera_rays', m, int(width), int(height), float(focal), o.data_ptr(),
              d.data_ptr(), torch.cuda.current_stream(o.device).cuda_stream)
    return o, d


def camera_rays_pinhole(c2w, width, height, fx, fy, cx, cy, device):
    """Rays of a general pinhole camera (utils/dataset/tanks.py:120-169): pixel (x, y) looks along
    ((x + 0.5 - cx)/fx, -(y + 0.5 - cy)/fy, -1) rotated by c2w[:, :3], from c2w[:, 3]."""
    o = torch.empty((height * width, 3), dtype=torch.float32, device=device)
    d = torch.empty((height * width, 3), dtype=torch.float32, device=device)
    m = (ctypes.c_float * 12)(*[float(x) for x in list(c2w.reshape(-1))])
    _lib.call('diver_camera_rays_pinhole', m, int(width), int(height), float(fx), float(fy),
              float(cx), float(cy), o.data_ptr(), d.data_ptr(),
              torch.cuda.current_stream(o.device).cuda_stream)
    return o, d
