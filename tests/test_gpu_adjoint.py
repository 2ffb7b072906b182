"""Hand-written adjoint of the fused render (integral + decoder + compositing) against
(a) FP64 central finite differences (h = 1e-3) on 256 sampled parameters of the FP64 oracle, and
(b) torch autograd in FP64 through the reference's offline path restated on the CPU
    (utils/integrator.py:62-113, model/diver.py:141-207, train.py:109 MSE), no early termination.

Tolerances: (a) |g - g_fd| <= 2e-3 * max|g_fd| + 1e-6 per parameter; (b) max-abs error of the grid
and decoder gradients <= 1e-4 * their max magnitude.
"""
import numpy as np
import pytest
import torch

import oracle
from oracle import cpu_path, fp64
from oracle.render import intervals
from diver_amd import synthetic as S
from diver_amd.adjoint import BenchmarkDecoder, FusedRenderer
from diver_amd.renderer import SceneGrid

pytestmark = pytest.mark.gpu


def _setup(dev, n_rays=512, hidden=64, seed=21):
    sc = S.config1_scene()
    gt = S.make_scene(64, 32, 6, 0, seed=S.SEED_GT)
    cams = S.hemisphere_cameras()
    o, d = S.random_rays(cams, n_rays, 64, 64, seed)
    layers = S.baseline_decoder_layers(32, hidden)
    return sc, gt, o, d, layers


def _fp64_loss(sc_vox, mask, layers, o, d, target, t_stop, rows=None):
    """FP64 render (offline walk) + MSE."""
    if rows is None:
        rows = intervals(o, d, mask, 64, -1.0, 2.0 / 64)
    ray, pin, pout, num = rows
    R = o.shape[0]
    F = fp64.integrate_einsum(sc_vox, pin, pout)
    s, c = fp64.baseline_decoder(layers, F, d.astype(np.float64)[ray])
    alpha = -np.expm1(-fp64.softplus(s))
    acc, T, used = fp64.render_front_to_back(ray, alpha, c, R, t_stop)
    rgb = acc + T[:, None]
    return float(np.mean((rgb - target) ** 2)), used


def test_adjoint_vs_finite_differences(dev):
    sc, gt, o, d, layers = _setup(dev)
    t_stop = 1e-4
    grid = SceneGrid.from_synthetic(sc, dev)
    dec = BenchmarkDecoder(layers, 32).to(dev)
    model = FusedRenderer(grid, dec, traversal='offline')
    og, dg = torch.from_numpy(o).to(dev), torch.from_numpy(d).to(dev)
    # target: the ground-truth grid rendered by the FP64 oracle
    target, _ = _fp64_target(gt, layers, o, d, t_stop)
    rgb = model(og, dg, t_stop)
    loss = torch.mean((rgb - torch.from_numpy(target).float().to(dev)) ** 2)
    loss.backward()
    gvox = model.voxels.grad.cpu().numpy().astype(np.float64)
    gdec = [p.grad.cpu().numpy().astype(np.float64) for p in
            (dec.w1, dec.b1, dec.w2, dec.b2, dec.w3, dec.b3)]
    vox = sc.dense_voxels().astype(np.float64)
    rows = intervals(o, d, sc.mask, 64, -1.0, 2.0 / 64)
    rng = np.random.default_rng(5)
    # 128 grid entries on corners the rays touch, 128 decoder entries
    touched = np.flatnonzero(np.abs(gvox).reshape(-1) > 0)
    gidx = rng.choice(touched, 128, replace=False)
    h = 1e-3
    checked = 0
    worst = 0.0
    fd_all, an_all = [], []
    for flat in gidx:
        vp = vox.copy(); vp.reshape(-1)[flat] += h
        vm = vox.copy(); vm.reshape(-1)[flat] -= h
        lp, up = _fp64_loss(vp, sc.mask, layers, o, d, target, t_stop, rows)
        lm, um = _fp64_loss(vm, sc.mask, layers, o, d, target, t_stop, rows)
        if not np.array_equal(up, um):
            continue  # a ray's termination point moved: the loss is not smooth there
        fd_all.append((lp - lm) / (2 * h))
        an_all.append(gvox.reshape(-1)[flat])
    for k in range(128):
        li = rng.integers(0, 6)
        arr = [w for pair in layers for w in pair][li]
        flat = rng.integers(0, arr.size)
        def with_delta(delta):
            ls = [[np.array(w, np.float64) for w in pair] for pair in layers]
            flat_layers = [w for pair in ls for w in pair]
            flat_layers[li].reshape(-1)[flat] += delta
            return [(flat_layers[0], flat_layers[1]), (flat_layers[2], flat_layers[3]),
                    (flat_layers[4], flat_layers[5])]
        lp, up = _fp64_loss(vox, sc.mask, with_delta(h), o, d, target, t_stop, rows)
        lm, um = _fp64_loss(vox, sc.mask, with_delta(-h), o, d, target, t_stop, rows)
        if not np.array_equal(up, um):
            continue
        fd_all.append((lp - lm) / (2 * h))
        an_all.append(gdec[li].reshape(-1)[flat])
    fd_all, an_all = np.array(fd_all), np.array(an_all)
    assert fd_all.size >= 200
    scale = np.abs(fd_all).max()
    err = np.abs(an_all - fd_all)
    assert np.all(err <= 2e-3 * scale + 1e-6), (err.max(), scale)


def _fp64_target(gt, layers, o, d, t_stop):
    vox = gt.dense_voxels().astype(np.float64)
    ray, pin, pout, num = intervals(o, d, gt.mask, 64, -1.0, 2.0 / 64)
    F = fp64.integrate_einsum(vox, pin, pout)
    # a different decoder bias makes the targets differ in colour as well as in shape
    s, c = fp64.baseline_decoder(layers, F, d.astype(np.float64)[ray])
    alpha = -np.expm1(-fp64.softplus(s))
    acc, T, used = fp64.render_front_to_back(ray, alpha, c, o.shape[0], t_stop)
    return acc + T[:, None], used


def test_adjoint_vs_reference_autograd(dev):
    """Reference offline path (no early termination) with torch autograd in float64."""
    sc, gt, o, d, layers = _setup(dev, n_rays=1024, seed=22)
    target, _ = _fp64_target(gt, layers, o, d, 0.0)
    # ours, t_stop = 0 (the offline path composites every interval)
    grid = SceneGrid.from_synthetic(sc, dev)
    dec = BenchmarkDecoder(layers, 32).to(dev)
    model = FusedRenderer(grid, dec, traversal='offline')
    rgb = model(torch.from_numpy(o).to(dev), torch.from_numpy(d).to(dev), 0.0)
    loss = torch.mean((rgb - torch.from_numpy(target).float().to(dev)) ** 2)
    loss.backward()
    # reference: torch float64 on the CPU through integrate / MLP / cumprod render
    vox = torch.tensor(sc.dense_voxels(), dtype=torch.float64, requires_grad=True)
    lt = [(torch.tensor(w, dtype=torch.float64, requires_grad=True),
           torch.tensor(b, dtype=torch.float64, requires_grad=True)) for w, b in layers]
    isect, num, tns = oracle.masked_intersect(o, d, np.float32(-1), np.float32(1), 64,
                                              np.float32(2.0 / 64), sc.mask, 1.0)
    k = int(num.max())
    coord = torch.from_numpy(isect[:, :k]).double()
    hit = torch.from_numpy(tns[:, :k]) >= 0
    rows_t = coord[hit]
    F = cpu_path.integrate(vox, rows_t[:, :3], rows_t[:, 3:])
    dd = torch.from_numpy(d).double()[torch.where(hit)[0]]
    x = torch.cat([F[:, 1:], cpu_path.positional_encoding(dd)], -1)
    for j, (w, b) in enumerate(lt):
        x = torch.nn.functional.linear(x, w, b)
        if j < 2:
            x = torch.relu(x)
    col_ = torch.sigmoid(x)
    B = o.shape[0]
    sigma = torch.zeros(B, k, dtype=torch.float64)
    sigma[hit] = F[:, 0]
    color = torch.zeros(B, k, 3, dtype=torch.float64)
    color[hit] = col_
    alpha = 1 - torch.exp(-torch.nn.functional.softplus(sigma) * hit)
    shifted = torch.nn.functional.pad(1 - alpha[None, :, :-1], (1, 0), value=1)[0]
    weight = alpha * torch.cumprod(shifted, -1)
    rgb_ref = (weight[:, :, None] * color).sum(1) + (1 - weight.sum(1, keepdim=True))
    loss_ref = torch.mean((rgb_ref - torch.from_numpy(target)) ** 2)
    loss_ref.backward()
    assert abs(loss.item() - loss_ref.item()) <= 1e-6 * max(1.0, loss_ref.item())
    g_ours = model.voxels.grad.cpu().double()
    g_ref = vox.grad
    assert (g_ours - g_ref).abs().max() <= 1e-4 * g_ref.abs().max()
    for p, (w, b) in zip([(dec.w1, dec.b1), (dec.w2, dec.b2), (dec.w3, dec.b3)], lt):
        for a_, r_ in zip(p, (w, b)):
            ga, gr = a_.grad.cpu().double(), r_.grad
            assert (ga - gr).abs().max() <= 1e-4 * gr.abs().max() + 1e-9


def test_training_steps_reduce_loss(dev):
    """A few Adam steps on the fused renderer lower the photometric loss."""
    sc, gt, o, d, layers = _setup(dev, n_rays=2048, seed=23)
    target, _ = _fp64_target(gt, layers, o, d, 1e-4)
    grid = SceneGrid.from_synthetic(sc, dev)
    dec = BenchmarkDecoder(layers, 32).to(dev)
    model = FusedRenderer(grid, dec)
    opt = torch.optim.Adam(model.parameters(), lr=1e-2)
    og, dg = torch.from_numpy(o).to(dev), torch.from_numpy(d).to(dev)
    tg = torch.from_numpy(target).float().to(dev)
    losses = []
    for _ in range(20):
        opt.zero_grad()
        loss = torch.mean((model(og, dg) - tg) ** 2)
        loss.backward()
        opt.step()
        losses.append(loss.item())
    assert losses[-1] < 0.7 * losses[0]


def test_ray_suffix_matches_segmented_sums(dev):
    """diver_ray_suffix against an FP64 loop over ragged rays (empty rays included), with and
    without the background term: the kernel sums in FP64 and rounds once to FP32."""
    from diver_amd import calls
    g = torch.Generator().manual_seed(0)
    R = 300
    cnt = torch.randint(0, 40, (R,), generator=g).to(torch.int32)
    cnt[::7] = 0
    off = torch.cumsum(cnt.long(), 0) - cnt.long()
    K = int(cnt.sum())
    w = torch.rand(K, generator=g)
    c = torch.rand(K, 3, generator=g)
    bg = torch.rand(R, generator=g)
    ref = torch.empty(K, 3, dtype=torch.float64)
    for r in range(R):
        o, n = int(off[r]), int(cnt[r])
        acc = torch.zeros(3, dtype=torch.float64)
        for k in range(n - 1, -1, -1):
            ref[o + k] = acc
            acc = acc + w[o + k].double() * c[o + k].double()
    got = calls.ray_suffix(off.to(dev), cnt.to(dev), w.to(dev), c.to(dev)).cpu().double()
    assert float((got - ref).abs().max()) <= 1e-5
    ray = torch.repeat_interleave(torch.arange(R), cnt.long())
    got_bg = calls.ray_suffix(off.to(dev), cnt.to(dev), w.to(dev), c.to(dev), bg.to(dev)).cpu().double()
    assert float((got_bg - (ref + bg.double()[ray][:, None])).abs().max()) <= 1e-5


class _ByteMaskGrid:
    """The same grid without its block/brick bits: diver_render_rows then takes the byte-mask
    path, which walks, integrates every channel and composites per lane in one kernel."""

    def __init__(self, grid):
        self.voxels, self.mask, self.octrees = grid.voxels, grid.mask, grid.octrees
        self.n, self.channels = grid.n, grid.channels
        self.xyzmin, self.voxel_size = grid.xyzmin, grid.voxel_size
        self.blocks = self.bricks = self.tail = None


@pytest.mark.parametrize('traversal,t_stop', [('offline', 0.0), ('offline', 1e-4),
                                              ('realtime', 1e-4)])
def test_render_rows_match_byte_mask_path(dev, traversal, t_stop):
    """The rows of the block/brick pipeline (walk writing segments, features gathered over all
    rows, then a per-ray pass for s, T_before and alpha) equal the byte-mask kernel's, which
    computes all of them on the walk: the same segments, integrals and transmittance, bit for
    bit; and each ray's last row leaves the transmittance the forward returned."""
    from diver_amd import calls
    sc, _gt, o, d, layers = _setup(dev, n_rays=2048, seed=5)
    grid = SceneGrid.from_synthetic(sc, dev)
    model = FusedRenderer(grid, BenchmarkDecoder(layers, 32).to(dev), traversal=traversal)
    og, dg = torch.from_numpy(o).to(dev), torch.from_numpy(d).to(dev)
    opts = calls.render_opts('fp32', traversal, t_stop, 0.0)
    rgba, count = calls.render_rays(grid, model.handle, og, dg, opts)
    cnt64 = count.to(torch.int64)
    off = torch.cumsum(cnt64, 0) - cnt64
    K = int(cnt64.sum())
    assert K > 1000
    rows = calls.render_rows(grid, model.handle, og, dg, opts, count, off, K)
    ref = calls.render_rows(_ByteMaskGrid(grid), model.handle, og, dg, opts, count, off, K)
    assert torch.equal(rows, ref)
    # T after each ray's last row, from its header (T_before, and sigma = softplus(s)) in the
    # forward's FP32 order, equals the forward's T wherever the ray has rows
    last = (off + cnt64 - 1)[cnt64 > 0]
    s, tb = rows[last, 6].double(), rows[last, 7].double()
    t_after = tb * torch.exp(-torch.nn.functional.softplus(s))
    t_fwd = rgba[cnt64 > 0, 3].double()
    assert float((t_after - t_fwd).abs().max()) <= 1e-5


@pytest.mark.parametrize('t_stop,hidden,seed', [(0.0, 64, 9), (0.0, 32, 9), (0.0, 64, 10)])
def test_recorded_segments_match_rewalk(dev, t_stop, hidden, seed):
    """The training forward that records each interval's entry and exit and evaluates densities in
    the shading batches (diver_render_rays_segments) returns the same rgba, counts and sparsity sums
    as the plain forward, which evaluates each on the walk, bit for bit; and the rows built from its
    segments (diver_render_rows_segments) equal the re-walk's rows (diver_render_rows) bit for bit.
    With early termination (t_stop > 0) the call is refused."""
    from diver_amd import calls
    sc, _gt, o, d, layers = _setup(dev, n_rays=4099, hidden=hidden, seed=seed)
    grid = SceneGrid.from_synthetic(sc, dev)
    model = FusedRenderer(grid, BenchmarkDecoder(layers, 32).to(dev), traversal='offline')
    og, dg = torch.from_numpy(o).to(dev), torch.from_numpy(d).to(dev)
    opts = calls.render_opts('fp32', 'offline', t_stop, 0.0)
    cap = 3 * grid.n + 8
    print(f'segments t_stop={t_stop} hidden={hidden}', flush=True)
    rgba0, count0, sreg0 = calls.render_rays_sparsity(grid, model.handle, og, dg, opts)
    rgba, count, sreg, seg = calls.render_rays_segments(grid, model.handle, og, dg, opts, cap)
    assert torch.equal(rgba, rgba0) and torch.equal(count, count0) and torch.equal(sreg, sreg0)
    _, count_ns, sreg_ns, seg_ns = calls.render_rays_segments(grid, model.handle, og, dg, opts, cap,
                                                              sparsity=False)
    assert sreg_ns is None and torch.equal(count_ns, count0)
    cnt64 = count.to(torch.int64)
    assert int(cnt64.max()) <= cap
    off = torch.cumsum(cnt64, 0) - cnt64
    K = int(cnt64.sum())
    assert K > 1000
    ref = calls.render_rows(grid, model.handle, og, dg, opts, count, off, K)
    rows, ray = calls.render_rows_segments(grid, model.handle, opts, count, off, K, seg, with_ray=True)
    assert torch.equal(rows, ref)
    assert torch.equal(ray, torch.repeat_interleave(torch.arange(count.shape[0], device=dev), cnt64))
    assert torch.equal(calls.render_rows_segments(grid, model.handle, opts, count, off, K, seg_ns), ref)
    from diver_amd._lib import DiverError
    with pytest.raises(DiverError):
        calls.render_rays_segments(grid, model.handle, og, dg,
                                   calls.render_opts('fp32', 'offline', 1e-4, 0.0), cap)
