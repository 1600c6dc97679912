"""The BF16 MFMA MLP kernels against the FP64 oracle (oracle/mlp.py)."""
import numpy as np
import pytest

from oracle import mlp

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def net(vt, ctx):
    m = vt.MLP(ctx, seed=3, max_batch=4096)
    yield m
    m.close()


def batch(rng, n):
    x = rng.normal(0, 1, (n, 12)).astype(np.float32)
    y = (rng.random(n) < 0.25).astype(np.float32)
    return x, y


@pytest.mark.parametrize("n", [1, 37, 64, 1000, 4096])
def test_forward_matches_oracle(vt, ctx, net, n):
    rng = np.random.default_rng(n)
    x, _ = batch(rng, n)
    p = net.params().astype(np.float64)
    d_x, d_z = ctx.to_device(x), ctx.alloc(n * 4)
    net.forward(d_x, n, d_z)
    z = d_z.download(np.float32, (n,)).astype(np.float64)
    z_ref, _ = mlp.forward(p, x, emulate_bf16=True)
    z_64, _ = mlp.forward(p, x, emulate_bf16=False)
    # against the BF16-faithful oracle: FP32 accumulation only
    assert np.max(np.abs(z - z_ref)) <= 2e-3 * (1 + np.max(np.abs(z_ref)))
    # against pure FP64: BF16 rounding of weights and activations
    if n > 1:
        assert np.corrcoef(z, z_64)[0, 1] > 0.999
    d_x.free()
    d_z.free()


def test_gradients_match_oracle(vt, ctx, net):
    rng = np.random.default_rng(11)
    n = 2048
    x, y = batch(rng, n)
    d_x, d_y = ctx.to_device(x), ctx.to_device(y)
    g, loss = net.gradients(d_x, d_y, n, pos_weight=2.5)
    p = net.params().astype(np.float64)
    loss_ref, g_ref = mlp.loss_and_grad(p, x, y, pos_weight=2.5, emulate_bf16=True)
    assert abs(loss - loss_ref) <= 1e-4 * (1 + abs(loss_ref))
    offs = np.cumsum([0] + [int(np.prod(s)) for _, s in mlp.SHAPES])
    for (name, _), a, b in zip(mlp.SHAPES, offs[:-1], offs[1:]):
        gg, gr = g[a:b].astype(np.float64), g_ref[a:b]
        scale = np.max(np.abs(gr)) + 1e-12
        err = np.max(np.abs(gg - gr)) / scale
        assert err < 2e-2, (name, err)
        if gr.size > 1:
            cos = gg @ gr / (np.linalg.norm(gg) * np.linalg.norm(gr) + 1e-30)
            assert cos > 0.9995, (name, cos)
    d_x.free()
    d_y.free()


def test_training_reduces_loss(vt, ctx):
    """A learnable rule (label = brick in front of the camera and near)
    must be fitted: the loss falls well below its starting value."""
    m = vt.MLP(ctx, seed=5, max_batch=4096)
    rng = np.random.default_rng(2)
    n = 4096
    x = rng.uniform(-1, 1, (n, 12)).astype(np.float32)
    y = ((x[:, 11] > 0.3) & (x[:, 10] < 0.5)).astype(np.float32)
    d_x, d_y = ctx.to_device(x), ctx.to_device(y)
    first = m.train_step(d_x, d_y, n, lr=1e-3)
    for _ in range(300):
        last = m.train_step(d_x, d_y, n, lr=1e-3)
    assert last < 0.25 * first, (first, last)
    d_z = ctx.alloc(n * 4)
    m.forward(d_x, n, d_z)
    acc = ((d_z.download(np.float32, (n,)) > 0) == (y > 0.5)).mean()
    assert acc > 0.97
    m.close()


def test_features_match_oracle(vt, vto, ctx, net):
    n_levels, level_c = 8, 3
    codes = np.array([vto.encode(x * 32, y * 32, z * 32, level_c, n_levels)
                      for x in range(8) for y in range(8) for z in range(8)], np.uint64)
    mins = np.array([vto.decode(c, level_c, n_levels) for c in codes])
    for cam in (vt.make_camera(8, 8, (128.0, 100.0, 700.0)),
                vt.make_camera(64, 48, (300.0, 20.0, -90.0), *vt.look_at((300.0, 20.0, -90.0), (120, 140, 100)))):
        d_c, d_x = ctx.to_device(codes), ctx.alloc(codes.size * 48)
        net.features(cam, n_levels, level_c, d_c, codes.size, d_x)
        got = d_x.download(np.float32, (codes.size, 12)).astype(np.float64)
        ref = mlp.features(cam, mins, 32, n_levels, level_c)
        assert np.allclose(got, ref, atol=2e-5)
        # (u, v) invert the camera's ray construction: a centre in front with
        # |u|, |v| < 4 lies on the ray through the screen point (u, v)
        pos = np.array(cam.pos[:], np.float64)
        look, up, right = (np.array(v[:], np.float64) for v in (cam.look, cam.up, cam.right))
        c = mins + 16.0
        u, v = got[:, 10], got[:, 11]
        inside = (np.abs(u) < 3.99) & (np.abs(v) < 3.99)
        A = (look[None] * cam.distance + right[None] * (u[:, None] * cam.w / 2) + up[None] * (v[:, None] * cam.h / 2))
        dvec = c - pos
        cosang = np.sum(A * dvec, 1) / np.linalg.norm(A, axis=1) / np.linalg.norm(dvec, axis=1)
        assert inside.sum() > 20 and np.all(cosang[inside] > 1 - 1e-8)
        assert np.all(got[(dvec @ look) <= 0][:, 10:12] == 4.0)
        d_c.free()
        d_x.free()


def test_fused_score_equals_layer_path(vt, ctx, net):
    """vt_mlp_score (features and forward pass in one launch) returns the
    logits of vt_mlp_features followed by vt_mlp_forward bit for bit: partial
    workgroups, N above max_batch, and after the parameters change (by a
    training step, then by vt_mlp_set_params)."""
    rng = np.random.default_rng(5)
    n_levels = 8
    x, y = batch(rng, 4096)
    d_x, d_y = ctx.to_device(x), ctx.to_device(y)
    p0 = net.params()
    cams = [vt.make_camera(64, 48, (300.0, 20.0, -90.0), *vt.look_at((300.0, 20.0, -90.0), (120, 140, 100))),
            vt.make_camera(8, 8, (128.0, 100.0, 700.0))]

    def check(level_c, count):
        base = np.uint64(1) << np.uint64(3 * level_c)
        every = base + np.arange(8 ** level_c, dtype=np.uint64)
        assert count <= every.size
        codes = np.sort(rng.choice(every, count, replace=False)) if count < every.size else every
        d_c, d_z = ctx.to_device(codes), ctx.alloc(count * 4)
        d_xb, d_fb = ctx.alloc(net.max_batch * 48), ctx.alloc(net.max_batch * 4)
        for cam in cams:
            net.score(cam, n_levels, level_c, d_c, count, d_z)
            got = d_z.download(np.float32, (count,))
            ref = np.empty(count, np.float32)
            for o in range(0, count, net.max_batch):
                k = min(net.max_batch, count - o)
                d_cb = ctx.to_device(codes[o:o + k])
                net.features(cam, n_levels, level_c, d_cb, k, d_xb)
                net.forward(d_xb, k, d_fb)
                ref[o:o + k] = d_fb.download(np.float32, (k,))
                d_cb.free()
            assert np.array_equal(got.view(np.uint32), ref.view(np.uint32)), (level_c, count)
        for b in (d_c, d_z, d_xb, d_fb):
            b.free()

    try:
        net.train_step(d_x, d_y, 4096, lr=1e-3)   # weights away from their initial values
        for level_c, count in ((3, 1), (3, 77), (3, 512), (5, 32768)):
            check(level_c, count)
        net.set_params(p0 * np.float32(0.5) + np.float32(0.01))
        check(5, 4096 + 65)
    finally:
        net.set_params(p0)
        d_x.free()
        d_y.free()
