"""The FP64 MLP oracle: its backward pass pinned by central finite differences."""
import numpy as np

from oracle import mlp


def test_oracle_gradient_matches_finite_differences():
    rng = np.random.default_rng(0)
    p = rng.normal(0, 0.2, mlp.NPARAMS)
    x = rng.normal(0, 1, (32, mlp.IN))
    y = (rng.random(32) < 0.3).astype(np.float64)
    loss, g = mlp.loss_and_grad(p, x, y, pos_weight=3.0)
    idx = rng.choice(mlp.NPARAMS, 300, replace=False)
    # every parameter block is probed
    offs = np.cumsum([0] + [int(np.prod(s)) for _, s in mlp.SHAPES])
    for a, b in zip(offs[:-1], offs[1:]):
        idx = np.concatenate([idx, rng.integers(a, b, 3)])
    eps = 1e-6
    for i in idx:
        pp, pm = p.copy(), p.copy()
        pp[i] += eps
        pm[i] -= eps
        fd = (mlp.loss_and_grad(pp, x, y, 3.0)[0] - mlp.loss_and_grad(pm, x, y, 3.0)[0]) / (2 * eps)
        assert abs(fd - g[i]) <= 1e-6 + 1e-5 * abs(fd), (i, fd, g[i])


def test_bf16_rounding():
    vals = np.array([1.0, 1.00390625, 1.0019531, 3.14159, -2.5e-3, 65504.0], np.float64)
    r = mlp.bf16(vals)
    # representable: 8 significant bits
    assert r[0] == 1.0 and r[1] == 1.0078125 or r[1] == 1.0   # ties to even
    assert abs(r[3] - 3.140625) < 1e-12
    assert np.all(np.abs(r - vals) <= np.abs(vals) * 2 ** -8)
