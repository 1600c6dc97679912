"""FP64 oracle of the brick-visibility MLP -- test infrastructure only.

Same network as vt_mlp (12 -> 256 -> 256 -> 256 -> 1, ReLU, weighted
BCE-with-logits).  With emulate_bf16=True the weights, inputs, hidden
activations and backpropagated activation gradients are rounded to BF16
(round to nearest even) where the GPU stores them in BF16, so the remaining
difference is FP32-vs-FP64 accumulation only.
"""
from __future__ import annotations

import numpy as np

IN, HID = 12, 256
SHAPES = [("W1", (IN, HID)), ("b1", (HID,)), ("W2", (HID, HID)), ("b2", (HID,)), ("W3", (HID, HID)),
          ("b3", (HID,)), ("w4", (HID,)), ("b4", (1,))]
NPARAMS = sum(int(np.prod(s)) for _, s in SHAPES)


def unpack(p):
    out, off = {}, 0
    for name, shape in SHAPES:
        n = int(np.prod(shape))
        out[name] = np.asarray(p[off: off + n], np.float64).reshape(shape)
        off += n
    return out


def pack(d):
    return np.concatenate([np.asarray(d[name], np.float64).ravel() for name, _ in SHAPES])


def bf16(x):
    """Round to the nearest BF16 value (ties to even), returned as float64."""
    f = np.asarray(x, np.float32)
    u = f.view(np.uint32).astype(np.uint64)
    rnd = ((u >> 16) & 1) + 0x7FFF
    u = ((u + rnd) >> 16) << 16
    return u.astype(np.uint32).view(np.float32).astype(np.float64)


def forward(p, x, emulate_bf16=False):
    w = unpack(p)
    r = bf16 if emulate_bf16 else (lambda a: np.asarray(a, np.float64))
    x = r(x)
    W1, W2, W3 = (r(w["W1"]), r(w["W2"]), r(w["W3"]))
    h1 = r(np.maximum(x @ W1 + w["b1"], 0.0))
    h2 = r(np.maximum(h1 @ W2 + w["b2"], 0.0))
    h3 = r(np.maximum(h2 @ W3 + w["b3"], 0.0))
    z = h3 @ w["w4"] + w["b4"][0]
    return z, (x, h1, h2, h3, W1, W2, W3)


def loss_and_grad(p, x, y, pos_weight=1.0, emulate_bf16=False):
    """Mean weighted BCE-with-logits and its gradient (packed like p)."""
    w = unpack(p)
    r = bf16 if emulate_bf16 else (lambda a: np.asarray(a, np.float64))
    z, (x, h1, h2, h3, W1, W2, W3) = forward(p, x, emulate_bf16)
    y = np.asarray(y, np.float64)
    N = z.size
    wt = np.where(y > 0.5, pos_weight, 1.0)
    loss = np.mean(wt * (np.logaddexp(0.0, z) - y * z))
    dz = wt * (1.0 / (1.0 + np.exp(-z)) - y) / N
    g = {"w4": h3.T @ dz, "b4": np.array([dz.sum()])}
    dh3 = r(np.outer(dz, w["w4"]) * (h3 > 0))
    g["W3"], g["b3"] = h2.T @ dh3, dh3.sum(0)
    dh2 = r((dh3 @ W3.T) * (h2 > 0))
    g["W2"], g["b2"] = h1.T @ dh2, dh2.sum(0)
    dh1 = r((dh2 @ W2.T) * (h1 > 0))
    g["W1"], g["b1"] = x.T @ dh1, dh1.sum(0)
    return loss, pack(g)


def features(cam, brick_min, brick_side, n, level_c):
    """12 features of vt_mlp_features for bricks with min corners brick_min[N,3]:
    camera position and look (the pose), brick centre, levelC/10, and the
    brick centre's image-plane position (u, v), 2*distance*(d.right, d.up) /
    ((d.look) * (w, h)), clamped to [-4, 4], 4 behind the camera."""
    D = float(1 << n)
    c = np.asarray(brick_min, np.float64) + 0.5 * brick_side
    p = np.asarray(cam.pos[:], np.float64)
    look, up, right = (np.asarray(v[:], np.float64) for v in (cam.look, cam.up, cam.right))
    d = c - p
    dl, dr, du = d @ look, d @ right, d @ up
    front = dl > 0
    safe = np.where(front, dl, 1.0)
    u = np.where(front, np.clip(dr / safe * (2.0 * cam.distance / cam.w), -4, 4), 4.0)
    v = np.where(front, np.clip(du / safe * (2.0 * cam.distance / cam.h), -4, 4), 4.0)
    out = np.zeros((c.shape[0], 12))
    out[:, 0:3] = p / D - 0.5
    out[:, 3:6] = look
    out[:, 6:9] = c / D - 0.5
    out[:, 9] = level_c / 10.0
    out[:, 10] = u
    out[:, 11] = v
    return out
