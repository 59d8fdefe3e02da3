"""CPU reference of the extension-cost MLP (numpy; float64 arithmetic with the
device's bf16 rounding points).

TEST INFRASTRUCTURE ONLY.  The MLP has no counterpart in the reference: it
only orders gapped extensions.  This module restates its definition so the
device kernels (forward, backward, Adam) can be checked:

  x[b, 8r + e] = bf16(embed[code[b, r], e])              r < 128, e < 8
  z1 = x bf16(W1)^T + b1,   h1 = bf16(gelu(z1))          1024 -> 256
  z2 = h1 bf16(W2)^T + b2,  h2 = bf16(gelu(z2))          256 -> 64
  y  = h2 . bf16(w3) + b3                                64 -> 1
  loss = mean_b (y_b - t_b)^2

gelu is the erf form.  bf16() rounds to nearest even; gradients pass through
the roundings unchanged (straight-through), as in the device backward.  The
hand-written backward here is itself checked against torch autograd by a CPU
test (torch is never imported by GPU tests: it would load its own HIP runtime
into the process).
"""
import numpy as np
from scipy.special import erf

SIZES = [("embed", (21, 8)), ("w1", (256, 1024)), ("b1", (256,)), ("w2", (64, 256)),
         ("b2", (64,)), ("w3", (64,)), ("b3", (1,))]


def unflatten(flat):
    out, o = {}, 0
    for name, shape in SIZES:
        n = int(np.prod(shape))
        out[name] = np.asarray(flat[o:o + n], np.float32).reshape(shape)
        o += n
    assert o == len(flat)
    return out


def bf16(x):
    """Round float32 values to bfloat16 (nearest even), returned as float32."""
    u = np.ascontiguousarray(x, np.float32).view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    return u.astype(np.uint32).view(np.float32)


def gelu(z):
    return 0.5 * z * (1.0 + erf(z / np.sqrt(2.0)))


def gelu_grad(z):
    return 0.5 * (1.0 + erf(z / np.sqrt(2.0))) + z * np.exp(-0.5 * z * z) / np.sqrt(2 * np.pi)


def forward(flat, codes, keep=False):
    p = unflatten(flat)
    c = np.minimum(np.asarray(codes, np.int64).reshape(-1, 128), 20)
    emb = bf16(p["embed"]).astype(np.float64)
    x = emb[c].reshape(len(c), 1024)
    z1 = x @ bf16(p["w1"]).astype(np.float64).T + p["b1"]
    h1 = bf16(gelu(z1).astype(np.float32)).astype(np.float64)
    z2 = h1 @ bf16(p["w2"]).astype(np.float64).T + p["b2"]
    h2 = bf16(gelu(z2).astype(np.float32)).astype(np.float64)
    y = h2 @ bf16(p["w3"]).astype(np.float64) + p["b3"][0]
    if keep:
        return y, dict(p=p, c=c, x=x, z1=z1, h1=h1, z2=z2, h2=h2)
    return y


def forward_backward(flat, codes, targets=None):
    """Returns (y, loss, flat_grad) -- loss/grad None without targets."""
    y, a = forward(flat, codes, keep=True)
    if targets is None:
        return y.astype(np.float32), None, None
    t = np.asarray(targets, np.float64)
    n = len(t)
    loss = float(((y - t) ** 2).mean())
    dy = 2.0 * (y - t) / n
    p = a["p"]
    w3 = bf16(p["w3"]).astype(np.float64)
    w2 = bf16(p["w2"]).astype(np.float64)
    w1 = bf16(p["w1"]).astype(np.float64)
    g = {}
    g["w3"] = a["h2"].T @ dy
    g["b3"] = np.array([dy.sum()])
    dz2 = np.outer(dy, w3) * gelu_grad(a["z2"])
    g["w2"] = dz2.T @ a["h1"]
    g["b2"] = dz2.sum(axis=0)
    dz1 = (dz2 @ w2) * gelu_grad(a["z1"])
    g["w1"] = dz1.T @ a["x"]
    g["b1"] = dz1.sum(axis=0)
    dx = (dz1 @ w1).reshape(n, 128, 8)
    ge = np.zeros((21, 8))
    np.add.at(ge, a["c"].ravel(), dx.reshape(-1, 8))
    g["embed"] = ge
    grad = np.concatenate([np.asarray(g[k], np.float64).ravel() for k, _ in SIZES])
    return y.astype(np.float32), loss, grad.astype(np.float32)


def adam_step(flat, grad, m, v, step, lr, b1=0.9, b2=0.999, eps=1e-8):
    """Adam with bias correction (Kingma & Ba, Algorithm 1), fp32."""
    m = b1 * m + (1 - b1) * grad
    v = b2 * v + (1 - b2) * grad * grad
    mh = m / (1 - b1 ** step)
    vh = v / (1 - b2 ** step)
    return (flat - lr * mh / (np.sqrt(vh) + eps)).astype(np.float32), m, v
