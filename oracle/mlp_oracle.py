"""CPU reference of the extension-cost MLP (torch fp32 autograd on the CPU).

TEST INFRASTRUCTURE ONLY.  The MLP has no counterpart in the reference: it
only orders gapped extensions.  This module restates its definition so the
device kernels (forward, backward, Adam) can be checked:

  x[b, 8r + e] = bf16(embed[code[b, r], e])              r < 128, e < 8
  z1 = x bf16(W1)^T + b1,   h1 = bf16(gelu(z1))          1024 -> 256
  z2 = h1 bf16(W2)^T + b2,  h2 = bf16(gelu(z2))          256 -> 64
  y  = h2 . bf16(w3) + b3                                64 -> 1
  loss = mean_b (y_b - t_b)^2

gelu is the erf form.  bf16() rounds to nearest even; gradients pass through
the roundings unchanged (straight-through), as in the device backward.
"""
import numpy as np
import torch

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


def _bf16(x):
    return x + (x.to(torch.bfloat16).to(torch.float32) - x).detach()


def forward_backward(flat, codes, targets=None):
    """Returns (y, loss, flat_grad) -- loss/grad None without targets."""
    p = {k: torch.tensor(v, requires_grad=True) for k, v in unflatten(flat).items()}
    c = torch.tensor(np.asarray(codes, np.int64).reshape(-1, 128))
    x = _bf16(p["embed"])[c].reshape(len(c), 1024)
    z1 = x @ _bf16(p["w1"]).T + p["b1"]
    h1 = _bf16(torch.nn.functional.gelu(z1))
    z2 = h1 @ _bf16(p["w2"]).T + p["b2"]
    h2 = _bf16(torch.nn.functional.gelu(z2))
    y = h2 @ _bf16(p["w3"]) + p["b3"]
    if targets is None:
        return y.detach().numpy(), None, None
    t = torch.tensor(np.asarray(targets, np.float32))
    loss = ((y - t) ** 2).mean()
    loss.backward()
    grad = np.concatenate([p[k].grad.numpy().ravel() for k, _ in SIZES])
    return y.detach().numpy(), float(loss.detach()), grad


def adam_step(flat, grad, m, v, step, lr, b1=0.9, b2=0.999, eps=1e-8):
    """Adam with bias correction (Kingma & Ba, Algorithm 1), fp32."""
    m = b1 * m + (1 - b1) * grad
    v = b2 * v + (1 - b2) * grad * grad
    mh = m / (1 - b1 ** step)
    vh = v / (1 - b2 ** step)
    return (flat - lr * mh / (np.sqrt(vh) + eps)).astype(np.float32), m, v
