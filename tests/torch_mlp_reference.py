"""torch autograd reference of the cost-model MLP, run as a child process by
test_mlp_oracle.py so that torch (and the HIP runtime it bundles) never loads
into a test process that uses the GPU.

  python torch_mlp_reference.py inputs.npz outputs.npz
"""
import sys

import numpy as np
import torch

SIZES = [("embed", (21, 8)), ("w1", (256, 1024)), ("b1", (256,)), ("w2", (64, 256)),
         ("b2", (64,)), ("w3", (64,)), ("b3", (1,))]


def r(x):
    """bf16 rounding (through float32) with a straight-through gradient."""
    return x + (x.float().to(torch.bfloat16).double() - x).detach()


def main(src, dst):
    d = np.load(src)
    flat, codes, t = d["flat"], d["codes"], d["targets"]
    p, o = {}, 0
    for k, shape in SIZES:
        n = int(np.prod(shape))
        p[k] = torch.tensor(flat[o:o + n].reshape(shape), dtype=torch.float64,
                            requires_grad=True)
        o += n
    c = torch.tensor(np.minimum(codes.astype(np.int64).reshape(-1, 128), 20))
    x = r(p["embed"])[c].reshape(len(c), 1024)
    z1 = x @ r(p["w1"]).T + p["b1"]
    h1 = r(torch.nn.functional.gelu(z1))
    z2 = h1 @ r(p["w2"]).T + p["b2"]
    h2 = r(torch.nn.functional.gelu(z2))
    y = h2 @ r(p["w3"]) + p["b3"]
    loss = ((y - torch.tensor(t, dtype=torch.float64)) ** 2).mean()
    loss.backward()
    grad = np.concatenate([p[k].grad.numpy().ravel() for k, _ in SIZES])
    probe = d["probe"]
    rounded = torch.tensor(probe).to(torch.bfloat16).float().numpy()
    np.savez(dst, y=y.detach().numpy(), loss=float(loss.detach()), grad=grad, rounded=rounded)


if __name__ == "__main__":
    main(sys.argv[1], sys.argv[2])
