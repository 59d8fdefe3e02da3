"""The numpy MLP reference (oracle/mlp_oracle.py) against torch autograd on the
CPU, and its bf16 rounding against torch's bfloat16 conversion.

torch runs in a child process (torch_mlp_reference.py): a test process that
uses the GPU must not load torch's bundled HIP runtime.
"""
import os
import subprocess
import sys

import numpy as np

import mlp_oracle

HERE = os.path.dirname(os.path.abspath(__file__))


def test_numpy_reference_matches_torch_autograd(tmp_path):
    rng = np.random.default_rng(2)
    flat = rng.normal(0, 0.05, sum(int(np.prod(s)) for _, s in mlp_oracle.SIZES)).astype(
        np.float32)
    codes = rng.integers(0, 22, (97, 128)).astype(np.uint8)
    t = rng.normal(5, 2, 97)
    probe = np.concatenate([rng.normal(0, 1, 10000), rng.normal(0, 1e-30, 100),
                            [0.0, -0.0, 1.0, 1.00390625, 1.01171875, 3.0e38]]).astype(np.float32)
    src, dst = tmp_path / "in.npz", tmp_path / "out.npz"
    np.savez(src, flat=flat, codes=codes, targets=t, probe=probe)
    subprocess.run([sys.executable, os.path.join(HERE, "torch_mlp_reference.py"), str(src),
                    str(dst)], check=True, timeout=600)
    ref = np.load(dst)
    assert np.array_equal(mlp_oracle.bf16(probe), ref["rounded"])
    y, loss, grad = mlp_oracle.forward_backward(flat, codes, t)
    assert np.allclose(y, ref["y"], rtol=1e-5, atol=1e-6)
    assert abs(loss - float(ref["loss"])) <= 1e-6 * abs(float(ref["loss"]))
    o = 0
    for name, shape in mlp_oracle.SIZES:
        k = int(np.prod(shape))
        g, r = grad[o:o + k], ref["grad"][o:o + k]
        assert np.linalg.norm(g - r) <= 1e-5 * (np.linalg.norm(r) + 1e-12), name
        o += k
