"""The CPU restatement of the learned stress update's network (oracle/
mlp_oracle.py), pinned: its analytic gradient in double precision equals
central finite differences of its loss (a wrong transpose, activation
derivative or bias sum fails this), its bfloat16 rounding is round-to-
nearest-even, and its single-precision form stays close to the double one.
The CPU build of the GPU kernels is checked against it in tests/test_emu.py."""
import os
import sys

import numpy as np
from conftest import ROOT

sys.path.insert(0, os.path.join(ROOT, "oracle"))
import mlp_oracle as mo  # noqa: E402


def random_theta(seed=0, scale=0.2):
    rng = np.random.default_rng(seed)
    theta = scale * rng.standard_normal(mo.NPARAM)
    P = mo.unpack(theta)
    P["W1"][mo.NIN:, :] = 0.0      # padded inputs / outputs as the GPU keeps them
    P["W4"][:, mo.NOUT:] = 0.0
    P["b4"][mo.NOUT:] = 0.0
    return mo.pack(P)


def test_gradient_equals_finite_differences():
    rng = np.random.default_rng(1)
    theta = random_theta()
    x = rng.standard_normal((7, mo.NIN))
    t = rng.standard_normal((7, mo.NOUT))
    _, g = mo.loss_and_grad(theta, x, t, bf16=False)
    # parameters of every block: W1, b1, W2, b3, W4, b4 (real outputs)
    offs = np.cumsum([0] + [int(np.prod(s)) for _, s in mo.SHAPES])
    idx = []
    for b in range(8):
        idx += list(rng.integers(offs[b], offs[b + 1], size=6))
    fd = mo.finite_difference(theta, x, t, idx)
    assert np.allclose(g[idx], fd, rtol=1e-5, atol=1e-9), np.abs(g[idx] - fd).max()


def test_bf16_rounding_is_nearest_even():
    a = np.array([1.0, 1.0 + 2 ** -8, 1.0 + 3 * 2 ** -8, 1.0 + 2 ** -9, -2.5, 3.0e38], dtype=np.float32)
    r = mo.to_bf16(a)
    # 1 + 2^-8 is halfway between 1 and 1 + 2^-7: ties to even -> 1
    assert r[1] == 1.0 and r[2] == np.float32(1.0 + 2 ** -6) and r[3] == 1.0
    assert r[4] == -2.5
    assert np.all(np.abs(r - a) <= np.abs(a) * 2 ** -8)


def test_bf16_form_close_to_double_form():
    rng = np.random.default_rng(2)
    theta = random_theta(3)
    x = rng.standard_normal((64, mo.NIN)).astype(np.float32)
    t = rng.standard_normal((64, mo.NOUT)).astype(np.float32)
    y16, _ = mo.forward(theta, x)
    y64, _ = mo.forward(theta, x, bf16=False)
    assert np.abs(y16 - y64).max() <= 2e-2 * np.abs(y64).max()
    _, g16 = mo.loss_and_grad(theta, x, t)
    _, g64 = mo.loss_and_grad(theta, x, t, bf16=False)
    assert np.linalg.norm(g16 - g64) <= 5e-2 * np.linalg.norm(g64)
