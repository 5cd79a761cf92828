"""Pins the CPU oracle of the initial-guess network: its gradient equals central
finite differences of its loss (float64 mode), and the bfloat16 rounding helper
matches the IEEE round-to-nearest-even definition on hand-picked values."""
import numpy as np

import mpnn_oracle as M


def small_graph(n=6):
    # ring of n nodes with unequal edge factors
    ptr, col, fac = [0], [], []
    for i in range(n):
        for j in sorted({(i - 1) % n, i, (i + 1) % n}):
            col.append(j)
            fac.append(0.0 if j == i else 1.0 + 0.1 * ((i + j) % 3))
        ptr.append(len(col))
    return M.graph_weights(np.array(ptr), np.array(col), np.array(fac))


def test_gradient_matches_finite_differences():
    rng = np.random.default_rng(0)
    g = small_graph(7)
    theta = M.init_params(3, nlayers=2)
    X = rng.normal(size=(2, 7, M.NFEAT))
    T = rng.normal(size=(2, 7))
    loss, grad = M.loss_and_grad(theta, X, T, g, nlayers=2, emulate_bf16=False)
    idx = rng.choice(theta.size, 60, replace=False)
    for k in idx:
        e = np.zeros_like(theta)
        e[k] = 1e-6
        lp, _ = M.loss_and_grad(theta + e, X, T, g, nlayers=2, emulate_bf16=False)
        lm, _ = M.loss_and_grad(theta - e, X, T, g, nlayers=2, emulate_bf16=False)
        fd = (lp - lm) / 2e-6
        assert abs(fd - grad[k]) <= 1e-6 * max(1.0, abs(fd)), (k, fd, grad[k])


def test_bf16_rounding():
    vals = np.array([1.0, 1.0 + 2 ** -7, 1.0 + 2 ** -8, 1.0 + 3 * 2 ** -8, -2.5, 3.0e38, 1e-40], dtype=np.float32)
    out = M.bf16(vals)
    # 7 explicit mantissa bits: 1 + 2^-7 is representable; 1 + 2^-8 is a tie -> even (1.0);
    # 1 + 3*2^-8 is a tie between 1 + 2^-7 (odd) and 1 + 2^-6 (even) -> 1 + 2^-6
    assert out[0] == 1.0 and out[1] == 1.0 + 2 ** -7 and out[2] == 1.0 and out[3] == 1.0 + 2 ** -6
    assert out[4] == -2.5
    assert abs(out[5] - 3.0e38) / 3.0e38 < 2 ** -8


def test_param_count():
    assert M.nparams(4, 64) == M.init_params(0).size == 34177
