"""The numpy restatement of the surrogate network (the checker of the
device kernels), pinned without the device: its FP64 gradient against
central finite differences of its FP64 loss, its forward pass against an
explicit per-sample loop, its BF16 rounding against hand-worked values."""
import numpy as np

from oracle import mlp as M


def params(seed=3, scale=0.3):
    rng = np.random.default_rng(seed)
    return rng.normal(scale=scale, size=M.N_PARAMS)


def test_layout():
    assert M.N_PARAMS == 52744
    assert M.W_OFF[1] == 128 * 16 + 128 and M.B_OFF[4] == M.N_PARAMS - 8


def test_bf16_rounding():
    x = np.array([1.0, 1.0 + 2.0 ** -8, 1.0 + 3 * 2.0 ** -9, 1.0 + 2.0 ** -7 + 2.0 ** -8, -2.5,
                  3.0e38, 1.0 + 2.0 ** -9], np.float32)
    want = np.array([1.0, 1.0, 1.0 + 2.0 ** -7, 1.0 + 2.0 ** -6, -2.5, 2.9980945e38, 1.0], np.float32)
    got = M.bf16(x)
    assert np.array_equal(got[:5], want[:5]) and got[6] == want[6]
    assert abs(got[5] / 3.0e38 - 1) < 2 ** -8
    # bf16 values are fixed points
    assert np.array_equal(M.bf16(got), got)


def test_forward_matches_per_sample_loop():
    p = params()
    rng = np.random.default_rng(4)
    x = rng.normal(size=(5, 16))
    y = M.forward(p, x, rounding=False)
    Ws, bs = M.unpack(p)
    for s in range(5):
        a = x[s]
        for l in range(M.LAYERS):
            z = np.array([sum(Ws[l][o, i] * a[i] for i in range(Ws[l].shape[1])) for o in range(Ws[l].shape[0])])
            z = z + bs[l]
            a = z / (1 + np.exp(-z)) if l < M.LAYERS - 1 else z
        np.testing.assert_allclose(y[s], a, rtol=1e-12, atol=1e-12)


def test_gradient_matches_finite_differences():
    p = params(seed=5, scale=0.2)
    rng = np.random.default_rng(6)
    x = rng.normal(size=(7, 16))
    x[:, 12:] = 0.0
    t = rng.normal(size=(7, 8))
    _, g = M.loss_and_grad(p, x, t, rounding=False)
    # a spread of weights and biases of every layer
    idx = []
    for l in range(M.LAYERS):
        idx += list(rng.integers(M.W_OFF[l], M.B_OFF[l], 6)) + list(rng.integers(M.B_OFF[l], M.B_OFF[l] + 6, 2))
    h = 1e-6
    for i in idx:
        pp, pm = p.copy(), p.copy()
        pp[i] += h
        pm[i] -= h
        fd = (M.loss_and_grad(pp, x, t, rounding=False)[0] - M.loss_and_grad(pm, x, t, rounding=False)[0]) / (2 * h)
        assert abs(fd - g[i]) <= 1e-6 * max(1.0, abs(g[i])), (i, fd, g[i])
    # the padded outputs carry no gradient
    assert np.all(g[M.W_OFF[4] + 6 * 128:M.B_OFF[4]] == 0) and np.all(g[M.B_OFF[4] + 6:] == 0)


def test_rounded_gradient_close_to_exact():
    """The BF16 rounding points move the gradient by BF16-level amounts only
    (a misplaced rounding point or a dropped term would not)."""
    p = params(seed=7, scale=0.1)
    rng = np.random.default_rng(8)
    x = rng.normal(size=(64, 16))
    t = rng.normal(size=(64, 8))
    l0, g0 = M.loss_and_grad(p, x, t, rounding=False)
    l1, g1 = M.loss_and_grad(p, x, t, rounding=True)
    assert abs(l1 / l0 - 1) < 2e-2
    assert np.linalg.norm(g1 - g0) / np.linalg.norm(g0) < 3e-2
