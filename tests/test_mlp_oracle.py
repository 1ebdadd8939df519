"""Pins for the FP64 MLP oracle (oracle/mlp.py): finite differences of its loss."""
import numpy as np

from oracle import mlp as om


def random_params(seed=0, scale=1.0):
    rng = np.random.default_rng(seed)
    parts = []
    rin = [12, 256, 256, 256, 256]
    rout = [256, 256, 256, 256, 9]
    for din, dout, ri, ro in zip(om.DIN, om.DOUT, rin, rout):
        W = np.zeros((dout, din))
        W[:ro, :ri] = rng.uniform(-1, 1, (ro, ri)) * scale * np.sqrt(6.0 / (ri + ro))
        b = np.zeros(dout)
        b[:ro] = rng.normal(0, 0.1, ro)
        parts += [W.ravel(), b]
    return np.concatenate(parts)


def test_gradients_match_finite_differences():
    rng = np.random.default_rng(1)
    p = random_params()
    X = np.zeros((16, 32))
    X[:, :12] = rng.normal(size=(16, 12))
    T = rng.normal(size=(16, 16))
    loss, g = om.loss_and_grads(p, X, T)
    idx = rng.choice(np.flatnonzero(np.abs(g) > 1e-6), 40, replace=False)
    for i in idx:
        h = 1e-6 * max(1.0, abs(p[i]))
        pp, pm = p.copy(), p.copy()
        pp[i] += h
        pm[i] -= h
        fd = (om.loss_and_grads(pp, X, T)[0] - om.loss_and_grads(pm, X, T)[0]) / (2 * h)
        assert abs(fd - g[i]) <= 1e-5 * max(abs(g[i]), 1e-3), (i, fd, g[i])


def test_gelu_closed_forms():
    """GELU's tanh form: zero at 0, identity far right, zero far left, within 2e-3 of the exact
    erf form everywhere, and its derivative matches finite differences."""
    from scipy.special import erf
    assert om.gelu(np.array([0.0]))[0] == 0.0
    assert abs(om.gelu(np.array([10.0]))[0] - 10.0) < 1e-12
    assert abs(om.gelu(np.array([-10.0]))[0]) < 1e-12
    z = np.linspace(-6, 6, 241)
    assert np.abs(om.gelu(z) - 0.5 * z * (1 + erf(z / np.sqrt(2)))).max() < 2e-3
    z = np.linspace(-4, 4, 33)
    fd = (om.gelu(z + 1e-6) - om.gelu(z - 1e-6)) / 2e-6
    assert np.allclose(fd, om.dgelu(z), atol=1e-8)


def test_bf16_rounding_model():
    x = np.array([1.0, 1.0 + 2 ** -8, 1.0 + 3 * 2 ** -9, -3.14159])
    r = om.bf16_round(x)
    # spacing 2^-7 on [1, 2) and 2^-6 on [2, 4): 1 + 2^-8 is a tie (to even: 1.0), 1 + 3 * 2^-9 is
    # above the midpoint, and 3.14159 is 201.06 steps of 2^-6
    assert list(r) == [1.0, 1.0, 1.0 + 2 ** -7, -201 * 2 ** -6]
    assert np.all(np.abs(r - x) <= np.abs(x) * 2 ** -8)
    assert om.bf16_round(np.array([1.0 + 2 ** -8]))[0] == 1.0        # tie -> even
    assert om.bf16_round(np.array([1.0 + 3 * 2 ** -8]))[0] == 1.0 + 2 ** -6
