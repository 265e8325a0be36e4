"""CPU checks of the shift-predictor oracle (oracle/mlp_oracle.py): its
gradient against central finite differences, Adam against its closed form
for a constant gradient, BF16 rounding against the definition, and the flat
parameter layout."""
import numpy as np

from oracle import mlp_oracle as mo


def test_layout_sizes_and_padding():
    lay = mo.Layout(5, 128, 3, 2)
    # W0 128x16 + b0 128, 2 x (128x128 + 128), Wo 16x128 + bo 16
    assert lay.nparams == 128 * 16 + 128 + 2 * (128 * 128 + 128) + 16 * 128 + 16 == 37264
    p = lay.random_params(np.random.default_rng(0))
    Ws, bs = lay.unpack(p)
    assert np.all(Ws[0][:, 5:] == 0) and np.all(Ws[-1][2:] == 0) and np.all(bs[-1][2:] == 0)


def test_gradient_matches_finite_differences():
    rng = np.random.default_rng(1)
    lay = mo.Layout(5, 32, 2, 2)
    p = lay.random_params(rng, 0.5)
    x = rng.uniform(-1, 1, (16, 5))
    t = rng.standard_normal((16, 2))
    loss, g = mo.loss_grad(lay, p, x, t)
    L = lay.layers()
    idx = []
    for off, o, i, ob in L:   # real weights and biases of every layer
        ro = 2 if (off, o, i, ob) == L[-1] else o
        ri = 5 if off == 0 else i
        idx += [off + (k % ro) * i + (k * 7) % ri for k in range(6)] + [ob + k % ro for k in range(2)]
    fd = mo.finite_difference_grad(lay, p, x, t, idx)
    assert np.max(np.abs(fd - g[idx])) <= 1e-7 * max(1.0, np.max(np.abs(fd)))
    # padded entries have zero gradient
    off, o, i, ob = L[0]
    assert np.all(g[off:off + o * i].reshape(o, i)[:, 5:] == 0)


def test_adam_constant_gradient_closed_form():
    """For a constant gradient the bias-corrected moments are g and g^2, so each
    step moves by lr g / (|g| + eps)."""
    g = np.array([0.3, -2.0, 1e-3, 0.0])
    p = np.array([1.0, 2.0, 3.0, 4.0])
    m = np.zeros(4)
    v = np.zeros(4)
    lr = 1e-2
    for t in range(1, 26):
        p, m, v = mo.adam_step(p, g, m, v, t, lr)
    expect = np.array([1.0, 2.0, 3.0, 4.0]) - 25 * lr * g / (np.abs(g) + 1e-8)
    assert np.max(np.abs(p - expect)) <= 1e-12


def test_bf16_round_to_nearest_even():
    one = 1.0
    u = 2.0 ** -7            # BF16 spacing at 1
    vals = np.array([one, one + u / 2, one + 3 * u / 2, one + u / 2 + 2.0 ** -20, -(one + 3 * u / 2), 3.0e-3])
    got = mo.bf16(vals)
    assert got[0] == 1.0
    assert got[1] == 1.0                      # tie -> even
    assert got[2] == one + 2 * u              # tie -> even (upwards)
    assert got[3] == one + u                  # above the tie
    assert got[4] == -(one + 2 * u)
    assert abs(got[5] - 3.0e-3) <= 3.0e-3 * 2.0 ** -8
