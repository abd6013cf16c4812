"""Pins of the learned-guess network's CPU oracle (oracle/cnn_oracle.py):
bf16 rounding on known values, adjoint identities of every layer type, closed
forms (identity and averaging kernels), and the full backward pass against
central finite differences of the float64 forward pass."""
import numpy as np
import pytest

from oracle import cnn_oracle as N


def test_bf16_rounding_known_values():
    x = np.array([1.0, 1.0 + 2.0 ** -8, 1.0 + 3 * 2.0 ** -8, 1.0 + 2.0 ** -9, -2.5, 3.140625, 0.0])
    # ties go to even: 1 + 2^-8 -> 1, 1 + 3*2^-8 -> 1 + 2^-6; 1 + 2^-9 -> 1
    expect = np.array([1.0, 1.0, 1.0 + 2.0 ** -6, 1.0, -2.5, 3.140625, 0.0])
    assert np.array_equal(N.bf16(x), expect)


@pytest.mark.parametrize("mode", [N.CONV3, N.DOWN, N.UP])
def test_layer_adjoints(mode):
    rng = np.random.RandomState(mode)
    G = 4
    ci, co = 3, 5
    x = rng.standard_normal((G, G, G, ci))
    W = rng.standard_normal((N.taps(mode), ci, co))
    y = N.FWD[mode](x, W)
    g = rng.standard_normal(y.shape)
    lhs = np.sum(y * g)
    # data adjoint and weight adjoint of the same bilinear form
    assert np.isclose(lhs, np.sum(x * N.DATA[mode](g, W)), rtol=1e-12)
    assert np.isclose(lhs, np.sum(W * N.WGT[mode](x, g)), rtol=1e-12)


def test_closed_forms():
    rng = np.random.RandomState(0)
    x = rng.standard_normal((4, 4, 4, 2))
    W = np.zeros((27, 2, 2))
    W[13] = np.eye(2)  # centre tap
    assert np.allclose(N.conv3(x, W), x, rtol=0, atol=0)
    Wd = np.stack([np.eye(2) / 8.0] * 8)
    pooled = x.reshape(2, 2, 2, 2, 2, 2, 2).mean(axis=(1, 3, 5))
    assert np.allclose(N.down(x, Wd), pooled, atol=1e-15)
    # a shifted tap moves the field by one voxel (zero padded)
    Ws = np.zeros((27, 2, 2))
    Ws[14] = np.eye(2)  # dx = 2: out[x] = x[x + 1]
    out = N.conv3(x, Ws)
    assert np.array_equal(out[:, :, :-1], x[:, :, 1:]) and np.all(out[:, :, -1] == 0)


def test_backward_matches_finite_differences():
    rng = np.random.RandomState(7)
    G = 8
    p = rng.standard_normal(N.n_params()) * 0.3
    x = rng.standard_normal((G, G, G, 8))
    dy = rng.standard_normal((G, G, G, 16))
    g = N.backward(p, x, dy, rounding=False)
    # a few coordinates in every layer's weights and biases
    offs = []
    o = 0
    for m, ci, co in N.SPECS:
        nw = N.taps(m) * ci * co
        offs += [o + rng.randint(nw), o + nw + rng.randint(co)]
        o += nw + co
    h = 1e-4  # the summed output cancels heavily: smaller steps lose digits to roundoff
    for i in offs:
        pp, pm = p.copy(), p.copy()
        pp[i] += h
        pm[i] -= h
        fd = (np.sum(N.forward(pp, x, False)[0] * dy) - np.sum(N.forward(pm, x, False)[0] * dy)) / (2 * h)
        assert abs(fd - g[i]) <= 1e-5 * max(1.0, abs(fd)), (i, fd, g[i])
