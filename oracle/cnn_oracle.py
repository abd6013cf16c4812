"""CPU oracle of the learned-guess network (test infrastructure only).

NumPy float64 restatement of the 3D U-Net in
pigfem_assignment_amd/csrc/kernels/pgf_cnn.hip, forward and backward, with
the device's bf16 rounding points reproduced when ``rounding=True``:
activations are stored in bf16, every MFMA operand (input, weights, gradient
of a layer's pre-activation) is bf16, accumulation is exact here (FP32 on the
device), biases and bias gradients are not rounded.  With ``rounding=False``
the same code is a plain float64 network used for finite-difference checks.

The network has no counterpart in the reference (it predicts an initial guess
for the solves of src/SolverNonlinearKSP.cpp:55-159); nothing here is
imported by the product path.

Arrays are channels-last with index [z, y, x, c] (voxel index (z*G + y)*G + x).
Layer weights are W[t][cin][cout]; tap t of a 3x3x3 kernel is
(dx, dy, dz) = (t % 3, t // 3 % 3, t // 9), of a 2x2x2 kernel
(t & 1, t >> 1 & 1, t >> 2).
"""
import numpy as np

CONV3, DOWN, UP = 0, 1, 2
# (mode, cin, cout) of layers 0..7; skips: layer 5 adds act2, layer 6 adds act0
SPECS = [(CONV3, 8, 16), (DOWN, 16, 32), (CONV3, 32, 32), (DOWN, 32, 64),
         (CONV3, 64, 64), (UP, 64, 32), (UP, 32, 16), (CONV3, 16, 16)]
OUT_CH = 3


def taps(mode):
    return 27 if mode == CONV3 else 8


def n_params():
    return sum(taps(m) * ci * co + co for m, ci, co in SPECS)


def unflatten(p):
    """Flat parameter vector -> [(W[t][cin][cout], b[cout]) per layer]."""
    out, o = [], 0
    for m, ci, co in SPECS:
        n = taps(m) * ci * co
        W = np.asarray(p[o:o + n], np.float64).reshape(taps(m), ci, co)
        o += n
        b = np.asarray(p[o:o + co], np.float64)
        o += co
        out.append((W, b))
    return out


def flatten(layers):
    return np.concatenate([np.concatenate([W.ravel(), b.ravel()]) for W, b in layers])


def bf16(x):
    """Round to bf16 (via float32, round to nearest even), returned as float64."""
    u = np.asarray(x, np.float32).view(np.uint32).astype(np.uint64)
    u = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16
    return u.astype(np.uint32).view(np.float32).astype(np.float64)


def _off3(t):
    return t % 3, (t // 3) % 3, t // 9


def _off2(t):
    return t & 1, (t >> 1) & 1, t >> 2


def conv3(x, W):
    """out[z,y,x] = sum_t xpad[z+dz, y+dy, x+dx] @ W[t] (zero padding 1)."""
    G = x.shape[0]
    xp = np.zeros((G + 2, G + 2, G + 2, x.shape[3]))
    xp[1:-1, 1:-1, 1:-1] = x
    out = np.zeros((G, G, G, W.shape[2]))
    for t in range(27):
        dx, dy, dz = _off3(t)
        out += xp[dz:dz + G, dy:dy + G, dx:dx + G] @ W[t]
    return out


def conv3_data(g, W):
    """Adjoint of conv3 in its input: dx[i] = sum_t g[i - off(t) + 1] @ W[t]^T."""
    G = g.shape[0]
    gp = np.zeros((G + 2, G + 2, G + 2, g.shape[3]))
    gp[1:-1, 1:-1, 1:-1] = g
    out = np.zeros((G, G, G, W.shape[1]))
    for t in range(27):
        dx, dy, dz = _off3(t)
        out += gp[2 - dz:2 - dz + G, 2 - dy:2 - dy + G, 2 - dx:2 - dx + G] @ W[t].T
    return out


def conv3_weight(x, g):
    G = x.shape[0]
    xp = np.zeros((G + 2, G + 2, G + 2, x.shape[3]))
    xp[1:-1, 1:-1, 1:-1] = x
    gf = g.reshape(-1, g.shape[3])
    dW = np.zeros((27, x.shape[3], g.shape[3]))
    for t in range(27):
        dx, dy, dz = _off3(t)
        dW[t] = xp[dz:dz + G, dy:dy + G, dx:dx + G].reshape(-1, x.shape[3]).T @ gf
    return dW


def down(x, W):
    """out[o] = sum_t x[2o + off(t)] @ W[t]."""
    out = 0.0
    for t in range(8):
        dx, dy, dz = _off2(t)
        out = out + x[dz::2, dy::2, dx::2] @ W[t]
    return out


def down_data(g, W):
    G = g.shape[0]
    out = np.zeros((2 * G, 2 * G, 2 * G, W.shape[1]))
    for t in range(8):
        dx, dy, dz = _off2(t)
        out[dz::2, dy::2, dx::2] = g @ W[t].T
    return out


def down_weight(x, g):
    gf = g.reshape(-1, g.shape[3])
    return np.stack([x[dz::2, dy::2, dx::2].reshape(-1, x.shape[3]).T @ gf
                     for dx, dy, dz in (_off2(t) for t in range(8))])


def up(x, W):
    """out[2i + off(t)] = x[i] @ W[t]."""
    G = x.shape[0]
    out = np.zeros((2 * G, 2 * G, 2 * G, W.shape[2]))
    for t in range(8):
        dx, dy, dz = _off2(t)
        out[dz::2, dy::2, dx::2] = x @ W[t]
    return out


def up_data(g, W):
    out = 0.0
    for t in range(8):
        dx, dy, dz = _off2(t)
        out = out + g[dz::2, dy::2, dx::2] @ W[t].T
    return out


def up_weight(x, g):
    xf = x.reshape(-1, x.shape[3])
    return np.stack([xf.T @ g[dz::2, dy::2, dx::2].reshape(-1, g.shape[3])
                     for dx, dy, dz in (_off2(t) for t in range(8))])


FWD = {CONV3: conv3, DOWN: down, UP: up}
DATA = {CONV3: conv3_data, DOWN: down_data, UP: up_data}
WGT = {CONV3: conv3_weight, DOWN: down_weight, UP: up_weight}


def forward(params, x, rounding=True):
    """Returns (y [G,G,G,16], activations act[0..6], rounded input)."""
    R = bf16 if rounding else (lambda v: np.asarray(v, np.float64))
    L = unflatten(params)
    x = R(x)
    act = []
    h = x
    for l, (m, ci, co) in enumerate(SPECS):
        W, b = L[l]
        pre = FWD[m](h, R(W)) + b
        if l == 5:
            pre = pre + act[2]
        if l == 6:
            pre = pre + act[0]
        if l == 7:
            return pre, act, x
        h = R(np.maximum(pre, 0.0))
        act.append(h)


def backward(params, x, dy, rounding=True):
    """Gradient of sum(dy . y) with respect to the flat parameter vector;
    dy has the shape of y (all 16 channels)."""
    R = bf16 if rounding else (lambda v: np.asarray(v, np.float64))
    L = unflatten(params)
    y, act, xr = forward(params, x, rounding)
    inputs = [xr] + act  # input of layer l is inputs[l]
    grads = [None] * 8
    g = np.asarray(dy, np.float64)  # gradient of layer 7's pre-activation
    skip_grad = {}
    for l in range(7, -1, -1):
        m, ci, co = SPECS[l]
        W, _ = L[l]
        gb = R(g)
        grads[l] = (WGT[m](inputs[l], gb), g.reshape(-1, co).sum(axis=0))
        if l == 0:
            break
        if l == 6:
            skip_grad[0] = g
        if l == 5:
            skip_grad[2] = g
        d = DATA[m](gb, R(W))  # gradient of act[l-1]
        if (l - 1) in skip_grad:
            d = d + skip_grad[l - 1]
        g = d * (act[l - 1] > 0)
    return flatten(grads)
