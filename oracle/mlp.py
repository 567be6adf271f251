"""TEST INFRASTRUCTURE ONLY: numpy restatement of the constitutive surrogate
network (forward pass, mean-squared-error loss and its gradient), the
checker of the device's BF16 matrix-core kernels.

Only tests/ use this module; the product package never imports it.

Network (no counterpart in the reference): 16 inputs (12 used) -> 4 x 128
SiLU -> 8 outputs (6 used); flat parameter vector with, per layer, W [out][in]
row-major then b [out].  With rounding=True the BF16 rounding points of the
device kernels are reproduced: inputs, weights, hidden activations and the
back-propagated pre-activation gradients are rounded to BF16 (round to
nearest even), sums are taken in FP64 and rounded to FP32 where the device
accumulates in FP32.  With rounding=False every quantity stays FP64 (used to
pin the gradient against finite differences of the loss).

Parity pin: tests/test_mlp_oracle.py checks the FP64 gradient against
central finite differences of the FP64 loss and the forward pass against an
explicit per-sample loop.
"""
import numpy as np

IN, HID, OUT, USED_IN, USED_OUT, LAYERS = 16, 128, 8, 12, 6, 5


def layer_in(l):
    return IN if l == 0 else HID


def layer_out(l):
    return OUT if l == LAYERS - 1 else HID


def offsets():
    w, b, o = [], [], 0
    for l in range(LAYERS):
        w.append(o)
        o += layer_out(l) * layer_in(l)
        b.append(o)
        o += layer_out(l)
    return w, b, o


W_OFF, B_OFF, N_PARAMS = offsets()


def bf16(x):
    """Round float32 values to bfloat16 (nearest, ties to even); returns FP32."""
    a = np.ascontiguousarray(x, dtype=np.float32)
    u = a.view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    return u.astype(np.uint32).view(np.float32).reshape(a.shape)


def unpack(p):
    p = np.asarray(p, np.float64)
    Ws = [p[W_OFF[l]:W_OFF[l] + layer_out(l) * layer_in(l)].reshape(layer_out(l), layer_in(l))
          for l in range(LAYERS)]
    bs = [p[B_OFF[l]:B_OFF[l] + layer_out(l)] for l in range(LAYERS)]
    return Ws, bs


def _silu(z):
    return z / (1.0 + np.exp(-z))


def _silu_grad(z):
    s = 1.0 / (1.0 + np.exp(-z))
    return s * (1.0 + z * (1.0 - s))


def _r(x, rounding, to):
    if not rounding:
        return np.asarray(x, np.float64)
    if to == "bf16":
        return bf16(np.asarray(x, np.float32)).astype(np.float64)
    return np.asarray(x, np.float32).astype(np.float64)


def forward(p, x, rounding=True, keep=False):
    """y [n][8] for features x [n][16]; keep=True also returns the
    per-layer (input activation, pre-activation) pairs."""
    Ws, bs = unpack(p)
    a = _r(x, rounding, "bf16")
    saved = []
    for l in range(LAYERS):
        W = _r(Ws[l], rounding, "bf16")
        z = _r(_r(a @ W.T, rounding, "f32") + bs[l], rounding, "f32")
        saved.append((a, z))
        if l < LAYERS - 1:
            a = _r(_silu(z), rounding, "bf16")
    return (z, saved) if keep else z


def loss_and_grad(p, x, t, rounding=True):
    """Mean squared error over the 6 used outputs and its gradient with
    respect to the flat parameter vector."""
    n = x.shape[0]
    y, saved = forward(p, x, rounding, keep=True)
    Ws, _ = unpack(p)
    r = y - np.asarray(t, np.float64)
    r[:, USED_OUT:] = 0.0
    loss = float(np.sum(r * r) / (n * USED_OUT))
    g = np.zeros(N_PARAMS)
    dz = _r(2.0 * r / (USED_OUT * n), rounding, "f32")
    for l in range(LAYERS - 1, -1, -1):
        a_in, _ = saved[l]
        dzb = _r(dz, rounding, "bf16")
        g[W_OFF[l]:W_OFF[l] + layer_out(l) * layer_in(l)] = _r(dzb.T @ a_in, rounding, "f32").reshape(-1)
        g[B_OFF[l]:B_OFF[l] + layer_out(l)] = _r(dz.sum(axis=0), rounding, "f32")
        if l == 0:
            break
        W = _r(Ws[l], rounding, "bf16")
        da = _r(dzb @ W, rounding, "f32")
        dz = _r(da * _silu_grad(saved[l - 1][1]), rounding, "f32")
    return loss, g
