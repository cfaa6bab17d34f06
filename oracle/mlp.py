"""FP64 CPU restatement of the SCF starting-density network (numpy).

TEST INFRASTRUCTURE ONLY: imported by tests/ as the checker of the device network
(rerdmft_amd/csrc/rr_mlp.hip); never by the library.

The network has no counterpart in the reference; what it restates is the textbook algorithm:
  forward   A_0 = X,  Z_l = A_{l-1} W_l^T + b_l,  A_l = tanh(Z_l) (hidden), Y = Z_L (output);
  loss      1/(2 B n_out) sum (Y - T)^2;
  backward  dZ_L = (Y - T)/(B n_out), dW_l = dZ_l^T A_{l-1}, db_l = sum_b dZ_l,
            dA_{l-1} = dZ_l W_l, dZ_{l-1} = dA_{l-1} (1 - A_{l-1}^2);
  Adam      (Kingma and Ba, arXiv:1412.6980, Algorithm 1) with bias-corrected moments.
Parameter vector layout (as on the device): per layer W_l [fan_out][fan_in] row-major, then b_l.

Pins: the analytic gradient is checked against central finite differences of forward() in
tests/test_oracle_mlp.py (and a planted sign error in the tanh derivative is shown to fail that
check); adam_step() is checked against the closed form of the first step (a step of exactly -lr
sign(g) for every parameter with g != 0).
"""
import numpy as np


def layer_dims(n_in, n_hidden, n_layers, n_out):
    return [n_in] + [n_hidden] * n_layers + [n_out]


def unpack(params, dims):
    """List of (W, b) views of a flat parameter vector."""
    out, off = [], 0
    p = np.asarray(params, dtype=np.float64)
    for fi, fo in zip(dims[:-1], dims[1:]):
        w = p[off:off + fi * fo].reshape(fo, fi)
        off += fi * fo
        b = p[off:off + fo]
        off += fo
        out.append((w, b))
    assert off == p.size, "parameter vector does not match the layer sizes"
    return out


def forward(params, dims, x, keep=False):
    layers = unpack(params, dims)
    a = [np.asarray(x, dtype=np.float64)]
    for l, (w, b) in enumerate(layers):
        z = a[-1] @ w.T + b
        a.append(np.tanh(z) if l + 1 < len(layers) else z)
    return a if keep else a[-1]


def loss(params, dims, x, t):
    y = forward(params, dims, x)
    return 0.5 * np.sum((y - t) ** 2) / y.size


def loss_grad(params, dims, x, t, planted=False):
    """(loss, flat gradient). planted=True uses 1 + A^2 for the tanh derivative (a planted mistake)."""
    a = forward(params, dims, x, keep=True)
    layers = unpack(params, dims)
    y = a[-1]
    t = np.asarray(t, dtype=np.float64)
    val = 0.5 * np.sum((y - t) ** 2) / y.size
    dz = (y - t) / y.size
    grads = []
    for l in range(len(layers) - 1, -1, -1):
        w, _ = layers[l]
        grads.append((dz.T @ a[l], dz.sum(axis=0)))
        if l > 0:
            da = dz @ w
            dz = da * ((1.0 + a[l] ** 2) if planted else (1.0 - a[l] ** 2))
    flat = []
    for gw, gb in reversed(grads):
        flat.append(gw.ravel())
        flat.append(gb)
    return val, np.concatenate(flat)


def adam_step(params, grad, m, v, step, lr, b1=0.9, b2=0.999, eps=1e-8):
    """One Adam update; returns (params, m, v) for step number `step` (1-based)."""
    m = b1 * m + (1 - b1) * grad
    v = b2 * v + (1 - b2) * grad * grad
    mh = m / (1 - b1 ** step)
    vh = v / (1 - b2 ** step)
    return params - lr * mh / (np.sqrt(vh) + eps), m, v
