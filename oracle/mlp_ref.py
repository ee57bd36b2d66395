"""TEST INFRASTRUCTURE ONLY: FP64 numpy restatement of the surrogate
network's arithmetic, the checker of the HIP BF16-MFMA kernels.

The network (umwm_amd/csrc/mlp.hpp) maps K0 inputs through hidden layers
z_l = W_l a_{l-1} + b_l, a_l = gelu(z_l) (tanh form) to a linear output
y = W_L a_{L-1} + b_L; W_l is stored [N_l][K_l].  The training loss is the
weighted mean squared error over rows and outputs,
    loss = sum_{r,o} w_o (y[r,o] - t[r,o])^2 / (rows * nout),
and its gradients are the textbook back-propagation below.  The reference
model (umwm/umwm) has no learned part; what pins this restatement is
mathematics: central finite differences of `loss` and closed forms of a
one-layer linear network (tests/test_surrogate_ref.py).

Only tests/ and bench.py's checks import this module.
"""
import numpy as np

C_GELU = np.sqrt(2.0 / np.pi)


def gelu(z):
    z = np.asarray(z, dtype=np.float64)
    return 0.5 * z * (1.0 + np.tanh(C_GELU * (z + 0.044715 * z ** 3)))


def gelu_grad(z):
    z = np.asarray(z, dtype=np.float64)
    t = np.tanh(C_GELU * (z + 0.044715 * z ** 3))
    return 0.5 * (1.0 + t) + 0.5 * z * (1.0 - t * t) * C_GELU * (1.0 + 3 * 0.044715 * z * z)


def forward(ws, bs, x):
    """Returns (zs, acts, y): pre-activations and activations of the hidden
    layers (acts[0] = x) and the linear output."""
    a = np.asarray(x, dtype=np.float64)
    zs, acts = [], [a]
    for l, (w, b) in enumerate(zip(ws, bs)):
        z = a @ np.asarray(w, dtype=np.float64).T + np.asarray(b, dtype=np.float64)
        if l + 1 < len(ws):
            zs.append(z)
            a = gelu(z)
            acts.append(a)
        else:
            return zs, acts, z
    raise ValueError("no layers")


def loss(ws, bs, x, t, wout=None, nout=None):
    _, _, y = forward(ws, bs, x)
    t = np.asarray(t, dtype=np.float64)
    nout = y.shape[1] if nout is None else nout
    wo = np.ones(nout) if wout is None else np.asarray(wout, dtype=np.float64)[:nout]
    d = (y - t)[:, :nout]
    return float(np.sum(wo * d * d) / (y.shape[0] * nout))


def gradients(ws, bs, x, t, wout=None, nout=None):
    """Loss and its gradients with respect to every W_l and b_l."""
    zs, acts, y = forward(ws, bs, x)
    t = np.asarray(t, dtype=np.float64)
    rows = y.shape[0]
    nout = y.shape[1] if nout is None else nout
    wo = np.ones(nout) if wout is None else np.asarray(wout, dtype=np.float64)[:nout]
    d = np.zeros_like(y)
    d[:, :nout] = (y - t)[:, :nout]
    L = float(np.sum(wo * d[:, :nout] ** 2) / (rows * nout))
    dz = np.zeros_like(y)
    dz[:, :nout] = 2.0 * wo * d[:, :nout] / (rows * nout)
    gws, gbs = [None] * len(ws), [None] * len(ws)
    for l in range(len(ws) - 1, -1, -1):
        gws[l] = dz.T @ acts[l]
        gbs[l] = dz.sum(0)
        if l > 0:
            dz = (dz @ np.asarray(ws[l], dtype=np.float64)) * gelu_grad(zs[l - 1])
    return L, gws, gbs


def bf16_round(a):
    """Rounds float values to bfloat16 (round to nearest even), returned as
    float64: the values the device kernels multiply."""
    a32 = np.asarray(a, dtype=np.float32)
    u = a32.view(np.uint32).astype(np.uint64)
    lsb = (u >> 16) & 1
    u = ((u + 0x7FFF + lsb) >> 16) << 16
    return u.astype(np.uint32).view(np.float32).astype(np.float64)
