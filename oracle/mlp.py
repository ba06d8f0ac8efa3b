"""FP64 restatement of the chemistry-surrogate MLP and its training step.

TEST INFRASTRUCTURE ONLY: imported by tests/ as the checker of the library's
BF16 MFMA forward, backward and Adam kernels; never by the product path.

Network (reference pattern: evalModel, reactions.jl:59-72, dense layers with
GELU; training loop lux-ck.jl:38-71, MSE loss and Adam): 11 -> 64 -> 64 -> 64
-> 9, GELU (tanh form) on the hidden layers, linear output.  Parameters are one
packed FP32 vector of 12512 floats, inputs padded to 32 and outputs to 32
(padding weights zero):

    W1 [64][32] | b1 [64] | W2 [64][64] | b2 [64] | W3 [64][64] | b3 [64] | W4 [32][64] | b4 [32]

Loss L = sum_{s, o < nout} (y_so - t_so)^2 / (B * nout), nout = 9 for H2/O2
(11 inputs) and 5 for 5-species air (7 inputs, the rest of the 11 input
slots zero).

`emulate_bf16=True` rounds (round to nearest even) exactly the values the GPU
kernels hold in BF16: the inputs, the weights, every hidden activation and
every back-propagated dL/dZ used by the matrix products (the bias gradients
use the unrounded dL/dZ, GELU' is taken in full precision).  With
`emulate_bf16=False` the network is plain FP64, which the finite-difference
checks pin.

Pinned by (tests/test_oracle_mlp.py): central finite differences of the FP64
loss for every layer's weights and biases; the tanh-form GELU against the
erf form (|difference| < 1e-3) and its derivative against finite
differences; the first Adam step in closed form (-lr g / (|g| + eps)); a
single-sample, single-hidden-unit reduction checked by hand.
"""
from __future__ import annotations

import numpy as np

NN_IN, NN_INP, NN_H, NN_OUT, NN_OUTP = 11, 32, 64, 9, 32
LAYERS = [(NN_INP, NN_H), (NN_H, NN_H), (NN_H, NN_H), (NN_H, NN_OUTP)]  # (in, out)


def _offsets():
    w, b, off = [], [], 0
    for ni, no in LAYERS:
        w.append(off)
        off += ni * no
        b.append(off)
        off += no
    return w, b, off


W_OFF, B_OFF, NPARAM = _offsets()
assert NPARAM == 12512


def unpack(params):
    p = np.asarray(params, dtype=np.float64)
    out = []
    for l, (ni, no) in enumerate(LAYERS):
        W = p[W_OFF[l]:W_OFF[l] + ni * no].reshape(no, ni)
        b = p[B_OFF[l]:B_OFF[l] + no]
        out.append((W, b))
    return out


def pack(layers):
    p = np.zeros(NPARAM)
    for l, (W, b) in enumerate(layers):
        ni, no = LAYERS[l]
        p[W_OFF[l]:W_OFF[l] + ni * no] = np.asarray(W).reshape(-1)
        p[B_OFF[l]:B_OFF[l] + no] = b
    return p


def bf16(x):
    """Round to the nearest BF16 value (ties to even), returned as float64."""
    f = np.asarray(x, dtype=np.float32)
    u = f.view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    return u.astype(np.uint32).view(np.float32).astype(np.float64)


_C = np.sqrt(2.0 / np.pi)
_K = 0.044715


def gelu(x):
    return 0.5 * x * (1.0 + np.tanh(_C * (x + _K * x ** 3)))


def gelu_grad(x):
    t = np.tanh(_C * (x + _K * x ** 3))
    return 0.5 * (1.0 + t) + 0.5 * x * (1.0 - t * t) * _C * (1.0 + 3.0 * _K * x * x)


def forward(params, X, emulate_bf16=False):
    """y [B, 9] and the cache (activations, pre-activations) for backward;
    X [B, nin] with nin <= 11 (missing inputs are zero)."""
    r = bf16 if emulate_bf16 else (lambda v: np.asarray(v, dtype=np.float64))
    B = X.shape[0]
    A = np.zeros((B, NN_INP))
    A[:, :X.shape[1]] = X
    A = r(A)
    acts, pre = [A], []
    layers = unpack(params)
    for l, (W, b) in enumerate(layers):
        Z = A @ r(W).T + b
        pre.append(Z)
        if l < 3:
            A = r(gelu(Z))
            acts.append(A)
    return Z[:, :NN_OUT], (acts, pre)


def loss_grad(params, X, T, emulate_bf16=False):
    """(loss, grad[NPARAM]) of L = sum (y - t)^2 / (B * nout), T [B, nout]."""
    r = bf16 if emulate_bf16 else (lambda v: np.asarray(v, dtype=np.float64))
    B, nout = T.shape
    y, (acts, pre) = forward(params, X, emulate_bf16)
    d = y[:, :nout] - T
    loss = float(np.sum(d * d) / (B * nout))
    g = np.zeros((B, NN_OUTP))
    g[:, :nout] = 2.0 * d / (B * nout)
    layers = unpack(params)
    grad = np.zeros(NPARAM)
    for l in range(3, -1, -1):
        ni, no = LAYERS[l]
        dZ = r(g)
        grad[B_OFF[l]:B_OFF[l] + no] = g.sum(axis=0)
        grad[W_OFF[l]:W_OFF[l] + ni * no] = (dZ.T @ acts[l]).reshape(-1)
        if l == 0:
            break
        dA = dZ @ r(layers[l][0])
        g = dA * gelu_grad(pre[l - 1])
    return loss, grad


def adam_step(params, m1, m2, grad, step, lr, b1=0.9, b2=0.999, eps=1e-8):
    """Adam (Kingma & Ba 2015), bias-corrected; step counts from 1."""
    m1 = b1 * m1 + (1 - b1) * grad
    m2 = b2 * m2 + (1 - b2) * grad * grad
    mh = m1 / (1 - b1 ** step)
    vh = m2 / (1 - b2 ** step)
    return params - lr * mh / (np.sqrt(vh) + eps), m1, m2
