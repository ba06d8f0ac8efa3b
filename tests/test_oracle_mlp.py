"""Pins the FP64 MLP oracle (oracle/mlp.py) that checks the GPU surrogate
kernels: finite differences of the loss, closed forms of GELU and of the
first Adam step, a hand-reduced tiny network, BF16 rounding of known values,
and the surrogate's parameter layout and persistence."""
import math

import numpy as np
import pytest

from oracle import mlp


def _rand_params(seed=0, scale=0.3):
    rng = np.random.default_rng(seed)
    layers = []
    for ni, no in mlp.LAYERS:
        W = np.zeros((no, ni))
        fi = 11 if ni == mlp.NN_INP else ni
        fo = 9 if no == mlp.NN_OUTP else no
        W[:fo, :fi] = rng.normal(0, scale, (fo, fi))
        b = np.zeros(no)
        b[:fo] = rng.normal(0, 0.1, fo)
        layers.append((W, b))
    return mlp.pack(layers)


def test_layout():
    assert mlp.NPARAM == 12512
    assert mlp.W_OFF == [0, 2112, 6272, 10432]
    assert mlp.B_OFF == [2048, 6208, 10368, 12480]
    p = _rand_params()
    for (W, b), (ni, no) in zip(mlp.unpack(p), mlp.LAYERS):
        assert W.shape == (no, ni) and b.shape == (no,)
    assert np.array_equal(mlp.pack(mlp.unpack(p)), p)


def test_gelu_closed_forms():
    x = np.linspace(-6, 6, 241)
    erf_gelu = 0.5 * x * (1 + np.vectorize(math.erf)(x / math.sqrt(2)))
    assert np.max(np.abs(mlp.gelu(x) - erf_gelu)) < 1e-3
    assert mlp.gelu(0.0) == 0.0
    h = 1e-6
    fd = (mlp.gelu(x + h) - mlp.gelu(x - h)) / (2 * h)
    assert np.max(np.abs(fd - mlp.gelu_grad(x))) < 1e-8
    assert abs(mlp.gelu_grad(0.0) - 0.5) < 1e-15
    # saturation: gelu(x) -> x for large x, -> 0 for very negative x
    assert abs(mlp.gelu(10.0) - 10.0) < 1e-12 and abs(mlp.gelu(-10.0)) < 1e-12


@pytest.mark.parametrize("layer", [0, 1, 2, 3])
@pytest.mark.parametrize("nin,nout", [(11, 9), (7, 5)])  # H2/O2, 5-species air
def test_gradient_finite_differences(layer, nin, nout):
    rng = np.random.default_rng(1 + layer)
    p = _rand_params(layer)
    X = rng.normal(0, 1, (16, nin))
    T = rng.normal(0, 1, (16, nout))
    _, g = mlp.loss_grad(p, X, T)
    ni, no = mlp.LAYERS[layer]
    fi = nin if ni == mlp.NN_INP else ni
    fo = nout if no == mlp.NN_OUTP else no
    idx = [mlp.W_OFF[layer] + o * ni + k for o, k in zip(rng.integers(0, fo, 12),
                                                          rng.integers(0, fi, 12))]
    idx += [mlp.B_OFF[layer] + o for o in rng.integers(0, fo, 6)]
    h = 1e-6
    for i in idx:
        pp, pm = p.copy(), p.copy()
        pp[i] += h
        pm[i] -= h
        fd = (mlp.loss_grad(pp, X, T)[0] - mlp.loss_grad(pm, X, T)[0]) / (2 * h)
        assert abs(fd - g[i]) <= 1e-7 * max(1.0, abs(g[i])), (i, fd, g[i])
    # padded weights carry no gradient
    if ni == mlp.NN_INP:
        pad = [mlp.W_OFF[0] + o * ni + k for o in range(no) for k in range(nin, ni)]
        assert np.all(g[pad] == 0)
    if no == mlp.NN_OUTP:
        pad = list(range(mlp.W_OFF[3] + nout * ni, mlp.W_OFF[3] + no * ni))
        pad += list(range(mlp.B_OFF[3] + nout, mlp.B_OFF[3] + no))
        assert np.all(g[pad] == 0)


def test_tiny_network_by_hand():
    """One sample, one active unit per layer: the network reduces to the
    scalar chain y = w4 g(w3 g(w2 g(w1 x + b1) + b2) + b3) + b4."""
    layers = []
    vals = [(0.7, 0.1), (-1.3, 0.2), (0.9, -0.3), (1.1, 0.05)]
    for (ni, no), (w, b) in zip(mlp.LAYERS, vals):
        W = np.zeros((no, ni))
        W[0, 0] = w
        bb = np.zeros(no)
        bb[0] = b
        layers.append((W, bb))
    p = mlp.pack(layers)
    x = np.zeros((1, 11))
    x[0, 0] = 0.8
    t = np.zeros((1, 9))
    t[0, 0] = 0.25
    z1 = 0.7 * 0.8 + 0.1
    z2 = -1.3 * mlp.gelu(z1) + 0.2
    z3 = 0.9 * mlp.gelu(z2) - 0.3
    y = 1.1 * mlp.gelu(z3) + 0.05
    yo, _ = mlp.forward(p, x)
    assert abs(yo[0, 0] - y) < 1e-15
    # other outputs are the biases (zero)
    assert np.all(yo[0, 1:] == 0)
    loss, g = mlp.loss_grad(p, x, t)
    assert abs(loss - ((y - 0.25) ** 2 + 0.0) / 9) < 1e-15
    dy = 2 * (y - 0.25) / 9
    dw4 = dy * mlp.gelu(z3)
    assert abs(g[mlp.W_OFF[3]] - dw4) < 1e-15
    dz3 = dy * 1.1 * mlp.gelu_grad(z3)
    assert abs(g[mlp.B_OFF[2]] - dz3) < 1e-15
    dz1 = dz3 * 0.9 * mlp.gelu_grad(z2) * -1.3 * mlp.gelu_grad(z1)
    assert abs(g[mlp.W_OFF[0]] - dz1 * 0.8) < 1e-15


def test_adam_first_step_closed_form():
    rng = np.random.default_rng(3)
    p = rng.normal(size=50)
    g = rng.normal(size=50)
    g[0] = 0.0
    lr = 1e-3
    p1, m1, m2 = mlp.adam_step(p, np.zeros(50), np.zeros(50), g, 1, lr)
    assert np.allclose(p1, p - lr * g / (np.abs(g) + 1e-8), rtol=0, atol=1e-15)
    assert p1[0] == p[0]
    # constant gradient: every step moves by lr * g/(|g| + eps) exactly
    p2, _, _ = mlp.adam_step(p1, m1, m2, g, 2, lr)
    assert np.allclose(p2 - p1, -lr * g / (np.abs(g) + 1e-8), atol=1e-15)


def test_bf16_rounding():
    ulp = 2.0 ** -7  # spacing of BF16 in [1, 2)
    assert mlp.bf16(1.0) == 1.0
    assert mlp.bf16(1.0 + ulp / 2) == 1.0            # tie -> even
    assert mlp.bf16(1.0 + 1.5 * ulp) == 1.0 + 2 * ulp  # tie -> even (up)
    assert mlp.bf16(1.0 + 0.51 * ulp) == 1.0 + ulp
    assert mlp.bf16(-3.0) == -3.0
    x = np.random.default_rng(0).normal(size=1000)
    r = mlp.bf16(x)
    assert np.max(np.abs(r - x) / np.abs(x)) <= 2.0 ** -8


def test_emulated_bf16_close_to_fp64():
    rng = np.random.default_rng(5)
    p = _rand_params(5)
    X = rng.normal(0, 1, (64, 11))
    T = rng.normal(0, 1, (64, 9))
    l64, g64 = mlp.loss_grad(p, X, T)
    l16, g16 = mlp.loss_grad(p, X, T, emulate_bf16=True)
    assert abs(l16 - l64) / l64 < 2e-2
    assert np.linalg.norm(g16 - g64) / np.linalg.norm(g64) < 3e-2


def test_surrogate_init_and_persistence(tmp_path, pkg):
    sur = pkg.surrogate
    p = sur.init_params(7)
    assert p.dtype == np.float32 and p.size == mlp.NPARAM
    layers = mlp.unpack(p)
    W1 = layers[0][0]
    assert np.all(W1[:, 11:] == 0) and np.any(W1[:, :11] != 0)
    W4, b4 = layers[3]
    assert np.all(W4[9:] == 0) and np.all(b4 == 0)
    cfg = pkg.abi.NNConfig()
    for k in range(11):
        cfg.in_mean[k] = k
        cfg.in_std[k] = 1 + k
    cfg.d_max = 4.5
    cfg.res_tol = 1e-6
    cfg.boxcox_lambda = 0.1
    cfg.n_elem = 3
    cfg.elem[1][2] = 2.0
    cfg.dt = 1e-7
    path = str(tmp_path / "s.npz")
    sur.save(path, cfg, p, {"note": "x"})
    c2, p2, meta = sur.load(path)
    assert np.array_equal(p2, p) and meta == {"note": "x"}
    assert list(c2.in_std) == list(cfg.in_std) and c2.d_max == pytest.approx(4.5)
    assert c2.elem[1][2] == 2.0 and c2.n_elem == 3 and c2.dt == 1e-7
