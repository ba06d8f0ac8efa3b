"""GPU tests of the learned chemistry surrogate (BF16 MFMA MLP).

  * forward, gradient and Adam kernels against the FP64 MLP oracle
    (oracle/mlp.py, pinned by finite differences), with the oracle emulating
    the kernels' BF16 roundings (bars written in each test);
  * training by the library's kernels reduces the loss of a teacher network;
  * the fused surrogate chemistry step: post_predict, the element-conserving
    projection and the FP64 energy bookkeeping against a host restatement
    (1e-12); cells that fail a gate end bit-identical to the Newton step;
    the backward-Euler residual gate agrees with the oracle's residual;
  * the trained H2 surrogate on fresh reactors: species and temperature error
    against the Newton step, the fallback fraction, and the Newton work it
    removes (what the network gains over the integrator alone);
  * the same kernels on a 5-species air context (7 inputs, 5 outputs), and
    the trained air surrogate on fresh dissociating-air reactors.
"""
import os

import numpy as np
import pytest
import torch

from oracle import mlp

pytestmark = pytest.mark.gpu
SIZE = (16, 16, 8)


def _abi():
    import flame3d_amd.abi as abi
    return abi


def _mech():
    import flame3d_amd.mech as mm
    return mm.load("h2_lidryer")


def _ctx(size=(8, 8, 8)):
    abi = _abi()
    ctx = abi.Context(*size, _mech(), "fp64")
    cfg = abi.NNConfig()
    for k in range(11):
        cfg.in_std[k] = 1.0
    for k in range(9):
        cfg.out_std[k] = 1.0
    cfg.d_max = 1e30
    cfg.elem_tol = 1e30
    cfg.neg_tol = 1.0
    cfg.boxcox_lambda = 0.1
    cfg.dt = 1e-7
    ctx.nn_init(cfg)
    return ctx, cfg


def _rand_params(seed, scale=0.3):
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
    return mlp.pack(layers).astype(np.float32)


def _rel(a, b):
    return float(np.linalg.norm(np.asarray(a) - np.asarray(b)) / max(np.linalg.norm(b), 1e-300))


def test_forward_matches_oracle():
    print("\n[surrogate] forward B=256", flush=True)
    ctx, _ = _ctx()
    p = _rand_params(1)
    ctx.nn_set_params(p)
    assert np.array_equal(ctx.nn_get_params(), p)
    rng = np.random.default_rng(2)
    X = rng.normal(0, 1, (256, 11)).astype(np.float32)
    y = ctx.nn_forward(torch.from_numpy(X).cuda()).cpu().numpy()
    ye, _ = mlp.forward(p.astype(np.float64), X.astype(np.float64), emulate_bf16=True)
    y64, _ = mlp.forward(p.astype(np.float64), X.astype(np.float64))
    # same BF16 roundings, FP32 vs FP64 accumulation: rare one-ulp flips
    assert _rel(y, ye) < 5e-3, _rel(y, ye)
    assert np.max(np.abs(y - ye)) < 3e-2 * np.sqrt(np.mean(ye ** 2))
    # against the exact FP64 network: the BF16 approximation itself
    assert _rel(y, y64) < 3e-2, _rel(y, y64)
    ctx.close()


def test_gradient_matches_oracle():
    print("\n[surrogate] gradient B=512", flush=True)
    ctx, _ = _ctx()
    p = _rand_params(3)
    ctx.nn_set_params(p)
    rng = np.random.default_rng(4)
    X = rng.normal(0, 1, (512, 11)).astype(np.float32)
    T = rng.normal(0, 1, (512, 9)).astype(np.float32)
    g = ctx.nn_grad(torch.from_numpy(X).cuda(), torch.from_numpy(T).cuda()).cpu().numpy()
    le, ge = mlp.loss_grad(p.astype(np.float64), X.astype(np.float64), T.astype(np.float64),
                           emulate_bf16=True)
    l64, g64 = mlp.loss_grad(p.astype(np.float64), X.astype(np.float64), T.astype(np.float64))
    assert abs(g[-1] - le) / le < 1e-3, (g[-1], le)
    for l, (ni, no) in enumerate(mlp.LAYERS):
        w = slice(mlp.W_OFF[l], mlp.W_OFF[l] + ni * no)
        b = slice(mlp.B_OFF[l], mlp.B_OFF[l] + no)
        assert _rel(g[w], ge[w]) < 2e-2, (l, "W", _rel(g[w], ge[w]))
        assert _rel(g[b], ge[b]) < 2e-2, (l, "b", _rel(g[b], ge[b]))
        assert _rel(g[w], g64[w]) < 6e-2, (l, "W64", _rel(g[w], g64[w]))
    # padded weights get exactly zero gradient
    pad = [mlp.W_OFF[0] + o * 32 + k for o in range(64) for k in range(11, 32)]
    pad += list(range(mlp.W_OFF[3] + 9 * 64, mlp.B_OFF[3])) + list(range(mlp.B_OFF[3] + 9, 12512))
    assert np.all(g[pad] == 0)
    ctx.close()


def _air_ctx(size=(8, 8, 8)):
    import flame3d_amd.mech as mm
    abi = _abi()
    ctx = abi.Context(*size, mm.load("air5"), "fp64")
    cfg = abi.NNConfig()
    for k in range(11):
        cfg.in_std[k] = 1.0
    for k in range(9):
        cfg.out_std[k] = 1.0
    cfg.d_max = 1e30
    cfg.elem_tol = 1e30
    cfg.neg_tol = 1.0
    cfg.boxcox_lambda = 0.1
    cfg.dt = 6e-9
    ctx.nn_init(cfg)
    return ctx, cfg


def test_air_forward_and_gradient_match_oracle():
    """Ns = 5: the kernels read 7 inputs per sample and write / train on 5
    outputs; the oracle sees the same 7 inputs (other input slots zero) and
    the loss over the first 5 outputs.  Bars as for H2/O2."""
    print("\n[surrogate] air forward B=256, gradient B=512", flush=True)
    ctx, _ = _air_ctx()
    p = _rand_params(11)
    ctx.nn_set_params(p)
    rng = np.random.default_rng(12)
    X = rng.normal(0, 1, (256, 7)).astype(np.float32)
    y = ctx.nn_forward(torch.from_numpy(X).cuda()).cpu().numpy()
    assert y.shape == (256, 5)
    ye = mlp.forward(p.astype(np.float64), X.astype(np.float64), emulate_bf16=True)[0][:, :5]
    assert _rel(y, ye) < 5e-3, _rel(y, ye)
    X = rng.normal(0, 1, (512, 7)).astype(np.float32)
    T = rng.normal(0, 1, (512, 5)).astype(np.float32)
    g = ctx.nn_grad(torch.from_numpy(X).cuda(), torch.from_numpy(T).cuda()).cpu().numpy()
    le, ge = mlp.loss_grad(p.astype(np.float64), X.astype(np.float64), T.astype(np.float64),
                           emulate_bf16=True)
    assert abs(g[-1] - le) / le < 1e-3, (g[-1], le)
    for l, (ni, no) in enumerate(mlp.LAYERS):
        w = slice(mlp.W_OFF[l], mlp.W_OFF[l] + ni * no)
        b = slice(mlp.B_OFF[l], mlp.B_OFF[l] + no)
        assert _rel(g[w], ge[w]) < 2e-2, (l, "W", _rel(g[w], ge[w]))
        assert _rel(g[b], ge[b]) < 2e-2, (l, "b", _rel(g[b], ge[b]))
    # unused input columns (7..31) and output rows (5..31) get exactly zero gradient
    pad = [mlp.W_OFF[0] + o * 32 + k for o in range(64) for k in range(7, 32)]
    pad += list(range(mlp.W_OFF[3] + 5 * 64, mlp.B_OFF[3])) + list(range(mlp.B_OFF[3] + 5, 12512))
    assert np.all(g[pad] == 0)
    ctx.close()


def test_adam_steps():
    print("\n[surrogate] adam", flush=True)
    ctx, _ = _ctx()
    p0 = _rand_params(5)
    ctx.nn_set_params(p0)
    rng = np.random.default_rng(6)
    X = torch.from_numpy(rng.normal(0, 1, (256, 11)).astype(np.float32)).cuda()
    T = torch.from_numpy(rng.normal(0, 1, (256, 9)).astype(np.float32)).cuda()
    lr = 1e-3
    g1 = ctx.nn_grad(X, T).cpu().numpy()[:-1].astype(np.float64)
    loss = torch.zeros(1, dtype=torch.float32, device="cuda")
    ctx.nn_train_step(X, T, lr, loss)
    p1 = ctx.nn_get_params().astype(np.float64)
    # first Adam step in closed form: -lr g / (|g| + eps)
    exp1 = p0 - lr * g1 / (np.abs(g1) + 1e-8)
    assert np.max(np.abs(p1 - exp1)) < 1e-6 * lr + 1e-6 * np.max(np.abs(p0))
    # second step against the oracle's Adam with the GPU's own gradient
    g2 = ctx.nn_grad(X, T).cpu().numpy()[:-1].astype(np.float64)
    _, m1, m2 = mlp.adam_step(p0.astype(np.float64), 0 * g1, 0 * g1, g1, 1, lr)
    exp2, _, _ = mlp.adam_step(p1, m1, m2, g2, 2, lr)
    ctx.nn_train_step(X, T, lr, loss)
    p2 = ctx.nn_get_params().astype(np.float64)
    assert np.max(np.abs(p2 - exp2)) < 1e-5 * lr + 1e-6 * np.max(np.abs(p0))
    pad = [mlp.W_OFF[0] + o * 32 + k for o in range(64) for k in range(11, 32)]
    assert np.all(p2[pad] == 0)
    ctx.close()


def test_training_reduces_loss():
    print("\n[surrogate] training on a teacher network", flush=True)
    ctx, _ = _ctx()
    rng = np.random.default_rng(7)
    teacher = _rand_params(8, scale=0.25).astype(np.float64)
    X = rng.normal(0, 1, (8192, 11))
    Y, _ = mlp.forward(teacher, X)
    Y = (Y - Y.mean(0)) / Y.std(0)
    import flame3d_amd.surrogate as sur
    ctx.nn_set_params(sur.init_params(9))
    Xt = torch.from_numpy(X.astype(np.float32)).cuda()
    Yt = torch.from_numpy(Y.astype(np.float32)).cuda()
    l0 = sur.mse(ctx, Xt, Yt)
    hist = sur.train(ctx, Xt, Yt, epochs=40, batch=1024, lr=3e-3, decay_every=20, log=None)
    l1 = sur.mse(ctx, Xt, Yt)
    print(f"  mse {l0:.3e} -> {l1:.3e}", flush=True)
    # 320 Adam steps fit a random 3-hidden-layer teacher to ~10% of its variance
    assert l1 < 0.25 * l0
    assert hist[-1] < 0.25 * hist[0]
    ctx.close()


# ---------------------------------------------------------------------------
# fused surrogate chemistry step
# ---------------------------------------------------------------------------
def _cells(seed=12345, burn_steps=0):
    import flame3d_amd.cases as cases
    m = _mech()
    Q, r = cases.chem_cells(SIZE, m, seed=seed, dev="cuda")
    U = cases.prim2c(Q.cpu()).cuda().contiguous()
    if burn_steps:
        abi = _abi()
        c = abi.Context(*SIZE, m, "fp64")
        for _ in range(burn_steps):
            c.chem_newton(U, Q, r, 1e-6, 50, 1e-10, 1e-6)
        c.close()
    return m, Q, r, U


def _projection(m, cfg):
    comp = np.asarray(m.composition, np.float64)
    E = comp / np.asarray(m.mw)[None, :]
    S = np.diag(np.array(list(cfg.out_std), np.float64) ** 2)
    G = E @ S @ E.T
    return np.eye(9) - S @ E.T @ np.linalg.inv(G) @ E


def _interior(t):
    import flame3d_amd.surrogate as sur
    return sur.interior(t).cpu().numpy()


def test_nn_chem_all_fallback_equals_newton():
    print(f"\n[surrogate] nn_chem all-fallback size={SIZE}", flush=True)
    abi = _abi()
    m, Q, r, U = _cells()
    ctx = abi.Context(*SIZE, m, "fp64")
    cfg = abi.NNConfig()
    for k in range(11):
        cfg.in_std[k] = 1.0
    for k in range(9):
        cfg.out_std[k] = 1.0
    cfg.d_max = -1.0  # every cell out of distribution
    cfg.elem_tol, cfg.neg_tol, cfg.boxcox_lambda, cfg.dt = 1e30, 1.0, 0.1, 1e-7
    ctx.nn_init(cfg)
    ctx.nn_set_params(_rand_params(10))
    tag = torch.zeros((Q.shape[1], Q.shape[2], Q.shape[3]), dtype=torch.int64, device="cuda")
    tag[4 + 3, 4 + 2, 4:4 + 5] = 1  # five tagged cells stay untouched
    Qn, rn, Un = Q.clone(), r.clone(), U.clone()
    ctx.chem_newton(Un, Qn, rn, 1e-7, 50, 1e-10, 1e-6, tag=tag)
    st = torch.zeros((SIZE[2], SIZE[1], SIZE[0]), dtype=torch.int8, device="cuda")
    Q0 = Q.clone()
    nfb = ctx.nn_chem(U, Q, r, 1e-7, 50, 1e-10, 1e-6, status=st, tag=tag)
    torch.cuda.synchronize()
    ncell = SIZE[0] * SIZE[1] * SIZE[2]
    assert nfb == ncell - 5
    assert int((st == 2).sum()) == ncell - 5 and int((st == 0).sum()) == 5
    assert torch.equal(Q, Qn) and torch.equal(r, rn) and torch.equal(U, Un)
    assert torch.equal(Q[:, 7, 6, 4:9], Q0[:, 7, 6, 4:9])
    with pytest.raises(abi.F3DError):
        ctx.nn_chem(U, Q, r, 2e-7)  # dt other than the trained step
    ctx.close()


@pytest.mark.parametrize("label_mode", [0, 3])
def test_nn_chem_post_predict_and_projection(label_mode):
    """With all weights zero the network outputs out_mean exactly; the
    committed state must equal the host restatement of the label inversion,
    the element projection (fixed out_std^2 weights in mode 0, per-cell
    (s_k + |dY_k|)^2 weights in mode 3), post_predict and the FP64 energy
    bookkeeping to 1e-12."""
    print(f"\n[surrogate] nn_chem post_predict label_mode={label_mode} size={SIZE}", flush=True)
    abi = _abi()
    from oracle import Oracle
    m, Q, r, U = _cells(seed=21)
    ctx = abi.Context(*SIZE, m, "fp64")
    cfg = abi.NNConfig()
    for k in range(11):
        cfg.in_std[k] = 1.0
    if label_mode == 0:
        mean = [-1.0e-4, -8.0e-4, 9.0e-4, 1.0e-5, 1.0e-5, 2.0e-5, 1.0e-6, 1.0e-7, 3.0e-5]
        std = [1e-3, 2e-3, 2e-3, 1e-4, 1e-4, 3e-4, 1e-5, 1e-6, 1e-5]
    else:
        mean = [-2.0, -1.5, 2.5, 1.2, 0.7, 1.9, 0.4, -0.3, 0.6]  # asinh(dY / s) labels
        std = [1.0] * 9
        cfg.label_mode = 3
        cfg.label_rel, cfg.label_floor = 0.1, 1e-14
        for k in range(9):
            cfg.label_scale[k] = 1e-6
    for k in range(9):
        cfg.out_mean[k], cfg.out_std[k] = mean[k], std[k]
    cfg.d_max, cfg.elem_tol, cfg.neg_tol, cfg.res_tol = 1e30, 1e30, 1.0, 0.0
    cfg.boxcox_lambda, cfg.dt = 0.1, 1e-7
    cfg.n_elem = 3
    for e in range(3):
        for k in range(9):
            cfg.elem[e][k] = float(m.composition[e][k])
    ctx.nn_init(cfg)
    ctx.nn_set_params(np.zeros(12512, np.float32))  # prediction = out_mean exactly
    q0, r0, u0 = _interior(Q), _interior(r), _interior(U)
    st = torch.zeros((SIZE[2], SIZE[1], SIZE[0]), dtype=torch.int8, device="cuda")
    nfb = ctx.nn_chem(U, Q, r, 1e-7, status=st)
    torch.cuda.synchronize()
    assert nfb == 0 and bool((st == 1).all())
    q1, r1, u1 = _interior(Q), _interior(r), _interior(U)
    comp = np.asarray(m.composition) / np.asarray(m.mw)[None, :]
    o = np.array(mean, np.float32).astype(np.float64)
    Pm = _projection(m, cfg)

    def increment(Y):
        if label_mode == 0:
            return Pm @ o
        f32 = [float(np.float32(v)) for v in (1e-6, 1e-14, 0.1)]
        sk = np.minimum(f32[0], np.maximum(f32[1], f32[2] * Y))
        d0 = sk * np.sinh(o)
        S = (sk + np.abs(d0)) ** 2 + 1e-60
        z = np.linalg.solve((comp * S[None, :]) @ comp.T, comp @ d0)
        return d0 - S * (comp.T @ z)

    # element conservation of the projected increment, to rounding of the
    # projection: 1e-12 of the gross elemental turnover, or 1e-15 of the cell's
    # elemental content (per-cell weights spanning 1e-28..1e-12 condition the
    # 3x3 solve at ~1e8 in cells that hold only traces of a species class)
    for c in (0, 100, 1000):
        Y = np.maximum(r0[c] / q0[c, 0], 0)
        dY = increment(Y)
        assert abs(dY[8]) <= 1e-12 * np.max(np.abs(dY))  # N2 is inert: projected out
        bar = 1e-12 * np.abs(comp) @ np.abs(dY) + 1e-15 * np.abs(comp) @ Y
        assert np.all(np.abs(comp @ dY) <= bar), (c, comp @ dY, bar)
    orc = Oracle(*SIZE, m)
    lo6 = np.asarray(m.coeffs_lo)[5] * m.Ru / np.asarray(m.mw)
    worst = 0.0
    for c in range(0, q0.shape[0], 7):
        rho = q0[c, 0]
        Y = np.maximum(r0[c] / rho, 0)
        dY = increment(Y)
        rn = np.maximum(Y + dY, 0) * rho
        dei = -np.sum(lo6 * (rn - r0[c]))
        rn = rn * (rho / rn.sum())
        ein = q0[c, 6] + dei
        T = orc.get_T_from(ein, rn, q0[c, 5])  # the kernel warm-starts GetT at the old T
        p = m.Ru * T * np.sum(rn / np.asarray(m.mw))
        e_sp = np.max(np.abs(r1[c] - rn)) / rho
        exp = np.array([T, p, ein, u0[c, 4] + dei])
        got = np.array([q1[c, 5], q1[c, 4], q1[c, 6], u1[c, 4]])
        worst = max(worst, e_sp, np.max(np.abs(got - exp) / np.abs(exp)))
    assert worst < 1e-12, worst
    ctx.close()


def test_nn_chem_residual_gate():
    print(f"\n[surrogate] nn_chem residual gate size={SIZE}", flush=True)
    abi = _abi()
    from oracle import Oracle
    m, Q, r, U = _cells(seed=33)
    ctx = abi.Context(*SIZE, m, "fp64")
    cfg = abi.NNConfig()
    for k in range(11):
        cfg.in_std[k] = 1.0
    for k in range(9):
        cfg.out_std[k] = 1.0
    tol = 1e-8
    cfg.d_max, cfg.elem_tol, cfg.neg_tol, cfg.res_tol = 1e30, 1e30, 1.0, tol
    cfg.boxcox_lambda, cfg.dt = 0.1, 1e-7
    ctx.nn_init(cfg)
    ctx.nn_set_params(np.zeros(12512, np.float32))  # identity prediction dY = 0
    Qn, rn, Un = Q.clone(), r.clone(), U.clone()
    ctx.chem_newton(Un, Qn, rn, 1e-7, 50, 1e-10, 1e-6)
    q0, r0 = _interior(Q), _interior(r)
    st = torch.zeros((SIZE[2], SIZE[1], SIZE[0]), dtype=torch.int8, device="cuda")
    nfb = ctx.nn_chem(U, Q, r, 1e-7, 50, 1e-10, 1e-6, status=st)
    torch.cuda.synchronize()
    s = st.cpu().numpy().reshape(-1)
    assert nfb == int((s >= 2).sum()) and 0 < nfb < s.size
    assert set(np.unique(s).tolist()) <= {1, 6}  # only the residual gate is active
    q1, r1, qn, rnn = _interior(Q), _interior(r), _interior(Qn), _interior(rn)
    # rejected cells are exactly the Newton result
    rej = s == 6
    assert np.array_equal(q1[rej], qn[rej]) and np.array_equal(r1[rej], rnn[rej])
    orc = Oracle(*SIZE, m)
    mw = np.asarray(m.mw)
    checked = 0
    for c in range(0, s.size, 3):
        rho = q0[c, 0]
        rr = np.maximum(r0[c], 0)
        rr = rr * (rho / rr.sum())
        T = orc.get_T_from(q0[c, 6], rr, q0[c, 5])
        w = orc.production_rate(rr / mw * 1e-6, T, jac=False)
        R = np.max(np.abs(rr - r0[c] - w * mw * 1e6 * 1e-7)) / rho
        if abs(R / tol - 1) < 1e-6:
            continue
        assert (s[c] == 1) == (R <= tol), (c, R, s[c])
        checked += 1
    assert checked > s.size // 4
    ctx.close()


def test_nn_chem_chain_gate():
    """Identity prediction (dY = 0) on fresh radical-free reactors with only
    the chain-explosiveness gate on: a cell falls back (status 8) exactly when
    the 3x3 block dt d wdot / d c over H, O, OH at the predicted state has an
    eigenvalue with real part >= chain_max; rejected cells equal the Newton
    step.  The kernel differentiates the rates analytically (mass action and
    [M], falloff factor held fixed); the host restatement takes forward
    differences of the oracle's rates (relative step 1e-8) and numpy's
    eigenvalues, so cells within 1e-3 of the threshold are not judged."""
    print(f"\n[surrogate] nn_chem chain gate size={SIZE}", flush=True)
    abi = _abi()
    from oracle import Oracle
    m, Q, r, U = _cells(seed=44)
    ctx = abi.Context(*SIZE, m, "fp64")
    cfg = abi.NNConfig()
    for k in range(11):
        cfg.in_std[k] = 1.0
    for k in range(9):
        cfg.out_std[k] = 1.0
    chain = [m.index("H"), m.index("O"), m.index("OH")]
    cmax = 0.5
    cfg.d_max, cfg.elem_tol, cfg.neg_tol, cfg.res_tol = 1e30, 1e30, 1.0, 0.0
    cfg.chain_max, cfg.chain_species = cmax, sum(1 << k for k in chain)
    cfg.boxcox_lambda, cfg.dt = 0.1, 1e-7
    ctx.nn_init(cfg)
    ctx.nn_set_params(np.zeros(12512, np.float32))  # identity prediction dY = 0
    Qn, rn, Un = Q.clone(), r.clone(), U.clone()
    ctx.chem_newton(Un, Qn, rn, 1e-7, 50, 1e-10, 1e-6)
    q0, r0 = _interior(Q), _interior(r)
    st = torch.zeros((SIZE[2], SIZE[1], SIZE[0]), dtype=torch.int8, device="cuda")
    nfb = ctx.nn_chem(U, Q, r, 1e-7, 50, 1e-10, 1e-6, status=st)
    torch.cuda.synchronize()
    s = st.cpu().numpy().reshape(-1)
    assert nfb == int((s >= 2).sum()) and 0 < nfb < s.size
    assert set(np.unique(s).tolist()) <= {1, 8}  # only the chain gate is active
    q1, r1, qn, rnn = _interior(Q), _interior(r), _interior(Qn), _interior(rn)
    rej = s == 8
    assert np.array_equal(q1[rej], qn[rej]) and np.array_equal(r1[rej], rnn[rej])
    orc = Oracle(*SIZE, m)
    mw = np.asarray(m.mw)
    checked = 0
    for c in range(0, s.size, 3):
        rho = q0[c, 0]
        rr = np.maximum(r0[c], 0)
        rr = rr * (rho / rr.sum())
        T = orc.get_T_from(q0[c, 6], rr, q0[c, 5])
        sc = rr / mw * 1e-6
        w0 = orc.production_rate(sc, T, jac=False)
        eps = 1e-8 * sc.sum()
        M = np.zeros((3, 3))
        for j, cj in enumerate(chain):
            s2 = sc.copy()
            s2[cj] += eps
            M[:, j] = 1e-7 * (orc.production_rate(s2, T, jac=False) - w0)[chain] / eps
        lam = np.linalg.eigvals(M).real.max()
        if abs(lam / cmax - 1) < 1e-3:
            continue
        assert (s[c] == 1) == (lam < cmax), (c, lam, s[c])
        checked += 1
    assert checked > s.size // 4
    ctx.close()


SURROGATE = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))),
                         "flame3d.jl_amd", "mechanisms", "h2_surrogate.npz")


@pytest.mark.skipif(not os.path.exists(SURROGATE), reason="trained surrogate not present")
def test_trained_surrogate_accuracy_and_gain():
    """Fresh reactors (seed unseen in training) at 64^3: surrogate step with
    gates and Newton fallback vs the Newton step alone.  Bars: species
    mass-fraction error <= 1e-3 and temperature error <= 1 K (Linf over all
    cells), and the step runs faster than Newton alone, for two settings of
    the chain gate: the shipped 0.05, set for flows in which a cell stays in
    the chain-branching induction phase for many steps (where the network's
    per-step radical error compounds into a shifted ignition), and 0.5, the
    setting for one step from a fresh state, at which the network must take
    over at least half of the cells and of the Newton iterations."""
    import flame3d_amd.cases as cases
    import flame3d_amd.surrogate as sur
    abi = _abi()
    shape = (64, 64, 64)
    print(f"\n[surrogate] trained network on {shape} fresh reactors", flush=True)
    m = _mech()
    ctx = abi.Context(*shape, m, "fp64")
    Q0, r0 = cases.chem_cells(shape, m, seed=2024, dev="cuda")
    U0 = torch.zeros((5,) + tuple(Q0.shape[1:]), dtype=torch.float64, device="cuda")
    Qn, rn, Un = Q0.clone(), r0.clone(), U0.clone()
    it = torch.zeros((shape[2], shape[1], shape[0]), dtype=torch.int32, device="cuda")
    ctx.chem_newton(Un, Qn, rn, sur.DT, 50, 1e-10, 1e-6, iters=it)
    its = it.reshape(-1).double()

    def best_ms(fn):
        ts = []
        for _ in range(3):  # three repetitions from the same fresh state
            Qf, rf, Uf = Q0.clone(), r0.clone(), torch.zeros_like(U0)
            torch.cuda.synchronize()
            e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            e0.record()
            fn(Uf, Qf, rf)
            e1.record()
            torch.cuda.synchronize()
            ts.append(e0.elapsed_time(e1))
        return min(ts)

    t_newton = best_ms(lambda U, Q, r: ctx.chem_newton(U, Q, r, sur.DT, 50, 1e-10, 1e-6))
    for chain_max in (None, 0.5):
        over = {} if chain_max is None else {"chain_max": chain_max}
        cfg, _ = sur.attach(ctx, SURROGATE, **over)
        Q, r, U = Q0.clone(), r0.clone(), U0.clone()
        st = torch.zeros_like(it, dtype=torch.int8)
        nfb = ctx.nn_chem(U, Q, r, sur.DT, 50, 1e-10, 1e-6, status=st)
        torch.cuda.synchronize()
        Ye = (sur.mass_fractions(Q, r) - sur.mass_fractions(Qn, rn)).abs().max().item()
        Te = (sur.interior(Q)[:, 5] - sur.interior(Qn)[:, 5]).abs().max().item()
        avoided = float(its[st.reshape(-1) == 1].sum() / its.sum())
        t_nn = best_ms(lambda U, Q, r: ctx.nn_chem(U, Q, r, sur.DT, 50, 1e-10, 1e-6))
        frac = nfb / st.numel()
        print(f"  chain_max {cfg.chain_max:.2f}: Y linf {Ye:.2e}  T linf {Te:.3f} K  fallback "
              f"{frac:.3f}  newton iterations avoided {avoided:.3f}  newton {t_newton:.2f} ms  "
              f"surrogate {t_nn:.2f} ms", flush=True)
        assert Ye <= 1e-3
        assert Te <= 1.0
        assert t_nn < t_newton
        if chain_max == 0.5:
            assert frac <= 0.5
            assert avoided >= 0.5
    ctx.close()


AIR_SURROGATE = os.path.join(os.path.dirname(SURROGATE), "air5_surrogate.npz")


@pytest.mark.skipif(not os.path.exists(AIR_SURROGATE), reason="trained air surrogate not present")
def test_trained_air_surrogate_accuracy():
    """Fresh dissociating-air reactors (seed unseen in training; T ~ U[1500,
    6000] K, P ~ U[0.02, 0.1] atm) at 64^3, one chemistry step of 6e-9 s:
    surrogate with gates and Newton fallback vs the Newton step alone.
    Bars: species error <= 1e-3 and temperature error <= 1 K (Linf), the
    network takes over at least half of the cells."""
    import flame3d_amd.cases as cases
    import flame3d_amd.mech as mm
    import flame3d_amd.surrogate as sur
    abi = _abi()
    shape = (64, 64, 64)
    print(f"\n[surrogate] trained air network on {shape} fresh reactors", flush=True)
    m = mm.load("air5")
    ctx = abi.Context(*shape, m, "fp64")
    cfg, _ = sur.attach(ctx, AIR_SURROGATE)
    # trained for the 6e-9 s step; of the chain carriers only O is in air
    assert abs(cfg.dt - sur.AIR_DT) < 1e-20 and cfg.chain_species == 1 << m.index("O")
    Q, r = cases.chem_cells_air(shape, m, seed=2024, T_range=(1500.0, 6000.0),
                                P_atm=(0.02, 0.1))
    Q, r = Q.cuda().contiguous(), r.cuda().contiguous()
    U = torch.zeros((5,) + tuple(Q.shape[1:]), dtype=torch.float64, device="cuda")
    Qn, rn, Un = Q.clone(), r.clone(), U.clone()
    Q0, r0 = Q.clone(), r.clone()
    ctx.chem_newton(Un, Qn, rn, sur.AIR_DT, 50, 1e-10, 1e-6)
    st = torch.zeros((shape[2], shape[1], shape[0]), dtype=torch.int8, device="cuda")
    nfb = ctx.nn_chem(U, Q, r, sur.AIR_DT, 50, 1e-10, 1e-6, status=st)
    torch.cuda.synchronize()
    Ye = (sur.mass_fractions(Q, r) - sur.mass_fractions(Qn, rn)).abs().max().item()
    Te = (sur.interior(Q)[:, 5] - sur.interior(Qn)[:, 5]).abs().max().item()
    dY0 = (sur.mass_fractions(Qn, rn) - sur.mass_fractions(Q0, r0)).abs().max().item()
    times = {}
    for name in ("newton", "surrogate"):
        ts = []
        for _ in range(3):
            Qf, rf, Uf = Q0.clone(), r0.clone(), torch.zeros_like(U)
            torch.cuda.synchronize()
            e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            e0.record()
            if name == "newton":
                ctx.chem_newton(Uf, Qf, rf, sur.AIR_DT, 50, 1e-10, 1e-6)
            else:
                ctx.nn_chem(Uf, Qf, rf, sur.AIR_DT, 50, 1e-10, 1e-6)
            e1.record()
            torch.cuda.synchronize()
            ts.append(e0.elapsed_time(e1))
        times[name] = min(ts)
    frac = nfb / st.numel()
    print(f"  Y linf {Ye:.2e} (step's own change {dY0:.2e})  T linf {Te:.3f} K  fallback "
          f"{frac:.4f}  newton {times['newton']:.2f} ms  surrogate {times['surrogate']:.2f} ms",
          flush=True)
    assert dY0 > 1e-4  # the step does move the composition
    assert Ye <= 1e-3
    assert Te <= 1.0
    assert frac <= 0.5
    ctx.close()


def test_train_on_integrator_data():
    """End to end at small scale: integrator data -> statistics -> training
    by the library's kernels -> surrogate step; the held-out loss must fall
    well below the untrained network's."""
    import flame3d_amd.surrogate as sur
    abi = _abi()
    m = _mech()
    print("\n[surrogate] small end-to-end training", flush=True)
    data = sur.generate(m, (16, 16, 8), steps=200, n_record=8, seed=5, log=None)
    assert data.X.shape == (8 * 2048, 11) and data.Y1.shape == (8 * 2048, 9)
    cfg = sur.make_config(m, data)
    assert cfg.d_max > 0 and cfg.n_elem == 3
    Xn, Yn = sur.normalize(cfg, data)
    ctx = abi.Context(8, 8, 8, m, "fp64")
    ctx.nn_init(cfg)
    ctx.nn_set_params(sur.init_params(1))
    Xv, Yv = Xn[:1024].contiguous(), Yn[:1024].contiguous()
    l0 = sur.mse(ctx, Xv, Yv)
    sur.train(ctx, Xn[1024:].contiguous(), Yn[1024:].contiguous(), epochs=30, batch=1024,
              lr=3e-3, decay_every=15, log=None)
    l1 = sur.mse(ctx, Xv, Yv)
    print(f"  held-out mse {l0:.3e} -> {l1:.3e}", flush=True)
    assert l1 < 0.35 * l0  # 30 epochs on 15k samples: a fourfold drop
    ctx.close()
