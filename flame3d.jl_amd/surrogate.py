"""Learned chemistry surrogate, host side: training data from the GPU
backward-Euler integrator, normalization and gate statistics, training by
the library's BF16 MFMA kernels, persistence, and the surrogate step.

Reference pattern:
  * NN/H2/gen_data.py: (state, state after dt) pairs sampled at random steps
    along 0-D reactor trajectories with dt = 1e-7 s and Box-Cox lambda = 0.1;
    here the trajectories come from the library's own FP64 backward-Euler
    Newton integrator (f3d_chem_newton) instead of Cantera.
  * NN/H2/lux-ck.jl:38-71: MSE on z-scored labels, Adam, step decay of the
    learning rate, seed 12345.
  * reactions.jl:2-56 (pre_input / post_predict) and solver.jl:204-253: the
    inference call, fused here into f3d_nn_chem with Newton fallback.

Inputs (11): T, log P, BoxCox(max(Y_k, 1e-12)) = (Y^lambda - 1)/lambda for
the 9 species; outputs (9): the increments dY_k over dt, learned as
asinh(dY_k / s_k) with the per-cell scale s_k = min(1e-6, max(1e-14,
0.1 Y_k)) (label_mode 3): linear below s_k and logarithmic above, so the
network's relative accuracy holds over the decades that matter for T, and a
trace species' increment stays relative to the trace (a fixed s = 1e-6 left
an absolute error floor of ~1e-7 that seeded radicals into the free streams
of a flow); the reference's Box-Cox-space increments, plain dY and the fixed
scale are the other label forms.  Inputs and labels are z-scored; the
Mahalanobis gate uses the mean and inverse covariance of the z-scored
training inputs and the training distances' 3-sigma quantile.
"""
from __future__ import annotations

import json
import math
import os
import time
from dataclasses import dataclass, field
from typing import Dict, List, Optional

import numpy as np

from . import abi, cases

NG = abi.NG
NN_IN, NN_OUT, NPARAM = 11, 9, abi.NN_NPARAM  # largest network (H2/O2); air uses 7 / 5
LAMBDA = 0.1   # NN/H2/gen_data.py: lamda = 0.1
DT = 1e-7      # NN/H2/gen_data.py: dt = 1e-7
SEED = 12345   # lux-ck.jl:35-36
HERE = os.path.dirname(os.path.abspath(__file__))
DEFAULT_PATH = os.path.join(HERE, "mechanisms", "h2_surrogate.npz")
AIR_PATH = os.path.join(HERE, "mechanisms", "air5_surrogate.npz")
AIR_DT = 6e-9  # the air network's chemistry step: the jet's half-step at its fixed dt of 1.2e-8 s

# layer shapes of the packed parameter vector (layout in csrc/f3d_nn.h)
_LAYERS = [(32, 64, 11, 64), (64, 64, 64, 64), (64, 64, 64, 64), (64, 32, 64, 9)]


def _offsets():
    w, b, off = [], [], 0
    for ni, no, _, _ in _LAYERS:
        w.append(off)
        off += ni * no
        b.append(off)
        off += no
    return w, b


W_OFF, B_OFF = _offsets()


def init_params(seed: int = SEED, nin: int = NN_IN, nout: int = NN_OUT) -> np.ndarray:
    """Glorot-uniform weights on the true fan-in/fan-out (nin = Ns + 2 inputs,
    nout = Ns outputs), zero biases and zero padding (Lux.Dense's default
    initialization)."""
    rng = np.random.default_rng(seed)
    p = np.zeros(NPARAM, np.float32)
    for l, (ni, no, fi, fo) in enumerate(_LAYERS):
        fi = nin if l == 0 else fi
        fo = nout if l == len(_LAYERS) - 1 else fo
        a = math.sqrt(6.0 / (fi + fo))
        W = np.zeros((no, ni), np.float32)
        W[:fo, :fi] = rng.uniform(-a, a, size=(fo, fi))
        p[W_OFF[l]:W_OFF[l] + ni * no] = W.reshape(-1)
    return p


def interior(t):
    """(NV, Nz+8, Ny+8, Nx+8) -> (C, NV) in interior cell-id order (x fastest)."""
    s = t[..., NG:-NG, NG:-NG, NG:-NG]
    return s.reshape(s.shape[0], -1).T


def features(Q, rhoi, lam: float = LAMBDA):
    """Raw surrogate inputs (C, 11) in FP64: T, log P, BoxCox(Y)."""
    import torch
    q = interior(Q)
    r = interior(rhoi)
    rho = q[:, 0:1].clamp_min(2.220446049250313e-16)
    Y = (r / rho).clamp_min(0.0)
    return torch.cat([q[:, 5:6], torch.log(q[:, 4:5]), (Y ** lam - 1.0) / lam], dim=1)


def mass_fractions(Q, rhoi):
    q = interior(Q)
    return (interior(rhoi) / q[:, 0:1].clamp_min(2.220446049250313e-16)).clamp_min(0.0)


def boxcox(Y, lam: float = LAMBDA):
    return (Y.clamp_min(0.0) ** lam - 1.0) / lam


@dataclass
class Dataset:
    X: "object"   # (M, 11) FP64 raw inputs, device
    Y0: "object"  # (M, 9) FP64 mass fractions before the step, device
    Y1: "object"  # (M, 9) FP64 mass fractions after one step of DT, device
    newton_iters: float = 0.0  # mean Newton iterations per sample

    def labels(self, label_mode: int, label_scale=None, label_rel: float = 0.0,
               label_floor: float = 0.0):
        """0: dY; 1: d BoxCox(Y) (the post_predict form, reactions.jl:26-56);
        2: asinh(dY / label_scale), logarithmic in |dY| above the scale;
        3: asinh(dY / s(Y0)), s = min(label_scale, max(label_floor, label_rel Y0))."""
        import torch
        if label_mode == 0:
            return self.Y1 - self.Y0
        if label_mode in (2, 3):
            sc = torch.as_tensor(list(label_scale)[:self.Y0.shape[1]], dtype=self.Y0.dtype,
                                 device=self.Y0.device)
            if label_mode == 3:
                sc = torch.minimum(sc, (label_rel * self.Y0).clamp_min(label_floor))
            return torch.asinh((self.Y1 - self.Y0) / sc)
        return boxcox(self.Y1) - boxcox(self.Y0)


def generate(mech, shape=(64, 64, 32), steps: int = 2000, n_record: int = 32, seed: int = SEED,
             maxiter: int = 50, rtol: float = 1e-10, ttol: float = 1e-6,
             T_range=(900.0, 2500.0), P_atm=(1.0, 10.0), z_pure: float = 0.1,
             trace: float = 0.0, dt: float = DT, log=print) -> Dataset:
    """Trajectory samples (NN/H2/gen_data.py pattern): Nx*Ny*Nz reactors drawn
    from the H2/O2/N2 distribution (T ~ U[900, 2500] K, P ~ U[1, 10] atm,
    mixture fraction ~ Beta(2, 2), plus a fraction z_pure each of pure
    oxidizer and pure fuel, the free streams of a non-premixed flow) are
    advanced by `steps` backward-Euler steps of DT on the GPU; at `n_record`
    steps (the first four, then random) every reactor contributes the pair
    (state, state after one step).  With trace > 0 every species of every
    reactor is seeded with a trace, log-uniform in [1e-14, trace] (the states
    that diffusion leaves in the free streams of a flow).  A mechanism
    without H2 (5-species air) draws the dissociating-air reactors of
    cases.chem_cells_air over T_range and P_atm instead."""
    import torch
    Nx, Ny, Nz = shape
    ctx = abi.Context(Nx, Ny, Nz, mech, precision="fp64")
    if "H2" in mech.species:
        Q, rhoi = cases.chem_cells(shape, mech, seed=seed, dev="cuda", T_range=T_range,
                                   P_atm=P_atm, z_pure=z_pure)
    else:
        Q, rhoi = cases.chem_cells_air(shape, mech, seed=seed, T_range=T_range, P_atm=P_atm)
        Q, rhoi = Q.cuda().contiguous(), rhoi.cuda().contiguous()
    if trace > 0:
        g = torch.Generator(device="cuda").manual_seed(seed + 1)
        rho = Q[0]
        lo, hi = math.log(1e-14), math.log(trace)
        tr = torch.exp(lo + (hi - lo) * torch.rand(rhoi.shape, generator=g, dtype=torch.float64,
                                                   device="cuda"))
        Y = rhoi / rho + tr
        Y = Y / Y.sum(0, keepdim=True)
        T, P = Q[5].clone(), Q[4].clone()
        zero = torch.zeros_like(T)
        Q, rhoi = cases.state_from_TPY(mech, T, P, Y, zero, zero, zero)
        Q, rhoi = Q.contiguous(), rhoi.contiguous()
    U = torch.zeros((5,) + tuple(Q.shape[1:]), dtype=torch.float64, device="cuda")
    rng = np.random.default_rng(seed)
    rec = set(range(min(4, n_record)))
    extra = rng.choice(np.arange(4, steps), size=max(0, n_record - len(rec)), replace=False)
    rec |= {int(s) for s in extra}
    iters = torch.zeros((Nz, Ny, Nx), dtype=torch.int32, device="cuda")
    Xs, Y0s, Y1s, it_sum = [], [], [], 0.0
    t0 = time.time()
    for step in range(steps):
        if step in rec:
            X = features(Q, rhoi)
            Y0 = mass_fractions(Q, rhoi)
            ctx.chem_newton(U, Q, rhoi, dt, maxiter, rtol, ttol, iters=iters)
            Xs.append(X)
            Y0s.append(Y0)
            Y1s.append(mass_fractions(Q, rhoi))
            it_sum += float(iters.double().mean())
        else:
            ctx.chem_newton(U, Q, rhoi, dt, maxiter, rtol, ttol)
        if log and (step + 1) % 500 == 0:
            torch.cuda.synchronize()
            log(f"  generate: step {step + 1}/{steps}, {len(Xs)} records, {time.time() - t0:.1f}s")
    ctx.close()
    return Dataset(torch.cat(Xs), torch.cat(Y0s), torch.cat(Y1s), it_sum / max(1, len(Xs)))


def generate_flow(mech, N=(192, 192, 192), steps: int = 8, per_call: int = 262144,
                  seed: int = 999, maxiter: int = 50, rtol: float = 1e-10, ttol: float = 1e-6,
                  dt: float = DT, log=print) -> Dataset:
    """Samples from a reacting flow itself (a different seed and resolution
    from the benchmark's): the reacting H2/air mixing layer, or for 5-species
    air the hot supersonic jet on a 2N x N x N block.  The FP64 flow with
    Strang Newton chemistry (half-steps of dt) is run for `steps` time steps,
    and at every chemistry half-step `per_call` random interior cells
    contribute (state, state after the backward-Euler step of dt).  This
    covers the states a flow produces and 0-D reactors do not (e.g. fresh gas
    with radicals diffused into it)."""
    import torch
    if "H2" in mech.species:
        case = cases.reacting_mixing_layer(N, mech, seed=seed, dev="cuda", dt=2 * dt)
        scheme = None
    else:
        N = (2 * N[0], N[1], N[2])
        case = cases.reacting_jet(N, mech, seed=seed, dev="cuda", dt=2 * dt)
        scheme = {"recon_kind": abi.RECON_WENO7}
    ctx = abi.Context(*N, mech, precision="fp64", scheme=scheme)
    ctx.set_metrics(case.grid.metrics, case.grid.J)
    ctx.set_boundary(case.bc)
    Q, r = case.Q.clone(), case.rhoi.clone()
    U = ctx.zeros(5)
    ctx.prim2c(U, Q)
    Un, rn = ctx.zeros(5), ctx.zeros(mech.Ns)
    ncell = N[0] * N[1] * N[2]
    g = torch.Generator(device="cuda").manual_seed(seed)
    Xs, Y0s, Y1s = [], [], []

    def chem():
        idx = torch.randint(0, ncell, (per_call,), generator=g, device="cuda")
        Xs.append(features(Q, r)[idx])
        Y0s.append(mass_fractions(Q, r)[idx])
        ctx.chem_newton(U, Q, r, dt, maxiter, rtol, ttol)
        Y1s.append(mass_fractions(Q, r)[idx])
        ctx.fill(Q, U, r)

    for step in range(steps):
        chem()
        for k in (1, 2, 3):
            ctx.rk3_stage(k, U, Un, r, rn, Q, case.dt)
        chem()
        if log:
            log(f"  generate_flow: step {step + 1}/{steps}")
    ctx.close()
    return Dataset(torch.cat(Xs), torch.cat(Y0s), torch.cat(Y1s), 0.0)


def concat(a: Dataset, b: Dataset) -> Dataset:
    import torch
    return Dataset(torch.cat([a.X, b.X]), torch.cat([a.Y0, b.Y0]), torch.cat([a.Y1, b.Y1]),
                   0.5 * (a.newton_iters + b.newton_iters))


def _elements(mech):
    names = list(mech.elements)
    comp = np.asarray(mech.composition, dtype=np.float64)  # [n_el, Ns]
    return names, comp


LABEL_SCALE = 1e-6  # asinh label scale: dY below it is linear, above it logarithmic
LABEL_REL = 0.1     # label_mode 3: below Y = LABEL_SCALE / LABEL_REL the scale follows 0.1 Y
LABEL_FLOOR = 1e-14  # ... down to this floor (species at zero)
Y_FLOOR = 1e-12     # inputs see Y below this as this (traces below it do not move the features)
CHAIN_MAX = 0.05    # chain-explosiveness gate: largest real eigenvalue of dt d wdot / d c
CHAIN_SPECIES = ("H", "O", "OH")  # ... over the chain carriers


def floor_inputs(X, y_floor: float, lam: float = LAMBDA):
    """Raw inputs with BoxCox(Y_k) taken at max(Y_k, y_floor) (BoxCox is
    increasing, so this is a clamp of the Box-Cox columns)."""
    if not y_floor > 0:
        return X
    import torch
    lo = (y_floor ** lam - 1.0) / lam
    return torch.cat([X[:, :2], X[:, 2:].clamp_min(lo)], dim=1)


def make_config(mech, data: Dataset, label_mode: int = 3, label_scale: float = LABEL_SCALE,
                elem_tol: float = 1e30, neg_tol: float = 1e-6, res_tol: float = 3e-5,
                res_T_tol: float = 0.25, dT_max: float = 20.0, chi_max: float = 0.0,
                chain_max: float = CHAIN_MAX, chain_species=CHAIN_SPECIES,
                label_rel: float = LABEL_REL, label_floor: float = LABEL_FLOOR,
                y_floor: float = Y_FLOOR, sigmas: float = 3.0,
                ridge: float = 1e-6, dt: float = DT) -> abi.NNConfig:
    """Normalization and gate statistics from the training data.  The
    Mahalanobis threshold is the training distances' quantile at the one-sided
    `sigmas`-sigma coverage of a normal law (0.99865 for 3 sigma)."""
    import torch
    X = floor_inputs(data.X, y_floor)
    nin, nout = X.shape[1], mech.Ns  # Ns + 2 inputs, Ns outputs
    lab = data.labels(label_mode, [label_scale] * nout, label_rel, label_floor)
    if not (bool(torch.isfinite(X).all()) and bool(torch.isfinite(lab).all())):
        raise ValueError("training data hold non-finite values")
    in_mean = X.mean(0)
    in_std = X.std(0).clamp_min(1e-12)
    out_mean = lab.mean(0)
    out_std = lab.std(0).clamp_min(1e-6 if label_mode == 2 else 1e-30)
    Xn = (X - in_mean) / in_std
    mu = Xn.mean(0)
    Xc = Xn - mu
    cov = (Xc.T @ Xc / (Xn.shape[0] - 1)).cpu().numpy() + ridge * np.eye(nin)
    inv_np = np.linalg.inv(cov)
    inv = torch.from_numpy(inv_np).to(X.device)
    d = torch.sqrt(torch.clamp(((Xc @ inv) * Xc).sum(1), min=0.0))
    cover = 0.5 * (1.0 + math.erf(sigmas / math.sqrt(2.0)))
    d_max = float(torch.quantile(d[:: max(1, d.numel() // 1_000_000)].float(), cover))
    cfg = abi.NNConfig()
    for k in range(nin):
        cfg.in_mean[k] = float(in_mean[k])
        cfg.in_std[k] = float(in_std[k])
        cfg.mahal_mean[k] = float(mu[k])
        for j in range(nin):
            cfg.mahal_inv[k * NN_IN + j] = float(inv[k, j])  # row stride 11 in every case
    for k in range(nout):
        cfg.out_mean[k] = float(out_mean[k])
        cfg.out_std[k] = float(out_std[k])
    cfg.d_max = d_max
    cfg.elem_tol = elem_tol
    cfg.neg_tol = neg_tol
    cfg.res_tol = res_tol
    cfg.res_T_tol = res_T_tol
    cfg.dT_max = dT_max
    cfg.chi_max = chi_max
    cfg.chain_max = chain_max
    cfg.chain_species = sum(1 << mech.index(n) for n in chain_species if n in mech.species)
    if cfg.chain_species == 0:  # carriers absent from the mechanism are dropped (air keeps O)
        cfg.chain_max = 0.0
    cfg.boxcox_lambda = LAMBDA
    cfg.label_mode = label_mode
    for k in range(nout):
        cfg.label_scale[k] = label_scale
    cfg.label_rel = label_rel
    cfg.label_floor = label_floor
    cfg.y_floor = y_floor
    names, comp = _elements(mech)
    cfg.n_elem = len(names)
    for e in range(len(names)):
        for k in range(nout):
            cfg.elem[e][k] = float(comp[e, k])
    cfg.dt = dt
    return cfg


def normalize(cfg: abi.NNConfig, data: Dataset):
    """z-scored float32 inputs and labels (label form from cfg.label_mode)."""
    import torch
    dev = data.X.device
    nin, nout = data.X.shape[1], data.Y0.shape[1]
    im = torch.tensor(list(cfg.in_mean)[:nin], dtype=torch.float64, device=dev)
    isd = torch.tensor(list(cfg.in_std)[:nin], dtype=torch.float64, device=dev)
    om = torch.tensor(list(cfg.out_mean)[:nout], dtype=torch.float64, device=dev)
    osd = torch.tensor(list(cfg.out_std)[:nout], dtype=torch.float64, device=dev)
    Xn = ((floor_inputs(data.X, cfg.y_floor) - im) / isd).float().contiguous()
    lab = data.labels(cfg.label_mode, list(cfg.label_scale), cfg.label_rel, cfg.label_floor)
    Yn = ((lab - om) / osd).float().contiguous()
    return Xn, Yn


def train(ctx: abi.Context, Xn, Yn, epochs: int = 60, batch: int = 8192, lr: float = 2e-3,
          decay: float = 0.3, decay_every: int = 20, seed: int = SEED, log=print) -> List[float]:
    """Adam on the MSE of z-scored labels with step decay of the learning
    rate (lux-ck.jl:20-71 schedule shape).  Returns the mean loss per epoch."""
    import torch
    M = (Xn.shape[0] // batch) * batch
    g = torch.Generator(device="cuda").manual_seed(seed)
    loss = torch.zeros(1, dtype=torch.float32, device="cuda")
    acc = torch.zeros(1, dtype=torch.float64, device="cuda")
    hist = []
    t0 = time.time()
    for ep in range(epochs):
        lr_ep = lr * decay ** (ep // decay_every)
        perm = torch.randperm(Xn.shape[0], generator=g, device="cuda")[:M]
        acc.zero_()
        nb = M // batch
        for b in range(nb):
            idx = perm[b * batch:(b + 1) * batch]
            xb = Xn.index_select(0, idx)
            yb = Yn.index_select(0, idx)
            ctx.nn_train_step(xb, yb, lr_ep, loss)
            acc += loss
        hist.append(float(acc) / nb)
        if log and (ep % 5 == 0 or ep == epochs - 1):
            log(f"  train: epoch {ep + 1}/{epochs} lr {lr_ep:.2e} loss {hist[-1]:.3e} "
                f"({time.time() - t0:.1f}s)")
    return hist


def mse(ctx: abi.Context, Xn, Yn) -> float:
    M = (Xn.shape[0] // 64) * 64
    Y = ctx.nn_forward(Xn[:M].contiguous())
    return float(((Y - Yn[:M]) ** 2).mean())


# ---- persistence (plain arrays: np.load with allow_pickle=False) ----------
_CFG_ARRAYS = ("in_mean", "in_std", "out_mean", "out_std", "mahal_mean", "mahal_inv",
               "label_scale")
_CFG_SCALARS = ("d_max", "elem_tol", "neg_tol", "res_tol", "res_T_tol", "dT_max", "chi_max",
                "chain_max", "chain_species",
                "boxcox_lambda", "label_mode", "label_rel", "label_floor", "y_floor", "n_elem",
                "dt")


def save(path: str, cfg: abi.NNConfig, params: np.ndarray, meta: Optional[Dict] = None):
    arrs = {k: np.array(list(getattr(cfg, k)), np.float32) for k in _CFG_ARRAYS}
    arrs["elem"] = np.array([list(r) for r in cfg.elem], np.float32)
    for k in _CFG_SCALARS:
        arrs[k] = np.array(getattr(cfg, k))
    arrs["params"] = np.asarray(params, np.float32)
    arrs["meta"] = np.array(json.dumps(meta or {}))
    np.savez(path, **arrs)


def load(path: str = DEFAULT_PATH):
    """(NNConfig, params float32[NPARAM], meta dict)."""
    z = np.load(path, allow_pickle=False)
    cfg = abi.NNConfig()
    for k in _CFG_ARRAYS:
        a = z[k]
        for i, v in enumerate(a):
            getattr(cfg, k)[i] = float(v)
    for e in range(4):
        for k in range(NN_OUT):
            cfg.elem[e][k] = float(z["elem"][e, k])
    for k in _CFG_SCALARS:
        if k in z.files:  # gates added later default to off
            setattr(cfg, k, type(getattr(cfg, k))(z[k][()]))
    return cfg, z["params"].astype(np.float32), json.loads(str(z["meta"][()]))


def attach(ctx: abi.Context, path: str = DEFAULT_PATH, **gate_overrides):
    """Load a trained surrogate into a context of the mechanism it was
    trained for (9-species H2/O2 or 5-species air)."""
    cfg, params, meta = load(path)
    for k, v in gate_overrides.items():
        setattr(cfg, k, v)
    ctx.nn_init(cfg)
    ctx.nn_set_params(params)
    return cfg, meta
