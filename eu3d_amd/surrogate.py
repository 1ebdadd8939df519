"""Learned chemistry surrogate: harvesting, normalisation, training and evaluation.

The network and all of its arithmetic (BF16 MFMA GEMMs, GELU, backward, Adam, the features and
the fallback checks) run in libeu3d.so; this module drives them. Training pairs are harvested
from the solver's own FP64 chemistry: states (rho_s, T, E) before and after one adaptive
trapezoid step of interval dt on the GPU. torch tensors hold device memory only.
"""
from __future__ import annotations

import math
import time

import numpy as np

from . import configs
from .api import MLP, Context, SurrogateParams
from .mech import R_UNIV, load_default

NIN_REAL, NOUT_REAL = 12, 9


def _torch():
    import torch
    return torch


def to_bf16_bits(x):
    """float tensor -> int16 tensor holding BF16 bits (round to nearest even)."""
    torch = _torch()
    return x.to(torch.bfloat16).view(torch.int16)


def raw_features(rho, T, dt):
    """Un-normalised features (T, ln p/atm, ln max(Y, 1e-20), ln dt) of batch cells, FP64 on the
    device, for the normalisation statistics; the kernels compute the same quantities."""
    torch = _torch()
    mw = torch.tensor(load_default().mw, dtype=torch.float64, device=rho.device)
    D = rho.sum(0)
    Y = rho / D
    sy = (Y / mw[:, None]).sum(0)
    p = T * D * (R_UNIV * 1000 * sy)
    f = torch.empty((T.shape[0], NIN_REAL), dtype=torch.float64, device=rho.device)
    f[:, 0] = T
    f[:, 1] = torch.log(p / 101325.0)
    f[:, 2:11] = torch.log(torch.clamp(Y, min=1e-20)).T
    f[:, 11] = math.log(dt) if not torch.is_tensor(dt) else torch.log(dt)
    return f


class Harvest:
    """FP64 increments: rho0 -> rho1 over dt for batches of cells at rest."""

    def __init__(self):
        self.rho0, self.rho1, self.T0, self.dt = [], [], [], []

    def add(self, rho0, rho1, T0, dt, keep=None):
        """Keep the pairs whose states before and after are finite (a few cells of the FP64
        chemistry leave the finite range at large dt), and where the boolean device mask keep
        holds, if given."""
        torch = _torch()
        ok = torch.isfinite(rho0).all(0) & torch.isfinite(rho1).all(0) & torch.isfinite(T0)
        if keep is not None:
            ok &= keep
        self.rho0.append(rho0[:, ok].contiguous())
        self.rho1.append(rho1[:, ok].contiguous())
        self.T0.append(T0[ok].contiguous())
        self.dt.append(torch.full_like(self.T0[-1], dt))

    def cat(self):
        torch = _torch()
        return (torch.cat(self.rho0, 1), torch.cat(self.rho1, 1), torch.cat(self.T0), torch.cat(self.dt))

    @property
    def n(self):
        return sum(t.shape[0] for t in self.T0)


def fp64_chem(ctx: Context, rho, T, E, dt, nchem_max=10000):
    """One FP64 adaptive-trapezoid chemistry step on device arrays (in place); returns stats."""
    n = T.shape[0]
    return ctx.chem_batch(n, rho.data_ptr(), T.data_ptr(), E.data_ptr(), dt, nchem_max)


def harvest_states(ctx: Context, H: Harvest, n: int, dts, seed: int = 0, T_range=(900.0, 3000.0)):
    """Seeded chemistry states (mixture fraction / progress variable, as configs.chem_states),
    one FP64 step per dt."""
    torch = _torch()
    for i, dt in enumerate(dts):
        rho, T, E, _ = configs.chem_states(n, seed=seed + 7919 * i)
        if T_range != (900.0, 3000.0):
            T = np.random.default_rng(seed + i).uniform(*T_range, n)
            mech = load_default()
            Y = rho / rho.sum(1, keepdims=True)
            E = configs._nasa_e(mech, T, Y)
        r = torch.tensor(rho.T.copy(), device="cuda")
        t = torch.tensor(T, device="cuda")
        e = torch.tensor(E, device="cuda")
        r0, t0 = r.clone(), t.clone()
        fp64_chem(ctx, r, t, e, dt)
        H.add(r0, r, t0, dt)


def harvest_trajectories(ctx: Context, H: Harvest, n: int, dts, nsteps: int, seed: int = 0, keep_every: int = 1):
    """FP64 chemistry trajectories: seeded states (configs.chem_states) advanced nsteps times by
    each dt, every keep_every-th (before, after) pair harvested, so that radical pools and the
    ignition transients enter the data."""
    torch = _torch()
    for i, dt in enumerate(dts):
        rho, T, E, _ = configs.chem_states(n, seed=seed + 7919 * i)
        r = torch.tensor(rho.T.copy(), device="cuda")
        t = torch.tensor(T, device="cuda")
        e = torch.tensor(E, device="cuda")
        for s in range(nsteps):
            r0, t0 = r.clone(), t.clone()
            fp64_chem(ctx, r, t, e, dt)
            if s % keep_every == 0:
                H.add(r0, r.clone(), t0, dt)


def harvest_run(H: Harvest, wl, nsteps: int, every: int, cfl=0.5, ramp=(0.05, 20), min_T=600.0, max_cells=1 << 20):
    """Snapshots of a plain FP64 run of a workload: every `every` steps the interior state after
    the flow update is taken as a chemistry input and advanced by the step's dt with the FP64
    chemistry (the solver's own increments)."""
    torch = _torch()
    g = Context(wl.xnode, wl.ynode, wl.znode, quirks=0)
    for k in ("T", "P", "U", "V", "W", "YI"):
        g.set(k, wl.prims[k])
    g.field_initial()
    kw = dict(wl.step)
    kw.pop("cfl", None)
    for s in range(nsteps):
        c = ramp[0] + (cfl - ramp[0]) * min(1.0, s / ramp[1])
        g.step(1, cfl=c, bc_kinds=wl.bc_kinds, **kw)
        if (s + 1) % every == 0:
            dt, _ = g.get_time()
            Tf = g.get("T")[2:-2, 2:-2, 2:-2].reshape(-1)
            rho = g.get("CS")[2:-2, 2:-2, 2:-2, :9].reshape(-1, 9)
            E = g.get("E")[2:-2, 2:-2, 2:-2].reshape(-1)
            keep = np.flatnonzero(Tf >= min_T)[:max_cells]
            if keep.size == 0:
                continue
            r = torch.tensor(rho[keep].T.copy(), device="cuda")
            t = torch.tensor(Tf[keep], device="cuda")
            e = torch.tensor(E[keep], device="cuda")
            r0, t0 = r.clone(), t.clone()
            fp64_chem(g, r, t, e, dt)
            H.add(r0, r, t0, dt)
    g.close()


def increments(rho0, rho1, out_eps=0.0):
    """Training outputs of harvested pairs [n x 9]: dY_k, or d ln(Y_k + eps) when eps > 0 (the
    quantity the kernels' targets and apply step use)."""
    torch = _torch()
    Y0, Y1 = (rho0 / rho0.sum(0)).T, (rho1 / rho1.sum(0)).T
    if out_eps > 0:   # the FP64 chemistry leaves some mass fractions a rounding error below zero
        return torch.log(Y1.clamp(min=0) + out_eps) - torch.log(Y0.clamp(min=0) + out_eps)
    return Y1 - Y0


def fit_params(rho0, rho1, T0, dt, hull_margin=0.25, dT_max=40.0, elem_tol=2e-3, y_neg_tol=2e-4, T_cold=500.0,
               nchem_threshold=0, out_eps=0.0, nchem_threshold_hi=0):
    """Normalisation and training hull from harvested pairs. nchem_threshold > 0 sends only the
    cells needing more FP64 sub-steps than that through the network (and, with
    nchem_threshold_hi > 0, no more than that many). out_eps > 0 makes the
    network's outputs log increments d ln(Y_k + out_eps), positive mass fractions by construction."""
    torch = _torch()
    f = raw_features(rho0, T0, dt)
    dY = increments(rho0, rho1, out_eps)
    p = SurrogateParams()
    # the kernels normalise with the FP32 copies of these statistics: round first, so the hull
    # below is taken on the features exactly as the kernels compute them
    fm = f.mean(0).float().double()
    fs = f.std(0).clamp(min=1e-3).float().double()
    om, osd = dY.mean(0), dY.std(0).clamp(min=1e-12)
    fn = (f - fm) / fs
    lo, hi = fn.min(0).values, fn.max(0).values
    span = (hi - lo).clamp(min=1e-3)
    for i in range(NIN_REAL):
        p.feat_mean[i], p.feat_std[i] = float(fm[i]), float(fs[i])
        p.hull_lo[i], p.hull_hi[i] = float(lo[i] - hull_margin * span[i]), float(hi[i] + hull_margin * span[i])
    for i in range(NOUT_REAL):
        p.out_mean[i], p.out_std[i] = float(om[i]), float(osd[i])
    p.dT_max, p.elem_tol, p.y_neg_tol, p.T_cold = dT_max, elem_tol, y_neg_tol, T_cold
    p.nchem_threshold = nchem_threshold
    p.nchem_threshold_hi = nchem_threshold_hi
    p.out_eps = out_eps
    return p


def make_dataset(mlp: MLP, rho0, rho1, T0, dt):
    """Normalised BF16 features [n x 32] and FP32 targets [n x 16] computed by the kernels."""
    torch = _torch()
    n = T0.shape[0]
    feats = torch.empty((n, MLP.NIN), dtype=torch.int16, device="cuda")
    tg = torch.empty((n, MLP.NOUT), dtype=torch.float32, device="cuda")
    # features depend on dt per group: process contiguous runs of equal dt
    dts = dt.cpu().numpy()
    starts = np.flatnonzero(np.r_[True, dts[1:] != dts[:-1]])
    ends = np.r_[starts[1:], n]
    for a, b in zip(starts, ends):
        r = rho0[:, a:b].contiguous()
        mlp.features(r, T0[a:b].contiguous(), float(dts[a]), feats[a:b])
    mlp.targets(rho0.contiguous(), rho1.contiguous(), tg)
    return feats, tg


def train(mlp: MLP, feats, tg, epochs=20, batch=16384, lr=2e-3, lr_end=1e-4, wd=0.0, seed=0, log=None):
    """Adam over shuffled minibatches with a cosine learning-rate decay; returns the loss history."""
    torch = _torch()
    n = feats.shape[0]
    g = torch.Generator(device="cuda")
    g.manual_seed(seed)
    steps = epochs * max(1, n // batch)
    hist, k = [], 0
    t0 = time.time()
    for ep in range(epochs):
        perm = torch.randperm(n, device="cuda", generator=g)
        tot, nb = 0.0, 0
        for i in range(0, n - batch + 1, batch):
            idx = perm[i:i + batch]
            x = feats[idx].contiguous()
            y = tg[idx].contiguous()
            a = lr_end + 0.5 * (lr - lr_end) * (1 + math.cos(math.pi * k / max(1, steps)))
            tot += mlp.train_step(x, y, a, wd)
            nb += 1
            k += 1
        hist.append(tot / max(1, nb))
        if log:
            log(f"epoch {ep} loss {hist[-1]:.4e} ({time.time() - t0:.1f} s)")
    return hist


def predict_dY(mlp: MLP, p: SurrogateParams, feats):
    """Network outputs [n x 9] (FP64, de-normalised): dY_k, or d ln(Y_k + eps) in the log mode."""
    torch = _torch()
    out = torch.empty((feats.shape[0], MLP.NOUT), dtype=torch.float32, device="cuda")
    mlp.forward(feats, out)
    om = torch.tensor(list(p.out_mean), dtype=torch.float64, device="cuda")
    os_ = torch.tensor(list(p.out_std), dtype=torch.float64, device="cuda")
    return out[:, :NOUT_REAL].double() * os_ + om
