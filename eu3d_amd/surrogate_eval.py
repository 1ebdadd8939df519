"""Train the chemistry surrogate on the solver's own FP64 increments and measure what it gains over
the FP64 chemistry alone, and what it costs in accuracy.

Two workloads, each a JSON report:
- batch: 2^22 seeded cells on the H2-air equilibrium lines, one chemistry step of dt = 1e-8 s.
  The FP64 adaptive trapezoid (persistent kernel, tuned schedule) is compared with the hybrid
  surrogate chemistry. In the hybrid, cells needing more than `threshold` FP64 sub-steps go
  through the network, the rest take the FP64 sub-steps, and the rejects fall back to FP64.
  Training pairs come from other seeds of the same generator.
- detonation: the 3-D planar detonation, plain FP64 steps against steps with the hybrid surrogate
  chemistry, run side by side from the same initial state. It reports:
  - step times;
  - the front position (mean over y and z of the largest x with T > 1500 K);
  - L-inf / L2 errors of T, p and Y_k;
  - the fraction of cells sent through the network and the fraction rejected.
  Training pairs are harvested from plain runs of a smaller transverse section of the same
  configuration.
- plain-detonation: the plain FP64 detonation step alone (its time per step over the last 40 of
  140 steps), the driver of the detonation kernel profile.

- stiff-gain: what the trained network gains on the stiff batched chemistry (stiff_gain), the
  bench's and the gain test's measurement.
- train-at-scale: the network trained on about 1e7 pairs harvested from the 1-D tube, the
  shock-bubble and the batched chemistry (train_at_scale).

Usage: python -m eu3d_amd.surrogate_eval batch|detonation|plain-detonation|stiff-gain|train-at-scale [options]
"""
from __future__ import annotations

import argparse
import json
import time

import numpy as np

from . import autotune, chemistry, configs, surrogate
from .api import MLP, Context


def _torch():
    import torch
    return torch


def _timed(ctx_stream, fn):
    torch = _torch()
    s = torch.cuda.ExternalStream(ctx_stream)
    torch.cuda.synchronize()
    a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    a.record(s)
    out = fn()
    b.record(s)
    b.synchronize()
    return out, a.elapsed_time(b)


def harvest_batch(ctx, seeds, n, dtm, min_nchem=0, log=print):
    """(rho0, rho1, T0, dt) pairs: one FP64 chemistry step of equilibrium-line states. With
    min_nchem > 0 only the cells needing more FP64 sub-steps than that are kept: the cells the
    hybrid sends through the network."""
    torch = _torch()
    H = surrogate.Harvest()
    for sd in seeds:
        log(f"harvest seed {sd}: {n} states")
        (r, t, e), _ = chemistry.states(ctx, n, seed=sd, burnt="equilibrium", log=log)
        r0, t0 = r.clone(), t.clone()
        nch = torch.empty(n, dtype=torch.int32, device="cuda")
        ctx.chem_batch(n, r.data_ptr(), t.data_ptr(), e.data_ptr(), dtm, 10000, nch.data_ptr())
        H.add(r0, r, t0, dtm, keep=(nch > min_nchem) if min_nchem > 0 else None)
        log(f"  kept {H.n} pairs so far")
    return H


def train_network(H, epochs=12, batch=16384, lr=2e-3, threshold=0, log=print, **fit):
    r0, r1, t0, dt = H.cat()
    m = MLP(seed=11)
    p = surrogate.fit_params(r0, r1, t0, dt, nchem_threshold=threshold, **fit)
    m.set_surrogate(p)
    X, Y = surrogate.make_dataset(m, r0, r1, t0, dt)
    t = time.time()
    hist = surrogate.train(m, X, Y, epochs=epochs, batch=batch, lr=lr, log=log)
    return m, p, {"pairs": int(X.shape[0]), "epochs": epochs, "loss_first": hist[0], "loss_last": hist[-1],
                  "train_s": time.time() - t}


def errors(T_a, T_b, r_a, r_b):
    """Errors of state a against reference b (device tensors; rho [ns][n])."""
    fin = np.isfinite(T_b) & np.isfinite(T_a)
    Ya = (r_a / r_a.sum(0))[:, fin]
    Yb = (r_b / r_b.sum(0))[:, fin]
    dT = np.abs(T_a[fin] - T_b[fin])
    return {"T_Linf_K": float(dT.max()), "T_Linf_rel": float((dT / T_b[fin]).max()),
            "T_L2_rel": float(np.sqrt((dT ** 2).mean()) / np.sqrt((T_b[fin] ** 2).mean())),
            "Y_Linf": float(np.abs(Ya - Yb).max()), "nonfinite": int((~fin).sum())}


def eval_batch(n=1 << 22, dtm=1e-8, threshold=8, train_seeds=(101, 102, 103), n_train=1 << 21, epochs=12,
               stiff_only=False, log=print, **fit):
    """stiff_only: train on the cells above the hybrid threshold only (the cells the network sees)."""
    torch = _torch()
    x = configs.make_nodes(8, 1e-4)
    ctx = Context(x, x, x, quirks=0)
    H = harvest_batch(ctx, train_seeds, n_train, dtm, min_nchem=threshold if stiff_only else 0, log=log)
    log(f"training on {H.n} pairs")
    m, p, tr = train_network(H, epochs=epochs, threshold=threshold, log=log, **fit)
    tr["stiff_only"] = stiff_only
    tr["checks"] = {"dT_max": p.dT_max, "y_neg_tol": p.y_neg_tol, "elem_tol": p.elem_tol, "hull_margin": fit.get("hull_margin", 0.25)}
    del H
    log(f"evaluation states: {n}")
    (r0, t0, e0), _ = chemistry.states(ctx, n, seed=2024, burnt="equilibrium", log=log)
    sched = autotune.load_or_tune(ctx, (r0, t0, e0), dtm)
    log(f"schedule {sched}")
    out = {"cells": n, "dt": dtm, "threshold": threshold, "training": tr,
           "schedule": {k: sched[k] for k in ("block", "blocks_per_cu", "waves", "tail")}}
    # FP64 reference
    best = None
    for _ in range(2):
        r, t, e = r0.clone(), t0.clone(), e0.clone()
        st, ms = _timed(ctx.stream(), lambda: ctx.chem_batch(n, r.data_ptr(), t.data_ptr(), e.data_ptr(), dtm, 10000))
        best = ms if best is None else min(best, ms)
    ref_r, ref_t = r.cpu().numpy(), t.cpu().numpy()
    out["fp64"] = {"ms": best, "cells_per_s": n / (best * 1e-3), "stats": st}
    # hybrid surrogate
    best = None
    for _ in range(2):
        r, t, e = r0.clone(), t0.clone(), e0.clone()
        sst, ms = _timed(ctx.stream(), lambda: ctx.surrogate_batch(m, n, r.data_ptr(), t.data_ptr(), e.data_ptr(), dtm,
                                                                    10000))
        best = ms if best is None else min(best, ms)
    out["surrogate"] = {"ms": best, "cells_per_s": n / (best * 1e-3), "stats": sst,
                        "network_fraction": sst["network"] / n,
                        "rejected_fraction_of_network": sst["fallback"] / max(1, sst["network"])}
    out["speedup"] = out["fp64"]["ms"] / best
    out["errors_vs_fp64"] = errors(t.cpu().numpy(), ref_t, r.cpu().numpy(), ref_r)
    # errors over the cells the network answered (routed, not rejected), the rest being FP64 anyway
    nch = torch.empty(n, dtype=torch.int32, device="cuda")
    r2, t2, e2 = r0.clone(), t0.clone(), e0.clone()
    ctx.chem_batch(n, r2.data_ptr(), t2.data_ptr(), e2.data_ptr(), dtm, 10000, nch.data_ptr())
    routed = (nch.cpu().numpy() > threshold) & (t0.cpu().numpy() >= p.T_cold)
    tn = t.cpu().numpy()
    diff = np.abs(tn - ref_t) > 0
    sel = routed & diff
    if sel.any():
        out["errors_network_cells"] = errors(tn[sel], ref_t[sel], r.cpu().numpy()[:, sel], ref_r[:, sel])
        out["errors_network_cells"]["cells"] = int(sel.sum())
    ctx.close()
    return out


STIFF_TAILS = (512, 256, 128, 64)


def stiff_gain(n=1 << 21, dtm=1e-8, lo=8, hi=512, train_seeds=(301, 302, 303), n_train=1 << 20, epochs=30,
               seed=2024, tails=STIFF_TAILS, log=print):
    """What the trained network gains on the stiff batched chemistry (equilibrium-line states, one
    step of dtm), and its error against the FP64 result cell by cell.

    The network (log-increment outputs, abundance-weighted element projection, |dT| <= 40 K) is
    trained on the FP64 increments of other seeds, on the cells needing more than 4 and at most
    2 hi FP64 sub-steps. In the hybrid the cells needing lo+1..hi sub-steps go through the network;
    the rest, and the rejects of the checks, take the FP64 path. Two cases: the whole batch, and
    the same states without the cells needing more than hi sub-steps (the igniting cells, whose
    FP64 chains up to the 10000 cap stay on FP64 in both runs and set the whole batch's time).
    Each leg is timed at each tail threshold in `tails` (the sub-step count above which a cell goes
    to the tail kernel) and reported at its own best, so neither is timed at a schedule tuned for
    the other."""
    torch = _torch()
    x = configs.make_nodes(8, 1e-4)
    ctx = Context(x, x, x, quirks=0)
    ctx.set_schedule(256, 4, 2, 512)
    t_start = time.time()
    H = surrogate.Harvest()
    for sd in train_seeds:
        (r, t, e), _ = chemistry.states(ctx, n_train, seed=sd, burnt="equilibrium")
        r0, t0 = r.clone(), t.clone()
        nch = torch.empty(n_train, dtype=torch.int32, device="cuda")
        ctx.chem_batch(n_train, r.data_ptr(), t.data_ptr(), e.data_ptr(), dtm, 10000, nch.data_ptr())
        H.add(r0, r, t0, dtm, keep=(nch > 4) & (nch <= 2 * hi))
    r0, r1, t0, dt = H.cat()
    m = MLP(seed=11)
    p = surrogate.fit_params(r0, r1, t0, dt, nchem_threshold=lo, nchem_threshold_hi=hi, out_eps=1e-10, dT_max=40.0)
    m.set_surrogate(p)
    X, Y = surrogate.make_dataset(m, r0, r1, t0, dt)
    hist = surrogate.train(m, X, Y, epochs=epochs, batch=4096, lr=2e-3)
    log(f"surrogate on stiff chemistry: {H.n} pairs, loss {hist[0]:.4f} -> {hist[-1]:.4f}")
    out = {"training": {"pairs": int(H.n), "epochs": epochs, "loss_first": hist[0], "loss_last": hist[-1],
                        "harvest_and_train_s": time.time() - t_start, "band_substeps": [lo + 1, hi],
                        "checks": {"dT_max_K": p.dT_max, "elem_tol": p.elem_tol, "y_neg_tol": p.y_neg_tol,
                                   "T_cold_K": p.T_cold}}}
    del H, r0, r1, t0, dt, X, Y
    (re, te, ee), _ = chemistry.states(ctx, n, seed=seed, burnt="equilibrium")
    nch = torch.empty(n, dtype=torch.int32, device="cuda")
    r, t, e = re.clone(), te.clone(), ee.clone()
    ctx.chem_batch(n, r.data_ptr(), t.data_ptr(), e.data_ptr(), dtm, 10000, nch.data_ptr())
    keep = (nch <= hi).nonzero().squeeze(1)
    cases = {"whole_batch": (re, te, ee),
             "band": (re[:, keep].contiguous(), te[keep].contiguous(), ee[keep].contiguous())}
    for name, (ra, ta, ea) in cases.items():
        k = ta.shape[0]
        rf, tf, ef = ra.clone(), ta.clone(), ea.clone()
        rs, ts, es = ra.clone(), ta.clone(), ea.clone()
        best_f = best_s = None
        for tail in tails:
            ctx.set_schedule(256, 4, 2, tail)
            for leg in ("fp64", "hybrid"):
                ms_best = None
                for _ in range(2):
                    if leg == "fp64":
                        rf.copy_(ra), tf.copy_(ta), ef.copy_(ea)
                        st, ms = _timed(ctx.stream(), lambda: ctx.chem_batch(k, rf.data_ptr(), tf.data_ptr(),
                                                                             ef.data_ptr(), dtm, 10000))
                    else:
                        rs.copy_(ra), ts.copy_(ta), es.copy_(ea)
                        st, ms = _timed(ctx.stream(), lambda: ctx.surrogate_batch(m, k, rs.data_ptr(), ts.data_ptr(),
                                                                                  es.data_ptr(), dtm, 10000))
                    ms_best = ms if ms_best is None else min(ms_best, ms)
                log(f"  {name}, tail threshold {tail}: {leg} {ms_best:.3f} ms")
                if leg == "fp64" and (best_f is None or ms_best < best_f[0]):
                    best_f = (ms_best, tail, st)
                if leg == "hybrid" and (best_s is None or ms_best < best_s[0]):
                    best_s = (ms_best, tail, st)
        T_f, T_s = tf.cpu().numpy(), ts.cpu().numpy()
        fin = np.isfinite(T_f)
        dT = np.abs(T_s - T_f)[fin]
        changed = dT > 1e-9 * T_f[fin]
        sts = best_s[2]
        res = {"cells": int(k), "fp64_ms": best_f[0], "fp64_tail": best_f[1], "hybrid_ms": best_s[0],
               "hybrid_tail": best_s[1], "speedup": best_f[0] / best_s[0],
               "fp64_cells_per_s": k / (best_f[0] * 1e-3), "hybrid_cells_per_s": k / (best_s[0] * 1e-3),
               "fp64_substeps": best_f[2]["substeps_total"], "network": sts["network"], "rejected": sts["fallback"],
               "cells_changed": int(changed.sum()), "fp64_nonfinite": int((~fin).sum()),
               "errors_vs_fp64": errors(T_s, T_f, rs.cpu().numpy(), rf.cpu().numpy())}
        if changed.any():
            res["T_err_network_cells_K"] = {"p50": float(np.median(dT[changed])),
                                            "p99": float(np.percentile(dT[changed], 99)), "max": float(dT.max()),
                                            "max_rel": float((dT / T_f[fin]).max())}
        out[name] = res
        log(f"  {name}: {res}")
    ctx.close()
    return out


def front_x(T, xnode, bc=2, T_front=1500.0):
    """Mean over (y, z) of the largest interior x with T > T_front."""
    Ti = T[bc:-bc, bc:-bc, bc:-bc]
    hot = Ti > T_front
    idx = np.where(hot.any(0), Ti.shape[0] - 1 - np.argmax(hot[::-1], axis=0), 0)
    return float(np.mean(xnode[bc + idx]))


def eval_detonation(grid=(512, 256, 128), nsteps=150, threshold=4, train_grid=(512, 32, 32), train_steps=150,
                    every=5, epochs=10, log=print):
    torch = _torch()
    # harvest from plain runs of a narrower section (same generator, other seed)
    H = surrogate.Harvest()
    wl_t = configs.det3d(*train_grid, seed=777)
    log(f"harvest: plain run {train_grid}, {train_steps} steps")
    surrogate.harvest_run(H, wl_t, train_steps, every, max_cells=1 << 20)
    log(f"harvested {H.n} pairs")
    m, p, tr = train_network(H, epochs=epochs, threshold=threshold, log=log)
    del H
    wl = configs.det3d(*grid)
    ctxs = []
    for _ in range(2):
        g = Context(wl.xnode, wl.ynode, wl.znode, quirks=0)
        for k in ("T", "P", "U", "V", "W", "YI"):
            g.set(k, wl.prims[k])
        g.field_initial()
        ctxs.append(g)
    plain, sur = ctxs
    sur.set_surrogate(m)
    kw = dict(wl.step)
    kw.pop("cfl")
    times = {"plain": 0.0, "surrogate": 0.0}
    net = rej = cells = 0
    for s in range(nsteps):
        c = 0.05 + (0.5 - 0.05) * min(1.0, s / 20)
        last = s == nsteps - 1   # the derived fields are refreshed once, after the last step
        _, ms = _timed(plain.stream(), lambda: plain.step(1, cfl=c, bc_kinds=wl.bc_kinds, skip_derive=not last, **kw))
        times["plain"] += ms
        kws = dict(kw, chem=4)
        _, ms = _timed(sur.stream(), lambda: sur.step(1, cfl=c, bc_kinds=wl.bc_kinds, skip_derive=not last, **kws))
        times["surrogate"] += ms
        st = sur.surrogate_stats()
        if (s + 1) % 10 == 0:
            log(f"step {s + 1}: plain {times['plain'] / (s + 1):.2f} ms/step, surrogate "
                f"{times['surrogate'] / (s + 1):.2f} ms/step, network {st['network']}, rejected {st['fallback']}")
        net += st["network"]
        rej += st["fallback"]
        cells += st["cells"]
    Tp, Ts = plain.get("T"), sur.get("T")
    Pp, Ps = plain.get("P"), sur.get("P")
    Yp, Ys = plain.get("YI"), sur.get("YI")
    sl = (slice(2, -2), slice(2, -2), slice(2, -2))
    fin = np.isfinite(Tp[sl]) & np.isfinite(Ts[sl])
    dT = np.abs(Ts[sl] - Tp[sl])[fin]
    dP = np.abs(Ps[sl] - Pp[sl])[fin]
    out = {"grid": list(grid), "steps": nsteps, "threshold": threshold, "training": tr,
           "ms_per_step": {k: v / nsteps for k, v in times.items()},
           "speedup": times["plain"] / times["surrogate"],
           "front_x_m": {"plain": front_x(Tp, wl.xnode), "surrogate": front_x(Ts, wl.xnode)},
           "T_Linf_rel": float((dT / np.abs(Tp[sl][fin])).max()),
           "T_L2_rel": float(np.sqrt((dT ** 2).mean() / (Tp[sl][fin] ** 2).mean())),
           "p_Linf_rel": float((dP / np.abs(Pp[sl][fin])).max()),
           "p_L2_rel": float(np.sqrt((dP ** 2).mean() / (Pp[sl][fin] ** 2).mean())),
           "Y_Linf": float(np.abs(Ys[sl] - Yp[sl])[fin].max()),
           "network_fraction": net / max(1, cells), "rejected_fraction": rej / max(1, cells),
           "nonfinite_cells": {"plain": int((~np.isfinite(Tp[sl])).sum()), "surrogate": int((~np.isfinite(Ts[sl])).sum())}}
    dx = wl.xnode[3] - wl.xnode[2]
    out["front_error_cells"] = abs(out["front_x_m"]["surrogate"] - out["front_x_m"]["plain"]) / dx
    for g in ctxs:
        g.close()
    return out


def time_plain_detonation(grid=(512, 256, 128), nsteps=140, window=40, log=print):
    """The plain FP64 detonation step alone: adaptive trapezoid chemistry, the CFL ramp of
    eval_detonation, one step per call with the derived fields refreshed after the last. Returns
    the mean ms/step of the last `window` steps (HIP events on the context stream)."""
    _torch()
    wl = configs.det3d(*grid)
    g = Context(wl.xnode, wl.ynode, wl.znode, quirks=0)
    for k in ("T", "P", "U", "V", "W", "YI"):
        g.set(k, wl.prims[k])
    g.field_initial()
    kw = dict(wl.step)
    kw.pop("cfl")
    ts = []
    for s in range(nsteps):
        c = 0.05 + (0.5 - 0.05) * min(1.0, s / 20)
        last = s == nsteps - 1
        _, ms = _timed(g.stream(), lambda: g.step(1, cfl=c, bc_kinds=wl.bc_kinds, skip_derive=not last, **kw))
        ts.append(ms)
        if (s + 1) % 20 == 0:
            log(f"step {s + 1}: {ms:.2f} ms")
    T = g.get("T")[2:-2, 2:-2, 2:-2]
    out = {"grid": list(grid), "steps": nsteps, "ms_per_step_last": float(np.mean(ts[-window:])),
           "window": window, "cells_per_s": float(np.prod(grid) / (np.mean(ts[-window:]) * 1e-3)),
           "T_finite": bool(np.isfinite(T).all())}
    g.close()
    return out


def train_at_scale(zn_steps=2000, bub_steps=500, bub_every=50, bub_cells=1 << 19, batch_seeds=(401, 402, 403),
                   n_batch=1 << 20, epochs=30, batch=1 << 16, holdout=1 << 16, log=print):
    """The network trained at the scale of the surrogate's configuration: about 1e7 (state,
    increment) pairs harvested from the solver's FP64 adaptive trapezoid on the other three
    workloads, trained by the GPU kernels (fused forward, backward, Adam).

    Sources: the 1-D detonation tube (2000 cells, a snapshot after every one of its `zn_steps`
    steps, cells at 600 K and above); the shock-bubble (384 x 128 x 128, every `bub_every`-th of
    its `bub_steps` steps, `bub_cells` cells a snapshot, every temperature); the batched chemistry
    (`n_batch` equilibrium-line states of each seed in `batch_seeds`, one step of dt = 1e-8 s).
    A snapshot's pairs advance the interior state after the step's flow and chemistry update by
    the step's dt. `holdout` random pairs are kept out of training; the report gives the pairs
    per source, the training loss per epoch, the held-out loss before and after, and the time."""
    torch = _torch()
    t_all = time.time()
    H = surrogate.Harvest()
    counts = {}
    t = time.time()
    surrogate.harvest_run(H, configs.zn1d(), zn_steps, 1, max_cells=1 << 20)
    counts["zn1d_2000"] = H.n
    log(f"harvest zn1d: {H.n} pairs ({time.time() - t:.1f} s)")
    t = time.time()
    surrogate.harvest_run(H, configs.bubble3d(384, 128, 128), bub_steps, bub_every, min_T=0.0, max_cells=bub_cells)
    counts["bubble3d_384x128x128"] = H.n - sum(counts.values())
    log(f"harvest bubble3d: {counts['bubble3d_384x128x128']} pairs ({time.time() - t:.1f} s)")
    t = time.time()
    x = configs.make_nodes(8, 1e-4)
    ctx = Context(x, x, x, quirks=0)
    for sd in batch_seeds:
        (r, tt, e), _ = chemistry.states(ctx, n_batch, seed=sd, burnt="equilibrium")
        r0, t0 = r.clone(), tt.clone()
        ctx.chem_batch(n_batch, r.data_ptr(), tt.data_ptr(), e.data_ptr(), 1e-8, 10000)
        H.add(r0, r, t0, 1e-8)
    ctx.close()
    counts["batch_equilibrium"] = H.n - sum(counts.values())
    log(f"harvest batch: {counts['batch_equilibrium']} pairs ({time.time() - t:.1f} s)")
    t_harvest = time.time() - t_all
    r0, r1, t0, dt = H.cat()
    del H
    m = MLP(seed=11)
    p = surrogate.fit_params(r0, r1, t0, dt)
    m.set_surrogate(p)
    X, Y = surrogate.make_dataset(m, r0, r1, t0, dt)
    del r0, r1, t0, dt
    n = X.shape[0]
    g = torch.Generator(device="cuda")
    g.manual_seed(5)
    perm = torch.randperm(n, device="cuda", generator=g)
    hv, tr_idx = perm[:holdout], perm[holdout:]
    Xh, Yh = X[hv].contiguous(), Y[hv].contiguous()
    Xt, Yt = X[tr_idx].contiguous(), Y[tr_idx].contiguous()
    del X, Y, perm

    def loss_of(Xs, Ys):
        """Inference MSE over the 9 outputs, and the share of it from the worst 0.1 % of pairs."""
        out = torch.empty((Xs.shape[0], MLP.NOUT), dtype=torch.float32, device="cuda")
        m.forward(Xs, out)
        se = ((out[:, :surrogate.NOUT_REAL] - Ys[:, :surrogate.NOUT_REAL]) ** 2).mean(1)
        top = se.topk(max(1, se.shape[0] // 1000)).values
        return {"mse": float(se.mean()), "share_of_worst_0.1pct": float(top.sum() / se.sum())}

    def held_out_loss():
        return loss_of(Xh, Yh)

    h0 = held_out_loss()
    torch.cuda.synchronize()
    t = time.time()
    hist = surrogate.train(m, Xt, Yt, epochs=epochs, batch=batch, lr=2e-3, log=log)
    torch.cuda.synchronize()
    t_train = time.time() - t
    h1 = held_out_loss()
    sub = torch.randperm(Xt.shape[0], device="cuda", generator=g)[:holdout]
    t1 = loss_of(Xt[sub].contiguous(), Yt[sub].contiguous())
    full = {"mse": 0.0}
    for a in range(0, Xt.shape[0], 1 << 20):   # the whole training set in inference, 2^20 rows at a time
        part = loss_of(Xt[a:a + (1 << 20)].contiguous(), Yt[a:a + (1 << 20)].contiguous())
        full["mse"] += part["mse"] * min(1 << 20, Xt.shape[0] - a) / Xt.shape[0]
    steps = epochs * (Xt.shape[0] // batch)
    return {"pairs": {"total": int(n), "training": int(Xt.shape[0]), "held_out": int(holdout), **counts},
            "harvest_s": t_harvest, "epochs": epochs, "minibatch": batch, "optimizer_steps": steps,
            "train_s": t_train, "pairs_per_s": steps * batch / t_train,
            "loss_per_epoch": hist, "held_out_loss": {"before": h0, "after": h1},
            "training_loss_after": {"random_subset_of_held_out_size": t1, "whole_set": full},
            "outputs": "normalised dY_k of the 9 species (MSE over the 9 outputs)"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("workload", choices=("batch", "detonation", "plain-detonation", "stiff-gain", "train-at-scale"))
    ap.add_argument("--threshold", type=int, default=None)
    ap.add_argument("--steps", type=int, default=150)
    ap.add_argument("--grid", type=str, default="512x256x128")
    ap.add_argument("--epochs", type=int, default=None)
    ap.add_argument("--stiff-only", action="store_true", help="batch: train on the cells above the threshold only")
    ap.add_argument("--y-neg-tol", type=float, default=None)
    ap.add_argument("--dT-max", type=float, default=None)
    ap.add_argument("--seeds", type=int, default=3, help="batch: training seeds of 2^21 states each")
    a = ap.parse_args()
    log = lambda s: print(s, flush=True)
    if a.workload == "batch":
        kw = {}
        if a.threshold is not None:
            kw["threshold"] = a.threshold
        if a.epochs:
            kw["epochs"] = a.epochs
        if a.y_neg_tol is not None:
            kw["y_neg_tol"] = a.y_neg_tol
        if a.dT_max is not None:
            kw["dT_max"] = a.dT_max
        kw["train_seeds"] = tuple(101 + i for i in range(a.seeds))
        res = eval_batch(log=log, stiff_only=a.stiff_only, **kw)
    elif a.workload == "stiff-gain":
        res = stiff_gain(log=log)
    elif a.workload == "train-at-scale":
        res = train_at_scale(log=log) if a.epochs is None else train_at_scale(epochs=a.epochs, log=log)
    elif a.workload == "plain-detonation":
        res = time_plain_detonation(grid=tuple(int(v) for v in a.grid.split("x")), nsteps=a.steps, log=log)
    else:
        kw = dict(grid=tuple(int(v) for v in a.grid.split("x")), nsteps=a.steps)
        if a.threshold is not None:
            kw["threshold"] = a.threshold
        if a.epochs:
            kw["epochs"] = a.epochs
        res = eval_detonation(log=log, **kw)
    print(json.dumps(res))


if __name__ == "__main__":
    main()
