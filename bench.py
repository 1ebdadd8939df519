"""Benchmark: plain FP64 cell-updates/s of the EU3D hot path on one MI355X.

Workload: the 3-D shock / H2-bubble interaction, 384 x 128 x 128 cells, dx = 1e-4 m, Mach-2 shock
in air, stoichiometric H2-air bubble of radius 25 cells, IMEX chemistry, CFL 0.5 (seeded
synthetic initial condition). A step is one full time step: CFL dt, ghost fill, MUSCL/AUSM
advection of all three directions with the explicit update, and the IMEX chemistry with the
temperature Newton. Inputs are resident in HBM before the timed region.

Prints one JSON line. `roofline` is for the dominant kernel of the step, k_imex_fused (the fused
Diagonalized + Update_IMEX + Update_after_CS cell kernel), timed live with HIP events on the
context's stream; `kernels` gives the same for the two advection passes. `cpu_baseline` times the
CPU restatement of the reference arithmetic (the oracle) on a bounded sample of the same workload.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "cell-updates/s (flux+chem), plain FP64 vs MLP-chem surrogate at stated error"
HBM_PEAK_GBS = 8000.0
FP64_PEAK_TFLOPS = 78.6
# Algorithmic bytes per interior cell of each kernel of the plain IMEX step (FP64):
#  k_imex_fused (Diagonalized + Update_IMEX + Update_after_CS of one cell): read the 13 conserved
#    values, the 13 RHS values and the temperature guess, write the 13 conserved values and T
#  k_flux_z / k_flux_xy (the two advection passes): read the 16 reconstructed primitives
#    (P, rho, u, v, w, H, gamma, Y_1..9) and write 13 values (the z term, then the RHS, which
#    the x/y pass also reads)
KERNEL_BYTES_PER_CELL = {
    "imex_fused": (13 + 13 + 1 + 13 + 1) * 8,
    "flux_z": (16 + 13) * 8,
    "flux_xy": (16 + 13 + 13) * 8,
}
KERNEL_NAMES = {"imex_fused": "k_imex_fused", "flux_z": "k_flux_z<8, MUSCL_2 minmod>",
                "flux_xy": "k_flux_xy<16,16,XCD order, MUSCL_2 minmod>"}
PROFILE = "imex_fused_bubble3d_384x128x128_pmc.json"


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--grid", type=str, default="384x128x128")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--no-extra", action="store_true",
                    help="skip the batched-chemistry and surrogate-network sections")
    ap.add_argument("--cpu-grid", type=str, default="384x128x128")
    ap.add_argument("--cpu-steps", type=int, default=3)
    ap.add_argument("--copy-calibration", action="store_true",
                    help="only copy 256 MiB of float64 three times (the known-size kernel that scales the "
                         "FETCH_SIZE / WRITE_SIZE counters to bytes) and exit")
    return ap.parse_args()


def hip_timer(stream_ptr):
    import torch
    s = torch.cuda.ExternalStream(stream_ptr)

    def t():
        e = torch.cuda.Event(enable_timing=True)
        e.record(s)
        return e
    return t


def cpu_baseline(grid, steps):
    """Oracle (CPU restatement, OpenMP over cells) on a bounded sample of the same workload."""
    from eu3d_amd import configs
    import oracle
    ni, nj, nk = (int(v) for v in grid.split("x"))
    cores = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else os.cpu_count()
    cores = min(cores, int(os.environ.get("OMP_NUM_THREADS", cores)))
    oracle.set_threads(cores)
    wl = configs.bubble3d(ni, nj, nk, radius_cells=25.0 * ni / 384)
    o = oracle.OracleFlow(wl.xnode, wl.ynode, wl.znode, quirks=0)
    for k in ("T", "P", "U", "V", "W", "YI"):
        o.set(k, wl.prims[k])
    o.field_initial()
    kw = dict(cfl=0.5, bc_kinds=wl.bc_kinds, rk=1, chem=1)
    o.step(1, **kw)
    t0 = time.perf_counter()
    o.step(steps, **kw)
    el = time.perf_counter() - t0
    res = {"value": wl.ncells * steps / el, "unit": "cell-updates/s", "cores": cores, "kind": "port",
           "sample": f"bubble3d {ni}x{nj}x{nk} (same generator; bubble radius 25 cells scaled with ni/384), "
                     f"{steps} IMEX steps after 1 warm-up step, {el:.1f} s"}
    # one core, on a quarter-size section of the same generator
    oracle.set_threads(1)
    w1 = configs.bubble3d(96, 64, 64, radius_cells=25.0 * 96 / 384)
    o1 = oracle.OracleFlow(w1.xnode, w1.ynode, w1.znode, quirks=0)
    for k in ("T", "P", "U", "V", "W", "YI"):
        o1.set(k, w1.prims[k])
    o1.field_initial()
    o1.step(1, **dict(kw, bc_kinds=w1.bc_kinds))
    t0 = time.perf_counter()
    o1.step(1, **dict(kw, bc_kinds=w1.bc_kinds))
    el1 = time.perf_counter() - t0
    oracle.set_threads(cores)
    res["single_core"] = {"value": w1.ncells / el1, "unit": "cell-updates/s", "cores": 1,
                          "sample": f"bubble3d 96x64x64, 1 IMEX step after 1 warm-up step, {el1:.1f} s"}
    return res


def bench_batched_chemistry(ctx, n=1 << 22, dtm=1e-8, seed=2024):
    """Batched FP64 chemistry: 2^22 seeded cells, one adaptive trapezoid step of dt = 1e-8 s,
    with the persistent kernel's schedule from the autotuner. Reported for both burnt
    compositions (complete combustion, and the relaxed equilibrium line with its radical pool)."""
    import torch
    from eu3d_amd import autotune, chemistry
    out = {}
    for burnt in ("complete", "equilibrium"):
        (r0, t0, e0), _ = chemistry.states(ctx, n, seed=seed, burnt=burnt)
        sched = autotune.load_or_tune(ctx, (r0, t0, e0), dtm)
        s = torch.cuda.ExternalStream(ctx.stream())
        times = []
        for _ in range(3):
            r, t, e = r0.clone(), t0.clone(), e0.clone()
            torch.cuda.synchronize()
            a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            a.record(s)
            st = ctx.chem_batch(n, r.data_ptr(), t.data_ptr(), e.data_ptr(), dtm, 10000)
            b.record(s)
            b.synchronize()
            times.append(a.elapsed_time(b))
        ms = min(times)
        out[burnt] = {"cells": n, "dt": dtm, "ms": ms, "cells_per_s": n / (ms * 1e-3),
                      "substeps_per_s": st["substeps_total"] / (ms * 1e-3), "active_cells": st["active_cells"],
                      "substeps_total": st["substeps_total"], "nchem_max": st["nchem_max_seen"],
                      "schedule": {k: sched[k] for k in ("block", "blocks_per_cu", "waves", "tail")},
                      "schedule_from_cache": sched.get("cached", False)}
        del r0, t0, e0
    return out


def bench_surrogate_network(n=1 << 20, reps=10):
    """Surrogate network throughput at batch 2^20: inference (5 BF16 MFMA GEMMs with fused
    bias/GELU epilogues) and one training step (forward, backward, Adam), HIP events on the
    network's stream. FLOPs per cell / sample: 403,968 (inference), 1,205,760 (training)."""
    import torch
    from eu3d_amd.api import MLP
    m = MLP(seed=3)
    x = (torch.randn(n, MLP.NIN, device="cuda") * (torch.arange(MLP.NIN, device="cuda") < 12)).to(torch.bfloat16)
    x = x.view(torch.int16)
    y = torch.randn(n, MLP.NOUT, device="cuda")
    out = torch.empty(n, MLP.NOUT, device="cuda")
    s = torch.cuda.ExternalStream(m.stream())
    m.forward(x, out)
    m.train_step(x, y, 1e-4)

    def timed(fn):
        a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        a.record(s)
        for _ in range(reps):
            fn()
        b.record(s)
        b.synchronize()
        return a.elapsed_time(b) / reps
    fwd = timed(lambda: m.forward(x, out))
    trn = timed(lambda: m.train_step(x, y, 1e-4))
    fi, ft = 403968.0 * n, 1205760.0 * n
    return {"batch": n, "inference_ms": fwd, "inference_cells_per_s": n / (fwd * 1e-3),
            "inference_TFLOPs": fi / (fwd * 1e-3) / 1e12, "inference_frac_bf16_peak": fi / (fwd * 1e-3) / 2.5e15,
            "train_step_ms": trn, "train_samples_per_s": n / (trn * 1e-3),
            "train_TFLOPs": ft / (trn * 1e-3) / 1e12, "train_frac_bf16_peak": ft / (trn * 1e-3) / 2.5e15,
            "note": "inference includes the copy of the features in and the outputs out; training includes "
                    "the loss read-back"}


def read_profile():
    """HBM bytes per launch (FETCH_SIZE + WRITE_SIZE, scaled by the copy-kernel calibration) and
    FP64 FLOPs per cell of k_imex_fused from the committed rocprofv3 --pmc summary, or {}."""
    p = os.path.join(ROOT, "profiles", PROFILE)
    if not os.path.exists(p):
        return {}
    with open(p) as f:
        return json.load(f).get("kernels", {}).get("k_imex_fused", {})


def main():
    a = parse()
    if a.gpus != 1:
        print(json.dumps({"error": "one GPU only: runs on several GPUs are not part of this implementation",
                          "n_gpus": a.gpus}))
        sys.exit(2)
    import torch
    from eu3d_amd import configs
    from eu3d_amd.api import Context
    assert torch.cuda.is_available(), "bench.py needs the GPU"
    if a.copy_calibration:
        x = torch.ones(256 * 1024 * 1024 // 8, dtype=torch.float64, device="cuda")
        y = torch.empty_like(x)
        for _ in range(3):
            y.copy_(x)
        torch.cuda.synchronize()
        print(json.dumps({"copied_bytes_per_launch": x.numel() * 8, "launches": 3}))
        return
    ni, nj, nk = (int(v) for v in a.grid.split("x"))
    wl = configs.bubble3d(ni, nj, nk)
    g = Context(wl.xnode, wl.ynode, wl.znode, quirks=0)
    for k in ("T", "P", "U", "V", "W", "YI"):
        g.set(k, wl.prims[k])
    g.field_initial()
    g.sync()
    kw = dict(cfl=0.5, bc_kinds=wl.bc_kinds, rk=1, chem=1, diff=1, limiter=0)
    ev = hip_timer(g.stream())
    g.step(max(a.warmup, 1), **kw)
    g.sync()
    torch.cuda.synchronize()
    t0 = ev()
    w0 = time.perf_counter()
    g.step(a.steps, **kw)
    t1 = ev()
    g.sync()
    torch.cuda.synchronize()
    wall = time.perf_counter() - w0
    ms = t0.elapsed_time(t1) / a.steps
    value = wl.ncells / (ms * 1e-3)
    rc = g.check_finite()
    # per-kernel launch durations, HIP events on the context stream around each launch
    kern = {}
    for k in ("imex_fused", "flux_z", "flux_xy"):
        kms = g.time_kernel(k, reps=10)
        gbs = KERNEL_BYTES_PER_CELL[k] * wl.ncells / (kms * 1e-3) / 1e9
        kern[KERNEL_NAMES[k]] = {"avg_launch_ms": kms, "algorithmic_bytes_per_cell": KERNEL_BYTES_PER_CELL[k],
                                 "achieved_GBs": gbs, "frac_hbm": gbs / HBM_PEAK_GBS}
    dom = kern["k_imex_fused"]
    prof = read_profile()
    alg = KERNEL_BYTES_PER_CELL["imex_fused"] * wl.ncells
    achieved = alg / (dom["avg_launch_ms"] * 1e-3) / 1e9
    roof = {"kernel": "k_imex_fused (Diagonalized + Update_IMEX + Update_after_CS, one cell per lane)",
            "bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS, "traffic": prof.get("hbm_bytes_per_launch"),
            "algorithmic_bytes_per_launch": alg, "avg_launch_ms": dom["avg_launch_ms"]}
    if prof.get("fp64_flop_per_cell"):
        tf = prof["fp64_flop_per_cell"] * wl.ncells / (dom["avg_launch_ms"] * 1e-3) / 1e12
        roof["fp64"] = {"flop_per_cell": prof["fp64_flop_per_cell"], "achieved_TFLOPs": tf,
                        "peak_TFLOPs": FP64_PEAK_TFLOPS, "frac": tf / FP64_PEAK_TFLOPS}
    res = {
        "metric": METRIC, "value": value, "unit": "cell-updates/s", "n_gpus": 1, "steps": a.steps,
        "warmup": a.warmup, "ms_per_step": ms, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f64", "data": "synthetic",
        "config": {"workload": f"bubble3d {ni}x{nj}x{nk} (3-D Mach-2 shock / H2-air bubble, IMEX, "
                               f"FP64, MUSCL-minmod + AUSM, CFL 0.5)", "cells": wl.ncells,
                   "chemistry": "IMEX", "quirks": "FIXED"},
        "roofline": roof,
        "kernels": kern,
        "finite": rc == 0, "wall_s_timed": wall,
    }
    if not a.no_extra:
        res["batched_chemistry"] = bench_batched_chemistry(g)
        res["surrogate_network"] = bench_surrogate_network()
    if not a.no_cpu:
        res["cpu_baseline"] = cpu_baseline(a.cpu_grid, a.cpu_steps)
    print(json.dumps(res))


if __name__ == "__main__":
    main()
