"""Benchmark: cell-updates/s of the EU3D hot path on one MI355X, plain FP64 and with the MLP
chemistry surrogate, on the 3-D planar H2-air detonation.

Workload (the configuration the metric compares plain and surrogate chemistry on, and the largest
single-GPU configuration): 512 x 256 x 128 cells, dx = 5e-5 m, stoichiometric H2-air premix at
1 atm / 300 K with a seeded 2 % RMS density perturbation of 10-cell correlation length, driver
slab of 16 x-cells at 30 atm / 3000 K, wall at x-, outflow at x+, periodic in y and z; MUSCL-
minmod + AUSM, explicit update, adaptive trapezoid chemistry (FIXED quirks), CFL ramp 0.05 -> 0.5
over the first 20 steps. A step is one full time step: CFL dt, ghost fill, both advection passes,
the fused Update_after_Adv + prediction + first sub-step pass, and the remaining sub-steps of the
cells that need them. Inputs are resident in HBM before the timed region. The initial state is
prepared by the 20 steps of the CFL ramp (untimed, part of the workload's start), then W warm-up
steps at CFL 0.5 let the front develop (untimed), then K steps run in one call.

The line carries:
- value: plain FP64 cell-updates/s over the K timed steps;
- surrogate: the same K steps with the hybrid surrogate chemistry (cells needing more than 4
  FP64 sub-steps go through the BF16-MFMA network, rejects fall back to FP64) run side by side
  from the same initial state, with its error against the plain run (T, p, Y, front position)
  and the fraction of cells through the network and rejected;
- roofline: the dominant kernel of the plain step (k_trap_first), timed live with HIP events on
  the context's stream;
- kernels, zn1d_rk2 (the 1-D detonation tube, 2000 steps, with its CPU rate beside it),
  bubble3d_imex (the 3-D shock / H2-bubble IMEX configuration), batched_chemistry,
  surrogate_network, surrogate_stiff_chemistry (the hybrid against plain FP64 chemistry on the
  stiff batched states, where the network takes cells): further sections;
- cpu_baseline: the CPU restatement of the reference arithmetic (the oracle) on a bounded sample
  of the same workload, on the host's cores.
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
BF16_PEAK_TFLOPS = 2500.0
# Algorithmic bytes per cell-update of the full plain step (FP64), as the hot-path contract
# counts them: one EE stage reads the 13 conserved values and the temperature guess and writes
# both back (224 B); the trapezoid chemistry reads rho_k[9] and T and writes rho_k[9] (152 B).
STEP_BYTES = {"det3d": 224 + 152, "bubble3d": 224}
# Algorithmic bytes per interior cell of each kernel on its own (FP64):
#  k_trap_first (Update_after_Adv + Update_after_CS + prediction + first sub-step): read the 13
#    conserved values, the 13 RHS values and the temperature guess; write 13 conserved values and T
#  k_imex_fused (Diagonalized + Update_IMEX + Update_after_CS): the same reads and writes
#  k_flux_z / k_flux_xy (the two advection passes): read the 16 reconstructed primitives
#    (P, rho, u, v, w, H, gamma, Y_1..9) and write 13 values (the z term, then the RHS, which the
#    x/y pass also reads)
KERNEL_BYTES_PER_CELL = {
    "trap_first": (13 + 13 + 1 + 13 + 1) * 8,
    "imex_fused": (13 + 13 + 1 + 13 + 1) * 8,
    "flux_z": (16 + 13) * 8,
    "flux_xy": (16 + 13 + 13) * 8,
}
KERNEL_NAMES = {"trap_first": "k_trap_first<false, true>", "imex_fused": "k_imex_fused",
                "flux_z": "k_flux_z<{zc}, true>", "flux_xy": "k_flux_xy<32, 8, true, true>"}
PROFILES = {"trap_first": "trap_first_det3d_512x256x128_pmc.json",
            "imex_fused": "imex_fused_bubble3d_384x128x128_pmc.json"}
SURROGATE_THRESHOLD = 4
# Steps of the CFL ramp that start the detonation (configs.cfl_ramp): the input state's preparation,
# run before the W warm-up steps whatever W is.
RAMP_STEPS = 20


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=40)
    ap.add_argument("--warmup", type=int, default=90,
                    help="untimed steps at CFL 0.5 after the 20-step CFL ramp")
    ap.add_argument("--grid", type=str, default="512x256x128")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--no-surrogate", action="store_true", help="skip the surrogate leg")
    ap.add_argument("--no-extra", action="store_true",
                    help="skip the bubble3d, batched-chemistry and surrogate-network sections")
    ap.add_argument("--cpu-grid", type=str, default="512x64x64")
    ap.add_argument("--cpu-steps", type=int, default=8)
    ap.add_argument("--network-only", action="store_true",
                    help="only the surrogate-network section (inference and training at batch 2^20), for profiling")
    ap.add_argument("--bubble3d-only", action="store_true",
                    help="only the bubble3d IMEX section (the configuration its profiles are taken on)")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="only copy 256 MiB of float64 three times (the known-size kernel that scales the "
                         "FETCH_SIZE / WRITE_SIZE counters to bytes) and exit")
    return ap.parse_args()


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def events(stream_ptr):
    import torch
    s = torch.cuda.ExternalStream(stream_ptr)

    def ev():
        e = torch.cuda.Event(enable_timing=True)
        e.record(s)
        return e
    return ev


def cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def host_cores():
    cores = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else os.cpu_count()
    return min(cores, int(os.environ.get("OMP_NUM_THREADS", cores)))


def cpu_baseline(grid, steps):
    """Oracle (CPU restatement, OpenMP over cells) on a bounded sample of the same workload: a
    section of the detonation with the full x extent, one warm-up step, then `steps` steps."""
    from eu3d_amd import configs
    import oracle
    ni, nj, nk = (int(v) for v in grid.split("x"))
    cores = host_cores()
    oracle.set_threads(cores)
    wl = configs.det3d(ni, nj, nk)
    o = oracle.OracleFlow(wl.xnode, wl.ynode, wl.znode, quirks=0)
    for k in ("T", "P", "U", "V", "W", "YI"):
        o.set(k, wl.prims[k])
    o.field_initial()
    kw = dict(bc_kinds=wl.bc_kinds, rk=1, chem=3, nchem_max=10000)
    o.step(1, cfl=configs.cfl_ramp(0), **kw)
    t0 = time.perf_counter()
    for s in range(1, steps + 1):
        o.step(1, cfl=configs.cfl_ramp(s), **kw)
    el = time.perf_counter() - t0
    res = {"value": wl.ncells * steps / el, "unit": "cell-updates/s", "cores": cores, "kind": "port",
           "cpu_model": cpu_model(),
           "sample": f"det3d {ni}x{nj}x{nk} (same generator), {steps} adaptive-trapezoid steps after 1 warm-up "
                     f"step, {el:.1f} s"}
    oracle.set_threads(1)
    w1 = configs.det3d(64, 32, 32)
    o1 = oracle.OracleFlow(w1.xnode, w1.ynode, w1.znode, quirks=0)
    for k in ("T", "P", "U", "V", "W", "YI"):
        o1.set(k, w1.prims[k])
    o1.field_initial()
    o1.step(1, cfl=configs.cfl_ramp(0), **dict(kw, bc_kinds=w1.bc_kinds))
    t0 = time.perf_counter()
    o1.step(1, cfl=configs.cfl_ramp(1), **dict(kw, bc_kinds=w1.bc_kinds))
    el1 = time.perf_counter() - t0
    oracle.set_threads(cores)
    res["single_core"] = {"value": w1.ncells / el1, "unit": "cell-updates/s", "cores": 1,
                          "sample": f"det3d 64x32x32, 1 step after 1 warm-up step, {el1:.2f} s"}
    return res


def cpu_chem_batch(rho, T, E, dtm, grid=(256, 128, 128)):
    """The oracle's adaptive trapezoid (Euler::Trapezoid(1) cell loop) on flat cells at rest, all
    of them: rho [n x 9], T, E laid out in the interior of a grid (the chemistry is cell-local;
    ghost cells hold a copy of the first cell and are not integrated), OpenMP over cells on the
    host cores. Returns the seconds of the chemistry call and its T."""
    import oracle
    from eu3d_amd import configs
    ni, nj, nk = grid
    assert ni * nj * nk == T.shape[0]
    o = oracle.OracleFlow(configs.make_nodes(ni, 1e-4), configs.make_nodes(nj, 1e-4), configs.make_nodes(nk, 1e-4),
                          quirks=0)
    sh = o.shape3
    cs = np.zeros(sh + (13,))
    cs[..., 0:9] = rho[0]
    cs[..., 12] = rho[0].sum() * E[0]
    cs[2:-2, 2:-2, 2:-2, :9] = rho.reshape(ni, nj, nk, 9)
    cs[2:-2, 2:-2, 2:-2, 12] = (rho.sum(1) * E).reshape(ni, nj, nk)
    Tg = np.full(sh, T[0])
    Tg[2:-2, 2:-2, 2:-2] = T.reshape(ni, nj, nk)
    o.set("CS", cs)
    o.set("T", Tg)
    o.update_after_cs()
    o.set_time(dtm)
    t0 = time.perf_counter()
    o.trapezoid_adaptive(10000)
    el = time.perf_counter() - t0
    return el, o.get("T")[2:-2, 2:-2, 2:-2].reshape(-1)


def cpu_batched_chemistry(n=1 << 22, dtm=1e-8, seed=2024):
    """The batched-chemistry workload in full on the CPU restatement: the same 2^22
    equilibrium-line states as the GPU section, one adaptive-trapezoid step of dt = 1e-8 s."""
    from eu3d_amd import chemistry, configs
    from eu3d_amd.api import Context
    import oracle
    x = configs.make_nodes(8, 1e-4)
    ctx = Context(x, x, x, quirks=0)
    (r, t, e), _ = chemistry.states(ctx, n, seed=seed, burnt="equilibrium")
    rho, T, E = r.cpu().numpy().T.copy(), t.cpu().numpy(), e.cpu().numpy()
    del r, t, e
    ctx.close()
    cores = host_cores()
    oracle.set_threads(cores)
    el, _ = cpu_chem_batch(rho, T, E, dtm)
    return {"value": n / el, "unit": "cells/s", "cores": cores, "kind": "port", "cells": n, "s": el,
            "sample": f"all {n} equilibrium-line cells (seed {seed}), one adaptive-trapezoid step of dt = {dtm:g} s"}


def read_profile(kernel):
    """HBM bytes per launch (FETCH_SIZE + WRITE_SIZE, scaled by the copy-kernel calibration) and
    FP64 FLOPs per cell of a kernel from its committed rocprofv3 --pmc summary, or {}."""
    p = os.path.join(ROOT, "profiles", PROFILES.get(kernel, ""))
    if not os.path.isfile(p):
        return {}
    with open(p) as f:
        return json.load(f).get("kernels", {}).get(KERNEL_NAMES[kernel].split("<")[0], {})


def load_state(g, wl):
    """The initial state handed over from host arrays (the reference's Array layout) and the
    conserved fields formed on the device; returns the seconds it took."""
    import torch
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for k in ("T", "P", "U", "V", "W", "YI"):
        g.set(k, wl.prims[k])
    g.field_initial()
    g.sync()
    return time.perf_counter() - t0


def new_context(wl):
    from eu3d_amd.api import Context
    g = Context(wl.xnode, wl.ynode, wl.znode, quirks=0)
    load_state(g, wl)
    return g


def z_chunk(wl):
    """The z planes per column chunk the library's z pass takes on this grid (its default): 64 when
    that still gives every CU four blocks of 64 x 4 columns, else 32."""
    import torch
    ni, nj, nk = (len(a) - 4 for a in (wl.xnode, wl.ynode, wl.znode))
    blocks = -(-ni // 64) * -(-nj // 4) * -(-nk // 64)
    return 64 if blocks >= 4 * torch.cuda.get_device_properties(0).multi_processor_count else 32


def kname(kernel, wl):
    """The kernel instance that runs for `kernel` on this workload's grid."""
    return KERNEL_NAMES[kernel].format(zc=z_chunk(wl)) if kernel == "flux_z" else KERNEL_NAMES[kernel]


def kernel_entry(g, name, ncells, reps=10):
    ms = g.time_kernel(name, reps=reps)
    gbs = KERNEL_BYTES_PER_CELL[name] * ncells / (ms * 1e-3) / 1e9
    return {"avg_launch_ms": ms, "algorithmic_bytes_per_cell": KERNEL_BYTES_PER_CELL[name], "achieved_GBs": gbs,
            "frac_hbm": gbs / HBM_PEAK_GBS}


def roofline(kernel, ms, ncells, step_bytes, step_ms):
    """The dominant kernel against the HBM roof. `achieved` counts the step's algorithmic bytes
    per cell-update (the contract's figure) over that kernel's launch; the kernel's own minimal
    bytes and the whole step's rate are given beside it."""
    prof = read_profile(kernel)
    alg = step_bytes * ncells
    achieved = alg / (ms * 1e-3) / 1e9
    own = KERNEL_BYTES_PER_CELL[kernel] * ncells / (ms * 1e-3) / 1e9
    r = {"kernel": KERNEL_NAMES[kernel], "bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
         "frac": achieved / HBM_PEAK_GBS, "traffic": prof.get("hbm_bytes_per_launch"),
         "algorithmic_bytes_per_launch": alg, "bytes_basis": f"{step_bytes} B per cell-update (full plain step)",
         "avg_launch_ms": ms,
         "kernel_own": {"bytes_per_cell": KERNEL_BYTES_PER_CELL[kernel], "achieved_GBs": own,
                        "frac": own / HBM_PEAK_GBS},
         "step": {"ms": step_ms, "achieved_GBs": alg / (step_ms * 1e-3) / 1e9,
                  "frac": alg / (step_ms * 1e-3) / 1e9 / HBM_PEAK_GBS}}
    if prof.get("fp64_flop_per_cell"):
        tf = prof["fp64_flop_per_cell"] * ncells / (ms * 1e-3) / 1e12
        r["fp64"] = {"flop_per_cell": prof["fp64_flop_per_cell"], "achieved_TFLOPs": tf,
                     "peak_TFLOPs": FP64_PEAK_TFLOPS, "frac": tf / FP64_PEAK_TFLOPS}
    return r


def run_steps(g, wl, start, nsteps, chem, stats=None):
    """nsteps steps one call each (the CFL ramp changes per step), derived fields refreshed after
    the last; with stats, the surrogate statistics of every step are summed into it."""
    from eu3d_amd import configs
    kw = dict(bc_kinds=wl.bc_kinds, rk=1, chem=chem, nchem_max=10000)
    for s in range(start, start + nsteps):
        g.step(1, cfl=configs.cfl_ramp(s), skip_derive=s < start + nsteps - 1, **kw)
        if stats is not None:
            st = g.surrogate_stats()
            for k in ("cells", "network", "fallback"):
                stats[k] += st[k]


def train_surrogate(threshold):
    """Harvest FP64 chemistry increments from a plain run of a 512 x 32 x 32 section of the
    detonation (another seed), fit the normalisation and hull, and train the network."""
    from eu3d_amd import configs, surrogate
    from eu3d_amd.api import MLP
    t0 = time.time()
    H = surrogate.Harvest()
    surrogate.harvest_run(H, configs.det3d(512, 32, 32, seed=777), 150, 5, max_cells=1 << 20)
    r0, r1, T0, dt = H.cat()
    m = MLP(seed=11)
    p = surrogate.fit_params(r0, r1, T0, dt, nchem_threshold=threshold)
    m.set_surrogate(p)
    X, Y = surrogate.make_dataset(m, r0, r1, T0, dt)
    hist = surrogate.train(m, X, Y, epochs=10, batch=16384, lr=2e-3)
    return m, {"pairs": int(X.shape[0]), "epochs": 10, "loss_first": hist[0], "loss_last": hist[-1],
               "harvest_and_train_s": time.time() - t0, "threshold_substeps": threshold,
               "checks": {"dT_max_K": p.dT_max, "elem_tol": p.elem_tol, "y_neg_tol": p.y_neg_tol, "T_cold_K": p.T_cold}}


def front_x(T, xnode, bc=2, T_front=1500.0):
    """Mean over (y, z) of the largest interior x with T > T_front."""
    Ti = T[bc:-bc, bc:-bc, bc:-bc]
    hot = Ti > T_front
    idx = np.where(hot.any(0), Ti.shape[0] - 1 - np.argmax(hot[::-1], axis=0), 0)
    return float(np.mean(xnode[bc + idx]))


def compare_runs(sur, plain, wl):
    Tp, Ts = plain.get("T"), sur.get("T")
    Pp, Ps = plain.get("P"), sur.get("P")
    Yp, Ys = plain.get("YI"), sur.get("YI")
    I = (slice(2, -2),) * 3
    fin = np.isfinite(Tp[I]) & np.isfinite(Ts[I])
    dT = np.abs(Ts[I] - Tp[I])[fin]
    dP = np.abs(Ps[I] - Pp[I])[fin]
    dx = wl.xnode[3] - wl.xnode[2]
    fp, fs = front_x(Tp, wl.xnode), front_x(Ts, wl.xnode)
    return {"T_Linf_rel": float((dT / np.abs(Tp[I][fin])).max()),
            "T_L2_rel": float(np.sqrt((dT ** 2).mean() / (Tp[I][fin] ** 2).mean())),
            "p_Linf_rel": float((dP / np.abs(Pp[I][fin])).max()),
            "p_L2_rel": float(np.sqrt((dP ** 2).mean() / (Pp[I][fin] ** 2).mean())),
            "Y_Linf": float(np.abs(Ys[I] - Yp[I])[fin].max()),
            "front_x_m": {"plain": fp, "surrogate": fs}, "front_error_cells": abs(fs - fp) / dx,
            "nonfinite_cells": {"plain": int((~np.isfinite(Tp[I])).sum()), "surrogate": int((~np.isfinite(Ts[I])).sum())}}


def bench_bubble3d(steps=20, warmup=3):
    """The 3-D shock / H2-bubble configuration, 384 x 128 x 128, IMEX chemistry, CFL 0.5."""
    import torch
    from eu3d_amd import configs
    wl = configs.bubble3d(384, 128, 128)
    g = new_context(wl)
    kw = dict(cfl=0.5, bc_kinds=wl.bc_kinds, rk=1, chem=1)
    g.step(warmup, **kw)
    g.sync()
    ev = events(g.stream())
    torch.cuda.synchronize()
    t0 = ev()
    g.step(steps, **kw)
    t1 = ev()
    g.sync()
    ms = t0.elapsed_time(t1) / steps
    kern = {kname(k, wl): kernel_entry(g, k, wl.ncells) for k in ("imex_fused", "flux_z", "flux_xy")}
    out = {"workload": "bubble3d 384x128x128 (Mach-2 shock / H2-air bubble, IMEX, CFL 0.5)", "cells": wl.ncells,
           "steps": steps, "ms_per_step": ms, "cells_per_s": wl.ncells / (ms * 1e-3), "finite": g.check_finite() == 0,
           "roofline": roofline("imex_fused", kern[KERNEL_NAMES["imex_fused"]]["avg_launch_ms"], wl.ncells,
                                STEP_BYTES["bubble3d"], ms),
           "kernels": kern}
    g.close()
    return out


def bench_zn1d(nsteps=2000, cpu_steps=100, cpu=True):
    """The 1-D H2-air detonation as stated: 2000 cells of 5e-5 m, 40-cell driver at 30 atm / 3000 K,
    SSP-RK2, adaptive trapezoid chemistry, the 20-step CFL ramp to 0.5 (one call each), then the
    remaining steps in one eu3d_step call, timed with HIP events; beside it the CPU restatement
    (OpenMP over cells on the host cores) timed over the first cpu_steps steps after the ramp."""
    import torch
    from eu3d_amd import configs
    from eu3d_amd.api import Context
    wl = configs.zn1d()
    kw = dict(bc_kinds=wl.bc_kinds, rk=2, chem=3, nchem_max=10000)
    g = Context(wl.xnode, wl.ynode, wl.znode, quirks=0)
    load_state(g, wl)
    for s in range(RAMP_STEPS):
        g.step(1, cfl=configs.cfl_ramp(s), skip_derive=True, **kw)
    g.sync()
    ev = events(g.stream())
    torch.cuda.synchronize()
    t0 = ev()
    g.step(nsteps - RAMP_STEPS, cfl=0.5, **kw)
    t1 = ev()
    g.sync()
    ms = t0.elapsed_time(t1) / (nsteps - RAMP_STEPS)
    out = {"workload": "zn1d 2000 cells (1-D H2-air detonation, RK2, adaptive trapezoid, CFL 0.5 after the ramp)",
           "cells": wl.ncells, "steps": nsteps, "timed_steps": f"{RAMP_STEPS + 1}..{nsteps}", "ms_per_step": ms,
           "cells_per_s": wl.ncells / (ms * 1e-3), "finite": g.check_finite() == 0, "time_s": g.get_time()[1]}
    g.close()
    if not cpu:
        return out
    import oracle
    cores = host_cores()
    oracle.set_threads(cores)
    o = oracle.OracleFlow(wl.xnode, wl.ynode, wl.znode, quirks=0)
    for k in ("T", "P", "U", "V", "W", "YI"):
        o.set(k, wl.prims[k])
    o.field_initial()
    for s in range(RAMP_STEPS):
        o.step(1, cfl=configs.cfl_ramp(s), **kw)
    c0 = time.perf_counter()
    o.step(cpu_steps, cfl=0.5, **kw)
    el = time.perf_counter() - c0
    out["cpu_baseline"] = {"value": wl.ncells * cpu_steps / el, "unit": "cell-updates/s", "cores": cores,
                           "kind": "port", "sample": f"steps {RAMP_STEPS + 1}..{RAMP_STEPS + cpu_steps}, {el:.2f} s"}
    return out


def bench_batched_chemistry(n=1 << 22, dtm=1e-8, seed=2024):
    """Batched FP64 chemistry: 2^22 seeded cells, one adaptive trapezoid step of dt = 1e-8 s,
    with the persistent kernel's schedule from the autotuner, for both burnt compositions."""
    import torch
    from eu3d_amd import autotune, chemistry, configs
    from eu3d_amd.api import Context
    x = configs.make_nodes(8, 1e-4)
    ctx = Context(x, x, x, quirks=0)
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
    ctx.close()
    return out


def bench_surrogate_network(n=1 << 20, reps=10):
    """Surrogate network throughput at batch 2^20: inference (the fused network kernel) and one
    training step (forward, backward, Adam), HIP events on the network's stream. FLOPs per cell /
    sample: 403,968 (inference), 1,205,760 (training)."""
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


def bench_surrogate_stiff():
    """The surrogate where the chemistry dominates: 2^21 equilibrium-line cells, one step of
    dt = 1e-8 s, plain FP64 chemistry against the hybrid (cells needing 9..512 FP64 sub-steps through
    the network trained first on other seeds), for the whole batch and for the same states without
    the igniting cells (more than 512 sub-steps), each leg at its best tail threshold; errors against
    the FP64 result (eu3d_amd.surrogate_eval.stiff_gain)."""
    from eu3d_amd.surrogate_eval import stiff_gain
    return stiff_gain(log=lambda s: None)


def main():
    a = parse()
    if a.gpus != 1:
        print(json.dumps({"error": "one GPU only: runs on several GPUs are not part of this implementation",
                          "n_gpus": a.gpus}))
        sys.exit(2)
    import torch
    from eu3d_amd import configs
    assert torch.cuda.is_available(), "bench.py needs the GPU"
    if a.copy_calibration:
        x = torch.ones(256 * 1024 * 1024 // 8, dtype=torch.float64, device="cuda")
        y = torch.empty_like(x)
        for _ in range(3):
            y.copy_(x)
        torch.cuda.synchronize()
        print(json.dumps({"copied_bytes_per_launch": x.numel() * 8, "launches": 3}))
        return
    if a.network_only:
        print(json.dumps({"surrogate_network": bench_surrogate_network()}))
        return
    if a.bubble3d_only:
        print(json.dumps({"bubble3d_imex": bench_bubble3d()}))
        return
    if a.warmup < 0 or a.steps < 1:
        raise SystemExit("--warmup must be >= 0 and --steps >= 1")
    prep = RAMP_STEPS + a.warmup
    ni, nj, nk = (int(v) for v in a.grid.split("x"))
    wl = configs.det3d(ni, nj, nk)
    kw = dict(bc_kinds=wl.bc_kinds, rk=1, nchem_max=10000)

    # plain FP64: W untimed steps (CFL ramp, front development), then K steps in one call
    log(f"plain: {wl.ncells} cells, {RAMP_STEPS} ramp + {a.warmup} warm-up steps")
    from eu3d_amd.api import Context
    g = Context(wl.xnode, wl.ynode, wl.znode, quirks=0)
    t_in = load_state(g, wl)
    run_steps(g, wl, 0, prep, chem=3)
    g.sync()
    torch.cuda.synchronize()
    ev = events(g.stream())
    w0 = time.perf_counter()
    t0 = ev()
    g.step(a.steps, cfl=0.5, chem=3, **kw)
    t1 = ev()
    g.sync()
    torch.cuda.synchronize()
    wall = time.perf_counter() - w0
    ms = t0.elapsed_time(t1) / a.steps
    value = wl.ncells / (ms * 1e-3)
    finite = g.check_finite() == 0
    log(f"plain: {ms:.3f} ms/step")
    # the state read back to host arrays as a caller of the boundary would (conserved fields and T)
    o0 = time.perf_counter()
    for k in ("CS", "T"):
        g.get(k)
    t_out = time.perf_counter() - o0
    res = {
        "metric": METRIC, "value": value, "unit": "cell-updates/s", "n_gpus": 1, "steps": a.steps,
        "warmup": a.warmup, "ms_per_step": ms, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f64", "data": "synthetic",
        "config": {"workload": f"det3d {ni}x{nj}x{nk} (3-D planar H2-air detonation, dx 5e-5 m, FP64 MUSCL-minmod "
                               f"+ AUSM, adaptive trapezoid chemistry, CFL 0.5 after a 20-step ramp)",
                   "cells": wl.ncells, "chemistry": "adaptive trapezoid (plain FP64)", "quirks": "FIXED",
                   "timed_steps": f"{prep + 1}..{prep + a.steps} (after {RAMP_STEPS} CFL-ramp steps and "
                                  f"{a.warmup} warm-up steps)"},
        "finite": finite, "wall_s_timed": wall,
        "host_handover": {
            "set_initial_state_s": t_in, "get_state_s": t_out,
            "value_including_one_handover_each_way": wl.ncells * a.steps / (a.steps * ms * 1e-3 + t_in + t_out),
            "note": "the timed steps run on state resident in HBM; this rate adds one hand-over of the initial "
                    "state from host arrays (T, P, U, V, W, Y in the reference's Array layout, then FieldInitial "
                    "on the device) and one read-back of CS and T to the K timed steps"},
    }

    if not a.no_surrogate:
        log("surrogate: harvest and training")
        m, tr = train_surrogate(SURROGATE_THRESHOLD)
        gs = new_context(wl)
        gs.set_surrogate(m)
        stats = {"cells": 0, "network": 0, "fallback": 0}
        run_steps(gs, wl, 0, prep, chem=4, stats=stats)
        gs.sync()
        torch.cuda.synchronize()
        evs = events(gs.stream())
        s0 = evs()
        gs.step(a.steps, cfl=0.5, chem=4, **kw)
        s1 = evs()
        gs.sync()
        torch.cuda.synchronize()
        sms = s0.elapsed_time(s1) / a.steps
        last = gs.surrogate_stats()
        log(f"surrogate: {sms:.3f} ms/step")
        res["surrogate"] = {
            "value": wl.ncells / (sms * 1e-3), "unit": "cell-updates/s", "ms_per_step": sms,
            "speedup_vs_plain": ms / sms, "dtype": "bf16 network (fp32 accumulation), fp64 elsewhere",
            "errors_vs_plain_after_all_steps": compare_runs(gs, g, wl),
            "network_fraction_of_cell_steps": stats["network"] / max(1, stats["cells"]),
            "rejected_fraction_of_network_cells": stats["fallback"] / max(1, stats["network"]),
            "stats_basis": f"steps 1..{prep} (one call each); last timed step: {last['network']} network "
                           f"cells, {last['fallback']} rejected",
            "training": tr}
        gs.close()
        del m
    # kernel timings last: each timed launch leaves its outputs in the plain context
    kern = {kname(k, wl): kernel_entry(g, k, wl.ncells) for k in ("trap_first", "flux_z", "flux_xy")}
    res["roofline"] = roofline("trap_first", kern[KERNEL_NAMES["trap_first"]]["avg_launch_ms"], wl.ncells,
                               STEP_BYTES["det3d"], ms)
    res["kernels"] = kern
    g.close()

    if not a.no_extra:
        log("zn1d")
        res["zn1d_rk2"] = bench_zn1d(cpu=not a.no_cpu)
        log("bubble3d")
        res["bubble3d_imex"] = bench_bubble3d()
        log("batched chemistry")
        res["batched_chemistry"] = bench_batched_chemistry()
        log("surrogate network")
        res["surrogate_network"] = bench_surrogate_network()
        log("surrogate on stiff chemistry")
        res["surrogate_stiff_chemistry"] = bench_surrogate_stiff()
    if not a.no_cpu:
        log("cpu baseline")
        res["cpu_baseline"] = cpu_baseline(a.cpu_grid, a.cpu_steps)
        if not a.no_extra:
            log("cpu baseline: batched chemistry")
            res["cpu_baseline"]["batched_chemistry_equilibrium"] = cpu_batched_chemistry()
    print(json.dumps(res))


if __name__ == "__main__":
    main()
