"""Autotuner for the persistent chemistry kernel's schedule.

The persistent kernel takes cells from a compacted work queue, and its cost per cell varies with
the sub-step count of that cell. The best launch therefore depends on the load: how many cells are
active and how uneven their sub-step counts are. The schedule has four parameters:
- threads per block;
- blocks per CU (with the block size, this sets how many lanes pull from the queue);
- the occupancy target of the register allocation (waves per SIMD: 1 keeps every value in
  registers, 2 and 4 trade spills for latency hiding);
- the tail threshold: cells needing more sub-steps than this run in the tail kernel, two cells
  per wavefront with each sub-step spread over the lanes (0 keeps every cell in the persistent
  kernel). Long chains of dependent sub-steps bound the batch's time, and the tail kernel
  shortens each sub-step's latency at the cost of throughput.

`tune` times the candidates on the caller's own batch with HIP events on the context stream,
restoring the inputs between runs. It first picks the tail threshold at the default launch, then
the launch at that threshold, and returns the fastest. Results are cached in a JSON file, keyed by
the GPU (architecture and CU count), the mechanism, the quirk set and a signature of the load
histogram.
"""
from __future__ import annotations

import hashlib
import json
import math
import os
import time

import numpy as np

from .api import Context
from .mech import load_default

BLOCKS = (64, 128, 256)
BLOCKS_PER_CU = (1, 2, 4, 8, 16)
WAVES = (1, 2, 4)
TAILS = (0, 128, 256, 512, 1024, 2048)
DEFAULT_LAUNCH = (256, 4, 2)
DEFAULT_CACHE = os.path.join(os.path.expanduser("~"), ".cache", "eu3d_amd", "chem_schedule.json")


def _torch():
    import torch
    return torch


def device_key() -> str:
    torch = _torch()
    p = torch.cuda.get_device_properties(0)
    arch = getattr(p, "gcnArchName", "gfx").split(":")[0]
    return f"{arch}-{p.multi_processor_count}cu"


def mech_key() -> str:
    m = load_default()
    h = hashlib.sha1()
    for a in (m.stoi_f, m.stoi_b, m.af, m.bf, m.eaf, m.react_tb, m.coeff0, m.coeff1, m.mw):
        h.update(np.ascontiguousarray(a, dtype=np.float64).tobytes())
    return h.hexdigest()[:12]


def load_signature(nchem: np.ndarray) -> str:
    """Bucketed histogram of the sub-step counts: the active fraction (1 / 16 steps of log2),
    and log2 of the median and of the 99th percentile of the active cells' counts."""
    n = max(1, nchem.size)
    act = nchem[nchem > 1]
    frac = act.size / n
    fb = -1 if frac == 0 else max(0, min(16, int(round(16 + math.log2(frac)))))
    if act.size == 0:
        return f"a{fb}"
    q50 = int(round(math.log2(float(np.median(act)))))
    q99 = int(round(math.log2(float(np.percentile(act, 99)))))
    return f"a{fb}-m{q50}-p{q99}"


def _time_batch(ctx: Context, inputs, dtm, nchem_max, reps):
    torch = _torch()
    r0, t0, e0 = inputs
    n = t0.shape[0]
    s = torch.cuda.ExternalStream(ctx.stream())
    best = float("inf")
    for _ in range(reps):
        r, t, e = r0.clone(), t0.clone(), e0.clone()
        torch.cuda.synchronize()
        a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        a.record(s)
        ctx.chem_batch(n, r.data_ptr(), t.data_ptr(), e.data_ptr(), dtm, nchem_max)
        b.record(s)
        b.synchronize()
        best = min(best, a.elapsed_time(b))
    return best


def tune(ctx: Context, inputs, dtm: float, nchem_max: int = 10000, reps: int = 2, candidates=None, tails=None,
         log=None):
    """Time the schedules on the batch `inputs` = (rho [ns][n], T, E) device tensors (left
    unchanged): the tail thresholds `tails` at the default launch, then every launch in
    `candidates` (block, blocks per CU, waves) at the best threshold. Returns (best schedule dict,
    table of {schedule: ms})."""
    table = {}

    def run(b, k, w, tail):
        ctx.set_schedule(b, k, w, tail)
        ms = _time_batch(ctx, inputs, dtm, nchem_max, reps)
        table[f"{b}x{k}/w{w}/t{tail}"] = ms
        if log:
            log(f"  schedule block {b} blocks/CU {k} waves {w} tail {tail}: {ms:.3f} ms")
        return ms

    b0, k0, w0 = DEFAULT_LAUNCH
    best_tail = min(tails or TAILS, key=lambda t: run(b0, k0, w0, t))
    for b, k, w in candidates or [(b, k, w) for w in WAVES for b in BLOCKS for k in BLOCKS_PER_CU]:
        if f"{b}x{k}/w{w}/t{best_tail}" not in table:
            run(b, k, w, best_tail)
    key = min(table, key=table.get)
    b, rest = key.split("x")
    k, rest = rest.split("/w")
    w, t = rest.split("/t")
    best = {"block": int(b), "blocks_per_cu": int(k), "waves": int(w), "tail": int(t), "ms": table[key]}
    ctx.set_schedule(best["block"], best["blocks_per_cu"], best["waves"], best["tail"])
    return best, table


def cache_key(ctx: Context, nchem: np.ndarray) -> str:
    return f"{device_key()}|{mech_key()}|q{ctx.quirks}|{load_signature(nchem)}|tail"


def load_or_tune(ctx: Context, inputs, dtm: float, nchem_max: int = 10000, cache: str | None = None, log=None):
    """Schedule for this load from the cache, tuning (and caching) on a miss. The load signature
    comes from one prediction pass on a copy of the batch."""
    torch = _torch()
    cache = cache or os.environ.get("EU3D_SCHEDULE_CACHE", DEFAULT_CACHE)
    r0, t0, e0 = inputs
    n = t0.shape[0]
    nch = torch.empty(n, dtype=torch.int32, device="cuda")
    r, t, e = r0.clone(), t0.clone(), e0.clone()
    ctx.chem_batch(n, r.data_ptr(), t.data_ptr(), e.data_ptr(), dtm, nchem_max, nch.data_ptr())
    key = cache_key(ctx, nch.cpu().numpy())
    db = {}
    if os.path.exists(cache):
        with open(cache) as f:
            db = json.load(f)
    if key in db:
        s = db[key]
        ctx.set_schedule(s["block"], s["blocks_per_cu"], s["waves"], s.get("tail", 0))
        return dict(s, cached=True, key=key)
    t_start = time.time()
    best, table = tune(ctx, inputs, dtm, nchem_max, log=log)
    best.update(key=key, tuned_s=time.time() - t_start, candidates=len(table))
    db[key] = {k: best[k] for k in ("block", "blocks_per_cu", "waves", "tail", "ms")}
    os.makedirs(os.path.dirname(cache), exist_ok=True)
    with open(cache, "w") as f:
        json.dump(db, f, indent=1)
    return dict(best, cached=False)
