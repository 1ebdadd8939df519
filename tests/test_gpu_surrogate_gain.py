"""What the trained chemistry network gains over the FP64 chemistry alone.

Workload: the batched-chemistry states on the H2-air equilibrium lines, one step of dt = 1e-8 s
(src/Euler.cpp:526-556 per cell, src/Reaction.cpp:489-670). The network (log-increment outputs,
abundance-weighted element projection) is trained on the solver's own FP64 increments of other
seeds, on the cells that need more than 4 FP64 sub-steps. In the hybrid, the cells needing 9..512
sub-steps go through the network; the rest, and the rejects of the checks, take the FP64 path.

Each run takes the best of four tail thresholds of the chemistry schedule (512, 256, 128, 64: the
sub-step count above which a cell goes to the 32-lanes-a-cell tail kernel), chosen per leg and per
case, so that neither side is timed at a schedule tuned for the other.

Two measurements:
- the whole batch: its time is set by the few cells that ignite within the step (up to the
  10000-sub-step cap), which stay on FP64 in both runs, so the network removes FP64 work
  without shortening that critical path (and its own stage adds to it);
- the same states without those igniting cells (at most 512 FP64 sub-steps): the FP64 time is the
  sub-step throughput of the stiff cells, which the network replaces.
The error is measured against the FP64 result cell by cell. The measurement itself is
eu3d_amd.surrogate_eval.stiff_gain, which bench.py also reports."""
import pytest

from tests.common import say

pytestmark = pytest.mark.gpu


def test_network_gain_on_stiff_chemistry():
    import torch
    assert torch.cuda.is_available()
    from eu3d_amd.surrogate_eval import stiff_gain
    say("case surrogate gain: harvest, train, time both legs")
    res = stiff_gain(log=say)
    tr = res["training"]
    assert tr["loss_last"] < 0.4 * tr["loss_first"]
    for name in ("whole_batch", "band"):
        r = res[name]
        # every cell the network did not answer is the FP64 result
        assert r["cells_changed"] <= r["network"] - r["rejected"], (name, r)
    whole, band = res["whole_batch"], res["band"]
    # measured on one MI355X: band 4.49 ms (FP64, tail 256) -> 2.72 ms (hybrid, tail 128), 1.65x;
    # the hybrid's own critical path is then the chain of the cells the |dT| check rejects, redone
    # on FP64 in the tail kernel. Whole batch 30.8 -> 29.6 ms (1.04x): the longest FP64 chains (up
    # to the 10000-sub-step cap) stay on FP64 in both runs and set the time.
    # T error of the network's cells: median 2.7 K, p99 24 K, largest 456 K (16 %)
    assert band["speedup"] > 1.3
    assert whole["speedup"] > 0.9
    assert band["network"] > 0.03 * band["cells"]
    err = band["T_err_network_cells_K"]
    assert err["p50"] < 10.0 and err["p99"] < 60.0 and err["max_rel"] < 0.35
