"""Parity at the full sizes of the two 3-D workloads, against the oracle run from the same state:
one step of the planar detonation (512 x 256 x 128, adaptive trapezoid) and two steps of the
shock-bubble (384 x 128 x 128, IMEX), both FIXED, MUSCL_2 minmod, CFL 0.5.

The small-grid parity tests pin the arithmetic; these check that nothing in the kernels' indexing,
launch shapes or buffer sizes changes with the grid: the x/y tiles and z chunks at their full
counts, 64-bit field offsets of 13-component fields, the CFL partials of 65,536 blocks and the
prediction pass over 16.8 M cells. (In this state no detonation cell needs a second sub-step;
the queue kernels' sub-stepping is compared on the equilibrium-line batch.)

The state is a developed one: the GPU runs the start of the workload (the detonation's 20-step
CFL ramp and 90 more steps, the state the benchmark's timed window starts from; ten bubble
steps), then hands every field a step reads to the oracle (conserved state, T, E, primitives,
derived fields, dt and time). Both then step with the same parameters. CS, T, D,
P, E and Y_k are compared over the interior at 1e-10 (normwise per component), the sub-step counts
of the adaptive trapezoid exactly (Euler::Computing, src/Euler.cpp:266-333; the chemistry of
src/Reaction.cpp:489-670 and :736-811)."""
import os

import pytest

from tests.common import compare, init_state, say

pytestmark = pytest.mark.gpu

# fields a step writes before it reads them; everything else is handed over
OUTPUTS = {"RHS", "F", "G", "Q", "CMS", "MD", "NCHEM"}


def _hand_over(g, o):
    from eu3d_amd.api import FIELDS
    for name in FIELDS:
        if name in OUTPUTS:
            continue
        o.set(name, g.get(name))
    dt, t = g.get_time()
    o.set_time(dt, t)


def _oracle_like(wl):
    import oracle
    cores = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else os.cpu_count()
    oracle.set_threads(max(1, min(16, cores, int(os.environ.get("OMP_NUM_THREADS", cores)))))
    return oracle.OracleFlow(wl.xnode, wl.ynode, wl.znode, quirks=0)


def test_det3d_full_size_step_vs_oracle():
    import torch
    from eu3d_amd import configs
    from eu3d_amd.api import Context
    assert torch.cuda.is_available()
    say("case det3d 512x256x128: 110 steps on the GPU, then one step on both")
    wl = configs.det3d(512, 256, 128)
    g = Context(wl.xnode, wl.ynode, wl.znode, quirks=0)
    init_state(g, wl)
    kw = dict(bc_kinds=wl.bc_kinds, rk=1, chem=3, nchem_max=10000)
    for s in range(110):
        g.step(1, cfl=configs.cfl_ramp(s), skip_derive=s < 109, **kw)
    assert g.check_finite() == 0
    o = _oracle_like(wl)
    _hand_over(g, o)
    say("  state handed over; stepping")
    g.step(1, cfl=0.5, **kw)
    o.step(1, cfl=0.5, **kw)
    say("  comparing")
    assert g.get_time() == pytest.approx(o.get_time(), rel=1e-13)
    nch_g, nch_o = g.get("NCHEM")[2:-2, 2:-2, 2:-2], o.get("NCHEM")[2:-2, 2:-2, 2:-2]
    assert (nch_g == nch_o).all(), f"{int((nch_g != nch_o).sum())} cells differ in Nchem"
    say(f"  Nchem equal; {int((nch_o > 1).sum())} cells sub-stepped, largest {int(nch_o.max())}")
    worst = compare(g, o, ["CS", "T", "D", "P", "E", "YI"])
    say("  worst relative differences:", {k: f"{v:.2e}" for k, v in worst.items()})
    g.close()


def test_bubble3d_full_size_steps_vs_oracle():
    import torch
    from eu3d_amd import configs
    from eu3d_amd.api import Context
    assert torch.cuda.is_available()
    say("case bubble3d 384x128x128: 10 steps on the GPU, then two steps on both")
    wl = configs.bubble3d(384, 128, 128)
    g = Context(wl.xnode, wl.ynode, wl.znode, quirks=0)
    init_state(g, wl)
    kw = dict(cfl=0.5, bc_kinds=wl.bc_kinds, rk=1, chem=1)
    g.step(10, **kw)
    assert g.check_finite() == 0
    o = _oracle_like(wl)
    _hand_over(g, o)
    say("  state handed over; stepping")
    g.step(2, **kw)
    o.step(2, **kw)
    say("  comparing")
    assert g.get_time() == pytest.approx(o.get_time(), rel=1e-13)
    worst = compare(g, o, ["CS", "T", "D", "P", "E", "YI"])
    say("  worst relative differences:", {k: f"{v:.2e}" for k, v in worst.items()})
    g.close()
