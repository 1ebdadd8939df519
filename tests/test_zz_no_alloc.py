"""eu3d_chem_batch_noalloc (no context, no device allocation; scratch in a static device array)
against eu3d_chem_batch on the same cells: bit for bit in rho_s, T, E and the sub-step counts,
and equal statistics, on two calls in a row, the second on fewer cells (the array keeps the first
call's lists and counters), with and without the statistics read-back."""
import numpy as np
import pytest

from tests.common import say

Q_C3 = 1 << 7


def _t():
    import torch
    assert torch.cuda.is_available()
    return torch


@pytest.mark.gpu
@pytest.mark.parametrize("quirks", [0, Q_C3])
def test_noalloc_batch_equals_context_batch(quirks):
    torch = _t()
    from eu3d_amd import api, configs
    from eu3d_amd.api import Context
    x = configs.make_nodes(8, 1e-4)
    g = Context(x, x, x, quirks=quirks)
    dtm = 7e-6                       # up to 4 sub-steps per cell: compaction, persistent queue, tail
    for call, (n, tail, stats) in enumerate([(60000, 1, True), (17000, 512, True), (9000, 1, False)]):
        say("case noalloc quirks", quirks, "call", call, "n", n, "tail", tail, "stats", stats)
        rho, T, E, _ = configs.chem_states(n, seed=31 + call)
        dev = lambda: (torch.tensor(rho.T.copy(), device="cuda"), torch.tensor(T, device="cuda"),
                       torch.tensor(E, device="cuda"), torch.empty(n, dtype=torch.int32, device="cuda"))
        ra, ta, ea, na = dev()
        rb, tb, eb, nb = dev()
        sched = (256, 4, 2, tail)
        g.set_schedule(*sched)
        sa = g.chem_batch(n, ra.data_ptr(), ta.data_ptr(), ea.data_ptr(), dtm, 10000, na.data_ptr())
        torch.cuda.synchronize()
        sb = api.chem_batch_noalloc(n, rb.data_ptr(), tb.data_ptr(), eb.data_ptr(), dtm, 10000, quirks, sched,
                                    torch.cuda.current_stream().cuda_stream, nb.data_ptr(), stats=stats)
        torch.cuda.synchronize()
        say("  context", sa, "noalloc", sb)
        assert sa["active_cells"] > 0
        if stats:
            assert sb == sa
        assert torch.equal(na, nb)
        for a, b in ((ra, rb), (ta, tb), (ea, eb)):
            fa, fb = torch.isfinite(a), torch.isfinite(b)
            assert torch.equal(fa, fb)
            assert torch.equal(a[fa], b[fb])


def test_noalloc_rejects_oversize_batch():
    """n above the static scratch's capacity is refused before any device work (no GPU needed)."""
    from eu3d_amd import api
    with pytest.raises(api.Eu3dError):
        api.chem_batch_noalloc(api.NOALLOC_MAX_CELLS + 1, 1, 1, 1, 1e-8)
    with pytest.raises(api.Eu3dError):
        api.chem_batch_noalloc(16, 1, 1, 1, 0.0)
    with pytest.raises(api.Eu3dError):
        api.chem_batch_noalloc(16, 1, 1, 1, 1e-8, schedule=(100, 4, 2, 0))
