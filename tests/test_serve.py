"""Serving: admission policy on the CPU, continuous batching on the GPU
(each result identical alone and in the mix)."""
import collections

import numpy as np
import pytest

from ramses_amd.serve import Request, admit


def test_admission_budget_and_slots():
    q = collections.deque(Request(i, n, 0.1) for i, n in enumerate([16, 64, 128, 16, 32]))
    per = {n: 2 * 5 * n ** 3 * 8 for n in (16, 32, 64, 128)}
    # budget fits the first two only: the 128^3 request waits at the head
    out = admit(q, 8, 0, per[16] + per[64] + 1)
    assert [r.n for r in out] == [16, 64] and q[0].n == 128
    # slot limit
    q2 = collections.deque(Request(i, 16, 0.1) for i in range(5))
    assert len(admit(q2, 3, 0, 1 << 40)) == 3
    # nothing fits
    q3 = collections.deque([Request(0, 128, 0.1)])
    assert admit(q3, 4, 0, per[64]) == []


def _ic(seed):
    def f(n):
        from ramses_amd import ics
        return ics.turbulent(n, 2.0, gamma=5.0 / 3.0, rho_sigma=0.3, seed=seed)
    return f


@pytest.mark.gpu
@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_alone_equals_mix(gpu, precision):
    import torch
    import ramses_amd as ra
    from ramses_amd.serve import HydroServer
    ra.configure(gamma=5 / 3, riemann="hllc", slope_type=1, courant_factor=0.5, precision=precision)
    sizes = [16, 32, 64, 16, 48, 32]
    t_end = 0.02
    mixed = [Request(i, n, t_end, arrival=0.002 * i, ic=_ic(1000 + i), keep_result=True)
             for i, n in enumerate(sizes)]
    srv = HydroServer(hbm_budget=1 << 30, max_slots=4)
    srv.run(mixed)
    assert srv.max_active >= 2, "requests did not overlap"
    for r in mixed:
        alone = Request(r.id, r.n, t_end, ic=_ic(1000 + r.id), keep_result=True)
        HydroServer(hbm_budget=1 << 30, max_slots=1).run([alone])
        assert alone.steps == r.steps
        assert torch.equal(alone.result, r.result), (precision, r.id)
    ra.configure()


@pytest.mark.gpu
def test_served_equals_uniform_level(gpu):
    """A served request reproduces the plain resident level step by step
    (dt capped at the end time like the server does)."""
    import torch
    import ramses_amd as ra
    from ramses_amd.serve import HydroServer
    ra.configure(gamma=5 / 3, riemann="hllc", slope_type=1, courant_factor=0.5)
    t_end = 0.01
    u0 = _ic(7)(32)
    r = Request(0, 32, t_end, ic=lambda n: u0, keep_result=True)
    HydroServer(hbm_budget=1 << 30, max_slots=2).run([r])
    lvl = ra.UniformLevel(u0)
    lvl.courant()
    t, steps = 0.0, 0
    while t < t_end:
        if float(lvl.dt[0]) > t_end - t:
            lvl.dt[0] = t_end - t
        t += float(lvl.step())
        steps += 1
    assert steps == r.steps
    assert torch.equal(lvl.u, r.result)
    ra.configure()


@pytest.mark.gpu
def test_budget_respected(gpu):
    import ramses_amd as ra
    from ramses_amd.serve import HydroServer
    ra.configure(gamma=5 / 3, riemann="llf", slope_type=1, courant_factor=0.5, precision="fp32")
    reqs = [Request(i, n, 0.01, arrival=0.0, ic=_ic(i)) for i, n in enumerate([32, 32, 64, 16, 64, 32, 16])]
    per64 = 2 * 5 * 64 ** 3 * 8
    srv = HydroServer(hbm_budget=per64 + 2 * 5 * 32 ** 3 * 8, max_slots=8)
    srv.run(reqs)
    assert srv.max_used <= srv.budget
    assert all(r.finished is not None and r.steps > 0 for r in reqs)
    ra.configure()


@pytest.mark.gpu
def test_graph_ticks_equal_direct(gpu):
    """Ticks replayed from captured hipGraphs give the same bits as ticks
    launched directly (eager work and syncs happen between replays)."""
    import torch
    import ramses_amd as ra
    from ramses_amd.serve import HydroServer
    ra.configure(gamma=5 / 3, riemann="hllc", slope_type=1, courant_factor=0.5, precision="fp32")
    out = []
    for use_graph in (False, True):
        reqs = [Request(i, n, 0.02, arrival=0.0, ic=_ic(50 + i), keep_result=True)
                for i, n in enumerate([32, 16, 32])]
        srv = HydroServer(hbm_budget=1 << 30, max_slots=4, steps_per_tick=2, use_graph=use_graph, graph_after=1)
        srv.run(reqs)
        if use_graph:
            assert srv.graph_replays > 2
        out.append(reqs)
    for a, b in zip(*out):
        assert a.steps == b.steps and torch.equal(a.result, b.result)
    ra.configure()
