"""Host logic of the serving engine (empm_amd.serving.Scheduler): admission
by HBM budget and capacity, first come first served, segment planning,
grouping of equal segments into batched calls, the checkpoint/segment
bookkeeping of gradient requests, and that every request finishes."""
import numpy as np
import pytest

from empm_amd import serving as sv


def req(n=1000, grid=32, steps=120, grad=False):
    x = np.full((n, 3), 0.5, np.float32)
    return sv.RolloutRequest(x=x, vol=np.ones(n, np.float32), material=0, E=1e4, nu=0.3, n_grid=grid, n_steps=steps,
                             want_grad=grad, target=x if grad else None)


def drive(sch, max_ticks=10_000):
    """Execute plans as the engine would (without device work); return the
    per-tick plans."""
    ticks = []
    for _ in range(max_ticks):
        if sch.idle():
            return ticks
        sch.admit(0.0)
        plan = sch.plan()
        assert plan, "active requests but nothing planned"
        for w in plan:
            for rid in w.rids:
                if w.kind == sv.FORWARD:
                    sch.forward_done(rid, w.length)
                else:
                    sch.backward_done(rid)
        ticks.append(plan)
    raise AssertionError("scheduler did not drain")


def test_request_bytes_grows_with_size_and_tape():
    a = sv.request_bytes(10_000, 32, 100, False, 50)
    assert sv.request_bytes(20_000, 32, 100, False, 50) > a
    g = sv.request_bytes(10_000, 32, 100, True, 50)
    assert g > a
    # one more segment checkpoint per chunk of horizon
    assert sv.request_bytes(10_000, 32, 200, True, 50) - g == 2 * 96 * 10_000


def test_admission_is_fifo_and_bounded_by_budget():
    one = sv.request_bytes(1000, 32, 120, False, 50)
    sch = sv.Scheduler(budget_bytes=int(2.5 * one), max_particles=10**6, max_cand=8, max_grid=64, chunk=50)
    for i in range(4):
        sch.submit(i, req(), 0.0)
    got = sch.admit(0.0)
    assert [s.rid for s in got] == [0, 1]
    assert list(sch.queue) == [2, 3]
    # finishing one frees room for the next in order
    for _ in range(3):
        for w in sch.plan():
            for rid in w.rids:
                sch.forward_done(rid, w.length)
    assert sch.slots[0].phase == sv.DONE and sch.slots[1].phase == sv.DONE
    assert [s.rid for s in sch.admit(0.0)] == [2, 3]


def test_admission_respects_simulation_capacity():
    sch = sv.Scheduler(budget_bytes=10**12, max_particles=10**6, max_cand=3, max_grid=64)
    for i in range(5):
        sch.submit(i, req(), 0.0)
    assert len(sch.admit(0.0)) == 3


def test_requests_that_can_never_fit_are_rejected():
    sch = sv.Scheduler(budget_bytes=10**9, max_particles=5000, max_cand=4, max_grid=64)
    assert sch.submit(0, req(n=6000), 0.0).phase == sv.REJECTED
    assert sch.submit(1, req(grid=128), 0.0).phase == sv.REJECTED
    assert sch.submit(2, req(n=4000), 0.0).phase == sv.QUEUED
    tiny = sv.Scheduler(budget_bytes=1000, max_particles=5000, max_cand=4, max_grid=64)
    assert tiny.submit(0, req(), 0.0).phase == sv.REJECTED


def test_segments_are_grouped_by_length_and_split_by_capacity():
    sch = sv.Scheduler(budget_bytes=10**12, max_particles=2500, max_cand=8, max_grid=64, chunk=50)
    sch.submit(0, req(steps=120), 0.0)
    sch.submit(1, req(steps=30), 0.0)
    sch.submit(2, req(steps=120), 0.0)
    sch.submit(3, req(steps=200), 0.0)
    sch.admit(0.0)
    plan = sch.plan()
    # three requests with a 50-substep segment: 3000 particles > 2500, two calls
    lens = sorted((w.length, len(w.rids)) for w in plan)
    assert lens == [(30, 1), (50, 1), (50, 2)]
    assert all(w.kind == sv.FORWARD for w in plan)


def test_forward_only_requests_finish_after_their_horizon():
    sch = sv.Scheduler(budget_bytes=10**12, max_particles=10**6, max_cand=8, max_grid=64, chunk=50)
    for i, T in enumerate((1, 49, 50, 51, 173)):
        sch.submit(i, req(steps=T), 0.0)
    ticks = drive(sch)
    assert len(ticks) == 4  # ceil(173 / 50)
    assert all(s.phase == sv.DONE and s.k == s.req.n_steps for s in sch.slots.values())
    assert sch.used == 0


def test_gradient_requests_sweep_their_segments_in_reverse():
    sch = sv.Scheduler(budget_bytes=10**12, max_particles=10**6, max_cand=8, max_grid=64, chunk=50)
    sch.submit(0, req(steps=120, grad=True), 0.0)
    seen = []
    while not sch.idle():
        sch.admit(0.0)
        for w in sch.plan():
            for rid in w.rids:
                s = sch.slots[rid]
                if w.kind == sv.FORWARD:
                    seen.append(("f", s.k, s.k + w.length))
                    sch.forward_done(rid, w.length)
                else:
                    seen.append(("b", s.starts[-1], s.seg_end))
                    assert w.length == s.seg_end - s.starts[-1]
                    sch.backward_done(rid)
    assert seen == [("f", 0, 50), ("f", 50, 100), ("f", 100, 120), ("b", 100, 120), ("b", 50, 100), ("b", 0, 50)]
    assert sch.slots[0].phase == sv.DONE and sch.used == 0


def test_mixed_stream_drains_and_later_arrivals_join_running_ticks():
    rng = np.random.Generator(np.random.PCG64(0))
    sch = sv.Scheduler(budget_bytes=10**12, max_particles=20_000, max_cand=6, max_grid=64, chunk=25)
    rid = 0
    for _ in range(4):
        sch.submit(rid, req(n=int(rng.integers(500, 5000)), steps=int(rng.integers(10, 140)),
                            grad=bool(rng.random() < 0.4)), 0.0)
        rid += 1
    joined = False
    for t in range(10_000):
        if t == 2:  # arrivals while the first requests are mid-horizon
            for _ in range(6):
                sch.submit(rid, req(n=int(rng.integers(500, 5000)), steps=int(rng.integers(10, 140)),
                                    grad=bool(rng.random() < 0.4)), 0.0)
                rid += 1
        if sch.idle():
            break
        sch.admit(0.0)
        plan = sch.plan()
        for w in plan:
            assert sum(sch.slots[r].req.n for r in w.rids) <= 20_000 and len(w.rids) <= 6
            ks = {sch.slots[r].k for r in w.rids if w.kind == sv.FORWARD}
            joined |= len(ks) > 1  # one call advances requests at different points of their horizons
            for r in w.rids:
                (sch.forward_done(r, w.length) if w.kind == sv.FORWARD else sch.backward_done(r))
    assert sch.idle() and all(s.phase == sv.DONE for s in sch.slots.values())
    assert joined
