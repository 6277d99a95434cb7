"""Serving with continuous batching (relax_amd.serve): every request's result
is the one it gets alone on the single-model path, whatever else shares the
device.  Members of a class with an afterslip request run the fault-creep
kernels, so results agree to rounding (relative L2 1e-5 on the surface
displacement, 1e-9 on the model time); otherwise they are bitwise equal."""
import numpy as np
import pytest

from relax_amd import serve

pytestmark = pytest.mark.gpu


def _stream(n, seed):
    reqs = serve.request_stream(n, seed=seed, steps=(3, 6))
    # keep the stream light: the 512^3 class is covered by the benchmark
    return [r for r in reqs if r.grid < 2]


def test_served_results_match_single_path():
    reqs = _stream(14, seed=7)
    assert len({r.grid for r in reqs}) == 2
    alone = {r.rid: r for r in serve.run_back_to_back(reqs, realtime=False)}
    mixed = {r.rid: r for r in serve.Server(admit_per_pass=3).run(reqs, realtime=False)}
    assert sorted(alone) == sorted(mixed) == sorted(r.rid for r in reqs)
    for q in reqs:
        a, b = alone[q.rid], mixed[q.rid]
        assert a.status == b.status, q.rid
        if a.status:
            continue
        assert b.nsteps == q.nsteps
        assert abs(a.t - b.t) <= 1e-9 * abs(a.t), q.rid
        for ua, ub in zip(a.surface, b.surface):
            na = np.linalg.norm(ua.astype(np.float64))
            assert np.linalg.norm(ua.astype(np.float64) - ub) <= 1e-5 * max(na, 1e-30), q.rid


def test_server_respects_the_memory_budget():
    """A budget of two small models forces queueing; all requests complete."""
    reqs = [r for r in _stream(10, seed=11) if r.grid == 0][:5]
    fp = serve._footprint(reqs[0].spec)
    srv = serve.Server(budget_bytes=2 * fp + 1)
    peak = []
    admit = srv._admit

    def tracked(r, now):
        m = admit(r, now)
        peak.append(srv.used)
        return m

    srv._admit = tracked
    res = srv.run(reqs, realtime=False)
    assert len(res) == len(reqs)
    assert max(peak) <= 2 * fp
    assert all(r.status == 0 for r in res)


def test_reset_model_matches_a_new_one():
    """A pooled model, reset and set up for another request, steps exactly
    as a freshly created one."""
    from relax_amd import workloads

    reqs = [r for r in _stream(30, seed=5) if r.grid == 1]
    assert len(reqs) >= 2
    a, b = reqs[0], reqs[1]
    fresh = workloads.build(b.spec)
    for _ in range(3):
        fresh.step(info=False)
    ref = fresh.download()
    fresh.close()
    m = workloads.build(a.spec)
    m.step()
    m.reset()
    workloads.configure(m, b.spec)
    for _ in range(3):
        m.step(info=False)
    out = m.download()
    m.close()
    for k in ("u1", "u2", "u3", "sig", "tau", "gamma"):
        np.testing.assert_array_equal(out[k], ref[k], err_msg=k)


def test_c4_stream_with_every_grid_class():
    """The serving stream at its sizes: Poisson arrivals over the three grid
    classes (128x128x64, 256x256x128 and 512x512x256) with 10-30 steps and
    afterslip requests; every served result equals the one it gets alone."""
    reqs = serve.request_stream(16, seed=3, steps=(10, 30))
    classes = {r.grid for r in reqs}
    if 2 not in classes:  # make sure the 512x512x256 class is present
        extra = serve.request_stream(40, seed=4, steps=(10, 30))
        big = next(r for r in extra if r.grid == 2)
        big.rid = len(reqs)
        reqs.append(big)
    assert {r.grid for r in reqs} >= {0, 1, 2}
    alone = {r.rid: r for r in serve.run_back_to_back(reqs, realtime=False)}
    mixed = {r.rid: r for r in serve.Server().run(reqs, realtime=False)}
    assert sorted(alone) == sorted(mixed) == sorted(r.rid for r in reqs)
    for q in reqs:
        a, b = alone[q.rid], mixed[q.rid]
        assert a.status == b.status, q.rid
        if a.status:
            continue
        assert b.nsteps == q.nsteps
        assert abs(a.t - b.t) <= 1e-9 * abs(a.t), q.rid
        for ua, ub in zip(a.surface, b.surface):
            na = np.linalg.norm(ua.astype(np.float64))
            assert np.linalg.norm(ua.astype(np.float64) - ub) <= 1e-5 * max(na, 1e-30), q.rid


def test_served_with_tuned_launch_shapes():
    """After serve.tune() (both autotuners for every grid class) the single
    path and the batches may launch different shapes (the Green variants
    differ in FP32 rounding), so each served result agrees with its result
    alone to FP32 rounding: the adaptive time to 1e-5, the surface
    displacement to relative L2 1e-5."""
    tuned = serve.tune()
    assert {g for g, _ in tuned} == {0, 1, 2}
    reqs = serve.request_stream(10, seed=5, steps=(10, 20))
    alone = {r.rid: r for r in serve.run_back_to_back(reqs, realtime=False)}
    mixed = {r.rid: r for r in serve.Server().run(reqs, realtime=False)}
    assert sorted(alone) == sorted(mixed) == sorted(r.rid for r in reqs)
    for q in reqs:
        a, b = alone[q.rid], mixed[q.rid]
        assert a.status == b.status, q.rid
        if a.status:
            continue
        assert b.nsteps == q.nsteps
        assert abs(a.t - b.t) <= 1e-5 * abs(a.t), q.rid
        for ua, ub in zip(a.surface, b.surface):
            na = np.linalg.norm(ua.astype(np.float64))
            assert np.linalg.norm(ua.astype(np.float64) - ub) <= 1e-5 * max(na, 1e-30), q.rid


def test_mixed_server_results_match_single_path():
    """Server(mixed=True): every request in one batch of uneven grids
    (work-list divergence and strain launches over all members); each
    served result agrees with its result alone.  Launch shapes tuned earlier
    in the process may differ between the paths (the Green variants differ
    in FP32 rounding), so the adaptive time is held to 1e-5 and the surface
    displacement to relative L2 1e-5."""
    reqs = serve.request_stream(12, seed=6, steps=(10, 20))
    assert len({r.grid for r in reqs}) >= 2
    alone = {r.rid: r for r in serve.run_back_to_back(reqs, realtime=False)}
    mixed = {r.rid: r for r in serve.Server(mixed=True).run(reqs, realtime=False)}
    assert sorted(alone) == sorted(mixed) == sorted(r.rid for r in reqs)
    for q in reqs:
        a, b = alone[q.rid], mixed[q.rid]
        assert a.status == b.status, q.rid
        if a.status:
            continue
        assert b.nsteps == q.nsteps
        assert abs(a.t - b.t) <= 1e-5 * abs(a.t), q.rid
        for ua, ub in zip(a.surface, b.surface):
            na = np.linalg.norm(ua.astype(np.float64))
            assert np.linalg.norm(ua.astype(np.float64) - ub) <= 1e-5 * max(na, 1e-30), q.rid
