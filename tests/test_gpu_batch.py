"""Batches: many models of one grid stepped by shared launches and a captured
hipGraph (relax_batch_* of relax_hip.h).  Each member's result must be the
one it gets when stepped alone -- bitwise, since the launches run the same
per-model work -- whatever the other members are, and across membership
changes and graph replays with eager work in between."""
import numpy as np
import pytest

from relax_amd import RelaxError, workloads
from relax_amd.model import Batch

from _cases import small_spec

pytestmark = pytest.mark.gpu

FIELDS = ("u1", "u2", "u3", "tau", "sig", "gamma")


def _variant(kind, seed, sx3=64):
    s = small_spec(kind, sx3=sx3, seed=seed)
    if kind == "maxwell":
        # a different viscosity per seed
        s["linear"] = [(6.0, 0.2 * (1 + seed), 1.0, 0.0, 0.6, 0.0)]
    return s


def _alone(s, nsteps):
    m = workloads.build(s)
    recs = [m.step() for _ in range(nsteps)]
    out = m.download()
    pat = m.get_patches() if s["planes"] else None
    m.close()
    return recs, out, pat


def _same(a, b):
    for k in FIELDS:
        np.testing.assert_array_equal(a[k], b[k], err_msg=k)


def test_batch_members_match_their_single_runs_bitwise():
    specs = [_variant("maxwell", 0), _variant("powerlaw", 1), _variant("maxwell", 2)]
    nsteps = 4
    ref = [_alone(s, nsteps) for s in specs]
    models = [workloads.build(s) for s in specs]
    with Batch(models) as b:
        assert len(b) == 3
        for k in range(nsteps):
            recs = b.step()
            for i in range(3):
                assert recs[i] == ref[i][0][k], (i, k)
            # eager work on one member between replays of the captured step
            models[1].download()
    for i, m in enumerate(models):
        _same(m.download(), ref[i][1])
        m.close()


@pytest.mark.parametrize("tuning", [dict(fft_cols=1, green=3, bodyforce=3), dict(fft_cols=2, green=1, bodyforce=2)])
def test_batch_with_tuned_launch_shapes(tuning):
    """A batch launches every kernel with its first member's launch shapes
    (the autotuner keeps one choice per grid): with the register Green pass,
    other y-transform tiles and the deeper divergence tiles the members still
    equal their single runs with the same shapes, bitwise."""
    specs = [_variant("powerlaw", 3, sx3=128), _variant("maxwell", 4, sx3=128)]
    nsteps = 3
    ref = []
    for s in specs:
        m = workloads.build(s)
        m.set_tuning(**tuning)
        recs = [m.step() for _ in range(nsteps)]
        ref.append((recs, m.download()))
        m.close()
    models = [workloads.build(s) for s in specs]
    for m in models:
        m.set_tuning(**tuning)
    with Batch(models) as b:
        for k in range(nsteps):
            recs = b.step()
            for i in range(2):
                assert recs[i] == ref[i][0][k], (i, k)
    for i, m in enumerate(models):
        _same(m.download(), ref[i][1])
        m.close()


def test_batch_with_fault_creep_member():
    """An afterslip member switches the batch to the fault-creep kernels; the
    viscous members then run those kernels too, so they agree to rounding."""
    specs = [_variant("afterslip", 0), _variant("maxwell", 1)]
    nsteps = 3
    ref = [_alone(s, nsteps) for s in specs]
    models = [workloads.build(s) for s in specs]
    with Batch(models) as b:
        for k in range(nsteps):
            recs = b.step()
            for i in range(2):
                assert abs(recs[i]["t"] - ref[i][0][k]["t"]) <= 1e-12 * abs(ref[i][0][k]["t"])
                assert abs(recs[i]["power"] / ref[i][0][k]["power"] - 1) < 1e-5
    for i, m in enumerate(models):
        out = m.download()
        for key in FIELDS:
            a, r = out[key].astype(np.float64), ref[i][1][key].astype(np.float64)
            nr = np.linalg.norm(r)
            assert np.linalg.norm(a - r) <= 1e-6 * (nr if nr > 0 else 1.0), (i, key)
        if i == 0:
            pat = m.get_patches()
            np.testing.assert_allclose(pat["slip"], ref[0][2]["slip"], rtol=1e-5, atol=1e-12)
        m.close()


def test_batch_membership_changes():
    """A member leaves after two steps and continues alone; a new one joins
    for the last steps.  Each trajectory equals its single run bitwise."""
    sa, sb, sc = _variant("maxwell", 0), _variant("powerlaw", 1), _variant("maxwell", 3)
    ref_a = _alone(sa, 4)
    ref_b = _alone(sb, 4)
    ref_c = _alone(sc, 2)
    a, bm, c = workloads.build(sa), workloads.build(sb), workloads.build(sc)
    with Batch([a, bm]) as b:
        b.step(info=False)
        b.step(info=False)
        b.remove(bm)
        b.add(c)
        bm.step()
        bm.step()
        b.step(info=False)
        b.step(info=False)
        b.synchronize()
        _same(a.download(), ref_a[1])
        _same(bm.download(), ref_b[1])
        _same(c.download(), ref_c[1])
    for m in (a, bm, c):
        m.close()


def test_batch_refuses_member_steps_and_second_batches():
    a = workloads.build(_variant("maxwell", 0))
    other = workloads.build(_variant("maxwell", 0, sx3=32))
    with Batch([a]) as b:
        b.add(other)  # another grid: the batch becomes a mixed one
        with pytest.raises(RelaxError):
            Batch([other])  # a model is in one batch at a time
        with pytest.raises(RelaxError):
            a.step()
        with pytest.raises(RelaxError):
            b.add(a)
        with pytest.raises(RelaxError):  # events are applied to models outside a batch
            a.event(**workloads.event_sources(_variant("maxwell", 0), 1))
    a.step()  # free again once the batch is gone
    a.close()
    other.close()


def test_batch_with_transient_member():
    """A member with transient creep and weak zones switches the batch to the
    extended rheology kernels; each member agrees with its single run to
    rounding, and the transient strain follows."""
    specs = [small_spec("transient", seed=0), _variant("maxwell", 1)]
    nsteps = 3
    ref = []
    for s in specs:
        m = workloads.build(s)
        recs = [m.step() for _ in range(nsteps)]
        out = m.download()
        if s.get("ltransient"):
            out["epsilonik"] = m.transient_strain()
        m.close()
        ref.append((recs, out))
    models = [workloads.build(s) for s in specs]
    with Batch(models) as b:
        for k in range(nsteps):
            recs = b.step()
            for i in range(2):
                assert abs(recs[i]["Dt"] / ref[i][0][k]["Dt"] - 1) < 1e-6
    for i, m in enumerate(models):
        out = m.download()
        if i == 0:
            out["epsilonik"] = m.transient_strain()
        for key in FIELDS + (("epsilonik",) if i == 0 else ()):
            a, r = out[key].astype(np.float64), ref[i][1][key].astype(np.float64)
            nr = np.linalg.norm(r)
            assert np.linalg.norm(a - r) <= 1e-6 * (nr if nr > 0 else 1.0), (i, key)
        m.close()


def test_c3_batch_of_64_models_of_128_cubed():
    """The batched workload at its size: 64 models of 128^3 (c3: 1-10
    patches, log-uniform viscosity, n in {1, 2, 3, 3.5}, 15-40 km lid), 10
    steps as one batch.  Every member equals its single run bitwise, and one
    member's 10-step trajectory follows the oracle (relative L2 1e-4)."""
    import oracle as O
    from _cases import oracle_initial_state, oracle_model, rel_l2

    nsteps = 10
    specs = [workloads.spec("c3%d" % i) for i in range(64)]
    models = [workloads.build(s) for s in specs]
    with Batch(models) as b:
        for _ in range(nsteps):
            b.step(info=False)
        recs = b.scalars()
    outs = [m.download() for m in models]
    for m in models:
        m.close()
    for i in range(0, 64, 9):  # a spread of members against their single runs
        r, out, _ = _alone(specs[i], nsteps)
        assert recs[i] == r[-1], i
        _same(outs[i], out)
    # one member against the oracle over the 10 steps
    i = 7
    s = specs[i]
    print("c3 member", i, "n", (s["linear"] or s["nonlinear"])[0][2], flush=True)
    st = oracle_initial_state(s)
    info = O.rk2_run(oracle_model(s), st, nsteps)
    # the adaptive step sequence (the relaxation time is an FP32 minimum)
    assert abs(recs[i]["t"] - info[-1, 7]) <= 1e-5 * abs(info[-1, 7])
    n1 = s["sx1"]
    for k in ("u1", "u2", "u3"):
        assert rel_l2(outs[i][k][..., :n1], st[k][..., :n1]) < 1e-4, k
    for k in ("sig", "tau"):
        assert rel_l2(outs[i][k], st[k]) < 1e-4, k


def test_batch_autotune_keeps_state_and_members_match_their_single_runs():
    """relax_batch_autotune times each variant as one launch over all members:
    the members' state does not change; stepped with the chosen shapes each
    member equals its single run with the same shapes, bitwise; and a later
    batch of the same grid and rheology class starts from the choice."""
    specs = [_variant("powerlaw", 5, sx3=256), _variant("powerlaw", 6, sx3=256), _variant("powerlaw", 7, sx3=256)]
    nsteps = 3
    models = [workloads.build(s) for s in specs]
    before = [m.download() for m in models]
    with Batch(models) as b:
        choice = b.autotune(times=True)
        print("batch autotune", choice, flush=True)
        assert 0 <= choice["fft_cols"] < 3 and 0 <= choice["green"] < 4 and 0 <= choice["bodyforce"] < 6
        for m, ref in zip(models, before):
            _same(m.download(), ref)
        tuned = [b.step() for _ in range(nsteps)]
    out = [m.download() for m in models]
    for m in models:
        m.close()
    shapes = {k: choice[k] for k in ("fft_cols", "green", "bodyforce")}
    for i, s in enumerate(specs):
        m = workloads.build(s)
        m.set_tuning(**shapes)
        recs = [m.step() for _ in range(nsteps)]
        for k in range(nsteps):
            assert recs[k] == tuned[k][i], (i, k)
        _same(m.download(), out[i])
        m.close()
    # a new batch of fresh models (default single-model shapes) uses the choice
    models = [workloads.build(s) for s in specs]
    with Batch(models) as b:
        for _ in range(nsteps):
            b.step(info=False)
    for i, m in enumerate(models):
        _same(m.download(), out[i])
        m.close()


def _mixed_specs():
    """Four grid classes: 64^3 (two members), 128x128x64, 64x64x128 and
    64x64x32."""
    return [_variant("maxwell", 0), small_spec("powerlaw", sx=128, sx3=64, seed=1), _variant("powerlaw", 2, sx3=128),
            _variant("maxwell", 3), _variant("powerlaw", 4, sx3=32)]


def test_mixed_grid_batch_matches_single_runs_bitwise():
    """Models of different grids in one batch: the divergence and strain
    passes are one work-list launch over all members, the transforms and
    the Green pass one launch per grid class.  Every member equals its
    single run bitwise, over graph replays with eager work in between."""
    specs = _mixed_specs()
    nsteps = 4
    ref = [_alone(s, nsteps) for s in specs]
    models = [workloads.build(s) for s in specs]
    with Batch(models) as b:
        for k in range(nsteps):
            recs = b.step()
            for i in range(len(specs)):
                assert recs[i] == ref[i][0][k], (i, k)
            models[2].download()
    for i, m in enumerate(models):
        _same(m.download(), ref[i][1])
        m.close()


def test_mixed_grid_batch_membership_changes():
    """A mixed batch loses the members of all but one class (a uniform batch
    again), then gains another grid's model; each trajectory equals its
    single run bitwise."""
    sa, sb, sc = _variant("maxwell", 0), small_spec("powerlaw", sx=128, sx3=64, seed=1), _variant("powerlaw", 2,
                                                                                                    sx3=128)
    ref_a, ref_b, ref_c = _alone(sa, 6), _alone(sb, 2), _alone(sc, 2)
    a, bm, c = workloads.build(sa), workloads.build(sb), workloads.build(sc)
    with Batch([a, bm]) as b:
        for _ in range(2):
            b.step(info=False)
        b.remove(bm)
        for _ in range(2):
            b.step(info=False)
        b.add(c)
        for _ in range(2):
            b.step(info=False)
        b.synchronize()
        _same(a.download(), ref_a[1])
        _same(bm.download(), ref_b[1])
        _same(c.download(), ref_c[1])
    for m in (a, bm, c):
        m.close()


def test_mixed_grid_batch_with_fault_creep_member():
    """An afterslip member of one grid switches every member of the mixed
    batch to the fault-creep kernels; the others then agree with their
    single runs to rounding."""
    specs = [_variant("afterslip", 0), _variant("maxwell", 1, sx3=128), small_spec("powerlaw", sx=128, sx3=64, seed=2)]
    nsteps = 3
    ref = [_alone(s, nsteps) for s in specs]
    models = [workloads.build(s) for s in specs]
    with Batch(models) as b:
        for k in range(nsteps):
            recs = b.step()
            for i in range(len(specs)):
                assert abs(recs[i]["t"] - ref[i][0][k]["t"]) <= 1e-12 * abs(ref[i][0][k]["t"])
    for i, m in enumerate(models):
        out = m.download()
        for key in FIELDS:
            a, r = out[key].astype(np.float64), ref[i][1][key].astype(np.float64)
            nr = np.linalg.norm(r)
            assert np.linalg.norm(a - r) <= 1e-6 * (nr if nr > 0 else 1.0), (i, key)
        m.close()
    models = [workloads.build(s) for s in specs[1:]]
    with Batch(models) as b:
        with pytest.raises(RelaxError):
            b.autotune()  # tune each grid class as a batch of its own
    for m in models:
        m.close()


def test_mixed_grid_batch_with_transient_member():
    """A transient-creep member with weak zones (extended rheology kernels)
    in a batch with models of two other grids: every member runs the
    extended kernels through the work lists and agrees with its single run
    to rounding; the transient strain follows."""
    specs = [small_spec("transient", seed=1), _variant("maxwell", 2, sx3=128),
             small_spec("powerlaw", sx=128, sx3=64, seed=3)]
    nsteps = 3
    ref = []
    for s in specs:
        m = workloads.build(s)
        recs = [m.step() for _ in range(nsteps)]
        out = m.download()
        if s.get("ltransient"):
            out["epsilonik"] = m.transient_strain()
        m.close()
        ref.append((recs, out))
    models = [workloads.build(s) for s in specs]
    with Batch(models) as b:
        for k in range(nsteps):
            recs = b.step()
            for i in range(len(specs)):
                assert abs(recs[i]["Dt"] / ref[i][0][k]["Dt"] - 1) < 1e-6, (i, k)
    for i, m in enumerate(models):
        out = m.download()
        if i == 0:
            out["epsilonik"] = m.transient_strain()
        for key in FIELDS + (("epsilonik",) if i == 0 else ()):
            a, r = out[key].astype(np.float64), ref[i][1][key].astype(np.float64)
            nr = np.linalg.norm(r)
            assert np.linalg.norm(a - r) <= 1e-6 * (nr if nr > 0 else 1.0), (i, key)
        m.close()
