"""The benchmark workloads at their full sizes -- c1, 512x512x256 with a
power-law mantle under 40 slip patches, and c2, 256x256x128 with afterslip on
a friction plane and a power-law mantle -- against the oracle.

* The coseismic set-up (relax.f90:316-371) on the device against the
  oracle's.
* The whole trajectory (20 steps of c1, 30 of c2) run by the device with the
  launch shapes the autotuner picks for the benchmark, in chunks: after each
  chunk the oracle, which keeps its own state from the set-up on, replays the
  device's Dt sequence.  The totals and each chunk's increments (u, sig, tau,
  gamma after the chunk minus before it) are compared, and the relaxation
  time, power and C:E^i of every step.
* Every launch-shape variant of the tunable kernels (Green column pass,
  divergence tile, y-transform columns) takes one step from the same state
  and is compared with the oracle's step, on the increments.
* Beside the FP32 oracle, the FP64 build of the same pipeline replays the same
  steps: each field's error against it is printed (pytest -s) for the HIP
  path and for the FP32 oracle.

Tolerances (relative L2, FP32 storage): 1e-4 for the totals and the
per-step scalars, the parity bound of the path.  An increment is the
difference of two FP32 totals, so where it is small against the field its
FP32 storage (the reference's own) limits how well any FP32 pipeline can
resolve it: the HIP increment must be within 1e-4 of the FP32 oracle's, or,
where the two FP32 pipelines part by more, no farther from the FP64 build's
increment than the FP32 oracle's is (times 1.5).  Against the FP64 build the
HIP totals must be no farther than the FP32 oracle's (times 1.5, plus
1e-6): the HIP path is not a worse FP32 pipeline than the reference's."""
import json
import time

import numpy as np
import pytest

import oracle as O
from _cases import oracle_initial_state, oracle_model, rel_l2
from relax_amd import RelaxError, workloads

pytestmark = pytest.mark.gpu

FIELDS = ("u1", "u2", "u3", "sig", "tau", "gamma")
# steps per chunk: c1 runs 20 steps, c2 30
CHUNKS = {"c1": (4, 4, 4, 4, 4), "c2": (10, 10, 10)}
TOL = 1e-4

_T = {}


def _view(st, k, n1):
    a = st[k]
    return a[..., :n1] if k[0] in "ug" else a


def _errors(hip, ref, n1):
    return {k: rel_l2(_view(hip, k, n1), _view(ref, k, n1)) for k in FIELDS if np.abs(_view(ref, k, n1)).max() > 0}


def _diff(after, before, n1):
    return {k: _view(after, k, n1).astype(np.float64) - _view(before, k, n1) for k in FIELDS}


def _inc_errors(dh, do):
    return {k: rel_l2(dh[k], do[k]) for k in FIELDS if np.abs(do[k]).max() > 0}


def _fmt(d):
    return {k: float("%.3g" % v) for k, v in d.items()}


def _copy(st):
    return {k: (v.copy() if isinstance(v, np.ndarray) else v) for k, v in st.items()}


def _okada_on_a_sample(s, g, before, after, n1):
    """The set-up's Okada part at full size: the displacement the device's
    relax_model_okada added, on a seeded sample of grid points (the whole
    surface plane and 20000 points below it), against the oracle's
    dislocations_disp at those points; then sig = -tau + C:eps(u) of the
    device's final u through the oracle's stressupdate."""
    sx1, sx2, h3 = s["sx1"], s["sx2"], s["sx3"] // 2
    dx1, dx2, dx3 = workloads.spacing(s)
    rng = np.random.default_rng(7)
    surf = np.array([(i1, i2, 0) for i2 in range(0, sx2, 2) for i1 in range(0, sx1, 2)])
    deep = np.stack([rng.integers(0, sx1, 20000), rng.integers(0, sx2, 20000), rng.integers(1, h3, 20000)], -1)
    idx = np.concatenate([surf, deep])
    sc, tsc, _ = workloads.okada_sources(s)
    t0 = time.perf_counter()
    ref = O.dislocations_disp_points(sc.astype(O.SOURCE), tsc.astype(O.SOURCE), workloads.LAM, workloads.MU,
                                     s.get("nyquist", workloads.NYQUIST), sx1, sx2, dx1, dx2, dx3, idx)
    dev = np.stack([after["u%d" % (c + 1)][idx[:, 2], idx[:, 1], idx[:, 0]].astype(np.float64) -
                    before["u%d" % (c + 1)][idx[:, 2], idx[:, 1], idx[:, 0]] for c in range(3)], -1)
    err = rel_l2(dev, ref)
    sig = -after["tau"].astype(np.float32)
    u = [np.ascontiguousarray(after["u%d" % (c + 1)]) for c in range(3)]
    O.stressupdate(*u, workloads.LAM, workloads.MU, dx1, dx2, dx3, sig)
    err_sig = rel_l2(after["sig"], sig)
    print("%s Okada part on %d points: rel L2 %.3g (|du| max %.3g), sig from u: %.3g; oracle %.1f s" %
          (s["name"], len(idx), err, np.abs(ref).max(), err_sig, time.perf_counter() - t0), flush=True)
    assert np.abs(ref).max() > 0
    assert err < 1e-5 and err_sig < 1e-5


@pytest.mark.parametrize("name", ["c1", "c2"])
def test_coseismic_setup_at_full_size(name):
    """The set-up against the oracle's.  Where the deck has sources below
    the Nyquist size (c2's 1 km patches), Okada's part is checked on a sample
    of points (the oracle takes minutes for it on the whole grid), the rest
    of the set-up in full, and the trajectory starts the oracle from the
    device's set-up."""
    s = workloads.spec(name)
    n1 = s["sx1"]
    _, _, nsmall = workloads.okada_sources(s)
    t0 = time.perf_counter()
    g = workloads.build(s, okada=False)
    tuning = g.autotune()
    st0 = g.download()
    if nsmall:
        sc, tsc, _ = workloads.okada_sources(s)
        g.okada(sc, tsc, nyquist=s.get("nyquist", workloads.NYQUIST))
        st1 = g.download()
    print("%s device: set-up and autotune in %.1f s, tuning %s, %d sources below the Nyquist size" %
          (name, time.perf_counter() - t0, tuning, nsmall), flush=True)
    t0 = time.perf_counter()
    st32 = oracle_initial_state(s, np.float32, okada=False)
    st64 = oracle_initial_state(s, np.float64, okada=False)
    print("%s oracle: set-up (FP32 and FP64) in %.1f s" % (name, time.perf_counter() - t0), flush=True)
    err = _errors(st0, st32, n1)
    err64 = _errors(st0, st64, n1)
    print(json.dumps({"case": name, "set_up_rel_l2_vs_oracle": _fmt(err), "vs_fp64": _fmt(err64)}), flush=True)
    for k in ("u1", "u2", "u3", "sig", "tau"):
        if nsmall and k.startswith("u") and k not in err:
            continue  # every source is below the Nyquist size: u is Okada's alone
        assert k in err, k
        assert err[k] < TOL, (k, err[k])
    if nsmall:
        _okada_on_a_sample(s, g, st0, st1, n1)
        st0 = st1
        st32 = {k: (np.array(v, np.float32) if isinstance(v, np.ndarray) else v) for k, v in st1.items()}
        st64 = {k: (np.array(v, np.float64) if isinstance(v, np.ndarray) else v) for k, v in st1.items()}
        for st in (st32, st64):
            st.update(t=0.0, oi=1, i=0)
    om = oracle_model(s)
    om64 = dict(om, patches=None if om.get("patches") is None else om["patches"].copy())
    _T[name] = dict(s=s, g=g, tuning=tuning, st32=st32, st64=st64, om=om, om64=om64, dev_prev=st0, setup=st0,
                    orc_prev={k: _view(st32, k, n1).astype(np.float64) for k in FIELDS},
                    orc64_prev={k: _view(st64, k, n1).copy() for k in FIELDS}, orc_setup=_copy(st32),
                    orc64_setup=_copy(st64), chunk=0)


def _check_increments(dh, do, d64, where):
    """dh, do, d64: the increments of the HIP path, the FP32 oracle and the
    FP64 oracle; returns the printed errors after checking them."""
    inc = _inc_errors(dh, do)
    inc64 = _inc_errors(dh, d64)
    orc_inc64 = _inc_errors(do, d64)
    for kk in FIELDS:
        if kk not in inc:
            continue
        assert inc[kk] < TOL or inc64[kk] <= 1.5 * orc_inc64[kk], (where, kk, "increment", inc[kk], inc64[kk],
                                                                  orc_inc64[kk])
    return {"rel_l2_increment": _fmt(inc), "increment_vs_fp64": {"hip": _fmt(inc64), "oracle_fp32": _fmt(orc_inc64)}}


@pytest.mark.parametrize("name,k", [("c1", k) for k in range(len(CHUNKS["c1"]))] +
                         [("c2", k) for k in range(len(CHUNKS["c2"]))])
def test_trajectory_at_full_size(name, k):
    if name not in _T:
        pytest.skip("the set-up test of %s did not pass" % name)
    T = _T[name]
    if T["chunk"] != k:
        pytest.skip("an earlier chunk of %s did not pass" % name)
    s, g, n1 = T["s"], T["g"], T["s"]["sx1"]
    nsteps = CHUNKS[name][k]
    t0 = time.perf_counter()
    recs = g.run(nsteps)
    cur = g.download()
    assert len(recs) == nsteps and all(r["status"] == 0 for r in recs)
    dts = [r["Dt"] for r in recs]
    t1 = time.perf_counter()
    info = O.rk2_run(T["om"], T["st32"], nsteps, dt_replay=dts)
    t2 = time.perf_counter()
    O.rk2_run(T["om64"], T["st64"], nsteps, dt_replay=dts)
    t3 = time.perf_counter()
    print("%s chunk %d (%d steps to t = %.4g): device %.1f s, oracle FP32 %.1f s, FP64 %.1f s" %
          (name, k, nsteps, recs[-1]["t"], t1 - t0, t2 - t1, t3 - t2), flush=True)
    tot = _errors(cur, T["st32"], n1)
    tot64 = _errors(cur, T["st64"], n1)
    orc64 = _errors(T["st32"], T["st64"], n1)
    orc_now = {kk: _view(T["st32"], kk, n1).astype(np.float64) for kk in FIELDS}
    orc64_now = {kk: _view(T["st64"], kk, n1).copy() for kk in FIELDS}
    do = {kk: orc_now[kk] - T["orc_prev"][kk] for kk in FIELDS}
    d64 = {kk: orc64_now[kk] - T["orc64_prev"][kk] for kk in FIELDS}
    dh = _diff(cur, T["dev_prev"], n1)
    T["orc_prev"], T["orc64_prev"], T["dev_prev"] = orc_now, orc64_now, cur
    rec = {"case": name, "chunk": k, "steps": nsteps, "t": recs[-1]["t"], "rel_l2_total": _fmt(tot),
           "total_vs_fp64": {"hip": _fmt(tot64), "oracle_fp32": _fmt(orc64)}}
    try:
        rec.update(_check_increments(dh, do, d64, (name, k)))
    finally:
        print(json.dumps(rec), flush=True)
    # the scalars of every step: the oracle's own relaxation time for its
    # state gives the device's step
    for r, o in zip(recs, info):
        assert o[1] == pytest.approx(r["tm"], rel=TOL), ("tm", r["i"])
        assert o[5] == pytest.approx(r["power"], rel=TOL), ("power", r["i"])
        assert o[6] == pytest.approx(r["cei"], rel=TOL), ("cei", r["i"])
    if T["om"].get("patches") is not None:
        pd = g.get_patches()
        po = T["om"]["patches"]
        assert rel_l2(pd["slip"], po["slip"]) < TOL
        assert rel_l2(pd["v"], po["v"]) < TOL
    for kk in FIELDS:
        assert tot[kk] < TOL, (kk, "total", tot[kk])
        assert tot64[kk] <= 1.5 * orc64[kk] + 1e-6, (kk, "total vs FP64", tot64[kk], orc64[kk])
    T["chunk"] = k + 1


VARIANTS = [("green", v) for v in range(4)] + [("bodyforce", v) for v in range(6)] + \
           [("fft_cols", v) for v in range(3)]


@pytest.mark.parametrize("name", ["c1", "c2"])
@pytest.mark.parametrize("kind,variant", VARIANTS)
def test_every_variant_at_full_size(name, kind, variant):
    """One step from the coseismic state with each launch shape, the other
    two kinds at the benchmark's choice, against the oracle's step."""
    if name not in _T:
        pytest.skip("the set-up test of %s did not pass" % name)
    T = _T[name]
    s, g, n1 = T["s"], T["g"], T["s"]["sx1"]
    st0 = T["setup"]
    if "one_step" not in T:
        st = _copy(T["orc_setup"])
        st64 = _copy(T["orc64_setup"])
        t0 = time.perf_counter()
        info = O.rk2_run(oracle_model(s), st, 1)
        O.rk2_run(oracle_model(s), st64, 1, dt_replay=info[:, 0])
        print("%s oracle: one step from the set-up (FP32, FP64) in %.1f s" % (name, time.perf_counter() - t0),
              flush=True)
        d64 = {kk: _view(st64, kk, n1) - _view(T["orc64_setup"], kk, n1) for kk in FIELDS}
        T["one_step"] = (st, info[0], d64)
        T["orc_setup_v"] = {kk: _view(T["orc_setup"], kk, n1).astype(np.float64) for kk in FIELDS}
    st, o, d64 = T["one_step"]
    tuning = dict(T["tuning"])
    tuning[kind] = variant
    # back to the set-up: tables, planes (zero patch slip) and the first
    # event, then the set-up state itself, bit for bit
    g.reset()
    workloads.configure(g, s)
    g.upload(st0["u1"], st0["u2"], st0["u3"], st0["tau"], st0["sig"], st0["gamma"], t=0.0, oi=1, i=0)
    g.set_tuning(**tuning)
    try:
        rec = g.step(dt=float(o[0]))
    except RelaxError as e:
        if e.status == 11:  # RELAX_ERR_SIZE: the launch shape does not fit this grid
            pytest.skip("variant %s=%d does not fit %s" % (kind, variant, name))
        raise
    out = g.download()
    do = {kk: _view(st, kk, n1).astype(np.float64) - T["orc_setup_v"][kk] for kk in FIELDS}
    tot = _errors(out, st, n1)
    rep = {"case": name, "tuning": tuning, "rel_l2_total": _fmt(tot)}
    try:
        rep.update(_check_increments(_diff(out, st0, n1), do, d64, (name, kind, variant)))
    finally:
        print(json.dumps(rep), flush=True)
    assert o[1] == pytest.approx(rec["tm"], rel=TOL)
    for kk in FIELDS:
        assert tot[kk] < TOL, (kk, "total", tot[kk])
