"""The benchmark workloads at their full sizes -- c1, 512x512x256 with a
power-law mantle, and c2, 256x256x128 with afterslip on a friction plane and
a power-law mantle: one adaptive RK2 step of the device from its coseismic
state against the oracle stepping the same state (FP32 storage, relative L2),
and the relaxation time and amplitudes the oracle finds for that state."""
import time

import numpy as np
import pytest

import oracle as O
from _cases import oracle_model, rel_l2
from relax_amd import workloads

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("name", ["c1", "c2"])
def test_full_size_step_matches_oracle(name):
    s = workloads.spec(name)
    t0 = time.perf_counter()
    g = workloads.build(s)
    st0 = g.download()
    rec = g.step()
    out = g.download()
    g.close()
    print("device: set-up, one step and downloads in %.1f s" % (time.perf_counter() - t0), flush=True)
    om = oracle_model(s)
    st = {k: np.ascontiguousarray(st0[k]) for k in ("u1", "u2", "u3", "tau", "sig", "gamma")}
    st.update(t=0.0, oi=1, i=0)
    t0 = time.perf_counter()
    info = O.rk2_run(om, st, 1, dt_replay=[rec["Dt"]])
    print("oracle: one step in %.1f s" % (time.perf_counter() - t0), flush=True)
    # the oracle's own relaxation time for this state gives the same step
    assert info[0, 1] == pytest.approx(rec["tm"], rel=1e-5)
    assert info[0, 5] == pytest.approx(rec["power"], rel=1e-5)
    assert info[0, 6] == pytest.approx(rec["cei"], rel=1e-5)
    n1 = s["sx1"]
    err = {k: rel_l2(out[k][..., :n1], st[k][..., :n1]) for k in ("u1", "u2", "u3")}
    err.update({k: rel_l2(out[k], st[k]) for k in ("sig", "tau", "gamma")})
    print("relative L2 against the oracle:", {k: "%.2e" % v for k, v in err.items()}, flush=True)
    for k in ("u1", "u2", "u3"):
        assert np.abs(st[k][..., :n1]).max() > 0, k
        assert err[k] < 1e-5, k
    assert err["sig"] < 1e-5 and err["tau"] < 1e-5
    assert err["gamma"] < 1e-4
