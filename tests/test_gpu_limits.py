"""The largest grids the model accepts along each axis (sx1 = sx2 = 1024 and
sx3 = 512, relax_model_create) against the oracle: 1024x1024x64, whose x
transform runs 1024-point rows and whose split y transform runs 128-point
columns, and 32x32x512, whose 256-deep z columns take the LDS form of the
Green pass.  Three RK2 steps of the power-law workload, with the default
launch shapes and with the autotuner's, at the path's parity bound
(relative L2 1e-4); and fault creep at 1024x1024x64."""
import numpy as np
import pytest

import oracle as O
from _cases import oracle_initial_state, oracle_model, rel_l2, small_spec
from relax_amd import workloads

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("sx,sx3", [(1024, 64), (32, 512)])
def test_largest_grids_follow_the_oracle(sx, sx3):
    s = small_spec("powerlaw", sx=sx, sx3=sx3)
    nsteps = 3
    st = oracle_initial_state(s)
    info = O.rk2_run(oracle_model(s), st, nsteps)
    for tuned in (False, True):
        g = workloads.build(s)
        if tuned:
            print("%dx%dx%d autotune" % (sx, sx, sx3), g.autotune(), flush=True)
        recs = g.run(nsteps, dt_replay=info[:, 0])
        out = g.download()
        g.close()
        assert len(recs) == nsteps and all(r["status"] == 0 for r in recs)
        for k in ("u1", "u2", "u3"):
            assert rel_l2(out[k][..., :sx], st[k][..., :sx]) < 1e-4, (k, tuned)
        for k in ("sig", "tau"):
            assert rel_l2(out[k], st[k]) < 1e-4, (k, tuned)
        assert np.abs(st["gamma"]).max() > 0
        assert rel_l2(out["gamma"][..., :sx], st["gamma"][..., :sx]) < 1e-4, tuned
        for k, r in enumerate(recs):
            assert abs(r["tm"] / info[k, 1] - 1) < 1e-4, (k, tuned)


def test_afterslip_at_the_largest_horizontal_grid():
    """Fault creep on a friction plane at 1024x1024x64 (the afterslip kernels,
    the patch velocities of monitorfriction and frictionadd) follows the
    oracle over 3 steps."""
    s = small_spec("afterslip", sx=1024, sx3=64)
    nsteps = 3
    st = oracle_initial_state(s)
    om = oracle_model(s)
    info = O.rk2_run(om, st, nsteps)
    g = workloads.build(s)
    recs = g.run(nsteps, dt_replay=info[:, 0])
    out = g.download()
    pat = g.get_patches()
    g.close()
    assert all(r["status"] == 0 for r in recs)
    for k in ("u1", "u2", "u3"):
        assert rel_l2(out[k][..., :1024], st[k][..., :1024]) < 1e-4, k
    for k in ("sig", "tau"):
        assert rel_l2(out[k], st[k]) < 1e-4, k
    ref = om["patches"]
    assert np.abs(ref["v"]).max() > 0
    assert rel_l2(pat["slip"], ref["slip"]) < 1e-4
    assert rel_l2(pat["v"], ref["v"]) < 1e-4
