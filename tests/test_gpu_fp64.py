"""Each field's error against the FP64 build of the same pipeline.

The oracle is templated on its storage type; its double build is the
all-FP64 reference of the RK2 loop.  For the reference's own CPU case
(128x128, Maxwell half-space below 20 km, 8 steps to 10 yr; the grid is
128 deep so that the layer lies inside the half grid) and the small
power-law and afterslip cases, the HIP result (FP32) and the FP32 oracle
are both compared with the FP64 oracle after the same Dt sequence.  The
errors are printed (pytest -s) so they can be quoted; the test requires
relative L2 <= 1e-4 against FP64 for every field, and the HIP error to stay
within 4x the FP32 oracle's plus 1e-6 (it is not a worse FP32 pipeline)."""
import json

import numpy as np
import pytest

import oracle as O
from _cases import oracle_initial_state, oracle_model, rel_l2, small_spec
from relax_amd import workloads

pytestmark = pytest.mark.gpu

FIELDS = ("u1", "u2", "u3", "sig", "tau", "gamma")


@pytest.mark.parametrize("kind,nsteps", [("c0b", 8), ("powerlaw", 4), ("afterslip", 4)])
def test_errors_against_fp64_build(kind, nsteps):
    s = workloads.spec(kind) if kind == "c0b" else small_spec(kind)
    om = oracle_model(s)
    st64 = oracle_initial_state(s, np.float64)
    st32 = oracle_initial_state(s, np.float32)
    info = O.rk2_run(dict(om, patches=None if om.get("patches") is None else om["patches"].copy()), st64, nsteps)
    O.rk2_run(dict(om, patches=None if om.get("patches") is None else om["patches"].copy()), st32, nsteps,
              dt_replay=info[:, 0])
    g = workloads.build(s)
    g.run(nsteps, dt_replay=info[:, 0])
    out = g.download()
    g.close()
    sx1 = s["sx1"]
    errs = {}
    for k in FIELDS:
        ref = st64[k][..., :sx1] if k[0] in "ug" else st64[k]
        hip = out[k][..., :sx1] if k[0] in "ug" else out[k]
        cpu = st32[k][..., :sx1] if k[0] in "ug" else st32[k]
        if np.abs(ref).max() == 0:
            continue
        errs[k] = (rel_l2(hip, ref), rel_l2(cpu, ref))
    print(json.dumps({"case": kind, "steps": nsteps,
                      "rel_l2_vs_fp64": {k: {"hip_fp32": float("%.3g" % a), "oracle_fp32": float("%.3g" % b)}
                                         for k, (a, b) in errs.items()}}))
    assert errs
    for k, (a, b) in errs.items():
        assert a <= 1e-4, (k, a)
        assert a <= 4 * b + 1e-6, (k, a, b)
