"""The autotuner (relax_model_autotune) picks launch shapes by timing them on
the model's scratch buffers: the model's state must not change, and a model
stepped with the chosen shapes agrees with the default shapes to FP32
rounding (the variants differ in FFT radix and tile, not in the algorithm)."""
import numpy as np
import pytest

from relax_amd import workloads

from _cases import rel_l2, small_spec

pytestmark = pytest.mark.gpu


def test_autotune_keeps_state_and_results():
    s = small_spec("powerlaw", sx=128, sx3=128)
    ref = workloads.build(s)
    before = ref.download()
    recs_ref = [ref.step() for _ in range(3)]
    out_ref = ref.download()
    ref.close()

    m = workloads.build(s)
    choice = m.autotune()
    assert 0 <= choice["fft_cols"] < 3 and 0 <= choice["green"] < 4 and 0 <= choice["bodyforce"] < 6
    after = m.download()
    for k in ("u1", "u2", "u3", "sig", "tau", "gamma"):
        np.testing.assert_array_equal(after[k], before[k], err_msg=k)
    recs = [m.step() for _ in range(3)]
    out = m.download()
    m.close()
    for k in ("u1", "u2", "u3"):
        assert rel_l2(out[k][..., :128], out_ref[k][..., :128]) < 1e-5, k
    for k in ("sig", "tau"):
        assert rel_l2(out[k], out_ref[k]) < 1e-5, k
    for a, b in zip(recs, recs_ref):
        assert abs(a["t"] - b["t"]) <= 1e-6 * abs(b["t"])


@pytest.mark.parametrize("sx3", [64, 128, 256])
@pytest.mark.parametrize("green", [1, 2, 3])
def test_every_green_variant_matches_the_default(green, sx3):
    """Each launch shape of the Green column pass (4 or 8 columns through LDS
    stages of radix 16 or 8, or 16 columns with register transforms) gives
    the default's trajectory to FP32 rounding."""
    s = small_spec("powerlaw", sx=64, sx3=sx3)
    out = {}
    for v in (0, green):
        m = workloads.build(s)
        m.set_tuning(green=v)
        for _ in range(3):
            m.step(info=False)
        out[v] = m.download()
        m.close()
    for k in ("u1", "u2", "u3"):
        assert rel_l2(out[green][k][..., :64], out[0][k][..., :64]) < 1e-5, k
    for k in ("sig", "tau"):
        assert rel_l2(out[green][k], out[0][k]) < 1e-5, k


@pytest.mark.parametrize("body,sx3,kind", [(1, 64, "afterslip"), (2, 64, "afterslip"), (3, 32, "afterslip"),
                                           (3, 64, "afterslip"), (3, 128, "afterslip"), (4, 32, "afterslip"),
                                           (4, 64, "afterslip"), (4, 128, "afterslip"), (5, 64, "afterslip"),
                                           (5, 128, "afterslip"), (4, 128, "powerlaw"), (5, 64, "transient")])
def test_every_bodyforce_tile_matches_the_default(body, sx3, kind):
    """The divergence tiles (32x16, 32x8, 64x8 16 or 32 planes deep, and 64x8
    or 64x16 with the x3 differences in registers) give the same trajectory,
    with the 32-plane chunk cut short (16 planes), whole, and repeated, for
    the fault-creep, viscous and extended rheology kernels."""
    s = small_spec(kind, sx=128, sx3=sx3)
    out = {}
    for v in (0, body):
        m = workloads.build(s)
        m.set_tuning(bodyforce=v)
        for _ in range(3):
            m.step(info=False)
        out[v] = m.download()
        m.close()
    for k in ("u1", "u2", "u3", "sig", "tau"):
        np.testing.assert_array_equal(out[body][k], out[0][k], err_msg=k)
