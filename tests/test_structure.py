"""Hand-computed pins of the host-side deck conventions (CPU only).

relax_amd.structure builds the per-depth rheology tables and the internal
fault geometry that both the HIP path and the oracle consume, so a mistake
there would pass every parity test.  These cases are worked out by hand
from the reference's definitions:

  fix                     elastic3d.f90:70-83 (rounds half down)
  viscoelasticstructure   elastic3d.f90:732-829 (interfaces projected onto
                          multiples of dx3, linear interpolation, the last
                          layer continued in single precision)
  wangconvention          input.f90:1780-1800 (upper corner -> centre,
                          right-hand-rule dip direction)
  friction_plane          input.f90:1079-1093 (rake turned by 180 degrees,
                          px2 = fix(L/dx2), px3 = fix(W/dx3))
"""
import math

import numpy as np
import pytest

from relax_amd import structure


@pytest.mark.parametrize("x,expect", [(2.5, 2), (2.500001, 3), (2.49, 2), (-2.5, -3), (-2.4, -2), (-2.6, -3),
                                      (3.0, 3), (0.5, 0), (0.51, 1), (-0.5, -1)])
def test_fix_rounds_half_down(x, expect):
    assert structure.fix(x) == expect


def test_viscoelasticstructure_hand_computed():
    # interfaces at 1.2 and 3.7 km on dx3 = 0.5: fix(2.4) = 2 and fix(7.4) = 7
    # put them at 1.0 and 3.5 km (indices 2 and 7); above the first the
    # defaults hold (gammadot0 = 0, friction 0.6, the first layer's exponent)
    # defaults hold (gammadot0 = 0, friction 0.6, the first layer's exponent);
    # from the last interface down (its own sample included: vstruct(i3e:sx3))
    # the last layer's values rounded to single precision
    lay = [(1.2, 1.0, 1.0, 0.0, 0.2, 0.0), (3.7, 2.0, 3.0, 0.4, 0.5, 7.0)]
    vs = structure.viscoelasticstructure(lay, 10, 0.5)
    f = lambda v: float(np.float32(v))  # noqa: E731
    np.testing.assert_allclose(vs["gammadot0"], [0, 0, 1.0, 1.2, 1.4, 1.6, 1.8, 2.0, 2.0, 2.0], rtol=1e-15, atol=0)
    np.testing.assert_allclose(vs["stressexponent"], [1, 1, 1.0, 1.4, 1.8, 2.2, 2.6, 3.0, 3.0, 3.0], rtol=1e-15)
    np.testing.assert_allclose(vs["cohesion"], [0, 0, 0.0, 0.08, 0.16, 0.24, 0.32] + [f(0.4)] * 3, rtol=1e-12,
                               atol=1e-15)
    assert np.all(vs["cohesion"][7:] == f(0.4)) and vs["cohesion"][6] == pytest.approx(0.32, rel=1e-14)
    np.testing.assert_allclose(vs["friction"], [0.6, 0.6, 0.2, 0.26, 0.32, 0.38, 0.44, 0.5, 0.5, 0.5], rtol=1e-12)
    np.testing.assert_allclose(vs["Gk"], [0, 0, 0.0, 1.4, 2.8, 4.2, 5.6, 7.0, 7.0, 7.0], rtol=1e-12, atol=1e-15)


def test_viscoelasticstructure_projection_ties_and_single_precision():
    # 1.25/0.5 = 2.5 rounds down to 2 (1.0 km); 1.26/0.5 = 2.52 rounds up to 3
    # (1.5 km); one interface gives a step at its projected depth
    for z, first in ((1.25, 2), (1.26, 3)):
        vs = structure.viscoelasticstructure([(z, 0.1, 1.0, 0.0, 0.6, 0.0)], 8, 0.5)
        assert np.all(vs["gammadot0"][:first] == 0)
        # the continued layer is REAL(gamma1): 0.1 rounded to single precision
        assert np.all(vs["gammadot0"][first:] == float(np.float32(0.1)))
        assert vs["gammadot0"][first] != 0.1
    # between two interfaces the samples are interpolated in double
    # precision; from the last interface on (its own sample included) they
    # are the single-precision value
    vs = structure.viscoelasticstructure([(0.0, 0.1, 1.0, 0, 0.6, 0), (2.0, 0.3, 1.0, 0, 0.6, 0)], 8, 1.0)
    assert vs["gammadot0"][0] == 0.1 and vs["gammadot0"][1] == pytest.approx(0.2, rel=1e-15)
    assert np.all(vs["gammadot0"][2:] == float(np.float32(0.3)))


def test_viscoelasticstructure_rejects_bad_structures():
    with pytest.raises(ValueError):
        structure.viscoelasticstructure([], 8, 1.0)
    with pytest.raises(ValueError):  # interfaces must not go up
        structure.viscoelasticstructure([(3.0, 1, 1, 0, 0.6, 0), (1.0, 1, 1, 0, 0.6, 0)], 8, 1.0)


def _centre(x, y, z, L, W, strike_deg, dip_deg):
    """The centre of a fault given by its upper corner at the start of the
    strike (x north, y east, z down), strike clockwise from north, dip down to
    the right of the strike direction."""
    st, dp = math.radians(strike_deg), math.radians(dip_deg)
    along = np.array([math.cos(st), math.sin(st), 0.0])
    right = np.array([math.cos(st + math.pi / 2), math.sin(st + math.pi / 2), 0.0])
    down = math.cos(dp) * right + math.sin(dp) * np.array([0.0, 0.0, 1.0])
    return np.array([x, y, z]) + 0.5 * L * along + 0.5 * W * down


@pytest.mark.parametrize("deck", [
    (1e-3, -10.0, 0.0, 0.0, 20.0, 10.0, 0.0, 90.0, 0.0),   # the vertical strike-slip fault of the CPU case
    (2e-3, 0.0, 0.0, 0.0, 10.0, 4.0, 0.0, 30.0, 90.0),     # dipping east under a north strike
    (1e-3, 0.0, 0.0, 0.0, 10.0, 4.0, 90.0, 30.0, 45.0),    # dipping south under an east strike
    (1e-3, 3.0, -2.0, 1.5, 6.0, 8.0, 217.0, 63.0, -20.0),
])
def test_wangconvention_centres_the_patch(deck):
    slip, x, y, z, L, W, st, dp, rk = deck
    s2, x2, y2, z2, st2, dp2, rk2 = structure.wangconvention(slip, x, y, z, L, W, st, dp, rk)
    np.testing.assert_allclose([x2, y2, z2], _centre(x, y, z, L, W, st, dp), atol=1e-12)
    assert s2 == -slip
    # internal angles: strike -90 - strike, dip from the vertical, radians
    assert st2 == pytest.approx(math.radians(-90.0 - st))
    assert dp2 == pytest.approx(math.radians(90.0 - dp))
    assert rk2 == pytest.approx(math.radians(rk))


def test_wangconvention_vertical_fault_of_the_cpu_case():
    # deck line "1 1e-3 -10 0 0 20 10 0 90 0": a 20 x 10 km vertical fault
    # from (-10, 0) along north, centred at the origin, 5 km deep
    s2, x2, y2, z2, st2, dp2, rk2 = structure.wangconvention(1e-3, -10.0, 0.0, 0.0, 20.0, 10.0, 0.0, 90.0, 0.0)
    np.testing.assert_allclose([x2, y2, z2], [0.0, 0.0, 5.0], atol=1e-12)
    assert dp2 == 0.0


@pytest.mark.parametrize("rake,flipped", [(90.0, -90.0), (0.0, -180.0), (-30.0, 150.0), (170.0, -10.0)])
def test_friction_plane_rake_flip(rake, flipped):
    p = structure.friction_plane(-32.0, 0.0, 0.0, 64.0, 32.0, 0.0, 20.0, rake, 1.0, 1.0)
    assert p["rake"][0] == pytest.approx(math.radians(flipped))


@pytest.mark.parametrize("L,W,dx2,dx3,px2,px3", [(64.0, 32.0, 1.0, 1.0, 64, 32), (64.0, 32.0, 0.75, 0.6, 85, 53),
                                                (10.0, 5.0, 4.0, 2.0, 2, 2), (10.0, 5.0, 4.0, 1.9, 2, 3)])
def test_friction_plane_patch_counts(L, W, dx2, dx3, px2, px3):
    # fix(10/4) = fix(2.5) = 2, fix(5/2) = 2, fix(5/1.9) = fix(2.63) = 3
    p = structure.friction_plane(0.0, 0.0, 0.0, L, W, 0.0, 20.0, 90.0, dx2, dx3)
    assert (int(p["px2"][0]), int(p["px3"][0])) == (px2, px3)
    np.testing.assert_allclose([p["x"][0], p["y"][0], p["z"][0]], _centre(0.0, 0.0, 0.0, L, W, 0.0, 20.0),
                               atol=1e-12)
