"""The deck reader (relax_amd.deck) against the reference's own input decks
and hand-written ones (CPU only).

The decks under tests/golden/decks are the reference's tutorial and Landers
inputs, kept verbatim; the values checked below were read off those files
by hand.  Section order and the optional sections follow input.f90 init
(:252-1754); comment and blank lines are skipped as getdata.f does.
"""
import math
import os

import numpy as np
import pytest

from relax_amd import deck, structure, workloads

DECKS = os.path.join(os.path.dirname(__file__), "golden", "decks")


def _read(name, **kw):
    return deck.read(os.path.join(DECKS, name + ".input"), **kw)


def test_run2_tutorial_deck():
    d = _read("run2")
    assert (d.sx1, d.sx2, d.sx3) == (256, 256, 256)
    assert (d.dx1, d.dx2, d.dx3, d.beta, d.nyquist) == (0.05, 0.05, 0.05, 0.2, 0.0)
    assert (d.x0, d.y0, d.rot, d.oz, d.ozs) == (0.0, 0.0, 0.0, 0.0, 0.5)
    assert d.wdir == "output2"
    assert (d.lam, d.mu, d.gam) == (30.0, 30.0, 8.33e-4)
    # "20 -1 0.5": a negative step is an output every ceil(1) step, with the
    # computed step scaled by 0.5
    assert (d.interval, d.odt, d.skip, d.tscale) == (20.0, -1.0, 1, 0.5)
    assert d.nonlinear == [(2.0, 5e3, 3.0, 0.0), (8.0, 5e3, 3.0, 0.0)]
    assert d.linear == [] and d.faultcreep == [] and d.prestress == []
    assert len(d.events) == 1
    assert d.events[0].shear == [(1.0, -1.0, 0.0, 0.0, 2.0, 1.0, 0.0, 90.0, 0.0)]
    assert d.events[0].tensile == [] and d.events[0].mogi == [] and d.events[0].loads == []


def test_run3_tutorial_deck_afterslip():
    d = _read("run3")
    assert (d.interval, d.odt, d.skip, d.tscale) == (0.5, -1.0, 1, 0.01)
    assert d.planes_obs == [(-3.0, 0.0, 0.5, 6.0, 2.5, 0.0, 90.0)]
    assert d.faultcreep == [(0.0, 1e3, 1e3, 0.0, 0.0)]
    assert d.afterslip == [(-2.0, 0.1, 0.0, 4.0, 2.0, 0.0, 90.0, 0.0), (-2.0, 1.4, 0.0, 4.0, 2.0, 0.0, 90.0, 0.0)]
    assert d.events[0].shear == [(1.0, -0.5, 0.1, 0.0, 1.0, 1.0, 0.0, 90.0, 0.0)]
    s = d.spec()
    assert s["friction"] == [(0.0, 1e3, 1e3, 0.0, 0.0, 0.0)]  # LAYER order: z, v0, (a-b)sigma, cohesion, friction
    assert len(s["planes"]) == 2 and s["planes"][1]["y"] == 1.4


def test_run4_tutorial_deck_segments_and_weak_zone():
    d = _read("run4")
    assert len(d.stress_segments) == 5
    assert d.stress_segments[3] == (0.5, 1.0, 0.5, 0.1, 0.1, 45.0, 90.0, 0.6)
    assert d.linear == [(2.0, 1.0, 0.0), (3.0, 1.0, 0.0), (3.0, 0.0, 0.0)]
    assert d.weak_linear == [(1.0, -1.0, 0.0, 1.1, 2.0, 0.9, 1.0, 0.0, 90.0)]
    assert d.events[0].shear == [(-1.0, -1.0, 0.0, 0.0, 2.0, 1.0, 0.0, 90.0, 0.0)]
    s = d.spec()
    t = workloads.tables(s)["linear"]
    z = np.arange(128) * 0.05
    # fluidity 1 between the interfaces at 2 and 3 km, 0 above and (after
    # the discontinuity at 3 km) below, exponent 1 throughout
    assert np.all(t["gammadot0"][z < 1.9] == 0.0)
    assert np.all(t["gammadot0"][(z > 2.05) & (z < 2.95)] == 1.0)
    assert np.all(t["gammadot0"][z > 3.05] == 0.0)
    assert np.all(t["stressexponent"] == 1.0)


def test_costress_deck_fortran_exponent():
    d = _read("co-stress")
    assert (d.lam, d.mu) == (3e4, 3e4)  # "3e4 3d4": the Fortran exponent letter
    assert (d.dx1, d.nyquist, d.interval) == (1.0, 2.0, 0.0)
    assert d.events[0].shear == [(1.0, -10.0, 0.0, 10.0, 20.0, 20.0, 0.0, 30.0, 90.0)]


def test_landers_deck():
    d = _read("landers")
    assert (d.sx1, d.sx2, d.sx3) == (512, 512, 512)
    assert (d.dx1, d.beta, d.nyquist) == (0.6, 0.3, 2.0)
    assert (d.lam, d.mu) == (1.0, 1.0)
    assert d.nonlinear == [(25.0, 0.3, 1.0, 0.0), (30.0, 0.3, 1.0, 0.0), (30.0, 0.0, 1.0, 0.0),
                           (60.0, 0.0, 1.0, 0.0), (60.0, 3.0, 1.0, 0.0)]
    sh = d.events[0].shear
    assert len(sh) == 179
    assert sh[0] == (1.3475, 14.246, -45.439, 10.056, 5.6, 4.94, 132.7, 91.0, -114.7)
    assert sh[35] == (1.4857, -13.605, -20.07, 10.056, 4.78, 4.94, 142.2, 91.0, -97.3)
    s = d.spec()
    src = workloads.event_sources(s, 1)["sources"]
    assert len(src) == 179 and np.all(src["beta"] == 0.3)


@pytest.mark.parametrize("name", ["run1", "mogi"])
def test_decks_of_older_layouts_are_rejected(name):
    """run1.input carries a geographic-origin line and mogi.input has no
    stress-segment count; input.f90's reader stops on both."""
    with pytest.raises(deck.DeckError):
        _read(name)


MINI = """# a comment
   # an indented comment

64 64 64
1 1 1 0.25 2
0 0 0
0 0
out
3e4 3e4 8.33e-4
10 0.5
0
0
0
0
0
0
0
0
0
1
1
1 1 -5 0 0 10 5 0 90 0
0
0
0
"""


# line numbers of MINI's data (0-based, in MINI.split("\n"))
FRAME, TIMES, PRESTRESS, POWERLAW, FAULTCREEP, NSHEAR, PATCH, MOGI, LOADS = 5, 9, 13, 15, 16, 20, 21, 23, 24


def _mini(at=None, put=(), drop=0):
    """MINI with the lines at..at+drop replaced by `put`."""
    lines = MINI.split("\n")
    if at is not None:
        lines = lines[:at] + list(put) + lines[at + drop:]
    return "\n".join(lines)


def test_comments_and_blank_lines_are_skipped():
    d = deck.read(MINI)
    assert (d.sx1, d.beta, d.interval, d.odt, d.skip) == (64, 0.25, 10.0, 0.5, 0)
    assert d.events[0].shear == [(1.0, -5.0, 0.0, 0.0, 10.0, 5.0, 0.0, 90.0, 0.0)]


def test_index_misfit_and_truncated_deck():
    with pytest.raises(deck.DeckError, match="index"):
        deck.read(_mini(PATCH, ["2 1 -5 0 0 10 5 0 90 0"], 1))
    with pytest.raises(deck.DeckError, match="ends"):
        deck.read("\n".join(MINI.split("\n")[:20]))
    with pytest.raises(deck.DeckError, match="positive"):
        deck.read(_mini(PATCH, ["1 1 -5 0 0 0 0 0 90 0"], 1))
    with pytest.raises(deck.DeckError, match="beta"):
        deck.read(_mini(PATCH, ["1 1 -5 0 0 10 5 0 90 0 1e9"], 1))


def test_first_event_needs_a_source_or_an_interval():
    text = _mini(NSHEAR, ["0"], 2).replace("\n10 0.5\n", "\n0 0.5\n")
    with pytest.raises(deck.DeckError, match="first event"):
        deck.read(text)
    assert deck.read(_mini(NSHEAR, ["0"], 2)).events[0].shear == []  # a loading interval is enough


def test_optional_taper_per_segment():
    d = deck.read(_mini(PATCH, ["1 1 -5 0 0 10 5 0 90 0 0.4"], 1))
    src = workloads.event_sources(d.spec(), 1)["sources"]
    assert src["beta"][0] == 0.4
    src = workloads.event_sources(deck.read(MINI).spec(), 1)["sources"]
    assert src["beta"][0] == 0.25  # the deck's beta


def test_event_time_is_a_multiple_of_the_output_step():
    """skip = 0: a later event's time is nint(time/odt)*odt (input.f90:1264-1267)."""
    text = _mini(LOADS + 1, ["1.3", "0", "0", "0", "1", "1 0 0 4 4 0.1 0 0"]).replace("\n1\n1\n1 1 -5", "\n2\n1\n1 1 -5")
    d = deck.read(text)
    assert len(d.events) == 2
    assert d.events[1].time == 1.5  # nint(2.6) = 3
    assert d.events[1].loads == [(0.0, 0.0, 4.0, 4.0, 0.1, 0.0, 0.0)]
    assert workloads.event_times(d.spec()) == (0.0, 1.5)
    with pytest.raises(deck.DeckError, match="increase"):
        deck.read(text.replace("\n1.3\n", "\n-0.2\n"))


def test_optional_sections():
    """--with-transient adds two interface lists after the power-law ones;
    --with-eigenstrain adds cuboids per event and --curvilinear tetrahedra
    and their vertices (input.f90:740-998, :1496-1671)."""
    text = _mini(FAULTCREEP, ["1", "1 5 0.5 2e4", "0", "1", "1 8 3e6 3 1e4", "0"])
    d = deck.read(text, with_transient=True)
    assert d.ltransient == [(5.0, 0.5, 2e4)] and d.nltransient == [(8.0, 3e6, 3.0, 1e4)]
    t = workloads.tables(d.spec())
    assert t["ltransient"]["Gk"][10] == pytest.approx(2e4) and t["nltransient"]["stressexponent"][10] == 3.0
    with pytest.raises(deck.DeckError):
        deck.read(text)  # without the option the lists are out of place
    text = _mini(LOADS, ["1", "1 1e-3 0 0 0 0 0 0 0 5 4 3 2 0 90", "1", "1 0 0 0 0 0 1e-3 1 2 3 4", "4",
                         "1 0 0 1", "2 1 0 1", "3 0 1 1", "4 0 0 2"])
    d = deck.read(text, with_eigenstrain=True, curvilinear=True)
    assert d.events[0].cuboids == [(1e-3, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 5.0, 4.0, 3.0, 2.0, 0.0, 90.0)]
    assert d.events[0].tetrahedra == [(0.0, 0.0, 0.0, 0.0, 0.0, 1e-3, 1, 2, 3, 4)]
    assert d.events[0].vertices[3] == (0.0, 0.0, 2.0)


def test_frame_of_the_deck():
    """x0, y0, rot: Wang's convention shifts by the origin and turns by rot;
    loads and point sources are only turned (input.f90:1364-1369, :1488,
    :1733).  The centre of a vertical patch of strike 0 (along x1) with its
    upper corner at (-5, 0, 0), length 10 and width 5 is (0, 0, 2.5) before
    the frame, (-10, 5, 2.5) after the shift by (10, -5)."""
    lines = MINI.split("\n")
    lines = (lines[:FRAME] + ["10 -5 30"] + lines[FRAME + 1:MOGI] +
             ["1", "1 1e-3 6 -2 1", "1", "1 2 3 4 4 0.1 0 0"] + lines[LOADS + 1:])
    s = deck.read("\n".join(lines)).spec()
    ev = workloads.event_sources(s, 1)
    a = math.radians(30.0)
    x, y = -10.0, 5.0
    assert ev["sources"]["x"][0] == pytest.approx(x * math.cos(a) + y * math.sin(a), abs=1e-12)
    assert ev["sources"]["y"][0] == pytest.approx(-x * math.sin(a) + y * math.cos(a), abs=1e-12)
    assert ev["sources"]["z"][0] == pytest.approx(2.5, abs=1e-12)
    assert ev["sources"]["strike"][0] == pytest.approx(math.radians(-90.0 + 30.0))
    assert ev["mogi"]["x"][0] == pytest.approx(6 * math.cos(a) - 2 * math.sin(a))
    assert ev["mogi"]["y"][0] == pytest.approx(-6 * math.sin(a) - 2 * math.cos(a))
    assert ev["loads"]["x"][0] == pytest.approx(2 * math.cos(a) + 3 * math.sin(a))
    assert ev["loads"]["y"][0] == pytest.approx(-2 * math.sin(a) + 3 * math.cos(a))
    # Okada's solution reads the deck's own copies, in the deck's frame
    sc, _ = structure.deck_copies(s["deck"])
    assert sc["x"][0] == -5.0 and sc["y"][0] == 0.0


def test_tensorstructure_hand_computed():
    """tensorstructure (elastic3d.f90:630-702): zero above the first
    interface, linear between interfaces projected onto multiples of dx3,
    the last interface continued downward.  Interfaces at 1.2 and 3.7 km on
    dx3 = 0.5 project to 1.0 and 3.5 km (samples 2 and 7); s11 goes 1 -> 3.5
    and s33 -2 -> 0.5 there, so s11 = z and s33 = z - 3 in between."""
    lay = [(1.2, 1.0, 0.5, 0.0, 0.0, 0.0, -2.0), (3.7, 3.5, 0.5, 0.0, 0.0, 0.0, 0.5)]
    t = structure.tensorstructure(lay, 10, 0.5)["t"]
    z = np.arange(10) * 0.5
    assert np.all(t[:2] == 0.0)
    np.testing.assert_allclose(t[2:8, 0], z[2:8], rtol=1e-6)
    np.testing.assert_allclose(t[2:8, 5], z[2:8] - 3.0, rtol=1e-6, atol=1e-6)
    np.testing.assert_allclose(t[2:8, 1], 0.5, rtol=1e-6)
    assert np.all(t[8:, 0] == np.float32(3.5)) and np.all(t[8:, 5] == np.float32(0.5))
    with pytest.raises(ValueError):
        structure.tensorstructure([(3.0, 0, 0, 0, 0, 0, 0), (1.0, 0, 0, 0, 0, 0, 0)], 10, 0.5)


def test_prestress_deck_to_tables():
    """A deck's prestress interfaces become the model's background stress."""
    d = deck.read(_mini(PRESTRESS, ["1", "1 2 0 1 0 0 0 -1"], 1))
    assert d.prestress == [(2.0, 0.0, 1.0, 0.0, 0.0, 0.0, -1.0)]
    pre = workloads.tables(d.spec())["prestress"]["t"]
    assert np.all(pre[:2] == 0.0) and np.all(pre[2:, 1] == 1.0) and np.all(pre[2:, 5] == -1.0)
