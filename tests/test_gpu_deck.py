"""RELAX input decks through the deck reader (relax_amd.deck) onto the device
model, against the oracle on the same deck.

The hand-written deck uses every deck feature the model consumes: an origin
and a rotation of the frame, a depth-dependent prestress (tensorstructure,
and tau = -prestress before the first event, relax.f90:286-297), linear and
power-law interfaces, a linear weak zone, a segment with its own taper, a
segment below the Nyquist size (Okada's displacement, in the deck's frame as
the reference computes it), a point dilatation source and a surface load.
The tutorial decks the reference ships run at their own sizes.
"""
import os

import numpy as np
import pytest

import oracle as O
from relax_amd import deck, serve, workloads

from _cases import oracle_initial_state, oracle_model, rel_l2

pytestmark = pytest.mark.gpu

DECKS = os.path.join(os.path.dirname(__file__), "golden", "decks")

DECK = """# 64^3 deck exercising the frame, prestress and every source kind
64 64 64
1 1 1 0.2 2
2 -3 20
0 0
out
3e4 3e4 8.33e-4
# interval, negative step (skip 1), scaling
100 -1 1
# observation planes, points, stress segments
0
0
0
# prestress interfaces: z s11 s12 s13 s22 s23 s33
2
1 0  0 0.5 0 0   0 0
2 20 0 1.5 0 0.3 0 0
# linear interfaces and weak zones
1
1 12 0.05 0
1
1 0.3 -4 -2 4 8 6 5 30 60
# power-law interfaces and weak zones
1
1 8 5e6 3 0
0
# fault creep
0
# inter-seismic shear and tensile
0
0
# events
1
2
1 1e-3 -6 0 0 10 6 10 80 20 0.3
2 5e-4  3 4 1 1.5 1.5 45 60 90
0
1
1 1e-4 5 -5 6
1
1 -3 4 6 5 0.2 0 0
"""


def _pair(s, nsteps):
    st = oracle_initial_state(s)
    om = oracle_model(s)
    ost = {k: (v.copy() if isinstance(v, np.ndarray) else v) for k, v in st.items()}
    info = O.rk2_run(om, ost, nsteps)
    g = workloads.build(s)
    first = g.download()
    recs = workloads.run(g, s, nsteps)
    out = g.download()
    patches = g.get_patches() if om.get("patches") is not None else None
    g.close()
    return st, first, ost, info, out, recs, om, patches


def _check(s, st, first, ost, info, out, recs, nsteps):
    n1 = s["sx1"]
    assert len(recs) == nsteps
    for k in ("u1", "u2", "u3"):
        assert rel_l2(first[k][..., :n1], st[k][..., :n1]) < 1e-4, k
        assert rel_l2(out[k][..., :n1], ost[k][..., :n1]) < 1e-4, k
    for k in ("sig", "tau"):
        assert rel_l2(first[k], st[k]) < 1e-4, k
        assert rel_l2(out[k], ost[k]) < 1e-4, k
    if np.abs(ost["gamma"]).max() > 0:
        assert rel_l2(out["gamma"][..., :n1], ost["gamma"][..., :n1]) < 1e-4
    for k, r in enumerate(recs):
        assert abs(r["Dt"] / info[k, 0] - 1) < 1e-5, k
        assert abs(r["power"] / info[k, 5] - 1) < 1e-4


def test_deck_model_matches_oracle():
    s = deck.read(DECK).spec()
    st, first, ost, info, out, recs, om, _ = _pair(s, 4)
    # the background stress is in the initial state: deep down, away from
    # the sources, sigma_12 is the prestress (1.5 from 20 km) and tau its
    # negative
    assert np.abs(first["tau"][25:, :, :, 1] + 1.5).max() < 1e-4
    assert abs(float(np.median(first["sig"][25:, :, :, 1])) - 1.5) < 0.02
    _check(s, st, first, ost, info, out, recs, 4)


@pytest.mark.parametrize("name", ["run2", "run3", "run4"])
def test_tutorial_deck_at_its_size(name):
    """The reference's tutorial decks at their own grids (256^3, dx 0.05):
    power law (run2), afterslip on two planes (run3), a linear layer with a
    weak zone (run4); four steps against the oracle."""
    s = deck.read(os.path.join(DECKS, name + ".input")).spec()
    print("deck", name, flush=True)
    st, first, ost, info, out, recs, om, patches = _pair(s, 4)
    _check(s, st, first, ost, info, out, recs, 4)
    if patches is not None:
        assert np.abs(om["patches"]["v"]).max() > 0
        assert rel_l2(patches["v"], om["patches"]["v"]) < 1e-4


def test_deck_requests_are_served_as_alone():
    """Deck requests of one batch class and of another (a different mu):
    served together, each result is the one it gets alone."""
    reqs = [serve.deck_request(0, DECK, 3), serve.deck_request(1, DECK.replace("1 1e-3 -6", "1 2e-3 -6"), 4),
            serve.deck_request(2, DECK.replace("3e4 3e4 8.33e-4", "3e4 2.5e4 8.33e-4"), 3)]
    assert reqs[0].cls == reqs[1].cls != reqs[2].cls
    alone = {r.rid: r for r in serve.run_back_to_back(reqs, realtime=False)}
    mixed = {r.rid: r for r in serve.Server().run(reqs, realtime=False)}
    for q in reqs:
        a, b = alone[q.rid], mixed[q.rid]
        assert a.status == b.status == 0
        assert b.nsteps == q.nsteps and a.t == b.t
        for ua, ub in zip(a.surface, b.surface):
            assert np.abs(ua).max() > 0
            np.testing.assert_array_equal(ua, ub)


EIGEN = """# 64^3 deck with eigenstrain volumes (--with-eigenstrain --curvilinear)
64 64 64
1 1 1 0.2 2
1 2 15
0 0
out
3e4 3e4 8.33e-4
100 -1 1
0
0
0
0
# linear interfaces, no weak zones
1
1 10 0.05 0
0
# power-law interfaces
0
# fault creep
0
0
0
# two events
2
# first: a shear segment, a vertical and a dipping cuboid, a tetrahedron
1
1 5e-4 -8 2 0 6 4 20 70 30
0
0
2
1 1e-3 -2e-4 3e-4 5e-4 -4e-4 2e-4 -3 -4 1 6 4 3 10 90
2 -5e-4 4e-4 0 3e-4 2e-4 -1e-4 4 3 2 5 3 2 -40 55
1
1 3e-4 -2e-4 1e-4 -3e-4 2e-4 4e-4 1 2 3 4
4
1 -4 -3 2
2 5 -2 3
3 0 6 4
4 1 0 11
0
# second, at the time the steps reach: one cuboid
8
0
0
0
1
1 2e-4 0 0 -3e-4 0 1e-4 6 -6 2 4 4 4 0 90
0
0
"""


def test_eigenstrain_volumes_match_oracle():
    """Cuboids (vertical and dipping) and a tetrahedron in the set-up, a
    cuboid in a later event: the set-up and the trajectory across the event
    follow the oracle's (relax.f90:1151-1223 through the model's event)."""
    s = deck.read(EIGEN, with_eigenstrain=True, curvilinear=True).spec()
    st, first, ost, info, out, recs, om, _ = _pair(s, 6)
    assert np.abs(first["tau"]).max() > 0
    assert om.get("event_cuboids") is not None and len(om["event_cuboids"][0]) == 1
    assert any(r["e"] > 0 for r in recs)  # the second event was applied within the steps
    _check(s, st, first, ost, info, out, recs, 6)
