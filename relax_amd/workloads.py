"""Seeded synthetic workloads of the RK2 step (units: km, MPa, yr; slip in km).

  c0   128x128x64, dx 0.5: one vertical strike-slip fault (20 x 10 km, 1 m),
       linear Maxwell below 20 km (eta 1e19 Pa s); 8 steps to 10 yr
  c0b  the same with sx3 = 128 so that the Maxwell layer is sampled
  c1   512x512x256, dx 0.5: 40 patches (5 x 5 km, slip 0-3 m, random
       orientation), power law n = 3 below a 25 km lid with the reference
       strain rate growing exponentially with depth (e-fold 10 km); 20 steps
  c2   256x256x128, dx 1: rate-strengthening afterslip plane (64 x 32 km,
       64 x 32 patches, dip 20, rake 90) loaded by a 32 x 16 km coseismic
       thrust up dip of it (slip 0-1 m; its 1 km patches are below the
       Nyquist size, so their displacement is Okada's), power law n = 3.5
       below 30 km (eta0 1e18); 30 steps.  With 3 m of slip the stress on
       the adjacent friction plane exceeds 20 (a-b)sigma, where the
       reference stops (friction3d.f90:200-215).
  c3   one of 64 batched models on 128^3, dx 1: 1-10 patches, log-uniform
       eta in [1e17, 1e20], n in {1, 2, 3, 3.5}, 15-40 km lid; 10 steps

Viscosity maps to the reference strain rate gammadot0 = mu/eta (per year);
for n > 1 eta is the effective viscosity at tau_ref = 1 MPa, so
gammadot0 = (mu/eta) (mu/tau_ref)^(n-1).
"""
import math

import numpy as np

from . import structure
from ._lib import LAYER, TENSOR_LAYER
from .model import FAULTCREEP, LINEAR, LTRANSIENT, NLTRANSIENT, POWERLAW, Model

MU = LAM = 3.0e4          # MPa
GAM = 8.33e-4             # Cowling coefficient, 1/km
BETA = 0.2                # taper width
NYQUIST = 2.0             # patches up to nyquist*dx get no source() body force (run1.input:5)
SEC_PER_YR = 3.15576e7
SEED0 = 20261202


def gdot0(eta, n=1.0, mu_pa=3.0e10, tau_ref=1.0):
    """Reference strain rate (1/yr) of viscosity eta (Pa s)."""
    return mu_pa / eta * SEC_PER_YR * (MU / tau_ref) ** (n - 1.0)


def _deck_patch(rng, lo=-40.0, hi=40.0, zmax=15.0, size=5.0, smax=3e-3):
    strike = rng.uniform(0, 360)
    dip = rng.uniform(30, 90)
    rake = (0.0 if rng.uniform() < 0.5 else 90.0) + rng.uniform(-30, 30)
    return (rng.uniform(0, smax), rng.uniform(lo, hi), rng.uniform(lo, hi), rng.uniform(0, zmax), size, size,
            strike, dip, rake)


def spec(name):
    """Description of one workload (a dict the builders below consume)."""
    if name in ("c0", "c0b"):
        sx3 = 64 if name == "c0" else 128
        return dict(name=name, sx1=128, sx2=128, sx3=sx3, dx=0.5, linear=[(20.0, gdot0(1e19), 1.0, 0, 0.6, 0)],
                    nonlinear=None, friction=None, planes=None,
                    deck=[(1e-3, -10.0, 0.0, 0.0, 20.0, 10.0, 0.0, 90.0, 0.0)],
                    interval=10.0, odt=1.25, skip=0, tscale=1.0, nsteps=8)
    if name == "c1":
        rng = np.random.default_rng(SEED0 + 1)
        deck = [_deck_patch(rng) for _ in range(40)]
        g25 = gdot0(1e19, 3.0)
        nonlin = [(25.0 + 0.5 * k, g25 * math.exp((0.5 * k) / 10.0), 3.0, 0.0, 0.6, 0.0) for k in range(78)]
        return dict(name=name, sx1=512, sx2=512, sx3=256, dx=0.5, linear=None, nonlinear=nonlin, friction=None,
                    planes=None, deck=deck, interval=1e9, odt=1.0, skip=1, tscale=1.0, nsteps=20)
    if name == "c2":
        rng = np.random.default_rng(SEED0 + 2)
        dip = 20.0
        # coseismic thrust: 32 x 16 patches of 1 km from the surface down dip
        deck = []
        for j3 in range(16):
            for j2 in range(32):
                z = j3 * math.sin(math.radians(dip))
                y = -(j3 + 0.0) * math.cos(math.radians(dip))
                deck.append((rng.uniform(0, 1e-3), -16.0 + j2, y, z, 1.0, 1.0, 0.0, dip, 90.0))
        ztop = 16 * math.sin(math.radians(dip))
        ytop = -16 * math.cos(math.radians(dip))
        plane = dict(x=-32.0, y=ytop, z=ztop, length=64.0, width=32.0, strike=0.0, dip=dip, rake=90.0)
        nonlin = [(30.0, gdot0(1e18, 3.5), 3.5, 0.0, 0.6, 0.0)]
        # nyquist 2 (the default): the 1 km patches on the 1 km grid are
        # below the Nyquist size, so they take Okada's displacement and add
        # their moment density to tau, as the reference's set-up does
        return dict(name=name, sx1=256, sx2=256, sx3=128, dx=1.0, linear=None, nonlinear=nonlin,
                    friction=[(0.0, 1e-3, 1.0, 0.0, 0.6, 0.0)], planes=[plane], deck=deck,
                    interval=1e9, odt=1.0, skip=1, tscale=1.0, nsteps=30)
    if name.startswith("c3"):
        idx = int(name[3:]) if len(name) > 3 else 0
        rng = np.random.default_rng(SEED0 + 3 + 1000 * idx)
        n = [1.0, 2.0, 3.0, 3.5][rng.integers(0, 4)]
        eta = 10 ** rng.uniform(17, 20)
        lid = rng.uniform(15, 40)
        npatch = int(rng.integers(1, 11))
        deck = []
        for _ in range(npatch):
            p = list(_deck_patch(rng, -30, 30, 10.0, 5.0, 5e-3))
            p[0] = rng.uniform(0.5e-3, 5e-3)
            deck.append(tuple(p))
        lay = [(lid, gdot0(eta, n), n, 0.0, 0.6, 0.0)]
        return dict(name=name, sx1=128, sx2=128, sx3=128, dx=1.0, linear=lay if n == 1.0 else None,
                    nonlinear=None if n == 1.0 else lay, friction=None, planes=None, deck=deck,
                    interval=1e9, odt=1.0, skip=1, tscale=1.0, nsteps=10)
    raise KeyError(name)


def constants(s):
    """(lambda, mu, gamma, beta) of a spec: its own, or the module's."""
    return (float(s.get("lam", LAM)), float(s.get("mu", MU)), float(s.get("gam", GAM)),
            float(s.get("beta", BETA)))


def frame(s):
    """(x0, y0, rot) of a spec: the deck's origin and rotation (input.f90:264)."""
    return float(s.get("x0", 0.0)), float(s.get("y0", 0.0)), float(s.get("rot", 0.0))


def spacing(s):
    """(dx1, dx2, dx3) of a spec, whose "dx" is one spacing or three."""
    dx = s["dx"]
    if np.ndim(dx) == 0:
        return float(dx), float(dx), float(dx)
    dx1, dx2, dx3 = (float(v) for v in dx)
    return dx1, dx2, dx3


def tables(s):
    """Per-depth tables (sx3/2 entries) of the mechanisms of a spec."""
    h3 = s["sx3"] // 2
    dx3 = spacing(s)[2]
    out = {}
    for key in ("linear", "nonlinear", "friction", "ltransient", "nltransient"):
        lay = s.get(key)
        out[key] = None if lay is None else structure.viscoelasticstructure(lay, h3, dx3)
    pre = s.get("prestress_layers")
    out["prestress"] = np.zeros(h3, TENSOR_LAYER) if not pre else structure.tensorstructure(pre, h3, dx3)
    return out


# weak-zone deck keys of a spec and the mechanism each perturbs
WEAK_KEYS = (("weak_linear", LINEAR), ("weak_nonlinear", POWERLAW), ("weak_ltransient", LTRANSIENT),
             ("weak_nltransient", NLTRANSIENT))


def weak_zones(s):
    """{mechanism: WEAK array} of the spec's weak zones."""
    return {mech: structure.weakzones(s[key], frame(s)) for key, mech in WEAK_KEYS if s.get(key)}


def planes(s):
    """Fault-creep planes of a spec (PLANE array) or None."""
    if not s["planes"]:
        return None
    _, dx2, dx3 = spacing(s)
    ps = [structure.friction_plane(p["x"], p["y"], p["z"], p["length"], p["width"], p["strike"], p["dip"],
                                   p["rake"], dx2, dx3, frame(s)) for p in s["planes"]]
    return np.concatenate(ps)


def event_times(s):
    """The spec's event times: the set-up at t = 0, then its later events
    (s["events"]: a list of (time, deck[, load deck[, tensile deck[, point
    source deck]]]))."""
    return (0.0,) + tuple(float(ev[0]) for ev in s.get("events", ()))


def build(s, okada=True):
    """Create the model of a spec on the GPU with its coseismic state
    (okada=False leaves out Okada's displacement of the sources below the
    Nyquist size)."""
    dx1, dx2, dx3 = spacing(s)
    m = Model(s["sx1"], s["sx2"], s["sx3"], dx1, dx2, dx3, *constants(s), interval=s["interval"], odt=s["odt"],
              tscale=s["tscale"], skip=s["skip"], event_times=event_times(s))
    return configure(m, s, okada)


def run(m, s, nsteps, dt_replay=None):
    """Up to nsteps RK2 steps of a model of spec s, applying each later event
    of the spec when a step reaches its time (the driver's side of the event
    switch, relax.f90:858-886).  Returns the step records."""
    recs = []
    while len(recs) < nsteps:
        k = len(recs)
        got = m.run(nsteps - k, dt_replay=None if dt_replay is None else dt_replay[k:])
        if not got:
            break
        e0 = recs[-1]["e"] if recs else 0
        recs.extend(got)
        if got[-1]["e"] != e0:
            # the step's event index is 0-based; event_sources counts from 1
            m.event(nyquist=s.get("nyquist", NYQUIST), **event_sources(s, got[-1]["e"] + 1))
    return recs


def interseismic(s):
    """The spec's inter-seismic loading as (shear, tensile) SOURCE arrays:
    s["inter"] are deck lines (slip rate, x, y, z, length, width, strike, dip,
    rake), s["inter_tensile"] (opening rate, x, y, z, length, width, strike,
    dip); None without any."""
    if not (s.get("inter") or s.get("inter_tensile")):
        return None
    return structure.interseismic(s.get("inter"), s.get("inter_tensile"), constants(s)[3], spacing(s),
                                  s.get("nyquist", NYQUIST), frame(s))


def event_sources(s, e):
    """The source lists of event e (1 = the set-up at t = 0): {sources,
    loads, tensile, mogi} as SOURCE arrays (shear patches, surface loads,
    tensile cracks, point dilatation sources) and the eigenstrain volumes
    {cuboids (WEAK), tetrahedra (TETRA), vertices ((n, 3) float64)}.  A later
    event of s["events"] is (time, deck[, loads[, tensile[, mogi[, cuboids[,
    tetrahedra[, vertices]]]]]])."""
    if e == 1:
        decks = (s["deck"], s.get("loads"), s.get("tensile"), s.get("mogi"), s.get("cuboids"), s.get("tetrahedra"),
                 s.get("vertices"))
    else:
        ev = tuple(s["events"][e - 2][1:])
        decks = ev + (None,) * (7 - len(ev))
    beta, fr = constants(s)[3], frame(s)
    vx = np.zeros((0, 3)) if not decks[6] else np.asarray(decks[6], np.float64).reshape(-1, 3)
    return dict(sources=structure.sources(decks[0], beta, fr), loads=structure.loads(decks[1], beta, fr[2]),
                tensile=structure.tensile(decks[2], beta, fr), mogi=structure.mogi(decks[3], fr[2]),
                cuboids=structure.cuboids(decks[4], fr), tetrahedra=structure.tetrahedra(decks[5]), vertices=vx)


def configure(m, s, okada=True):
    """Set up a new (or reset) model of the spec's grid: rheology, the
    prestress (its tables, and tau = -prestress before the first event,
    relax.f90:286-297), planes, coseismic state.  The time line (interval,
    odt, skip) is the model's."""
    t = tables(s)
    m.set_prestress(t["prestress"])
    m.set_structure(LINEAR, t["linear"])
    m.set_structure(POWERLAW, t["nonlinear"])
    if t["ltransient"] is not None:
        m.set_structure(LTRANSIENT, t["ltransient"])
    if t["nltransient"] is not None:
        m.set_structure(NLTRANSIENT, t["nltransient"])
    for mech, zones in weak_zones(s).items():
        m.set_weakzones(mech, zones)
    pl = planes(s)
    if pl is not None:
        m.set_planes(pl)
    m.set_structure(FAULTCREEP, t["friction"])
    inter = interseismic(s)
    if inter is not None:
        m.set_interseismic(*inter)
    # the first event at t = 0 (relax.f90:316-371), with Okada's displacement
    # of its sources below the Nyquist size (:357-360)
    nyq = s.get("nyquist", NYQUIST)
    m.event(nyquist=nyq, **event_sources(s, 1))
    if okada:
        sc, tsc = structure.deck_copies(s["deck"], s.get("tensile"))
        m.okada(sc, tsc, nyquist=nyq)
    return m


def okada_sources(s):
    """The deck copies (shear, tensile) of the spec's first event and the
    number of them at or below the Nyquist size (Okada's solution)."""
    sc, tsc = structure.deck_copies(s["deck"], s.get("tensile"))
    dmin = s.get("nyquist", NYQUIST) * min(spacing(s))
    n = sum(1 for a in (sc, tsc) for r in a if dmin >= r["length"] or dmin >= r["width"])
    return sc, tsc, n
