"""Shared helpers of the GPU parity tests (not collected by pytest)."""
import numpy as np

import oracle as O
from relax_amd import structure, workloads
from relax_amd.model import LINEAR, LTRANSIENT, NLTRANSIENT, POWERLAW


def rel_l2(a, b):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    nb = np.linalg.norm(b)
    return np.linalg.norm(a - b) / (nb if nb > 0 else 1.0)


def small_spec(kind, sx=64, sx3=64, dx=1.0, seed=0):
    """Oracle-sized variants of the workloads: same physics, smaller grid."""
    rng = np.random.default_rng(1000 + seed)
    if kind == "maxwell":
        s = workloads.spec("c0b")
        s.update(sx1=sx, sx2=sx, sx3=sx3, dx=dx, linear=[(6.0, 0.2, 1.0, 0.0, 0.6, 0.0)], interval=1e9, skip=1,
                 deck=[(1e-3, -5.0, 0.0, 0.0, 10.0, 6.0, 0.0, 90.0, 0.0)])
        return s
    if kind == "powerlaw":
        deck = [workloads._deck_patch(rng, -10, 10, 5.0, 4.0, 3e-3) for _ in range(4)]
        lay = [(8.0 + k, 5e6 * np.exp(k / 10.0), 3.0, 0.0, 0.6, 0.0) for k in range(20)]
        return dict(name="powerlaw", sx1=sx, sx2=sx, sx3=sx3, dx=dx, linear=None, nonlinear=lay, friction=None,
                    planes=None, deck=deck, interval=1e9, odt=1.0, skip=1, tscale=1.0, nsteps=4)
    if kind == "afterslip":
        dip = 20.0
        deck = [(rng.uniform(0, 1e-4), -6.0 + j2, -j3 * np.cos(np.radians(dip)), j3 * np.sin(np.radians(dip)),
                 1.0, 1.0, 0.0, dip, 90.0) for j3 in range(6) for j2 in range(12)]
        plane = dict(x=-12.0, y=-6 * np.cos(np.radians(dip)), z=6 * np.sin(np.radians(dip)), length=24.0,
                     width=12.0, strike=0.0, dip=dip, rake=90.0)
        return dict(name="afterslip", sx1=sx, sx2=sx, sx3=sx3, dx=dx, linear=None,
                    nonlinear=[(12.0, 4e7, 3.5, 0.0, 0.6, 0.0)], friction=[(0.0, 1e-3, 1.0, 0.0, 0.6, 0.0)],
                    planes=[plane], deck=deck, interval=1e9, odt=1.0, skip=1, tscale=1.0, nsteps=4, nyquist=0.5)
    if kind == "transient":
        # linear Maxwell below 20 km, linear Burgers transient below 6 km and
        # a power-law Burgers transient below 12 km, with weak zones on the
        # Maxwell and the linear transient mechanisms
        deck = [workloads._deck_patch(rng, -10, 10, 5.0, 5.0, 2e-3) for _ in range(3)]
        return dict(name="transient", sx1=sx, sx2=sx, sx3=sx3, dx=dx, linear=[(20.0, 0.05, 1.0, 0.0, 0.6, 0.0)],
                    nonlinear=None, friction=None, planes=None, deck=deck, interval=1e9, odt=1.0, skip=1,
                    tscale=1.0, nsteps=4,
                    ltransient=[(6.0, 0.4, 1.0, 0.0, 0.6, 1.5e4)],
                    nltransient=[(12.0, 3e7, 3.0, 0.0, 0.6, 2.0e4)],
                    weak_linear=[(0.3, -4.0, -2.0, 4.0, 8.0, 6.0, 5.0, 30.0, 60.0)],
                    weak_ltransient=[(0.8, 3.0, 2.0, 2.0, 6.0, 4.0, 6.0, 0.0, 90.0)])
    if kind == "ltransient":
        deck = [workloads._deck_patch(rng, -10, 10, 5.0, 5.0, 2e-3) for _ in range(3)]
        return dict(name="ltransient", sx1=sx, sx2=sx, sx3=sx3, dx=dx, linear=None, nonlinear=None,
                    friction=None, planes=None, deck=deck, interval=1e9, odt=1.0, skip=1, tscale=1.0, nsteps=4,
                    ltransient=[(4.0, 0.5, 1.0, 0.0, 0.6, 2.0e4)])
    raise KeyError(kind)


def oracle_model(s):
    t = workloads.tables(s)
    pl = workloads.planes(s)
    dx1, dx2, dx3 = workloads.spacing(s)
    lam, mu, gam, beta = workloads.constants(s)
    m = dict(sx1=s["sx1"], sx2=s["sx2"], sx3=s["sx3"], dx1=dx1, dx2=dx2, dx3=dx3, lam=lam,
             mu=mu, gam=gam, beta=beta, linear=t["linear"], nonlinear=t["nonlinear"],
             faultcreep=t["friction"], prestress=t["prestress"], interval=s["interval"], odt=s["odt"],
             tscale=s["tscale"], skip=s["skip"], event_times=list(workloads.event_times(s)),
             ltransient=t["ltransient"], nltransient=t["nltransient"])
    m["nyquist"] = s.get("nyquist", workloads.NYQUIST)
    inter = workloads.interseismic(s)
    if inter is not None:
        m["inter"], m["inter_tensile"] = inter
    m["loads0"] = workloads.event_sources(s, 1)["loads"].astype(O.SOURCE)
    if s.get("events"):
        srcs = [workloads.event_sources(s, e) for e in range(2, len(s["events"]) + 2)]
        for key, name in (("event_sources", "sources"), ("event_loads", "loads"), ("event_tensile", "tensile"),
                          ("event_mogi", "mogi")):
            m[key] = [d[name].astype(O.SOURCE) for d in srcs]
        m["event_cuboids"] = [d["cuboids"].astype(O.WEAK) for d in srcs]
        m["event_tetrahedra"] = [d["tetrahedra"].astype(O.TETRA) for d in srcs]
        m["event_vertices"] = [d["vertices"] for d in srcs]
    h3 = s["sx3"] // 2
    keys = {LINEAR: "dg_linear", POWERLAW: "dg_nonlinear", LTRANSIENT: "dg_ltransient",
            NLTRANSIENT: "dg_nltransient"}
    for mech, zones in workloads.weak_zones(s).items():
        # the model builds the field with taper 0, as relax.f90:487-541 does
        m[keys[mech]] = O.builddgammadot0(s["sx1"], s["sx2"], h3, dx1, dx2, dx3, 0.0,
                                          zones.astype(O.WEAK), np.float32)
    if pl is not None:
        m["planes"] = np.ascontiguousarray(pl.astype(O.PLANE))
        m["patches"] = np.zeros(int(sum(int(p["px2"]) * int(p["px3"]) for p in pl)), O.PATCH)
    return m


def oracle_initial_state(s, dtype=np.float32, okada=True):
    """The first event at t = 0 (relax.f90:316-371) on tau = -prestress
    (:286-297): dislocations and traction of the deck's patches and loads, u
    from the Green's function of their body force, sig = -tau + C:eps(u) --
    the set-up the model's event performs on the device."""
    sx1, sx2, sx3 = s["sx1"], s["sx2"], s["sx3"]
    dx1, dx2, dx3 = workloads.spacing(s)
    lam, mu, gam, beta = workloads.constants(s)
    h3 = sx3 // 2
    tau = np.zeros((h3, sx2, sx1, 6), dtype)
    pre = workloads.tables(s)["prestress"]["t"]
    tau[...] = -pre[:, None, None, :]
    c = [np.zeros((sx3, sx2, sx1 + 2), dtype) for _ in range(3)]
    t = [np.zeros((sx2, sx1 + 2), dtype) for _ in range(3)]
    ev = workloads.event_sources(s, 1)
    src = {k: ev[k].astype(O.SOURCE) for k in ("sources", "loads", "tensile", "mogi")}
    O.dislocations(src["sources"], src["loads"], mu, s.get("nyquist", workloads.NYQUIST), 0.0, dx1, dx2, dx3, c,
                   t, tau, lam=lam, beta=beta, tensile=src["tensile"], mogi=src["mogi"],
                   cuboids=ev["cuboids"].astype(O.WEAK), tetrahedra=ev["tetrahedra"].astype(O.TETRA),
                   vertices=ev["vertices"])
    O.greenfunctioncowling(c[0], c[1], c[2], t[0], t[1], t[2], dx1, dx2, dx3, lam, mu, gam)
    u = [np.ascontiguousarray(a[:h3]) for a in c]
    # Okada's displacement of the sources below the Nyquist size (relax.f90:357-360)
    if okada:
        sc, tsc = structure.deck_copies(s["deck"], s.get("tensile"))
        O.dislocations_disp(sc.astype(O.SOURCE), tsc.astype(O.SOURCE), lam, mu,
                            s.get("nyquist", workloads.NYQUIST), dx1, dx2, dx3, *u)
    sig = -tau
    sig = np.ascontiguousarray(sig)
    O.stressupdate(u[0], u[1], u[2], lam, mu, dx1, dx2, dx3, sig)
    st = dict(u1=u[0], u2=u[1], u3=u[2], tau=tau, sig=sig, gamma=np.zeros_like(u[0]), t=0.0, oi=1, i=0)
    if s.get("ltransient") or s.get("nltransient"):
        st["epsilonik"] = np.zeros((h3, sx2, sx1, 6), dtype)
    return st
