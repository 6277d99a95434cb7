"""Reader of RELAX's input deck (input.f90:43-1767, init): the same sections in
the same order, so that a forward-model request can be a RELAX deck.

Lines whose first non-blank character is '#' are comments (getdata.f:1-30);
blank lines are skipped as well.  Each data line is read list-directed, as
the reference's READ (dataline,*): the leading fields are taken and the rest
of the line is ignored.  Section order (input.f90:252-1754):

  sx1 sx2 sx3 / dx1 dx2 dx3 beta nyquist / x0 y0 rot / oz ozs / output dir /
  lambda mu gamma / interval odt [tscale] /
  observation planes / observation points / stress segments /
  prestress interfaces / linear viscous interfaces [+ weak zones] /
  power-law interfaces [+ weak zones] /
  (--with-transient) linear transient [+ weak zones], power-law transient
  [+ weak zones] /
  fault-creep interfaces [+ afterslip planes] /
  inter-seismic shear segments / inter-seismic tensile segments /
  number of events, then per event: [time (events after the first)],
  shear segments (optional taper beta), tensile segments, point sources,
  (--with-eigenstrain) cuboids [(curvilinear) tetrahedra and vertices],
  surface loads (optional taper beta).

Indices at the start of each item line must run 1, 2, ... as the reference
checks them (STOP 1 there); here they raise DeckError.  Deck.spec() turns
the deck into the workload description relax_amd.workloads builds a model
from.
"""
import dataclasses
import math
import os

import numpy as np

from . import structure


class DeckError(ValueError):
    pass


def data_lines(text):
    """getdata.f: the data lines of a deck, comments (first non-blank
    character '#') and blank lines skipped; yields (line number, text)."""
    for no, line in enumerate(text.splitlines(), 1):
        s = line.strip()
        if not s or s.startswith("#"):
            continue
        yield no, line


class _Reader:
    def __init__(self, text):
        self._it = data_lines(text)
        self.no = 0

    def line(self, what):
        try:
            self.no, s = next(self._it)
        except StopIteration:
            raise DeckError("deck ends before %s" % what)
        return s

    def values(self, what, n, types=float, optional=0):
        """The first n fields of the next data line (n - optional of them
        required), converted; list-directed, extra fields ignored."""
        s = self.line(what)
        f = s.replace(",", " ").split()
        if len(f) < n - optional:
            raise DeckError("line %d (%s): %d values expected, got %r" % (self.no, what, n - optional, s))
        ts = types if isinstance(types, (list, tuple)) else [types] * n
        out = []
        for k in range(min(n, len(f))):
            try:
                out.append(ts[k](_fortran_real(f[k]) if ts[k] is float else f[k]))
            except ValueError:
                raise DeckError("line %d (%s): cannot read %r" % (self.no, what, f[k]))
        return out

    def count(self, what):
        v = self.values(what, 1, int)[0]
        if v < 0:
            raise DeckError("line %d: negative count for %s" % (self.no, what))
        return v

    def items(self, what, n, nfields, types=float, optional=0):
        """n item lines 'i f1 f2 ...' with i = 1..n; returns the field
        tuples (index dropped)."""
        out = []
        ts = [int] + (list(types) if isinstance(types, (list, tuple)) else [types] * nfields)
        for k in range(1, n + 1):
            v = self.values(what, nfields + 1, ts, optional)
            if v[0] != k:
                raise DeckError("line %d (%s): index %d, expected %d" % (self.no, what, v[0], k))
            out.append(tuple(v[1:]))
        return out


def _fortran_real(s):
    """A Fortran real literal (1.0d-3, 2e5, 3.)"""
    return float(s.replace("d", "e").replace("D", "E"))


@dataclasses.dataclass
class Event:
    time: float = 0.0
    shear: list = dataclasses.field(default_factory=list)     # (slip, x, y, z, length, width, strike, dip, rake[, beta])
    tensile: list = dataclasses.field(default_factory=list)   # (opening, x, y, z, length, width, strike, dip)
    mogi: list = dataclasses.field(default_factory=list)      # (volume change, x, y, z)
    cuboids: list = dataclasses.field(default_factory=list)   # (e11, e12, e13, e22, e23, e33, x, y, z, L, W, T, strike, dip)
    tetrahedra: list = dataclasses.field(default_factory=list)  # (e11..e33, i1, i2, i3, i4)
    vertices: list = dataclasses.field(default_factory=list)  # (x1, x2, x3)
    loads: list = dataclasses.field(default_factory=list)     # (x, y, length, width, t3, period, phase[, beta])


@dataclasses.dataclass
class Deck:
    sx1: int = 0
    sx2: int = 0
    sx3: int = 0
    dx1: float = 0.0
    dx2: float = 0.0
    dx3: float = 0.0
    beta: float = 0.2
    nyquist: float = 2.0
    x0: float = 0.0
    y0: float = 0.0
    rot: float = 0.0
    oz: float = 0.0
    ozs: float = 0.0
    wdir: str = ""
    lam: float = 0.0
    mu: float = 0.0
    gam: float = 0.0
    interval: float = 0.0
    odt: float = 0.0
    skip: int = 0
    tscale: float = 1.0
    planes_obs: list = dataclasses.field(default_factory=list)   # (x, y, z, length, width, strike, dip)
    points: list = dataclasses.field(default_factory=list)       # (name, x1, x2, x3)
    stress_segments: list = dataclasses.field(default_factory=list)  # (x, y, z, L, W, strike, dip, friction)
    prestress: list = dataclasses.field(default_factory=list)    # (z, s11, s12, s13, s22, s23, s33)
    linear: list = dataclasses.field(default_factory=list)       # (z, gammadot0, cohesion)
    weak_linear: list = dataclasses.field(default_factory=list)  # (dg, x, y, z, L, W, T, strike, dip)
    nonlinear: list = dataclasses.field(default_factory=list)    # (z, gammadot0, power, cohesion)
    weak_nonlinear: list = dataclasses.field(default_factory=list)
    ltransient: list = dataclasses.field(default_factory=list)   # (z, gammadot0, Gk)
    weak_ltransient: list = dataclasses.field(default_factory=list)
    nltransient: list = dataclasses.field(default_factory=list)  # (z, gammadot0, power, Gk)
    weak_nltransient: list = dataclasses.field(default_factory=list)
    faultcreep: list = dataclasses.field(default_factory=list)   # (z, v0, (a-b)sigma, friction, cohesion)
    afterslip: list = dataclasses.field(default_factory=list)    # (x, y, z, length, width, strike, dip, rake)
    inter: list = dataclasses.field(default_factory=list)        # (slip rate, x, y, z, L, W, strike, dip, rake)
    inter_tensile: list = dataclasses.field(default_factory=list)  # (opening rate, x, y, z, L, W, strike, dip)
    events: list = dataclasses.field(default_factory=list)

    # ------------------------------------------------------------------
    def spec(self, name="deck", nsteps=None):
        """The workload description (relax_amd.workloads) of the deck: grid,
        spacing, elastic constants, taper, Nyquist gate, frame, rheology
        interfaces in the layer form viscoelasticstructure takes, prestress,
        planes, inter-seismic loading and the events."""
        def layers(rows, kind):
            out = []
            for r in rows:
                if kind == "linear":        # z, gammadot0, cohesion; exponent 1 (input.f90:511)
                    z, g, c = r
                    out.append((z, g, 1.0, c, 0.6, 0.0))
                elif kind == "nonlinear":   # z, gammadot0, power, cohesion
                    z, g, n, c = r
                    out.append((z, g, n, c, 0.6, 0.0))
                elif kind == "ltransient":  # z, gammadot0, Gk; exponent 1
                    z, g, gk = r
                    out.append((z, g, 1.0, 0.0, 0.6, gk))
                elif kind == "nltransient":  # z, gammadot0, power, Gk
                    z, g, n, gk = r
                    out.append((z, g, n, 0.0, 0.6, gk))
                else:                       # fault creep: z, v0, (a-b)sigma, friction, cohesion
                    z, v0, ab, fr, c = r
                    out.append((z, v0, ab, c, fr, 0.0))
            return out or None

        s = dict(name=name, sx1=self.sx1, sx2=self.sx2, sx3=self.sx3, dx=(self.dx1, self.dx2, self.dx3),
                 lam=self.lam, mu=self.mu, gam=self.gam, beta=self.beta, nyquist=self.nyquist, x0=self.x0,
                 y0=self.y0, rot=self.rot, interval=self.interval, odt=self.odt,
                 skip=self.skip, tscale=self.tscale, nsteps=nsteps,
                 linear=layers(self.linear, "linear"), nonlinear=layers(self.nonlinear, "nonlinear"),
                 ltransient=layers(self.ltransient, "ltransient"),
                 nltransient=layers(self.nltransient, "nltransient"),
                 friction=layers(self.faultcreep, "faultcreep") if self.afterslip else None,
                 prestress_layers=list(self.prestress) or None,
                 planes=[dict(x=x, y=y, z=z, length=L, width=W, strike=st, dip=dp, rake=rk)
                         for (x, y, z, L, W, st, dp, rk) in self.afterslip] or None,
                 inter=list(self.inter) or None, inter_tensile=list(self.inter_tensile) or None,
                 weak_linear=list(self.weak_linear) or None, weak_nonlinear=list(self.weak_nonlinear) or None,
                 weak_ltransient=list(self.weak_ltransient) or None,
                 weak_nltransient=list(self.weak_nltransient) or None,
                 points=[(x1, x2, x3) for (_, x1, x2, x3) in self.points] or None)
        e0 = self.events[0] if self.events else Event()
        s.update(deck=list(e0.shear), loads=list(e0.loads) or None, tensile=list(e0.tensile) or None,
                 mogi=list(e0.mogi) or None, cuboids=list(e0.cuboids) or None,
                 tetrahedra=list(e0.tetrahedra) or None, vertices=list(e0.vertices) or None)
        s["events"] = [(e.time, list(e.shear), list(e.loads), list(e.tensile), list(e.mogi), list(e.cuboids),
                        list(e.tetrahedra), list(e.vertices)) for e in self.events[1:]]
        if not s["events"]:
            del s["events"]
        return s


def read(source, with_transient=False, with_eigenstrain=False, curvilinear=False):
    """Read a deck (its text, or the path of a file holding it).  The
    options are the command-line switches that add sections: --with-transient
    (input.f90:740-998), --with-eigenstrain (:1496-1671) and --curvilinear
    (tetrahedra, :1586-1668)."""
    text = source
    if "\n" not in source and os.path.exists(source):
        with open(source) as f:
            text = f.read()
    r = _Reader(text)
    d = Deck()
    d.sx1, d.sx2, d.sx3 = r.values("grid dimension", 3, int)
    d.dx1, d.dx2, d.dx3, d.beta, d.nyquist = r.values("sampling and smoothing", 5)
    d.x0, d.y0, d.rot = r.values("origin and rotation", 3)
    d.oz, d.ozs = r.values("observation depths", 2)
    d.wdir = r.line("output directory").strip()
    d.lam, d.mu, d.gam = r.values("elastic parameters", 3)
    t = r.values("time interval", 3, optional=1)
    d.interval, d.odt = t[0], t[1]
    if d.odt < 0:
        if len(t) < 3:
            raise DeckError("line %d: a negative step needs the time scaling" % r.no)
        d.skip = int(math.ceil(-d.odt))
        d.tscale = t[2]
    if min(d.sx1, d.sx2, d.sx3) <= 0 or min(d.dx1, d.dx2, d.dx3) <= 0:
        raise DeckError("grid dimensions and sampling must be positive")
    d.planes_obs = r.items("observation plane", r.count("observation planes"), 7)
    d.points = r.items("observation point", r.count("observation points"), 4, [str, float, float, float])
    d.stress_segments = r.items("stress segment", r.count("stress segments"), 8)
    d.prestress = r.items("prestress interface", r.count("prestress interfaces"), 7)
    nv = r.count("linear viscous interfaces")
    if nv:
        d.linear = _rates(r.items("linear interface", nv, 3), r)
        d.weak_linear = r.items("linear weak zone", r.count("linear weak zones"), 9)
    npl = r.count("power-law interfaces")
    if npl:
        d.nonlinear = _rates(r.items("power-law interface", npl, 4), r)
        d.weak_nonlinear = r.items("power-law weak zone", r.count("power-law weak zones"), 9)
    if with_transient:
        nlt = r.count("linear transient interfaces")
        if nlt:
            d.ltransient = _rates(r.items("linear transient interface", nlt, 3), r)
            d.weak_ltransient = r.items("linear transient weak zone", r.count("linear transient weak zones"), 9)
        nnlt = r.count("power-law transient interfaces")
        if nnlt:
            d.nltransient = _rates(r.items("power-law transient interface", nnlt, 4), r)
            d.weak_nltransient = r.items("power-law transient weak zone", r.count("power-law transient weak zones"),
                                         9)
    nfc = r.count("fault creep interfaces")
    if nfc:
        d.faultcreep = _rates(r.items("fault creep interface", nfc, 5), r)
        d.afterslip = r.items("afterslip plane", r.count("afterslip planes"), 8)
    d.inter = r.items("inter-seismic shear segment", r.count("inter-seismic shear segments"), 9)
    d.inter_tensile = r.items("inter-seismic tensile segment", r.count("inter-seismic tensile segments"), 8)
    ne = r.count("events")
    for e in range(ne):
        ev = Event()
        if e > 0:
            ev.time = r.values("time of the next event", 1)[0]
            if d.skip == 0:
                # a multiple of the output step (input.f90:1264-1267)
                ev.time = _nint(ev.time / d.odt) * d.odt
            if ev.time <= d.events[-1].time:
                raise DeckError("line %d: event times must increase" % r.no)
        ev.shear = r.items("shear segment", r.count("shear segments"), 10, optional=1)
        for line in ev.shear:
            if len(line) > 9 and line[9] > 0.5e8:
                raise DeckError("invalid smoothing parameter (beta)")
            if max(line[4], line[5]) <= 0:
                raise DeckError("shear segment lengths must be positive")
        ev.tensile = r.items("tensile segment", r.count("tensile segments"), 8)
        ev.mogi = r.items("point source", r.count("point sources"), 4)
        if with_eigenstrain:
            ev.cuboids = r.items("cuboid", r.count("cuboids"), 14)
            if curvilinear:
                ntet = r.count("tetrahedra")
                if ntet:
                    ev.tetrahedra = [tuple(v[:6]) + tuple(int(i) for i in v[6:]) for v in
                                     r.items("tetrahedron", ntet, 10)]
                    nvx = r.count("tetrahedron vertices")
                    if nvx == 0:
                        raise DeckError("line %d: tetrahedra need vertices" % r.no)
                    ev.vertices = r.items("vertex", nvx, 3)
        ev.loads = r.items("surface load", r.count("surface loads"), 8, optional=1)
        for line in ev.loads:
            if len(line) > 7 and line[7] > 0.5e8:
                raise DeckError("invalid smoothing parameter (beta)")
        d.events.append(ev)
    # something to deform the first event, or a loading interval
    # (input.f90:1740-1754)
    e0 = d.events[0] if d.events else Event()
    if not (e0.shear or e0.tensile or e0.mogi or e0.loads or e0.cuboids or e0.tetrahedra) and d.interval <= 0:
        raise DeckError("no input dislocations or dilatation point sources or surface tractions for first event")
    return d


def _nint(x):
    """Fortran NINT: to the nearest integer, halves away from zero."""
    return math.floor(x + 0.5) if x >= 0 else -math.floor(-x + 0.5)


def _rates(rows, r):
    """Interfaces whose rate (second field) must not be negative
    (input.f90:522-527 and the like)."""
    for row in rows:
        if row[1] < 0:
            raise DeckError("near line %d: strain rates must be positive" % r.no)
    return rows
