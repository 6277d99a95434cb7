"""Input-side helpers that turn a RELAX description into the per-depth tables
and internal-convention fault geometry the hot path consumes."""
import math

import numpy as np

from ._lib import LAYER, PLANE, SOURCE, TENSOR_LAYER, TETRA, WEAK

DEG2RAD = 0.01745329251994329547437168059786927


def fix(number):
    """fix, elastic3d.f90:70-83: rounds half down."""
    f, c = math.floor(number), math.ceil(number)
    return c if (number - f) > 0.5 else f


def viscoelasticstructure(layers, sx3, dx3):
    """viscoelasticstructure, elastic3d.f90:732-829.

    layers: sequence of (z, gammadot0, stressexponent, cohesion, friction, Gk)
    interfaces; returns a LAYER array of sx3 entries sampled every dx3 with
    linear interpolation between interfaces projected onto multiples of dx3
    and the last layer continued downward (in single precision, as the
    reference's REAL() there)."""
    lay = np.asarray([tuple(l) for l in layers], LAYER)
    nv = len(lay)
    if nv <= 0:
        raise ValueError("invalid elastic structure")
    vs = np.zeros(sx3, LAYER)
    vs["friction"] = 0.6
    vs["stressexponent"] = lay[0]["stressexponent"]
    z0 = fix(lay[0]["z"] / dx3) * dx3
    i3e = 1
    keys = ("gammadot0", "stressexponent", "friction", "cohesion", "Gk")
    v0 = dict.fromkeys(keys, 0.0)
    v1 = dict.fromkeys(keys, 0.0)
    z1 = z0
    for k in range(nv):
        z1 = fix(lay[k]["z"] / dx3) * dx3
        if z1 < z0:
            raise ValueError("invalid mechanical structure")
        for key in keys:
            v1[key] = float(lay[k][key])
        if z1 == z0:
            i3e = fix(z1 / dx3 + 1)
        else:
            i3s = fix(z0 / dx3) + 1
            i3e = min(fix(z1 / dx3 + 1), sx3)
            for i3 in range(i3s, i3e + 1):
                z = (i3 - 1.0) * dx3
                for key in keys:
                    vs[i3 - 1][key] = ((z - z0) * v1[key] + (z1 - z) * v0[key]) / (z1 - z0)
        z0 = z1
        v0 = dict(v1)
    if fix(z1 / dx3) < sx3 - 1:
        for key in keys:
            vs[key][i3e - 1:] = float(np.float32(v1[key]))
    return vs


def wangconvention(slip, x, y, z, length, width, strike, dip, rake, x0=0.0, y0=0.0, rot=0.0):
    """wangconvention, input.f90:1780-1800: deck geometry (upper corner,
    angles in degrees from north / horizontal) -> internal convention (patch
    centre, radians, dip measured from the vertical)."""
    slip = -slip
    strike = -90.0 - strike
    dip = 90.0 - dip
    strike, dip, rake = strike * DEG2RAD, dip * DEG2RAD, rake * DEG2RAD
    x = x - x0 - length / 2.0 * math.sin(strike) + width / 2.0 * math.sin(dip) * math.cos(strike)
    y = y - y0 - length / 2.0 * math.cos(strike) - width / 2.0 * math.sin(dip) * math.sin(strike)
    z = z + width / 2.0 * math.cos(dip)
    x, y = rotation(x, y, rot)
    strike = strike + rot * DEG2RAD
    return slip, x, y, z, strike, dip, rake


def rotation(x, y, rot):
    """rotation, input.f90:1809-1822: a point into the computational frame
    turned by rot degrees."""
    if rot == 0.0:
        return x, y
    a = rot * DEG2RAD
    return x * math.cos(a) + y * math.sin(a), -x * math.sin(a) + y * math.cos(a)


def tensorstructure(layers, sx3, dx3):
    """tensorstructure, elastic3d.f90:630-702: a prestress profile sampled
    every dx3.

    layers: sequence of (z, s11, s12, s13, s22, s23, s33) interfaces; returns
    a TENSOR_LAYER array of sx3 entries, zero above the first interface,
    interpolated linearly between interfaces projected onto multiples of dx3
    (in single precision, as the reference's REAL() and TENSOR arithmetic
    there) and the last interface continued downward."""
    lay = [tuple(float(v) for v in l) for l in layers]
    if not lay:
        raise ValueError("invalid tensor structure")
    f32 = np.float32
    vs = np.zeros(sx3, TENSOR_LAYER)
    z0 = fix(lay[0][0] / dx3) * dx3
    t0 = t1 = np.zeros(6, f32)
    i3e = 1
    z1 = z0
    for l in lay:
        z1 = fix(l[0] / dx3) * dx3
        if z1 < z0:
            raise ValueError("invalid mechanical structure")
        t1 = np.asarray(l[1:7], f32)
        if z1 == z0:
            i3e = fix(z1 / dx3 + 1)
        else:
            i3s = fix(z0 / dx3) + 1
            i3e = min(fix(z1 / dx3 + 1), sx3)
            r = f32(1.0 / (z1 - z0))
            for i3 in range(i3s, i3e + 1):
                z = (i3 - 1.0) * dx3
                vs["t"][i3 - 1] = r * (f32(z - z0) * t1 + f32(z1 - z) * t0)
        z0 = z1
        t0 = t1
    if fix(z1 / dx3) < sx3 - 1:
        vs["t"][i3e - 1:] = t1
    return vs


def patches(deck):
    """Convert deck fault lines (slip, x, y, z, length, width, strike, dip,
    rake) to a PLANE array in the internal convention and the slips."""
    out = np.zeros(len(deck), PLANE)
    slips = np.zeros(len(deck))
    for k, (s, x, y, z, L, W, st, dp, rk) in enumerate(deck):
        s2, x2, y2, z2, st2, dp2, rk2 = wangconvention(s, x, y, z, L, W, st, dp, rk)
        out[k] = (x2, y2, z2, W, L, st2, dp2, rk2, 0, 0)
        slips[k] = s2
    return out, slips


def sources(deck, beta, frame=(0.0, 0.0, 0.0)):
    """Convert deck fault lines (slip, x, y, z, length, width, strike, dip,
    rake[, beta]) to the SOURCE array of an event as the parser leaves it
    (input.f90:1297-1370: Wang's convention in the frame (x0, y0, rot), the
    line's own taper or else the default beta)."""
    out = np.zeros(len(deck or ()), SOURCE)
    for k, line in enumerate(deck or ()):
        s, x, y, z, L, W, st, dp, rk = line[:9]
        b = line[9] if len(line) > 9 else beta
        s2, x2, y2, z2, st2, dp2, rk2 = wangconvention(s, x, y, z, L, W, st, dp, rk, *frame)
        out[k] = (s2, x2, y2, z2, W, L, st2, dp2, rk2, 0.0, 0.0, b, 0.0)
    return out


def antialiasingfilter(slip, length, width, dx1, dx2, dx3, nyquist):
    """antialiasingfilter, input.f90:1836-1856: a segment shorter or narrower
    than nyquist*min(dx) is widened to that size with its slip scaled down in
    proportion.  Returns (slip, length, width)."""
    dx = min(dx1, dx2, dx3) * nyquist
    if length < dx:
        slip = slip * length / dx
        length = dx
    if width < dx:
        slip = slip * width / dx
        width = dx
    return slip, length, width


def interseismic(deck, tensile_deck, beta, dx, nyquist, frame=(0.0, 0.0, 0.0)):
    """The inter-seismic loading of a deck (input.f90:1114-1250): shear
    segment lines (slip rate, x, y, z, length, width, strike, dip, rake) and
    tensile segment lines (opening rate, x, y, z, length, width, strike, dip)
    as SOURCE arrays.  Each segment goes through antialiasingfilter, then
    Wang's convention.  For a tensile segment the filter scales its slip
    field, which is zero, so only its size changes and the opening is kept
    (input.f90:1205, :1233-1236).  The reference never sets the taper of
    these segments (types.f90 SOURCE_STRUCT%beta is left as ALLOCATE leaves
    it); they take the deck's taper beta here, as the coseismic segments do
    (input.f90:1316).  dx is (dx1, dx2, dx3)."""
    dx1, dx2, dx3 = dx
    shear = np.zeros(len(deck or ()), SOURCE)
    for k, (s, x, y, z, L, W, st, dp, rk) in enumerate(deck or ()):
        s, L, W = antialiasingfilter(s, L, W, dx1, dx2, dx3, nyquist)
        s2, x2, y2, z2, st2, dp2, rk2 = wangconvention(s, x, y, z, L, W, st, dp, rk, *frame)
        shear[k] = (s2, x2, y2, z2, W, L, st2, dp2, rk2, 0.0, 0.0, beta, 0.0)
    tens = np.zeros(len(tensile_deck or ()), SOURCE)
    for k, (op, x, y, z, L, W, st, dp) in enumerate(tensile_deck or ()):
        _, L, W = antialiasingfilter(0.0, L, W, dx1, dx2, dx3, nyquist)
        _, x2, y2, z2, st2, dp2, _ = wangconvention(0.0, x, y, z, L, W, st, dp, 0.0, *frame)
        tens[k] = (0.0, x2, y2, z2, W, L, st2, dp2, 0.0, 0.0, 0.0, beta, op)
    return shear, tens


def deck_copies(deck, tensile_deck=None):
    """The deck's own copies of an event's shear and tensile sources, as
    input.f90 keeps them beside the converted ones (in%events(e)%sc and %tsc,
    :1319, :1410): upper corner at the start of the strike, north-east-down,
    strike, dip and rake in degrees, the slip as given.  Okada's solution for
    the sources below the Nyquist size reads these (relax.f90:1300-1349).
    Returns (shear, tensile) SOURCE arrays."""
    sh = np.zeros(len(deck or ()), SOURCE)
    for k, line in enumerate(deck or ()):
        s, x, y, z, L, W, st, dp, rk = line[:9]
        sh[k] = (s, x, y, z, W, L, st, dp, rk, 0.0, 0.0, 0.0, 0.0)
    te = np.zeros(len(tensile_deck or ()), SOURCE)
    for k, (op, x, y, z, L, W, st, dp) in enumerate(tensile_deck or ()):
        te[k] = (0.0, x, y, z, W, L, st, dp, 0.0, 0.0, 0.0, 0.0, op)
    return sh, te


def loads(deck, beta, rot=0.0):
    """Convert deck surface-load lines (x, y, length, width, t3, period,
    phase[, beta]) to the SOURCE array of an event (input.f90:1686-1733: the
    line's own taper or else the default beta; the centre turned by the
    frame's rotation rot, with no shift of origin)."""
    out = np.zeros(len(deck or ()), SOURCE)
    for k, line in enumerate(deck or ()):
        x, y, L, W, amp, period, phase = line[:7]
        b = line[7] if len(line) > 7 else beta
        x, y = rotation(x, y, rot)
        out[k] = (amp, x, y, 0.0, W, L, 0.0, 0.0, 0.0, period, phase, b, 0.0)
    return out


def tensile(deck, beta, frame=(0.0, 0.0, 0.0)):
    """Convert deck tensile-crack lines (opening, x, y, z, length, width,
    strike, dip) to a SOURCE array as the parser leaves them
    (input.f90:1398-1442: Wang's convention of the geometry in the frame
    (x0, y0, rot); slip holds the negated opening, the opening is kept)."""
    out = np.zeros(len(deck or ()), SOURCE)
    for k, (op, x, y, z, L, W, st, dp) in enumerate(deck or ()):
        s2, x2, y2, z2, st2, dp2, _ = wangconvention(op, x, y, z, L, W, st, dp, 0.0, *frame)
        out[k] = (s2, x2, y2, z2, W, L, st2, dp2, 0.0, 0.0, 0.0, beta, op)
    return out


def mogi(deck, rot=0.0):
    """Convert deck point-source lines (volume change, x, y, z) to a SOURCE
    array (input.f90:1470-1490: the position turned by the frame's rotation
    rot, with no shift of origin)."""
    out = np.zeros(len(deck or ()), SOURCE)
    for k, (o, x, y, z) in enumerate(deck or ()):
        x, y = rotation(x, y, rot)
        out[k]["slip"], out[k]["x"], out[k]["y"], out[k]["z"] = o, x, y, z
    return out


def friction_plane(x, y, z, length, width, strike, dip, rake, dx2, dx3, frame=(0.0, 0.0, 0.0)):
    """A fault-creep plane as the parser builds it (input.f90:1079-1093): the
    rake is turned by 180 degrees, the geometry follows Wang's convention in
    the frame (x0, y0, rot) and the patch grid is px2 = fix(length/dx2) by
    px3 = fix(width/dx3)."""
    rake = rake - 180.0 if rake >= 0.0 else rake + 180.0
    _, x2, y2, z2, st2, dp2, rk2 = wangconvention(0.0, x, y, z, length, width, strike, dip, rake, *frame)
    p = np.zeros(1, PLANE)
    p[0] = (x2, y2, z2, width, length, st2, dp2, rk2, fix(length / dx2), fix(width / dx3))
    return p


def weakzones(deck, frame=(0.0, 0.0, 0.0)):
    """Convert deck weak-zone lines (dgammadot0, x, y, z, length, width,
    thickness, strike, dip) to a WEAK array in the internal convention, as
    the parser does (input.f90:550-583: Wang's convention of the position,
    strike and dip in the frame (x0, y0, rot); the thickness is kept)."""
    out = np.zeros(len(deck), WEAK)
    for k, (dg, x, y, z, L, W, T, st, dp) in enumerate(deck):
        _, x2, y2, z2, st2, dp2, _ = wangconvention(0.0, x, y, z, L, W, st, dp, 0.0, *frame)
        out[k]["dgammadot0"] = dg
        out[k]["x"], out[k]["y"], out[k]["z"] = x2, y2, z2
        out[k]["length"], out[k]["width"], out[k]["thickness"] = L, W, T
        out[k]["strike"], out[k]["dip"] = st2, dp2
    return out


def cuboids(deck, frame=(0.0, 0.0, 0.0)):
    """Convert deck cuboid lines (e11, e12, e13, e22, e23, e33, x, y, z,
    length, width, thickness, strike, dip) to a WEAK array as the parser
    leaves them (input.f90:1507-1562: the eigenstrain in e, Wang's convention
    of the position, strike and dip in the frame (x0, y0, rot))."""
    out = np.zeros(len(deck or ()), WEAK)
    for k, line in enumerate(deck or ()):
        e = line[:6]
        x, y, z, L, W, T, st, dp = line[6:14]
        _, x2, y2, z2, st2, dp2, _ = wangconvention(0.0, x, y, z, L, W, st, dp, 0.0, *frame)
        out[k]["x"], out[k]["y"], out[k]["z"] = x2, y2, z2
        out[k]["length"], out[k]["width"], out[k]["thickness"] = L, W, T
        out[k]["strike"], out[k]["dip"] = st2, dp2
        out[k]["e"] = e
    return out


def tetrahedra(deck):
    """Convert deck tetrahedron lines (e11, e12, e13, e22, e23, e33, i1, i2,
    i3, i4) to a TETRA array (input.f90:1598-1627; the indices are 1-based
    into the event's vertices, which the frame does not move)."""
    out = np.zeros(len(deck or ()), TETRA)
    for k, line in enumerate(deck or ()):
        out[k]["e"] = line[:6]
        out[k]["i1"], out[k]["i2"], out[k]["i3"], out[k]["i4"] = (int(i) for i in line[6:10])
    return out
