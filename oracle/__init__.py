"""ctypes front end of the CPU oracle (test infrastructure only).

The oracle is a C++ restatement of RELAX's per-time-step solve
(oracle/relax_oracle.cpp).  It is the checker the HIP library is compared
against in tests/, in __graft_entry__.smoke() and in the cpu_baseline leg of
bench.py; the product never imports it.

Array conventions (numpy, C order, so the flat memory equals the reference's
Fortran column-major arrays):
  vector field  (sx3, sx2, sx1+2) float32/float64   -- REAL*4 (sx1+2,sx2,sx3)
  plane         (sx2, sx1+2)
  tensor field  (sx3, sx2, sx1, 6)                  -- TYPE(TENSOR) (sx1,sx2,sx3)
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None

LAYER = np.dtype([("z", "f8"), ("gammadot0", "f8"), ("stressexponent", "f8"),
                  ("cohesion", "f8"), ("friction", "f8"), ("Gk", "f8")])
TENSOR_LAYER = np.dtype([("z", "f4"), ("dum", "f4"), ("t", "f4", (6,))])
PATCH = np.dtype([("x1", "f8"), ("x2", "f8"), ("x3", "f8"), ("lx", "f8"), ("lz", "f8"),
                  ("slip", "f8"), ("ss", "f8"), ("ds", "f8"), ("v", "f8"), ("vss", "f8"),
                  ("vds", "f8"), ("taus", "f8"), ("sig", "f4", (6,))])
PLANE = np.dtype([("x", "f8"), ("y", "f8"), ("z", "f8"), ("width", "f8"), ("length", "f8"),
                  ("strike", "f8"), ("dip", "f8"), ("rake", "f8"), ("px2", "i4"), ("px3", "i4")])
# WEAK_STRUCT, types.f90:46-50
WEAK = np.dtype([("dgammadot0", "f8"), ("x", "f8"), ("y", "f8"), ("z", "f8"), ("width", "f8"),
                 ("length", "f8"), ("thickness", "f8"), ("strike", "f8"), ("dip", "f8"), ("e", "f4", (6,))])
# SOURCE_STRUCT, types.f90:26-29 (shear patches and surface loads of an event)
SOURCE = np.dtype([(k, "f8") for k in ("slip", "x", "y", "z", "width", "length", "strike", "dip", "rake", "period",
                                       "phase", "beta", "opening")])
TETRA = np.dtype([("e", "f4", (6,)), ("i1", "i4"), ("i2", "i4"), ("i3", "i4"), ("i4", "i4")])


class OrcModel(ctypes.Structure):
    _fields_ = [("sx1", ctypes.c_int), ("sx2", ctypes.c_int), ("sx3", ctypes.c_int),
                ("dx1", ctypes.c_double), ("dx2", ctypes.c_double), ("dx3", ctypes.c_double),
                ("lam", ctypes.c_double), ("mu", ctypes.c_double), ("gam", ctypes.c_double),
                ("beta", ctypes.c_double),
                ("linear", ctypes.c_void_p), ("nonlinear", ctypes.c_void_p),
                ("faultcreep", ctypes.c_void_p), ("prestress", ctypes.c_void_p),
                ("np", ctypes.c_int), ("planes", ctypes.c_void_p), ("patches", ctypes.c_void_p),
                ("interval", ctypes.c_double), ("odt", ctypes.c_double), ("tscale", ctypes.c_double),
                ("skip", ctypes.c_int), ("ne", ctypes.c_int), ("event_times", ctypes.c_void_p),
                ("ltransient", ctypes.c_void_p), ("nltransient", ctypes.c_void_p),
                ("dg0", ctypes.c_void_p * 4),
                ("ev_src", ctypes.c_void_p), ("ev_off", ctypes.c_void_p), ("ev_loads", ctypes.c_void_p),
                ("ev_loff", ctypes.c_void_p), ("nyquist", ctypes.c_double),
                ("ev_ts", ctypes.c_void_p), ("ev_tsoff", ctypes.c_void_p), ("ev_mg", ctypes.c_void_p),
                ("ev_mgoff", ctypes.c_void_p),
                ("inter_s", ctypes.c_void_p), ("inter_ns", ctypes.c_int),
                ("inter_ts", ctypes.c_void_p), ("inter_nt", ctypes.c_int),
                ("ev_cb", ctypes.c_void_p), ("ev_cboff", ctypes.c_void_p), ("ev_th", ctypes.c_void_p),
                ("ev_thoff", ctypes.c_void_p), ("ev_vx", ctypes.c_void_p), ("ev_vxoff", ctypes.c_void_p)]


def lib():
    global _LIB
    if _LIB is None:
        path = os.path.join(_HERE, "liboracle.so")
        if not os.path.exists(path):
            raise RuntimeError("oracle/liboracle.so is not built (run make -C oracle)")
        _LIB = ctypes.CDLL(path)
        _LIB.orc_tensoramplitude_f.restype = ctypes.c_double
        _LIB.orc_tensoramplitude_d.restype = ctypes.c_double
        assert _LIB.orc_sizeof_patch() == PATCH.itemsize
        assert _LIB.orc_sizeof_model() == ctypes.sizeof(OrcModel)
        assert _LIB.orc_sizeof_source() == SOURCE.itemsize
        assert _LIB.orc_sizeof_weak() == WEAK.itemsize
        assert _LIB.orc_sizeof_tetra() == TETRA.itemsize
    return _LIB


def _p(a):
    return None if a is None else a.ctypes.data_as(ctypes.c_void_p)


def _sfx(a):
    return "f" if a.dtype == np.float32 else "d"


def _d(x):
    return ctypes.c_double(float(x))


def _chk(rc, what):
    if rc != 0:
        raise RuntimeError("%s failed with status %d" % (what, rc))


# --- fourier.f90 -----------------------------------------------------------
def fft3(data, sx1, sx2, sx3, dx1, dx2, dx3, direction):
    _chk(getattr(lib(), "orc_fft3_" + _sfx(data))(_p(data), sx1, sx2, sx3, _d(dx1), _d(dx2), _d(dx3), direction), "fft3")


def fft2(data, sx1, sx2, dx1, dx2, direction):
    _chk(getattr(lib(), "orc_fft2_" + _sfx(data))(_p(data), sx1, sx2, _d(dx1), _d(dx2), direction), "fft2")


# --- green.f90 -------------------------------------------------------------
def elasticresponse(lam, mu, f1, f2, f3, dx1, dx2, dx3):
    sx3, sx2, p = f1.shape
    getattr(lib(), "orc_elasticresponse_" + _sfx(f1))(_d(lam), _d(mu), _p(f1), _p(f2), _p(f3), p - 2, sx2, sx3,
                                                     _d(dx1), _d(dx2), _d(dx3))


def surfacetractioncowling(lam, mu, gamma, u1, u2, u3, dx1, dx2, dx3):
    sx3, sx2, p = u1.shape
    out = [np.zeros((sx2, p), u1.dtype) for _ in range(3)]
    getattr(lib(), "orc_surfacetractioncowling_" + _sfx(u1))(
        _d(lam), _d(mu), _d(gamma), _p(u1), _p(u2), _p(u3), p - 2, sx2, sx3, _d(dx1), _d(dx2), _d(dx3),
        _p(out[0]), _p(out[1]), _p(out[2]))
    return out


def cerrutimodified(p1, p2, p3, lam, mu, gamma, u1, u2, u3, dx1, dx2, dx3):
    sx3, sx2, p = u1.shape
    getattr(lib(), "orc_cerrutimodified_" + _sfx(u1))(
        _p(p1), _p(p2), _p(p3), _d(lam), _d(mu), _d(gamma), _p(u1), _p(u2), _p(u3), p - 2, sx2, sx3,
        _d(dx1), _d(dx2), _d(dx3))


def greenfunctioncowling(c1, c2, c3, t1, t2, t3, dx1, dx2, dx3, lam, mu, gamma):
    sx3, sx2, p = c1.shape
    _chk(getattr(lib(), "orc_greenfunctioncowling_" + _sfx(c1))(
        _p(c1), _p(c2), _p(c3), _p(t1), _p(t2), _p(t3), p - 2, sx2, sx3, _d(dx1), _d(dx2), _d(dx3),
        _d(lam), _d(mu), _d(gamma)), "greenfunctioncowling")


# --- elastic3d.f90 ---------------------------------------------------------
def stressupdate(v1, v2, v3, lam, mu, dx1, dx2, dx3, sig):
    sx3, sx2, sx1, _ = sig.shape
    _chk(getattr(lib(), "orc_stressupdate_" + _sfx(sig))(
        _p(v1), _p(v2), _p(v3), _d(lam), _d(mu), _d(dx1), _d(dx2), _d(dx3), sx1, sx2, sx3, _p(sig)), "stressupdate")


def equivalentbodyforce(sig, dx1, dx2, dx3, c1, c2, c3, t1, t2, t3, mask=None):
    sx3, sx2, sx1, _ = sig.shape
    _chk(getattr(lib(), "orc_equivalentbodyforce_" + _sfx(sig))(
        _p(sig), _d(dx1), _d(dx2), _d(dx3), sx1, sx2, sx3, _p(c1), _p(c2), _p(c3), _p(t1), _p(t2), _p(t3),
        _p(mask)), "equivalentbodyforce")


def tensoramplitude(t, dx1, dx2, dx3):
    n = t.size // 6
    return getattr(lib(), "orc_tensoramplitude_" + _sfx(t))(_p(t), ctypes.c_long(n), _d(dx1), _d(dx2), _d(dx3))


def momentdensityshear(mu, slip, x, y, z, L, W, strike, dip, rake, beta, dx1, dx2, dx3, sig):
    sx3, sx2, sx1, _ = sig.shape
    getattr(lib(), "orc_momentdensityshear_" + _sfx(sig))(
        _d(mu), _d(slip), _d(x), _d(y), _d(z), _d(L), _d(W), _d(strike), _d(dip), _d(rake), _d(beta),
        sx1, sx2, sx3, _d(dx1), _d(dx2), _d(dx3), _p(sig))


# --- viscoelastic3d.f90 / friction3d.f90 -----------------------------------
def viscouseigenstress(mu, structure, prestress, sig, moment, maxwelltime=None, dgammadot0=None, gamma=None):
    sx3, sx2, sx1, _ = sig.shape
    mw = None if maxwelltime is None else np.array([maxwelltime], np.float64)
    _chk(getattr(lib(), "orc_viscouseigenstress_" + _sfx(sig))(
        _d(mu), _p(structure), _p(prestress), _p(sig), sx1, sx2, sx3, _p(moment), _p(mw), _p(dgammadot0),
        _p(gamma)), "viscouseigenstress")
    return None if mw is None else float(mw[0])


def transienteigenstress(mu, structure, sig, epsilonik, moment, epsilonikdot, maxwelltime=None, dgammadot0=None):
    """viscoelastic3d.f90:123-202; epsilonik may be epsilonikdot itself."""
    sx3, sx2, sx1, _ = sig.shape
    mw = None if maxwelltime is None else np.array([maxwelltime], np.float64)
    _chk(getattr(lib(), "orc_transienteigenstress_" + _sfx(sig))(
        _d(mu), _p(structure), _p(sig), _p(epsilonik), sx1, sx2, sx3, _p(moment), _p(epsilonikdot), _p(mw),
        _p(dgammadot0)), "transienteigenstress")
    return None if mw is None else float(mw[0])


def builddgammadot0(sx1, sx2, sx3, dx1, dx2, dx3, beta, zones, dtype=np.float32):
    """viscoelastic3d.f90:320-435: the weak-zone fluidity field (sx3, sx2, sx1)."""
    zones = np.ascontiguousarray(zones, WEAK)
    out = np.zeros((sx3, sx2, sx1), dtype)
    getattr(lib(), "orc_builddgammadot0_" + ("f" if out.dtype == np.float32 else "d"))(
        sx1, sx2, sx3, _d(dx1), _d(dx2), _d(dx3), _d(beta), len(zones), _p(zones), _p(out))
    return out


def frictioneigenstress(x, y, z, L, W, strike, dip, rake, beta, sig, mu, structure, dx1, dx2, dx3, moment,
                        maxwelltime=None, vel=None):
    sx3, sx2, sx1, _ = sig.shape
    mw = None if maxwelltime is None else np.array([maxwelltime], np.float64)
    _chk(getattr(lib(), "orc_frictioneigenstress_" + _sfx(sig))(
        _d(x), _d(y), _d(z), _d(L), _d(W), _d(strike), _d(dip), _d(rake), _d(beta), _p(sig), _d(mu),
        _p(structure), sx1, sx2, sx3, _d(dx1), _d(dx2), _d(dx3), _p(moment), _p(mw), _p(vel)),
        "frictioneigenstress")
    return None if mw is None else float(mw[0])


def monitorfriction(x, y, z, L, W, strike, dip, rake, beta, sx3_full, dx1, dx2, dx3, sig, structure, patch):
    sx3, sx2, sx1, _ = sig.shape
    px3, px2 = patch.shape
    _chk(getattr(lib(), "orc_monitorfriction_" + _sfx(sig))(
        _d(x), _d(y), _d(z), _d(L), _d(W), _d(strike), _d(dip), _d(rake), _d(beta), sx1, sx2, sx3_full, sx3,
        _d(dx1), _d(dx2), _d(dx3), _p(sig), _p(structure), _p(patch), px2, px3), "monitorfriction")


def source(mu, slip, x, y, z, L, W, strike, dip, rake, beta, dx1, dx2, dx3, f1, f2, f3, t1, t2, t3):
    """elastic3d.f90:2064-2252: adds the body force of a shear dislocation to
    the padded fields f (sx3, sx2, sx1+2) (top half) and its surface
    traction to the padded planes t (sx2, sx1+2)."""
    sx3, sx2, p1 = f1.shape
    getattr(lib(), "orc_source_" + _sfx(f1))(
        _d(mu), _d(slip), _d(x), _d(y), _d(z), _d(L), _d(W), _d(strike), _d(dip), _d(rake), _d(beta), p1 - 2, sx2,
        sx3, _d(dx1), _d(dx2), _d(dx3), _p(f1), _p(f2), _p(f3), _p(t1), _p(t2), _p(t3))


def traction(loads, dx1, dx2, t, t3, Dt=0.0, rate=False):
    """elastic3d.f90:2517-2590: subtracts the surface loads (SOURCE array) at
    time t (or their rate over [t, t+Dt]) from the padded plane t3."""
    sx2, p1 = t3.shape
    loads = np.ascontiguousarray(loads, SOURCE)
    getattr(lib(), "orc_traction_" + _sfx(t3))(_p(loads), len(loads), p1 - 2, sx2, _d(dx1), _d(dx2), _d(t), _d(Dt),
                                               _p(t3), int(rate))


def dislocations(sources, loads, mu, nyquist, t, dx1, dx2, dx3, v, tr, tau, lam=None, beta=0.2, tensile=None,
                 mogi=None, cuboids=None, tetrahedra=None, vertices=None):
    """relax.f90:1116-1289 then traction: the source() body force of the
    shear patches (SOURCE array) longer or wider than nyquist*min(dx) into
    the padded fields v (3 arrays (sx3, sx2, sx1+2)) and planes tr, every
    patch's moment density into tau; the same for the eigenstrain cuboids
    (WEAK) and tetrahedra (TETRA with vertices (n, 3)), the tensile cracks
    (SOURCE, opening) and the point dilatation sources (SOURCE, slip = volume
    change), whose body forces fill the full depth of v; the surface loads
    into tr[2].  lam defaults to mu."""
    sx3, sx2, p1 = v[0].shape
    arr = [np.ascontiguousarray(np.zeros(0, SOURCE) if a is None else a, SOURCE)
           for a in (sources, tensile, mogi, loads)]
    cb = np.ascontiguousarray(np.zeros(0, WEAK) if cuboids is None else cuboids, WEAK)
    th = np.ascontiguousarray(np.zeros(0, TETRA) if tetrahedra is None else tetrahedra, TETRA)
    vx = np.ascontiguousarray(np.zeros((0, 3)) if vertices is None else vertices, np.float64).reshape(-1, 3)
    getattr(lib(), "orc_dislocations_" + _sfx(tau))(
        _p(arr[0]), len(arr[0]), _p(arr[1]), len(arr[1]), _p(arr[2]), len(arr[2]), _p(arr[3]), len(arr[3]),
        _d(mu if lam is None else lam), _d(mu), _d(beta), _d(nyquist), _d(t), p1 - 2, sx2, sx3, _d(dx1), _d(dx2),
        _d(dx3), _p(v[0]), _p(v[1]), _p(v[2]), _p(tr[0]), _p(tr[1]), _p(tr[2]), _p(tau), _p(cb), len(cb), _p(th),
        len(th), _p(vx), len(vx))


def eigenstrain(cuboids, tetrahedra, vertices, lam, mu, beta, dx1, dx2, dx3, v, tau):
    """relax.f90:1151-1178, :1192-1223: the body forces of the cuboids (WEAK
    arrays, the eigenstrain in e) and tetrahedra (TETRA, 1-based indices into
    vertices, (n, 3) float64) into the top half of the padded fields v (3
    arrays (sx3, sx2, sx1+2)), then their moment densities into tau."""
    sx3, sx2, p1 = v[0].shape
    cb = np.ascontiguousarray(np.zeros(0, WEAK) if cuboids is None else cuboids, WEAK)
    th = np.ascontiguousarray(np.zeros(0, TETRA) if tetrahedra is None else tetrahedra, TETRA)
    vx = np.ascontiguousarray(np.zeros((0, 3)) if vertices is None else vertices, np.float64).reshape(-1, 3)
    getattr(lib(), "orc_eigenstrain_" + _sfx(tau))(
        _p(cb), len(cb), _p(th), len(th), _p(vx), len(vx), _d(lam), _d(mu), _d(beta), p1 - 2, sx2, sx3, _d(dx1),
        _d(dx2), _d(dx3), _p(v[0]), _p(v[1]), _p(v[2]), _p(tau))


def _e6(e):
    return np.ascontiguousarray(e, np.float32).reshape(6)


def cuboidsource(lam, mu, e, x, y, z, L, W, T, strike, dip, beta, dx1, dx2, dx3, f1, f2, f3):
    """elastic3d.f90:1680-1884 on padded fields (sx3, sx2, sx1+2), top half."""
    sx3, sx2, p1 = f1.shape
    e = _e6(e)
    getattr(lib(), "orc_cuboidsource_" + _sfx(f1))(
        _d(lam), _d(mu), _p(e), _d(x), _d(y), _d(z), _d(L), _d(W), _d(T), _d(strike), _d(dip), _d(beta), p1 - 2,
        sx2, sx3, _d(dx1), _d(dx2), _d(dx3), _p(f1), _p(f2), _p(f3))


def momentdensitycuboid(mu, lam, e, x, y, z, L, W, T, strike, dip, beta, dx1, dx2, dx3, sig):
    """elastic3d.f90:2927-3029, added to the tensor field sig (sx3, sx2, sx1, 6)."""
    sx3, sx2, sx1, _ = sig.shape
    e = _e6(e)
    getattr(lib(), "orc_momentdensitycuboid_" + _sfx(sig))(
        _d(mu), _d(lam), _p(e), _d(x), _d(y), _d(z), _d(L), _d(W), _d(T), _d(strike), _d(dip), _d(beta), sx1, sx2,
        sx3, _d(dx1), _d(dx2), _d(dx3), _p(sig))


def _v3(a):
    return np.ascontiguousarray(a, np.float64).reshape(3)


def tetrahedronsource(lam, mu, e, A, B, C, D, dx1, dx2, dx3, f1, f2, f3):
    """elastic3d.f90:1916-2039 on padded fields (sx3, sx2, sx1+2), top half."""
    sx3, sx2, p1 = f1.shape
    e, A, B, C, D = _e6(e), _v3(A), _v3(B), _v3(C), _v3(D)
    getattr(lib(), "orc_tetrahedronsource_" + _sfx(f1))(
        _d(lam), _d(mu), _p(e), _p(A), _p(B), _p(C), _p(D), p1 - 2, sx2, sx3, _d(dx1), _d(dx2), _d(dx3), _p(f1),
        _p(f2), _p(f3))


def momentdensitytetrahedron(mu, lam, e, A, B, C, D, dx1, dx2, dx3, sig):
    """elastic3d.f90:2788-2873, added to the tensor field sig (sx3, sx2, sx1, 6)."""
    sx3, sx2, sx1, _ = sig.shape
    e, A, B, C, D = _e6(e), _v3(A), _v3(B), _v3(C), _v3(D)
    getattr(lib(), "orc_momentdensitytetrahedron_" + _sfx(sig))(
        _d(mu), _d(lam), _p(e), _p(A), _p(B), _p(C), _p(D), sx1, sx2, sx3, _d(dx1), _d(dx2), _d(dx3), _p(sig))


def tensilesource(lam, mu, s, x, y, z, L, W, strike, dip, beta, dx1, dx2, dx3, f1, f2, f3):
    """elastic3d.f90:2315-2411 on padded full-depth fields (sx3, sx2, sx1+2)."""
    sx3, sx2, p1 = f1.shape
    getattr(lib(), "orc_tensilesource_" + _sfx(f1))(
        _d(lam), _d(mu), _d(s), _d(x), _d(y), _d(z), _d(L), _d(W), _d(strike), _d(dip), _d(beta), p1 - 2, sx2, sx3,
        _d(dx1), _d(dx2), _d(dx3), _p(f1), _p(f2), _p(f3))


def mogisource(lam, mu, o, x, y, z, dx1, dx2, dx3, f1, f2, f3):
    """elastic3d.f90:2458-2509 on padded full-depth fields (sx3, sx2, sx1+2)."""
    sx3, sx2, p1 = f1.shape
    getattr(lib(), "orc_mogisource_" + _sfx(f1))(
        _d(lam), _d(mu), _d(o), _d(x), _d(y), _d(z), p1 - 2, sx2, sx3, _d(dx1), _d(dx2), _d(dx3), _p(f1), _p(f2),
        _p(f3))


def momentdensitytensile(lam, mu, s, x, y, z, L, W, strike, dip, beta, dx1, dx2, dx3, sig):
    """elastic3d.f90:3137-3215, added to the tensor field sig (sx3, sx2, sx1, 6)."""
    sx3, sx2, sx1, _ = sig.shape
    getattr(lib(), "orc_momentdensitytensile_" + _sfx(sig))(
        _d(lam), _d(mu), _d(s), _d(x), _d(y), _d(z), _d(L), _d(W), _d(strike), _d(dip), _d(beta), sx1, sx2, sx3,
        _d(dx1), _d(dx2), _d(dx3), _p(sig))


def momentdensitymogi(lam, mu, o, x, y, z, dx1, dx2, dx3, sig):
    """elastic3d.f90:3240-3286, added to the tensor field sig (sx3, sx2, sx1, 6)."""
    sx3, sx2, sx1, _ = sig.shape
    getattr(lib(), "orc_momentdensitymogi_" + _sfx(sig))(
        _d(lam), _d(mu), _d(o), _d(x), _d(y), _d(z), sx1, sx2, sx3, _d(dx1), _d(dx2), _d(dx3), _p(sig))


def shiftedindex(x1, x2, x3, sx1, sx2, sx3, dx1, dx2, dx3):
    """elastic3d.f90:4063-4101, 0-based (i1, i2, i3); None out of range."""
    idx = np.zeros(3, np.int32)
    rc = lib().orc_shiftedindex(_d(x1), _d(x2), _d(x3), sx1, sx2, sx3, _d(dx1), _d(dx2), _d(dx3), _p(idx))
    return None if rc else tuple(int(v) for v in idx)


def exportpoints(u1, u2, u3, sig, dx1, dx2, dx3, pts, rot=0.0):
    """util.f90:59-139: (n, 9) array of v1 v2 v3 (user frame) and sig at the
    observation points pts (n, 3) of the internal frame."""
    sx3, sx2, sx1, _ = sig.shape
    pts = np.ascontiguousarray(pts, np.float64).reshape(-1, 3)
    out = np.zeros((len(pts), 9), np.float64)
    _chk(getattr(lib(), "orc_exportpoints_" + _sfx(sig))(
        _p(u1), _p(u2), _p(u3), _p(sig), sx1, sx2, sx3, _d(dx1), _d(dx2), _d(dx3), _p(pts), len(pts), _d(rot),
        _p(out)), "exportpoints")
    return out


# --- okada/dc3d.f, okada/green_space.f90, relax.f90:1300-1349 -------------
def dc3d_core(alpha, x, y, z, depth, dip, al1, al2, aw1, aw2, U):
    """DC3D's displacement in double precision (dip in degrees, z <= 0 up);
    returns (u (3,), singular flag)."""
    Ua = np.ascontiguousarray(U, np.float64)
    u = np.zeros(3)
    rc = lib().orc_dc3d_core(*[_d(v) for v in (alpha, x, y, z, depth, dip, al1, al2, aw1, aw2)], _p(Ua), _p(u))
    return u, rc


def dc3d(alpha, x, y, z, depth, dip, al1, al2, aw1, aw2, d1, d2, d3):
    """DC3D with single-precision arguments and results, as dc3d.f."""
    u = np.zeros(3, np.float32)
    f = ctypes.c_float
    rc = lib().orc_dc3d(*[f(v) for v in (alpha, x, y, z, depth, dip, al1, al2, aw1, aw2, d1, d2, d3)], _p(u))
    return u, rc


def okada_point(lam, mu, slip, opening, x, y, z, width, length, strike, dip, rake, xrec, yrec, zrec):
    """green_space.f90:29-105: displacement (north, east, down) at one point."""
    d = np.zeros(3)
    lib().orc_okada_point(*[_d(v) for v in (lam, mu, slip, opening, x, y, z, width, length, strike, dip, rake, xrec,
                                            yrec, zrec)], _p(d))
    return d


def grnfct_okada(lam, mu, slip, opening, x, y, z, width, length, strike, dip, rake, dx1, dx2, dx3, u1, u2, u3):
    """green_space.f90:119-146 on padded fields (sx3, sx2, sx1+2) (added to)."""
    sx3, sx2, p1 = u1.shape
    getattr(lib(), "orc_grnfct_okada_" + _sfx(u1))(
        *[_d(v) for v in (lam, mu, slip, opening, x, y, z, width, length, strike, dip, rake)], p1 - 2, sx2, sx3,
        _d(dx1), _d(dx2), _d(dx3), _p(u1), _p(u2), _p(u3))


def dislocations_disp(sc, tsc, lam, mu, nyquist, dx1, dx2, dx3, u1, u2, u3):
    """relax.f90:1300-1349: Okada's displacement of the shear patches sc and
    tensile cracks tsc (SOURCE arrays in the deck's convention) at or below
    the nyquist size, added to the half-grid fields u (h3, sx2, sx1+2)."""
    h3, sx2, p1 = u1.shape
    arr = [np.ascontiguousarray(np.zeros(0, SOURCE) if a is None else a, SOURCE) for a in (sc, tsc)]
    getattr(lib(), "orc_dislocations_disp_" + _sfx(u1))(
        _p(arr[0]), len(arr[0]), _p(arr[1]), len(arr[1]), _d(lam), _d(mu), _d(nyquist), p1 - 2, sx2, h3, _d(dx1),
        _d(dx2), _d(dx3), _p(u1), _p(u2), _p(u3))


def dislocations_disp_points(sc, tsc, lam, mu, nyquist, sx1, sx2, dx1, dx2, dx3, idx):
    """The displacement dislocations_disp adds at the grid points idx ((n, 3)
    0-based i1, i2, i3), summed from zero in single precision: (n, 3)."""
    idx = np.ascontiguousarray(idx, np.int32).reshape(-1, 3)
    out = np.zeros((len(idx), 3), np.float32)
    arr = [np.ascontiguousarray(np.zeros(0, SOURCE) if a is None else a, SOURCE) for a in (sc, tsc)]
    lib().orc_dislocations_disp_points(_p(arr[0]), len(arr[0]), _p(arr[1]), len(arr[1]), _d(lam), _d(mu),
                                       _d(nyquist), sx1, sx2, _d(dx1), _d(dx2), _d(dx3), _p(idx),
                                       ctypes.c_long(len(idx)), _p(out))
    return out


def frictionadd(patch, dt):
    lib().orc_frictionadd(_p(patch), ctypes.c_long(patch.size), _d(dt))


# --- relax.f90 RK2 loop ----------------------------------------------------
def rk2_run(model, state, nsteps, dt_replay=None):
    """Advance `state` (dict of u1,u2,u3,tau,sig,gamma arrays, t, oi, i) in place.

    `model` is a dict: sx1,sx2,sx3,dx1,dx2,dx3,lam,mu,gam,beta, linear/nonlinear/
    faultcreep (LAYER arrays of sx3/2 or None), prestress (TENSOR_LAYER sx3/2),
    planes (PLANE array or None), patches (PATCH array or None), interval, odt,
    tscale, skip, event_times; optional ltransient/nltransient (LAYER arrays)
    with state["epsilonik"] (tensor field), and weak-zone fields dg_linear,
    dg_nonlinear, dg_ltransient, dg_nltransient ((sx3/2, sx2, sx1) arrays of
    the state's dtype), and the later events: event_sources, a list (one
    entry per event after the first) of SOURCE arrays, event_loads,
    event_tensile and event_mogi likewise (surface loads, tensile cracks,
    point dilatation sources; optional), event_cuboids (WEAK arrays),
    event_tetrahedra (TETRA arrays) and event_vertices ((n, 3) arrays), loads0 (the first event's surface
    loads, whose rate the steps add until the next event), nyquist
    (dislocations' size gate), and the inter-seismic loading: inter and
    inter_tensile, SOURCE arrays of shear patches and tensile cracks whose
    slip or opening is a rate (relax.f90:669-674, :795-799).
    state["e"] is the current event (1-based, default 1).  Returns an
    (nsteps, 10) array of
    Dt, tm, maxwell(1:3), power, C:E^i, t, maxwell(4:5).
    """
    keep = []

    def ptr(a):
        if a is None:
            return None
        a = np.ascontiguousarray(a)
        keep.append(a)
        return a.ctypes.data

    ev = np.asarray(model.get("event_times", [0.0]), np.float64)
    m = OrcModel(model["sx1"], model["sx2"], model["sx3"], model["dx1"], model["dx2"], model["dx3"],
                 model["lam"], model["mu"], model["gam"], model.get("beta", 0.2),
                 ptr(model.get("linear")), ptr(model.get("nonlinear")), ptr(model.get("faultcreep")),
                 ptr(model["prestress"]), 0 if model.get("planes") is None else len(model["planes"]),
                 ptr(model.get("planes")), None if model.get("patches") is None else model["patches"].ctypes.data,
                 model["interval"], model.get("odt", 1.0), model.get("tscale", 1.0), model.get("skip", 1),
                 len(ev), ptr(ev), ptr(model.get("ltransient")), ptr(model.get("nltransient")))
    for key, fs, fo in (("event_sources", "ev_src", "ev_off"), ("event_loads", "ev_loads", "ev_loff"),
                        ("event_tensile", "ev_ts", "ev_tsoff"), ("event_mogi", "ev_mg", "ev_mgoff")):
        evs = model.get(key) or []
        first = model.get("loads0") if key == "event_loads" else None
        if evs or first is not None:
            # entry 0: the first event (only its loads are used, for their rate)
            lists = [np.asarray(np.zeros(0, SOURCE) if first is None else first, SOURCE)]
            lists += [np.asarray(a, SOURCE) for a in evs]
            off = np.zeros(len(lists) + 1, np.int32)
            off[1:] = np.cumsum([len(a) for a in lists])
            setattr(m, fs, ptr(np.concatenate(lists).astype(SOURCE)))
            setattr(m, fo, ptr(off))
    for key, fs, fo, dt in (("event_cuboids", "ev_cb", "ev_cboff", WEAK),
                            ("event_tetrahedra", "ev_th", "ev_thoff", TETRA)):
        evs = model.get(key) or []
        if evs:
            lists = [np.zeros(0, dt)] + [np.asarray(a, dt) for a in evs]
            off = np.zeros(len(lists) + 1, np.int32)
            off[1:] = np.cumsum([len(a) for a in lists])
            setattr(m, fs, ptr(np.concatenate(lists).astype(dt)))
            setattr(m, fo, ptr(off))
    evs = model.get("event_vertices") or []
    if evs:
        lists = [np.zeros((0, 3))] + [np.asarray(a, np.float64).reshape(-1, 3) for a in evs]
        off = np.zeros(len(lists) + 1, np.int32)
        off[1:] = np.cumsum([len(a) for a in lists])
        m.ev_vx = ptr(np.ascontiguousarray(np.concatenate(lists)))
        m.ev_vxoff = ptr(off)
    m.nyquist = model.get("nyquist", 1.0)
    for key, fp, fn in (("inter", "inter_s", "inter_ns"), ("inter_tensile", "inter_ts", "inter_nt")):
        a = model.get(key)
        if a is not None and len(a):
            setattr(m, fp, ptr(np.asarray(a, SOURCE)))
            setattr(m, fn, len(a))
    for k, key in enumerate(("dg_linear", "dg_nonlinear", "dg_ltransient", "dg_nltransient")):
        a = model.get(key)
        if a is not None:
            assert a.dtype == state["u1"].dtype and a.size == state["tau"].size // 6, key
        m.dg0[k] = ptr(a)
    sfx = _sfx(state["u1"])
    t = ctypes.c_double(state.get("t", 0.0))
    oi = ctypes.c_int(state.get("oi", 1))
    it = ctypes.c_int(state.get("i", 0))
    ev = ctypes.c_int(state.get("e", 1))
    info = np.zeros((nsteps, 10), np.float64)
    dtr = None if dt_replay is None else np.ascontiguousarray(dt_replay, np.float64)
    rc = getattr(lib(), "orc_rk2_run_" + sfx)(
        ctypes.byref(m), _p(state["u1"]), _p(state["u2"]), _p(state["u3"]), _p(state["tau"]), _p(state["sig"]),
        _p(state["gamma"]), _p(state.get("epsilonik")), ctypes.byref(t), ctypes.byref(oi), ctypes.byref(it), ctypes.byref(ev), nsteps, _p(dtr), _p(info))
    _chk(rc, "rk2_run")
    done = it.value - state.get("i", 0)
    state["t"], state["oi"], state["i"], state["e"] = t.value, oi.value, it.value, ev.value
    return info[:done]
