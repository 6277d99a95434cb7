"""elastic3d.f90 on the GPU: FIR stencils, field algebra, diagnostics."""
import ctypes

import numpy as np

from ._lib import SOURCE, check, f32, load, ptr


def stressupdate(v1, v2, v3, lam, mu, dx1, dx2, dx3, sig, align_data=1):
    """stressupdate, elastic3d.f90:956-1180: sig += C:eps(v) (in place).
    v: (sx3, sx2, sx1+2) (align_data=1) or (sx3, sx2, sx1); sig: (sx3, sx2, sx1, 6)."""
    sx3, sx2, sx1, _ = sig.shape
    check(load().relax_stressupdate(ptr(f32(v1, "v1")), ptr(f32(v2, "v2")), ptr(f32(v3, "v3")), lam, mu, dx1, dx2,
                                    dx3, sx1, sx2, sx3, ptr(f32(sig, "sig")), align_data), "stressupdate")


def equivalentbodyforce(sig, dx1, dx2, dx3, c1, c2, c3, t1, t2, t3, mask=None, align_data=1):
    """equivalentbodyforce, elastic3d.f90:1214-1506: c -= div(sig), t += sig(:,:,1).e3."""
    sx3, sx2, sx1, _ = sig.shape
    check(load().relax_equivalentbodyforce(ptr(f32(sig, "sig")), dx1, dx2, dx3, sx1, sx2, sx3, ptr(f32(c1, "c1")),
                                           ptr(f32(c2, "c2")), ptr(f32(c3, "c3")), ptr(f32(t1, "t1")),
                                           ptr(f32(t2, "t2")), ptr(f32(t3, "t3")),
                                           None if mask is None else ptr(f32(mask, "mask")), align_data),
          "equivalentbodyforce")


def _opt(c):
    return None if c is None else ctypes.byref(ctypes.c_float(c))


def fieldadd(u, v, c1=None, c2=None):
    """fieldadd, elastic3d.f90:3774-3812: u = c1*u + c2*v."""
    check(load().relax_fieldadd(ptr(f32(u, "u")), ptr(f32(v, "v")), u.size, _opt(c1), _opt(c2)), "fieldadd")


def tensorfieldadd(t1, t2, c1=None, c2=None):
    """tensorfieldadd, elastic3d.f90:3880-3959: t1 = c1*t1 + c2*t2."""
    check(load().relax_tensorfieldadd(ptr(f32(t1, "t1")), ptr(f32(t2, "t2")), t1.size // 6, _opt(c1), _opt(c2)),
          "tensorfieldadd")


def tensoramplitude(t, dx1, dx2, dx3):
    """tensoramplitude, elastic3d.f90:446-466."""
    out = np.zeros(1, np.float64)
    check(load().relax_tensoramplitude(ptr(f32(t, "t")), t.size // 6, dx1, dx2, dx3, ptr(out)), "tensoramplitude")
    return float(out[0])


def source(mu, slip, x, y, z, L, W, strike, dip, rake, beta, dx1, dx2, dx3, f1, f2, f3, t1, t2, t3, align_data=1):
    """source, elastic3d.f90:2064-2252: adds the body force of a tapered shear
    dislocation to the top half of f1..f3 ((sx3, sx2, sx1+2) with
    align_data=1, else (sx3, sx2, sx1)) and its surface traction to t1..t3
    ((sx2, sx1+2) or (sx2, sx1)), in place.  L is the down-dip extent (the
    patch width) and W the along-strike one (its length)."""
    sx3, sx2, p1 = f1.shape
    sx1 = p1 - 2 if align_data else p1
    check(load().relax_source(mu, slip, x, y, z, L, W, strike, dip, rake, beta, sx1, sx2, sx3, dx1, dx2, dx3,
                              ptr(f32(f1, "f1")), ptr(f32(f2, "f2")), ptr(f32(f3, "f3")), ptr(f32(t1, "t1")),
                              ptr(f32(t2, "t2")), ptr(f32(t3, "t3")), int(bool(align_data))), "source")


def traction(mu, loads, dx1, dx2, t, Dt, t3, rate=False, align_data=1):
    """traction, elastic3d.f90:2517-2590: subtracts the surface loads (SOURCE
    array) at time t, or their rate over [t, t+Dt], from t3 ((sx2, sx1+2)
    with align_data=1, else (sx2, sx1)), in place."""
    sx2, p1 = t3.shape
    sx1 = p1 - 2 if align_data else p1
    ld = np.ascontiguousarray(loads, SOURCE)
    check(load().relax_traction(mu, ptr(ld), len(ld), sx1, sx2, dx1, dx2, t, Dt, ptr(f32(t3, "t3")), int(rate),
                                int(bool(align_data))), "traction")


def tensilesource(lam, mu, s, x, y, z, L, W, strike, dip, beta, dx1, dx2, dx3, f1, f2, f3, align_data=1):
    """tensilesource, elastic3d.f90:2315-2411: adds the body force of an
    opening crack to the whole depth of f1..f3 ((sx3, sx2, sx1+2) with
    align_data=1, else (sx3, sx2, sx1)), in place."""
    sx3, sx2, p1 = f1.shape
    sx1 = p1 - 2 if align_data else p1
    check(load().relax_tensilesource(lam, mu, s, x, y, z, L, W, strike, dip, beta, sx1, sx2, sx3, dx1, dx2, dx3,
                                     ptr(f32(f1, "f1")), ptr(f32(f2, "f2")), ptr(f32(f3, "f3")),
                                     int(bool(align_data))), "tensilesource")


def mogisource(lam, mu, o, x, y, z, dx1, dx2, dx3, f1, f2, f3, align_data=1):
    """mogisource, elastic3d.f90:2458-2509: adds the body force of a point
    dilatation source (volume change o) to the whole depth of f1..f3."""
    sx3, sx2, p1 = f1.shape
    sx1 = p1 - 2 if align_data else p1
    check(load().relax_mogisource(lam, mu, o, x, y, z, sx1, sx2, sx3, dx1, dx2, dx3, ptr(f32(f1, "f1")),
                                  ptr(f32(f2, "f2")), ptr(f32(f3, "f3")), int(bool(align_data))), "mogisource")


def momentdensitytensile(lam, mu, s, x, y, z, L, W, strike, dip, beta, dx1, dx2, dx3, sig):
    """momentdensitytensile, elastic3d.f90:3137-3215 (adds to sig in place)."""
    sx3, sx2, sx1, _ = sig.shape
    check(load().relax_momentdensitytensile(lam, mu, s, x, y, z, L, W, strike, dip, beta, sx1, sx2, sx3, dx1, dx2,
                                            dx3, ptr(f32(sig, "sig"))), "momentdensitytensile")


def momentdensitymogi(lam, mu, o, x, y, z, dx1, dx2, dx3, sig):
    """momentdensitymogi, elastic3d.f90:3240-3286 (adds to sig in place)."""
    sx3, sx2, sx1, _ = sig.shape
    check(load().relax_momentdensitymogi(lam, mu, o, x, y, z, sx1, sx2, sx3, dx1, dx2, dx3, ptr(f32(sig, "sig"))),
          "momentdensitymogi")


def momentdensityshear(mu, slip, x, y, z, L, W, strike, dip, rake, beta, dx1, dx2, dx3, sig):
    """momentdensityshear, elastic3d.f90:2661-2750 (adds to sig in place)."""
    sx3, sx2, sx1, _ = sig.shape
    check(load().relax_momentdensityshear(mu, slip, x, y, z, L, W, strike, dip, rake, beta, sx1, sx2, sx3, dx1, dx2,
                                          dx3, ptr(f32(sig, "sig"))), "momentdensityshear")
