"""viscoelastic3d.f90 on the GPU."""
import ctypes

import numpy as np

from ._lib import LAYER, TENSOR_LAYER, WEAK, check, f32, load, ptr


def viscouseigenstress(mu, structure, prestress, sig, moment, maxwelltime=None, dgammadot0=None, gamma=None,
                       align_data=1):
    """viscouseigenstress, viscoelastic3d.f90:217-314.

    structure: LAYER array (sx3 entries), prestress: TENSOR_LAYER array;
    moment is accumulated in place; returns the updated maxwelltime."""
    assert structure.dtype == LAYER and prestress.dtype == TENSOR_LAYER
    sx3, sx2, sx1, _ = sig.shape
    mw = None if maxwelltime is None else ctypes.c_double(maxwelltime)
    check(load().relax_viscouseigenstress(mu, ptr(structure), len(structure), ptr(prestress), len(prestress),
                                          ptr(f32(sig, "sig")), sx1, sx2, sx3, 1.0, 1.0, 1.0,
                                          ptr(f32(moment, "moment")), None if mw is None else ctypes.byref(mw),
                                          None if dgammadot0 is None else ptr(f32(dgammadot0, "dgammadot0")),
                                          None if gamma is None else ptr(f32(gamma, "gamma")), align_data),
          "viscouseigenstress")
    return None if mw is None else mw.value


def transienteigenstress(mu, structure, sig, epsilonik, moment, epsilonikdot, maxwelltime=None, dgammadot0=None):
    """transienteigenstress, viscoelastic3d.f90:123-202.

    structure: LAYER array (sx3 entries, Gk the Kelvin modulus); moment is
    accumulated and epsilonikdot written in place (epsilonik may be
    epsilonikdot itself); returns the updated maxwelltime."""
    assert structure.dtype == LAYER
    sx3, sx2, sx1, _ = sig.shape
    mw = None if maxwelltime is None else ctypes.c_double(maxwelltime)
    eik = ptr(f32(epsilonik, "epsilonik"))
    check(load().relax_transienteigenstress(mu, ptr(structure), len(structure), ptr(f32(sig, "sig")), eik, sx1,
                                            sx2, sx3, 1.0, 1.0, 1.0, ptr(f32(moment, "moment")),
                                            ptr(f32(epsilonikdot, "epsilonikdot")),
                                            None if mw is None else ctypes.byref(mw),
                                            None if dgammadot0 is None else ptr(f32(dgammadot0, "dgammadot0"))),
          "transienteigenstress")
    return None if mw is None else mw.value


def builddgammadot0(sx1, sx2, sx3, dx1, dx2, dx3, beta, zones):
    """builddgammadot0, viscoelastic3d.f90:320-435: the weak-zone fluidity
    field, a (sx3, sx2, sx1) float32 array."""
    zones = np.ascontiguousarray(zones, WEAK)
    out = np.zeros((sx3, sx2, sx1), np.float32)
    check(load().relax_builddgammadot0(sx1, sx2, sx3, dx1, dx2, dx3, beta, len(zones),
                                       ptr(zones) if len(zones) else None, ptr(out)), "builddgammadot0")
    return out
