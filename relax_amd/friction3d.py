"""friction3d.f90 on the GPU."""
import ctypes

import numpy as np

from ._lib import LAYER, PATCH, TENSOR_LAYER, check, f32, load, ptr


def frictioneigenstress(x, y, z, L, W, strike, dip, rake, beta, sig, mu, structure, dx1, dx2, dx3, moment,
                        maxwelltime=None, vel=None, prestress=None, align_data=1):
    """frictioneigenstress, friction3d.f90:82-241; returns the updated maxwelltime."""
    assert structure.dtype == LAYER
    sx3, sx2, sx1, _ = sig.shape
    mw = None if maxwelltime is None else ctypes.c_double(maxwelltime)
    pre = np.zeros(sx3, TENSOR_LAYER) if prestress is None else prestress
    check(load().relax_frictioneigenstress(x, y, z, L, W, strike, dip, rake, beta, ptr(f32(sig, "sig")), ptr(pre),
                                           mu, ptr(structure), sx1, sx2, sx3, dx1, dx2, dx3,
                                           ptr(f32(moment, "moment")), None if mw is None else ctypes.byref(mw),
                                           None if vel is None else ptr(f32(vel, "vel")), align_data),
          "frictioneigenstress")
    return None if mw is None else mw.value


def monitorfriction(x, y, z, L, W, strike, dip, rake, beta, sx3_full, dx1, dx2, dx3, sig, structure, patch,
                    prestress=None):
    """monitorfriction, friction3d.f90:263-416; patch is a (px3, px2) PATCH array."""
    assert patch.dtype == PATCH and structure.dtype == LAYER
    sx3, sx2, sx1, _ = sig.shape
    px3, px2 = patch.shape
    pre = np.zeros(sx3, TENSOR_LAYER) if prestress is None else prestress
    check(load().relax_monitorfriction(x, y, z, L, W, strike, dip, rake, beta, sx1, sx2, sx3_full, sx3, dx1, dx2,
                                       dx3, ptr(f32(sig, "sig")), ptr(pre), ptr(structure), ptr(patch), px2, px3),
          "monitorfriction")


def frictionadd(patch, dt):
    """frictionadd, friction3d.f90:424-453 (one plane's patches)."""
    assert patch.dtype == PATCH
    check(load().relax_frictionadd(ptr(patch), patch.size, dt), "frictionadd")
