"""ctypes binding of librelax_hip.so (the C ABI declared in include/relax_hip.h).

The library is loaded from this package directory only; there is no CPU
fallback: every call goes to the HIP kernels and a missing library or device
raises.
"""
import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIBPATH = os.path.join(HERE, "librelax_hip.so")

c_int, c_long, c_double, c_float, c_void_p, c_size_t = (ctypes.c_int, ctypes.c_long, ctypes.c_double,
                                                        ctypes.c_float, ctypes.c_void_p, ctypes.c_size_t)
P = c_void_p  # every array argument is passed as a raw pointer

# numpy views of the ABI structures (types.f90 SEQUENCE layouts)
LAYER = np.dtype([("z", "f8"), ("gammadot0", "f8"), ("stressexponent", "f8"),
                  ("cohesion", "f8"), ("friction", "f8"), ("Gk", "f8")])
TENSOR_LAYER = np.dtype([("z", "f4"), ("dum", "f4"), ("t", "f4", (6,))])
PATCH = np.dtype([("x1", "f8"), ("x2", "f8"), ("x3", "f8"), ("lx", "f8"), ("lz", "f8"),
                  ("slip", "f8"), ("ss", "f8"), ("ds", "f8"), ("v", "f8"), ("vss", "f8"),
                  ("vds", "f8"), ("taus", "f8"), ("sig", "f4", (6,))])
PLANE = np.dtype([("x", "f8"), ("y", "f8"), ("z", "f8"), ("width", "f8"), ("length", "f8"),
                  ("strike", "f8"), ("dip", "f8"), ("rake", "f8"), ("px2", "i4"), ("px3", "i4")])
WEAK = np.dtype([("dgammadot0", "f8"), ("x", "f8"), ("y", "f8"), ("z", "f8"), ("width", "f8"),
                 ("length", "f8"), ("thickness", "f8"), ("strike", "f8"), ("dip", "f8"), ("e", "f4", (6,))])
SOURCE = np.dtype([(k, "f8") for k in ("slip", "x", "y", "z", "width", "length", "strike", "dip", "rake", "period",
                                       "phase", "beta", "opening")])


TETRA = np.dtype([("e", "f4", (6,)), ("i1", "i4"), ("i2", "i4"), ("i3", "i4"), ("i4", "i4")])


class EventDesc(ctypes.Structure):
    """relax_event_t: the source lists of one event (SOURCE arrays), its
    cuboids (WEAK), tetrahedra (TETRA) and their vertices ((n, 3) float64)."""
    _fields_ = [("s", c_void_p), ("ns", c_int), ("ts", c_void_p), ("nt", c_int), ("m", c_void_p), ("nm", c_int),
                ("l", c_void_p), ("nl", c_int), ("c", c_void_p), ("nc", c_int), ("th", c_void_p), ("nth", c_int),
                ("v", c_void_p), ("nv", c_int)]


class ModelDesc(ctypes.Structure):
    _fields_ = [("sx1", c_int), ("sx2", c_int), ("sx3", c_int),
                ("dx1", c_double), ("dx2", c_double), ("dx3", c_double),
                ("lam", c_double), ("mu", c_double), ("gamma", c_double), ("beta", c_double),
                ("interval", c_double), ("odt", c_double), ("tscale", c_double),
                ("skip", c_int), ("ne", c_int), ("event_times", c_void_p)]


class StepInfo(ctypes.Structure):
    _fields_ = [("i", c_int), ("Dt", c_double), ("tm", c_double), ("maxwell", c_double * 5),
                ("power", c_double), ("cei", c_double), ("t", c_double), ("oi", c_int), ("status", c_int),
                ("e", c_int)]


# name -> (restype, argtypes); the list is the ABI contract checked by the tests
SIGNATURES = {
    "relax_hip_version": (ctypes.c_char_p, []),
    "relax_hip_status_string": (ctypes.c_char_p, [c_int]),
    "relax_hip_init": (c_int, [c_int]),
    "relax_fft3": (c_int, [P, c_int, c_int, c_int, c_double, c_double, c_double, c_int, c_int]),
    "relax_fft2": (c_int, [P, c_int, c_int, c_double, c_double, c_int, c_int]),
    "relax_elasticresponse": (c_int, [c_double, c_double, P, P, P, c_int, c_int, c_int, c_double, c_double,
                                      c_double]),
    "relax_surfacetractioncowling": (c_int, [c_double, c_double, c_double, P, P, P, c_int, c_int, c_int, c_double,
                                             c_double, c_double, P, P, P]),
    "relax_cerrutimodified": (c_int, [P, P, P, c_double, c_double, c_double, P, P, P, c_int, c_int, c_int,
                                      c_double, c_double, c_double]),
    "relax_greenfunctioncowling": (c_int, [P, P, P, P, P, P, c_int, c_int, c_int, c_double, c_double, c_double,
                                           c_double, c_double, c_double, c_int]),
    "relax_greenfunctioncowling_noalloc": (c_int, [P, P, P, P, P, P, c_int, c_int, c_int, c_double, c_double,
                                                   c_double, c_double, c_double, c_double, P]),
    "relax_stressupdate": (c_int, [P, P, P, c_double, c_double, c_double, c_double, c_double, c_int, c_int, c_int,
                                   P, c_int]),
    "relax_equivalentbodyforce": (c_int, [P, c_double, c_double, c_double, c_int, c_int, c_int, P, P, P, P, P, P,
                                          P, c_int]),
    "relax_fieldadd": (c_int, [P, P, c_long, P, P]),
    "relax_tensorfieldadd": (c_int, [P, P, c_long, P, P]),
    "relax_tensoramplitude": (c_int, [P, c_long, c_double, c_double, c_double, P]),
    "relax_source": (c_int, [c_double] * 11 + [c_int, c_int, c_int, c_double, c_double, c_double] + [P] * 6
                     + [c_int]),
    "relax_traction": (c_int, [c_double, P, c_int, c_int, c_int, c_double, c_double, c_double, c_double, P, c_int,
                               c_int]),
    "relax_momentdensityshear": (c_int, [c_double, c_double, c_double, c_double, c_double, c_double, c_double,
                                         c_double, c_double, c_double, c_double, c_int, c_int, c_int, c_double,
                                         c_double, c_double, P]),
    "relax_viscouseigenstress": (c_int, [c_double, P, c_int, P, c_int, P, c_int, c_int, c_int, c_double, c_double,
                                         c_double, P, P, P, P, c_int]),
    "relax_transienteigenstress": (c_int, [c_double, P, c_int, P, P, c_int, c_int, c_int, c_double, c_double,
                                           c_double, P, P, P, P]),
    "relax_builddgammadot0": (c_int, [c_int, c_int, c_int, c_double, c_double, c_double, c_double, c_int, P, P]),
    "relax_frictioneigenstress": (c_int, [c_double] * 9 + [P, P, c_double, P, c_int, c_int, c_int, c_double,
                                                           c_double, c_double, P, P, P, c_int]),
    "relax_monitorfriction": (c_int, [c_double] * 9 + [c_int, c_int, c_int, c_int, c_double, c_double, c_double,
                                                       P, P, P, P, c_int, c_int]),
    "relax_frictionadd": (c_int, [P, c_long, c_double]),
    "relax_model_footprint": (c_size_t, [P]),
    "relax_model_create": (c_int, [P, P]),
    "relax_model_destroy": (c_int, [P]),
    "relax_model_reset": (c_int, [P]),
    "relax_model_set_structure": (c_int, [P, c_int, P, c_int]),
    "relax_model_set_prestress": (c_int, [P, P, c_int]),
    "relax_model_set_weakzones": (c_int, [P, c_int, P, c_int]),
    "relax_model_set_transient_strain": (c_int, [P, P]),
    "relax_model_get_transient_strain": (c_int, [P, P]),
    "relax_model_set_planes": (c_int, [P, P, c_int, P]),
    "relax_model_get_patches": (c_int, [P, P]),
    "relax_model_upload": (c_int, [P, P, P, P, P, P, P, c_double, c_int, c_int, c_int]),
    "relax_model_download": (c_int, [P, P, P, P, P, P, P, P, c_int]),
    "relax_model_step": (c_int, [P, P, P]),
    "relax_model_run": (c_int, [P, c_int, P, P, P]),
    "relax_model_synchronize": (c_int, [P]),
    "relax_model_stream": (c_void_p, [P]),
    "relax_model_add_shear": (c_int, [P, P, P, c_int]),
    "relax_model_coseismic": (c_int, [P]),
    "relax_model_event": (c_int, [P, P, c_double]),
    "relax_model_set_interseismic": (c_int, [P, P]),
    "relax_model_okada": (c_int, [P, P, c_int, P, c_int, c_double]),
    "relax_grnfct_okada": (c_int, [c_double] * 12 + [c_int, c_int, c_int, c_double, c_double, c_double, P, P, P,
                                                      c_int]),
    "relax_tensilesource": (c_int, [c_double] * 11 + [c_int, c_int, c_int, c_double, c_double, c_double] + [P] * 3
                            + [c_int]),
    "relax_mogisource": (c_int, [c_double] * 6 + [c_int, c_int, c_int, c_double, c_double, c_double] + [P] * 3
                         + [c_int]),
    "relax_momentdensitytensile": (c_int, [c_double] * 11 + [c_int, c_int, c_int, c_double, c_double, c_double, P]),
    "relax_momentdensitymogi": (c_int, [c_double] * 6 + [c_int, c_int, c_int, c_double, c_double, c_double, P]),
    "relax_model_profile": (c_int, [P, c_int, P, P]),
    "relax_model_autotune": (c_int, [P, P, P]),
    "relax_model_set_tuning": (c_int, [P, P]),
    "relax_hip_meminfo": (c_int, [P, P]),
    "relax_batch_create": (c_int, [P]),
    "relax_batch_destroy": (c_int, [P]),
    "relax_batch_add": (c_int, [P, P]),
    "relax_batch_remove": (c_int, [P, P]),
    "relax_batch_size": (c_int, [P]),
    "relax_batch_step": (c_int, [P, P]),
    "relax_batch_synchronize": (c_int, [P]),
    "relax_batch_busy": (c_int, [P]),
    "relax_batch_scalars": (c_int, [P, P]),
    "relax_batch_autotune": (c_int, [P, P, P]),
    "relax_model_surface": (c_int, [P, P, P, P, c_int]),
    "relax_model_points": (c_int, [P, P, c_int, c_double, P, P]),
}

_LIB = None


class RelaxError(RuntimeError):
    def __init__(self, status, what):
        self.status = status
        super().__init__("%s: %s (status %d)" % (what, status_string(status), status))


def load():
    """Load the HIP library; raises if it has not been built."""
    global _LIB
    if _LIB is None:
        if not os.path.exists(LIBPATH):
            raise RuntimeError("relax_amd/librelax_hip.so is missing: build it with `python -m relax_amd.build`")
        lib = ctypes.CDLL(LIBPATH)
        for name, (res, args) in SIGNATURES.items():
            f = getattr(lib, name)
            f.restype = res
            f.argtypes = args
        _LIB = lib
    return _LIB


def status_string(s):
    return load().relax_hip_status_string(int(s)).decode()


def check(rc, what):
    if rc != 0:
        raise RelaxError(rc, what)


def ptr(a):
    if a is None:
        return None
    if not a.flags["C_CONTIGUOUS"]:
        raise ValueError("arrays must be C-contiguous")
    return a.ctypes.data


def f32(a, what):
    if a.dtype != np.float32:
        raise TypeError("%s must be float32" % what)
    return a
