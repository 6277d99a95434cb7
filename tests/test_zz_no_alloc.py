"""relax_greenfunctioncowling_noalloc: greenfunctioncowling on device buffers
with its traction planes in the library's static device array, checked
through the C ABI against relax_greenfunctioncowling (host buffers, the
existing entry point) on the same seeded inputs: two calls in a row, the
second on a smaller grid, so the second must not depend on what the first
left in the shared array.  The device buffers and the stream come from the
HIP runtime the library itself is linked against (plumbing)."""
import ctypes

import numpy as np
import pytest

from relax_amd import green, load

pytestmark = pytest.mark.gpu

LAM = MU = 3.0e4
GAM = 8.33e-4
RELAX_ERR_SIZE = 11
H2D, D2H = 1, 2  # hipMemcpyHostToDevice, hipMemcpyDeviceToHost


def _hip():
    load()  # the library loads the runtime; the same one is bound here
    hip = ctypes.CDLL("libamdhip64.so.7")
    hip.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
    hip.hipFree.argtypes = [ctypes.c_void_p]
    hip.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
    hip.hipStreamCreate.argtypes = [ctypes.POINTER(ctypes.c_void_p)]
    hip.hipStreamSynchronize.argtypes = [ctypes.c_void_p]
    hip.hipStreamDestroy.argtypes = [ctypes.c_void_p]
    return hip


def _ok(rc):
    assert rc == 0, "HIP call failed with %d" % rc


def _inputs(sx1, sx2, sx3, seed):
    rng = np.random.default_rng(seed)
    c = [np.zeros((sx3, sx2, sx1 + 2), np.float32) for _ in range(3)]
    for a in c:
        a[: sx3 // 2, :, :sx1] = rng.standard_normal((sx3 // 2, sx2, sx1))
    t = [np.zeros((sx2, sx1 + 2), np.float32) for _ in range(3)]
    for a in t:
        a[:, :sx1] = rng.standard_normal((sx2, sx1))
    return c, t


def _noalloc(c, t, dx):
    hip = _hip()
    sx3, sx2, p = c[0].shape
    sx1 = p - 2
    arrays = [np.ascontiguousarray(a) for a in c + t]
    dev = []
    stream = ctypes.c_void_p()
    _ok(hip.hipStreamCreate(ctypes.byref(stream)))
    try:
        for a in arrays:
            d = ctypes.c_void_p()
            _ok(hip.hipMalloc(ctypes.byref(d), a.nbytes))
            dev.append(d)
            _ok(hip.hipMemcpy(d, a.ctypes.data, a.nbytes, H2D))
        rc = load().relax_greenfunctioncowling_noalloc(*dev, sx1, sx2, sx3, dx[0], dx[1], dx[2], LAM, MU, GAM, stream)
        _ok(hip.hipStreamSynchronize(stream))
        assert rc == 0, rc
        out = [np.empty_like(a) for a in arrays]
        for o, d in zip(out, dev):
            _ok(hip.hipMemcpy(o.ctypes.data, d, o.nbytes, D2H))
    finally:
        for d in dev:
            hip.hipFree(d)
        hip.hipStreamDestroy(stream)
    return out[:3], out[3:]


def _rel(a, b):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    return np.linalg.norm(a - b) / np.linalg.norm(b)


def test_noalloc_green_matches_the_host_entry_twice():
    for (sx1, sx2, sx3), dx, seed in (((64, 64, 64), (1.0, 0.75, 0.5), 1), ((32, 32, 16), (0.5, 0.5, 0.5), 2)):
        c, t = _inputs(sx1, sx2, sx3, seed)
        ref_c = [a.copy() for a in c]
        ref_t = [a.copy() for a in t]
        green.greenfunctioncowling(*ref_c, *ref_t, *dx, LAM, MU, GAM)
        out_c, out_t = _noalloc(c, t, dx)
        h3 = sx3 // 2
        for k in range(3):
            assert np.abs(ref_c[k][:h3, :, :sx1]).max() > 0
            err = _rel(out_c[k][:h3, :, :sx1], ref_c[k][:h3, :, :sx1])
            assert err <= 1e-6, (sx1, k, err)
            assert _rel(out_c[k][..., :sx1], ref_c[k][..., :sx1]) <= 1e-6


def test_noalloc_green_rejects_grids_beyond_its_scratch():
    hip = _hip()
    d = ctypes.c_void_p()
    _ok(hip.hipMalloc(ctypes.byref(d), 64))
    try:
        rc = load().relax_greenfunctioncowling_noalloc(d, d, d, d, d, d, 2048, 64, 64, 1.0, 1.0, 1.0, LAM, MU, GAM,
                                                       None)
    finally:
        hip.hipFree(d)
    assert rc == RELAX_ERR_SIZE
