"""green.f90 on the GPU (Fourier-domain arrays are (sx3, sx2, sx1+2) float32
holding sx1/2+1 interleaved complex values per row)."""
import numpy as np

from ._lib import check, f32, load, ptr


def _dims(a):
    sx3, sx2, p = a.shape
    return p - 2, sx2, sx3


def elasticresponse(lam, mu, f1, f2, f3, dx1, dx2, dx3):
    """elasticresponse, green.f90:51-101 (in place)."""
    sx1, sx2, sx3 = _dims(f1)
    check(load().relax_elasticresponse(lam, mu, ptr(f32(f1, "f1")), ptr(f32(f2, "f2")), ptr(f32(f3, "f3")),
                                       sx1, sx2, sx3, dx1, dx2, dx3), "elasticresponse")


def surfacetractioncowling(lam, mu, gamma, u1, u2, u3, dx1, dx2, dx3):
    """surfacetractioncowling, green.f90:327-377; returns (p1, p2, p3)."""
    sx1, sx2, sx3 = _dims(u1)
    p = [np.zeros((sx2, sx1 + 2), np.float32) for _ in range(3)]
    check(load().relax_surfacetractioncowling(lam, mu, gamma, ptr(f32(u1, "u1")), ptr(f32(u2, "u2")),
                                              ptr(f32(u3, "u3")), sx1, sx2, sx3, dx1, dx2, dx3,
                                              ptr(p[0]), ptr(p[1]), ptr(p[2])), "surfacetractioncowling")
    return p


def cerrutimodified(p1, p2, p3, lam, mu, gamma, u1, u2, u3, dx1, dx2, dx3):
    """cerrutimodified, green.f90:520-612 (adds to u in place)."""
    sx1, sx2, sx3 = _dims(u1)
    check(load().relax_cerrutimodified(ptr(f32(p1, "p1")), ptr(f32(p2, "p2")), ptr(f32(p3, "p3")), lam, mu, gamma,
                                       ptr(f32(u1, "u1")), ptr(f32(u2, "u2")), ptr(f32(u3, "u3")), sx1, sx2, sx3,
                                       dx1, dx2, dx3), "cerrutimodified")


def greenfunctioncowling(c1, c2, c3, t1, t2, t3, dx1, dx2, dx3, lam, mu, gamma, align_data=1):
    """greenfunctioncowling, green.f90:1008-1129 (in place on c and t).

    With align_data=1 the arrays are (sx3, sx2, sx1+2) / (sx2, sx1+2); with
    align_data=0 (ctfft build) the real data is packed at stride sx1 inside
    the same buffers."""
    sx3, sx2, p = c1.shape
    sx1 = p - 2
    check(load().relax_greenfunctioncowling(ptr(f32(c1, "c1")), ptr(f32(c2, "c2")), ptr(f32(c3, "c3")),
                                            ptr(f32(t1, "t1")), ptr(f32(t2, "t2")), ptr(f32(t3, "t3")),
                                            sx1, sx2, sx3, dx1, dx2, dx3, lam, mu, gamma, align_data),
          "greenfunctioncowling")
