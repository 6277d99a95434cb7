"""Grid limits of the model handle (host logic, no device needed):
relax_model_create rejects sizes outside powers of two in 32..1024 (sx1,
sx2) and 32..512 (sx3), and non-positive spacing or rigidity, before it
touches a device; the largest grid's footprint fits one 288 GB device."""
import ctypes

import pytest

from relax_amd import load
from relax_amd._lib import ModelDesc

RELAX_ERR_ARG, RELAX_ERR_SIZE = 10, 11


def _desc(sx1=64, sx2=64, sx3=64, dx=1.0, mu=3e4):
    d = ModelDesc()
    d.sx1, d.sx2, d.sx3 = sx1, sx2, sx3
    d.dx1 = d.dx2 = d.dx3 = dx
    d.lam, d.mu = 3e4, mu
    return d


def _create(d):
    h = ctypes.c_void_p()
    return load().relax_model_create(ctypes.byref(d), ctypes.byref(h))


@pytest.mark.parametrize("sizes", [(2048, 64, 64), (64, 2048, 64), (64, 64, 1024), (16, 64, 64), (64, 64, 16),
                                   (96, 64, 64), (64, 48, 64), (64, 64, 40), (0, 64, 64)])
def test_create_rejects_grid_sizes(sizes):
    assert _create(_desc(*sizes)) == RELAX_ERR_SIZE


@pytest.mark.parametrize("kw", [dict(dx=0.0), dict(dx=-1.0), dict(mu=0.0)])
def test_create_rejects_spacing_and_rigidity(kw):
    assert _create(_desc(**kw)) == RELAX_ERR_ARG


def test_largest_grid_fits_one_device():
    lib = load()
    small = lib.relax_model_footprint(ctypes.byref(_desc(512, 512, 256)))
    big = lib.relax_model_footprint(ctypes.byref(_desc(1024, 1024, 512)))
    assert 0 < small < big
    # eight times the points: the footprint grows with the grid
    assert 7.5 * small < big < 8.5 * small
    assert big < 288e9
