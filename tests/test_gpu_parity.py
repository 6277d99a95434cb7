"""Parity of the HIP library with the oracle, through the C ABI (GPU only).

Tolerances (relative L2, FP32): transforms and pointwise spectral routines
1e-5; stencils and rheology 1e-5; composed RK2 trajectories 1e-4 (the
tolerance stated for the fields after N steps); integer and index work
(patch sampling indices, counters) exact.
"""
import numpy as np
import pytest

import oracle as O
from _cases import oracle_initial_state, oracle_model, rel_l2, small_spec
from relax_amd import (LAYER, PATCH, TENSOR_LAYER, WEAK, RelaxError, elastic3d, fourier, friction3d, green,
                       structure, viscoelastic3d, workloads)
from relax_amd.model import Model

pytestmark = pytest.mark.gpu
LAM = MU = 3.0e4


def rnd(rng, *shape):
    return rng.standard_normal(shape).astype(np.float32)


# ------------------------------------------------------------- fourier ----
@pytest.mark.parametrize("shape", [(16, 32, 64), (64, 64, 32), (32, 128, 16), (128, 16, 256)])
@pytest.mark.parametrize("align", [1, 0])
def test_fft3_roundtrip_and_oracle(shape, align):
    rng = np.random.default_rng(11)
    sx3, sx2, sx1 = shape
    dx = (0.5, 0.75, 1.25)
    x = rnd(rng, sx3, sx2, sx1)
    ref = np.zeros((sx3, sx2, sx1 + 2), np.float32)
    ref[..., :sx1] = x
    buf = ref.copy() if align else np.concatenate([x.ravel(), np.zeros(2 * sx2 * sx3, np.float32)]).reshape(ref.shape)
    fourier.fft3(buf, sx1, sx2, sx3, *dx, -1, align_data=align)
    O.fft3(ref, sx1, sx2, sx3, *dx, -1)
    assert rel_l2(buf, ref) < 2e-6
    fourier.fft3(buf, sx1, sx2, sx3, *dx, 1, align_data=align)
    back = buf[..., :sx1] if align else buf.ravel()[:x.size].reshape(x.shape)
    assert rel_l2(back, x) < 2e-6


@pytest.mark.parametrize("shape", [(16, 32), (256, 64), (512, 512)])
def test_fft2_matches_oracle(shape):
    rng = np.random.default_rng(12)
    sx2, sx1 = shape
    a = np.zeros((sx2, sx1 + 2), np.float32)
    a[:, :sx1] = rnd(rng, sx2, sx1)
    b = a.copy()
    fourier.fft2(a, sx1, sx2, 0.5, 0.5, -1)
    O.fft2(b, sx1, sx2, 0.5, 0.5, -1)
    assert rel_l2(a, b) < 2e-6


# --------------------------------------------------------------- green ----
def spectra(rng, sx1, sx2, sx3):
    return [rnd(rng, sx3, sx2, sx1 + 2) for _ in range(3)]


@pytest.mark.parametrize("shape", [(32, 16, 64), (64, 64, 64)])
def test_elasticresponse(shape):
    rng = np.random.default_rng(13)
    sx1, sx2, sx3 = shape
    a = spectra(rng, sx1, sx2, sx3)
    b = [x.copy() for x in a]
    green.elasticresponse(LAM, MU, *a, 0.5, 0.5, 0.5)
    O.elasticresponse(LAM, MU, *b, 0.5, 0.5, 0.5)
    for x, y in zip(a, b):
        assert rel_l2(x, y) < 1e-5


@pytest.mark.parametrize("gam", [0.0, 8.33e-4, 0.05])
def test_surfacetraction_and_cerruti(gam):
    rng = np.random.default_rng(14)
    sx1, sx2, sx3 = 32, 64, 64
    u = spectra(rng, sx1, sx2, sx3)
    p = green.surfacetractioncowling(LAM, MU, gam, *u, 0.5, 0.5, 0.5)
    q = O.surfacetractioncowling(LAM, MU, gam, *u, 0.5, 0.5, 0.5)
    for x, y in zip(p, q):
        assert rel_l2(x, y) < 1e-5
    v = [x.copy() for x in u]
    green.cerrutimodified(*q, LAM, MU, gam, *u, 0.5, 0.5, 0.5)
    O.cerrutimodified(*q, LAM, MU, gam, *v, 0.5, 0.5, 0.5)
    for x, y in zip(u, v):
        assert rel_l2(x, y) < 1e-5


def body_force_case(sx1, sx2, sx3, dx, seed=15):
    s = small_spec("maxwell", sx1, sx3, dx, seed)
    s["sx2"] = sx2
    st = oracle_initial_state(dict(s, sx1=sx1, sx2=sx2, sx3=sx3))
    h3 = sx3 // 2
    c = [np.zeros((sx3, sx2, sx1 + 2), np.float32) for _ in range(3)]
    t = [np.zeros((sx2, sx1 + 2), np.float32) for _ in range(3)]
    O.equivalentbodyforce(st["tau"], dx, dx, dx, c[0][:h3], c[1][:h3], c[2][:h3], t[0], t[1], t[2])
    return c, t


@pytest.mark.parametrize("shape,align", [((64, 64, 64), 1), ((64, 32, 128), 1), ((32, 64, 64), 0)])
def test_greenfunctioncowling(shape, align):
    sx1, sx2, sx3 = shape
    dx = 0.5
    c, t = body_force_case(sx1, sx2, sx3, dx)
    rc = [x.copy() for x in c]
    rt = [x.copy() for x in t]
    O.greenfunctioncowling(*rc, *rt, dx, dx, dx, LAM, MU, 8.33e-4)
    if align:
        gc, gt = [x.copy() for x in c], [x.copy() for x in t]
    else:  # packed rows inside the same buffers (ctfft build)
        gc = [np.concatenate([x[..., :sx1].ravel(), np.zeros(2 * sx2 * sx3, np.float32)]).reshape(x.shape) for x in c]
        gt = [np.concatenate([x[:, :sx1].ravel(), np.zeros(2 * sx2, np.float32)]).reshape(x.shape) for x in t]
    green.greenfunctioncowling(*gc, *gt, dx, dx, dx, LAM, MU, 8.33e-4, align_data=align)
    for k in range(3):
        out = gc[k][..., :sx1] if align else gc[k].ravel()[:sx1 * sx2 * sx3].reshape(sx3, sx2, sx1)
        assert rel_l2(out, rc[k][..., :sx1]) < 1e-5
        tt = gt[k][:, :sx1] if align else gt[k].ravel()[:sx1 * sx2].reshape(sx2, sx1)
        assert rel_l2(tt, rt[k][:, :sx1]) < 1e-5


# ------------------------------------------------------------ stencils ----
@pytest.mark.parametrize("shape", [(64, 32, 24), (8, 16, 12), (128, 64, 32)])
@pytest.mark.parametrize("align", [1, 0])
def test_stressupdate(shape, align):
    rng = np.random.default_rng(16)
    sx1, sx2, h3 = shape
    p = sx1 + 2 if align else sx1
    v = [rnd(rng, h3, sx2, p) for _ in range(3)]
    sig = rnd(rng, h3, sx2, sx1, 6)
    ref = sig.copy()
    vp = [np.ascontiguousarray(np.pad(a[..., :sx1], ((0, 0), (0, 0), (0, 2)))) for a in v]
    elastic3d.stressupdate(*v, LAM, MU, 0.5, 0.7, 0.9, sig, align_data=align)
    O.stressupdate(*vp, LAM, MU, 0.5, 0.7, 0.9, ref)
    assert rel_l2(sig, ref) < 1e-6


@pytest.mark.parametrize("shape", [(64, 32, 24), (16, 8, 10)])
def test_equivalentbodyforce(shape):
    rng = np.random.default_rng(17)
    sx1, sx2, h3 = shape
    sig = rnd(rng, h3, sx2, sx1, 6)
    c = [rnd(rng, h3, sx2, sx1 + 2) for _ in range(3)]
    t = [rnd(rng, sx2, sx1 + 2) for _ in range(3)]
    mask = (rng.uniform(size=h3) > 0.3).astype(np.float32)
    rc, rt = [x.copy() for x in c], [x.copy() for x in t]
    elastic3d.equivalentbodyforce(sig, 0.5, 0.6, 0.7, *c, *t, mask=mask)
    O.equivalentbodyforce(sig, 0.5, 0.6, 0.7, *rc, *rt, mask=mask)
    for a, b in zip(c + t, rc + rt):
        assert rel_l2(a, b) < 1e-6


def test_field_algebra_and_amplitude():
    rng = np.random.default_rng(18)
    u, v = rnd(rng, 1000), rnd(rng, 1000)
    for c1, c2 in [(None, None), (0.5, None), (None, -2.0), (0.25, 3.0)]:
        a = u.copy()
        elastic3d.fieldadd(a, v, c1, c2)
        ref = (u if c1 is None else np.float32(c1) * u) + (v if c2 is None else np.float32(c2) * v)
        np.testing.assert_array_equal(a, ref)
    t1, t2 = rnd(rng, 50, 6), rnd(rng, 50, 6)
    for c1, c2, ref in [(0.0, 0.0, np.zeros_like(t1)), (0.0, 2.0, np.float32(2) * t2),
                        (1.5, -1.0, np.float32(1.5) * t1 + np.float32(-1) * t2), (None, None, t2 + t2)]:
        a = t1.copy()
        elastic3d.tensorfieldadd(a, t2, c1, c2)
        np.testing.assert_array_equal(a, ref)
    t = rnd(rng, 4, 5, 6, 6)
    assert abs(elastic3d.tensoramplitude(t, 0.5, 0.5, 2.0) / O.tensoramplitude(t.astype(np.float64), 0.5, 0.5,
                                                                                  2.0) - 1) < 1e-6


def test_momentdensityshear():
    sig = np.zeros((24, 64, 64, 6), np.float32)
    ref = sig.copy()
    args = (MU, 2e-3, 1.0, -2.0, 4.0, 6.0, 8.0, 0.4, 0.3, 1.1, 0.2, 0.5, 0.5, 0.5)
    elastic3d.momentdensityshear(*args, sig)
    O.momentdensityshear(*args, ref)
    assert np.abs(ref).max() > 0 and rel_l2(sig, ref) < 1e-6


@pytest.mark.parametrize("case", ["dipping", "vertical", "shallow"])
@pytest.mark.parametrize("align", [1, 0])
def test_source(case, align):
    """source(): body force of the top half and surface traction, added to
    nonzero fields; the padding columns stay as they were."""
    sx1, sx2, sx3, dx = 64, 48, 48, 0.5
    kw = dict(dipping=(1.3e-3, 0.4, -0.8, 6.0, 4.0, 6.0, 0.5, 0.6, 1.0),
              vertical=(2e-3, -1.0, 0.5, 3.0, 5.0, 7.0, 1.2, 0.01, 0.3),
              shallow=(1e-3, 0.0, 0.0, 1.0, 3.0, 5.0, -0.7, 0.9, -1.4))[case]
    slip, x, y, z, L, W, strike, dip, rake = kw
    rng = np.random.default_rng(5)
    p1 = sx1 + 2 if align else sx1
    f = [rnd(rng, sx3, sx2, p1) * 1e-3 for _ in range(3)]
    t = [rnd(rng, sx2, p1) * 1e-3 for _ in range(3)]
    fo = [np.zeros((sx3, sx2, sx1 + 2), np.float32) for _ in range(3)]
    to = [np.zeros((sx2, sx1 + 2), np.float32) for _ in range(3)]
    for a, b in zip(fo + to, f + t):
        a[..., :sx1] = b[..., :sx1]
    f0 = [a.copy() for a in f + t]
    elastic3d.source(MU, slip, x, y, z, L, W, strike, dip, rake, 0.25, dx, dx, dx, *f, *t, align_data=align)
    O.source(MU, slip, x, y, z, L, W, strike, dip, rake, 0.25, dx, dx, dx, *fo, *to)
    for a, b, c in zip(f + t, fo + to, f0):
        assert np.abs(b[..., :sx1] - c[..., :sx1]).max() > 0 or b.ndim == 2
        assert rel_l2(a[..., :sx1], b[..., :sx1]) < 1e-6
        if align:
            np.testing.assert_array_equal(a[..., sx1:], c[..., sx1:])
    np.testing.assert_array_equal(f[0][sx3 // 2:], f0[0][sx3 // 2:])  # the bottom half is not touched


@pytest.mark.parametrize("align", [1, 0])
def test_tensile_and_mogi_sources(align):
    """tensilesource and mogisource over the whole depth (their images land
    in the lower half of the grid), added to nonzero fields."""
    sx1, sx2, sx3, dx = 64, 48, 48, 0.5
    rng = np.random.default_rng(11)
    p1 = sx1 + 2 if align else sx1
    f = [rnd(rng, sx3, sx2, p1) * 1e-3 for _ in range(3)]
    fo = [np.zeros((sx3, sx2, sx1 + 2), np.float32) for _ in range(3)]
    for a, b in zip(fo, f):
        a[..., :sx1] = b[..., :sx1]
    f0 = [a.copy() for a in f]
    elastic3d.tensilesource(2.0e4, MU, 1.2e-3, 0.4, -0.8, 3.0, 4.0, 6.0, 0.5, 0.6, 0.25, dx, dx, dx, *f,
                            align_data=align)
    O.tensilesource(2.0e4, MU, 1.2e-3, 0.4, -0.8, 3.0, 4.0, 6.0, 0.5, 0.6, 0.25, dx, dx, dx, *fo)
    elastic3d.mogisource(2.0e4, MU, 2e-3, -1.0, 2.0, 2.5, dx, dx, dx, *f, align_data=align)
    O.mogisource(2.0e4, MU, 2e-3, -1.0, 2.0, 2.5, dx, dx, dx, *fo)
    for a, b, c in zip(f, fo, f0):
        assert np.abs(b[sx3 // 2:, :, :sx1] - c[sx3 // 2:, :, :sx1]).max() > 0  # the images
        assert rel_l2(a[..., :sx1], b[..., :sx1]) < 1e-6
        if align:
            np.testing.assert_array_equal(a[..., sx1:], c[..., sx1:])


def test_tensile_and_mogi_moment_densities():
    sig = np.zeros((24, 64, 64, 6), np.float32)
    ref = sig.copy()
    elastic3d.momentdensitytensile(2.0e4, MU, 2e-3, 1.0, -2.0, 4.0, 6.0, 8.0, 0.4, 0.3, 0.2, 0.5, 0.5, 0.5, sig)
    O.momentdensitytensile(2.0e4, MU, 2e-3, 1.0, -2.0, 4.0, 6.0, 8.0, 0.4, 0.3, 0.2, 0.5, 0.5, 0.5, ref)
    elastic3d.momentdensitymogi(2.0e4, MU, 3e-3, -2.0, 1.0, 5.0, 0.5, 0.5, 0.5, sig)
    O.momentdensitymogi(2.0e4, MU, 3e-3, -2.0, 1.0, 5.0, 0.5, 0.5, 0.5, ref)
    assert np.abs(ref).max() > 0 and rel_l2(sig, ref) < 1e-6


@pytest.mark.parametrize("rate", [False, True])
def test_traction(rate):
    """traction(): constant and periodic surface loads, the load and its rate."""
    sx1, sx2, dx = 96, 64, 0.5
    loads = structure.loads([(1.0, -2.0, 8.0, 5.0, 0.4, 0.0, 0.0), (-6.0, 3.0, 4.0, 6.0, -0.2, 2.5, 0.3),
                             (20.0, 14.0, 6.0, 6.0, 0.1, 0.7, -1.0)], 0.3)
    rng = np.random.default_rng(9)
    t3 = rnd(rng, sx2, sx1 + 2)
    ref = t3.copy()
    elastic3d.traction(MU, loads, dx, dx, 1.7, 0.2, t3, rate=rate)
    O.traction(loads.astype(O.SOURCE), dx, dx, 1.7, ref, Dt=0.2, rate=rate)
    assert np.abs(ref - rnd(np.random.default_rng(9), sx2, sx1 + 2)).max() > 0.01
    assert rel_l2(t3, ref) < 1e-6


# ------------------------------------------------------------ rheology ----
def layers(h3, g0=0.2, n=3.0, coh=0.1):
    st = np.zeros(h3, LAYER)
    st["gammadot0"] = np.linspace(0, 1, h3) * g0
    st["stressexponent"] = n
    st["cohesion"] = coh
    st["friction"] = 0.6
    return st


@pytest.mark.parametrize("n", [1.0, 3.0, 3.5])
def test_viscouseigenstress(n):
    rng = np.random.default_rng(19)
    h3, sx2, sx1 = 12, 16, 32
    sig = rnd(rng, h3, sx2, sx1, 6) * 50
    st = layers(h3, 0.2 if n == 1 else 2e6, n)
    pre = np.zeros(h3, TENSOR_LAYER)
    pre["t"][:, 1] = 3.0
    pre["t"][:, 0] = -1.0
    dg = (rng.uniform(size=(h3, sx2, sx1)) * 0.01).astype(np.float32)
    m, g = np.zeros_like(sig), np.zeros((h3, sx2, sx1 + 2), np.float32)
    rm, rg = m.copy(), g.copy()
    tm = viscoelastic3d.viscouseigenstress(MU, st, pre, sig, m, 1e7, dg, g)
    rtm = O.viscouseigenstress(MU, st, pre, sig, rm, 1e7, dg, rg)
    assert rel_l2(m, rm) < 1e-5 and rel_l2(g, rg) < 1e-5
    assert abs(tm / rtm - 1) < 1e-5


def test_viscous_errors_and_ignored_structure():
    sig = np.ones((4, 16, 16, 6), np.float32)
    with pytest.raises(RelaxError) as e:
        viscoelastic3d.viscouseigenstress(MU, layers(4, n=0.5), np.zeros(4, TENSOR_LAYER), sig, np.zeros_like(sig))
    assert e.value.status == 1
    m = np.zeros_like(sig)  # a structure of the wrong size is ignored (viscoelastic3d.f90:240-241)
    viscoelastic3d.viscouseigenstress(MU, layers(3), np.zeros(3, TENSOR_LAYER), sig, m)
    assert np.abs(m).max() == 0


@pytest.mark.parametrize("n", [1.0, 3.0])
def test_transienteigenstress(n):
    """Rate, moment and relaxation time of the Burgers element with a weak
    zone field; the corrector's call with one array as strain and rate."""
    rng = np.random.default_rng(23)
    h3, sx2, sx1 = 12, 16, 32
    sig = rnd(rng, h3, sx2, sx1, 6) * (50 if n == 1 else 2e3)
    e = rnd(rng, h3, sx2, sx1, 6) * 1e-3
    st = layers(h3, 0.3 if n == 1 else 4e6, n)
    st["Gk"] = np.linspace(1e4, 3e4, h3)
    st["gammadot0"][3] = 0.0
    dg = (rng.uniform(size=(h3, sx2, sx1)) * (0.01 if n == 1 else 1e5)).astype(np.float32)
    dg[3] = 0.0
    m, ed = rnd(rng, h3, sx2, sx1, 6), np.full_like(sig, 5.0)
    rm, red = m.copy(), ed.copy()
    tm = viscoelastic3d.transienteigenstress(MU, st, sig, e, m, ed, 1e7, dg)
    rtm = O.transienteigenstress(MU, st, sig, e, rm, red, 1e7, dg)
    assert np.all(ed[3] == 0) and np.abs(ed).max() > 0
    assert rel_l2(m, rm) < 1e-5 and rel_l2(ed, red) < 1e-5
    assert abs(tm / rtm - 1) < 1e-5
    both, rboth = e.copy(), e.copy()
    m2, rm2 = np.zeros_like(sig), np.zeros_like(sig)
    viscoelastic3d.transienteigenstress(MU, st, sig, both, m2, both)
    O.transienteigenstress(MU, st, sig, rboth, rm2, rboth)
    assert rel_l2(both, rboth) < 1e-5 and rel_l2(m2, rm2) < 1e-5
    with pytest.raises(RelaxError) as err:
        viscoelastic3d.transienteigenstress(MU, layers(h3, n=0.5), sig, e, m, ed)
    assert err.value.status == 1
    m3 = np.zeros_like(sig)  # a structure of the wrong size is ignored (viscoelastic3d.f90:144)
    viscoelastic3d.transienteigenstress(MU, layers(h3 - 1), sig, e, m3, ed)
    assert np.abs(m3).max() == 0


def test_builddgammadot0():
    """Boxes (tapered, rotated), cylinders and a sphere: the device field
    equals the oracle's to FP32 rounding, and with no taper the box
    indicator agrees point for point."""
    zones = np.zeros(5, WEAK)
    zones[0] = (0.5, 1.0, -2.0, 5.0, 4.0, 6.0, 3.0, 0.4, 0.3, (0,) * 6)
    zones[1] = (0.2, -6.0, 4.0, 8.0, 5.0, 7.0, -4.0, 1.2, 0.0, (0,) * 6)
    zones[2] = (1.5, 3.0, 3.0, 6.0, -6.0, 4.0, -5.0, 0.0, 0.7, (0,) * 6)
    zones[3] = (0.7, 0.0, 0.0, 4.0, -6.0, -6.0, -6.0, 0.0, 0.0, (0,) * 6)
    zones[4] = (2.0, -3.0, -5.0, 10.0, 6.0, 3.0, 5.0, 2.5, 1.1, (0,) * 6)
    sx1, sx2, sx3, dx = 64, 32, 16, (0.5, 0.75, 1.0)
    for beta in (0.2, 0.0):
        f = viscoelastic3d.builddgammadot0(sx1, sx2, sx3, *dx, beta, zones)
        r = O.builddgammadot0(sx1, sx2, sx3, *dx, beta, zones.astype(O.WEAK), np.float32)
        assert np.abs(r).max() > 0
        np.testing.assert_allclose(f, r, rtol=1e-6, atol=1e-7)
    # the transient workload's deck zones, as the model builds them
    s = small_spec("transient")
    for key, mech in workloads.WEAK_KEYS:
        if not s.get(key):
            continue
        z = structure.weakzones(s[key])
        f = viscoelastic3d.builddgammadot0(s["sx1"], s["sx2"], s["sx3"] // 2, s["dx"], s["dx"], s["dx"], 0.0, z)
        r = O.builddgammadot0(s["sx1"], s["sx2"], s["sx3"] // 2, s["dx"], s["dx"], s["dx"], 0.0,
                              z.astype(O.WEAK), np.float32)
        assert (r > 0).sum() > 50
        np.testing.assert_array_equal(f, r)


def friction_layers(h3):
    st = np.zeros(h3, LAYER)
    st["gammadot0"] = 1e-3
    st["stressexponent"] = 1.0
    st["friction"] = 0.6
    return st


def test_frictioneigenstress():
    rng = np.random.default_rng(20)
    h3, sx = 32, 64
    sig = rnd(rng, h3, sx, sx, 6) * 2.0
    st = friction_layers(h3)
    args = (1.0, -2.0, 6.0, 10.0, 16.0, 0.3, 0.8, 1.2, 0.2)
    m, v = np.zeros_like(sig), np.zeros((h3, sx, sx + 2), np.float32)
    rm, rv = m.copy(), v.copy()
    tm = friction3d.frictioneigenstress(*args, sig, MU, st, 0.5, 0.5, 0.5, m, 1e7, v)
    rtm = O.frictioneigenstress(*args, sig, MU, st, 0.5, 0.5, 0.5, rm, 1e7, rv)
    assert np.abs(rm).max() > 0
    assert rel_l2(m, rm) < 1e-5 and rel_l2(v, rv) < 1e-5 and abs(tm / rtm - 1) < 1e-5
    # shear traction of 50 MPa along the rake: tau / tauc > 20 (STOP 5)
    strike, dip, rake = args[5], args[6], args[7]
    n = np.array([np.cos(dip) * np.cos(strike), -np.cos(dip) * np.sin(strike), -np.sin(dip)])
    r = np.array([np.sin(strike) * np.cos(rake) + np.cos(strike) * np.sin(dip) * np.sin(rake),
                  np.cos(strike) * np.cos(rake) - np.sin(strike) * np.sin(dip) * np.sin(rake), np.cos(dip) * np.sin(rake)])
    S = 25.0 * (np.outer(r, n) + np.outer(n, r))
    sig[...] = np.float32([S[0, 0], S[0, 1], S[0, 2], S[1, 1], S[1, 2], S[2, 2]])
    with pytest.raises(RelaxError) as e:
        friction3d.frictioneigenstress(*args, sig, MU, st, 0.5, 0.5, 0.5, np.zeros_like(sig))
    assert e.value.status == 5


def test_monitorfriction_frictionadd():
    rng = np.random.default_rng(21)
    h3, sx = 32, 64
    sig = rnd(rng, h3, sx, sx, 6)
    st = friction_layers(h3)
    patch = np.zeros((6, 8), PATCH)
    ref = np.zeros((6, 8), O.PATCH)
    args = (1.0, -2.0, 5.0, 6.0, 8.0, 0.3, 0.8, 1.2, 0.2, 2 * h3, 0.5, 0.5, 0.5)
    friction3d.monitorfriction(*args, sig, st, patch)
    O.monitorfriction(*args, sig.astype(np.float64), st, ref)
    for f in ("x1", "x2", "x3", "lx", "lz"):
        np.testing.assert_allclose(patch[f], ref[f], rtol=1e-12, atol=1e-12)
    np.testing.assert_array_equal(patch["sig"], ref["sig"])  # the same grid samples
    for f in ("v", "vss", "vds", "taus"):
        np.testing.assert_allclose(patch[f], ref[f], rtol=1e-9, atol=1e-15)
    friction3d.frictionadd(patch, 0.3)
    O.frictionadd(ref, 0.3)
    for f in ("ss", "ds", "slip"):
        np.testing.assert_allclose(patch[f], ref[f], rtol=1e-12, atol=1e-18)


# --------------------------------------------------------- RK2 (model) ----
def run_pair(s, nsteps, replay=True):
    st = oracle_initial_state(s)
    om = oracle_model(s)
    ost = {k: (v.copy() if isinstance(v, np.ndarray) else v) for k, v in st.items()}
    info = O.rk2_run(om, ost, nsteps)
    g = workloads.build(s)
    recs = g.run(nsteps, dt_replay=info[:, 0] if replay else None)
    out = g.download()
    if "epsilonik" in st:
        out["epsilonik"] = g.transient_strain()
    patches = g.get_patches() if om.get("patches") is not None else None
    g.close()
    return st, ost, info, out, recs, om, patches


@pytest.mark.parametrize("kind,nsteps", [("maxwell", 6), ("powerlaw", 4), ("afterslip", 4), ("ltransient", 4),
                                         ("transient", 4)])
def test_rk2_trajectory_matches_oracle(kind, nsteps):
    s = small_spec(kind)
    st, ost, info, out, recs, om, patches = run_pair(s, nsteps)
    assert len(recs) == nsteps
    for k in ("u1", "u2", "u3"):
        assert rel_l2(out[k][..., :s["sx1"]], ost[k][..., :s["sx1"]]) < 1e-4, k
    for k in ("sig", "tau"):
        assert rel_l2(out[k], ost[k]) < 1e-4, k
    if np.abs(ost["gamma"]).max() > 0:
        assert rel_l2(out["gamma"][..., :s["sx1"]], ost["gamma"][..., :s["sx1"]]) < 1e-4
    if "epsilonik" in ost:
        assert np.abs(ost["epsilonik"]).max() > 0
        assert rel_l2(out["epsilonik"], ost["epsilonik"]) < 1e-4
    for k, r in enumerate(recs):
        assert abs(r["power"] / info[k, 5] - 1) < 1e-4
        assert abs(r["cei"] / info[k, 6] - 1) < 1e-4
        assert abs(r["t"] - info[k, 7]) <= 1e-9 * abs(info[k, 7])
    if patches is not None:
        ref = om["patches"]
        assert np.abs(ref["v"]).max() > 0
        assert rel_l2(patches["slip"], ref["slip"]) < 1e-4
        assert rel_l2(patches["v"], ref["v"]) < 1e-4


@pytest.mark.parametrize("kind,dx", [("powerlaw", (0.8, 1.0, 0.6)), ("afterslip", (1.0, 0.75, 0.5))])
def test_rk2_trajectory_with_unequal_spacing(kind, dx):
    """dx1, dx2 and dx3 all different: the strain and divergence stencils,
    the wavenumbers, the per-depth tables (sampled every dx3), the friction
    plane's patch grid (px2 = fix(L/dx2), px3 = fix(W/dx3)) and the fault
    creep relaxation time (sqrt(dx1 dx3), friction3d.f90:223) each see their
    own axis; the adaptive trajectory follows the oracle's."""
    s = small_spec(kind, dx=dx)
    st, ost, info, out, recs, om, patches = run_pair(s, 4, replay=False)
    assert len(recs) == 4
    for k, r in enumerate(recs):
        assert abs(r["Dt"] / info[k, 0] - 1) < 1e-5, k
    for k in ("u1", "u2", "u3"):
        assert rel_l2(out[k][..., :s["sx1"]], ost[k][..., :s["sx1"]]) < 1e-4, k
    for k in ("sig", "tau", "gamma"):
        assert rel_l2(out[k][..., :s["sx1"]] if k == "gamma" else out[k],
                      ost[k][..., :s["sx1"]] if k == "gamma" else ost[k]) < 1e-4, k
    if patches is not None:
        ref = om["patches"]
        assert np.abs(ref["v"]).max() > 0
        assert rel_l2(patches["v"], ref["v"]) < 1e-4


@pytest.mark.parametrize("kind", ["maxwell", "elastic"])
def test_interseismic_loading_trajectory_matches_oracle(kind):
    """Inter-seismic loading (relax.f90:669-674, :795-799): a shear and a
    tensile loading segment under a Maxwell layer, and with nothing relaxing
    (tm = interval/20); the adaptive trajectory, the power (which includes the
    loading moment) and C:E^i follow the oracle's."""
    s = small_spec("maxwell")
    s["inter"] = [(1e-3, -8.0, 2.0, 0.0, 12.0, 5.0, 10.0, 70.0, 30.0), (5e-4, 3.0, -6.0, 1.0, 0.5, 6.0, 90.0, 50.0, 0.0)]
    s["inter_tensile"] = [(4e-4, 4.0, 5.0, 2.0, 5.0, 3.0, 0.0, 90.0)]
    if kind == "elastic":
        s["linear"] = None
        s["interval"] = 40.0
    st, ost, info, out, recs, om, patches = run_pair(s, 4, replay=False)
    assert len(recs) == 4
    if kind == "elastic":
        np.testing.assert_allclose([r["tm"] for r in recs], s["interval"] / 20.0)
    for k, r in enumerate(recs):
        assert abs(r["Dt"] / info[k, 0] - 1) < 1e-5, k
        assert abs(r["power"] / info[k, 5] - 1) < 1e-4
        assert abs(r["cei"] / info[k, 6] - 1) < 1e-4
    for k in ("u1", "u2", "u3"):
        assert rel_l2(out[k][..., :s["sx1"]], ost[k][..., :s["sx1"]]) < 1e-4, k
        d, do = (out[k] - st[k])[..., :s["sx1"]], (ost[k] - st[k])[..., :s["sx1"]]
        assert rel_l2(d, do) < 1e-4, k
    for k in ("sig", "tau"):
        assert rel_l2(out[k], ost[k]) < 1e-4, k
        assert rel_l2(out[k] - st[k], ost[k] - st[k]) < 1e-4, k


@pytest.mark.parametrize("kind", ["maxwell", "powerlaw", "afterslip", "transient"])
def test_rk2_adaptive_step_matches_oracle(kind):
    """Without replay the device-side Dt (min 1/gammadot over the five
    mechanisms, integrationstep) follows the oracle's to FP32 accuracy."""
    s = small_spec(kind)
    st, ost, info, out, recs, om, patches = run_pair(s, 3, replay=False)
    for k, r in enumerate(recs):
        assert abs(r["Dt"] / info[k, 0] - 1) < 1e-5
        for j, col in enumerate((2, 3, 4, 8, 9)):
            assert abs(r["maxwell"][j] / info[k, col] - 1) < 1e-5, (k, j)
    assert rel_l2(out["sig"], ost["sig"]) < 1e-4


OKADA_SOURCES = [
    # slip, opening, x, y, z, width, length, strike, dip, rake (deck convention, degrees)
    (1e-3, 0.0, -3.0, 2.0, 0.0, 4.0, 6.0, 0.0, 90.0, 0.0),        # vertical, corners on grid points
    (2e-3, 0.0, 1.5, -2.5, 1.0, 3.0, 5.0, 37.0, 25.0, 80.0),      # shallow thrust
    (0.0, 5e-4, -4.0, -4.0, 2.0, 2.0, 2.0, 120.0, 60.0, 0.0),     # tensile crack
    (1e-3, 0.0, 2.0, 3.0, 1.0, 2.0, 3.0, 90.0, 89.99999, 30.0),   # cos(dip) below 1e-6
]


@pytest.mark.parametrize("align", [1, 0])
def test_grnfct_okada_matches_oracle(align):
    """grnfct_okada (green_space.f90:119-146) through the per-routine entry
    against the oracle's restatement of dc3d.f, on a grid whose points fall on
    fault edges and edge extensions (dc3d.f's singular cases)."""
    from relax_amd import okada

    sx1 = sx2 = 32
    sx3 = 16
    dx = (1.0, 1.0, 0.5)
    for src in OKADA_SOURCES:
        ref = [np.zeros((sx3, sx2, sx1 + 2), np.float32) for _ in range(3)]
        O.grnfct_okada(LAM, MU, *src, *dx, *ref)
        u = [np.zeros((sx3, sx2, sx1 + 2 if align else sx1), np.float32) for _ in range(3)]
        okada.grnfct_okada(LAM, MU, *src, *dx, *u, align_data=align)
        for k in range(3):
            a, b = u[k][..., :sx1], ref[k][..., :sx1]
            assert np.abs(b).max() > 0
            assert rel_l2(a, b) < 1e-6, (src, k, rel_l2(a, b))
            # the same points are singular (no displacement) on both sides
            np.testing.assert_array_equal(a == 0, b == 0)


@pytest.mark.parametrize("kind", ["powerlaw", "loads", "tensile", "okada"])
def test_coseismic_setup_matches_oracle(kind):
    """The first event on the device (dislocations, traction, Green, stress)
    against the oracle's; "loads" adds a patch below the nyquist size (moment
    density only) and two surface loads; "tensile" adds two tensile cracks
    (one below the nyquist size) and two point dilatation sources, which
    take the full-depth Green solve."""
    s = small_spec("powerlaw")
    if kind in ("loads", "tensile"):
        s["deck"] = s["deck"] + [(2e-3, 3.0, -4.0, 2.0, 0.8, 0.9, 40.0, 60.0, 20.0)]
        s["loads"] = [(2.0, 1.0, 10.0, 6.0, 0.3, 0.0, 0.0), (-8.0, -5.0, 6.0, 6.0, -0.1, 4.0, 0.5)]
    if kind == "tensile":
        s["tensile"] = [(1.5e-3, -6.0, 4.0, 1.0, 6.0, 4.0, 30.0, 70.0), (1e-3, 5.0, 5.0, 3.0, 0.9, 0.8, 0.0, 90.0)]
        s["mogi"] = [(4e-3, 8.0, -6.0, 4.0), (-2e-3, -9.0, -9.0, 1.5)]
    if kind == "okada":
        # a 6 x 4 km thrust of 1 km patches on the 1 km grid, below the
        # Nyquist size: their displacement is Okada's (relax.f90:357-360)
        rng = np.random.default_rng(5)
        dip = 25.0
        s["deck"] = [(rng.uniform(0, 2e-3), -3.0 + j2, -j3 * np.cos(np.radians(dip)), 1.0 + j3 * np.sin(np.radians(dip)),
                      1.0, 1.0, 10.0, dip, 90.0) for j3 in range(4) for j2 in range(6)]
    st = oracle_initial_state(s)
    g = workloads.build(s)
    out = g.download()
    g.close()
    assert rel_l2(out["tau"], st["tau"]) < 1e-6
    for k in ("u1", "u2", "u3"):
        assert rel_l2(out[k][..., :s["sx1"]], st[k][..., :s["sx1"]]) < 1e-5
    assert rel_l2(out["sig"], st["sig"]) < 1e-5


def test_periodic_load_trajectory_matches_oracle():
    """Time-dependent surface loads: each stage of the step adds the rate of
    the current event's periodic loads to the surface traction (relax.f90:
    684-685, 805-806); the trajectory with a Maxwell layer follows the
    oracle's.  The periodic load has phase 0, so it adds nothing at t = 0:
    the same model without it has the same coseismic state, and its
    trajectory differs only through the rate terms of the steps."""
    s = small_spec("maxwell")
    s["loads"] = [(2.0, -1.0, 12.0, 8.0, 0.5, 7.0, 0.0), (-6.0, 5.0, 6.0, 6.0, 0.2, 0.0, 0.0)]
    nsteps = 5
    st = oracle_initial_state(s)
    om = oracle_model(s)
    info = O.rk2_run(om, st, nsteps)
    g = workloads.build(s)
    setup = g.download()
    recs = g.run(nsteps, dt_replay=info[:, 0])
    out = g.download()
    g.close()
    np.testing.assert_allclose([r["t"] for r in recs], info[:, 7], rtol=1e-12)
    for k in ("u1", "u2", "u3"):
        assert rel_l2(out[k][..., :s["sx1"]], st[k][..., :s["sx1"]]) < 1e-4, k
    for k in ("sig", "tau"):
        assert rel_l2(out[k], st[k]) < 1e-4, k
    # without the periodic load: the same set-up, bit for bit ...
    s2 = dict(s, loads=[s["loads"][1]])
    g = workloads.build(s2)
    setup2 = g.download()
    for k in ("u1", "u2", "u3", "sig", "tau"):
        np.testing.assert_array_equal(setup2[k], setup[k], err_msg=k)
    # ... and a different trajectory: the load's rate moved the surface
    g.run(nsteps, dt_replay=info[:, 0])
    u3 = g.download()["u3"]
    g.close()
    du = out["u3"][..., :s["sx1"]] - setup["u3"][..., :s["sx1"]]
    du2 = u3[..., :s["sx1"]] - setup["u3"][..., :s["sx1"]]
    assert rel_l2(du2, du) > 1e-2


def test_output_time_truncation_c0():
    """c0 literally: the Maxwell layer lies below the grid, so tm = STEP_MAX
    and Dt lands on the output times: 8 steps of 1.25 yr to 10 yr."""
    s = workloads.spec("c0")
    g = workloads.build(s)
    recs = g.run(20)
    g.close()
    assert len(recs) == 8
    np.testing.assert_allclose([r["t"] for r in recs], 1.25 * np.arange(1, 9), rtol=1e-12)
    np.testing.assert_allclose([r["maxwell"][0] for r in recs], 1e7)


def test_rk2_is_bitwise_reproducible():
    s = small_spec("afterslip")
    outs = []
    for _ in range(2):
        g = workloads.build(s)
        g.run(3)
        outs.append(g.download())
        g.close()
    for k in ("u1", "sig", "tau", "gamma"):
        np.testing.assert_array_equal(outs[0][k], outs[1][k])


def test_uniform_decay_at_full_size():
    """512x512x256: a uniform deviatoric stress has no divergence, so u stays
    zero and sigma decays by exactly (1 - h l + (h l)^2/2) per step, h l = 0.2."""
    sx1 = sx2 = 512
    sx3 = 256
    h3 = sx3 // 2
    with Model(sx1, sx2, sx3, 0.5, 0.5, 0.5, LAM, MU, 8.33e-4) as g:
        st = np.zeros(h3, LAYER)
        st["gammadot0"] = 0.5
        st["stressexponent"] = 1.0
        g.set_prestress(np.zeros(h3, TENSOR_LAYER))
        g.set_structure(0, st)
        z = np.zeros((h3, sx2, sx1 + 2), np.float32)
        tau = np.zeros((h3, sx2, sx1, 6), np.float32)
        tau[..., 1] = -2.0
        g.upload(z, z, z, tau)
        recs = g.run(3)
        out = g.download()
    f = (1 - 0.2 + 0.02) ** 3
    np.testing.assert_allclose(out["sig"][..., 1], 2.0 * f, rtol=2e-6)
    assert np.abs(out["u1"]).max() == 0 and np.abs(out["u3"]).max() == 0
    np.testing.assert_allclose([r["Dt"] for r in recs], 0.2, rtol=1e-6)


def test_model_rejects_bad_input():
    with pytest.raises(RelaxError) as e:
        Model(48, 64, 64, 1.0, 1.0, 1.0, LAM, MU)
    assert e.value.status == 11
    with Model(64, 64, 64, 1.0, 1.0, 1.0, LAM, MU) as g:
        with pytest.raises(RelaxError) as e:
            g.set_structure(1, layers(32, n=0.5))
        assert e.value.status == 1


def test_observation_points_match_oracle():
    """relax_model_points samples u and sig where exportpoints does, and
    turns u into the user frame; a point outside the grid is an error."""
    s = small_spec("powerlaw")
    g = workloads.build(s)
    g.run(2)
    out = g.download()
    rng = np.random.default_rng(31)
    dx, sx, h3 = s["dx"], s["sx1"], s["sx3"] // 2
    pts = np.column_stack([rng.uniform(-sx / 2 * dx, (sx / 2 - 1) * dx, 40),
                           rng.uniform(-sx / 2 * dx, (sx / 2 - 1) * dx, 40),
                           rng.uniform(0, (h3 - 1) * dx, 40)])
    pts[0] = (0.25 * dx, -0.5 * dx, 1.5 * dx)  # ties in x2 and x3
    for rot in (0.0, 30.0):
        u, sig = g.points(pts, rot)
        ref = O.exportpoints(out["u1"], out["u2"], out["u3"], out["sig"], dx, dx, dx, pts, rot)
        np.testing.assert_array_equal(sig, ref[:, 3:].astype(np.float32))
        np.testing.assert_allclose(u, ref[:, :3], rtol=1e-15, atol=0)
    with pytest.raises(RelaxError) as e:
        g.points([[0.0, 0.0, h3 * dx]])
    assert e.value.status == 15
    g.close()


@pytest.mark.parametrize("kind", ["shear", "tensile"])
def test_event_switch_matches_oracle(kind):
    """A second coseismic event at t = 1.3: integrationstep cuts Dt to reach
    it, the step's event index advances, relax_model_run stops there, and
    relax_model_event applies its patches and surface load (dislocations,
    traction, u += G[f, t]); the trajectory through and after the event
    follows the oracle."""
    s = small_spec("maxwell")
    rng = np.random.default_rng(77)
    s["events"] = [(1.3, [workloads._deck_patch(rng, -8, 8, 5.0, 4.0, 2e-3) for _ in range(2)],
                    [(3.0, -2.0, 8.0, 6.0, 0.2, 5.0, 0.3)])]
    if kind == "tensile":  # a crack and a point source as well: the full-depth solve
        s["events"][0] += ([(1e-3, 4.0, -3.0, 2.0, 5.0, 3.0, 10.0, 80.0)], [(3e-3, -5.0, 6.0, 3.0)])
    nsteps = 6
    st = oracle_initial_state(s)
    om = oracle_model(s)
    info = O.rk2_run(om, st, nsteps)
    assert st["e"] == 2
    g = workloads.build(s)
    first = g.run(nsteps)
    assert first[-1]["e"] == 1 and len(first) < nsteps  # stopped at the event
    g.close()
    g = workloads.build(s)
    recs = workloads.run(g, s, nsteps, dt_replay=info[:, 0])
    out = g.download()
    g.close()
    assert [r["e"] for r in recs] == [0 if r["t"] < 1.3 - 1e-9 else 1 for r in recs]
    np.testing.assert_allclose([r["t"] for r in recs], info[:, 7], rtol=1e-12)
    for k in ("u1", "u2", "u3"):
        assert rel_l2(out[k][..., :s["sx1"]], st[k][..., :s["sx1"]]) < 1e-4, k
    for k in ("sig", "tau"):
        assert rel_l2(out[k], st[k]) < 1e-4, k
