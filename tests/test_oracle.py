"""Pins of the CPU oracle (oracle/relax_oracle.cpp) -- CPU only.

The reference cannot be built in this image (every Fortran module includes the
waf-generated config.h, src/include.f90:1), and it ships no tests or golden
vectors for this path.  Each oracle function is therefore pinned by
mathematics that is independent of its code: closed forms, library routines
(numpy.fft), physical identities and invariants.  Every pin is also run with
a planted mistake switched on in the oracle (conftest.mutation) to show that
the pin fails when a plausible error is present.

Parity status per function:
  fft3 / fft2 ............... pinned (numpy.fft, delta closed form)
  elasticresponse ........... pinned (Navier operator inverts it exactly)
  surfacetractioncowling .... pinned (space-domain Hooke's law route)
  cerrutimodified ........... pinned (its spectrum is the continuous Fourier
                              transform of a depth profile (A + B z) e^(-beta z)
                              that solves the homogeneous Navier equation and
                              carries the applied traction, Cowling buoyancy
                              included, at z = 0+; plus k1=k2=0, mirror
                              symmetry, linearity)
  greenfunctioncowling ...... pinned through its parts + interior Navier
                              equation + free-surface traction reduction
  stressupdate / equivalentbodyforce ... pinned (FIR closed forms)
  viscouseigenstress ........ pinned (power-law closed forms)
  transienteigenstress ...... pinned (Kelvin-element closed forms: the
                              linear exponential approach to s'/(2 Gk) and
                              the n = 3 algebraic decay; aliasing of the
                              strain and rate arrays)
  builddgammadot0 ........... pinned (brute-force point counts of boxes and
                              spheres, rotation by the strike, linearity
                              in the zones)
  frictioneigenstress ....... pinned (invariants, delta normalisation with image)
  monitorfriction/frictionadd pinned (invariants, closed form)
  tensoramplitude ........... pinned (closed form)
  shiftedindex, exportpoints  pinned (grid points map to themselves, ties
                              round to the lower index, range; the user
                              frame at 90 degrees swaps the components)
  source .................... pinned (first moment of the body force is the
                              seismic moment mu s A (n b + b n); the
                              surface traction is half the moment density's
                              traction; the image makes a vertical
                              strike-slip source even in its depth)
  traction .................. pinned (plateau = -load; each taper integrates
                              to its length; periodic load and its rate)
  dislocations .............. pinned (the Nyquist gate: small patches add
                              only their moment density, large ones exactly
                              source()'s force; tau = momentdensityshear)
  tensilesource, mogisource . pinned (first moment of the body force is the
                              moment tensor s L W (lambda I + 2 mu n n), or
                              kappa o I; the image half is the mirror of the
                              source half)
  momentdensitytensile/mogi . pinned (the integral of the moment density:
                              closed forms of the Gaussian and taper
                              integrals; kappa o for a point source)
  event switch .............. pinned (elastic superposition: an event of
                              slip s1 at t1 on top of a set-up of slip s0
                              equals a set-up of slip s0 + s1)
  RK2 loop .................. pinned (uniform-stress closed-form decay,
                              Burgers transient midpoint recurrence, weak
                              zones equal to the same fluidity in layers)
"""
import numpy as np
import pytest

import oracle as O

LAM = MU = 3.0e4


def band_limited(rng, shape, keep=0.25):
    """Random real field with no content near the Nyquist frequency."""
    a = rng.standard_normal(shape)
    f = np.fft.fftn(a)
    for ax, n in enumerate(shape):
        k = np.abs(np.fft.fftfreq(n))
        sl = [None] * len(shape)
        sl[ax] = slice(None)
        f = f * (k <= keep)[tuple(sl)]
    return np.fft.ifftn(f).real


def padded(real3):
    s3, s2, s1 = real3.shape
    out = np.zeros((s3, s2, s1 + 2), real3.dtype)
    out[..., :s1] = real3
    return out


def as_complex(pad):
    return pad[..., 0::2] + 1j * pad[..., 1::2]


def wavenumbers(sx1, sx2, sx3, dx1, dx2, dx3):
    k1 = np.arange(sx1 // 2 + 1) / (sx1 * dx1)
    k2 = np.fft.fftfreq(sx2, dx2)
    k3 = np.fft.fftfreq(sx3, dx3)
    return np.meshgrid(k3, k2, k1, indexing="ij")[::-1]


# ---------------------------------------------------------------- FFT -----
@pytest.mark.parametrize("shape", [(16, 8, 32), (8, 16, 16)])
def test_fft3_matches_numpy(shape, mutation):
    rng = np.random.default_rng(1)
    sx3, sx2, sx1 = shape
    dx1, dx2, dx3 = 0.5, 0.7, 1.3
    x = rng.standard_normal(shape)
    d = padded(x)
    O.fft3(d, sx1, sx2, sx3, dx1, dx2, dx3, -1)
    ref = np.fft.rfftn(x) * dx1 * dx2 * dx3
    np.testing.assert_allclose(as_complex(d), ref, rtol=1e-12, atol=1e-12)
    O.fft3(d, sx1, sx2, sx3, dx1, dx2, dx3, 1)
    np.testing.assert_allclose(d[..., :sx1], x, rtol=1e-12, atol=1e-12)
    mutation(12)
    d = padded(x)
    O.fft3(d, sx1, sx2, sx3, dx1, dx2, dx3, -1)
    assert not np.allclose(as_complex(d), ref)


def test_fft3_delta_closed_form():
    sx1, sx2, sx3, dx = 8, 8, 8, 0.25
    x = np.zeros((sx3, sx2, sx1))
    x[0, 0, 0] = 1.0
    d = padded(x)
    O.fft3(d, sx1, sx2, sx3, dx, dx, dx, -1)
    np.testing.assert_allclose(as_complex(d), np.full((sx3, sx2, sx1 // 2 + 1), dx ** 3), atol=1e-15)
    # a cosine along x2 gives two spikes at +-k2 of height N*dV/2
    x = np.cos(2 * np.pi * 2 * np.arange(sx2) / sx2)[None, :, None] * np.ones((sx3, sx2, sx1))
    d = padded(x)
    O.fft3(d, sx1, sx2, sx3, dx, dx, dx, -1)
    c = as_complex(d)
    spike = sx1 * sx2 * sx3 * dx ** 3 / 2
    assert abs(c[0, 2, 0] - spike) < 1e-12 and abs(c[0, sx2 - 2, 0] - spike) < 1e-12
    c[0, 2, 0] = c[0, sx2 - 2, 0] = 0
    assert np.abs(c).max() < 1e-12


def test_fft2_matches_numpy():
    rng = np.random.default_rng(2)
    sx1, sx2, dx1, dx2 = 16, 8, 0.5, 2.0
    x = rng.standard_normal((sx2, sx1))
    d = np.zeros((sx2, sx1 + 2))
    d[:, :sx1] = x
    O.fft2(d, sx1, sx2, dx1, dx2, -1)
    np.testing.assert_allclose(as_complex(d), np.fft.rfft2(x) * dx1 * dx2, rtol=1e-12, atol=1e-12)
    O.fft2(d, sx1, sx2, dx1, dx2, 1)
    np.testing.assert_allclose(d[:, :sx1], x, atol=1e-12)


def test_fft_float_storage_precision():
    rng = np.random.default_rng(3)
    x = rng.standard_normal((16, 16, 16)).astype(np.float32)
    d = padded(x)
    O.fft3(d, 16, 16, 16, 1.0, 1.0, 1.0, -1)
    ref = np.fft.rfftn(x.astype(np.float64))
    err = np.linalg.norm(as_complex(d) - ref) / np.linalg.norm(ref)
    assert err < 1e-7


# ------------------------------------------------------ elasticresponse ----
def navier(k1, k2, k3, u1, u2, u3, lam, mu):
    """(2 pi)^2 [mu |k|^2 I + (lam+mu) k k^T] u  -- the elastostatic operator."""
    kk = k1 ** 2 + k2 ** 2 + k3 ** 2
    kdu = k1 * u1 + k2 * u2 + k3 * u3
    c = (2 * np.pi) ** 2
    return [c * (mu * kk * u + (lam + mu) * k * kdu) for u, k in ((u1, k1), (u2, k2), (u3, k3))]


@pytest.mark.parametrize("mut", [0, 1, 2])
def test_elasticresponse_inverts_navier(mut, mutation):
    rng = np.random.default_rng(4)
    sx1, sx2, sx3, dx1, dx2, dx3 = 16, 8, 32, 0.5, 0.5, 1.0
    f = [rng.standard_normal((sx3, sx2, sx1 + 2)) for _ in range(3)]
    u = [a.copy() for a in f]
    mutation(mut)
    O.elasticresponse(LAM, 2 * MU, u[0], u[1], u[2], dx1, dx2, dx3)
    k1, k2, k3 = wavenumbers(sx1, sx2, sx3, dx1, dx2, dx3)
    U = [as_complex(a) for a in u]
    F = [as_complex(a) for a in f]
    back = navier(k1, k2, k3, *U, LAM, 2 * MU)
    assert all(abs(U[c][0, 0, 0]) == 0 for c in range(3))
    back[0][0, 0, 0] = back[1][0, 0, 0] = back[2][0, 0, 0] = 0
    for c in range(3):
        F[c][0, 0, 0] = 0
    err = max(np.abs(back[c] - F[c]).max() / np.abs(F[c]).max() for c in range(3))
    if mut == 0:
        assert err < 1e-12
    else:
        assert err > 1e-3


# ----------------------------------------------- surfacetractioncowling ---
@pytest.mark.parametrize("mut,gam", [(0, 0.0), (0, 0.05), (3, 0.0), (4, 0.05)])
def test_surfacetraction_equals_space_domain_hooke(mut, gam, mutation):
    rng = np.random.default_rng(5)
    sx1, sx2, sx3, dx1, dx2, dx3 = 16, 16, 32, 0.5, 0.8, 1.0
    u = [band_limited(rng, (sx3, sx2, sx1)) for _ in range(3)]
    # independent route: spectral derivatives of the space-domain field,
    # Hooke's law at x3 = 0, the buoyancy term, then a 2-D transform.
    K1, K2, K3 = np.meshgrid(np.fft.fftfreq(sx1, dx1), np.fft.fftfreq(sx2, dx2), np.fft.fftfreq(sx3, dx3),
                             indexing="ij")
    K1, K2, K3 = (np.transpose(a, (2, 1, 0)) for a in (K1, K2, K3))

    def d(f, K):
        return np.fft.ifftn(np.fft.fftn(f) * 2j * np.pi * K).real

    du = [[d(u[i], K) for K in (K1, K2, K3)] for i in range(3)]
    alpha = (LAM + MU) / (LAM + 2 * MU)
    s13 = MU * (du[0][2] + du[2][0])
    s23 = MU * (du[1][2] + du[2][1])
    s33 = (LAM + 2 * MU) * du[2][2] + LAM * (du[0][0] + du[1][1]) - 2 * MU * alpha * gam * u[2]
    ref = [np.fft.rfft2(s[0]) * dx1 * dx2 for s in (s13, s23, s33)]
    uh = [padded(a) for a in u]
    for a in uh:
        O.fft3(a, sx1, sx2, sx3, dx1, dx2, dx3, -1)
    mutation(mut)
    p = O.surfacetractioncowling(LAM, MU, gam, uh[0], uh[1], uh[2], dx1, dx2, dx3)
    err = max(np.abs(as_complex(p[c]) - ref[c]).max() / np.abs(ref[c]).max() for c in range(3))
    if mut == 0:
        assert err < 1e-10
    else:
        assert err > 1e-3


# ------------------------------------------------------ cerrutimodified ----
def _cerruti(P, sx3, dx, gam, sx1=16, sx2=16):
    u = [np.zeros((sx3, sx2, sx1 + 2)) for _ in range(3)]
    O.cerrutimodified(P[0], P[1], P[2], LAM, MU, gam, u[0], u[1], u[2], dx, dx, dx)
    return [as_complex(a) for a in u]


def test_cerruti_zero_horizontal_wavenumber_is_untouched():
    rng = np.random.default_rng(6)
    P = [rng.standard_normal((16, 18)) for _ in range(3)]
    U = _cerruti(P, 16, 1.0, 0.01)
    for c in range(3):
        assert np.all(U[c][:, 0, 0] == 0)
        assert np.abs(U[c][:, 1:, :]).max() > 0


def test_cerruti_mirror_symmetry_x2():
    """Reflecting x2 -> -x2 maps (p1,p2,p3)(k2) -> (p1,-p2,p3)(-k2) and the
    correction must follow: (u1,-u2,u3)(k1,-k2,k3)."""
    rng = np.random.default_rng(7)
    sx2 = 16
    P = [rng.standard_normal((sx2, 18)) for _ in range(3)]
    for a in P:
        a[sx2 // 2] = 0  # the Nyquist row has no mirror image on the grid
    idx = (-np.arange(sx2)) % sx2
    Pr = [P[0][idx], -P[1][idx], P[2][idx]]
    U = _cerruti(P, 16, 1.0, 0.02)
    Ur = _cerruti(Pr, 16, 1.0, 0.02)
    for c, s in ((0, 1), (1, -1), (2, 1)):
        np.testing.assert_allclose(Ur[c], s * U[c][:, idx, :], rtol=1e-12, atol=1e-20)


def _cerruti_surface_solution(P, gam, i1=1, i2=2, sx1=8, sx2=8, sx3=4096, dx=1.0, dx3=0.01):
    """Depth profile of the correction for one horizontal wavenumber.

    cerrutimodified returns samples of the continuous Fourier transform over
    x3 (the depth z) of a profile that, for z > 0, must be (A + B z) e^(-beta z)
    with beta = 2 pi |k_h|: the decaying solutions of the homogeneous Navier
    equation.  The profile is recovered by an inverse DFT on a fine, long
    depth grid and fitted by that family; returns (A, B, fit residual, beta,
    a1, a2) with a_j = 2 pi k_j."""
    p = [np.zeros((sx2, sx1 + 2)) for _ in range(3)]
    for c in range(3):
        p[c][i2, 2 * i1], p[c][i2, 2 * i1 + 1] = P[c].real, P[c].imag
    u = [np.zeros((sx3, sx2, sx1 + 2)) for _ in range(3)]
    O.cerrutimodified(p[0], p[1], p[2], LAM, MU, gam, u[0], u[1], u[2], dx, dx, dx3)
    prof = [np.fft.ifft(a[:, i2, 2 * i1] + 1j * a[:, i2, 2 * i1 + 1]) / dx3 for a in u]
    a1, a2 = 2 * np.pi * i1 / (sx1 * dx), 2 * np.pi * i2 / (sx2 * dx)
    beta = np.hypot(a1, a2)
    z = np.arange(sx3) * dx3
    sel = (z > 0.5) & (z < 5.0)
    M = np.stack([np.exp(-beta * z[sel]), z[sel] * np.exp(-beta * z[sel])], 1)
    AB = [np.linalg.lstsq(M, c[sel], rcond=None)[0] for c in prof]
    res = max(np.abs(M @ AB[c] - prof[c][sel]).max() / np.abs(prof[c][sel]).max() for c in range(3))
    A = np.array([x[0] for x in AB])
    B = np.array([x[1] for x in AB])
    return A, B, res, beta, a1, a2


@pytest.mark.parametrize("mut,gam", [(0, 0.0), (0, 0.05), (5, 0.05), (6, 0.0), (7, 0.0)])
def test_cerruti_solves_the_surface_traction_problem(mut, gam, mutation):
    """Boussinesq-Cerruti check: the corrected displacement obeys
    mu (u'' - beta^2 u) + (lam+mu) grad div u = 0 for z > 0, and its traction
    at z = 0+, with the buoyancy term -2 mu alpha gamma u3 of the surface
    traction (green.f90:327-377), equals the applied p."""
    P = np.array([0.3 + 0.2j, -0.7 + 0.1j, 0.5 - 0.4j])
    mutation(mut)
    A, B, res, beta, a1, a2 = _cerruti_surface_solution(P, gam)
    mutation(0)
    assert res < 1e-5  # the profile is of the form (A + B z) e^(-beta z)
    alpha = (LAM + MU) / (LAM + 2 * MU)
    d = B - beta * A  # u'(0+)
    t = np.array([MU * (d[0] + 1j * a1 * A[2]), MU * (d[1] + 1j * a2 * A[2]),
                  (LAM + 2 * MU) * d[2] + LAM * 1j * (a1 * A[0] + a2 * A[1]) - 2 * MU * alpha * gam * A[2]])
    # Navier operator on (A + B z) e^(-beta z): e^(-beta z) and z e^(-beta z) parts
    ia = np.array([1j * a1, 1j * a2])
    divA, divB = ia @ A[:2] + d[2], ia @ B[:2] - beta * B[2]
    nav0 = -2 * beta * MU * B + (LAM + MU) * np.array([1j * a1 * divA, 1j * a2 * divA, divB - beta * divA])
    nav1 = (LAM + MU) * np.array([1j * a1 * divB, 1j * a2 * divB, -beta * divB])
    nav = max(np.abs(nav0).max(), np.abs(nav1).max()) / (MU * beta * np.abs(A).max())
    err = np.abs(t - P).max() / np.abs(P).max()
    if mut == 0:
        assert nav < 1e-5
        assert err < 1e-4
    else:
        assert err > 1e-2


def test_cerruti_linear_in_traction():
    rng = np.random.default_rng(8)
    P = [rng.standard_normal((16, 18)) for _ in range(3)]
    Q = [rng.standard_normal((16, 18)) for _ in range(3)]
    a = _cerruti(P, 16, 1.0, 0.02)
    b = _cerruti(Q, 16, 1.0, 0.02)
    ab = _cerruti([P[c] + 2 * Q[c] for c in range(3)], 16, 1.0, 0.02)
    for c in range(3):
        np.testing.assert_allclose(ab[c], a[c] + 2 * b[c], rtol=1e-10, atol=1e-18)


# ------------------------------------------------- greenfunctioncowling ----
def _moment_source(sx1, sx2, h3, dx, depth=8.0, wh=4.0, wv=2.0):
    z = np.arange(h3) * dx
    x = (np.arange(sx1) - sx1 // 2) * dx
    X3, X2, X1 = np.meshgrid(z, x, x, indexing="ij")
    g = np.exp(-(X1 ** 2 + X2 ** 2) / (2 * wh ** 2) - (X3 - depth) ** 2 / (2 * wv ** 2))
    g = np.roll(np.roll(g, -sx1 // 2, 2), -sx2 // 2, 1)
    m = np.zeros((h3, sx2, sx1, 6))
    m[..., 1] = g
    m[..., 2] = 0.3 * g
    m[..., 5] = -0.2 * g
    return m


def _green_surface_ratio(mut, mutation, full=True):
    sx1 = sx2 = 64
    sx3, dx, gam = 64, 0.5, 0.01
    h3 = sx3 // 2
    m = _moment_source(sx1, sx2, h3, dx, depth=4.0, wh=2.0, wv=1.0)
    c = [np.zeros((sx3, sx2, sx1 + 2)) for _ in range(3)]
    t = [np.zeros((sx2, sx1 + 2)) for _ in range(3)]
    O.equivalentbodyforce(m, dx, dx, dx, c[0], c[1], c[2], t[0], t[1], t[2])
    mutation(mut)
    if full:
        O.greenfunctioncowling(c[0], c[1], c[2], t[0], t[1], t[2], dx, dx, dx, LAM, MU, gam)
    else:
        for a in c:
            O.fft3(a, sx1, sx2, sx3, dx, dx, dx, -1)
        O.elasticresponse(LAM, MU, c[0], c[1], c[2], dx, dx, dx)
        for a in c:
            O.fft3(a, sx1, sx2, sx3, dx, dx, dx, 1)
    mutation(0)
    sig = -m.copy()
    O.stressupdate(c[0][:h3], c[1][:h3], c[2][:h3], LAM, MU, dx, dx, dx, sig)
    return np.abs(sig[0][..., [2, 4, 5]]).max() / np.abs(sig).max()


def test_green_reduces_free_surface_traction(mutation):
    """The surface correction must at least halve the free-surface traction
    left by the full-space solution; a sign error in p = t - p or in the
    traction makes it worse."""
    base = _green_surface_ratio(0, mutation, full=False)
    good = _green_surface_ratio(0, mutation)
    bad = _green_surface_ratio(3, mutation)
    assert good < 0.6 * base
    assert bad > base


def test_green_satisfies_navier_in_the_interior():
    """Away from the surface the displacement obeys
    mu lap u + (lam+mu) grad div u + f = 0 (spectral derivatives)."""
    sx1 = sx2 = 32
    sx3, dx = 64, 1.0
    h3 = sx3 // 2
    m = _moment_source(sx1, sx2, h3, dx, depth=12.0, wh=3.0, wv=2.5)
    c = [np.zeros((sx3, sx2, sx1 + 2)) for _ in range(3)]
    t = [np.zeros((sx2, sx1 + 2)) for _ in range(3)]
    O.equivalentbodyforce(m, dx, dx, dx, c[0], c[1], c[2], t[0], t[1], t[2])
    f = [a[..., :sx1].copy() for a in c]
    O.greenfunctioncowling(c[0], c[1], c[2], t[0], t[1], t[2], dx, dx, dx, LAM, MU, 0.0)
    u = [a[..., :sx1] for a in c]
    k1 = np.fft.fftfreq(sx1, dx)[None, None, :]
    k2 = np.fft.fftfreq(sx2, dx)[None, :, None]
    k3 = np.fft.fftfreq(sx3, dx)[:, None, None]
    U = [np.fft.fftn(a) for a in u]
    L = navier(k1, k2, k3, *U, LAM, MU)
    res = [np.fft.ifftn(L[i]).real - f[i] for i in range(3)]
    inner = slice(8, h3 - 4)
    r = max(np.abs(res[i][inner]).max() for i in range(3))
    fmax = max(np.abs(f[i]).max() for i in range(3))
    assert r < 0.05 * fmax


# ----------------------------------------------------------- stencils -----
def _sin_fields(sx1, sx2, h3, dx1, dx2, dx3, m1=1, m2=2):
    i1 = np.arange(sx1)[None, None, :]
    i2 = np.arange(sx2)[None, :, None]
    z = (np.arange(h3) * dx3)[:, None, None]
    th1, th2 = 2 * np.pi * m1 / sx1, 2 * np.pi * m2 / sx2
    return i1, i2, z, th1, th2


def fir_gain(n, th, dx):
    """Closed-form response of the differentiator chosen for n samples to
    sin(theta i): D = 2 sum_l h_l sin(l theta) / dx (antisymmetric FIR)."""
    fir7 = np.array([8.77856e-01, -2.81913e-01, +6.22696e-02, +2.82441e-02, -5.09029e-02, +4.20471e-02,
                     -1.59409e-02], np.float32).astype(np.float64)
    h = fir7 if 5 <= n <= 14 else np.array([0.5])
    return 2 * np.sum(h * np.sin(np.arange(1, len(h) + 1) * th)) / dx


@pytest.mark.parametrize("sx1,sx2,h3,mut", [(16, 32, 24, 0), (8, 16, 20, 0), (16, 32, 24, 8)])
def test_stressupdate_closed_form(sx1, sx2, h3, mut, mutation):
    dx1, dx2, dx3 = 0.5, 0.75, 1.25
    i1, i2, z, th1, th2 = _sin_fields(sx1, sx2, h3, dx1, dx2, dx3)
    a, b = 0.3, -0.7
    # u1 = sin(th1 i1)(a + b z), u2 = cos(th2 i2) z, u3 = sin(th1 i1) sin(th2 i2) (1 + 0.5 z)
    u1 = np.sin(th1 * i1) * (a + b * z) * np.ones((h3, sx2, sx1))
    u2 = np.cos(th2 * i2) * z * np.ones((h3, sx2, sx1))
    u3 = np.sin(th1 * i1) * np.sin(th2 * i2) * (1 + 0.5 * z)
    g1, g2 = fir_gain(sx1, th1, dx1), fir_gain(sx2, th2, dx2)
    e11 = np.cos(th1 * i1) * g1 * (a + b * z)
    e22 = -np.sin(th2 * i2) * g2 * z
    e33 = np.sin(th1 * i1) * np.sin(th2 * i2) * 0.5 * np.ones_like(z)
    e12 = 0.5 * (0 + 0) * z  # du1/dx2 = 0, du2/dx1 = 0
    e13 = 0.5 * (np.sin(th1 * i1) * b + np.cos(th1 * i1) * g1 * np.sin(th2 * i2) * (1 + 0.5 * z))
    e23 = 0.5 * (np.cos(th2 * i2) + np.sin(th1 * i1) * np.cos(th2 * i2) * g2 * (1 + 0.5 * z))
    sh = (h3, sx2, sx1)
    E = [np.broadcast_to(e, sh) for e in (e11, e12, e13, e22, e23, e33)]
    tr = E[0] + E[3] + E[5]
    ref = np.stack([2 * MU * E[0] + LAM * tr, 2 * MU * E[1], 2 * MU * E[2], 2 * MU * E[3] + LAM * tr,
                    2 * MU * E[4], 2 * MU * E[5] + LAM * tr], -1)
    sig = np.zeros(sh + (6,))
    mutation(mut)
    O.stressupdate(padded(u1), padded(u2), padded(u3), LAM, MU, dx1, dx2, dx3, sig)
    err = np.abs(sig - ref).max() / np.abs(ref).max()
    if mut == 0:
        assert err < 1e-12
    else:
        assert err > 1e-3


@pytest.mark.parametrize("sx1,sx2,h3,mut", [(16, 32, 24, 0), (32, 8, 20, 0), (16, 32, 24, 9)])
def test_equivalentbodyforce_closed_form(sx1, sx2, h3, mut, mutation):
    dx1, dx2, dx3 = 0.5, 0.75, 1.25
    i1, i2, z, th1, th2 = _sin_fields(sx1, sx2, h3, dx1, dx2, dx3)
    sh = (h3, sx2, sx1)
    one = np.ones(sh)
    s = np.zeros(sh + (6,))
    s[..., 0] = np.sin(th1 * i1) * (1 + z) * one          # s11
    s[..., 1] = np.cos(th2 * i2) * z * one                 # s12
    s[..., 2] = np.sin(th1 * i1) * np.cos(th2 * i2) * z * one  # s13
    s[..., 3] = np.sin(th2 * i2) * one                     # s22
    s[..., 4] = (2.0 - 0.5 * z) * np.cos(th1 * i1) * one   # s23
    s[..., 5] = np.cos(th2 * i2) * (3 * z + 1) * one       # s33
    g1, g2 = fir_gain(sx1, th1, dx1), fir_gain(sx2, th2, dx2)
    f1 = np.cos(th1 * i1) * g1 * (1 + z) - np.sin(th2 * i2) * g2 * z * 0 + \
        np.sin(th1 * i1) * np.cos(th2 * i2) * one  # d s11/dx1 + d s12/dx2 + d s13/dx3
    f1 = np.cos(th1 * i1) * g1 * (1 + z) + (-np.sin(th2 * i2) * g2 * z) + np.sin(th1 * i1) * np.cos(th2 * i2)
    f2 = 0 * one + np.cos(th2 * i2) * g2 + (-0.5) * np.cos(th1 * i1)
    f3 = np.cos(th1 * i1) * g1 * np.cos(th2 * i2) * z + 0 * one + np.cos(th2 * i2) * 3
    # d s12/dx1 = 0 and d s23/dx2 = 0 for these fields
    f2 = f2 + 0 * one
    f3 = f3 + 0 * one
    c = [np.zeros((h3, sx2, sx1 + 2)) for _ in range(3)]
    t = [np.zeros((sx2, sx1 + 2)) for _ in range(3)]
    mutation(mut)
    O.equivalentbodyforce(s, dx1, dx2, dx3, c[0], c[1], c[2], t[0], t[1], t[2])
    err = max(np.abs(c[k][..., :sx1] + np.broadcast_to(f, sh)).max() for k, f in enumerate((f1, f2, f3)))
    scale = max(np.abs(f).max() for f in (f1, f2, f3))
    if mut == 0:
        assert err < 1e-12 * max(scale, 1)
        for k, comp in enumerate((2, 4, 5)):
            np.testing.assert_allclose(t[k][:, :sx1], s[0, :, :, comp], atol=1e-14)
    else:
        assert err > 1e-3 * scale


# -------------------------------------------------------------- rheology --
def _layers(h3, gammadot0=0.1, n=1.0, cohesion=0.0):
    st = np.zeros(h3, O.LAYER)
    st["gammadot0"] = gammadot0
    st["stressexponent"] = n
    st["cohesion"] = cohesion
    st["friction"] = 0.6
    return st


@pytest.mark.parametrize("n,mut", [(1.0, 0), (3.0, 0), (3.0, 10)])
def test_viscous_power_law_closed_form(n, mut, mutation):
    h3, sx2, sx1 = 4, 4, 4
    rng = np.random.default_rng(9)
    sig = rng.standard_normal((h3, sx2, sx1, 6)) * 5.0
    st = _layers(h3, 0.2, n, cohesion=0.5)
    st["gammadot0"][0] = 0.0  # inviscid layer is skipped
    pre = np.zeros(h3, O.TENSOR_LAYER)
    mom = np.zeros_like(sig)
    gam = np.zeros((h3, sx2, sx1 + 2))
    mutation(mut)
    tm = O.viscouseigenstress(MU, st, pre, sig, mom, maxwelltime=1e7, gamma=gam)
    tr = (sig[..., 0] + sig[..., 3] + sig[..., 5]) / 3
    dev = sig.copy()
    for k in (0, 3, 5):
        dev[..., k] -= tr
    tau = np.sqrt((dev[..., 0] ** 2 + dev[..., 3] ** 2 + dev[..., 5] ** 2 +
                   2 * (dev[..., 1] ** 2 + dev[..., 2] ** 2 + dev[..., 4] ** 2)) / 2)
    tauc = np.maximum(0, tau - 0.5)
    gd = 0.2 * (tauc / MU) ** (n - 1)
    gd[0] = 0
    ref = 2 * gd[..., None] * dev
    ref[tauc <= 1e-20] = 0
    err = np.abs(mom - ref).max() / np.abs(ref).max()
    if mut:
        assert err > 1e-3
        return
    assert err < 1e-12
    np.testing.assert_allclose(mom[..., 0] + mom[..., 3] + mom[..., 5], 0, atol=1e-12)
    np.testing.assert_allclose(tm, min(1e7, (1 / gd[1:][tauc[1:] > 0]).min()), rtol=1e-6)
    np.testing.assert_allclose(gam[..., :sx1], gd * tauc / MU, rtol=1e-12, atol=1e-30)


def test_viscous_invalid_power_is_an_error():
    sig = np.ones((2, 2, 2, 6))
    st = _layers(2, 0.1, 0.5)
    with pytest.raises(RuntimeError):
        O.viscouseigenstress(MU, st, np.zeros(2, O.TENSOR_LAYER), sig, np.zeros_like(sig))


def test_viscous_prestress_cancels_itself():
    """sigma = prestress gives s + sp ... the prestress term removes the
    background: with sig = 0 the moment is -2 gdot_p sp."""
    h3 = 3
    st = _layers(h3, 0.3, 1.0)
    pre = np.zeros(h3, O.TENSOR_LAYER)
    pre["t"][:, 1] = 2.0  # s12 prestress
    sig = np.zeros((h3, 2, 2, 6))
    mom = np.zeros_like(sig)
    O.viscouseigenstress(MU, st, pre, sig, mom)
    # tau = |s + sp| = 2, gammadot = gammadot_p = 0.3 (n = 1): m = 2*0.3*0 - 2*0.3*sp
    np.testing.assert_allclose(mom[..., 1], -2 * 0.3 * 2.0, rtol=1e-12)


def _friction_setup(h3=40, sx=64, dx=0.5, dip=0.0):
    # internal convention: dip is the deviation from the vertical (input.f90:1780-1800)
    st = np.zeros(h3, O.LAYER)
    st["gammadot0"] = 1e-3   # v0
    st["stressexponent"] = 1.0  # (a-b) sigma
    st["friction"] = 0.0
    return st, dict(x=0.0, y=0.0, z=5.0, L=6.0, W=8.0, strike=0.0, dip=dip, rake=0.0, beta=0.2)


@pytest.mark.parametrize("mut", [0, 11])
def test_friction_invariants_and_image(mut, mutation):
    h3, sx, dx = 40, 64, 0.5
    st, P = _friction_setup(h3, sx, dx)
    P["z"] = 0.0  # fault reaching the surface: the image restores the lost half
    strike, dip, rake = P["strike"], P["dip"], P["rake"]
    n = np.array([np.cos(dip) * np.cos(strike), -np.cos(dip) * np.sin(strike), -np.sin(dip)])
    r = np.array([np.sin(strike) * np.cos(rake) + np.cos(strike) * np.sin(dip) * np.sin(rake),
                  np.cos(strike) * np.cos(rake) - np.sin(strike) * np.sin(dip) * np.sin(rake),
                  np.cos(dip) * np.sin(rake)])
    # uniform stress with traction along the rake: sigma = 0.5 (r n^T + n r^T)
    S = 0.5 * (np.outer(r, n) + np.outer(n, r))
    sig = np.zeros((h3, sx, sx, 6))
    sig[...] = [S[0, 0], S[0, 1], S[0, 2], S[1, 1], S[1, 2], S[2, 2]]
    mom = np.zeros_like(sig)
    vel = np.zeros((h3, sx, sx + 2))
    mutation(mut)
    O.frictioneigenstress(P["x"], P["y"], P["z"], P["L"], P["W"], strike, dip, rake, P["beta"], sig, MU, st,
                          dx, dx, dx, mom, maxwelltime=1e7, vel=vel)
    mutation(0)
    # traction t = S n = 0.5 r; shear part 0.5 r (r is in-plane); tau = 0.5
    gd = 1e-3 * 2 * np.sinh(0.5 / 1.0)
    impulse = vel[..., :sx] / (gd * np.sqrt(2 * np.pi) * dx)
    total = impulse.sum() * dx ** 3
    # the regularised delta integrates to 1 across the fault and the taper to
    # W along strike; along dip the image folds the part above the surface
    # back, so the grid sum is W * dx3 * sum_i3 2 omega(x3 / L)
    def omega(x, b=0.2):
        x = abs(x)
        if x <= (1 - 2 * b) / (1 - b) / 2:
            return 1.0
        if x < 1 / (1 - b) / 2:
            return np.cos(np.pi * ((1 - b) * x - 0.5 + b) / 2 / b) ** 2
        return 0.0

    area = P["W"] * dx * sum(2 * omega(i3 * dx / P["L"]) for i3 in range(h3))
    if mut:
        assert abs(total - area) / area > 0.2
        return
    assert abs(total - area) / area < 0.05
    tr = mom[..., 0] + mom[..., 3] + mom[..., 5]
    assert np.abs(tr).max() < 1e-12 * np.abs(mom).max()
    # m . n is parallel to the shear traction direction with size mu*impulse*gdot
    mn = np.stack([mom[..., 0] * n[0] + mom[..., 1] * n[1] + mom[..., 2] * n[2],
                   mom[..., 1] * n[0] + mom[..., 3] * n[1] + mom[..., 4] * n[2],
                   mom[..., 2] * n[0] + mom[..., 4] * n[1] + mom[..., 5] * n[2]], -1)
    np.testing.assert_allclose(mn, (MU * impulse * gd)[..., None] * r, rtol=1e-9, atol=1e-12)


def test_friction_rake_constraint_and_overflow():
    h3, sx, dx = 20, 32, 0.5
    st, P = _friction_setup(h3, sx, dx)
    sig = np.zeros((h3, sx, sx, 6))
    sig[..., 1] = -0.5  # shear traction opposite to the rake (strike-slip, rake 0)
    mom = np.zeros_like(sig)
    O.frictioneigenstress(P["x"], P["y"], P["z"], P["L"], P["W"], 0.0, 0.0, 0.0, 0.2, sig, MU, st,
                          dx, dx, dx, mom)
    assert np.abs(mom).max() == 0
    sig[..., 1] = 50.0  # tau/tauc > 20 is the reference's STOP 5
    with pytest.raises(RuntimeError):
        O.frictioneigenstress(P["x"], P["y"], P["z"], P["L"], P["W"], 0.0, 0.0, 0.0, 0.2, sig, MU, st,
                              dx, dx, dx, mom)


def test_monitorfriction_and_frictionadd():
    h3, sx, dx = 32, 32, 1.0
    st, P = _friction_setup(h3, sx, dx)
    rng = np.random.default_rng(10)
    sig = rng.standard_normal((h3, sx, sx, 6)) * 0.3
    patch = np.zeros((4, 6), O.PATCH)
    strike, dip, rake = 0.3, 1.0, 1.2
    O.monitorfriction(0.0, 0.0, 6.0, 4.0, 6.0, strike, dip, rake, 0.2, 2 * h3, dx, dx, dx, sig, st, patch)
    ok = patch["v"] != 0
    assert ok.sum() > 0
    np.testing.assert_allclose(patch["vss"][ok] ** 2 + patch["vds"][ok] ** 2, patch["v"][ok] ** 2, rtol=1e-10)
    # the sampled stress is the grid value at the nearest index of the centre
    for p in patch.ravel()[ok.ravel()][:3]:
        i1 = int(round(p["x1"] / dx)) % sx
        i2 = int(round(p["x2"] / dx)) % sx
        i3 = int(np.floor(p["x3"] / dx + 0.5 - 1e-12))
        np.testing.assert_allclose(p["sig"], sig[i3, i2, i1].astype(np.float32), rtol=1e-6)
    before = patch.copy()
    O.frictionadd(patch, 0.25)
    np.testing.assert_allclose(patch["ss"], before["ss"] + 0.25 * before["vss"])
    np.testing.assert_allclose(patch["slip"], np.hypot(patch["ss"], patch["ds"]))


def test_tensoramplitude_closed_form():
    t = np.zeros((3, 4, 5, 6))
    t[..., 1] = 2.0  # pure shear: norm = |s12| = 2
    assert abs(O.tensoramplitude(t, 0.5, 0.5, 2.0) - 60 * 2 * 0.5) < 1e-12


# ------------------------------------------------------------- RK2 loop ---
def _uniform_model(n=1.0, gd0=0.2, skip=1, odt=1.0, interval=100.0, sx=16, sx3=32):
    h3 = sx3 // 2
    st = _layers(h3, gd0, n)
    return dict(sx1=sx, sx2=sx, sx3=sx3, dx1=1.0, dx2=1.0, dx3=1.0, lam=LAM, mu=MU, gam=0.0, beta=0.2,
                linear=st if n == 1.0 else None, nonlinear=None if n == 1.0 else st,
                prestress=np.zeros(h3, O.TENSOR_LAYER), interval=interval, odt=odt, tscale=1.0, skip=skip,
                event_times=[0.0])


def _uniform_state(sx, sx3, s12, dtype=np.float64):
    h3 = sx3 // 2
    z = dict(u1=np.zeros((h3, sx, sx + 2), dtype), u2=np.zeros((h3, sx, sx + 2), dtype),
             u3=np.zeros((h3, sx, sx + 2), dtype), gamma=np.zeros((h3, sx, sx + 2), dtype),
             tau=np.zeros((h3, sx, sx, 6), dtype), sig=np.zeros((h3, sx, sx, 6), dtype))
    z["tau"][..., 1] = -s12
    z["sig"][..., 1] = s12
    return z


@pytest.mark.parametrize("mut", [0, 13])
def test_rk2_uniform_maxwell_decay(mut, mutation):
    """Uniform shear stress has no divergence, so u stays 0 and the midpoint
    scheme gives sigma_k = sigma_0 (1 - h l + (h l)^2 / 2)^k, l = 2 gdot0,
    h = Dt = tm/10 = 1/(10 gdot0)."""
    m = _uniform_model()
    s = _uniform_state(16, 32, 3.0)
    mutation(mut)
    info = O.rk2_run(m, s, 5)
    mutation(0)
    g = 1 - 0.2 + 0.02
    ref = 3.0 * g ** 5
    if mut:
        assert abs(s["sig"][..., 1].mean() - ref) > 1e-3
        return
    np.testing.assert_allclose(s["sig"][..., 1], ref, rtol=1e-12)
    np.testing.assert_allclose(s["tau"][..., 1], -ref, rtol=1e-12)
    assert np.abs(s["u1"]).max() == 0 and np.abs(s["u3"]).max() == 0
    np.testing.assert_allclose(info[:, 0], 0.5, rtol=1e-6)    # Dt = 1/(10*0.2)
    np.testing.assert_allclose(s["t"], 2.5)
    # gamma += Dt * gdot * tau_c / mu with the corrector stress, each step
    acc = sum(0.5 * 0.2 * 3.0 * g ** k * (1 - 0.2 / 2 * 2 * 0.5 / 1) / MU for k in range(5))
    acc = sum(0.5 * 0.2 * abs(3.0 * g ** k * (1 - 0.5 * 0.2)) / MU for k in range(5))
    np.testing.assert_allclose(s["gamma"][..., :16], acc, rtol=1e-10)
    # power = integral of |moment| over the half grid (corrector moment)
    vol = 16 * 16 * 16
    np.testing.assert_allclose(info[-1, 5], 2 * 0.2 * 3.0 * g ** 4 * 0.9 * vol, rtol=1e-10)


def test_rk2_power_law_matches_scalar_midpoint():
    """n = 3: the uniform field follows the scalar midpoint recurrence with
    the adaptive step Dt = 1/(10 gdot(s))."""
    m = _uniform_model(n=3.0, gd0=0.5, interval=1e9)
    s0 = 300.0
    s = _uniform_state(16, 32, s0)
    info = O.rk2_run(m, s, 4)
    x, mu = s0, MU
    for k in range(4):
        gd = 0.5 * (x / mu) ** 2
        dt = np.float64(np.float32(1 / gd)) / 10
        xp = x - dt / 2 * 2 * gd * x
        x = x - dt * 2 * 0.5 * (xp / mu) ** 2 * xp
        np.testing.assert_allclose(info[k, 0], dt, rtol=1e-12)
    np.testing.assert_allclose(s["sig"][..., 1], x, rtol=1e-10)


def test_rk2_output_time_truncation():
    """skip = 0: Dt is cut so that t lands on multiples of odt
    (integrationstep, relax.f90:1432-1439)."""
    m = _uniform_model(skip=0, odt=0.7, gd0=0.2)
    s = _uniform_state(16, 32, 3.0)
    info = O.rk2_run(m, s, 4)
    np.testing.assert_allclose(info[:, 7], [0.5, 0.7, 1.2, 1.4], rtol=1e-12)


def test_fp32_oracle_tracks_fp64():
    """The FP32-storage oracle stays within 1e-5 of the FP64 build on a
    non-uniform coupled trajectory (statement of the FP32 error floor)."""
    sx, sx3, dx = 16, 32, 1.0
    h3 = sx3 // 2
    st = _layers(h3, 0.0, 1.0)
    st["gammadot0"][2:] = 0.5
    m = dict(sx1=sx, sx2=sx, sx3=sx3, dx1=dx, dx2=dx, dx3=dx, lam=LAM, mu=MU, gam=0.001, beta=0.2,
             linear=st, prestress=np.zeros(h3, O.TENSOR_LAYER), interval=100.0, tscale=1.0, skip=1,
             event_times=[0.0])
    out = {}
    for dt in (np.float32, np.float64):
        s = _uniform_state(sx, sx3, 0.0, dt)
        tau = np.zeros((h3, sx, sx, 6), dt)
        O.momentdensityshear(MU, 1.0, 0.0, 0.0, 2.0, 4.0, 6.0, 0.3, 1.2, 0.5, 0.2, dx, dx, dx, tau)
        s["tau"][:] = tau
        s["sig"][:] = -tau
        O.rk2_run(m, s, 3, dt_replay=[0.2, 0.2, 0.2])
        out[dt] = s
    for k in ("sig", "tau", "u1", "u3"):
        a, b = out[np.float32][k].astype(np.float64), out[np.float64][k]
        assert np.linalg.norm(a - b) / np.linalg.norm(b) < 1e-5, k


# ------------------------------------------------- transient creep (Burgers) ---
def _transient_layers(h3, gd0, n, gk):
    st = _layers(h3, gd0, n)
    st["Gk"] = gk
    return st


def _shear_field(shape, s12, dtype=np.float64):
    t = np.zeros(shape + (6,), dtype)
    t[..., 1] = s12
    return t


def _integrate_transient(st, s12, e0, h, nsteps):
    """Classical RK4 in time of dE/dt = rate(E) at fixed stress, the rate
    taken from the oracle routine (one point)."""
    sig = _shear_field((1, 1, 1), s12)
    e = _shear_field((1, 1, 1), e0)

    def rate(ev):
        ed = np.zeros_like(ev)
        mom = np.zeros_like(ev)
        O.transienteigenstress(MU, st, sig, ev, mom, ed)
        return ed

    for _ in range(nsteps):
        k1 = rate(e)
        k2 = rate(e + h / 2 * k1)
        k3 = rate(e + h / 2 * k2)
        k4 = rate(e + h * k3)
        e = e + h / 6 * (k1 + 2 * k2 + 2 * k3 + k4)
    return e[0, 0, 0, 1]


@pytest.mark.parametrize("mut", [0, 14, 15])
def test_transient_linear_kelvin_closed_form(mut, mutation):
    """n = 1 at constant shear stress s: E(t) = s/(2 Gk) (1 - exp(-2 Gk gd0 t / mu))."""
    gd0, gk, s0 = 0.3, 2.0e4, 5.0
    st = _transient_layers(1, gd0, 1.0, gk)
    t_end, nsteps = 40.0, 400
    mutation(mut)
    e = _integrate_transient(st, s0, 0.0, t_end / nsteps, nsteps)
    mutation(0)
    ref = s0 / (2 * gk) * (1 - np.exp(-2 * gk * gd0 * t_end / MU))
    if mut:
        assert abs(e / ref - 1) > 1e-3
    else:
        np.testing.assert_allclose(e, ref, rtol=1e-9)


@pytest.mark.parametrize("mut", [0, 15])
def test_transient_power_law_closed_form(mut, mutation):
    """n = 3: q = s - 2 Gk E obeys dq/dt = -(2 Gk gd0 / mu^3) q^3, so
    q(t) = q0 / sqrt(1 + 2 (2 Gk gd0 / mu^3) q0^2 t)."""
    gd0, gk, s0 = 2.0, 1.0e4, 3.0e3
    st = _transient_layers(1, gd0, 3.0, gk)
    t_end, nsteps = 30.0, 600
    mutation(mut)
    e = _integrate_transient(st, s0, 0.0, t_end / nsteps, nsteps)
    mutation(0)
    c = 2 * gk * gd0 / MU ** 3
    q = s0 / np.sqrt(1 + 2 * c * s0 ** 2 * t_end)
    ref = (s0 - q) / (2 * gk)
    if mut:
        assert abs(e / ref - 1) > 1e-3
    else:
        np.testing.assert_allclose(e, ref, rtol=1e-8)


def test_transient_rate_moment_and_alias():
    """moment = 2 mu Edot; a point with gammadot0 < 1e-20 gets Edot = 0 and
    no moment; passing one array as strain and rate (the corrector's call)
    equals passing a copy; a power exponent below 1 is status 1."""
    rng = np.random.default_rng(3)
    h3, sx = 3, 4
    st = _transient_layers(h3, 0.4, 3.0, 1.5e4)
    st["gammadot0"][1] = 0.0
    sig = rng.standard_normal((h3, sx, sx, 6)) * 1e3
    e = rng.standard_normal((h3, sx, sx, 6)) * 1e-2
    mom = np.zeros_like(sig)
    ed = np.full_like(sig, 7.0)
    tm = O.transienteigenstress(MU, st, sig, e, mom, ed, maxwelltime=1e7)
    np.testing.assert_allclose(mom, 2 * MU * ed, rtol=1e-12)
    assert np.all(ed[1] == 0) and np.all(mom[1] == 0)
    assert np.all(np.abs(ed[0]) > 0)
    # 1/gammadot of every active point, gammadot = gd0 (|Q|/mu)^2
    sd = sig.copy()
    tr = (sd[..., 0] + sd[..., 3] + sd[..., 5]) / 3
    for k in (0, 3, 5):
        sd[..., k] -= tr
    q = sd - 2 * 1.5e4 * e
    nq = np.sqrt((q[..., 0] ** 2 + q[..., 3] ** 2 + q[..., 5] ** 2
                  + 2 * (q[..., 1] ** 2 + q[..., 2] ** 2 + q[..., 4] ** 2)) / 2)
    gd = 0.4 * (nq / MU) ** 2
    np.testing.assert_allclose(tm, np.float32(1 / gd[[0, 2]].max()), rtol=1e-6)
    both = e.copy()
    mom2 = np.zeros_like(sig)
    O.transienteigenstress(MU, st, sig, both, mom2, both)
    np.testing.assert_array_equal(both, ed)
    np.testing.assert_array_equal(mom2, mom)
    bad = st.copy()
    bad["stressexponent"][2] = 0.5
    with pytest.raises(RuntimeError):
        O.transienteigenstress(MU, bad, sig, e, mom, ed)


# ------------------------------------------------------------- weak zones ---
def _weak(**kw):
    z = np.zeros(1, O.WEAK)
    for k, v in kw.items():
        z[k] = v
    return z


def _coords(sx, dx):
    i = np.arange(sx)
    return np.where(i < sx // 2, i, i - sx) * dx


@pytest.mark.parametrize("mut", [0, 16])
def test_weak_box_counts_points(mut, mutation):
    """beta = 0: a box zone is the indicator of |x - c| <= extent/2 along each
    of its axes (x1: thickness, x2: length, x3: width), times dgammadot0."""
    sx1, sx2, sx3, dx = 32, 32, 16, 0.5
    z = _weak(dgammadot0=2.5, x=1.0, y=-2.0, z=3.0, thickness=3.0, length=5.0, width=2.0)
    mutation(mut)
    f = O.builddgammadot0(sx1, sx2, sx3, dx, dx, dx, 0.0, z, np.float64)
    mutation(0)
    x1, x2, x3 = _coords(sx1, dx), _coords(sx2, dx), np.arange(sx3) * dx
    ref = (np.abs(x3 - 3.0) <= 1.0)[:, None, None] * (np.abs(x2 + 2.0) <= 2.5)[None, :, None] \
        * (np.abs(x1 - 1.0) <= 1.5)[None, None, :] * 2.5
    if mut:
        assert np.abs(f - ref).max() > 1
    else:
        np.testing.assert_array_equal(f, ref)


def test_weak_sphere_and_cylinder_count_points():
    """Negative extents select spheres and cylinders, but every point is
    first gated by |x_i - c_i| <= LM, LM the largest tapered half extent
    (viscoelastic3d.f90:381-387).  A sphere (all three negative) has LM < 0
    and adds nothing; a cylinder with D, W < 0 and L > 0 is the disc
    rho <= 1/2 in x1-x2 times |x3 - z| <= L/2, clipped to the LM box."""
    sx, dx = 32, 0.5
    x1, x2, x3 = _coords(sx, dx), _coords(sx, dx), np.arange(sx // 2) * dx
    X3, X2, X1 = np.meshgrid(x3, x2, x1, indexing="ij")
    z = _weak(dgammadot0=1.0, x=0.5, y=1.0, z=4.0, thickness=-4.0, length=-4.0, width=-4.0)
    f = O.builddgammadot0(sx, sx, sx // 2, dx, dx, dx, 0.0, z, np.float64)
    assert np.all(f == 0)
    z = _weak(dgammadot0=1.0, x=0.5, y=1.0, z=4.0, thickness=-4.0, length=-4.0, width=3.0)
    f = O.builddgammadot0(sx, sx, sx // 2, dx, dx, dx, 0.0, z, np.float64)
    rho = np.sqrt(((X1 - 0.5) / 4) ** 2 + ((X2 - 1.0) / 4) ** 2)
    lm = 1.5
    gate = (np.abs(X1 - 0.5) <= lm) & (np.abs(X2 - 1.0) <= lm) & (np.abs(X3 - 4.0) <= lm)
    ref = (rho <= 0.5) * (np.abs(X3 - 4.0) <= 1.5) * gate * 1.0
    assert ref.sum() > 20
    np.testing.assert_array_equal(f, ref)


def test_weak_strike_rotates_and_zones_add():
    """strike = 90 degrees maps the zone's x1 axis onto -x2: the field is
    the unrotated one with x1 and x2 exchanged (centred zone, symmetric
    grid range); the field of two zones is the sum of their fields; the
    taper keeps dgammadot0 at the centre and decreases outwards."""
    sx, dx, beta = 32, 0.5, 0.25
    a = _weak(dgammadot0=1.0, x=0.0, y=0.0, z=3.0, thickness=2.0, length=5.0, width=2.0)
    b = a.copy()
    b["strike"] = np.pi / 2
    fa = O.builddgammadot0(sx, sx, 16, dx, dx, dx, beta, a, np.float64)
    fb = O.builddgammadot0(sx, sx, 16, dx, dx, dx, beta, b, np.float64)
    # indices 1..sx-1 hold the symmetric range -(sx/2-1)..(sx/2-1)
    sym = np.r_[np.arange(1, sx // 2), np.arange(sx // 2 + 1, sx)]
    ca = fa[:, sym][:, :, sym]
    cb = fb[:, sym][:, :, sym]
    np.testing.assert_allclose(cb, np.swapaxes(ca, 1, 2), atol=1e-12)
    c = _weak(dgammadot0=0.5, x=-3.0, y=2.0, z=5.0, thickness=3.0, length=3.0, width=3.0, dip=0.4)
    fc = O.builddgammadot0(sx, sx, 16, dx, dx, dx, beta, c, np.float64)
    fac = O.builddgammadot0(sx, sx, 16, dx, dx, dx, beta, np.concatenate([a, c]), np.float64)
    np.testing.assert_allclose(fac, fa + fc, rtol=0, atol=1e-15)
    assert fa[6, 0, 0] == 1.0  # centre (x3 = 3 at i3 = 6)
    line = fa[6, 0, : sx // 2]
    assert np.all(np.diff(line) <= 0) and line[-1] == 0.0


# ------------------------------------------------ RK2 with transient creep ---
def test_rk2_transient_midpoint_recurrence():
    """Uniform shear s12 over a linear Burgers (Kelvin) element, no Maxwell
    mechanism: u stays 0 and (sigma, E) follow the midpoint scheme of
    dsigma/dt = -2 gd0 q, dE/dt = gd0/mu q, q = sigma - 2 Gk E, with
    Dt = tm/10 = 1/(10 gd0); the scheme converges to the exact solution."""
    sx, sx3 = 8, 16
    h3 = sx3 // 2
    gd0, gk, s0 = 0.2, 1.5e4, 4.0
    m = _uniform_model(sx=sx, sx3=sx3)
    m["linear"] = None
    m["ltransient"] = _transient_layers(h3, gd0, 1.0, gk)
    s = _uniform_state(sx, sx3, s0)
    s["epsilonik"] = np.zeros((h3, sx, sx, 6))
    info = O.rk2_run(m, s, 6)
    sig, e = s0, 0.0
    dt = np.float64(np.float32(1 / gd0)) / 10
    for k in range(6):
        q = sig - 2 * gk * e
        em = e + dt / 2 * gd0 / MU * q
        sp = sig - dt / 2 * 2 * gd0 * q
        qc = sp - 2 * gk * em
        e = e + dt * gd0 / MU * qc
        sig = sig - dt * 2 * gd0 * qc
        np.testing.assert_allclose(info[k, 0], dt, rtol=1e-12)
        np.testing.assert_allclose(info[k, 8], np.float32(1 / gd0), rtol=1e-7)
    np.testing.assert_allclose(s["sig"][..., 1], sig, rtol=1e-12)
    np.testing.assert_allclose(s["epsilonik"][..., 1], e, rtol=1e-12)
    assert np.abs(s["u1"]).max() == 0 and np.abs(s["u3"]).max() == 0
    lam = 2 * gd0 * (1 + gk / MU)
    t = info[:, 7][-1]
    exact = s0 - 2 * gd0 * s0 * (1 - np.exp(-lam * t)) / lam
    # second-order scheme at lambda Dt = 0.3: about 1e-2 after six steps
    assert abs(sig / exact - 1) < 2e-2


def test_rk2_weak_zone_equals_layered_fluidity():
    """A linear structure with no fluidity plus a weak-zone field that holds
    0.2 everywhere steps exactly as the structure with gammadot0 = 0.2."""
    sx, sx3 = 8, 16
    h3 = sx3 // 2
    m = _uniform_model(sx=sx, sx3=sx3)
    s = _uniform_state(sx, sx3, 3.0)
    ref = O.rk2_run(m, s, 3)
    m2 = _uniform_model(sx=sx, sx3=sx3, gd0=0.0)
    m2["dg_linear"] = np.full((h3, sx, sx), 0.2)
    s2 = _uniform_state(sx, sx3, 3.0)
    got = O.rk2_run(m2, s2, 3)
    np.testing.assert_array_equal(got, ref)
    np.testing.assert_array_equal(s2["sig"], s["sig"])


# ------------------------------------------------ observation points ---
@pytest.mark.parametrize("mut", [0, 17])
def test_shiftedindex_nearest_point_ties_down(mut, mutation):
    """Coordinates on grid points map to their index; between two points to
    the nearer one, ties to the lower index (fix, elastic3d.f90:70-83), with
    the x1/x2 wrap of the shifted coordinates; out of range is refused."""
    sx1, sx2, sx3, dx = 32, 16, 8, 0.5
    mutation(mut)
    got = []
    for k3 in range(sx3 - 1):
        for f in (0.0, 0.25, 0.5, 0.75):
            got.append((O.shiftedindex(0.0, 0.0, (k3 + f) * dx, sx1, sx2, sx3, dx, dx, dx)[2],
                        int(np.ceil(k3 + f - 0.5))))
    mutation(0)
    bad = [(a, b) for a, b in got if a != b]
    if mut:
        assert bad
        return
    assert not bad
    for i1 in range(sx1):
        x1 = (i1 if i1 < sx1 // 2 else i1 - sx1) * dx
        assert O.shiftedindex(x1, 0.0, 0.0, sx1, sx2, sx3, dx, dx, dx)[0] == i1
    assert O.shiftedindex(-0.25 * dx, 0.0, 0.0, sx1, sx2, sx3, dx, dx, dx)[0] == 0
    assert O.shiftedindex(-0.75 * dx, 0.0, 0.0, sx1, sx2, sx3, dx, dx, dx)[0] == sx1 - 1
    assert O.shiftedindex(sx1 // 2 * dx, 0.0, 0.0, sx1, sx2, sx3, dx, dx, dx) is None
    assert O.shiftedindex(0.0, 0.0, -0.1, sx1, sx2, sx3, dx, dx, dx) is None


def test_exportpoints_user_frame():
    """At rot = 90 degrees the user frame is (v1, v2) = (-u2, u1); at 0 it is
    the computational frame; sig is sampled as it is."""
    rng = np.random.default_rng(4)
    h3, sx = 4, 8
    u = [rng.standard_normal((h3, sx, sx + 2)) for _ in range(3)]
    sig = rng.standard_normal((h3, sx, sx, 6))
    pts = np.array([[0.5, -1.0, 0.5], [1.5, 1.0, 1.0]])
    a = O.exportpoints(*u, sig, 0.5, 0.5, 0.5, pts, 0.0)
    b = O.exportpoints(*u, sig, 0.5, 0.5, 0.5, pts, 90.0)
    i = [O.shiftedindex(*p, sx, sx, h3, 0.5, 0.5, 0.5) for p in pts]
    for k, (i1, i2, i3) in enumerate(i):
        np.testing.assert_array_equal(a[k, :3], [u[0][i3, i2, i1], u[1][i3, i2, i1], u[2][i3, i2, i1]])
        np.testing.assert_array_equal(a[k, 3:], sig[i3, i2, i1])
    np.testing.assert_allclose(b[:, 0], -a[:, 1], atol=1e-15)
    np.testing.assert_allclose(b[:, 1], a[:, 0], atol=1e-15)
    np.testing.assert_array_equal(b[:, 2:], a[:, 2:])


# --------------------------------------------------------- event switch ---
def rel_l2(a, b):
    a, b = np.asarray(a, np.float64), np.asarray(b, np.float64)
    return np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300)


def _sources(patches, slips):
    out = np.zeros(len(patches), O.SOURCE)
    for k, (p, sl) in enumerate(zip(patches, slips)):
        out[k] = (sl, p["x"], p["y"], p["z"], p["width"], p["length"], p["strike"], p["dip"], p["rake"], 0, 0, 0.2, 0)
    return out


def _setup(sx, sx3, dx, patches, slips, dtype=np.float64):
    """The first event: dislocations (source() body force, moment density),
    u = G[f, t], sig = -tau + C:eps(u)."""
    h3 = sx3 // 2
    tau = np.zeros((h3, sx, sx, 6), dtype)
    c = [np.zeros((sx3, sx, sx + 2), dtype) for _ in range(3)]
    t = [np.zeros((sx, sx + 2), dtype) for _ in range(3)]
    O.dislocations(_sources(patches, slips), None, MU, 1.0, 0.0, dx, dx, dx, c, t, tau)
    O.greenfunctioncowling(c[0], c[1], c[2], t[0], t[1], t[2], dx, dx, dx, LAM, MU, 0.0)
    u = [np.ascontiguousarray(a[:h3]) for a in c]
    sig = np.ascontiguousarray(-tau)
    O.stressupdate(u[0], u[1], u[2], LAM, MU, dx, dx, dx, sig)
    return dict(u1=u[0], u2=u[1], u3=u[2], tau=tau, sig=sig, gamma=np.zeros_like(u[0]), t=0.0, oi=1, i=0)


@pytest.mark.parametrize("mut", [0, 18])
def test_event_switch_superposes_elastically(mut, mutation):
    """Without relaxation, a second event of slip s1 on the first event's
    patches, reached by the Dt truncation of integrationstep, leaves the
    state of a single set-up of slip s0 + s1; the output counter advances at
    the event."""
    sx, sx3, dx = 16, 32, 1.0
    patches = np.zeros(2, O.PLANE)
    patches[0] = (0.0, -1.0, 3.0, 4.0, 6.0, 0.3, 0.5, 1.0, 0, 0)
    patches[1] = (2.0, 2.0, 5.0, 3.0, 3.0, 1.2, 0.9, -0.4, 0, 0)
    s0, s1 = np.array([1.0, 0.5]), np.array([-0.3, 2.0])
    h3 = sx3 // 2
    m = dict(sx1=sx, sx2=sx, sx3=sx3, dx1=dx, dx2=dx, dx3=dx, lam=LAM, mu=MU, gam=0.0, beta=0.2,
             prestress=np.zeros(h3, O.TENSOR_LAYER), interval=100.0, odt=1.0, tscale=1.0, skip=1,
             event_times=[0.0, 2.5], event_sources=[_sources(patches, s1)], nyquist=1.0)
    st = _setup(sx, sx3, dx, patches, s0)
    mutation(mut)
    info = O.rk2_run(m, st, 1)
    mutation(0)
    ref = _setup(sx, sx3, dx, patches, s0 + s1)
    assert info[0, 0] == 2.5 and st["t"] == 2.5 and st["e"] == 2 and st["oi"] == 2
    err = max(rel_l2(st[k], ref[k]) for k in ("u1", "u2", "u3", "tau", "sig"))
    if mut:
        assert err > 1e-3
    else:
        assert err < 1e-12


# ------------------------------------------------------------- source() ---
def _source(mut, mutation, sx=64, sx3=64, dx=0.5, **kw):
    a = dict(slip=1.3, x=0.3, y=-0.7, z=7.0, L=4.0, W=6.0, strike=np.radians(30), dip=np.radians(40),
             rake=np.radians(60), beta=0.3)
    a.update(kw)
    f = [np.zeros((sx3, sx, sx + 2)) for _ in range(3)]
    t = [np.zeros((sx, sx + 2)) for _ in range(3)]
    mutation(mut)
    O.source(MU, a["slip"], a["x"], a["y"], a["z"], a["L"], a["W"], a["strike"], a["dip"], a["rake"], a["beta"],
             dx, dx, dx, *f, *t)
    mutation(0)
    return a, f, t


@pytest.mark.parametrize("case", ["dipping", "vertical"])
def test_source_first_moment_is_the_seismic_moment(case, mutation):
    """For a buried patch, sum x_j f_i dV over the grid is the seismic moment
    tensor mu s L W (n b + b n) (f = -div M, integrated by parts) and the
    net force vanishes; the residual is the quadrature of the tapers."""
    kw = dict(dip=np.radians(40)) if case == "dipping" else dict(dip=0.0, L=8.0, z=16.0, beta=0.45)
    sx3 = 64 if case == "dipping" else 128
    dx = 0.5
    a, f, t = _source(0, mutation, sx3=sx3, **kw)
    h3, sx = sx3 // 2, 64
    F = np.array([c[:h3, :, :sx] for c in f])
    X1 = (np.arange(sx) - sx * (np.arange(sx) >= sx // 2)) * dx
    xx = [X1[None, None, :], X1[None, :, None], (np.arange(h3) * dx)[:, None, None]]
    Mf = np.array([[(xx[j] * F[i]).sum() * dx ** 3 for j in range(3)] for i in range(3)])
    cs, ss, cd, sd = np.cos(a["strike"]), np.sin(a["strike"]), np.cos(a["dip"]), np.sin(a["dip"])
    cr, sr = np.cos(a["rake"]), np.sin(a["rake"])
    n = np.array([cd * cs, -cd * ss, -sd])
    b = np.array([ss * cr + cs * sd * sr, cs * cr - ss * sd * sr, cd * sr])
    M0 = MU * a["slip"] * a["L"] * a["W"] * (np.outer(n, b) + np.outer(b, n))
    assert np.abs(Mf - M0).max() / np.abs(M0).max() < (1e-3 if case == "dipping" else 2e-3)
    assert np.abs(F.sum(axis=(1, 2, 3)) * dx ** 3).max() < 1e-3 * np.abs(M0).max() / a["L"]
    assert all(np.abs(c).max() == 0 for c in t)  # buried: no surface traction


def test_source_surface_traction_is_half_the_moment_traction(mutation):
    """t = (n sdyad mu s b) e3 at x3 = 0: half of momentdensityshear's
    M e3 there (the same aperture, source and image)."""
    sx, sx3, dx = 64, 64, 0.5
    a, f, t = _source(0, mutation, z=2.0, dip=np.radians(30))
    M = np.zeros((sx3 // 2, sx, sx, 6))
    O.momentdensityshear(MU, a["slip"], a["x"], a["y"], a["z"], a["L"], a["W"], a["strike"], a["dip"], a["rake"],
                         a["beta"], dx, dx, dx, M)
    for i, k in enumerate((2, 4, 5)):
        assert np.abs(M[0, :, :, k]).max() > 0
        np.testing.assert_allclose(t[i][:, :sx], M[0, :, :, k] / 2, rtol=0, atol=1e-12 * np.abs(M).max())


@pytest.mark.parametrize("mut", [0, 19])
def test_source_vertical_image_symmetry(mut, mutation):
    """A vertical strike-slip patch (dip < 2.01 degrees) near the surface adds
    its image at -z: the strike-slip body force of depth z then equals that
    of depth -z exactly (source and image swap)."""
    kw = dict(z=2.0, dip=0.0, rake=0.0, L=4.0)
    _, f1, _ = _source(mut, mutation, **kw)
    _, f2, _ = _source(mut, mutation, **dict(kw, z=-2.0))
    diff = max(np.abs(a - b).max() for a, b in zip(f1, f2))
    if mut:
        assert diff > 1e-3 * max(np.abs(a).max() for a in f1)
    else:
        assert diff == 0 and max(np.abs(a).max() for a in f1) > 0


# ------------------------------------------------- traction, dislocations ---
def _load(**kw):
    a = dict(slip=0.7, x=1.5, y=-2.0, z=0.0, width=6.0, length=10.0, strike=0, dip=0, rake=0, period=0.0, phase=0.0,
             beta=0.25, opening=0)
    a.update(kw)
    out = np.zeros(1, O.SOURCE)
    out[0] = tuple(a[k] for k in O.SOURCE.names)
    return out


@pytest.mark.parametrize("mut", [0, 21])
def test_traction_plateau_and_load_integral(mut, mutation):
    """A constant load presses down: t3 = -amplitude on the plateau of the
    tapers, and its sum is -amplitude L W (each raised-cosine taper
    integrates to its length)."""
    sx, dx = 64, 0.5
    l = _load()
    t3 = np.zeros((sx, sx + 2))
    mutation(mut)
    O.traction(l, dx, dx, 0.0, t3)
    mutation(0)
    i1, i2 = int(round(1.5 / dx)), sx - int(round(2.0 / dx))
    if mut:
        assert t3[i2, i1] == pytest.approx(0.7)
        return
    assert t3[i2, i1] == -0.7
    assert t3.sum() * dx * dx == pytest.approx(-0.7 * 10.0 * 6.0, rel=1e-6)
    assert np.all(t3[:, sx:] == 0)


def test_traction_periodic_load_and_rate():
    """A periodic load is amp sin(2 pi t / T + phi); the rate form is the
    difference quotient over [t, t + Dt]."""
    sx, dx = 32, 1.0
    l = _load(period=3.0, phase=0.4, x=0.0, y=0.0)
    t, Dt = 1.1, 0.25
    a = np.zeros((sx, sx + 2))
    O.traction(l, dx, dx, t, a)
    b = np.zeros((sx, sx + 2))
    O.traction(l, dx, dx, t + Dt, b)
    r = np.zeros((sx, sx + 2))
    O.traction(l, dx, dx, t, r, Dt=Dt, rate=True)
    assert a[0, 0] == pytest.approx(-0.7 * np.sin(2 * np.pi * t / 3.0 + 0.4), rel=1e-14)
    np.testing.assert_allclose(r, (b - a) / Dt, rtol=1e-12, atol=1e-14)


@pytest.mark.parametrize("mut", [0, 20])
def test_dislocations_nyquist_gate(mut, mutation):
    """Patches up to nyquist*min(dx) in length and width get no source()
    body force (only their moment density); larger ones get exactly
    source()'s force and traction."""
    sx, sx3, dx = 32, 32, 1.0
    big = np.zeros(1, O.PLANE)
    big[0] = (0.0, 1.0, 5.0, 4.0, 6.0, 0.3, 0.5, 1.0, 0, 0)
    small = np.zeros(1, O.PLANE)
    small[0] = (-3.0, -2.0, 4.0, 1.5, 2.0, 1.0, 0.2, 0.4, 0, 0)
    h3 = sx3 // 2
    res = {}
    for name, pa in (("big", big), ("small", small)):
        tau = np.zeros((h3, sx, sx, 6))
        v = [np.zeros((sx3, sx, sx + 2)) for _ in range(3)]
        t = [np.zeros((sx, sx + 2)) for _ in range(3)]
        mutation(mut)
        O.dislocations(_sources(pa, [1.0]), None, MU, 2.0, 0.0, dx, dx, dx, v, t, tau)
        mutation(0)
        res[name] = (v, t, tau)
        M = np.zeros_like(tau)
        p = pa[0]
        O.momentdensityshear(MU, 1.0, p["x"], p["y"], p["z"], p["width"], p["length"], p["strike"], p["dip"],
                             p["rake"], 0.2, dx, dx, dx, M)
        np.testing.assert_array_equal(tau, M)
    v, t, _ = res["small"]
    if mut:
        assert max(np.abs(a).max() for a in v) > 0
        return
    assert max(np.abs(a).max() for a in v + t) == 0
    v, t, _ = res["big"]
    f = [np.zeros((sx3, sx, sx + 2)) for _ in range(3)]
    tt = [np.zeros((sx, sx + 2)) for _ in range(3)]
    p = big[0]
    O.source(MU, 1.0, p["x"], p["y"], p["z"], p["width"], p["length"], p["strike"], p["dip"], p["rake"], 0.2, dx, dx,
             dx, *f, *tt)
    for a, b in zip(v + t, f + tt):
        np.testing.assert_array_equal(a, b)
    assert max(np.abs(a).max() for a in v) > 0


# ------------------------------------------------- tensile cracks, Mogi ---
def _first_moment(F, sx, h3, dx):
    X = (np.arange(sx) - sx * (np.arange(sx) >= sx // 2)) * dx
    xx = [X[None, None, :], X[None, :, None], (np.arange(h3) * dx)[:, None, None]]
    return np.array([[(xx[j] * F[i]).sum() * dx ** 3 for j in range(3)] for i in range(3)])


@pytest.mark.parametrize("mut", [0, 24])
def test_mogisource_moment_and_mirror_image(mut, mutation):
    """A deep point dilatation source: the first moment of its body force in
    the lower half-space is (lambda + 2 mu / 3) o I, and its image fills the
    upper half of the grid (x3 < 0) as the mirror of the source."""
    sx, sx3, dx, lam, o, zs = 48, 64, 0.5, 2.0e4, 2.5e-3, 7.0
    f = [np.zeros((sx3, sx, sx + 2)) for _ in range(3)]
    mutation(mut)
    O.mogisource(lam, MU, o, 0.5, -1.0, zs, dx, dx, dx, *f)
    mutation(0)
    h3 = sx3 // 2
    F = np.array([c[:h3, :, :sx] for c in f])
    Mf = _first_moment(F, sx, h3, dx)
    # the source sits at (0.5, -1.0): shift the moment to its centre
    assert np.abs(F.sum(axis=(1, 2, 3))).max() * dx ** 3 < 1e-9
    kappa = lam + 2.0 * MU / 3.0
    np.testing.assert_allclose(np.diag(Mf), [kappa * o] * 3, rtol=1e-6)
    assert np.abs(Mf - np.diag(np.diag(Mf))).max() < 1e-9 * kappa * o
    mirror = [np.concatenate([c[:1], c[:0:-1]]) for c in f]  # x3 -> -x3
    diff = max(np.abs(a[h3 + 1:] - b[h3 + 1:]).max() for a, b in zip(f, mirror))
    if mut:
        assert diff > 1e-3 * np.abs(F).max()
    else:
        assert diff == 0


def test_momentdensitymogi_integral():
    """The isotropic moment density of a point source integrates to
    kappa o on the diagonal (normalised Gaussians of width dx)."""
    sx, h3, dx, lam, o = 48, 32, 0.5, 2.0e4, 2.5e-3
    sig = np.zeros((h3, sx, sx, 6))
    O.momentdensitymogi(lam, MU, o, 0.5, -1.0, 7.0, dx, dx, dx, sig)
    tot = sig.sum(axis=(0, 1, 2)) * dx ** 3
    kappa = lam + 2.0 / 3.0 * MU
    np.testing.assert_allclose(tot[[0, 3, 5]], kappa * o, rtol=1e-7)
    assert np.abs(tot[[1, 2, 4]]).max() == 0


@pytest.mark.parametrize("case", ["dipping", "vertical"])
def test_tensilesource_first_moment(case, mutation):
    """A deep opening crack: the first moment of its body force in the lower
    half-space is the moment tensor s L W (lambda I + 2 mu n n) (to the
    quadrature of the tapers)."""
    sx, sx3, dx, lam, s = 64, 128, 0.5, 2.0e4, 1.5e-3
    dip = np.radians(35) if case == "dipping" else 0.0
    # a vertical crack's x3 taper is sampled along its width: a longer, softer
    # taper keeps the quadrature error of its derivative small
    L, W, strike, z, beta = (4.0, 6.0, np.radians(25), 14.0, 0.3) if case == "dipping" else \
        (8.0, 6.0, np.radians(25), 16.0, 0.45)
    f = [np.zeros((sx3, sx, sx + 2)) for _ in range(3)]
    O.tensilesource(lam, MU, s, 0.3, -0.7, z, L, W, strike, dip, beta, dx, dx, dx, *f)
    h3 = sx3 // 2
    F = np.array([c[:h3, :, :sx] for c in f])
    Mf = _first_moment(F, sx, h3, dx)
    cs, ss, cd, sd = np.cos(strike), np.sin(strike), np.cos(dip), np.sin(dip)
    n = np.array([cd * cs, -cd * ss, -sd])
    M0 = s * L * W * (lam * np.eye(3) + 2 * MU * np.outer(n, n))
    assert np.abs(Mf - M0).max() / np.abs(M0).max() < 2e-3
    assert np.abs(F.sum(axis=(1, 2, 3)) * dx ** 3).max() < 1e-3 * np.abs(M0).max() / L


@pytest.mark.parametrize("mut", [0, 22])
def test_tensilesource_vertical_mirror_image(mut, mutation):
    """A deep vertical crack's image in the upper half of the grid (x3 < 0)
    is the mirror of its body force."""
    sx, sx3, dx = 48, 64, 0.5
    f = [np.zeros((sx3, sx, sx + 2)) for _ in range(3)]
    mutation(mut)
    O.tensilesource(2.0e4, MU, 1e-3, 0.0, 0.0, 8.0, 4.0, 5.0, np.radians(40), 0.0, 0.25, dx, dx, dx, *f)
    mutation(0)
    h3 = sx3 // 2
    mirror = [np.concatenate([c[:1], c[:0:-1]]) for c in f]
    diff = max(np.abs(a[h3 + 1:] - b[h3 + 1:]).max() for a, b in zip(f, mirror))
    scale = max(np.abs(c[:h3]).max() for c in f)
    if mut:
        assert diff > 1e-3 * scale
    else:
        assert diff <= 1e-12 * scale


@pytest.mark.parametrize("mut", [0, 23])
def test_momentdensitytensile_integral(mut, mutation):
    """Hooke's law of s (n a)(n a) for a deep crack, a = the aperture:
    integrates to s (2 mu n n + lambda I) times the integral of a^2, i.e.
    1/(2 dx sqrt(pi)) for the Gaussian and W (1 - 1.25 beta)/(1 - beta),
    L (1 - 1.25 beta)/(1 - beta) for the squared tapers."""
    sx, h3, dx, lam, s, beta = 64, 64, 0.25, 2.0e4, 1.5e-3, 0.3
    L, W, strike, dip, z = 3.0, 4.0, np.radians(25), np.radians(35), 8.0
    sig = np.zeros((h3, sx, sx, 6))
    mutation(mut)
    O.momentdensitytensile(lam, MU, s, 0.3, -0.7, z, L, W, strike, dip, beta, dx, dx, dx, sig)
    mutation(0)
    tot = sig.sum(axis=(0, 1, 2)) * dx ** 3
    cs, ss, cd, sd = np.cos(strike), np.sin(strike), np.cos(dip), np.sin(dip)
    n = np.array([cd * cs, -cd * ss, -sd])
    a2 = 1.0 / (2 * dx * np.sqrt(np.pi)) * W * (1 - 1.25 * beta) / (1 - beta) * L * (1 - 1.25 * beta) / (1 - beta)
    M = s * a2 * (2 * MU * np.outer(n, n) + lam * np.eye(3))
    ref = M[[0, 0, 0, 1, 1, 2], [0, 1, 2, 1, 2, 2]]
    err = np.abs(tot - ref).max() / np.abs(ref).max()
    if mut:
        assert err > 1e-2
    else:
        assert err < 2e-3


def test_rk2_periodic_load_rate_telescopes():
    """Time-dependent surface loads in the RK2 loop (relax.f90:805-806): with
    no relaxation, the corrector's u += Dt G[(load(t+Dt) - load(t))/Dt]
    telescopes, so after any step sequence u is the elastic response to
    load(t_n) - load(0)."""
    sx, sx3, dx = 32, 32, 1.0
    h3 = sx3 // 2
    load = _load(period=3.0, phase=0.4, x=1.0, y=-2.0, width=6.0, length=8.0, slip=0.5)
    m = dict(sx1=sx, sx2=sx, sx3=sx3, dx1=dx, dx2=dx, dx3=dx, lam=LAM, mu=MU, gam=8.33e-4, beta=0.2,
             prestress=np.zeros(h3, O.TENSOR_LAYER), interval=100.0, odt=1.0, tscale=1.0, skip=1,
             event_times=[0.0], loads0=load)
    st = dict(u1=np.zeros((h3, sx, sx + 2)), u2=np.zeros((h3, sx, sx + 2)), u3=np.zeros((h3, sx, sx + 2)),
              tau=np.zeros((h3, sx, sx, 6)), sig=np.zeros((h3, sx, sx, 6)), gamma=np.zeros((h3, sx, sx + 2)),
              t=0.0, oi=1, i=0)
    dts = [0.3, 0.7, 0.45, 1.1]
    O.rk2_run(m, st, len(dts), dt_replay=dts)
    tn = sum(dts)
    c = [np.zeros((sx3, sx, sx + 2)) for _ in range(3)]
    t = [np.zeros((sx, sx + 2)) for _ in range(3)]
    a, b = np.zeros((sx, sx + 2)), np.zeros((sx, sx + 2))
    O.traction(load, dx, dx, tn, a)
    O.traction(load, dx, dx, 0.0, b)
    t[2][:] = a - b
    O.greenfunctioncowling(c[0], c[1], c[2], t[0], t[1], t[2], dx, dx, dx, LAM, MU, 8.33e-4)
    assert np.abs(c[2][:h3, :, :sx]).max() > 0
    for k in range(3):
        np.testing.assert_allclose(st["u%d" % (k + 1)][..., :sx], c[k][:h3, :, :sx], rtol=0,
                                   atol=1e-9 * np.abs(c[2][:h3]).max())


# ------------------------------------------- depth stencils on quadratics --
def _depth_derivative(b, c, z, dx3):
    """What the reference's depth differences give for f = a + b z + c z^2
    (elastic3d.f90:977-1002, :1259-1270, fir1 in the interior): centred,
    exact, inside; forward at the top layer (b + c dx3); backward at the
    bottom layer (b + 2 c z - c dx3)."""
    d = b + 2 * c * z
    d[0] = b + c * dx3
    d[-1] = b + 2 * c * z[-1] - c * dx3
    return d


@pytest.mark.parametrize("mut", [0, 8, 25])
def test_stressupdate_quadratic_in_depth(mut, mutation):
    """Quadratic depth profiles tell a centred difference from a one-sided
    one: the interior must be exact and the two boundary layers off by
    exactly +-c dx3."""
    sx1, sx2, h3 = 8, 16, 24
    dx1, dx2, dx3 = 0.5, 0.75, 1.25
    z = np.arange(h3) * dx3
    coef = [(0.3, -0.2, 0.05), (-1.0, 0.4, -0.03), (0.5, 0.7, 0.02)]  # (a, b, c) of u1, u2, u3
    u = [np.broadcast_to((a + b * z + c * z * z)[:, None, None], (h3, sx2, sx1)) for a, b, c in coef]
    D = [_depth_derivative(b, c, z, dx3) for _, b, c in coef]
    e13, e23, e33 = 0.5 * D[0], 0.5 * D[1], D[2]
    ref = np.zeros((h3, 6))
    ref[:, 0] = LAM * e33
    ref[:, 2] = 2 * MU * e13
    ref[:, 3] = LAM * e33
    ref[:, 4] = 2 * MU * e23
    ref[:, 5] = (2 * MU + LAM) * e33
    sig = np.zeros((h3, sx2, sx1, 6))
    mutation(mut)
    O.stressupdate(*(padded(np.ascontiguousarray(a)) for a in u), LAM, MU, dx1, dx2, dx3, sig)
    err = np.abs(sig - ref[:, None, None, :]).max() / np.abs(ref).max()
    if mut == 0:
        assert err < 1e-12
    else:
        assert err > 1e-3


@pytest.mark.parametrize("mut", [0, 26])
def test_equivalentbodyforce_quadratic_in_depth(mut, mutation):
    sx1, sx2, h3 = 16, 8, 20
    dx1, dx2, dx3 = 0.5, 0.75, 1.25
    z = np.arange(h3) * dx3
    s = np.zeros((h3, sx2, sx1, 6))
    coef = {2: (0.1, 0.5, -0.04), 4: (-0.3, 0.2, 0.06), 5: (1.0, -0.8, 0.03)}  # s13, s23, s33
    for comp, (a, b, c) in coef.items():
        s[..., comp] = (a + b * z + c * z * z)[:, None, None]
    f = [_depth_derivative(b, c, z, dx3) for _, b, c in (coef[2], coef[4], coef[5])]
    cc = [np.zeros((h3, sx2, sx1 + 2)) for _ in range(3)]
    t = [np.zeros((sx2, sx1 + 2)) for _ in range(3)]
    mutation(mut)
    O.equivalentbodyforce(s, dx1, dx2, dx3, cc[0], cc[1], cc[2], t[0], t[1], t[2])
    err = max(np.abs(cc[k][..., :sx1] + f[k][:, None, None]).max() for k in range(3))
    if mut == 0:
        assert err < 1e-12
        for k, comp in enumerate((2, 4, 5)):
            np.testing.assert_allclose(t[k][:, :sx1], coef[comp][0], rtol=1e-14)
    else:
        assert err > 1e-3


# ------------------------------------------------ fault creep closed forms --
def _plane_frame(strike, dip):
    """Unit vectors of a fault plane in the internal convention, from the
    geometry alone: the strike direction is horizontal, the normal is
    perpendicular to it and tilted by the dip from the horizontal, and the
    down-dip direction completes the frame with a positive x3 (depth)
    component for |dip| < 90 degrees."""
    s = np.array([np.sin(strike), np.cos(strike), 0.0])
    h = np.array([np.cos(strike), -np.sin(strike), 0.0])  # horizontal, perpendicular to the strike
    n = np.cos(dip) * h + np.sin(dip) * np.array([0.0, 0.0, -1.0])
    d = np.cross(n, s)
    if d[2] < 0:
        d = -d
    return s, n, d


def _uniform_traction_stress(n, e, tau0, p):
    """sigma = tau0 (n e^T + e n^T) + p I: on the plane of normal n (|n| = 1,
    e in the plane, |e| = 1) the traction is tau0 e + p n."""
    S = tau0 * (np.outer(n, e) + np.outer(e, n)) + p * np.eye(3)
    return np.array([S[0, 0], S[0, 1], S[0, 2], S[1, 1], S[1, 2], S[2, 2]])


@pytest.mark.parametrize("mut", [0, 27])
def test_friction_relaxation_time_unequal_spacing(mut, mutation):
    """tm = tau / mu / gammadot * min(L, W) / sqrt(dx1 dx3) (friction3d.f90:
    223) with dx1, dx2, dx3 all different, for a uniform shear traction
    along the rake: every point near the plane gives the same candidate."""
    h3, sx = 32, 48
    dx1, dx2, dx3 = 0.5, 0.8, 0.3
    st = np.zeros(h3, O.LAYER)
    st["gammadot0"], st["stressexponent"], st["friction"], st["cohesion"] = 2e-3, 1.5, 0.4, 0.05
    strike, dip, rake = 0.3, 0.2, 0.0
    s, n, d = _plane_frame(strike, dip)
    tau0, p = 0.9, -0.25
    sig = np.zeros((h3, sx, sx, 6))
    sig[...] = _uniform_traction_stress(n, s, tau0, p)
    mom = np.zeros_like(sig)
    L, W = 4.0, 6.0
    mutation(mut)
    tm = O.frictioneigenstress(0.0, 0.0, 3.0, L, W, strike, dip, rake, 0.2, sig, MU, st, dx1, dx2, dx3, mom,
                               maxwelltime=1e7)
    tau = tau0 + 0.4 * p - 0.05
    gd = 2e-3 * 2 * np.sinh(tau / 1.5)
    expect = float(np.float32(tau / MU / gd * min(L, W) / np.sqrt(dx1 * dx3)))
    if mut == 0:
        assert np.abs(mom).max() > 0
        assert tm == pytest.approx(expect, rel=1e-6)
    else:
        assert abs(tm / expect - 1) > 0.1


def _monitor_setup():
    h3, sx = 32, 48
    dx1, dx2, dx3 = 1.0, 0.75, 0.5
    st = np.zeros(h3, O.LAYER)
    st["gammadot0"] = 1e-3 * (1.0 + np.arange(h3))        # v0 by depth
    st["stressexponent"] = 1.0 + 0.1 * np.arange(h3)       # (a-b) sigma by depth
    st["friction"] = 0.3
    st["cohesion"] = 0.05
    P = dict(x=0.3, y=-0.4, z=4.0, L=6.0, W=4.0, strike=0.4, dip=0.5)
    return h3, sx, (dx1, dx2, dx3), st, P


@pytest.mark.parametrize("mut", [0, 28, 30])
def test_monitorfriction_patch_geometry(mut, mutation):
    """The patch centres (friction3d.f90:318-340) lie on the fault plane,
    px2 x px3 of them spaced dx2 along the (horizontal) strike and dx3 down
    dip, patch (px2/2+1, px3/2+1) at the plane's centre; lx and lz are the
    in-plane coordinates of each centre."""
    h3, sx, (dx1, dx2, dx3), st, P = _monitor_setup()
    s, n, d = _plane_frame(P["strike"], P["dip"])
    px2, px3 = 6, 4
    patch = np.zeros((px3, px2), O.PATCH)
    sig = np.zeros((h3, sx, sx, 6))
    mutation(mut)
    O.monitorfriction(P["x"], P["y"], P["z"], P["L"], P["W"], P["strike"], P["dip"], 0.0, 0.2, 2 * h3,
                      dx1, dx2, dx3, sig, st, patch)
    C = np.array([P["x"], P["y"], P["z"]])
    X = np.stack([patch["x1"], patch["x2"], patch["x3"]], -1)
    j2 = np.arange(1, px2 + 1) - px2 / 2 - 1
    j3 = np.arange(1, px3 + 1) - px3 / 2 - 1
    expect = C + j2[None, :, None] * dx2 * s + j3[:, None, None] * dx3 * d
    err = np.abs(X - expect).max()
    if mut == 0:
        assert err < 1e-5  # the in-plane offsets are rounded to single precision
        assert np.abs((X - C) @ n).max() < 1e-5
        np.testing.assert_allclose(patch["lx"] - patch["lx"][:, :1],
                                   np.broadcast_to((j2 - j2[0]) * dx2, (px3, px2)), atol=1e-5)
        np.testing.assert_allclose(patch["lz"] - patch["lz"][:1, :],
                                   np.broadcast_to(((j3 - j3[0]) * dx3)[:, None], (px3, px2)), atol=1e-5)
    else:
        assert err > 1e-2


@pytest.mark.parametrize("mut", [0, 29])
def test_monitorfriction_velocity_law(mut, mutation):
    """v = 2 v0 sinh(max(0, tau_s + mu_f tau_n - c) / (a-b)sigma) with the
    friction parameters of the depth index of each patch centre (nearest
    grid point, ties down), split into the strike and (up-)dip directions;
    the rake test skips patches whose shear traction opposes the rake unless
    |rake| >= 3 pi (friction3d.f90:341-407)."""
    h3, sx, (dx1, dx2, dx3), st, P = _monitor_setup()
    s, n, d = _plane_frame(P["strike"], P["dip"])
    px2, px3 = 6, 4
    tau0, p = 0.8, -0.3
    # rake 0 points along the strike, rake +90 degrees down dip (the vector r
    # of friction3d.f90:303-305); the dip-direction rate (dv, :291) is positive up dip
    cases = [("strike", s, 0.0, 1.0, 0.0), ("updip", -d, -np.pi / 2, 0.0, 1.0), ("downdip", d, np.pi / 2, 0.0, -1.0)]
    for name, e, rake, fss, fds in cases:
        sig = np.zeros((h3, sx, sx, 6))
        sig[...] = _uniform_traction_stress(n, e, tau0, p)
        patch = np.zeros((px3, px2), O.PATCH)
        mutation(mut)
        O.monitorfriction(P["x"], P["y"], P["z"], P["L"], P["W"], P["strike"], P["dip"], rake, 0.2, 2 * h3,
                          dx1, dx2, dx3, sig, st, patch)
        mutation(0)
        i3 = np.ceil(patch["x3"] / dx3 - 0.5).astype(int)  # nearest, ties to the lower index
        v0, tc = st["gammadot0"][i3], st["stressexponent"][i3]
        tau = max(0.0, tau0 + 0.3 * p - 0.05)
        v = 2 * v0 * np.sinh(tau / tc)
        if mut:
            assert np.abs(patch["v"] / v - 1).max() > 1e-2
            return
        np.testing.assert_allclose(patch["v"], v, rtol=1e-6, err_msg=name)
        np.testing.assert_allclose(patch["vss"], fss * v, rtol=1e-6, atol=1e-12, err_msg=name)
        np.testing.assert_allclose(patch["vds"], fds * v, rtol=1e-6, atol=1e-12, err_msg=name)
        np.testing.assert_allclose(patch["taus"], tau0, rtol=1e-6)
    # shear traction against the rake: the patch keeps its previous values ...
    sig = np.zeros((h3, sx, sx, 6))
    sig[...] = _uniform_traction_stress(n, -s, tau0, p)
    patch = np.zeros((px3, px2), O.PATCH)
    patch["v"] = 12345.0
    O.monitorfriction(P["x"], P["y"], P["z"], P["L"], P["W"], P["strike"], P["dip"], 0.0, 0.2, 2 * h3,
                      dx1, dx2, dx3, sig, st, patch)
    assert np.all(patch["v"] == 12345.0)
    # ... unless |rake| >= 3 pi, where the test is off
    O.monitorfriction(P["x"], P["y"], P["z"], P["L"], P["W"], P["strike"], P["dip"], 3 * np.pi + 0.1, 0.2,
                      2 * h3, dx1, dx2, dx3, sig, st, patch)
    assert np.all(patch["v"] > 0) and np.all(patch["v"] < 1.0)


# ------------------------------------------------ inter-seismic loading --
@pytest.mark.parametrize("mut", [0, 31])
def test_rk2_interseismic_loading_closed_form(mut, mutation):
    """Inter-seismic loading with nothing relaxing (relax.f90:669-674,
    :795-799): tm = interval/20, so every step is Dt = interval/200; each
    corrector's moment is the loading moment density m, so u grows linearly,
    u(t) - u(0) = t G[-div m, m.e3] (the elastic response to m's body force
    and surface traction), and tau gains m unscaled plus Dt m per step (the
    moment-density loops of dislocations run in its eigenstress form too,
    relax.f90:1199-1206): tau_n - tau_0 = (n + t_n) m; the power of each step
    is |m|."""
    from relax_amd import structure

    sx, sx3, dx = 32, 32, 1.0
    h3 = sx3 // 2
    shear, tens = structure.interseismic([(1e-3, -6.0, 2.0, 0.0, 10.0, 4.0, 10.0, 70.0, 30.0)],
                                         [(5e-4, 4.0, -5.0, 2.0, 5.0, 3.0, 0.0, 90.0)], 0.2, (dx, dx, dx), 2.0)
    pre = np.zeros(h3, O.TENSOR_LAYER)
    interval = 40.0
    model = dict(sx1=sx, sx2=sx, sx3=sx3, dx1=dx, dx2=dx, dx3=dx, lam=LAM, mu=MU, gam=8.33e-4, beta=0.2,
                 linear=None, nonlinear=None, faultcreep=None, prestress=pre, interval=interval, odt=1.0,
                 tscale=1.0, skip=1, event_times=[0.0], inter=shear, inter_tensile=tens)
    rng = np.random.default_rng(3)
    st = dict(u1=np.zeros((h3, sx, sx + 2)), u2=np.zeros((h3, sx, sx + 2)), u3=np.zeros((h3, sx, sx + 2)),
              tau=rng.standard_normal((h3, sx, sx, 6)) * 0.1, sig=np.zeros((h3, sx, sx, 6)),
              gamma=np.zeros((h3, sx, sx + 2)), t=0.0, oi=1, i=0)
    st0 = {k: v.copy() for k, v in st.items() if isinstance(v, np.ndarray)}
    # the loading moment density m and its elastic response per unit time
    m = np.zeros((h3, sx, sx, 6))
    # (dislocations passes a segment's width as L, the down-dip size, and
    # its length as W, relax.f90:1183-1187)
    for s in shear:
        O.momentdensityshear(MU, s["slip"], s["x"], s["y"], s["z"], s["width"], s["length"], s["strike"], s["dip"],
                             s["rake"], s["beta"], dx, dx, dx, m)
    for s in tens:
        O.momentdensitytensile(LAM, MU, s["opening"], s["x"], s["y"], s["z"], s["width"], s["length"], s["strike"],
                               s["dip"], 0.2, dx, dx, dx, m)
    assert np.abs(m).max() > 0
    c = [np.zeros((sx3, sx, sx + 2)) for _ in range(3)]
    tr = [np.zeros((sx, sx + 2)) for _ in range(3)]
    O.equivalentbodyforce(m, dx, dx, dx, c[0][:h3], c[1][:h3], c[2][:h3], tr[0], tr[1], tr[2])
    O.greenfunctioncowling(c[0], c[1], c[2], tr[0], tr[1], tr[2], dx, dx, dx, LAM, MU, 8.33e-4)
    U = [a[:h3] for a in c]
    mutation(mut)
    info = O.rk2_run(model, st, 4)
    mutation(0)
    np.testing.assert_allclose(info[:, 1], interval / 20.0)
    np.testing.assert_allclose(info[:, 0], interval / 200.0)
    tn = info[-1, 7]
    assert tn == pytest.approx(4 * interval / 200.0)
    for k in range(3):
        du = st["u%d" % (k + 1)] - st0["u%d" % (k + 1)]
        np.testing.assert_allclose(du[..., :sx], tn * U[k][..., :sx], rtol=1e-9, atol=1e-12 * np.abs(U[k]).max())
    np.testing.assert_allclose(info[:, 5], O.tensoramplitude(m, dx, dx, dx), rtol=1e-12)
    err = np.abs(st["tau"] - st0["tau"] - (4 + tn) * m).max() / np.abs(m).max()
    if mut == 0:
        assert err < 1e-12
    else:
        assert err > 0.5


# ------------------------------------------- Okada (1992), okada/dc3d.f --
ALPHA = (LAM + MU) / (LAM + 2 * MU)


def _okada85_surface(x, y, d, dipdeg, L, W, U, lam=LAM, mu=MU):
    """Okada (1985, BSSA 75:1135-1154) surface displacement of a rectangular
    fault (x along strike, z up; the fault spans 0..L along strike and 0..W
    up dip from its reference point at depth d), a formulation independent
    of the 1992 one the oracle restates: at z = 0 the two must agree."""
    dl = np.radians(dipdeg)
    sd, cd = np.sin(dl), np.cos(dl)
    p, q = y * cd + d * sd, y * sd - d * cd
    k = mu / (lam + mu)

    def f(xi, et):
        R = np.sqrt(xi * xi + et * et + q * q)
        yt, dt = et * cd + q * sd, et * sd - q * cd
        X = np.sqrt(xi * xi + q * q)
        th = np.arctan(xi * et / (q * R))
        I5 = k * 2 / cd * np.arctan((et * (X + q * cd) + X * (R + X) * sd) / (xi * (R + X) * cd))
        I4 = k / cd * (np.log(R + dt) - sd * np.log(R + et))
        I3 = k * (yt / (cd * (R + dt)) - np.log(R + et)) + sd / cd * I4
        I2 = k * (-np.log(R + et)) - I3
        I1 = k * (-xi / (cd * (R + dt))) - sd / cd * I5
        ss = -np.array([xi * q / (R * (R + et)) + th + I1 * sd, yt * q / (R * (R + et)) + q * cd / (R + et) + I2 * sd,
                        dt * q / (R * (R + et)) + q * sd / (R + et) + I4 * sd])
        ds = -np.array([q / R - I3 * sd * cd, yt * q / (R * (R + xi)) + cd * th - I1 * sd * cd,
                        dt * q / (R * (R + xi)) + sd * th - I5 * sd * cd])
        tf = np.array([q * q / (R * (R + et)) - I3 * sd * sd,
                       -dt * q / (R * (R + xi)) - sd * (xi * q / (R * (R + et)) - th) - I1 * sd * sd,
                       yt * q / (R * (R + xi)) + cd * (xi * q / (R * (R + et)) - th) - I5 * sd * sd])
        return (U[0] * ss + U[1] * ds + U[2] * tf) / (2 * np.pi)

    return f(x, p) - f(x, p - W) - f(x - L, p) + f(x - L, p - W)


@pytest.mark.parametrize("dip,U", [(40.0, (0.7, -0.4, 0.3)), (70.0, (0.0, 1.0, 0.0)), (25.0, (0.0, 0.0, 1.0)),
                                   (60.0, (1.0, 0.0, 0.0)), (89.0, (0.3, 0.5, 0.0))])
def test_dc3d_surface_is_okada_1985(dip, U):
    for x, y in ((1.3, 2.1), (-3.0, -1.2), (6.0, 0.5), (2.0, -7.0)):
        u, rc = O.dc3d_core(ALPHA, x, y, 0.0, 3.0, dip, 0.0, 4.0, 0.0, 2.5, U)
        assert rc == 0
        np.testing.assert_allclose(u, _okada85_surface(x, y, 3.0, dip, 4.0, 2.5, U), rtol=1e-10, atol=1e-14)


def _fd_grad(fn, P, h):
    G = np.zeros((3, 3))
    for j in range(3):
        e = np.zeros(3)
        e[j] = h
        G[:, j] = (fn(P + e) - fn(P - e)) / (2 * h)
    return G


@pytest.mark.parametrize("mut", [0, 32])
@pytest.mark.parametrize("dip,U", [(40.0, (0.7, -0.4, 0.3)), (90.0, (1.0, 0.0, 0.0)), (15.0, (0.0, 0.0, 1.0)),
                                   (70.0, (0.0, 1.0, 0.0))])
def test_dc3d_solves_the_half_space_problem(dip, U, mut, mutation):
    """Away from the fault the displacement satisfies Navier's equation
    mu lap u + (lambda + mu) grad div u = 0, and the surface z = 0 is free of
    traction (second and first finite differences of the FP64 core)."""
    geo = (3.0, dip, -2.0, 3.0, -2.5, 0.5)
    mutation(mut)

    def fn(P):
        return O.dc3d_core(ALPHA, P[0], P[1], P[2], *geo, U)[0]

    worst = 0.0
    for P in ((1.3, 2.1, -1.7), (-0.7, -1.9, -4.2), (4.5, 0.3, -2.6)):
        P = np.array(P)
        h = 2e-3
        lap = sum((fn(P + h * e) - 2 * fn(P) + fn(P - h * e)) / h ** 2 for e in np.eye(3))
        gd = np.array([(np.trace(_fd_grad(fn, P + h * e, h)) - np.trace(_fd_grad(fn, P - h * e, h))) / (2 * h)
                       for e in np.eye(3)])
        res = MU * lap + (LAM + MU) * gd
        worst = max(worst, np.abs(res).max() / (MU * np.abs(lap).max()))
    for x, y in ((0.4, 1.1), (-2.4, -0.6), (5.0, 3.0)):
        G = _fd_grad(fn, np.array([x, y, 0.0]), 1e-4)
        eps = 0.5 * (G + G.T)
        S = LAM * np.trace(eps) * np.eye(3) + 2 * MU * eps
        worst = max(worst, np.abs(S[:, 2]).max() / np.abs(S).max())
    mutation(0)
    if mut == 0:
        assert worst < 1e-3
    else:
        assert worst > 1e-2


@pytest.mark.parametrize("dip", [40.0, 90.0, 20.0])
def test_dc3d_jump_is_the_dislocation(dip):
    """Across the fault the displacement jumps by the dislocation: the hanging
    wall moves U1 along the strike, U2 up the dip and U3 along the normal
    into the hanging wall, relative to the foot wall."""
    c = 3.0
    sd, cd = np.sin(np.radians(dip)), np.cos(np.radians(dip))
    P = np.array([1.0, -1.0 * cd, -c - 1.0 * sd])  # strike 1, 1 down dip from the reference point
    to_footwall = np.array([0.0, sd, -cd])
    e = 1e-5
    basis = (np.array([1.0, 0.0, 0.0]), np.array([0.0, cd, sd]), np.array([0.0, -sd, cd]))
    for k in range(3):
        U = np.zeros(3)
        U[k] = 1.0
        uh, _ = O.dc3d_core(ALPHA, *(P - e * to_footwall), c, dip, -2.0, 3.0, -2.5, 0.5, U)
        uf, _ = O.dc3d_core(ALPHA, *(P + e * to_footwall), c, dip, -2.0, 3.0, -2.5, 0.5, U)
        np.testing.assert_allclose(uh - uf, basis[k], atol=1e-4)


@pytest.mark.parametrize("mut", [0, 33])
def test_dc3d_singular_points(mut, mutation):
    """dc3d.f's special points: on a fault edge the result is flagged and
    zero; on the negative extension of an edge (R + xi = 0) it is the limit
    of the neighbouring values; |q| below 1e-6 counts as on the plane."""
    geo = (3.0, 90.0, 0.0, 4.0, -2.5, 0.0)  # vertical, top edge at depth 3
    U = (1.0, 0.0, 0.0)
    u, rc = O.dc3d_core(ALPHA, 2.0, 0.0, -3.0, *geo, U)  # on the top edge
    assert rc == 1 and np.all(u == 0)
    u, rc = O.dc3d_core(ALPHA, 0.0, 0.0, -4.0, *geo, U)  # on the start edge
    assert rc == 1 and np.all(u == 0)
    mutation(mut)
    on, rc = O.dc3d_core(ALPHA, -2.0, 0.0, -3.0, *geo, U)  # beyond the start, on the top edge's line
    near = [O.dc3d_core(ALPHA, -2.0, dy, -3.0 + dz, *geo, U)[0] for dy, dz in ((1e-3, 0), (-1e-3, 0), (0, 1e-3),
                                                                                (0, -1e-3))]
    mutation(0)
    avg = np.mean(near, axis=0)
    if mut == 0:
        assert rc == 0 and np.all(np.isfinite(on))
        np.testing.assert_allclose(on, avg, atol=2e-3 * np.abs(avg).max())
    else:
        assert not (np.all(np.isfinite(on)) and np.allclose(on, avg, atol=2e-3 * np.abs(avg).max()))
    # |q| < 1e-6 is on the plane: the same as q = 0 (inside the fault, theta = 0)
    a, _ = O.dc3d_core(ALPHA, 2.0, 5e-7, -4.0, *geo, U)
    b, _ = O.dc3d_core(ALPHA, 2.0, 0.0, -4.0, *geo, U)
    np.testing.assert_array_equal(a, b)


def test_okada_point_frame():
    """green_space.f90's frame change: a fault given in RELAX's frame
    (north, east, down; upper corner at the start of the strike) at strike 0
    is DC3D's fault with x north and y west; turning the fault and the
    receiver by 90 degrees about the vertical turns the displacement."""
    args = dict(lam=LAM, mu=MU, slip=1.0, opening=0.2, x=-5.0, y=1.0, z=2.0, width=4.0, length=10.0, dip=50.0,
                rake=30.0)

    def pt(strike, xr, yr, zr, **kw):
        a = dict(args, **kw)
        return O.okada_point(a["lam"], a["mu"], a["slip"], a["opening"], a["x"], a["y"], a["z"], a["width"],
                             a["length"], strike, a["dip"], a["rake"], xr, yr, zr)

    d = pt(0.0, 1.5, -2.0, 1.0)
    r = np.radians(30.0)
    U = (np.float32(np.cos(r)), np.float32(np.sin(r)), np.float32(0.2))
    u, rc = O.dc3d(ALPHA, 1.5 + 5.0, -(-2.0 - 1.0), -1.0, 2.0, 50.0, 0.0, 10.0, -4.0, 0.0, *U)
    np.testing.assert_allclose(d, [u[0], -u[1], -u[2]], rtol=2e-6, atol=1e-9)
    # strike 90: the fault runs east from (x, y); rotate the receiver with it
    d90 = pt(90.0, -5.0 + 2.0, 1.0 + 6.5, 1.0)
    d0 = pt(0.0, -5.0 + 6.5, 1.0 - 2.0, 1.0)
    np.testing.assert_allclose([d90[1], -d90[0], d90[2]], d0, rtol=1e-5, atol=1e-7)


def test_dislocations_disp_takes_the_small_sources_only():
    """relax.f90:1300-1349: only the sources at or below nyquist*min(dx) add
    Okada's displacement, each once, in order."""
    sx, h3, dx = 16, 8, 1.0
    from relax_amd import structure  # noqa: F401  (deck conventions are the test's input format)
    small = np.zeros(1, O.SOURCE)
    small[0] = (1e-3, -1.0, 0.5, 1.0, 1.0, 1.5, 30.0, 60.0, 20.0, 0, 0, 0.2, 0)
    big = np.zeros(1, O.SOURCE)
    big[0] = (1e-3, -3.0, 0.5, 1.0, 3.0, 6.0, 30.0, 60.0, 20.0, 0, 0, 0.2, 0)
    tens = np.zeros(1, O.SOURCE)
    tens[0] = (0, 2.0, -2.0, 2.0, 1.0, 1.0, 0.0, 90.0, 0.0, 0, 0, 0.2, 5e-4)
    u = [np.zeros((h3, sx, sx + 2)) for _ in range(3)]
    O.dislocations_disp(np.concatenate([small, big]), tens, LAM, MU, 2.0, dx, dx, dx, *u)
    ref = [np.zeros((h3, sx, sx + 2)) for _ in range(3)]
    s = small[0]
    O.grnfct_okada(LAM, MU, s["slip"], 0.0, s["x"], s["y"], s["z"], s["width"], s["length"], s["strike"], s["dip"],
                   s["rake"], dx, dx, dx, *ref)
    t = tens[0]
    O.grnfct_okada(LAM, MU, 0.0, t["opening"], t["x"], t["y"], t["z"], t["width"], t["length"], t["strike"],
                   t["dip"], 0.0, dx, dx, dx, *ref)
    for a, b in zip(u, ref):
        assert np.abs(b).max() > 0
        np.testing.assert_array_equal(a, b)


# ------------------------------------------------ eigenstrain volumes -----
# cuboidSource / momentDensityCuboid (elastic3d.f90:1680-1884, :2927-3029)
# and tetrahedronSource / momentDensityTetrahedron (:1916-2039, :2788-2873).
EIG_SX, EIG_H3, EIG_DX = 128, 64, 0.25


def _spectral_grad(a, dx=EIG_DX):
    """Spectral gradient (d/dx1, d/dx2, d/dx3) of a (h3, sx2, sx1) field that
    vanishes towards every edge."""
    h3, s2, s1 = a.shape
    F = np.fft.fftn(a)
    ks = (2j * np.pi * np.fft.fftfreq(s1, dx)[None, None, :], 2j * np.pi * np.fft.fftfreq(s2, dx)[None, :, None],
          2j * np.pi * np.fft.fftfreq(h3, dx)[:, None, None])
    return [np.real(np.fft.ifftn(k * F)) for k in ks]


_T6 = {(0, 0): 0, (0, 1): 1, (0, 2): 2, (1, 1): 3, (1, 2): 4, (2, 2): 5}


def _minus_div(m):
    """-div of a symmetric tensor field (h3, sx2, sx1, 6), spectrally."""
    out = np.zeros(m.shape[:3] + (3,))
    for i in range(3):
        for j in range(3):
            out[..., i] -= _spectral_grad(m[..., _T6[(min(i, j), max(i, j))]])[j]
    return out


def _hooke(E):
    """lambda tr(E) I + 2 mu E of a 3x3 strain."""
    return LAM * np.trace(E) * np.eye(3) + 2 * MU * E


def _t6(M):
    return np.array([M[0, 0], M[0, 1], M[0, 2], M[1, 1], M[1, 2], M[2, 2]])


def _cuboid(e, strike, dip, z=5.0, beta=0.3, want=("f", "m")):
    """A 4 x 3 x 2 km cuboid (deck length 4, width 3, thickness 2) with its
    upper corner at (-2, -1, z): its body force (top half) and moment
    density, on a 128 x 128 x 64 grid of 0.25 km."""
    from relax_amd import structure
    _, x, y, zc, st, dp, _ = structure.wangconvention(0.0, -2.0, -1.0, z, 4.0, 3.0, strike, dip, 0.0)
    out = {"geom": (x, y, zc, st, dp)}
    if "f" in want:
        v = [np.zeros((2 * EIG_H3, EIG_SX, EIG_SX + 2)) for _ in range(3)]
        # the driver passes (width, length) = (3, 4) as (L, W)
        O.cuboidsource(LAM, MU, e, x, y, zc, 3.0, 4.0, 2.0, st, dp, beta, EIG_DX, EIG_DX, EIG_DX, *v)
        out["f"] = np.stack([a[:EIG_H3, :, :EIG_SX] for a in v], -1)
    if "m" in want:
        m = np.zeros((EIG_H3, EIG_SX, EIG_SX, 6))
        O.momentdensitycuboid(MU, LAM, e, x, y, zc, 3.0, 4.0, 2.0, st, dp, beta, EIG_DX, EIG_DX, EIG_DX, m)
        out["m"] = m
    return out


@pytest.mark.parametrize("mut", [0, 36])
def test_cuboid_moment_density_is_hooke_of_the_rotated_eigenstrain(mut, mutation):
    """Inside the cuboid the moment density is C:E with E the eigenstrain in
    the (strike, normal, dip) frame, and its integral is the volume times
    that.  A vertical cuboid of strike 0 has the frame (x1, x2, x3), so
    E = e; with strike 90 the strike runs along x2, the normal along -x1:
    E = [[e22, -e12, -e23], [-e12, e11, e13], [-e23, e13, e33]]."""
    mutation(mut)
    rng = np.random.default_rng(5)
    e = rng.uniform(-1e-3, 1e-3, 6)
    e11, e12, e13, e22, e23, e33 = np.float32(e).astype(np.float64)
    ok = True
    for strike, E in ((0.0, np.array([[e11, e12, e13], [e12, e22, e23], [e13, e23, e33]])),
                      (90.0, np.array([[e22, -e12, -e23], [-e12, e11, e13], [-e23, e13, e33]]))):
        c = _cuboid(e, strike, 90.0, want=("m",))
        x, y, zc = c["geom"][:3]
        i1, i2, i3 = int(round(x / EIG_DX)) % EIG_SX, int(round(y / EIG_DX)) % EIG_SX, int(round(zc / EIG_DX))
        want = _t6(_hooke(E))
        ok &= np.allclose(c["m"][i3, i2, i1], want, rtol=1e-5, atol=1e-6 * np.abs(want).max())
        ok &= np.allclose(c["m"].sum((0, 1, 2)) * EIG_DX ** 3, 24.0 * want, rtol=1e-3, atol=1e-3 * np.abs(want).max())
    assert ok == (mut == 0)


def _deviation(e, geom, m, k):
    """The body force's departure from -div(m) in the reference's dipping
    terms of e13 and e23 (elastic3d.f90:1851-1872): the e13 dip derivatives
    carry an extra cos(dip), and f3 lacks 2 mu e23 sin(dip) d.grad(A), where
    A is the cuboid's aperture (read off the moment density)."""
    x, y, zc, st, dp = geom
    cs, ss, cd, sd = np.cos(st), np.sin(st), np.cos(dp), np.sin(dp)
    s = np.array([-ss, -cs, 0.0])
    n = np.array([cd * cs, -cd * ss, -sd])
    d = np.array([cs * sd, -ss * sd, cd])
    a, b = (s, d) if k == 2 else (n, d)
    A = m[..., 2] / (2 * MU * e[k] * (a[0] * b[2] + a[2] * b[0]))
    g = _spectral_grad(A)
    dipcs = d[0] * g[0] + d[1] * g[1] + d[2] * g[2]
    dev = np.zeros(m.shape[:3] + (3,))
    if k == 2:
        dev[..., 0] = 2 * MU * e[k] * (cd - 1) * ss * dipcs
        dev[..., 1] = 2 * MU * e[k] * (cd - 1) * cs * dipcs
    else:
        dev[..., 2] = -2 * MU * e[k] * sd * dipcs
    return dev


@pytest.mark.parametrize("strike,dip", [(0.0, 90.0), (30.0, 90.0), (30.0, 60.0)])
def test_cuboid_body_force_is_minus_the_divergence_of_its_moment(strike, dip):
    """A buried cuboid's body force is -div of its moment density (spectral
    derivative, relative L2 within 12 % per component on this grid, the
    discretisation of the tapers), for each eigenstrain component -- except
    that for a dipping cuboid the e13 and e23 parts follow the reference's
    formula, which departs from -div(m) by the terms of _deviation (with
    them the agreement is again within 12 %, without them it is not)."""
    for k in range(6):
        e = np.zeros(6)
        e[k] = 1e-3
        c = _cuboid(e, strike, dip)
        f, want = c["f"], _minus_div(c["m"])
        dev = _deviation(e, c["geom"], c["m"], k) if (dip != 90.0 and k in (2, 4)) else 0.0
        for i in range(3):
            nf = np.linalg.norm(f[..., i])
            if nf < 1e-9 * np.abs(f).max() * f[..., i].size:
                continue
            err = np.linalg.norm(f[..., i] - want[..., i] - (dev[..., i] if np.ndim(dev) else 0.0)) / nf
            assert err < 0.12, (k, i, err)
            if np.ndim(dev) and np.abs(dev[..., i]).max() > 0:
                assert np.linalg.norm(f[..., i] - want[..., i]) / nf > 0.12


@pytest.mark.parametrize("mut", [0, 35])
def test_cuboid_image_mirrors_a_subvertical_source(mut, mutation):
    """A subvertical cuboid's body force carries the image above the surface
    of its horizontal-shear part (e11, e12, e22; elastic3d.f90:1814-1831;
    the vertical-shear part, which includes lambda tr(e), has none): with
    e13 = e23 = e33 = tr(e) = 0, f1 and f2 of a cuboid centred at depth z
    equal those of the same cuboid centred at -z."""
    mutation(mut)
    e = np.array([0.5, 0.7, 0.0, -0.5, 0.0, 0.0]) * 1e-3
    v = {}
    for z in (1.0, -1.0):
        w = [np.zeros((2 * EIG_H3, EIG_SX, EIG_SX + 2)) for _ in range(3)]
        O.cuboidsource(LAM, MU, e, 0.5, -0.25, z, 3.0, 4.0, 2.0, np.radians(-120.0), 0.0, 0.3, EIG_DX, EIG_DX,
                       EIG_DX, *w)
        v[z] = [a[:EIG_H3, :, :EIG_SX] for a in w]
    scale = max(np.abs(a).max() for a in v[1.0])
    ok = (np.allclose(v[1.0][0], v[-1.0][0], atol=1e-9 * scale) and
          np.allclose(v[1.0][1], v[-1.0][1], atol=1e-9 * scale))
    assert ok == (mut == 0)


TET = np.array([[-4.5, -3.0, 3.0], [6.0, -1.5, 4.5], [0.0, 7.5, 6.0], [0.75, 0.0, 14.0]])


def _tetra(e, P=TET, want=("f", "m")):
    out = {}
    if "f" in want:
        v = [np.zeros((2 * EIG_H3, EIG_SX, EIG_SX + 2)) for _ in range(3)]
        O.tetrahedronsource(LAM, MU, e, *P, EIG_DX, EIG_DX, EIG_DX, *v)
        out["f"] = np.stack([a[:EIG_H3, :, :EIG_SX] for a in v], -1)
    if "m" in want:
        m = np.zeros((EIG_H3, EIG_SX, EIG_SX, 6))
        O.momentdensitytetrahedron(MU, LAM, e, *P, EIG_DX, EIG_DX, EIG_DX, m)
        out["m"] = m
    return out


@pytest.mark.parametrize("mut", [0, 34])
def test_tetrahedron_moment_density_and_body_force(mut, mutation):
    """Inside a tetrahedron the moment density is C:e, its integral is the
    volume (|det|/6) times C:e (within 2 %: the Gaussian smoothing of the
    faces and edges), and the body force is -div of the moment density
    (spectral derivative, within 2 %): the faces' tractions C:e.n times a
    smoothed delta.  Its total vanishes."""
    mutation(mut)
    e = np.array([0.4, -0.3, 0.2, 0.7, -0.5, -0.1]) * 1e-3
    E = np.float32(e).astype(np.float64)
    want = _t6(_hooke(np.array([[E[0], E[1], E[2]], [E[1], E[3], E[4]], [E[2], E[4], E[5]]])))
    t = _tetra(e)
    m, f = t["m"], t["f"]
    c = TET.mean(0)
    inside = np.allclose(m[int(round(c[2] / EIG_DX)), int(round(c[1] / EIG_DX)), int(round(c[0] / EIG_DX))],
                         want, rtol=1e-5)
    vol = abs(np.linalg.det(TET[1:] - TET[0])) / 6.0
    total = np.allclose(m.sum((0, 1, 2)) * EIG_DX ** 3, vol * want, rtol=0.02, atol=0.02 * vol * np.abs(want).max())
    div = np.linalg.norm(f - _minus_div(m)) / np.linalg.norm(f) < 0.02
    zero = np.abs(f.sum((0, 1, 2))).max() < 1e-6 * np.abs(f).max() * f.shape[0] * f.shape[1] * f.shape[2]
    assert (inside and total and div and zero) == (mut == 0)


def test_tetrahedron_moment_density_skips_its_deep_inside_of_small_ones():
    """momentDensityTetrahedron skips the points that lie within 7 dx1 of
    the inner side of all four faces (elastic3d.f90:2838-2841, as written),
    so a tetrahedron smaller than that gets no moment density inside, while
    its body force is there; a degenerate (flat) one gets neither."""
    small = np.array([[0.0, 0.0, 3.0], [2.0, 0.0, 3.0], [0.0, 2.0, 3.0], [0.0, 0.0, 5.0]])
    e = np.array([1.0, 0, 0, 1.0, 0, 1.0]) * 1e-3
    t = _tetra(e, small)
    assert np.abs(t["f"]).max() > 0
    i3 = int(round(3.5 / EIG_DX))
    assert np.all(t["m"][i3, 1:3, 1:3] == 0)
    assert np.abs(t["m"]).max() > 0  # outside, where some face is farther than 7 dx1
    flat = small.copy()
    flat[3, 2] = 3.0
    t = _tetra(e, flat)
    assert np.abs(t["f"]).max() == 0 and np.abs(t["m"]).max() == 0


def test_eigenstrain_volumes_in_dislocations():
    """dislocations adds the volumes' body forces to v and their moment
    densities to tau, as the routines do one by one; vertex indices are
    1-based."""
    e = np.array([0.2, 0.1, -0.3, 0.4, 0.0, -0.2]) * 1e-3
    cb = np.zeros(1, O.WEAK)
    cb[0]["x"], cb[0]["y"], cb[0]["z"], cb[0]["width"], cb[0]["length"], cb[0]["thickness"] = -3.0, 2.0, 4.0, 3.0, 4.0, 2.0
    cb[0]["strike"], cb[0]["dip"] = np.radians(-110.0), np.radians(20.0)
    cb[0]["e"] = e
    th = np.zeros(1, O.TETRA)
    th[0]["e"] = e[::-1]
    th[0]["i1"], th[0]["i2"], th[0]["i3"], th[0]["i4"] = 4, 2, 3, 1
    vx = TET[::-1].copy()  # vertex k+1 of vx is TET[3-k]
    v = [np.zeros((2 * EIG_H3, EIG_SX, EIG_SX + 2)) for _ in range(3)]
    t = [np.zeros((EIG_SX, EIG_SX + 2)) for _ in range(3)]
    tau = np.zeros((EIG_H3, EIG_SX, EIG_SX, 6))
    O.dislocations(None, None, MU, 2.0, 0.0, EIG_DX, EIG_DX, EIG_DX, v, t, tau, lam=LAM, beta=0.3, cuboids=cb,
                   tetrahedra=th, vertices=vx)
    w = [np.zeros_like(a) for a in v]
    m = np.zeros_like(tau)
    c = cb[0]
    O.cuboidsource(LAM, MU, e, c["x"], c["y"], c["z"], 3.0, 4.0, 2.0, c["strike"], c["dip"], 0.3, EIG_DX, EIG_DX,
                   EIG_DX, *w)
    O.tetrahedronsource(LAM, MU, e[::-1], TET[0], TET[2], TET[1], TET[3], EIG_DX, EIG_DX, EIG_DX, *w)
    O.momentdensitycuboid(MU, LAM, e, c["x"], c["y"], c["z"], 3.0, 4.0, 2.0, c["strike"], c["dip"], 0.3, EIG_DX,
                          EIG_DX, EIG_DX, m)
    O.momentdensitytetrahedron(MU, LAM, e[::-1], TET[0], TET[2], TET[1], TET[3], EIG_DX, EIG_DX, EIG_DX, m)
    for a, b in zip(v, w):
        np.testing.assert_array_equal(a, b)
    np.testing.assert_array_equal(tau, m)
    assert np.abs(tau).max() > 0 and all(np.abs(a).max() == 0 for a in t)
