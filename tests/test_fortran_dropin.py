"""The Fortran drop-in (MODULE relax_hip) called by a Fortran driver with the
reference's routine names, checked against the oracle (GPU only)."""
import os
import subprocess

import numpy as np
import pytest

import oracle as O
from _cases import rel_l2

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(ROOT, "tests", "fortran", "_bin", "drop_in")


@pytest.mark.gpu
def test_fortran_driver_green_and_stress(tmp_path):
    if not os.path.exists(EXE):
        pytest.skip("the Fortran driver is built by __graft_entry__.build() when flang is present")
    out = tmp_path / "drop_in.bin"
    r = subprocess.run([EXE, str(out)], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    sx1 = sx2 = sx3 = 32
    h3 = sx3 // 2
    raw = np.fromfile(out, np.float32)
    nv, nt = (sx1 + 2) * sx2 * sx3, (sx1 + 2) * sx2
    o = 0
    fields = []
    for n, shape in [(nv, (sx3, sx2, sx1 + 2))] * 3 + [(nt, (sx2, sx1 + 2))] * 3 + \
                    [(nv, (sx3, sx2, sx1 + 2))] * 3 + [(sx1 * sx2 * h3 * 6, (h3, sx2, sx1, 6))]:
        fields.append(raw[o:o + n].reshape(shape).copy())
        o += n
    assert o == raw.size
    c = fields[0:3]
    t = fields[3:6]
    O.greenfunctioncowling(*c, *t, 1.0, 1.0, 1.0, 3e4, 3e4, 8.33e-4)
    for k in range(3):
        assert rel_l2(fields[6 + k][..., :sx1], c[k][..., :sx1]) < 1e-5
    sig = np.zeros((h3, sx2, sx1, 6), np.float32)
    O.stressupdate(*[np.ascontiguousarray(a[:h3]) for a in c], 3e4, 3e4, 1.0, 1.0, 1.0, sig)
    assert rel_l2(fields[9], sig) < 1e-5


@pytest.mark.gpu
def test_fortran_model_loop_matches_python_driver(tmp_path):
    """The device-resident model driven from Fortran (create, structure,
    first event with a patch and a load, 3 RK2 steps, download) gives bitwise the fields and
    scalars of the same model driven through the Python binding."""
    exe = os.path.join(ROOT, "tests", "fortran", "_bin", "model_loop")
    if not os.path.exists(exe):
        pytest.skip("the Fortran driver is built by __graft_entry__.build() when flang is present")
    from relax_amd import LAYER, SOURCE, TENSOR_LAYER
    from relax_amd.model import LINEAR, Model

    out = tmp_path / "model_loop.bin"
    r = subprocess.run([exe, str(out)], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    sx1 = sx2 = sx3 = 32
    h3 = sx3 // 2
    nsteps = 3
    raw = open(out, "rb").read()
    scal = np.frombuffer(raw[:nsteps * 4 * 8], np.float64).reshape(nsteps, 4)
    f = np.frombuffer(raw[nsteps * 4 * 8:], np.float32)
    nv = (sx1 + 2) * sx2 * h3
    u = [f[k * nv:(k + 1) * nv].reshape(h3, sx2, sx1 + 2) for k in range(3)]
    tau = f[3 * nv:].reshape(h3, sx2, sx1, 6)

    with Model(sx1, sx2, sx3, 1.0, 1.0, 1.0, 3e4, 3e4, 8.33e-4, 0.2, interval=1e9, odt=1.0, tscale=1.0,
               skip=1, event_times=()) as m:
        lin = np.zeros(h3, LAYER)
        lin["z"] = np.arange(h3) * 1.0
        lin["gammadot0"] = np.where(np.arange(1, h3 + 1) > 6, 0.2, 0.0)
        lin["stressexponent"] = 1.0
        lin["friction"] = 0.6
        pre = np.zeros(h3, TENSOR_LAYER)
        pre["z"] = np.arange(h3)
        m.set_prestress(pre)
        m.set_structure(LINEAR, lin)
        pat = np.zeros(1, SOURCE)
        pat[0] = (1e-3, 0.0, 0.0, 1.0, 6.0, 10.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.2, 0.0)
        load = np.zeros(1, SOURCE)
        load[0] = (0.5, 3.0, -2.0, 0.0, 3.0, 4.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.2, 0.0)
        m.event(pat, load, nyquist=1.0)  # shear patch and load
        recs = [m.step() for _ in range(nsteps)]
        ref = m.download()
    np.testing.assert_array_equal(scal, [[r["Dt"], r["t"], r["power"], r["cei"]] for r in recs])
    for k in range(3):
        np.testing.assert_array_equal(u[k][..., :sx1], ref["u%d" % (k + 1)][..., :sx1])
    np.testing.assert_array_equal(tau, ref["tau"])
    assert np.abs(u[0][..., :sx1]).max() > 0  # the padding columns are not written


def _python_model(sx3, g0, nsteps):
    """The model batch_loop.f90 sets up, driven alone through Python."""
    from relax_amd import LAYER, SOURCE, TENSOR_LAYER
    from relax_amd.model import LINEAR, Model

    h3 = sx3 // 2
    with Model(32, 32, sx3, 1.0, 1.0, 1.0, 3e4, 3e4, 8.33e-4, 0.2, interval=1e9, odt=1.0, tscale=1.0, skip=1,
               event_times=()) as m:
        lin = np.zeros(h3, LAYER)
        lin["z"] = np.arange(h3) * 1.0
        lin["gammadot0"] = np.where(np.arange(1, h3 + 1) > 6, g0, 0.0)
        lin["stressexponent"] = 1.0
        lin["friction"] = 0.6
        pre = np.zeros(h3, TENSOR_LAYER)
        pre["z"] = np.arange(h3)
        m.set_prestress(pre)
        m.set_structure(LINEAR, lin)
        pat = np.zeros(1, SOURCE)
        pat[0] = (1e-3, 0.0, 0.0, 1.0, 6.0, 10.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.2, 0.0)
        load = np.zeros(1, SOURCE)
        load[0] = (0.5, 3.0, -2.0, 0.0, 3.0, 4.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.2, 0.0)
        m.event(pat, load, nyquist=1.0)
        recs = [m.step() for _ in range(nsteps)]
        return recs, m.download()["u1"]


@pytest.mark.gpu
def test_fortran_batch_matches_single_python_runs(tmp_path):
    """Three models (two grids) stepped as one batch from Fortran
    (relax_hip_batch_*) give bitwise the scalars and u1 of each model run
    alone through the Python binding."""
    exe = os.path.join(ROOT, "tests", "fortran", "_bin", "batch_loop")
    if not os.path.exists(exe):
        pytest.skip("the Fortran driver is built by __graft_entry__.build() when flang is present")
    out = tmp_path / "batch_loop.bin"
    r = subprocess.run([exe, str(out)], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    nsteps, depth, g0 = 3, (32, 32, 64), (0.2, 0.05, 0.1)
    raw = open(out, "rb").read()
    nscal = nsteps * len(depth) * 4
    scal = np.frombuffer(raw[:nscal * 8], np.float64).reshape(nsteps, len(depth), 4)
    f = np.frombuffer(raw[nscal * 8:], np.float32)
    o = 0
    for j, (sx3, g) in enumerate(zip(depth, g0)):
        n = 34 * 32 * (sx3 // 2)
        u1 = f[o:o + n].reshape(sx3 // 2, 32, 34)
        o += n
        recs, ref = _python_model(sx3, g, nsteps)
        np.testing.assert_array_equal(scal[:, j], [[q["Dt"], q["t"], q["power"], q["cei"]] for q in recs])
        np.testing.assert_array_equal(u1[..., :32], ref[..., :32])
        assert np.abs(u1[..., :32]).max() > 0
    assert o == f.size
