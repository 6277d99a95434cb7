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
