"""CPU checks of the C ABI: the library loads, exports every entry point the
header declares, and the structure layouts agree between C and Python.
No compute call is made (no GPU here)."""
import ctypes
import os
import re
import subprocess

import numpy as np
import pytest

from relax_amd import _lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "relax_hip.h")


def header_functions():
    src = open(HEADER).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"\b(relax_[a-z0-9_]+)\s*\(", src)))


def test_library_exports_every_header_function():
    lib = _lib.load()
    names = header_functions()
    assert len(names) >= 30
    for n in names:
        assert hasattr(lib, n), n


def test_python_binding_covers_the_header():
    assert sorted(_lib.SIGNATURES) == header_functions()


def test_status_strings_without_device():
    assert "gfx950" in _lib.load().relax_hip_version().decode()
    assert _lib.status_string(0) == "ok"
    assert "power" in _lib.status_string(1)
    assert "(a-b)sigma" in _lib.status_string(5)


def test_struct_layouts_match_header(tmp_path):
    prog = tmp_path / "sizes.c"
    prog.write_text(r'''
#include <stdio.h>
#include <stddef.h>
#include "relax_hip.h"
int main(void) {
  printf("%zu %zu %zu %zu %zu %zu %zu\n", sizeof(relax_tensor_t), sizeof(relax_layer_t),
         sizeof(relax_tensor_layer_t), sizeof(relax_slippatch_t), sizeof(relax_plane_t),
         sizeof(relax_model_desc_t), sizeof(relax_step_info_t));
  printf("%zu %zu %zu\n", offsetof(relax_slippatch_t, sig), offsetof(relax_model_desc_t, event_times),
         offsetof(relax_step_info_t, status));
  return 0;
}
''')
    exe = tmp_path / "sizes"
    subprocess.run(["gcc", "-I" + os.path.join(ROOT, "include"), str(prog), "-o", str(exe)], check=True)
    out = subprocess.run([str(exe)], capture_output=True, text=True, check=True).stdout.split("\n")
    sizes = [int(x) for x in out[0].split()]
    offs = [int(x) for x in out[1].split()]
    assert sizes == [24, 48, _lib.TENSOR_LAYER.itemsize, _lib.PATCH.itemsize, _lib.PLANE.itemsize,
                     ctypes.sizeof(_lib.ModelDesc), ctypes.sizeof(_lib.StepInfo)]
    assert offs == [_lib.PATCH.fields["sig"][1], _lib.ModelDesc.event_times.offset, _lib.StepInfo.status.offset]
    # and the oracle's copies of the reference types agree with the ABI
    import oracle
    assert oracle.PATCH.itemsize == _lib.PATCH.itemsize and oracle.LAYER.itemsize == 48


def test_product_has_no_cpu_fallback():
    """The package never imports the oracle, and a missing library raises."""
    pkg = os.path.join(ROOT, "relax_amd")
    for dirpath, _, files in os.walk(pkg):
        for f in files:
            if f.endswith(".py"):
                txt = open(os.path.join(dirpath, f)).read()
                assert "import oracle" not in txt and "from oracle" not in txt, f
    saved = _lib.LIBPATH
    try:
        _lib.LIBPATH = saved + ".absent"
        old, _lib._LIB = _lib._LIB, None
        with pytest.raises(RuntimeError):
            _lib.load()
    finally:
        _lib.LIBPATH = saved
        _lib._LIB = old
