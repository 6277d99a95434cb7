"""Build the gfx950 HIP library relax_amd/librelax_hip.so with hipcc.

Usage: python -m relax_amd.build   (or relax_amd.build.build() from Python)
"""
import concurrent.futures
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
OUT = os.path.join(HERE, "librelax_hip.so")
OBJ = os.path.join(HERE, "_build")
SOURCES = ["fft.hip", "green.hip", "stencil.hip", "model.hip", "compat.hip", "okada.hip", "eigen.hip"]
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
FLAGS = ["--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC", "-Wall", "-Wno-unused-function",
         "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
         "-I" + os.path.join(HERE, "..", "include")]


def _compile(src):
    obj = os.path.join(OBJ, src.replace(".hip", ".o"))
    cmd = [HIPCC] + FLAGS + ["-c", os.path.join(CSRC, src), "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    return src, obj, r.returncode, r.stdout + r.stderr


def build(verbose=False):
    os.makedirs(OBJ, exist_ok=True)
    with concurrent.futures.ThreadPoolExecutor(max_workers=min(8, len(SOURCES))) as ex:
        results = list(ex.map(_compile, SOURCES))
    errs = [(s, log) for s, _, rc, log in results if rc != 0]
    for s, _, rc, log in results:
        if verbose or rc != 0:
            filtered = "\n".join(l for l in log.splitlines() if "packed-fp32-ops" not in l)
            if filtered.strip():
                sys.stderr.write("== %s ==\n%s\n" % (s, filtered))
    if errs:
        raise RuntimeError("hipcc failed for: " + ", ".join(s for s, _ in errs))
    tmp = OUT + ".tmp"
    cmd = [HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", tmp] + [o for _, o, _, _ in results]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("link failed:\n" + r.stdout + r.stderr)
    os.replace(tmp, OUT)
    return OUT


def build_fortran():
    """Compile MODULE relax_hip and the Fortran test drivers (per-routine
    drop-in, RK2 model loop, batch) with flang (skipped when the image has no
    flang)."""
    import shutil
    flang = shutil.which("flang") or "/opt/rocm/llvm/bin/flang"
    if not os.path.exists(flang):
        return None
    root = os.path.dirname(HERE)
    bdir = os.path.join(root, "tests", "fortran", "_bin")
    os.makedirs(bdir, exist_ok=True)
    mod_obj = os.path.join(bdir, "relax_hip.o")
    r = subprocess.run([flang, "-cpp", "-O2", "-J", bdir, "-c", os.path.join(HERE, "fortran", "relax_hip.f90"),
                        "-o", mod_obj], capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("flang failed:\n" + r.stdout + r.stderr)
    exes = []
    for name in ("drop_in", "model_loop", "batch_loop"):
        exe = os.path.join(bdir, name)
        cmd = [flang, "-cpp", "-O2", "-J", bdir, mod_obj, os.path.join(root, "tests", "fortran", name + ".f90"),
               "-L" + HERE, "-lrelax_hip", "-Wl,-rpath,$ORIGIN/../../../relax_amd", "-o", exe]
        r = subprocess.run(cmd, capture_output=True, text=True)
        if r.returncode != 0:
            raise RuntimeError("flang failed:\n" + r.stdout + r.stderr)
        exes.append(exe)
    return exes


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
