"""Build the gfx950 shared library (and the Fortran driver) in-tree.

``python -m exciting_amd.build`` compiles every ``exciting_amd/csrc/*.hip`` with hipcc
for gfx950 and links ``exciting_amd/_lib/libexcgpu.so``; then it compiles the Fortran
iso_c_binding module and driver with amdflang into ``exciting_amd/_lib/``.
"""
from __future__ import annotations

import concurrent.futures as cf
import os
import shutil
import subprocess
import sys
from pathlib import Path

PKG = Path(__file__).resolve().parent
ROOT = PKG.parent
CSRC = PKG / "csrc"
FSRC = PKG / "fortran"
LIBDIR = PKG / "_lib"
OBJDIR = LIBDIR / "obj"
LIB = LIBDIR / "libexcgpu.so"

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
FLANG = os.environ.get("FC", "/opt/rocm/bin/amdflang")
ARCH = "gfx950"
HIPFLAGS = [
    "-O3", "-std=c++17", "-fPIC", f"--offload-arch={ARCH}",
    # packed FP32 instructions are kept out of every kernel
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-I", str(ROOT / "include"), "-Wno-unused-result",
]


def _run(cmd):
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("command failed: " + " ".join(cmd) + "\n" + r.stdout + r.stderr)
    return r


def build_hip(verbose: bool = False) -> Path:
    OBJDIR.mkdir(parents=True, exist_ok=True)
    srcs = sorted(CSRC.glob("*.hip"))
    objs = [OBJDIR / (s.stem + ".o") for s in srcs]

    def one(pair):
        s, o = pair
        _run([HIPCC, *HIPFLAGS, "-c", str(s), "-o", str(o)])
        return o

    with cf.ThreadPoolExecutor(max_workers=min(8, len(srcs))) as ex:
        list(ex.map(one, zip(srcs, objs)))
    tmp = LIB.with_suffix(".so.tmp")
    _run([HIPCC, "-shared", "-fPIC", f"--offload-arch={ARCH}", *map(str, objs), "-L/opt/rocm/lib", "-lrocsolver", "-lrocblas", "-lhipfft", "-lpthread", "-o", str(tmp)])
    os.replace(tmp, LIB)
    if verbose:
        print("built", LIB)
    return LIB


def build_fortran(verbose: bool = False):
    """Compile the iso_c_binding module, the reference-signature driver and its test program."""
    if not FSRC.exists() or not any(FSRC.glob("*.f90")):
        return None
    flang = shutil.which(FLANG) or FLANG
    if not Path(flang).exists():
        raise RuntimeError("amdflang not found")
    mod = LIBDIR / "mod"
    mod.mkdir(parents=True, exist_ok=True)
    order = ["exc_lapw_c.f90", "exc_evsystem.f90", "exc_setup_driver.f90"]
    objs = []
    for name in order:
        src = FSRC / name
        if not src.exists():
            continue
        obj = OBJDIR / (src.stem + ".o")
        _run([flang, "-O2", "-fPIC", "-c", str(src), "-module-dir", str(mod), "-I", str(mod), "-o", str(obj)])
        objs.append(obj)
    prog = FSRC / "exc_driver_main.f90"
    exe = None
    if prog.exists():
        exe = LIBDIR / "exc_driver_main"
        _run([flang, "-O2", str(prog), *map(str, objs), "-I", str(mod), "-o", str(exe),
              f"-L{LIBDIR}", "-lexcgpu", f"-Wl,-rpath,{LIBDIR}", "-Wl,-rpath,$ORIGIN"])
    if verbose:
        print("built fortran driver", exe)
    return exe


def build(verbose: bool = False):
    lib = build_hip(verbose)
    build_fortran(verbose)
    return lib


if __name__ == "__main__":
    build(verbose=True)
    sys.exit(0)
