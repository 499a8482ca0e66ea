"""The C-ABI library: it loads, exports every function include/*.h declares, the Fortran
interface module binds exactly those names, and argument errors are reported without a
GPU (no compute call is made here)."""
import ctypes as C
import os
import re
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "exc_lapw.h")
FMOD = os.path.join(ROOT, "exciting_amd", "fortran", "exc_lapw_c.f90")


def declared():
    txt = open(HEADER).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\bint\s+(exc_\w+)\s*\(", txt)))


def _lib():
    from exciting_amd import lapw
    try:
        return lapw.load_library()
    except OSError as e:
        pytest.fail(f"library not built: {e}")


def test_header_declares_entry_points():
    names = declared()
    for must in ["exc_match", "exc_match_dev", "exc_hamiltonandoverlapsetup", "exc_hamiltonandoverlapsetup_dev",
                 "exc_ctx_create", "exc_ctx_destroy"]:
        assert must in names


def test_library_exports_every_declared_symbol():
    lib = _lib()
    path = lib._name
    out = subprocess.run(["nm", "-D", "--defined-only", path], capture_output=True, text=True, check=True).stdout
    exported = set(re.findall(r"\sT\s(exc_\w+)", out))
    missing = [n for n in declared() if n not in exported]
    assert not missing, missing
    for n in declared():
        getattr(lib, n)


def test_fortran_interface_matches_header():
    binds = set(re.findall(r'bind\(c,\s*name="(exc_\w+)"\)', open(FMOD).read()))
    assert binds, "no bind(c) names found"
    assert binds <= set(declared()), binds - set(declared())


def test_argument_errors_without_gpu():
    from exciting_amd import lapw
    lib = _lib()
    h = C.c_void_p()
    assert lib.exc_ctx_create(None, C.byref(h)) == lapw.EXC_ERR_ARG
    d = lapw.ExcDims()
    assert lib.exc_ctx_create(C.byref(d), C.byref(h)) == lapw.EXC_ERR_ARG          # all-zero dims
    d.natmtot, d.nspecies, d.lmaxapw, d.lmaxmat, d.apwordmax = 2, 1, 8, 8, 3
    d.ngrtot, d.ngvec, d.omega = 10, 5, 100.0
    for i in range(3):
        d.intgv[i], d.intgv[3 + i] = -2, 2
    assert lib.exc_ctx_create(C.byref(d), C.byref(h)) == lapw.EXC_ERR_UNSUPPORTED    # apwordmax != 2
    d.apwordmax, d.lmaxmat = 2, 6
    assert lib.exc_ctx_create(C.byref(d), C.byref(h)) == lapw.EXC_ERR_UNSUPPORTED    # lmaxmat != lmaxapw
    # null-context calls
    assert lib.exc_match(None, 1, 1, None, None, None, None) == lapw.EXC_ERR_ARG
    assert lib.exc_hamiltonandoverlapsetup(None, 1, 1, None, None, None, 1, None, None) == lapw.EXC_ERR_ARG
    assert lib.exc_ctx_set_timing(None, 1) == lapw.EXC_ERR_ARG
    assert lib.exc_rhovalk(None, 1, None, None, 1, 1, None, 1, 1, None, None, None, C.c_double(1.0),
                           C.c_double(1e-8), None) == lapw.EXC_ERR_ARG
    assert lib.exc_rhovalk_dev(None, 1, None, None, 1, 1, None, 1, 1, None, None, None, C.c_double(1.0),
                               C.c_double(1e-8), None) == lapw.EXC_ERR_ARG
    assert lib.exc_eigen_batched_dev(None, 1, None, None, None, None, 1, None, None, None) == lapw.EXC_ERR_ARG
    assert lib.exc_ctx_destroy(None) == 0
    assert lib.exc_ctx_num_variants(2) >= 2
    assert lib.exc_ctx_num_variants(10) == 9          # eigensolve paths 0-8
    assert lib.exc_ctx_num_variants(15) == 11         # residue GEMM variants of the modular split
    assert lib.exc_ctx_num_variants(9) == 4           # thread blockings of the addition-theorem overlap
    assert lib.exc_ctx_num_variants(5) == 2           # grouped graph: overlap on a side branch or in line
    assert lib.exc_ctx_num_variants(17) == 2          # modular split: A and Yt, or the Gram form


def test_fortran_driver_built():
    assert os.path.exists(os.path.join(ROOT, "exciting_amd", "_lib", "exc_driver_main"))
