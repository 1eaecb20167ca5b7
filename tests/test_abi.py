"""CPU tests of the C-ABI boundary: the library loads, exports every entry
point include/*.h declares, keeps the reference's status codes, and rejects
bad arguments before touching a device (no compute call runs here)."""
import ctypes
import glob
import os
import re
import subprocess

import pytest

import pykblas

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADERS = sorted(glob.glob(os.path.join(ROOT, "include", "*.h")))


def declared_c_functions():
    names = set()
    for h in HEADERS:
        txt = open(h).read()
        # drop the inline C++ overload section
        txt = re.sub(r"#ifdef __cplusplus\s*/\*.*?#endif", "", txt, flags=re.S)
        for m in re.finditer(r"^\s*(?:int|void|double|const char\*|hipStream_t|void\*)\s+(kblas\w+)\s*\(", txt, re.M):
            names.add(m.group(1))
    return sorted(names)


def exported():
    out = subprocess.run(["nm", "-D", "--defined-only", pykblas.LIB_PATH], capture_output=True, text=True, check=True)
    return {l.split()[-1] for l in out.stdout.splitlines() if l.strip()}


def test_headers_parse():
    names = declared_c_functions()
    assert "kblasDpotrf_batch_strided" in names
    assert "kblasDposv_batch" in names
    assert "kblasDsposv_batch_strided" in names
    assert len(names) > 80


def test_every_declared_symbol_is_exported():
    missing = [n for n in declared_c_functions() if n not in exported()]
    assert not missing, missing


def test_library_loads_and_declares():
    lib = pykblas.load()
    assert lib.kblasGetErrorString(1) == b"success"
    assert lib.kblasGetErrorString(-6).startswith(b"insufficient workspace")


def test_status_codes_match_reference():
    txt = open(os.path.join(ROOT, "include", "kblas_defs.h")).read()
    codes = dict(re.findall(r"#define (KBLAS_\w+) (-?\d+)", txt))
    assert codes["KBLAS_Success"] == "1"
    assert codes["KBLAS_NotImplemented"] == "-2"
    assert codes["KBLAS_InsufficientWorkspace"] == "-6"
    assert codes["KBLAS_Error_WrongInput"] == "-10"


def test_null_handle_rejected_without_device():
    lib = pykblas.load()
    P = ctypes.c_void_p
    st = lib.kblasDpotrf_batch_strided(None, b"L", 8, P(0), 8, 64, 10, P(0))
    assert st == pykblas.KBLAS_Error_NotInitialized
    st = lib.kblasDposv_batch(None, b"R", b"L", 1, 8, P(0), 8, P(0), 1, 10, P(0))
    assert st == pykblas.KBLAS_Error_NotInitialized


def test_kernels_are_gfx950_code_objects():
    out = subprocess.run(["strings", pykblas.LIB_PATH], capture_output=True, text=True).stdout
    assert "gfx950" in out


def test_ara_argument_checks_without_device():
    lib = pykblas.load()
    P = ctypes.c_void_p
    z = P(0)
    assert lib.kblas_dara_fused_potrf_batch(z, z, z, z, z, z, 24, z, 4, z) == pykblas.KBLAS_NotImplemented
    assert lib.kblas_sara_fused_potrf_batch(z, z, z, z, z, z, 16, z, 0, z) == pykblas.KBLAS_Success
    assert lib.kblas_dara_fused_potrf_batch(z, z, z, z, z, z, 32, z, -1, z) == pykblas.KBLAS_Error_WrongInput
    assert lib.kblas_dara_fused_potrf_batch(z, z, z, z, z, z, 32, z, 3, z) == pykblas.KBLAS_Error_WrongInput
    assert lib.kblas_sara_trsm_batch(None, z, z, z, z, z, z, 4, 10, 16) == pykblas.KBLAS_Error_NotInitialized
    assert lib.kblas_sara_mp_syrk_batch(None, z, z, 10, 16, z, z, z, z, 4) == pykblas.KBLAS_Error_NotInitialized
