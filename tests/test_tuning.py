"""CPU tests of the per-size-bucket kernel table (host logic of kb_tune.c):
entries set and queried through the C ABI, later entries win, variants that
cannot serve an order are rejected, JSON tables load (the committed one
included) and a malformed one leaves the built-in rule."""
import json
import os

import pytest

import pykblas

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
TABLE = os.path.join(ROOT, "kblas-gpu_amd", "tuning.json")
VARS = ["auto", "rg16", "rg32", "rg4", "tile", "reg", "blk"]


@pytest.fixture()
def lib():
    L = pykblas.load()
    L.kblas_tuning_clear()
    yield L
    L.kblas_tuning_clear()


def q(L, prec, op, n):
    return L.kblas_tuning_query(prec.encode(), op.encode(), n)


def test_set_query_and_override(lib):
    assert q(lib, "d", "posv", 10) == 0                       # built-in rule
    assert lib.kblas_tuning_set(b"d", b"posv", 1, 16, b"rg16") == pykblas.KBLAS_Success
    assert lib.kblas_tuning_set(b"d", b"posv", 8, 12, b"blk") == pykblas.KBLAS_Success
    assert VARS[q(lib, "d", "posv", 5)] == "rg16"
    assert VARS[q(lib, "d", "posv", 10)] == "blk"             # the later entry wins
    assert VARS[q(lib, "d", "posv", 16)] == "rg16"
    assert q(lib, "d", "posv", 17) == 0
    assert q(lib, "s", "posv", 5) == 0                        # per precision
    assert q(lib, "d", "potrf", 5) == 0                       # per operation
    assert lib.kblas_tuning_variant_name(4) == b"tile"


def test_illegal_variants_rejected(lib):
    bad = pykblas.KBLAS_Error_WrongInput
    assert lib.kblas_tuning_set(b"d", b"posv", 10, 20, b"rg16") == bad      # rg16 serves n <= 16
    assert lib.kblas_tuning_set(b"d", b"posv", 33, 40, b"rg4") == bad       # rg4 is POTRF only
    assert lib.kblas_tuning_set(b"d", b"potrf", 30, 40, b"rg4") == bad      # and 32 < n <= 64
    assert lib.kblas_tuning_set(b"d", b"potrf", 60, 70, b"tile") == bad     # tile serves n <= 64
    assert lib.kblas_tuning_set(b"d", b"sposv", 1, 8, b"blk") == bad        # mixed keeps refinement
    assert lib.kblas_tuning_set(b"d", b"sposv", 30, 40, b"reg") == bad
    assert lib.kblas_tuning_set(b"d", b"nosuch", 1, 8, b"tile") == bad
    assert lib.kblas_tuning_set(b"x", b"posv", 1, 8, b"tile") == bad
    assert lib.kblas_tuning_set(b"d", b"vsplit", 0, 100, b"65") == bad
    assert lib.kblas_tuning_set(b"d", b"potrf", 33, 64, b"rg4") == pykblas.KBLAS_Success
    assert lib.kblas_tuning_set(b"d", b"potrf", 65, 4000, b"blk") == pykblas.KBLAS_Success
    assert lib.kblas_tuning_set(b"d", b"sposv", 1, 32, b"reg") == pykblas.KBLAS_Success


def test_vsplit(lib):
    assert q(lib, "d", "vsplit", 256) == -1
    assert lib.kblas_tuning_set(b"d", b"vsplit", 0, 1 << 30, b"48") == pykblas.KBLAS_Success
    assert q(lib, "d", "vsplit", 256) == 48
    assert q(lib, "s", "vsplit", 256) == -1


def test_load_json(lib, tmp_path):
    t = {"device": "x", "entries": [
        {"prec": "d", "op": "potrf", "n_lo": 1, "n_hi": 16, "variant": "rg16", "ms": {"1": {"rg16": 0.1}}},
        {"prec": "d", "op": "potrf", "n_lo": 17, "n_hi": 64, "variant": "tile"},
        {"prec": "s", "op": "vsplit", "n_lo": 0, "n_hi": 1000000, "variant": 32},
    ]}
    p = tmp_path / "t.json"
    p.write_text(json.dumps(t))
    assert lib.kblas_tuning_load(str(p).encode()) == pykblas.KBLAS_Success
    assert VARS[q(lib, "d", "potrf", 9)] == "rg16"
    assert VARS[q(lib, "d", "potrf", 40)] == "tile"
    assert q(lib, "s", "vsplit", 100) == 32
    # malformed or illegal tables are refused and leave the built-in rule
    p.write_text('{"entries": [{"prec": "d", "op": "potrf", "n_lo": 1, "n_hi": 40, "variant": "rg16"}]}')
    assert lib.kblas_tuning_load(str(p).encode()) == pykblas.KBLAS_Error_WrongInput
    assert q(lib, "d", "potrf", 9) == 0
    p.write_text('{"entries": [{"prec": "d", "op": ')
    assert lib.kblas_tuning_load(str(p).encode()) == pykblas.KBLAS_Error_WrongInput
    assert lib.kblas_tuning_load(str(tmp_path / "missing.json").encode()) == pykblas.KBLAS_Error_WrongInput


@pytest.mark.skipif(not os.path.exists(TABLE), reason="no measured table committed")
def test_committed_table_loads_and_covers_the_register_range(lib):
    assert lib.kblas_tuning_load(TABLE.encode()) == pykblas.KBLAS_Success
    t = json.load(open(TABLE))
    for e in t["entries"]:
        n = (e["n_lo"] + e["n_hi"]) // 2
        got = q(lib, e["prec"], e["op"], n)
        want = e["variant"] if e["op"] == "vsplit" else VARS.index(e["variant"])
        assert got == want, e
    for prec, ops in (("d", ("potrf", "posv", "potrs", "sposv")), ("s", ("potrf", "posv", "potrs"))):
        for op in ops:
            for n in range(1, 65):
                assert q(lib, prec, op, n) > 0, (prec, op, n)
