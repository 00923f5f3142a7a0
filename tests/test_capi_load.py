"""The C-ABI library loads on the CPU, exports every symbol include/voxscatter.h declares, and reports
a missing device as an error code (no compute on the GPU here)."""
import ctypes as C
import math
import re
from pathlib import Path

import numpy as np
import pytest

from voxscatter_amd import capi

HEADER = Path(__file__).resolve().parents[1] / "include" / "voxscatter.h"


def declared_symbols():
    text = HEADER.read_text()
    return sorted(set(re.findall(r"\b(vs_[A-Za-z0-9_]+)\s*\(", text)))


def test_header_and_binding_agree():
    assert set(declared_symbols()) == set(capi.SIGNATURES), "ctypes signatures must cover the header exactly"


def test_library_exports_every_declared_symbol():
    lib = capi.load()
    for name in declared_symbols():
        assert hasattr(lib, name), name


def test_no_device_is_an_error_code():
    lib = capi.load()
    n = C.c_int(-1)
    assert lib.vs_device_count(C.byref(n)) == capi.VS_OK
    if n.value > 0:
        pytest.skip("a GPU is present")
    h = C.c_void_p()
    rc = lib.vs_ctx_create(0, C.byref(h))
    assert rc == capi.VS_ERR_NODEV
    assert b"no HIP device" in lib.vs_last_error()


def test_alpha_ldr_host_entry_matches_oracle():
    from oracle.operator import alpha_ldr
    from voxscatter_amd.api import alpha_LDR
    for m, kv, e in [(1.31, [0, 0, 1], [1, 0, 0]), (1.31 + 0.01j, [1, 0, 0], [0, 0, 1]),
                     (1.5 + 0.2j, [1, 1, 0], [1, 0, 0])]:
        k = 2 * math.pi
        kv = k * np.asarray(kv, dtype=float) / np.linalg.norm(kv)
        got = alpha_LDR(m, k, 0.07, kv, e)
        ref = alpha_ldr(m, k, 0.07, kv, e)
        assert abs(got - ref) <= 1e-15 * abs(ref)


def test_argument_errors_are_reported():
    lib = capi.load()
    out = np.zeros(2)
    rc = lib.vs_alpha_ldr(None, 1.0, 1.0, None, None, capi.dptr(out))
    assert rc == capi.VS_ERR_ARG
