"""C-ABI library: loads without a GPU, exports every symbol the header
declares, and its host-side helpers agree with the fixtures."""
import json
import os
import re

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def header_symbols():
    text = open(os.path.join(ROOT, "include", "sdot_mi355x.h")).read()
    return sorted(set(re.findall(r"\b(sdot_[a-z0-9_]+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    import ctypes
    import sdot_amd
    L = ctypes.CDLL(sdot_amd.LIB_PATH)
    syms = header_symbols()
    assert len(syms) >= 19
    for s in syms:
        assert hasattr(L, s), s
    # and the Python front-end binds all of them
    assert set(syms) <= set(sdot_amd.SIGNATURES)


def test_rng_matches_fixture():
    import sdot_amd
    ref = json.load(open(os.path.join(ROOT, "tests", "golden", "rng_first8.json")))
    for seed, vals in ref.items():
        assert np.array_equal(sdot_amd.rng_uniform(int(seed), 8), np.array(vals))


def test_status_strings():
    import sdot_amd
    L = sdot_amd.lib()
    for st in range(9):
        assert len(L.sdot_status_string(st)) > 0


def test_invalid_arguments_are_status_codes():
    import ctypes
    import sdot_amd
    L = sdot_amd.lib()
    h = ctypes.c_void_p()
    assert L.sdot_pd_create(ctypes.byref(h), 4, 0) == sdot_amd.SDOT_INVALID_ARG
    assert L.sdot_pd_create(ctypes.byref(h), 2, 9) == sdot_amd.SDOT_INVALID_ARG
    assert L.sdot_pd_destroy(None) == sdot_amd.SDOT_INVALID_ARG
    assert L.sdot_pd_compute(None) == sdot_amd.SDOT_INVALID_ARG
    # batches: 2D only; null handles are argument errors
    assert L.sdot_batch_create(ctypes.byref(h), 3, 0) == sdot_amd.SDOT_NOT_SUPPORTED
    assert L.sdot_batch_create(ctypes.byref(h), 2, 9) == sdot_amd.SDOT_INVALID_ARG
    assert L.sdot_batch_create(None, 2, 0) == sdot_amd.SDOT_INVALID_ARG
    assert L.sdot_batch_destroy(None) == sdot_amd.SDOT_INVALID_ARG
    assert L.sdot_batch_set_tuning(None, 0, 1) == sdot_amd.SDOT_INVALID_ARG
    assert L.sdot_batch_set_problems(None, 1, None, None, None, None, None, 0) == sdot_amd.SDOT_INVALID_ARG
    assert L.sdot_batch_solve(None, 1e-6, 0, 10, None, None, None, None, None) == sdot_amd.SDOT_INVALID_ARG
    assert L.sdot_batch_last_error(None) == b"null handle"


def test_host_layer_header_compiles():
    import subprocess
    src = os.path.join(ROOT, "sdot_amd", "host", "host_example.cpp")
    r = subprocess.run(["g++", "-std=c++17", "-fsyntax-only", "-I" + os.path.join(ROOT, "sdot_amd", "host"), src],
                       capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
