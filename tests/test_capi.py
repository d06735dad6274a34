"""The C-ABI library loads and exports every entry point include/gpsat.h declares
(no compute calls here: this file runs without a GPU)."""
import ctypes
import os
import re

import pytest

import gpupsat_amd as G


def declared_symbols():
    text = open(G.HEADER_PATH).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(gpsat_[a-z0-9_]+)\s*\(", text)))


def test_library_built_and_loads():
    assert os.path.exists(G.LIB_PATH)
    G.lib()


def test_every_declared_symbol_is_exported():
    syms = declared_symbols()
    assert len(syms) > 10
    lib = G.lib()
    missing = [s for s in syms if not hasattr(lib, s)]
    assert not missing, missing


def test_version_and_errors():
    assert G.lib().gpsat_version() == b"v0.0.1"
    # an argument error sets the message and returns GPSAT_ERR_ARG without a device
    rc = G.lib().gpsat_batch_info(None, None, None, None, None)
    assert rc == -1
    assert b"null" in G.lib().gpsat_last_error()


def test_no_device_is_an_error_not_a_fallback():
    if G.device_count() > 0:
        pytest.skip("a GPU is visible")
    with pytest.raises(G.GpsatError):
        G.GpuBatch(G.Batch.from_formulas([G.Formula.from_clauses(2, [[1, 3]])]))
