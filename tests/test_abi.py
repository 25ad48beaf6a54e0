"""The C-ABI library loads and exports every symbol include/sdft_hip.h
declares; host-only entry points (no GPU compute) agree with the oracle."""
import ctypes
import json
import os
import re

import numpy as np

import pyoracle
from sdft_amd import native
from sdft_amd import synthetic as syn

HEADER = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "include", "sdft_hip.h")


def declared_symbols():
    text = open(HEADER).read()
    return sorted(set(re.findall(r"\b(sdft_hip_[a-z0-9_]+)\s*\(", text)))


def test_every_declared_symbol_is_exported():
    lib = native.lib()
    names = declared_symbols()
    assert len(names) >= 15
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing


def test_capabilities_json():
    caps = json.loads(native.capabilities())
    assert caps["hip_compiled"] is True and caps["schema_version"] == "sparse-dft-compute-capabilities-1"


def test_invalid_arguments_map_to_status_2():
    lib = native.lib()
    rc = lib.sdft_hip_zgemm(None, 0, 0, ctypes.c_size_t(0), ctypes.c_size_t(1), ctypes.c_size_t(1), None,
                            ctypes.c_size_t(1), None, ctypes.c_size_t(1), None, ctypes.c_size_t(1), 0,
                            ctypes.c_uint64(0), 0)
    assert rc == 2
    assert "invalid complex GEMM" in native.lib().sdft_hip_last_error().decode()


def test_product_sphere_is_bit_exact_with_oracle():
    for cfg, k in [(0, (0, 0, 0)), (0, (.125, .375, -.125)), (1, (0, 0, 0))]:
        cell = syn.CONFIGS[cfg]["cell"]()
        cutoff = syn.CONFIGS[cfg]["ecut"] * syn.HARTREE_EV + 1e-8
        a = native.enumerate_g_vectors(cell.lattice, cutoff, k)
        b = pyoracle.enumerate_g(cell.lattice, cutoff, k)
        assert np.array_equal(a[0], b[0])
        assert np.array_equal(a[1].view(np.uint64), b[1].view(np.uint64))  # identical FP64 sort keys
        assert np.array_equal(a[2].view(np.uint64), b[2].view(np.uint64))


def test_product_index_map_is_bit_exact_with_oracle(si8, si8_oracle):
    idx = native.grid_indices(si8_oracle.shape, si8.hkl)
    assert np.array_equal(idx, si8_oracle.indices)
    # closed form of src/plane_wave_operator.cpp:94-97
    s = si8_oracle.shape
    h = si8.hkl.astype(np.int64)
    expect = ((h[:, 0] % s[0]) * s[1] + (h[:, 1] % s[1])) * s[2] + (h[:, 2] % s[2])
    assert np.array_equal(idx.astype(np.int64), expect)
