"""The C-ABI library loads and exports every symbol include/*.h declares
(no compute calls: runs without a GPU)."""
import ctypes
import glob
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    names = set()
    for h in glob.glob(os.path.join(ROOT, "include", "*.h")):
        text = open(h).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        for m in re.finditer(r"\b(kspir_[a-z0-9_]+)\s*\(", text):
            names.add(m.group(1))
    return names


def test_headers_declare_the_reference_entry_points():
    names = declared_symbols()
    for n in ("kspir_matrix_vector_mul_bsgs_rns_crt_large", "kspir_external_product",
              "kspir_packing_rlwes", "kspir_answer", "kspir_database_create", "kspir_keys_create"):
        assert n in names


def test_library_exports_every_declared_symbol():
    path = os.path.join(ROOT, "kspir_amd", "libkspir_amd.so")
    if not os.path.exists(path):
        pytest.fail("libkspir_amd.so not built (run python kspir_amd/build.py)")
    lib = ctypes.CDLL(path)
    missing = [n for n in sorted(declared_symbols()) if not hasattr(lib, n)]
    assert not missing, missing


def test_python_binding_lists_every_symbol():
    import kspir_amd
    assert set(kspir_amd.EXPORTS) == declared_symbols()


def test_oracle_library_exports():
    from oracle import oracle as O
    lib = O.lib()
    hdr = open(os.path.join(ROOT, "oracle", "kspir_oracle.h")).read()
    for m in re.finditer(r"\b(ko_[a-z0-9_]+)\s*\(", hdr):
        assert hasattr(lib, m.group(1)), m.group(1)


def test_product_permutation_matrix_matches_oracle():
    """kspir_compute_permutation_matrix (host-side, no GPU) == utils.cpp:1591"""
    import kspir_amd
    from oracle import oracle as O
    for n1 in (4, 128, 512):
        assert (kspir_amd.compute_permutation_matrix(n1) == O.perm_matrix(n1)).all()
