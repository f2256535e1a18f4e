"""The C-ABI library loads and exports every symbol include/dmdgpu.h declares
(no GPU needed: loading libdmdgpu.so does not touch the device)."""
import os
import re

from conftest import ROOT

import dmdgpu


def header_symbols():
    txt = open(os.path.join(ROOT, "include", "dmdgpu.h")).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(dmdgpu_[A-Za-z_0-9]+)\s*\(", txt)))


def test_header_matches_binding_list():
    assert header_symbols() == sorted(dmdgpu.EXPORTS)


def test_library_exports_all_symbols():
    L = dmdgpu.load_library()
    for name in header_symbols():
        assert hasattr(L, name), name


def test_kernel_names_without_device():
    L = dmdgpu.load_library()
    names = [L.dmdgpu_kernel_name(i).decode() for i in range(L.dmdgpu_kernel_count())]
    assert "phase2_gradient" in names and "phase1_site_sums" in names


def test_no_oracle_in_product_sources():
    """The product path must not reach the oracle (test infrastructure)."""
    pkg = os.path.join(ROOT, "bt-am_amd")
    for dirpath, _, files in os.walk(pkg):
        for f in files:
            if f.endswith((".py", ".cpp", ".hip", ".h")):
                src = open(os.path.join(dirpath, f)).read()
                assert "oracle" not in src.lower(), f
