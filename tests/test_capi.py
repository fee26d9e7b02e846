"""The C-ABI library loads on the CPU host and exports every entry point
include/pulsesuite_amd.h declares (no GPU calls are made)."""

import os
import re

from pulsesuite_amd import capi

HEADER = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))),
                      "include", "pulsesuite_amd.h")


def declared_functions():
    txt = open(HEADER).read()
    return sorted(set(re.findall(r"^\s*(?:int|const char \*)\s*(psa_\w+)\s*\(", txt, re.M)))


def test_header_declares_the_bound_functions():
    names = declared_functions()
    assert names, "no declarations parsed"
    assert set(names) == set(capi.SIGNATURES)


def test_library_exports_every_declared_symbol():
    lib = capi.load()
    for n in declared_functions():
        assert hasattr(lib, n), n


def test_struct_mirror_matches():
    lib = capi.load()
    import ctypes
    assert lib.psa_group_size() == ctypes.sizeof(capi.PsaGroup)
    assert lib.psa_abi_version() == capi.PSA_ABI_VERSION
    assert lib.psa_pstd_size() == ctypes.sizeof(capi.PsaPstd)
