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


def test_entry_points_reject_bad_arguments_without_touching_the_gpu():
    """Argument checks return PSA_E_INVALID with a message before any HIP
    call, so they run on a host without a GPU."""
    import ctypes
    lib = capi.load()
    g = capi.PsaGroup()
    g.N, g.Nr, g.Nw, g.Nx, g.prec = 16, 34, 2, 64, 0
    gp = ctypes.byref(g)
    assert lib.psa_rhs(gp, 0, 2, 7, 1e-17, None) == -1
    assert b"mode" in lib.psa_last_error()
    assert lib.psa_rhs(gp, 0, 2, 3, 1e-17, None) == -1  # RK4 stage without pR, CR, DR
    assert lib.psa_step_rk4(gp, 0, 2, 1e-17, None) == -1
    assert lib.psa_step(gp, 1, 2, 1e-17, None) == -1  # wire range outside the group
    p = capi.PsaPstd()
    p.N = 100  # not a power of two
    assert lib.psa_pstd_advance(ctypes.byref(p), None) == -1
    assert b"power of two" in lib.psa_last_error()
    p.N = 256
    assert lib.psa_pstd_advance(ctypes.byref(p), None) == -1  # no buffers
    p.nw = 300
    assert lib.psa_pstd_gather(ctypes.byref(p), gp, 0, 1, 0, None) == -1
    q = capi.PsaQw3d()
    q.nx, q.ny, q.nz = 64, 12, 7  # 7 is not 2^a 3^b
    assert lib.psa_qw3d_drive(ctypes.byref(q), gp, None) == -1
    assert lib.psa_xqw_table(0, None, None, None, None, None, None, 1.0, 1.0, 1.0, 1.0, 1.0, 1, None,
                             None) == -1
