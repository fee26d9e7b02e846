"""The C-ABI library loads on the CPU host and exports every entry point
include/pulsesuite_amd.h declares (no GPU calls are made)."""

import os
import re

from pulsesuite_amd import capi

HEADER = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))),
                      "include", "pulsesuite_amd.h")


def declared_functions():
    txt = open(HEADER).read()
    return sorted(set(re.findall(r"^\s*(?:int|int64_t|const char \*)\s*(psa_\w+)\s*\(", txt, re.M)))


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


def test_multi_rhs_packing_checks_without_the_gpu():
    """psa_rhs_multi_pack validates every group and orders them by
    decreasing N; psa_rhs_multi refuses a description it did not write."""
    import ctypes
    import struct
    lib = capi.load()
    gs = []
    for N, prec in ((16, 0), (32, 0), (24, 0)):
        g = capi.PsaGroup()
        g.N, g.Nr, g.Nw, g.Nx, g.prec = N, 2 * N + 2, 2, 64, prec
        gs.append(g)
    n = len(gs)
    ptrs = (ctypes.c_void_p * n)(*[ctypes.addressof(g) for g in gs])
    w0 = (ctypes.c_int32 * n)(0, 0, 0)
    nw = (ctypes.c_int32 * n)(2, 1, 2)
    cp = lambda a: ctypes.cast(a, ctypes.c_void_p)  # noqa: E731
    assert lib.psa_rhs_multi_bytes(cp(nw), 0) == -1
    assert lib.psa_rhs_multi_bytes(cp(nw), capi.PSA_RHS_MULTI_MAX + 1) == -1
    buf = ctypes.create_string_buffer(int(lib.psa_rhs_multi_bytes(cp(nw), n)))
    assert lib.psa_rhs_multi_pack(cp(ptrs), cp(w0), cp(nw), n, 0, 0.0, cp(buf)) == 0
    M = capi.PSA_RHS_MULTI_MAX
    hdr = struct.unpack_from("4i", buf.raw)
    xstart = struct.unpack_from("9i", buf.raw, 16)
    xcnt = struct.unpack_from("8i", buf.raw, 16 + 36)
    gx = struct.unpack_from(f"{M}i", buf.raw, 16 + 36 + 32)
    # wires in decreasing N: N = 32 (1 wire), 24 (2), 16 (2), dealt to XCD lists
    # 0..4; one 64-tile per equation, 3 items per wire
    assert hdr[0] == 3 and hdr[3] == 5
    assert xstart == (0, 1, 2, 3, 4, 5, 5, 5, 5) and xcnt == (3, 3, 3, 3, 3, 0, 0, 0)
    assert gx[:3] == (1, 1, 1)
    assert lib.psa_rhs_multi_pack(cp(ptrs), cp(w0), cp(nw), n, 9, 0.0, cp(buf)) == -1  # mode
    gs[1].prec = 2
    assert lib.psa_rhs_multi_pack(cp(ptrs), cp(w0), cp(nw), n, 0, 0.0, cp(buf)) == -1
    assert b"precision" in lib.psa_last_error()
    gs[1].prec = 0
    nw[0] = 3  # outside the group
    assert lib.psa_rhs_multi_pack(cp(ptrs), cp(w0), cp(nw), n, 0, 0.0, cp(buf)) == -1
    zero = ctypes.create_string_buffer(len(buf.raw))
    q = (ctypes.c_uint32 * 2)()
    assert lib.psa_rhs_multi(cp(zero), cp(zero), cp(q), None) == -1


def test_noalloc_step_refuses_calls_beyond_its_scratch_without_touching_the_gpu():
    """psa_step_noalloc states the split-K scratch a call needs and refuses
    one above the library's maximum before any HIP call."""
    import ctypes
    lib = capi.load()
    g = capi.PsaGroup()
    g.N, g.Nr, g.Nw, g.Nx, g.prec, g.bnx = 256, 514, 64, 4096, 0, 600
    gp = ctypes.byref(g)
    need = lib.psa_step_noalloc_scratch(gp, 64)
    assert need == 16 * max(514 * 4 * 64, 511 * 3 * 64, 600 * 5 * 64)
    assert need <= capi.PSA_NOALLOC_SCRATCH_DOUBLES
    g.bnx, g.Nw = 4096, 200
    assert lib.psa_step_noalloc_scratch(gp, 200) > capi.PSA_NOALLOC_SCRATCH_DOUBLES
    assert lib.psa_step_noalloc(gp, 0, 200, 1e-17, None) == -1
    assert b"above the library's maximum" in lib.psa_last_error()
    assert lib.psa_step_noalloc(gp, 5, 200, 1e-17, None) == -1  # wire range outside the group
