"""ctypes binding of the C ABI in include/pulsesuite_amd.h.

The library is loaded from the package's ``_lib`` directory.  Loading fails
loudly when it is missing: there is no CPU fallback for the device path.
"""

import ctypes
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "_lib", "libpulsesuite_amd.so")

PSA_ABI_VERSION = 5
PREC = {"fp64": 0, "fp32": 1, "bf16x3": 2}

c_i32 = ctypes.c_int32
c_dbl = ctypes.c_double
c_ptr = ctypes.c_void_p

_INTS = ["N", "Nr", "Nw", "Nx", "Nhw", "flags", "prec", "gx0", "gnx", "bx0", "bnx"]
_DBLS = ["hbar", "e0", "gap", "L", "ehint", "dcv", "area", "dkr", "dr_r", "drr",
         "gam_e", "gam_h", "gam_eh", "Edc", "Vscale", "n1d_max", "dhw"]
_TABS = ["kr", "Ee", "Eh", "vtab", "chi", "qeh", "EP", "HP", "EPT", "HPT", "xe", "xh", "didx",
         "iq", "homega", "hsq", "window", "rq", "ytr", "gmat", "m2", "phi", "twid", "lor"]
_WIRE = ["wire_id", "pA", "pB", "pX", "CA", "CB", "CX", "DA", "DB", "DX", "Heh",
         "Heo", "Hho", "hdiag", "prep", "dedh", "wmb", "exq", "efield", "src",
         "out", "obs", "tphase", "part", "spl", "kscr"]


class PsaGroup(ctypes.Structure):
    """Mirror of ``psa_group``; the size is checked against the library."""

    _fields_ = ([(n, c_i32) for n in _INTS]
                + [("nnz_mb", c_i32 * 8), ("tune", c_i32 * 4), ("pad0", c_i32)]
                + [(n, c_dbl) for n in _DBLS]
                + [(n, c_ptr) for n in _TABS]
                + [("mb_off", c_ptr * 8), ("mb_idx", c_ptr * 8), ("mb_val", c_ptr * 8)]
                + [(n, c_ptr) for n in _WIRE]
                + [("kscr_len", ctypes.c_int64), ("wid_host", c_ptr)]
                + [(n, c_ptr) for n in ("pR", "CR", "DR", "wtag", "lazy")])


_GP = ctypes.POINTER(PsaGroup)


class PsaPstd(ctypes.Structure):
    """Mirror of ``psa_pstd``; the size is checked against the library."""

    _fields_ = ([(n, c_i32) for n in ("N", "nw", "flags", "nobs")]
                + [(n, c_dbl) for n in ("dy", "eps", "mu", "dt", "sheet")]
                + [(n, c_ptr) for n in ("state", "mode", "abc", "twid", "jw", "obs", "pold",
                                        "jsrc", "ew", "rec")]
                + [("nrec", ctypes.c_int64), ("count", c_ptr)])


_PP = ctypes.POINTER(PsaPstd)


class PsaQw3d(ctypes.Structure):
    """Mirror of ``psa_qw3d``; the size is checked against the library."""

    _fields_ = ([(n, c_i32) for n in ("nx", "ny", "nz", "flags", "i0", "j0", "k0", "pad")]
                + [(n, c_dbl) for n in ("ax", "ay", "az", "dt", "dx", "eps")]
                + [(n, c_ptr) for n in ("qx", "qy", "qz", "gy", "gz", "tw", "Ex", "Ey", "Ez", "Jx",
                                        "Jy", "Jz", "Rho", "Fk", "El", "lines", "old")])


_QP = ctypes.POINTER(PsaQw3d)
PSA_QW3D_CURR, PSA_QW3D_RHO, PSA_QW3D_POL, PSA_QW3D_KEEP_OLD = 1, 2, 4, 8
PSA_PSTD_ABC = 1
PSA_PSTD_MAXW = 256
PSA_RHS_MULTI_MAX = 16
PSA_RHS_MULTI_QUEUE = 16
PSA_NOALLOC_SCRATCH_DOUBLES = 4 << 20

# name -> (restype, argtypes)
SIGNATURES = {
    "psa_abi_version": (ctypes.c_int, []),
    "psa_last_error": (ctypes.c_char_p, []),
    "psa_group_size": (ctypes.c_int, []),
    "psa_vint_table": (ctypes.c_int, [c_ptr, c_i32, c_ptr, c_i32, c_dbl, c_dbl,
                                      c_dbl, c_dbl, c_ptr, c_ptr]),
    "psa_prep": (ctypes.c_int, [_GP, c_i32, c_i32, c_ptr]),
    "psa_calc_h": (ctypes.c_int, [_GP, c_i32, c_i32, c_ptr]),
    "psa_lorentz_table": (ctypes.c_int, [_GP, c_ptr]),
    "psa_mb_list": (ctypes.c_int, [c_i32, c_i32, c_ptr, c_ptr, c_dbl, c_dbl, c_dbl, c_i32, c_ptr,
                                   c_ptr, c_ptr, c_ptr, c_ptr]),
    "psa_offdiag_tables": (ctypes.c_int, [_GP, c_i32, c_i32, c_ptr]),
    "psa_rhs": (ctypes.c_int, [_GP, c_i32, c_i32, c_i32, c_dbl, c_ptr]),
    "psa_relaxation": (ctypes.c_int, [_GP, c_i32, c_i32, c_dbl, c_ptr]),
    "psa_transport": (ctypes.c_int, [_GP, c_i32, c_i32, c_dbl, c_ptr]),
    "psa_finish": (ctypes.c_int, [_GP, c_i32, c_i32, c_ptr]),
    "psa_sources": (ctypes.c_int, [_GP, c_i32, c_i32, c_ptr]),
    "psa_prop2qw": (ctypes.c_int, [_GP, c_i32, c_i32, c_ptr]),
    "psa_qw2prop": (ctypes.c_int, [_GP, c_i32, c_i32, c_ptr]),
    "psa_step": (ctypes.c_int, [_GP, c_i32, c_i32, c_dbl, c_ptr]),
    "psa_step_noalloc": (ctypes.c_int, [_GP, c_i32, c_i32, c_dbl, c_ptr]),
    "psa_step_noalloc_scratch": (ctypes.c_int64, [_GP, c_i32]),
    "psa_step_rk4": (ctypes.c_int, [_GP, c_i32, c_i32, c_dbl, c_ptr]),
    "psa_drive": (ctypes.c_int, [_GP, c_i32, c_i32, c_ptr, c_dbl, c_ptr]),
    "psa_record": (ctypes.c_int, [_GP, c_i32, c_i32, c_i32, c_ptr, c_ptr, c_i32, c_ptr]),
    "psa_offdiag_extrema": (ctypes.c_int, [_GP, c_i32, c_i32, c_ptr]),
    "psa_xqw_table": (ctypes.c_int, [c_i32, c_ptr, c_ptr, c_ptr, c_ptr, c_ptr, c_ptr, c_dbl, c_dbl,
                                     c_dbl, c_dbl, c_dbl, c_i32, c_ptr, c_ptr]),
    "psa_qw3d_size": (ctypes.c_int, []),
    "psa_qw3d_drive": (ctypes.c_int, [_QP, _GP, c_ptr]),
    "psa_qw3d_deposit": (ctypes.c_int, [_QP, _GP, c_ptr]),
    "psa_pstd_size": (ctypes.c_int, []),
    "psa_pstd_to_k": (ctypes.c_int, [_PP, c_ptr]),
    "psa_pstd_to_real": (ctypes.c_int, [_PP, c_ptr, c_ptr]),
    "psa_pstd_sample": (ctypes.c_int, [_PP, c_ptr]),
    "psa_pstd_advance": (ctypes.c_int, [_PP, c_ptr]),
    "psa_pstd_gather": (ctypes.c_int, [_PP, _GP, c_i32, c_i32, c_i32, c_ptr]),
    "psa_pstd_scatter": (ctypes.c_int, [_PP, _GP, c_i32, c_i32, c_i32, c_ptr]),
    "psa_rhs_multi_bytes": (ctypes.c_int64, [c_ptr, c_i32]),
    "psa_rhs_multi_pack": (ctypes.c_int, [c_ptr, c_ptr, c_ptr, c_i32, c_i32, c_dbl, c_ptr]),
    "psa_rhs_multi": (ctypes.c_int, [c_ptr, c_ptr, c_ptr, c_ptr]),
    "psa_step_multi": (ctypes.c_int, [c_ptr, c_ptr, c_ptr, c_dbl, c_ptr]),
    "psa_average_levels": (ctypes.c_int, [_GP, c_i32, c_i32, c_ptr]),
    "psa_mod_mb_arrays": (ctypes.c_int, [c_i32, c_ptr, c_ptr, c_dbl, c_dbl, c_dbl, c_i32, c_ptr, c_ptr,
                                         c_ptr, c_ptr, c_ptr]),
    "psa_mod_screen": (ctypes.c_int, [c_i32, c_i32, c_dbl, c_dbl, c_ptr, c_ptr, c_dbl, c_ptr, c_ptr,
                                      c_ptr, c_ptr, c_ptr, c_ptr]),
    "psa_mod_mbc": (ctypes.c_int, [c_i32, c_i32, c_ptr, c_ptr, c_ptr, c_ptr, c_ptr, c_ptr, c_ptr, c_ptr,
                                   c_ptr]),
    "psa_mod_mbp": (ctypes.c_int, [c_i32, c_i32, c_ptr, c_ptr, c_ptr, c_ptr, c_ptr, c_dbl, c_ptr, c_ptr,
                                   c_ptr]),
    "psa_mod_dft": (ctypes.c_int, [c_i32, c_i32, c_i32, c_dbl, c_ptr, c_ptr, c_ptr]),
    "psa_mod_resample": (ctypes.c_int, [c_i32, c_i32, c_i32, c_ptr, c_ptr, c_i32, c_ptr, c_ptr, c_ptr]),
    "psa_mod_pol3": (ctypes.c_int, [c_i32, c_i32, c_ptr, c_ptr, c_ptr, c_ptr, c_ptr, c_ptr, c_dbl, c_dbl,
                                    c_dbl, c_ptr, c_ptr]),
    "psa_mod_rho5": (ctypes.c_int, [c_i32, c_i32, c_ptr, c_ptr, c_ptr, c_ptr, c_dbl, c_dbl, c_ptr, c_ptr,
                                    c_ptr, c_ptr]),
    "psa_mod_plasmonics": (ctypes.c_int, [c_i32, c_dbl, c_i32, c_ptr, c_ptr, c_ptr, c_ptr]),
    "psa_mod_gamma": (ctypes.c_int, [c_i32, c_i32, c_ptr, c_ptr, c_ptr, c_ptr, c_ptr, c_ptr, c_ptr, c_dbl,
                                     c_ptr, c_ptr]),
    "psa_mod_offdiag2": (ctypes.c_int, [c_i32, c_ptr, c_ptr, c_ptr, c_ptr, c_ptr, c_ptr, c_dbl, c_dbl, c_dbl,
                                        c_dbl, c_ptr, c_ptr, c_ptr]),
    "psa_mod_spont": (ctypes.c_int, [c_i32, c_i32, c_ptr, c_ptr, c_ptr, c_ptr, c_ptr, c_dbl, c_ptr, c_ptr,
                                     c_dbl, c_dbl, c_ptr, c_ptr, c_ptr]),
    "psa_mod_mveh": (ctypes.c_int, [c_i32, c_i32, c_ptr, c_ptr, c_ptr, c_ptr, c_ptr, c_ptr]),
    "psa_mod_renorm": (ctypes.c_int, [c_i32, c_i32, c_ptr, c_ptr, c_ptr, c_ptr, c_ptr, c_ptr]),
    "psa_mod_vn1n2": (ctypes.c_int, [c_i32, c_ptr, c_ptr, c_ptr, c_ptr, c_dbl, c_ptr, c_ptr, c_ptr, c_ptr,
                                     c_ptr]),
    "psa_mod_plspectrum": (ctypes.c_int, [c_i32, c_i32, c_ptr, c_ptr, c_ptr, c_ptr, c_dbl, c_dbl, c_ptr, c_ptr,
                                          c_dbl, c_dbl, c_dbl, c_dbl, c_ptr, c_ptr, c_ptr]),
    "psa_mod_vehint": (ctypes.c_int, [c_dbl, c_ptr, c_i32, c_dbl, c_dbl, c_dbl, c_ptr, c_ptr]),
    "psa_mod_offdiag1": (ctypes.c_int, [c_i32, c_i32, c_ptr, c_ptr, c_ptr, c_ptr, c_ptr, c_ptr, c_ptr, c_dbl,
                                        c_dbl, c_dbl, c_ptr, c_ptr]),
    "psa_mod_offdiag1_apply": (ctypes.c_int, [c_i32, c_ptr, c_ptr, c_ptr, c_ptr, c_ptr, c_dbl, c_ptr, c_ptr]),
}

_lib = None


class PsaError(RuntimeError):
    pass


def load(path=None):
    """Load the HIP library (once) and declare every entry point."""
    global _lib
    if _lib is not None:
        return _lib
    path = path or LIB_PATH
    # One HIP runtime per process: torch ships its own libamdhip64
    # (soname libamdhip64.so.7).  Importing torch first makes this library's
    # libamdhip64.so.7 dependency bind to that already-loaded copy instead of
    # pulling in a second runtime that would not see the device.
    import torch  # noqa: F401
    if not os.path.exists(path):
        raise PsaError(
            f"{path} is missing: build it with `python -m pulsesuite_amd._build` "
            "(the device path has no CPU fallback)")
    lib = ctypes.CDLL(path)
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    if lib.psa_abi_version() != PSA_ABI_VERSION:
        raise PsaError("ABI version mismatch")
    if lib.psa_group_size() != ctypes.sizeof(PsaGroup):
        raise PsaError(f"psa_group size mismatch: C {lib.psa_group_size()} vs "
                       f"ctypes {ctypes.sizeof(PsaGroup)}")
    if lib.psa_qw3d_size() != ctypes.sizeof(PsaQw3d):
        raise PsaError("psa_qw3d size mismatch")
    if lib.psa_pstd_size() != ctypes.sizeof(PsaPstd):
        raise PsaError(f"psa_pstd size mismatch: C {lib.psa_pstd_size()} vs "
                       f"ctypes {ctypes.sizeof(PsaPstd)}")
    _lib = lib
    return lib


def check(rc):
    if rc != 0:
        msg = _lib.psa_last_error().decode() if _lib is not None else ""
        raise PsaError(f"pulsesuite_amd error {rc}: {msg}")
    return rc
