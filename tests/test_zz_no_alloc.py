"""psa_step_noalloc (no counterpart in the reference; for callers that
cannot allocate) through the C ABI against psa_step with the caller's
full-size split-K workspace: bit-identical state and outputs, on two calls
in a row (the second with a smaller input: fewer wires of a smaller N, so
the library's scratch holds the first call's larger partial sums), in the
FP64 and split-BF16 modes."""

import ctypes

import numpy as np
import pytest

from tests.helpers import params_pair, random_state, wire_length

pytestmark = pytest.mark.gpu


def _pair(N, nw, prec, seed):
    from pulsesuite_amd.engine import WireGroup
    L = wire_length(N)
    prm, _ = params_pair(L)
    rr = np.linspace(-1.2 * L, 1.2 * L, 160)
    st = random_state(np.random.default_rng(seed), N, nw)
    E = 3e7 * np.cos(2 * np.pi * rr / (3 * L)) + 0j
    out = []
    for _ in range(2):
        wg = WireGroup(prm, nw, rr, prec=prec)
        wg.set_state(*st)
        wg.set_fields(E, 2.0 * E, 0.3 * E)
        out.append(wg)
    return out


@pytest.mark.parametrize("prec", ["fp64", "bf16x3"])
def test_noalloc_step_equals_step(prec):
    import torch
    from pulsesuite_amd import capi
    lib = capi.load()
    dt = 10e-18
    for N, nw, seed in ((256, 6, 1), (64, 2, 2)):
        print(f"[noalloc] prec={prec} N={N} nw={nw}", flush=True)
        a, b = _pair(N, nw, prec, seed)
        need = lib.psa_step_noalloc_scratch(ctypes.byref(b.S), nw)
        assert 0 < need <= capi.PSA_NOALLOC_SCRATCH_DOUBLES
        # the caller's workspace is full size, so psa_step uses every split
        assert a.S.kscr_len >= need
        b.S.kscr = None  # psa_step_noalloc ignores it
        b.S.kscr_len = 0
        a.step(dt)
        stream = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
        capi.check(lib.psa_step_noalloc(ctypes.byref(b.S), 0, nw, dt, stream))
        b.advance_levels()
        torch.cuda.synchronize()
        for x, y in zip(a.get_state(), b.get_state()):
            assert np.array_equal(x, y), (N, nw)
        for x, y in zip(a.outputs(), b.outputs()):
            assert np.array_equal(x, y), (N, nw)
        assert np.abs(a.outputs()[0]).max() > 0
        del a, b
