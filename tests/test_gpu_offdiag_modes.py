"""Off-diagonal dephasing tables (psa_offdiag_tables) in the reduced modes
against FP64 on the same state.  Both reduced modes run the FP32 kernel
(the Lorentzian made in FP32 registers, FP32 MFMA on 2^100-scaled
occupation columns).  They stay within 1e-5 of the FP64 tables, relative to
the largest entry; the terms are all positive, so the sums have no
cancellation.  The step error of each mode is covered by test_gpu_parity."""

import numpy as np
import pytest

from tests.helpers import params_pair, random_state, rel, wire_length

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("N", [64, 256])
def test_reduced_mode_offdiag_tables_close_to_fp64(N):
    import torch
    from pulsesuite_amd.engine import WireGroup
    print(f"[offdiag modes] N={N}", flush=True)
    prm, _ = params_pair(wire_length(N))
    rr = np.linspace(-1.2 * wire_length(N), 1.2 * wire_length(N), 96)
    st = random_state(np.random.default_rng(N), N, 5)
    tabs = {}
    for prec in ("fp64", "fp32", "bf16x3"):
        wg = WireGroup(prm, 5, rr, prec=prec)
        wg.set_state(*st)
        wg.prep()
        wg.offdiag()
        torch.cuda.synchronize()
        tabs[prec] = wg.s["dedh"].cpu().numpy()
    ref = tabs["fp64"]
    assert np.abs(ref).max() > 0
    e32, e16 = rel(tabs["fp32"], ref), rel(tabs["bf16x3"], ref)
    print(f"[offdiag modes] N={N} fp32 {e32:.2e} bf16x3 {e16:.2e}", flush=True)
    assert e32 < 1e-5 and e16 < 1e-5
