"""Off-diagonal dephasing tables (psa_offdiag_tables) in the reduced modes
against FP64 on the same state.  FP32 runs the FP32 MFMA kernel (the
Lorentzian made in FP32 registers, 2^100-scaled occupation columns); the
split-BF16 mode runs the bf16 MFMA kernel with both operands split hi/lo
(three products, FP32 accumulation, product error ~2^-16).  Both stay
within 1e-5 of the FP64 tables, relative to the largest entry; the terms
are all positive, so the sums have no cancellation.  The step error of each
mode is covered by test_gpu_parity."""

import numpy as np
import pytest

from tests.helpers import params_pair, random_state, rel, wire_length

pytestmark = pytest.mark.gpu


# (N, wires): 70 wires take two column blocks of the split-BF16 kernel, N =
# 512 two row blocks and up to 32 K chunks
@pytest.mark.parametrize("N,nw", [(64, 5), (256, 5), (512, 3), (64, 70)])
def test_reduced_mode_offdiag_tables_close_to_fp64(N, nw):
    import torch
    from pulsesuite_amd.engine import WireGroup
    print(f"[offdiag modes] N={N} nw={nw}", flush=True)
    prm, _ = params_pair(wire_length(N))
    rr = np.linspace(-1.2 * wire_length(N), 1.2 * wire_length(N), 96)
    st = random_state(np.random.default_rng(N), N, min(nw, 5))
    if nw > 5:
        st = [np.concatenate([a] * (nw // 5 + 1), axis=2)[:, :, :nw] for a in st]
    tabs = {}
    for prec in ("fp64", "fp32", "bf16x3"):
        wg = WireGroup(prm, nw, rr, prec=prec)
        wg.set_state(*st)
        wg.prep()
        wg.offdiag()
        torch.cuda.synchronize()
        tabs[prec] = wg.s["dedh"].cpu().numpy()
    ref = tabs["fp64"]
    assert np.abs(ref).max() > 0
    e32, e16 = rel(tabs["fp32"], ref), rel(tabs["bf16x3"], ref)
    print(f"[offdiag modes] N={N} nw={nw} fp32 {e32:.2e} bf16x3 {e16:.2e}", flush=True)
    assert e32 < 1e-5 and e16 < 1e-5
    # the entries the reference leaves at zero (q = 0, k + q outside [0, N))
    # stay exactly zero in every mode
    k = np.arange(N)[:, None]
    q = np.arange(2 * N - 1)[None, :] - (N - 1)
    unused = (q == 0) | (k + q < 0) | (k + q >= N)
    for prec in ("fp32", "bf16x3"):
        assert not np.any(tabs[prec][:, :, unused]), prec
