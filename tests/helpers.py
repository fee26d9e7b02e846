"""Shared helpers for the parity tests (the oracle is the checker)."""

import os

import numpy as np

from oracle import sbe_oracle as O

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def rel(a, b):
    a = np.asarray(a)
    b = np.asarray(b)
    return float(np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300))


def wire_length(nk, me_rel=0.067, gap_ev=1.5):
    """Wire length giving exactly nk k-points (SBEs.py:4754-4758)."""
    from scipy.constants import e, hbar, m_e
    kmax = np.sqrt(1.2 * gap_ev * e * 2.0 * me_rel * m_e / hbar**2)
    return (nk / 2 + 0.5) * np.pi / kmax


def load_golden(tag):
    return np.load(os.path.join(GOLDEN, f"ref_{tag}.npz"))


def params_pair(L, flags=O.SHIPPED_FLAGS, me_rel=0.067):
    from pulsesuite_amd.tables import QWParams
    return (QWParams(L=L, me_rel=me_rel, flags=flags),
            O.Params(L=L, me_rel=me_rel, flags=flags))


def random_state(rng, n, nw, occ=0.2, coh=0.02, pol=0.01):
    """Levels (C2, D2, p2, C3, D3, p3) in the reference layout [n, n, nw]:
    Hermitian C, D with occupations on the diagonal, dense p."""
    out = []
    for _ in range(2):
        lv = []
        for kind in ("C", "D", "p"):
            a = np.zeros((n, n, nw), dtype=complex)
            for w in range(nw):
                m = rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n))
                if kind == "p":
                    a[:, :, w] = pol * m
                else:
                    h = (m + m.conj().T) * 0.5 * coh
                    np.fill_diagonal(h, rng.uniform(0.0, occ, n))
                    a[:, :, w] = h
            lv.append(a)
        out.extend(lv)
    return tuple(out)


def pulse(t, E0, tw, tp, w0):
    """sbetestprop.py:187-192."""
    x = w0 * (t - tp)
    return (E0 * np.exp(-(x**2) / (w0 * tw) ** 2) * np.cos(x)
            * np.exp(-(x**20) / (2 * tw * w0) ** 20))
