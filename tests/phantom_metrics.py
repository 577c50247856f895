"""Phantom-recovery measures shared by the oracle and GPU tests (not a test module).

The recovered RI volume is compared with the phantom's own refractive index on the output grid
(ctdm_synth_phantom): the contrast inside the cell, the nucleoli against the surrounding cytoplasm, and
the correlation over the volume. The bars and their reasons are stated in test_oracle_phantom.py.
"""
import numpy as np

NM = 1.337


def recovery_metrics(n, phantom, p, dx):
    """(contrast ratio inside, nucleolus excess over cytoplasm / true excess, correlation)."""
    t = phantom.real.astype(np.float64)
    n = n.real.astype(np.float64)
    V = n.shape[0]
    c = (np.arange(V) - V // 2) * dx * 1e6
    Z, Y, X = np.meshgrid(c, c, c, indexing="ij")  # volume x = image y, volume y = image x
    inner = (X / (p.semi_axes[1] * 1e6)) ** 2 + (Y / (p.semi_axes[0] * 1e6)) ** 2 + (Z / (p.semi_axes[2] * 1e6)) ** 2 < 0.7 ** 2
    nuc = t > p.n_cell + 1e-6
    ratio = (n[inner].mean() - NM) / (t[inner].mean() - NM)
    excess = (n[nuc].mean() - n[inner & ~nuc].mean()) / (p.n_nucleoli - p.n_cell)
    corr = float(np.corrcoef(n.ravel(), t.ravel())[0, 1])
    return float(ratio), float(excess), corr


def check_recovery(n, phantom, p, dx):
    ratio, excess, corr = recovery_metrics(n, phantom, p, dx)
    print(f"contrast ratio {ratio:.3f}, nucleolus excess {excess:.3f}, correlation {corr:.3f}")
    return abs(ratio - 1.0) <= 0.15 and excess >= 0.5 and corr >= 0.9
