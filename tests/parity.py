"""Error measures shared by the parity tests."""
import numpy as np


def rel_linf(a, b, p, floor=1e-6):
    """Relative L-inf over interior cells: per variable, max |a - b| / max |b|, the max over the
    variables.  A variable whose magnitude is below floor * (largest variable magnitude) -- an
    exactly-zero momentum component, say -- is measured against that floor instead."""
    ii = p.interior()
    x, y = a[ii], b[ii]
    top = float(np.abs(y).max())
    worst = 0.0
    for v in range(8):
        s = max(float(np.abs(y[v]).max()), floor * top, 1e-300)
        worst = max(worst, float(np.abs(x[v] - y[v]).max()) / s)
    return worst


def rel_l1(a, b, p, var=None, floor=1e-6):
    """Relative L1 over interior cells: sum |a - b| / sum |b| per variable (max over variables,
    or one variable), with the same floor for near-zero variables."""
    ii = p.interior()
    x, y = a[ii], b[ii]
    top = float(np.abs(y).sum(axis=(1, 2, 3)).max())
    vs = range(8) if var is None else [var]
    worst = 0.0
    for v in vs:
        s = max(float(np.abs(y[v]).sum()), floor * top, 1e-300)
        worst = max(worst, float(np.abs(x[v] - y[v]).sum()) / s)
    return worst
