"""Error measures shared by the parity tests.

Relative errors are taken per physical field: density, total energy, the momentum vector
(IU, IV, IW) and the magnetic field vector (IA, IB, IC).  A vector field is measured against
its largest component, so roundoff carried in from the O(1) terms of a flux into a component
that happens to be near zero (w-momentum of a 2D-like flow, say) is judged on the field's scale,
as a relative norm of the vector field is.
"""
import numpy as np

GROUPS = ((0,), (1,), (2, 3, 4), (5, 6, 7))


def _scales(y, reduce, floor):
    per_var = np.array([reduce(np.abs(y[v])) for v in range(8)])
    top = float(per_var.max())
    out = np.zeros(8)
    for g in GROUPS:
        s = max(float(per_var[list(g)].max()), floor * top, 1e-300)
        out[list(g)] = s
    return out


def rel_linf(a, b, p, floor=1e-6):
    """Relative L-inf over interior cells, max over the 8 variables of max|a - b| divided by the
    max |b| of the variable's physical field (floored at floor * the largest field)."""
    ii = p.interior()
    x, y = a[ii], b[ii]
    s = _scales(y, lambda z: float(z.max()), floor)
    return max(float(np.abs(x[v] - y[v]).max()) / s[v] for v in range(8))


def rel_l1(a, b, p, var=None, floor=1e-6):
    """Relative L1 over interior cells: sum |a - b| over sum |b| of the variable's physical
    field (max over variables, or one variable), same floor."""
    ii = p.interior()
    x, y = a[ii], b[ii]
    s = _scales(y, lambda z: float(z.sum()), floor)
    vs = range(8) if var is None else [var]
    return max(float(np.abs(x[v] - y[v]).sum()) / s[v] for v in vs)
