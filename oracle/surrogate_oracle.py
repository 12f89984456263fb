"""CPU restatement of the surrogate form (test infrastructure only: tests and the benchmark's
CPU leg use it as the checker, never the library).

* ``mlp_forward`` / ``mlp_grad``: the 16 -> 128 -> 128 -> 8 ReLU network and the gradient of
  its training loss, in float64, either exact or with the library's bf16 rounding points
  (standardised inputs, hidden activations, weights, output and pre-activation gradients
  rounded to bf16 with round-to-nearest-even before each matrix product).
* ``face_sample``: the features and training target of one face from its two primitive states,
  with the fluxes from the oracle's own Riemann solvers (HLLD and LLF of
  src/shared/RiemannSolvers_MHD.h:81-111, 133-367).

There is no reference counterpart (ppkMHD has no learned flux): these functions are pinned by
mathematics in tests/test_surrogate_cpu.py (finite differences, similarity invariance of the
dimensionless targets, LLF admissibility).
"""
import numpy as np

W1, B1 = 0, 128 * 16
W2 = B1 + 128
B2 = W2 + 128 * 128
W3 = B2 + 128
B3 = W3 + 8 * 128
N_TRAINED = B3 + 8
MU_IN, ISIG_IN = N_TRAINED, N_TRAINED + 16
MU_OUT, SIG_OUT = N_TRAINED + 32, N_TRAINED + 40
N_PARAMS = N_TRAINED + 48
GAMMA_REF = 5.0 / 3.0


def bf16(x):
    """Round to bf16 (round to nearest even) and return as float64."""
    f = np.ascontiguousarray(np.asarray(x, dtype=np.float32))
    u = f.view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    return u.astype(np.uint32).view(np.float32).astype(np.float64)


def params(P):
    P = np.asarray(P, dtype=np.float64)
    return (P[W1:B1].reshape(128, 16), P[B1:W2], P[W2:B2].reshape(128, 128), P[B2:W3],
            P[W3:B3].reshape(8, 128), P[B3:N_TRAINED], P[MU_IN:ISIG_IN], P[ISIG_IN:MU_OUT],
            P[MU_OUT:SIG_OUT], P[SIG_OUT:N_PARAMS])


def standardize_inputs(P, X):
    """(x - mu_in) * isig_in in float32, as the kernels compute it."""
    w = np.asarray(P, dtype=np.float32)
    mu, isig = w[MU_IN:ISIG_IN], w[ISIG_IN:MU_OUT]
    return ((np.asarray(X, dtype=np.float32) - mu) * isig).astype(np.float32)


def standardize_targets(P, T):
    w = np.asarray(P, dtype=np.float32)
    return ((np.asarray(T, dtype=np.float32) - w[MU_OUT:SIG_OUT]) / w[SIG_OUT:N_PARAMS]).astype(
        np.float32)


def mlp_forward(P, X, rounding=True):
    """Standardised outputs Y[n][8] and the intermediate values."""
    W1_, b1, W2_, b2, W3_, b3 = params(P)[:6]
    r = bf16 if rounding else (lambda a: np.asarray(a, dtype=np.float64))
    xs = r(standardize_inputs(P, X))
    z1 = xs @ r(W1_).T + b1
    h1 = r(np.maximum(z1, 0.0))
    z2 = h1 @ r(W2_).T + b2
    h2 = r(np.maximum(z2, 0.0))
    y = h2 @ r(W3_).T + b3
    return y, (xs, z1, h1, z2, h2)


def mlp_loss(P, X, T, rounding=False):
    y, _ = mlp_forward(P, X, rounding)
    t = standardize_targets(P, T).astype(np.float64)
    return np.mean((y - t) ** 2)


def mlp_grad(P, X, T, rounding=True):
    """Gradient of mean((y - t_std)^2) w.r.t. the 19720 trained parameters (flat)."""
    W1_, b1, W2_, b2, W3_, b3 = params(P)[:6]
    r = bf16 if rounding else (lambda a: np.asarray(a, dtype=np.float64))
    y, (xs, z1, h1, z2, h2) = mlp_forward(P, X, rounding)
    t = standardize_targets(P, T).astype(np.float64)
    n = X.shape[0]
    dy = 2.0 * (y - t) / (n * 8)
    if rounding:  # the kernel forms dy in float32 before rounding it
        dy = 2.0 * (y.astype(np.float32) - t.astype(np.float32)).astype(np.float64)
        dy = (dy.astype(np.float32) * np.float32(1.0 / (n * 8))).astype(np.float64)
    dyr = r(dy)
    dz2 = r((dyr @ r(W3_)) * (z2 > 0))
    dz1 = r((dz2 @ r(W2_)) * (z1 > 0))
    g = np.zeros(N_TRAINED)
    g[W1:B1] = (dz1.T @ xs).ravel()
    g[B1:W2] = dz1.sum(0)
    g[W2:B2] = (dz2.T @ h1).ravel()
    g[B2:W3] = dz2.sum(0)
    g[W3:B3] = (dyr.T @ h2).ravel()
    g[B3:N_TRAINED] = dyr.sum(0)
    return g


# ---- per-face pieces ------------------------------------------------------------------------
ID, IP, IU, IV, IW, IA, IB, IC = range(8)


def mhd_flux(q, gamma0):
    """Conservative state and ideal MHD flux of a primitive state (src/shared/mhd_utils.h:173-231)."""
    d, p, u, v, w, a, b, c = q
    ecin = 0.5 * (u * u + v * v + w * w) * d
    emag = 0.5 * (a * a + b * b + c * c)
    etot = p / (gamma0 - 1.0) + ecin + emag
    ptot = p + emag
    U = np.array([d, etot, d * u, d * v, d * w, a, b, c])
    F = np.array([d * u, (etot + ptot) * u - a * (a * u + b * v + c * w), d * u * u - a * a + ptot,
                  d * u * v - a * b, d * u * w - a * c, 0.0, b * u - a * v, c * u - a * w])
    return U, F


def signal_speed(q, gamma0):
    d, p, u, a, b, c = q[ID], q[IP], q[IU], q[IA], q[IB], q[IC]
    b2 = a * a + b * b + c * c
    c2 = gamma0 * p / d
    d2 = 0.5 * (b2 / d + c2)
    return np.sqrt(d2 + np.sqrt(d2 * d2 - c2 * a * a / d)) + abs(u)


def llf(ql, qr, gamma0):
    """(F_LLF, s, UL, UR, FL, FR) with Bx averaged, src/shared/RiemannSolvers_MHD.h:81-111."""
    ql, qr = np.array(ql, float), np.array(qr, float)
    bx = 0.5 * (ql[IA] + qr[IA])
    ql[IA] = qr[IA] = bx
    UL, FL = mhd_flux(ql, gamma0)
    UR, FR = mhd_flux(qr, gamma0)
    s = max(signal_speed(ql, gamma0), signal_speed(qr, gamma0))
    return 0.5 * (FL + FR) - 0.5 * s * (UR - UL), s, UL, UR, FL, FR


def features(ql, qr, gamma0):
    """16 dimensionless features (means and jumps of the two states), rho_bar and s."""
    _, s, *_ = llf(ql, qr, gamma0)
    bx = 0.5 * (ql[IA] + qr[IA])
    rb = 0.5 * (ql[ID] + qr[ID])
    ib = 1.0 / (np.sqrt(rb) * s)
    ip = 1.0 / (rb * s * s)
    m = lambda v: 0.5 * (ql[v] + qr[v])
    d = lambda v: qr[v] - ql[v]
    x = [d(ID) / rb, m(IU) / s, m(IV) / s, m(IW) / s, bx * ib, m(IB) * ib, m(IC) * ib,
         m(IP) * ip, gamma0 - GAMMA_REF, d(IU) / s, d(IV) / s, d(IW) / s, d(IB) * ib,
         d(IC) * ib, d(IP) * ip, 0.0]
    return np.array(x), rb, s


def scales(rb, s):
    m, b = rb * s, np.sqrt(rb) * s * s
    return np.array([m, m * s * s, m * s, m * s, m * s, 0.0, b, b])


def face_sample(hlld, ql, qr, gamma0):
    """(features, target) of a face; hlld(ql, qr) -> the HLLD flux (8 components)."""
    F, s, *_ = llf(ql, qr, gamma0)
    x, rb, s = features(ql, qr, gamma0)
    sc = scales(rb, s)
    fh = np.asarray(hlld(np.array(ql, float), np.array(qr, float)))
    t = np.where(np.arange(8) == IA, 0.0, (fh - F) / np.where(sc == 0, 1.0, sc))
    return x, t


def states_from_features(x):
    """The normalised face (rho_bar = 1, s = 1) the features describe, and gamma."""
    x = np.asarray(x, dtype=np.float64)
    mean = np.array([1.0, x[7], x[1], x[2], x[3], x[4], x[5], x[6]])
    jump = np.array([x[0], x[14], x[9], x[10], x[11], 0.0, x[12], x[13]])
    return mean - 0.5 * jump, mean + 0.5 * jump, float(x[8]) + GAMMA_REF


def admissible(ql, qr, F, gamma0):
    """The check of a candidate flux: U_L - (F - F_L)/s and U_R + (F - F_R)/s physical."""
    _, s, UL, UR, FL, FR = llf(ql, qr, gamma0)
    a = UL - (F - FL) / s
    b = UR + (F - FR) / s

    def pres(u):
        ke = 0.5 * (u[IU] ** 2 + u[IV] ** 2 + u[IW] ** 2) / u[ID]
        me = 0.5 * (u[IA] ** 2 + u[IB] ** 2 + u[IC] ** 2)
        return (gamma0 - 1.0) * (u[IP] - ke - me)

    return a[ID] > 0 and b[ID] > 0 and pres(a) > 0 and pres(b) > 0
