"""Python face of the surrogate form: the network (``MLP``), its training samples
(``FaceSet``) and the helpers that train it from a solver's own faces.  Every call goes through
the C ABI (include/ppkmhd_amd.h, ppk_mlp_* / ppk_faceset_*); the GPU does the work.
"""
import ctypes as C

import numpy as np

from ._lib import check, check_with, lib

N_IN, N_HID, N_OUT = 16, 128, 8
# offsets in the FP32 parameter vector (mlp.hpp)
W1, B1 = 0, 128 * 16
W2 = B1 + 128
B2 = W2 + 128 * 128
W3 = B2 + 128
B3 = W3 + 8 * 128
N_TRAINED = B3 + 8
MU_IN, ISIG_IN = N_TRAINED, N_TRAINED + 16
MU_OUT, SIG_OUT = N_TRAINED + 32, N_TRAINED + 40
N_PARAMS = N_TRAINED + 48

_FPP = C.POINTER(C.c_float)


def _fp(a):
    return a.ctypes.data_as(_FPP)


def unpack(P):
    """dict of the parameter arrays of a flat FP32 parameter vector."""
    P = np.asarray(P, dtype=np.float32)
    return {"W1": P[W1:B1].reshape(128, 16), "b1": P[B1:W2], "W2": P[W2:B2].reshape(128, 128),
            "b2": P[B2:W3], "W3": P[W3:B3].reshape(8, 128), "b3": P[B3:N_TRAINED],
            "mu_in": P[MU_IN:ISIG_IN], "isig_in": P[ISIG_IN:MU_OUT], "mu_out": P[MU_OUT:SIG_OUT],
            "sig_out": P[SIG_OUT:N_PARAMS]}


class MLP:
    """The surrogate network on the GPU (16 -> 128 -> 128 -> 8, ReLU, bf16 MFMA)."""

    def __init__(self, seed=0):
        assert lib().ppk_mlp_param_count() == N_PARAMS
        h = C.c_void_p()
        check(lib().ppk_mlp_create(C.c_uint64(seed), C.byref(h)))
        self._h = h

    def close(self):
        if getattr(self, "_h", None):
            lib().ppk_mlp_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _check(self, rc):
        return check_with(rc, self._h, lib().ppk_mlp_last_error)

    @property
    def handle(self):
        return self._h

    def get_params(self):
        P = np.empty(N_PARAMS, dtype=np.float32)
        self._check(lib().ppk_mlp_get_params(self._h, _fp(P), N_PARAMS))
        return P

    def set_params(self, P):
        P = np.ascontiguousarray(P, dtype=np.float32)
        assert P.size == N_PARAMS
        self._check(lib().ppk_mlp_set_params(self._h, _fp(P), N_PARAMS))

    def standardize(self, faceset):
        self._check(lib().ppk_mlp_standardize(self._h, faceset.handle))

    def train(self, faceset, steps, batch=4096, lr=1e-3, seed=1):
        self._check(lib().ppk_mlp_train(self._h, faceset.handle, int(steps), int(batch),
                                        C.c_float(lr), C.c_uint64(seed)))

    def gradient(self, X, T):
        X = np.ascontiguousarray(X, dtype=np.float32)
        T = np.ascontiguousarray(T, dtype=np.float32)
        g = np.empty(N_TRAINED, dtype=np.float32)
        self._check(lib().ppk_mlp_gradient(self._h, _fp(X), _fp(T), X.shape[0], _fp(g), N_TRAINED))
        return g

    def infer(self, X):
        X = np.ascontiguousarray(X, dtype=np.float32)
        Y = np.empty((X.shape[0], N_OUT), dtype=np.float32)
        self._check(lib().ppk_mlp_infer(self._h, _fp(X), X.shape[0], _fp(Y)))
        return Y

    def evaluate(self, faceset):
        mse = (C.c_double * 8)()
        self._check(lib().ppk_mlp_evaluate(self._h, faceset.handle, mse))
        return np.array(list(mse))


class FaceSet:
    """Training samples on the GPU: features X[n][16] and targets T[n][8]."""

    def __init__(self, capacity):
        h = C.c_void_p()
        check(lib().ppk_faceset_create(C.c_size_t(int(capacity)), C.byref(h)))
        self._h = h
        self.capacity = int(capacity)

    def close(self):
        if getattr(self, "_h", None):
            lib().ppk_faceset_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _check(self, rc):
        return check_with(rc, self._h, lib().ppk_faceset_last_error)

    @property
    def handle(self):
        return self._h

    def __len__(self):
        n = C.c_size_t()
        self._check(lib().ppk_faceset_size(self._h, C.byref(n)))
        return n.value

    def clear(self):
        self._check(lib().ppk_faceset_clear(self._h))

    def append(self, X, T):
        X = np.ascontiguousarray(X, dtype=np.float32)
        T = np.ascontiguousarray(T, dtype=np.float32)
        self._check(lib().ppk_faceset_append(self._h, _fp(X), _fp(T), X.shape[0]))

    def read(self, first=0, n=None):
        if n is None:
            n = len(self) - first
        X = np.empty((n, N_IN), dtype=np.float32)
        T = np.empty((n, N_OUT), dtype=np.float32)
        self._check(lib().ppk_faceset_read(self._h, first, n, _fp(X), _fp(T)))
        return X, T


def dump_faces(solver, faceset, stride=1, phase=0):
    """Append the samples of the faces the solver's next step will solve; returns the count."""
    n = C.c_size_t()
    solver._check(lib().ppk_mhd_dump_faces(solver._h, faceset.handle, int(stride), int(phase),
                                           C.byref(n)))
    return n.value


def attach(solver, net, gate_rho=0.1, gate_p=0.1):
    """Use `net` for the face solves of `solver` (None: plain HLLD).

    The library keeps only the network's device weights, so the solver holds a reference to
    `net` until it is detached: the weights cannot be freed while steps may still read them."""
    solver._check(lib().ppk_mhd_set_surrogate(solver._h, net.handle if net is not None else None,
                                              float(gate_rho), float(gate_p)))
    solver._sur_net = net


def counts(solver, reset=False):
    """(faces solved by the network, faces solved by HLLD) since the last reset."""
    a, b = C.c_int64(), C.c_int64()
    solver._check(lib().ppk_mhd_surrogate_counts(solver._h, C.byref(a), C.byref(b),
                                                 1 if reset else 0))
    return a.value, b.value


def gated(X, gate_rho, gate_p):
    """Mask of the samples whose face passes the gate of the face kernels (the relative jumps
    are scale-free: x0 is the density jump over the mean density, x7 and x14 the mean and the
    jump of the pressure in the same unit)."""
    X = np.asarray(X, dtype=np.float64)
    dr = np.abs(X[:, 0]) < gate_rho * (1.0 - 0.5 * np.abs(X[:, 0]))
    dp = np.abs(X[:, 14]) < gate_p * (X[:, 7] - 0.5 * np.abs(X[:, 14]))
    return dr & dp


def select_gated(faceset, gate_rho, gate_p, chunk=1 << 22):
    """A new FaceSet with the samples of `faceset` that pass the gate (the faces the network
    will be asked about)."""
    n = len(faceset)
    keep_x, keep_t = [], []
    for first in range(0, n, chunk):
        X, T = faceset.read(first, min(chunk, n - first))
        m = gated(X, gate_rho, gate_p)
        keep_x.append(X[m])
        keep_t.append(T[m])
    X = np.concatenate(keep_x) if keep_x else np.zeros((0, N_IN), np.float32)
    T = np.concatenate(keep_t) if keep_t else np.zeros((0, N_OUT), np.float32)
    out = FaceSet(max(1, X.shape[0]))
    if X.shape[0]:
        out.append(X, T)
    return out


def train_schedule(net, faceset, steps, batch=4096, lr=1e-3, seed=1, phases=4):
    """`steps` Adam steps in `phases` equal parts, the learning rate divided by 3 per part."""
    per = max(1, steps // phases)
    for k in range(phases):
        net.train(faceset, per, batch=batch, lr=lr / 3 ** k, seed=seed + k)
