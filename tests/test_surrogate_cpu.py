"""Pinning of the surrogate oracle (no GPU): the network gradient against finite differences,
bf16 rounding against known values, the dimensionless face features and targets against the
similarity invariance of ideal MHD, the LLF restatement against the oracle's LLF solver, and
the admissibility check against the property LLF has by construction."""
import numpy as np
import pytest

from ppkmhd_amd import _abi as A
from ppkmhd_amd import make_params

from oracle import surrogate_oracle as S


def random_params(seed, scale=1.0):
    rng = np.random.default_rng(seed)
    P = np.zeros(S.N_PARAMS, dtype=np.float32)
    P[S.W1:S.B1] = rng.uniform(-1, 1, 128 * 16) * np.sqrt(6 / 16) * scale
    P[S.B1:S.W2] = rng.uniform(-0.1, 0.1, 128)
    P[S.W2:S.B2] = rng.uniform(-1, 1, 128 * 128) * np.sqrt(6 / 128) * scale
    P[S.B2:S.W3] = rng.uniform(-0.1, 0.1, 128)
    P[S.W3:S.B3] = rng.uniform(-1, 1, 8 * 128) * np.sqrt(6 / 128)
    P[S.B3:S.N_TRAINED] = rng.uniform(-0.1, 0.1, 8)
    P[S.MU_IN:S.ISIG_IN] = rng.uniform(-0.5, 0.5, 16)
    P[S.ISIG_IN:S.MU_OUT] = rng.uniform(0.5, 2.0, 16)
    P[S.MU_OUT:S.SIG_OUT] = rng.uniform(-0.1, 0.1, 8)
    P[S.SIG_OUT:S.N_PARAMS] = rng.uniform(0.5, 2.0, 8)
    return P


def test_bf16_rounding_known_values():
    one = 1.0
    assert S.bf16(one) == 1.0
    assert S.bf16(1 + 2.0 ** -8) == 1.0                    # tie -> even mantissa
    assert S.bf16(1 + 3 * 2.0 ** -8) == 1 + 2.0 ** -6      # tie -> even (up)
    assert S.bf16(1 + 2.0 ** -8 + 2.0 ** -12) == 1 + 2.0 ** -7
    assert S.bf16(-3.0) == -3.0
    assert S.bf16(np.float32(3.14159274)) == 3.140625


def test_mlp_gradient_matches_finite_differences():
    rng = np.random.default_rng(3)
    P = random_params(1).astype(np.float64)
    X = rng.normal(size=(6, 16))
    T = rng.normal(size=(6, 8))
    g = S.mlp_grad(P, X, T, rounding=False)
    idx = np.concatenate([rng.choice(np.arange(S.W1, S.B1), 12, replace=False),
                          rng.choice(np.arange(S.W2, S.B2), 12, replace=False),
                          rng.choice(np.arange(S.W3, S.B3), 12, replace=False),
                          np.arange(S.B1, S.B1 + 4), np.arange(S.B2, S.B2 + 4),
                          np.arange(S.B3, S.N_TRAINED)])
    h = 1e-6
    for i in idx:
        Pp, Pm = P.copy(), P.copy()
        Pp[i] += h
        Pm[i] -= h
        fd = (S.mlp_loss(Pp, X, T) - S.mlp_loss(Pm, X, T)) / (2 * h)
        assert abs(fd - g[i]) <= 1e-6 * max(1.0, abs(fd)), (i, fd, g[i])


def test_mlp_gradient_rounding_is_small_perturbation():
    """bf16 rounding points change the gradient by O(2^-8) relative, not more."""
    rng = np.random.default_rng(4)
    P = random_params(2)
    X = rng.normal(size=(512, 16)).astype(np.float32)
    T = rng.normal(size=(512, 8)).astype(np.float32)
    g0 = S.mlp_grad(P, X, T, rounding=False)
    g1 = S.mlp_grad(P, X, T, rounding=True)
    rel = np.linalg.norm(g1 - g0) / np.linalg.norm(g0)
    assert 1e-5 < rel < 3e-2, rel


def physical_state(rng):
    return np.array([rng.uniform(0.2, 3.0), rng.uniform(0.05, 3.0), *rng.normal(size=3),
                     *rng.normal(size=3)])


def hlld_of(gamma0):
    p = make_params(2, 8, 8, gamma0=gamma0, riemann=A.RIEMANN_HLLD, float_round=False)
    from oracle import pyoracle

    pyoracle.build()
    return lambda ql, qr: pyoracle.riemann(p, ql, qr)


def test_llf_restatement_matches_oracle_llf():
    from oracle import pyoracle

    pyoracle.build()
    rng = np.random.default_rng(5)
    for _ in range(50):
        g = rng.uniform(1.4, 2.0)
        p = make_params(2, 8, 8, gamma0=g, riemann=A.RIEMANN_LLF, float_round=False)
        ql, qr = physical_state(rng), physical_state(rng)
        F, *_ = S.llf(ql, qr, g)
        np.testing.assert_allclose(F, pyoracle.riemann(p, ql, qr), rtol=1e-13, atol=1e-13)


def scale_state(q, a, b):
    d, p, u, v, w, bx, by, bz = q
    return np.array([a * d, a * b * b * p, b * u, b * v, b * w,
                     np.sqrt(a) * b * bx, np.sqrt(a) * b * by, np.sqrt(a) * b * bz])


@pytest.mark.parametrize("a,b", [(2.7, 0.6), (0.05, 13.0), (400.0, 1e-2)])
def test_features_and_targets_are_similarity_invariant(a, b):
    """rho -> a rho, v -> b v, B -> sqrt(a) b B, p -> a b^2 p leaves features and targets
    unchanged (a wrong exponent in the scales would not)."""
    rng = np.random.default_rng(6)
    for _ in range(20):
        g = rng.uniform(1.4, 2.0)
        hl = hlld_of(g)
        ql, qr = physical_state(rng), physical_state(rng)
        x0, t0 = S.face_sample(hl, ql, qr, g)
        x1, t1 = S.face_sample(hl, scale_state(ql, a, b), scale_state(qr, a, b), g)
        np.testing.assert_allclose(x1, x0, rtol=1e-11, atol=1e-12)
        np.testing.assert_allclose(t1, t0, rtol=1e-8, atol=1e-10)


def test_target_rebuilt_from_features():
    """The normalised face a feature vector describes reproduces the target."""
    rng = np.random.default_rng(7)
    for _ in range(20):
        g = rng.uniform(1.4, 2.0)
        hl = hlld_of(g)
        ql, qr = physical_state(rng), physical_state(rng)
        x, t = S.face_sample(hl, ql, qr, g)
        nl, nr, g2 = S.states_from_features(x)
        assert g2 == pytest.approx(g, abs=1e-15)
        F, s, *_ = S.llf(nl, nr, g)
        assert s == pytest.approx(1.0, rel=1e-12)
        x2, t2 = S.face_sample(hl, nl, nr, g)
        np.testing.assert_allclose(t2, t, rtol=1e-9, atol=1e-11)


def test_llf_flux_is_admissible_and_a_bad_flux_is_not():
    rng = np.random.default_rng(8)
    for _ in range(200):
        g = rng.uniform(1.4, 2.0)
        ql, qr = physical_state(rng), physical_state(rng)
        F, s, *_ = S.llf(ql, qr, g)
        assert S.admissible(ql, qr, F, g)
        bad = F.copy()
        bad[S.ID] += 2.0 * s * max(ql[S.ID], qr[S.ID])
        assert not S.admissible(ql, qr, bad, g)


def test_attached_network_outlives_its_last_user_reference(monkeypatch):
    """attach() keeps the network alive while the solver may use its device weights: the library
    stores only a raw pointer to them, so dropping the caller's last reference must not free
    the network; detaching (or closing the solver) releases it."""
    import gc
    import weakref

    from ppkmhd_amd import surrogate as SG

    class StubLib:
        def ppk_mhd_set_surrogate(self, h, net, gate_rho, gate_p):
            return 0

    class StubNet:
        handle = object()

    class StubSolver:
        _h = object()

        def _check(self, rc):
            assert rc == 0

    monkeypatch.setattr(SG, "lib", lambda: StubLib())
    solver, net = StubSolver(), StubNet()
    ref = weakref.ref(net)
    SG.attach(solver, net, 0.01, 0.01)
    del net
    gc.collect()
    assert ref() is not None, "the attached network was released while still in use"
    SG.attach(solver, None)
    gc.collect()
    assert ref() is None, "detaching should release the network"


def test_layer_slips_exceed_the_gpu_parity_bars():
    """The GPU network is held to 1e-3 of the bf16 oracle (inference) and 2e-4 (gradient
    blocks).  Plausible slips of one layer move the oracle's own output and gradient by more
    than those bars, so a kernel carrying one of them would fail the GPU tests."""
    P = random_params(11)
    X = np.random.default_rng(12).normal(size=(1000, 16)).astype(np.float32)
    T = np.random.default_rng(14).normal(size=(1000, 8)).astype(np.float32)[:512]
    ref, _ = S.mlp_forward(P, X, rounding=True)
    gref = S.mlp_grad(P, X[:512], T, rounding=True)

    def slips():
        for a, b in ((S.B1, S.W2), (S.B2, S.W3), (S.B3, S.N_TRAINED)):
            q = P.copy()
            q[a:b] = 0.0  # a dropped bias
            yield q
        for a, b in ((S.W1, S.B1), (S.W2, S.B2), (S.W3, S.B3)):
            q = P.copy()
            q[a:b] *= 1.01  # one layer's weights 1 % off
            yield q
        q = P.copy()  # layer 2 applied transposed
        q[S.W2:S.B2] = P[S.W2:S.B2].reshape(128, 128).T.ravel()
        yield q

    for q in slips():
        y, _ = S.mlp_forward(q, X, rounding=True)
        assert np.abs(y - ref).max() / np.abs(ref).max() > 1e-3
        g = S.mlp_grad(q, X[:512], T, rounding=True)
        assert np.linalg.norm(g - gref) / np.linalg.norm(gref) > 2e-4
