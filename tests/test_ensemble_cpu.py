"""Seeded ensemble definition (host only): SplitMix64 known answers and the member draws."""
from ppkmhd_amd import _abi as A
from ppkmhd_amd.ensemble import SIDES, draw_members, ot_ensemble, splitmix64, unit


def test_splitmix64_known_answers():
    # first outputs of SplitMix64 seeded with 0 and with 1234567 (published test values of the
    # generator, e.g. the reference C implementation by S. Vigna)
    s, x = splitmix64(0)
    assert x == 0xE220A8397B1DCDAF
    s, x = splitmix64(s)
    assert x == 0x6E789E6AA1B965F4
    s, x = splitmix64(s)
    assert x == 0x06C45D188009454F
    s, x = splitmix64(1234567)
    assert x == 6457827717110365317
    s, x = splitmix64(s)
    assert x == 3203168211198807973


def test_unit_interval():
    assert unit(0) == 0.0
    assert unit((1 << 64) - 1) < 1.0


def test_ensemble_draws():
    d = draw_members(64, 0)
    assert len(d) == 64
    assert all(side in SIDES for side, _, _ in d)
    assert all(1.4 <= g < 2.0 and 0.5 <= b < 2.0 for _, g, b in d)
    assert len({side for side, _, _ in d}) == len(SIDES)
    assert d == draw_members(64, 0) and d != draw_members(64, 1)
    ps, prs, _ = ot_ensemble(64, 0, steps=500)
    assert all(p.n_step_max == 500 and p.dim == 2 and p.nx == p.ny for p in ps)
    assert all(pr.kind == A.PROBLEM_ORSZAG_TANG for pr in prs)
    assert [p.gamma0 for p in ps] == [g for _, g, _ in d]
