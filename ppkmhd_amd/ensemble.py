"""Seeded ensembles of 2D Orszag-Tang runs for the batched solver.

The ensemble workload: ``count`` independent runs whose side length is drawn uniformly from
{64, 128, 192, 256, 384, 512}, gamma uniformly in [1.4, 2.0] and the field amplitude B0
multiplied by a factor drawn uniformly in [0.5, 2], FP64, periodic, HLLD + CT, minmod slopes,
CFL 0.8.  The draws come from SplitMix64 (Steele, Lea and Flood, "Fast splittable pseudorandom
number generators", OOPSLA 2014) so that any language reproduces the same members from the
seed: per member, three consecutive outputs give the size index (output mod 6), gamma and the
B0 factor (the top 53 bits as a double in [0, 1), mapped affinely).
"""
from . import _abi as A
from .params import default_problem, make_params

SIDES = (64, 128, 192, 256, 384, 512)
_MASK = (1 << 64) - 1


def splitmix64(state):
    """One SplitMix64 step: (new state, output)."""
    state = (state + 0x9E3779B97F4A7C15) & _MASK
    z = state
    z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & _MASK
    z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & _MASK
    return state, z ^ (z >> 31)


def unit(x):
    """Top 53 bits of a 64-bit output as a double in [0, 1)."""
    return (x >> 11) * (1.0 / (1 << 53))


def draw_members(count=64, seed=0):
    """[(side, gamma, b0_factor)] of the ensemble."""
    s = seed & _MASK
    out = []
    for _ in range(count):
        s, a = splitmix64(s)
        s, b = splitmix64(s)
        s, c = splitmix64(s)
        out.append((SIDES[a % len(SIDES)], 1.4 + 0.6 * unit(b), 0.5 + 1.5 * unit(c)))
    return out


def ot_ensemble(count=64, seed=0, steps=500, precision=A.FP64, tEnd=1.0e9):
    """(params list, problem list, draws): each member runs exactly `steps` steps (n_step_max)
    unless it reaches tEnd first."""
    ps, prs, draws = [], [], draw_members(count, seed)
    for side, gamma, b0 in draws:
        p = make_params(2, side, side, gamma0=gamma, cfl=0.8, slope_type=1.0,
                        riemann=A.RIEMANN_HLLD, bc=A.BC_PERIODIC, tEnd=tEnd, n_step_max=steps,
                        precision=precision, float_round=False)
        pr = default_problem(A.PROBLEM_ORSZAG_TANG, p)
        pr.ot_b0_scale = b0
        ps.append(p)
        prs.append(pr)
    return ps, prs, draws
