"""Ridders' extrapolation used by the adjoint's finite-difference check:
against closed-form derivatives, with and without round-off-like noise."""
import math

import numpy as np


def central(f, x0):
    return lambda h: (f(x0 + h) - f(x0 - h)) / (2 * h)


def test_ridders_smooth_functions():
    from sgipc.sim import ridders
    cases = [(lambda x: math.sin(x) * math.exp(x), lambda x: math.exp(x) * (math.sin(x) + math.cos(x)), 0.7),
             (lambda x: math.log(1 + x * x), lambda x: 2 * x / (1 + x * x), 1.3),
             (lambda x: x ** 5 - 3 * x, lambda x: 5 * x ** 4 - 3, -0.4)]
    for f, df, x0 in cases:
        est, err, hmin = ridders(central(f, x0), 1e-1)
        assert abs(est - df(x0)) <= 1e-10 * max(1.0, abs(df(x0))), (est, df(x0))
        assert err <= 1e-8


def test_ridders_beats_fixed_step_under_noise():
    """f has a deterministic 1e-15-relative roughness (as a rollout's
    round-off); a 1e-6 central step then carries ~1e-9 absolute error, the
    extrapolation from a larger step does not, and a small derivative
    (the slope of a nearly flat direction) keeps its relative accuracy."""
    from sgipc.sim import ridders
    rng = np.random.default_rng(0)
    table = rng.uniform(-1, 1, size=4096)

    def noisy(g):
        return lambda x: g(x) * (1 + 1e-15 * table[int(abs(x) * 1e9) % 4096])
    for slope in (1e-2, 4e-7):
        g = lambda x, s=slope: 1.0 + s * x + 0.5 * x * x * 1e-3 + (1 - math.cos(3 * x)) * 1e-4 + x ** 3
        dg = lambda x, s=slope: s + 1e-3 * x + 3e-4 * math.sin(3 * x) + 3 * x * x
        x0 = 0.0
        f = noisy(g)
        fixed = central(f, x0)(1e-6)
        est, err, _ = ridders(central(f, x0), 1e-4)
        exact = dg(x0)
        assert abs(est - exact) <= 1e-4 * abs(exact), (slope, est, exact)
        assert abs(est - exact) <= abs(fixed - exact) + 1e-13
