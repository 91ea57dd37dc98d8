"""pyOptimizer.MMA (MethodOfMovingAsymptotes.hh, binding python_bindings/Optimizer.cc:11-22), host C++, against
the oracle's numpy restatement (oracle/mma.py) iterate by iterate, and both against what the mathematics fixes:
Svanberg's two-constraint toy problem converges to its published optimum x = (2.0175, 1.7800, 1.2375), and a
separable volume-constrained compliance-like problem min sum c_j / x_j s.t. mean(x) <= V converges to its KKT
point x_j proportional to sqrt(c_j) (clipped to the bounds). Runs on the CPU (the optimiser has no GPU code)."""
import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


@pytest.fixture(scope="module")
def pyOptimizer():
    sys.path.insert(0, os.path.join(ROOT, "voxelfem_amd"))
    try:
        import pyOptimizer as mod
    finally:
        sys.path.pop(0)
    return mod


def toy():
    c1, c2 = np.array([5.0, 2.0, 1.0]), np.array([3.0, 4.0, 3.0])

    def f(x):
        return np.array([x @ x, (x - c1) @ (x - c1) - 9, (x - c2) @ (x - c2) - 9])

    def df(x):
        return np.stack([2 * x, 2 * (x - c1), 2 * (x - c2)])

    return f, df


@pytest.mark.parametrize("gcmma", [False, True])
def test_toy_problem_known_optimum_and_oracle(pyOptimizer, gcmma):
    from oracle.mma import MMA as OMMA

    f, df = toy()
    a = pyOptimizer.MMA(3, 2, np.zeros(3), np.full(3, 5.0), f, df)
    b = OMMA(3, 2, np.zeros(3), np.full(3, 5.0), f, df)
    for o in (a, b):
        o.enableGCMMA(gcmma)
        o.setInitialVar(np.array([4.0, 3.0, 2.0]))
    # the subproblem's interior-point iteration stops at eps < 1e-7 with its KKT residual <= 0.9 eps, so two
    # implementations that round differently can take a different number of Newton steps: the iterates agree to
    # that accuracy. Near the optimum the GCMMA feasibility test (f(x_inner) below its approximation) compares
    # quantities that vanish there, so the number of inner iterations depends on rounding: its iterates are
    # compared until then, and both runs must reach the optimum.
    for k in range(40):
        a.step()
        b.step()
        if not gcmma or k < 6:
            assert np.allclose(a.getOptimalVar(), b.x(), rtol=0, atol=1e-7), k
    assert np.allclose(b.x(), [2.0175, 1.7800, 1.2375], atol=2e-4)
    x = a.getOptimalVar()
    assert np.allclose(x, [2.0175, 1.7800, 1.2375], atol=2e-4)
    fx = f(x)
    assert fx[1] <= 1e-5 and fx[2] <= 1e-5


def test_separable_volume_problem_kkt(pyOptimizer):
    from oracle.mma import MMA as OMMA

    rng = np.random.default_rng(3)
    n, V = 2000, 0.3
    c = rng.uniform(0.5, 2.0, n)

    # objective scaled to O(1): MMA's elastic variables y cost c = 1000 (MethodOfMovingAsymptotes.hh:36), which
    # must exceed the constraint's multiplier for the constraint to be met exactly
    def f(x):
        return np.array([np.mean(c / x), np.mean(x) / V - 1])

    def df(x):
        return np.stack([-c / x ** 2 / n, np.full(n, 1.0 / (V * n))])

    lo, hi = np.full(n, 1e-3), np.ones(n)
    a = pyOptimizer.MMA(n, 1, lo, hi, f, df)
    b = OMMA(n, 1, lo, hi, f, df)
    x0 = np.full(n, V)
    a.setInitialVar(x0)
    b.setInitialVar(x0)
    for k in range(30):
        a.step()
        b.step()
        assert np.abs(a.getOptimalVar() - b.x()).max() <= 1e-7 * max(1.0, np.abs(b.x()).max()), k
    x = a.getOptimalVar()
    # KKT point: x_j = sqrt(c_j / lambda) with the volume constraint active (no bound is active here)
    s = np.sqrt(c)
    xs = s * (V * n / s.sum())
    assert np.abs(np.mean(x) - V) <= 1e-4
    assert np.abs(x - xs).max() <= 1e-3 * xs.max()


def test_errors(pyOptimizer):
    f, df = toy()
    o = pyOptimizer.MMA(3, 2, np.zeros(3), np.full(3, 5.0), f, df)
    with pytest.raises(RuntimeError, match="initial value"):
        o.step()
    with pytest.raises(RuntimeError):
        o.setInitialVar(np.zeros(4))
