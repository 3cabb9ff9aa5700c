"""Pins the oracle's direct sums (oracle_direct, restating direct_sum_pyccel.py:15-72 and
direct_log_sum_pyccel.py:15-70) against closed forms evaluated by SymPy.

The formulas are the reference's method-card rules (flat_tree.py:78-107): one source and
one target, potential q/sqrt(r^2+eps^2) and -q ln(r^2+eps^2), fields minus their
gradients, differentiated symbolically here (not restated), over the rule's ranges.
"""

import numpy as np
import pytest
import sympy


def rules():
    x, y, x1, y1, q1, eps = sympy.symbols("x y x1 y1 q1 eps", real=True)
    r2 = (x - x1) ** 2 + (y - y1) ** 2 + eps**2
    coulomb = q1 / sympy.sqrt(r2)
    log = -q1 * sympy.log(r2)
    out = {}
    for name, pot in (("coulomb", coulomb), ("log", log)):
        ex = -sympy.diff(pot, x)
        ey = -sympy.diff(pot, y)
        out[name] = [sympy.lambdify((x, y, x1, y1, q1, eps), e, "mpmath") for e in (pot, ex, ey)]
    return out


RULES = rules()


@pytest.mark.parametrize("kernel", ["coulomb", "log"])
def test_direct_matches_sympy_closed_form(oracle, kernel):
    rng = np.random.default_rng(0)
    for _ in range(40):
        x, y = rng.uniform(1.0, 2.0), rng.uniform(-1.0, 1.0)
        x1, y1 = rng.uniform(-1.0, 0.5), rng.uniform(-1.0, 1.0)
        q1, eps = rng.uniform(-2.0, 2.0), rng.uniform(0.0, 0.5)
        phi, field = oracle.direct_field([[x], [y]], [[x1], [y1]], [q1], softening=eps,
                                         kernel=kernel)
        want = [float(f(x, y, x1, y1, q1, eps)) for f in RULES[kernel]]
        np.testing.assert_allclose([phi[0], field[0, 0], field[1, 0]], want, rtol=1e-13,
                                   atol=1e-15)


def test_direct_is_a_sum_over_sources(oracle):
    # superposition: the sum over N sources equals the sum of N one-source sums, with the
    # excluded source and zero charges dropping out
    rng = np.random.default_rng(1)
    pos = rng.uniform(-1, 1, (2, 7))
    q = rng.normal(size=7)
    q[3] = 0.0
    targets = rng.uniform(-1, 1, (2, 5))
    exclude = np.array([2, -1, 0, 6, -1])
    for kernel in ("coulomb", "log"):
        phi, field = oracle.direct_field(targets, pos, q, softening=0.1, exclude=exclude,
                                         kernel=kernel)
        for i in range(5):
            p_sum, f_sum = 0.0, np.zeros(2)
            for j in range(7):
                if j == exclude[i]:
                    continue
                f = [float(g(*targets[:, i], *pos[:, j], q[j], 0.1)) for g in RULES[kernel]]
                p_sum += f[0]
                f_sum += f[1:]
            assert phi[i] == pytest.approx(p_sum, rel=1e-12, abs=1e-14)
            np.testing.assert_allclose(field[:, i], f_sum, rtol=1e-12, atol=1e-14)


def test_direct_known_answer_and_failures(oracle):
    # test_log_tree.py:31-38: phi = -2 * 2 ln 5 at (3, 4) from q = 2 at the origin
    phi, field = oracle.direct_field([[3.0], [4.0]], [[0.0], [0.0]], [2.0], kernel="log")
    assert phi[0] == pytest.approx(-2.0 * np.log(25.0), rel=1e-15)
    np.testing.assert_allclose(field[:, 0], [0.48, 0.64], rtol=1e-15)
    with pytest.raises(ValueError, match="Coincident"):
        oracle.direct_field([[0.5], [0.5]], [[0.5], [0.5]], [1.0])
    phi, _ = oracle.direct_field([[0.5], [0.5]], [[0.5], [0.5]], [0.0])
    assert phi[0] == 0.0
    phi, field = oracle.direct_field(np.zeros((2, 0)), [[0.0], [0.0]], [1.0])
    assert phi.shape == (0,) and field.shape == (2, 0)
