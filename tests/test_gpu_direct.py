"""GPU direct sums against the CPU oracle (flat_tree.py:125-160 semantics)."""

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def relative(a, b):
    scale = np.max(np.abs(b)) if np.size(b) else 0.0
    diff = np.max(np.abs(a - b)) if np.size(b) else 0.0
    return diff / scale if scale > 0 else diff


def cloud(n, seed):
    rng = np.random.default_rng(seed)
    pos = np.concatenate(
        (rng.normal(0.2, 0.05, (2, n // 2)), rng.uniform(-0.99, 0.99, (2, n - n // 2))), axis=1
    ).clip(-0.99, 0.99)
    return pos, rng.normal(0.3, 1.0, n)


@pytest.mark.parametrize("kernel", ["coulomb", "log"])
@pytest.mark.parametrize("n,m", [(1, 1), (257, 3), (5000, 5000), (20000, 700)])
def test_direct_fp64_matches_oracle(td3, oracle, kernel, n, m):
    pos, q = cloud(n, 1)
    q[::5] = 0.0
    rng = np.random.default_rng(2)
    targets = rng.uniform(-1, 1, (2, m))
    exclude = rng.integers(-1, n, m)
    print(f"direct fp64 kernel={kernel} n={n} m={m}", flush=True)
    got = td3.direct_field(targets, pos, q, softening=1e-3, exclude=exclude, kernel=kernel)
    ref = oracle.direct_field(targets, pos, q, softening=1e-3, exclude=exclude, kernel=kernel)
    for a, b in zip(got, ref):
        assert relative(a, b) < 1e-12


@pytest.mark.parametrize("kernel", ["coulomb", "log"])
def test_direct_mixed_is_close(td3, oracle, kernel):
    pos, q = cloud(30000, 3)
    targets = pos[:, :4096]
    exclude = np.arange(4096)
    got = td3.direct_field(targets, pos, q, softening=1e-4, exclude=exclude, kernel=kernel,
                           precision="mixed")
    ref = oracle.direct_field(targets, pos, q, softening=1e-4, exclude=exclude, kernel=kernel)
    for a, b in zip(got, ref):
        assert relative(a, b) < 1e-5


def test_direct_coincident_and_empty(td3):
    pos = np.array([[0.1, 0.1, 0.5], [0.2, 0.2, -0.3]])
    q = np.array([1.0, 2.0, -1.0])
    with pytest.raises(ValueError, match="Coincident"):
        td3.direct_field(pos, pos, q, exclude=[0, -1, 2])
    # an excluded or uncharged coincident source is fine
    q0 = np.array([1.0, 0.0, -1.0])
    phi, field = td3.direct_field(pos, pos, q0, exclude=[0, 0, 2], kernel="log")
    assert np.all(np.isfinite(phi)) and np.all(np.isfinite(field))
    phi, field = td3.direct_field(np.zeros((2, 0)), pos, q)
    assert phi.shape == (0,) and field.shape == (2, 0)
    phi, field = td3.direct_field(pos, np.zeros((2, 0)), np.zeros(0))
    assert not np.any(phi) and field.shape == (2, 3)
    with pytest.raises(ValueError, match=r"\(2, N\)"):
        td3.direct_field(np.zeros((5, 2)), pos, q)


def test_direct_known_answer_line_charge(td3):
    # test_log_tree.py:31-38: one line charge q=2 at the origin, target (3, 4)
    phi, field = td3.direct_field([[3.0], [4.0]], [[0.0], [0.0]], [2.0], kernel="log")
    assert phi[0] == pytest.approx(-2.0 * np.log(25.0), rel=1e-15)
    np.testing.assert_allclose(field[:, 0], 2 * 2.0 * np.array([3.0, 4.0]) / 25, rtol=1e-15)
