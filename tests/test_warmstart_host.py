"""Host logic of the sweep warm starts (no GPU): the minimal-residual
combination of candidate solutions and the neighbour lookup."""

import numpy as np

from sonium_amd.warmstart import minimal_residual_combination, nearest_solved


def test_minimal_residual_combination_recovers_a_solution_in_the_span():
    rng = np.random.default_rng(5)
    n = 40
    A = rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n)) + 8 * np.eye(n)
    x1 = rng.standard_normal(n) + 1j * rng.standard_normal(n)
    x2 = rng.standard_normal(n) + 1j * rng.standard_normal(n)
    x = (0.3 - 0.7j) * x1 + (1.1 + 0.2j) * x2
    b = A @ x
    x0, rel = minimal_residual_combination(b, [x1, x2], [A @ x1, A @ x2])
    assert rel < 1e-13
    np.testing.assert_allclose(x0, x, rtol=0, atol=1e-12 * np.abs(x).max())
    # outside the span: the residual is the least-squares one, below the best single candidate's
    x3 = rng.standard_normal(n) + 0j
    b3 = A @ x3
    x0, rel = minimal_residual_combination(b3, [x1, x2], [A @ x1, A @ x2])
    P = np.stack([A @ x1, A @ x2], 1)
    c = np.linalg.lstsq(P, b3, rcond=None)[0]
    assert abs(rel - np.linalg.norm(b3 - P @ c) / np.linalg.norm(b3)) < 1e-12
    for xc in (x1, x2):
        Ax = A @ xc
        best = np.linalg.norm(b3 - Ax * (np.vdot(Ax, b3) / np.vdot(Ax, Ax))) / np.linalg.norm(b3)
        assert rel <= best + 1e-12
    # an all-zero candidate takes no weight
    z = np.zeros(n, complex)
    x0, rel = minimal_residual_combination(b, [x1, z, x2], [A @ x1, z, A @ x2])
    assert rel < 1e-13 and np.all(np.isfinite(x0))


def test_nearest_solved():
    solved = {0: None, 4: None, 8: None}
    assert nearest_solved(solved, 2) == (0, 4)
    assert nearest_solved(solved, 9) == (8, None)
    assert nearest_solved({}, 3) == (None, None)
    assert nearest_solved({5: None}, 1) == (None, 5)


def test_sweep_group_order_and_neighbours():
    from sonium_amd.sweep import group_neighbours, group_order

    assert group_order(1) == [0]
    assert group_order(4) == [0, 2, 1, 3]
    assert group_order(8) == [0, 4, 2, 6, 1, 3, 5, 7]
    for stride in (2, 4, 8, 16):
        order = group_order(stride)
        assert sorted(order) == list(range(stride))
        nb = group_neighbours(stride)
        assert nb[0] is None
        for pos, g in enumerate(order[1:], start=1):
            lo, hi = nb[g]
            # both bracketing groups come earlier in the order and are the nearest ones
            assert lo in order[:pos] and (hi == stride or hi in order[:pos])
            assert lo < g < hi
            assert not any(lo < e < hi for e in order[:pos])
    try:
        group_order(6)
    except ValueError:
        pass
    else:
        raise AssertionError("a stride that is not a power of two must be refused")


def test_incident_phases_recurrence_and_direct():
    from sonium_amd.warmstart import incident_phases

    rng = np.random.default_rng(2)
    c = rng.standard_normal((300, 3))
    d = np.array([0.0, 0.6, 0.8])
    for ks in (np.linspace(1.0, 12.0, 64), np.array([1.0, 2.5, 2.7]), np.array([3.0])):
        ref = np.exp(1j * ks[:, None] * (c @ d)[None, :])
        assert np.abs(incident_phases(c, ks, d) - ref).max() < 1e-13
