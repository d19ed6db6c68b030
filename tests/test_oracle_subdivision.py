"""Pins of generate_subelements (singular.rs:1372-1528) in the oracle, and the
product's host route builder against it.

What the algorithm fixes, checked independently of the restatement:
* a pair whose collocation point is 3 sqrt(area) or more from the panel's
  centre is one unsubdivided TR13 leaf (factor 1);
* every kept leaf is at distance >= 3 sqrt(leaf area) (ratio test), its area
  in local coordinates is 0.5 * 4^-depth, its Gauss order is 4;
* without truncation the leaves tile the panel exactly (areas sum to 0.5 and
  no two leaves overlap);
* the caps hold: at most 110 leaves, at most 15 splits per level.
"""

import numpy as np
import pytest

import oracle
import sonium_amd as sa
from sonium_amd.mesh import icosphere


def _global(coords, s, t):
    c = np.asarray(coords).reshape(3, 3)
    return (1.0 - s - t) * c[0] + s * c[1] + t * c[2]


def _leaf_area(leaf):
    (x0, y0, x1, y1, x2, y2) = leaf[:6]
    return 0.5 * abs((x1 - x0) * (y2 - y0) - (x2 - x0) * (y1 - y0))


@pytest.fixture(scope="module")
def sphere():
    m = icosphere(1.0, 3)
    c, n, a = m.element_geometry()
    return m, c, a, m.nodes[m.triangles].reshape(-1, 9)


def _near_pairs(c, limit, rng):
    out = []
    for i in rng.choice(len(c), 25, replace=False):
        d = np.linalg.norm(c - c[i], axis=1)
        for j in np.argsort(d)[1:limit]:
            out.append((int(i), int(j)))
    return out


def test_far_pair_is_one_full_leaf(sphere):
    m, c, a, coords = sphere
    j = int(np.argmax(np.linalg.norm(c - c[0], axis=1)))
    leaves = oracle.generate_subelements(c[0], coords[j], a[j])
    assert leaves.shape[0] == 1
    assert np.array_equal(leaves[0], [0.0, 0.0, 1.0, 0.0, 0.0, 1.0, 1.0, 4.0])


def test_leaves_satisfy_ratio_and_tile_the_panel(sphere):
    m, c, a, coords = sphere
    rng = np.random.default_rng(1)
    checked = 0
    for i, j in _near_pairs(c, 20, rng):
        leaves = oracle.generate_subelements(c[i], coords[j], a[j])
        assert 1 <= leaves.shape[0] <= 110
        total = 0.0
        for leaf in leaves:
            la = _leaf_area(leaf)
            depth = round(np.log(0.5 / la) / np.log(4.0))
            assert abs(la - 0.5 * 4.0 ** -depth) < 1e-15
            assert abs(leaf[6] - 0.5 ** depth) < 1e-15
            assert leaf[7] == 4.0
            cs = (leaf[0] + leaf[2] + leaf[4]) / 3.0
            ct = (leaf[1] + leaf[3] + leaf[5]) / 3.0
            p = _global(coords[j], cs, ct)
            ratio = np.linalg.norm(p - c[i]) / np.sqrt(a[j] * 4.0 ** -depth)
            assert ratio >= 3.0 - 1e-12
            total += la
        if leaves.shape[0] < 110:
            # untruncated tiling of the reference triangle (area 1/2)
            assert abs(total - 0.5) < 1e-14 or total < 0.5
            if abs(total - 0.5) < 1e-14:
                checked += 1
    assert checked > 100


def test_truncation_caps(sphere):
    """A collocation point very close to a large panel hits the 110-leaf / ndie > 15 caps."""
    coords = np.array([0.0, 0.0, 0.0, 1.0, 0.0, 0.0, 0.0, 1.0, 0.0])
    area = 0.5
    src = np.array([0.3, 0.3, 1e-4])
    leaves = oracle.generate_subelements(src, coords, area)
    assert leaves.shape[0] <= 110
    total = sum(_leaf_area(l) for l in leaves)
    assert total < 0.5  # truncated: part of the panel is dropped, as in the reference


@pytest.mark.parametrize("src", [[0.3, 0.3, 1e-4], [0.05, 0.05, 1e-6], [0.5, 0.25, 3e-5], [0.9, 0.05, 0.01]])
def test_truncation_keeps_level_order_and_the_per_level_cap(src):
    """generate_subelements walks the tree level by level (singular.rs:1408-1540):
    a level keeps its cells in order, subdivides at most 15 of them (ndie > 15
    stops the level: the cells after the 16th that needs subdividing are
    dropped) and stores the rest as leaves; the result stops at 110 leaves.  No
    reference test covers the truncation, so these invariants pin it against
    plausible mistakes: a depth-first walk breaks the level order, a missing
    cap puts more than 4 * 15 = 60 cells on a level."""
    coords = np.array([0.0, 0.0, 0.0, 1.0, 0.0, 0.0, 0.0, 1.0, 0.0])
    leaves = oracle.generate_subelements(np.array(src, dtype=float), coords, 0.5)
    depth = np.array([int(round(np.log(0.5 / _leaf_area(l)) / np.log(4.0))) for l in leaves])
    assert np.all(np.diff(depth) >= 0)                      # level order
    counts = np.bincount(depth)
    assert counts.max() <= 60                               # at most 15 cells subdivided per level
    assert leaves.shape[0] <= 110
    total = sum(_leaf_area(l) for l in leaves)
    assert total <= 0.5 + 1e-15


def test_product_route_builder_matches_oracle(sphere):
    """The library's host route builder (no GPU needed) reproduces the oracle's leaves bit for bit."""
    m, c, a, coords = sphere
    rng = np.random.default_rng(5)
    for i, j in _near_pairs(c, 30, rng):
        lo = oracle.generate_subelements(c[i], coords[j], a[j])
        lp = sa.debug_subelements(c[i], coords[j], a[j])
        assert lo.shape[0] == lp.shape[0]
        assert np.array_equal(lo[:, :6], lp)
    for src in (np.array([0.3, 0.3, 1e-4]), np.array([0.9, 0.05, 0.01]), np.array([-0.2, 0.5, 0.02])):
        co = np.array([0.0, 0.0, 0.0, 1.0, 0.0, 0.0, 0.0, 1.0, 0.0])
        lo = oracle.generate_subelements(src, co, 0.5)
        lp = sa.debug_subelements(src, co, 0.5)
        assert np.array_equal(lo[:, :6], lp)
