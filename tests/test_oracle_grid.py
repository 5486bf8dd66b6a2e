"""Pins the oracle's grid FRNN and RNG restatements (CPU only).

- The closed-form counts of the reference's lattice tests
  (tests/grid2d_frnn.cu:62-139, 141-203, 280-340): unit lattice, r = 1 ->
  4 / 3 / 2 neighbours (interior / edge / corner); half-spaced lattice, r = 1
  -> 12 interior; quarter-spaced lattice, r = 0.25 -> 4 interior.
- The random test (tests/grid2d_frnn.cu:205-278: 10,000 points from
  mt19937(10), uniform_real_distribution<double>(-50, 50), cells of 1 from
  -50): grid search == O(n^2) brute force with the same FP32 predicate.
- The libstdc++ mt19937 / uniform_real_distribution draws the reference tests
  use, against fixtures written by g++ (tests/golden/mt_vectors.cpp) and the
  C++ standard's known answer (10000th draw of a default-seeded mt19937 is
  4123659995).
"""
import os

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
RNG = np.load(os.path.join(HERE, "golden", "rng.npz"))


def _lattice(side, step):
    g = np.stack(np.meshgrid(np.arange(side), np.arange(side), indexing="ij"), -1).reshape(-1, 2)
    return (g * step).astype(np.float32), g


def _closed_form(g, side, offsets):
    """neighbour count of each lattice node from the list of lattice offsets within r"""
    out = np.zeros(len(g), np.int32)
    for dx, dy in offsets:
        x, y = g[:, 0] + dx, g[:, 1] + dy
        out += ((x >= 0) & (x < side) & (y >= 0) & (y < side)).astype(np.int32)
    return out


def test_unit_lattice(oracle):
    xy, g = _lattice(10, 1.0)
    counts, _ = oracle.grid_frnn(xy, (0, 0), (1, 1), (10, 10), 1.0)
    want = _closed_form(g, 10, [(1, 0), (-1, 0), (0, 1), (0, -1)])
    assert np.array_equal(counts, want)
    assert sorted(set(want.tolist())) == [2, 3, 4]


def test_half_spaced_lattice(oracle):
    xy, g = _lattice(20, 0.5)
    counts, _ = oracle.grid_frnn(xy, (0, 0), (1, 1), (10, 10), 1.0)
    offs = [(dx, dy) for dx in range(-2, 3) for dy in range(-2, 3) if 0 < dx * dx + dy * dy <= 4]
    assert len(offs) == 12
    want = _closed_form(g, 20, offs)
    assert np.array_equal(counts, want) and want.max() == 12


def test_quarter_spaced_lattice(oracle):
    xy, g = _lattice(40, 0.25)
    counts, _ = oracle.grid_frnn(xy, (0, 0), (1, 1), (10, 10), 0.25)
    want = _closed_form(g, 40, [(1, 0), (-1, 0), (0, 1), (0, -1)])
    assert np.array_equal(counts, want) and want.max() == 4


def test_random_points_seed10(oracle):
    O = oracle
    d = O.mt_uniform_doubles(10, -50.0, 50.0, 20000)
    xy = d.reshape(-1, 2).astype(np.float32)  # glm::vec2(double, double)
    counts, nbr = O.grid_frnn(xy, (-50, -50), (1, 1), (100, 100), 1.0, cap=64)
    bc, bn = O.brute_frnn(xy, 1.0, cap=64)
    assert np.array_equal(counts, bc)
    for i in range(0, len(xy), 37):
        assert sorted(nbr[i, : counts[i]].tolist()) == sorted(bn[i, : bc[i]].tolist())


def test_periodic_grid_equals_brute_force(oracle):
    O = oracle
    rng = np.random.default_rng(7)
    xy = rng.random((3000, 2)).astype(np.float32)
    r = 0.03
    nc = int(1.0 / r)
    counts, nbr = O.grid_frnn(xy, (0, 0), (1.0 / nc, 1.0 / nc), (nc, nc), r, periodic=True, cap=64)
    bc, bn = O.brute_frnn(xy, r, periodic=True, cap=64)
    assert np.array_equal(counts, bc)
    # points near the seam see neighbours across it
    seam = (xy[:, 0] < 0.01) | (xy[:, 0] > 0.99)
    assert counts[seam].mean() > 0.8 * counts.mean()


def test_grid_sort_is_a_stable_counting_sort(oracle):
    xy = np.random.default_rng(8).random((5000, 2)).astype(np.float32)
    h, perm, start = oracle.grid_sort(xy, (0, 0), (0.05, 0.05), (20, 20))
    assert sorted(perm.tolist()) == list(range(len(xy)))
    hs = h[perm]
    assert np.all(np.diff(hs) >= 0)
    # within a cell, original order is kept
    for c in range(0, 400, 13):
        ids = perm[start[c]:start[c + 1]]
        assert np.all(np.diff(ids) > 0)
    # cell of a point: (int)((p - min) / cs) % n, key = y * nx + x
    cx = ((xy[:, 0] / np.float32(0.05)).astype(np.int32)) % 20
    cy = ((xy[:, 1] / np.float32(0.05)).astype(np.int32)) % 20
    assert np.array_equal(h, cy * 20 + cx)


def test_mt19937_against_libstdcxx(oracle):
    O = oracle
    assert np.array_equal(O.mt_raw(3, 2000), RNG["mt_raw_3"])
    assert np.array_equal(O.mt_raw(10, 2000), RNG["mt_raw_10"])
    assert int(O.mt_raw(5489, 10000)[-1]) == int(RNG["mt_default_10000th"][0]) == 4123659995


def test_uniform_real_distribution_against_libstdcxx(oracle):
    O = oracle
    assert np.array_equal(O.mt_uniform_floats(3, -1000.0, 1000.0, 2000), RNG["uniform_f32_seed3"])
    assert np.array_equal(O.mt_uniform_doubles(10, -50.0, 50.0, 4000), RNG["uniform_f64_seed10"])
