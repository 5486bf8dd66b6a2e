"""GPU grid FRNN (GridCount2d) against the oracle.

- The reference's lattice tests (tests/grid2d_frnn.cu:62-203, 280-340) run
  through update() + the FRNN query: counts equal the closed forms.
- Its random test (mt19937(10), uniform(-50, 50), 10,000 points): neighbour
  SETS keyed by original id equal the oracle's grid search and brute force
  (the reference checks counts to +-1; the FP32 predicate here is the same
  on both sides, so the sets are exact).
- Config-1 shape (N = 10,000 in [0,1), r = 0.02) with row-major and Morton
  keys, and the torus grid (periodic cells, minimum-image distances).
- update() permutes the caller's positions (and velocities) by cell, stable
  within a cell; d_ind maps old index -> new slot.
- mean_num_particle_in_cell.
"""
import numpy as np
import pytest

import mcleap_amd as M

pytestmark = pytest.mark.gpu


def _frnn_sets(g, pos_dev, rad):
    """neighbour lists per sorted slot, mapped to original ids"""
    import torch
    n = g.numP
    off = torch.zeros(n + 1, dtype=torch.int32, device="cuda")
    tot = g.frnn_csr(pos_dev, rad, off)
    nbr = torch.zeros(max(tot, 1), dtype=torch.int32, device="cuda")
    assert g.frnn_csr(pos_dev, rad, off, nbr) == tot
    off, nbr = off.cpu().numpy(), nbr.cpu().numpy()[:tot]
    perm = g.perm().cpu().numpy()
    out = {}
    for s in range(n):
        out[int(perm[s])] = sorted(int(perm[j]) for j in nbr[off[s]:off[s + 1]])
    return out


def _oracle_sets(counts, nbr):
    return {i: sorted(nbr[i, :counts[i]].tolist()) for i in range(len(counts))}


def _run_grid(xy, minxy, cs, nc, rad, flags):
    import torch
    g = M.GridCount2d(len(xy), minxy, cs, nc, flags)
    pos = torch.from_numpy(xy.copy()).cuda()
    g.update(pos)
    return g, pos


@pytest.mark.parametrize("side,step,rad,want_max", [(10, 1.0, 1.0, 4), (20, 0.5, 1.0, 12), (40, 0.25, 0.25, 4)])
def test_reference_lattices(oracle, side, step, rad, want_max):
    gi = np.stack(np.meshgrid(np.arange(side), np.arange(side), indexing="ij"), -1).reshape(-1, 2)
    xy = (gi * step).astype(np.float32)
    g, pos = _run_grid(xy, (0, 0), (1, 1), (10, 10), rad, M.GRID_ROWMAJOR)
    sets = _frnn_sets(g, pos, rad)
    counts, nbr = oracle.grid_frnn(xy, (0, 0), (1, 1), (10, 10), rad)
    assert max(len(v) for v in sets.values()) == want_max
    assert sets == _oracle_sets(counts, nbr)


def test_reference_random_seed10(oracle):
    O = oracle
    xy = O.mt_uniform_doubles(10, -50.0, 50.0, 20000).reshape(-1, 2).astype(np.float32)
    g, pos = _run_grid(xy, (-50, -50), (1, 1), (100, 100), 1.0, M.GRID_ROWMAJOR)
    sets = _frnn_sets(g, pos, 1.0)
    counts, nbr = O.grid_frnn(xy, (-50, -50), (1, 1), (100, 100), 1.0, cap=64)
    assert sets == _oracle_sets(counts, nbr)
    bc, _ = O.brute_frnn(xy, 1.0, cap=64)
    assert np.array_equal(counts, bc)


@pytest.mark.parametrize("flags", [M.GRID_ROWMAJOR, M.GRID_MORTON])
def test_config1_box(oracle, flags):
    rng = np.random.default_rng(1)
    xy = rng.random((10000, 2)).astype(np.float32)
    g, pos = _run_grid(xy, (0, 0), (0.02, 0.02), (50, 50), 0.02, flags)
    sets = _frnn_sets(g, pos, 0.02)
    counts, nbr = oracle.grid_frnn(xy, (0, 0), (0.02, 0.02), (50, 50), 0.02, cap=128)
    assert sets == _oracle_sets(counts, nbr)
    assert abs(np.mean(counts) - np.pi * 0.02 ** 2 * 10000) < 1.0


def test_torus_grid(oracle):
    rng = np.random.default_rng(2)
    n = 1 << 14
    xy = rng.random((n, 2)).astype(np.float32)
    r = 2.0 / np.sqrt(n)
    nc = int(np.floor(1.0 / r))
    cs = 1.0 / nc
    g, pos = _run_grid(xy, (0, 0), (cs, cs), (nc, nc), r, M.GRID_MORTON | M.GRID_PERIODIC)
    sets = _frnn_sets(g, pos, r)
    counts, nbr = oracle.grid_frnn(xy, (0, 0), (cs, cs), (nc, nc), r, periodic=True, cap=128)
    assert sets == _oracle_sets(counts, nbr)
    bc, _ = oracle.brute_frnn(xy, r, periodic=True, cap=128)
    assert np.array_equal(counts, bc)


@pytest.mark.parametrize("flags", [M.GRID_MORTON | M.GRID_PERIODIC, M.GRID_MORTON])
def test_async_csr_equals_counted(oracle, flags):
    """The asynchronous CSR form (no host read-back) gives the same offsets
    and neighbour lists as the counted form, also with a dense cluster of
    points that have hundreds of neighbours each."""
    import torch
    rng = np.random.default_rng(8)
    n0 = 20000
    r = 2.0 / np.sqrt(n0)
    xy = np.concatenate([rng.random((n0, 2)), 0.5 + 0.3 * r * (rng.random((300, 2)) - 0.5)]).astype(np.float32)
    xy[xy >= 1.0] = 0.0
    nc = int(np.floor(1.0 / r))
    cs = 1.0 / nc
    g, pos = _run_grid(xy, (0, 0), (cs, cs), (nc, nc), r, flags)
    n = g.numP
    off = torch.zeros(n + 1, dtype=torch.int32, device="cuda")
    tot = g.frnn_csr(pos, r, off)
    nbr = torch.zeros(tot, dtype=torch.int32, device="cuda")
    assert g.frnn_csr(pos, r, off, nbr) == tot
    assert (off[1:] - off[:-1]).max().item() > 200
    off2 = torch.zeros(n + 1, dtype=torch.int32, device="cuda")
    nbr2 = torch.full((tot,), -1, dtype=torch.int32, device="cuda")
    g.frnn_csr_async(pos, r, off2, nbr2)
    torch.cuda.synchronize()
    assert torch.equal(off, off2) and torch.equal(nbr, nbr2)


def test_update_permutes_in_place(oracle):
    import torch
    rng = np.random.default_rng(3)
    xy = rng.random((5000, 2)).astype(np.float32)
    vel = rng.standard_normal((5000, 2)).astype(np.float32)
    g = M.GridCount2d(len(xy), (0, 0), (0.05, 0.05), (20, 20), M.GRID_ROWMAJOR)
    pos = torch.from_numpy(xy.copy()).cuda()
    v = torch.from_numpy(vel.copy()).cuda()
    g.update(pos, v)
    h, perm_o, start = oracle.grid_sort(xy, (0, 0), (0.05, 0.05), (20, 20))
    perm = g.perm().cpu().numpy()
    assert np.array_equal(perm, perm_o)  # stable counting sort: same order as the oracle
    assert np.array_equal(pos.cpu().numpy(), xy[perm])
    assert np.array_equal(v.cpu().numpy(), vel[perm])
    ind = g.d_ind().cpu().numpy()
    assert np.array_equal(ind[perm], np.arange(len(xy)))
    occ = g.mean_num_particle_in_cell()
    nz = np.diff(start)
    # gridCount2d.h:313-334: particles / non-empty cells
    assert occ == np.float32(len(xy)) / np.float32((nz > 0).sum())


def test_build_leaves_source_untouched(oracle):
    import torch
    rng = np.random.default_rng(4)
    xy = rng.random((8000, 2)).astype(np.float32)
    g = M.GridCount2d(len(xy), (0, 0), (0.025, 0.025), (40, 40), M.GRID_MORTON)
    src = torch.from_numpy(xy.copy()).cuda()
    g.build(src)
    torch.cuda.synchronize()
    assert np.array_equal(src.cpu().numpy(), xy)
    sets = _frnn_sets(g, M.triangulation._tensor_from_ptr(g.sorted_ptr(), (len(xy), 2), torch.float32), 0.025)
    counts, nbr = oracle.grid_frnn(xy, (0, 0), (0.025, 0.025), (40, 40), 0.025, cap=128)
    assert sets == _oracle_sets(counts, nbr)


def test_apply_f_frnn_includes_self_pairs(oracle):
    import torch
    rng = np.random.default_rng(5)
    xy = rng.random((2000, 2)).astype(np.float32)
    g, pos = _run_grid(xy, (0, 0), (0.05, 0.05), (20, 20), 0.05, M.GRID_ROWMAJOR)
    seen = {}

    def f(i, j, d, dist):
        seen["i"], seen["j"], seen["dist"] = i.cpu().numpy(), j.cpu().numpy(), dist.cpu().numpy()

    g.apply_f_frnn(f, pos, 0.05)
    assert np.sum(seen["i"] == seen["j"]) == len(xy)
    assert np.all(seen["dist"] <= 0.05 + 1e-7)


def _brute_sets(xy, rad):
    """pairs with dx*dx + dy*dy <= rad*rad in FP32, separately rounded"""
    xy = np.asarray(xy, np.float32)
    r2 = np.float32(rad) * np.float32(rad)
    out = {}
    for i in range(len(xy)):
        d = xy - xy[i]
        q = d[:, 0] * d[:, 0] + d[:, 1] * d[:, 1]
        js = np.nonzero(q <= r2)[0]
        out[i] = sorted(int(j) for j in js if j != i)
    return out


def test_batch_grid_frnn(oracle):
    """GridCount2dBatch: every instance's neighbour sets (its own radius
    2/sqrt(n_i), its own cells) equal brute force on that instance alone;
    no pair crosses instances (all instances share [0,1)^2)."""
    import torch

    from mcleap_amd.workloads import ensemble_points
    sizes = [1, 2, 40, 300, 1000, 2500, 700, 5000, 64, 3]
    pts = ensemble_points(seed=21, sizes=sizes)
    B = M.HostTriangulationBatch(pts)
    rad = np.array([2.0 / np.sqrt(max(len(p), 1)) for p in pts], np.float32)
    rad = np.minimum(rad, np.float32(1.0 / 3.0))
    g = M.GridCount2dBatch(B.vstart, rad)
    assert g.numP == len(B.m_pos) - 4 * B.ninst and g.first == 4 * B.ninst
    pos = torch.from_numpy(B.m_pos).cuda()
    g.build(pos)
    n = g.numP
    off = torch.zeros(n + 1, dtype=torch.int32, device="cuda")
    tot = g.frnn_csr(off)
    nbr = torch.zeros(max(tot, 1), dtype=torch.int32, device="cuda")
    assert g.frnn_csr(off, nbr) == tot
    off, nbr = off.cpu().numpy(), nbr.cpu().numpy()[:tot]
    perm = g.perm().cpu().numpy()
    spos = g.sorted_pos().cpu().numpy()
    assert np.array_equal(spos, B.m_pos[g.first + perm])
    got = {}
    for s in range(n):
        got[int(perm[s]) + g.first] = sorted(int(perm[j]) + g.first for j in nbr[off[s]:off[s + 1]])
    for i in range(B.ninst):
        vs = np.arange(B.vstart[i], B.vstart[i + 1])
        want = _brute_sets(B.m_pos[vs], rad[i])
        for k, v in enumerate(vs):
            assert got[int(v)] == [int(vs[j]) for j in want[k]], (i, k)
    # the async form gives the same CSR
    off2 = torch.zeros(n + 1, dtype=torch.int32, device="cuda")
    nbr2 = torch.zeros(max(tot, 1), dtype=torch.int32, device="cuda")
    g.frnn_csr_async(off2, nbr2)
    torch.cuda.synchronize()
    assert np.array_equal(off2.cpu().numpy(), off) and np.array_equal(nbr2.cpu().numpy()[:tot], nbr)


def _pairs_gpu(g, pos_dev, rad):
    """all (i, j) pairs of the CSR query as original ids, sorted, int64 keys"""
    import torch
    n = g.numP
    off = torch.zeros(n + 1, dtype=torch.int32, device="cuda")
    tot = g.frnn_csr(pos_dev, rad, off)
    nbr = torch.zeros(max(tot, 1), dtype=torch.int32, device="cuda")
    assert g.frnn_csr(pos_dev, rad, off, nbr) == tot
    off, nbr = off.cpu().numpy(), nbr.cpu().numpy()[:tot].astype(np.int64)
    perm = g.perm().cpu().numpy().astype(np.int64)
    src = np.repeat(perm, np.diff(off))
    return np.sort(src * n + perm[nbr])


def _pairs_oracle(counts, nbr):
    n = len(counts)
    src = np.repeat(np.arange(n, dtype=np.int64), counts)
    mask = np.arange(nbr.shape[1])[None, :] < counts[:, None]
    return np.sort(src * n + nbr[mask].astype(np.int64))


def test_c1_frnn_after_every_step(oracle):
    """C1 (10,000 points, box, sigma = 1e-3 clamped, r = 0.02, the reference's
    row-major cell keys): after each of the 50 kinetic steps the grid FRNN
    pairs of the moved points equal the oracle's grid search; every 10 steps
    they are also checked against brute force"""
    import torch
    from tests.conftest import uniform_points
    O = oracle
    H = M.HostTriangulation()
    H.addDelaunayPoints(uniform_points(10000, 1))
    dt = M.DeviceTriangulation(H)
    n = dt.m_pos_size - 4
    g = M.GridCount2d(n, (0.0, 0.0), (0.02, 0.02), (50, 50), M.GRID_ROWMAJOR)
    for k in range(50):
        dt.step(M.MOTION_GAUSS, 1e-3, seed=1, step=k)
        pts = dt.download_positions()[4:].copy()
        pos = torch.from_numpy(pts.copy()).cuda()  # update() sorts a copy (moving_device_mesh.cu:125-130)
        g.update(pos)
        got = _pairs_gpu(g, pos, 0.02)
        counts, nbr = O.grid_frnn(pts, (0, 0), (0.02, 0.02), (50, 50), 0.02, cap=128)
        assert counts.max() < 128
        assert np.array_equal(got, _pairs_oracle(counts, nbr)), "step %d" % k
        if k % 10 == 9:
            # brute force: the reference's cell formula (gridCount2d.h:94-102,
            # FP32 divide, truncate, % nx) misses two kinds of pairs, both
            # reproduced here: a point clamped to 1 - 2^-24 divides to exactly
            # 50.0 in FP32 and wraps into cell 0, and with cells exactly r
            # wide a pair at distance ~r can fall two cells apart
            bc, bn = O.brute_frnn(pts, 0.02, cap=128)
            brute = _pairs_oracle(bc, bn)
            assert np.all(np.isin(got, brute))
            miss = np.setdiff1d(brute, got)
            i, j = miss // n, miss % n
            d = pts[j] - pts[i]
            r2 = d[:, 0] * d[:, 0] + d[:, 1] * d[:, 1]
            wraps = np.any(pts / np.float32(0.02) >= np.float32(50.0), axis=1)
            assert np.all(wraps[i] | wraps[j] | (r2 > np.float32(0.0199) ** 2)), "step %d" % k
            assert np.all(np.isin(miss, brute))


def test_c2_torus_frnn_1m(oracle):
    """C2 at full size: 2^20 points on the torus, r = 1/512 on the 512 x 512
    periodic grid (Morton keys): the ~13M pairs equal the oracle's periodic
    grid search"""
    import torch
    from mcleap_amd.workloads import grid_shape, initial_points
    O = oracle
    pts = initial_points("c2")
    n = len(pts)
    r = np.float32(2.0 / np.sqrt(n))
    nc, cs = grid_shape(float(r), "torus")
    g = M.GridCount2d(n, (0.0, 0.0), (cs, cs), (nc, nc), M.GRID_MORTON | M.GRID_PERIODIC)
    pos = torch.from_numpy(pts.copy()).cuda()
    g.update(pos)
    got = _pairs_gpu(g, pos, float(r))
    counts, nbr = O.grid_frnn(pts, (0, 0), (cs, cs), (nc, nc), float(r), periodic=True, cap=64)
    assert counts.max() < 64
    want = _pairs_oracle(counts, nbr)
    print("pairs", len(got), "mean", len(got) / n)
    assert len(got) > 12 * n and np.array_equal(got, want)


def test_c3_density_frnn(oracle):
    """C3's clustered density (64 Gaussian blobs, 4.2M points, r = 3/sqrt(N)):
    the points of a window around the densest blob centre (over 80 neighbours
    per point on average, up to ~550), queried on the C3 grid geometry, give the same pairs as the
    oracle's grid search and brute force"""
    import torch
    from mcleap_amd.workloads import CONFIGS, grid_shape, initial_points
    O = oracle
    allp = initial_points("c3")
    r = np.float32(3.0 / np.sqrt(CONFIGS["c3"]["n"]))
    nc, cs = grid_shape(float(r), "box")
    cell = np.floor(allp / np.float32(0.01)).astype(np.int64)
    key, cnt = np.unique(cell[:, 0] * 1000 + cell[:, 1], return_counts=True)
    cx, cy = divmod(int(key[np.argmax(cnt)]), 1000)
    c = np.array([(cx + 0.5) * 0.01, (cy + 0.5) * 0.01], np.float32)
    pts = allp[np.all(np.abs(allp - c) < 0.012, axis=1)]
    n = len(pts)
    g = M.GridCount2d(n, (0.0, 0.0), (cs, cs), (nc, nc), M.GRID_MORTON)
    pos = torch.from_numpy(pts.copy()).cuda()
    g.update(pos)
    got = _pairs_gpu(g, pos, float(r))
    counts, nbr = O.grid_frnn(pts, (0, 0), (cs, cs), (nc, nc), float(r), cap=1536)
    assert counts.max() < 1536
    print("window points", n, "mean neighbours", len(got) / n)
    assert len(got) / n > 80
    assert np.array_equal(got, _pairs_oracle(counts, nbr))
    bc, bn = O.brute_frnn(pts, float(r), cap=1536)
    assert np.array_equal(got, _pairs_oracle(bc, bn))


@pytest.mark.parametrize("periodic", [False, True])
@pytest.mark.parametrize("block", [64, 256])
def test_half_radius_cells_equal_brute_force(oracle, periodic, block):
    """cells of half the radius (the tuner's frnn_cells_per_radius = 2: a
    5 x 5 search) and any block size give exactly the brute-force pairs"""
    import torch
    from mcleap_amd.workloads import grid_shape
    O = oracle
    rng = np.random.default_rng(23)
    xy = rng.random((20000, 2)).astype(np.float32)
    r = 0.015
    nc, cs = grid_shape(r / 2, "torus" if periodic else "box")
    g = M.GridCount2d(len(xy), (0.0, 0.0), (cs, cs), (nc, nc), M.GRID_MORTON | (M.GRID_PERIODIC if periodic else 0))
    g.set_block(block)
    assert g.block() == block
    pos = torch.from_numpy(xy.copy()).cuda()
    g.update(pos)
    got = _pairs_gpu(g, pos, r)
    bc, bn = O.brute_frnn(xy, r, periodic=periodic, cap=128)
    assert np.array_equal(got, _pairs_oracle(bc, bn))
