"""The allocation-free grid FRNN (mcl_grid_frnn_csr_noalloc), called through
the C ABI, against the handle path it replaces (mcl_grid_create +
mcl_grid_build + mcl_grid_frnn_csr): equal offsets, neighbour slots and slot
permutation, on two calls in a row with the second on a smaller input (the
shared static buffer keeps the first call's contents), in the synchronous and
the asynchronous form; sizes beyond the buffer are refused."""
import ctypes

import numpy as np
import pytest

import mcleap_amd as M
from mcleap_amd import _lib
from mcleap_amd.workloads import grid_shape

pytestmark = pytest.mark.gpu


def _ptr(t):
    return ctypes.c_void_p(t.data_ptr())


def _handle_path(L, torch, pos, nc, cs, flags, rad, cap):
    n = pos.shape[0]
    g = ctypes.c_void_p()
    _lib.check(L.mcl_grid_create(n, 0.0, 0.0, cs, cs, nc, nc, flags, ctypes.byref(g)), "mcl_grid_create")
    try:
        _lib.check(L.mcl_grid_build(g, _ptr(pos), None), "mcl_grid_build")
        off = torch.zeros(n + 1, dtype=torch.int32, device="cuda")
        nbr = torch.full((cap,), -1, dtype=torch.int32, device="cuda")
        tot = ctypes.c_longlong()
        _lib.check(L.mcl_grid_frnn_csr(g, L.mcl_grid_sorted_ptr(g), rad, _ptr(off), _ptr(nbr), cap,
                                       ctypes.byref(tot), None), "mcl_grid_frnn_csr")
        d_ind, d_perm = ctypes.c_void_p(), ctypes.c_void_p()
        _lib.check(L.mcl_grid_index_ptrs(g, ctypes.byref(d_ind), ctypes.byref(d_perm), None, None),
                   "mcl_grid_index_ptrs")
        perm = M.triangulation._tensor_from_ptr(d_perm.value, (n,), torch.int32).clone()
        torch.cuda.synchronize()
        return off.cpu().numpy(), nbr.cpu().numpy()[: tot.value], perm.cpu().numpy(), tot.value
    finally:
        L.mcl_grid_destroy(g)


def _noalloc(L, torch, pos, nc, cs, flags, rad, cap, sync):
    n = pos.shape[0]
    off = torch.full((n + 1,), -7, dtype=torch.int32, device="cuda")
    nbr = torch.full((cap,), -1, dtype=torch.int32, device="cuda")
    perm = torch.full((n,), -1, dtype=torch.int32, device="cuda")
    tot = ctypes.c_longlong(-1)
    rc = L.mcl_grid_frnn_csr_noalloc(_ptr(pos), n, 0.0, 0.0, cs, cs, nc, nc, flags, rad, _ptr(perm), _ptr(off),
                                     _ptr(nbr), cap, ctypes.byref(tot) if sync else None, None)
    _lib.check(rc, "mcl_grid_frnn_csr_noalloc")
    torch.cuda.synchronize()
    o = off.cpu().numpy()
    total = tot.value if sync else int(o[-1])
    return o, nbr.cpu().numpy()[:total], perm.cpu().numpy(), total


@pytest.mark.parametrize("sync", [True, False])
def test_noalloc_frnn_equals_handle_path(sync):
    import torch
    L = M.lib()
    rng = np.random.default_rng(17)
    cases = [
        # (points, periodic): first a C2-shaped torus at 400k points, then a smaller box
        (400_000, True),
        (30_000, False),
    ]
    for n, periodic in cases:
        xy = rng.random((n, 2)).astype(np.float32)
        if not periodic:  # a dense cluster, so that cells hold many points
            xy[: n // 4] = (0.5 + 0.03 * rng.standard_normal((n // 4, 2))).astype(np.float32)
            xy = np.clip(xy, 0.0, np.float32(1.0 - 2.0 ** -24))
        rad = np.float32(2.0 / np.sqrt(n))
        nc, cs = grid_shape(float(rad), "torus" if periodic else "box")
        flags = M.GRID_MORTON | (M.GRID_PERIODIC if periodic else 0)
        pos = torch.from_numpy(xy).cuda()
        cap = n * 256
        o1, b1, p1, t1 = _handle_path(L, torch, pos, nc, cs, flags, float(rad), cap)
        o2, b2, p2, t2 = _noalloc(L, torch, pos, nc, cs, flags, float(rad), cap, sync)
        print("n %d periodic %s pairs %d" % (n, periodic, t1), flush=True)
        assert t1 == t2 and t1 > 0
        assert np.array_equal(o1, o2)
        assert np.array_equal(b1, b2)
        assert np.array_equal(p1, p2)
        assert np.array_equal(np.sort(p2), np.arange(n))


def test_noalloc_frnn_refuses_oversize():
    import torch
    L = M.lib()
    pos = torch.zeros((16, 2), dtype=torch.float32, device="cuda")
    off = torch.zeros(17, dtype=torch.int32, device="cuda")
    # more points than the static buffer holds (refused before any access)
    rc = L.mcl_grid_frnn_csr_noalloc(_ptr(pos), (1 << 20) + 1, 0.0, 0.0, 0.01, 0.01, 100, 100, 0, 0.01, None,
                                     _ptr(off), None, 0, None, None)
    assert rc == 5
    # more cell keys than it holds
    rc = L.mcl_grid_frnn_csr_noalloc(_ptr(pos), 16, 0.0, 0.0, 1e-4, 1e-4, 2000, 2000, M.GRID_MORTON, 1e-4, None,
                                     _ptr(off), None, 0, None, None)
    assert rc == 5
    # an empty input gives an empty CSR
    off.fill_(5)
    tot = ctypes.c_longlong(-1)
    rc = L.mcl_grid_frnn_csr_noalloc(_ptr(pos), 0, 0.0, 0.0, 0.1, 0.1, 10, 10, 0, 0.1, None, _ptr(off), None, 0,
                                     ctypes.byref(tot), None)
    assert rc == 0 and tot.value == 0 and int(off[0].item()) == 0
