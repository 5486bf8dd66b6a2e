"""Writes the golden predicate vectors (tests/golden/predicates.npz).

    python tests/golden/make_vectors.py

Inputs are FP32 coordinates; every expected sign is computed here with exact
rational arithmetic (fractions.Fraction), independently of the oracle and of
the GPU code:

  orient(a, b, c)      sign of (a-c) x (b-c)
  incircle(a, b, c, d) sign of the 3x3 lifted determinant relative to d
                       (positive = d inside the circle of CCW a, b, c; the
                       convention of Helpers_Triangulation.h:136-151)
  incircle_sos         the same sign with exact ties broken by symbolic
                       perturbation of the lifts: D(w + eps) for the 4x4
                       matrix with rows (x, y, w, 1), eps_i decreasing with the
                       rank of the point in decreasing lexicographic (x, y)
                       order; the tie sign is the sign of the cofactor dD/dw_i
                       of the highest-ranked point whose cofactor is nonzero
  periodic rows        differences taken as the minimum image on the unit
                       torus: b - a shifted by -1 if > 1/2, +1 if < -1/2

Input families: uniform random; near-collinear and exactly collinear; points
with a +-1e7 corner (the reference's outer vertices); the cocircular square
a=(-1,0), b=(0,-1), c=(1,0), d=(0,1) with d moved on a 2^-23 grid (the
construction of experiments/floating_point_predicates_precision.cu:184-207);
lattice quads k/1000 + U(+-1e-7) (near-cocircular); axis-aligned rectangles
in random vertex order (exactly cocircular); points straddling the torus seam.
"""
import os
from fractions import Fraction as F

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))


def _fr(v):
    return F(float(v))


def _minimg(p, base):
    """p shifted by an integer vector to the minimum image relative to base"""
    out = []
    for k in range(2):
        d = p[k] - base[k]
        s = -1 if d > F(1, 2) else (1 if d < -F(1, 2) else 0)
        out.append(p[k] + s)
    return out


def orient_exact(a, b, c, periodic=False):
    a, b, c = [[_fr(x) for x in p] for p in (a, b, c)]
    if periodic:
        a, b = _minimg(a, c), _minimg(b, c)
    det = (a[0] - c[0]) * (b[1] - c[1]) - (a[1] - c[1]) * (b[0] - c[0])
    return (det > 0) - (det < 0)


def _det(m):
    n = len(m)
    if n == 1:
        return m[0][0]
    s = F(0)
    for j in range(n):
        if m[0][j] == 0:
            continue
        minor = [row[:j] + row[j + 1:] for row in m[1:]]
        s += (-1) ** j * m[0][j] * _det(minor)
    return s


def incircle_exact(a, b, c, d, periodic=False, sos=False):
    a, b, c, d = [[_fr(x) for x in p] for p in (a, b, c, d)]
    if periodic:
        a, b, c = _minimg(a, d), _minimg(b, d), _minimg(c, d)
    rows = []
    for p in (a, b, c):
        dx, dy = p[0] - d[0], p[1] - d[1]
        rows.append([dx, dy, dx * dx + dy * dy])
    det = _det(rows)
    if det != 0 or not sos:
        return (det > 0) - (det < 0)
    pts = [a, b, c, d]
    m4 = [[p[0], p[1], p[0] * p[0] + p[1] * p[1], F(1)] for p in pts]
    assert _det(m4) == 0
    cof = []
    for i in range(4):
        mi = [row[:] for row in m4]
        for r in range(4):
            mi[r][2] = F(1) if r == i else F(0)
        cof.append(_det(mi))
    # rank: canonical coordinates are the stored ones (before the shift)
    raw = [[_fr(x) for x in p] for p in SOS_RAW]
    order = sorted(range(4), key=lambda i: (raw[i][0], raw[i][1]), reverse=True)
    for i in order:
        if cof[i] != 0:
            return (cof[i] > 0) - (cof[i] < 0)
    return 0


SOS_RAW = None


def f32(x):
    return np.asarray(x, dtype=np.float32)


def gen(seed=2024):
    rng = np.random.default_rng(seed)
    O, OP = [], []
    I, IP = [], []
    # --- orient
    for _ in range(1500):
        O.append(f32(rng.random((3, 2)))); OP.append(False)
    for _ in range(600):  # near-collinear
        a, b = rng.random(2), rng.random(2)
        t = rng.random() * 1.4 - 0.2
        c = a + t * (b - a)
        tri = f32([a, b, c])
        k = rng.integers(-2, 3, size=2)
        tri[2] = tri[2] + k.astype(np.float32) * np.spacing(tri[2])
        O.append(tri); OP.append(False)
    for _ in range(200):  # exactly collinear on a dyadic line
        s = rng.integers(1, 1000, size=3) / 1024.0
        O.append(f32([[s[0], s[0]], [s[1], s[1]], [s[2], s[2]]])); OP.append(False)
    for _ in range(300):  # a reference corner
        cx, cy = rng.choice([-1e7, 1e7]), rng.choice([-1e7, 1e7])
        tri = [[cx, cy]] + rng.random((2, 2)).tolist()
        O.append(f32(tri)); OP.append(False)
    for _ in range(600):  # straddling the torus seam
        p = rng.random((3, 2)) * 0.02
        p[rng.random((3, 2)) < 0.5] += 0.98
        O.append(f32(np.mod(p, 1.0))); OP.append(True)
    # --- incircle
    for _ in range(1500):
        I.append(f32(rng.random((4, 2)))); IP.append(False)
    sq = np.array([[-1, 0], [0, -1], [1, 0], [0, 1]], np.float64)
    for i in range(-8, 9):
        for j in range(-8, 9):
            q = sq.copy()
            q[3] += np.array([i, j]) * 2.0 ** -23
            I.append(f32(q)); IP.append(False)
    for _ in range(800):  # lattice quads with +-1e-7 jitter (near-cocircular)
        k = rng.integers(0, 999, size=2)
        base = np.array([[k[0], k[1]], [k[0] + 1, k[1]], [k[0] + 1, k[1] + 1], [k[0], k[1] + 1]], np.float64) / 1000.0
        base += rng.uniform(-1e-7, 1e-7, base.shape)
        I.append(f32(base)); IP.append(False)
    for _ in range(400):  # exact rectangles, random vertex order -> exactly cocircular
        x0, x1 = np.sort(rng.random(2)); y0, y1 = np.sort(rng.random(2))
        r = f32([[x0, y0], [x1, y0], [x1, y1], [x0, y1]])
        I.append(r[rng.permutation(4)]); IP.append(False)
    for _ in range(300):  # one reference corner as the tested point or a circle point
        q = rng.random((4, 2))
        q[rng.integers(0, 4)] = [rng.choice([-1e7, 1e7]), rng.choice([-1e7, 1e7])]
        I.append(f32(q)); IP.append(False)
    for _ in range(600):  # torus seam, random and rectangles
        p = rng.random((4, 2)) * 0.02
        p[rng.random((4, 2)) < 0.5] += 0.98
        I.append(f32(np.mod(p, 1.0))); IP.append(True)
    for _ in range(200):
        x0, y0 = rng.random(2) * 0.01 + 0.985
        w, h = rng.random(2) * 0.01 + 0.001
        r = np.mod(np.array([[x0, y0], [x0 + w, y0], [x0 + w, y0 + h], [x0, y0 + h]]), 1.0)
        r = f32(r)
        I.append(r[rng.permutation(4)]); IP.append(True)
    return np.stack(O), np.array(OP), np.stack(I), np.array(IP)


def rng_fixtures():
    """compile mt_vectors.cpp with g++ (libstdc++ <random>) and pack its draws"""
    import subprocess
    import tempfile
    with tempfile.TemporaryDirectory() as tmp:
        exe = os.path.join(tmp, "mt_vectors")
        subprocess.check_call(["g++", "-O2", "-o", exe, os.path.join(HERE, "mt_vectors.cpp")])
        kat = int(subprocess.check_output([exe, tmp]).decode().split()[0])
        rd = lambda name, dt: np.fromfile(os.path.join(tmp, name), dtype=dt)
        np.savez_compressed(os.path.join(HERE, "rng.npz"), mt_raw_3=rd("mt_raw_3.u32", np.uint32),
                            mt_raw_10=rd("mt_raw_10.u32", np.uint32),
                            uniform_f32_seed3=rd("uniform_f32_seed3.f32", np.float32),
                            uniform_f64_seed10=rd("uniform_f64_seed10.f64", np.float64)[:4000],
                            mt_default_10000th=np.array([kat], np.uint32))
    print("rng fixtures: 10000th default draw", kat)


def main():
    global SOS_RAW
    rng_fixtures()
    O, OP, I, IP = gen()
    os_ = np.array([orient_exact(*t, periodic=p) for t, p in zip(O, OP)], np.int8)
    is_, iss = [], []
    for q, p in zip(I, IP):
        SOS_RAW = q
        is_.append(incircle_exact(*q, periodic=p))
        iss.append(incircle_exact(*q, periodic=p, sos=True))
    np.savez_compressed(os.path.join(HERE, "predicates.npz"), orient_in=O, orient_periodic=OP,
                        orient_sign=os_, incircle_in=I, incircle_periodic=IP,
                        incircle_sign=np.array(is_, np.int8), incircle_sos=np.array(iss, np.int8))
    print("orient: %d cases (%d zero), incircle: %d cases (%d zero)"
          % (len(O), int((os_ == 0).sum()), len(I), int((np.array(is_) == 0).sum())))


if __name__ == "__main__":
    main()
