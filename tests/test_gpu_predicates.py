"""GPU predicates against the golden exact signs.

The device functions the flip, repair and relocation kernels use are run
through mcl_eval_predicates on every golden case (tests/golden/
predicates.npz, exact rational signs): zero sign disagreements are allowed,
for the FP32-first cascade of the fast path and the FP64-first cascade of
the plain path.  The plain FP64 determinant (no filter, no exact stage) is
run on the same cases and its disagreements are counted: they are the
errors the cascade removes.
"""
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu
HERE = os.path.dirname(os.path.abspath(__file__))
GOLD = np.load(os.path.join(HERE, "golden", "predicates.npz"))


def _eval(kind, pts, periodic):
    import torch

    import mcleap_amd as M
    n = len(pts)
    d = torch.from_numpy(np.ascontiguousarray(pts, np.float32).reshape(-1, 2)).cuda()
    s = torch.zeros(n, dtype=torch.int32, device="cuda")
    st = torch.zeros(n, dtype=torch.int32, device="cuda")
    rc = M.lib().mcl_eval_predicates(kind, n, d.data_ptr(), int(periodic), s.data_ptr(), st.data_ptr(), None)
    assert rc == 0
    return s.cpu().numpy(), st.cpu().numpy()


def _split(name):
    x, p = GOLD[name + "_in"], GOLD[name + "_periodic"]
    return (x[~p], GOLD[name + "_sign"][~p], False), (x[p], GOLD[name + "_sign"][p], True)


@pytest.mark.parametrize("kind", [0, 2])
def test_orient_cascade_exact(kind):
    stages = np.zeros(4, np.int64)
    for x, want, per in _split("orient"):
        got, st = _eval(kind, x, per)
        assert np.array_equal(got, want), (per, np.nonzero(got != want)[0][:10])
        stages += np.bincount(st, minlength=4)[:4]
    print("orient kind %d stages fp32/fp64/exact-small/exact-big:" % kind, stages.tolist())
    assert stages[2] + stages[3] > 0  # the exact stage was exercised


@pytest.mark.parametrize("kind", [1, 3])
def test_incircle_cascade_exact_with_tie_rule(kind):
    stages = np.zeros(4, np.int64)
    p = GOLD["incircle_periodic"]
    for per in (False, True):
        x = GOLD["incircle_in"][p == per]
        want = GOLD["incircle_sos"][p == per]
        got, st = _eval(kind, x, per)
        assert np.array_equal(got, want), (per, np.nonzero(got != want)[0][:10])
        stages += np.bincount(st, minlength=4)[:4]
    print("incircle kind %d stages:" % kind, stages.tolist())
    assert stages[2] + stages[3] > 0


def test_plain_fp64_forms_disagree_on_adversarial_cases():
    wrong_o = wrong_i = 0
    for x, want, per in _split("orient"):
        got, _ = _eval(4, x, per)
        wrong_o += int((got != want).sum())
    for x, want, per in _split("incircle"):
        got, _ = _eval(5, x, per)
        wrong_i += int((got != want).sum())
    print("plain FP64: %d orient and %d incircle signs differ from exact" % (wrong_o, wrong_i))
    assert wrong_i > 0


def test_exact_only_kinds_equal_golden():
    """the exact-only kinds (no filter: every case through the expansion
    arithmetic, the reference for the sign-disagreement counts bench.py
    reports) give the golden exact signs, with the tie rule for incircle"""
    for x, want, per in _split("orient"):
        got, _ = _eval(7, x, per)
        assert np.array_equal(got, want), (per, np.nonzero(got != want)[0][:10])
    p = GOLD["incircle_periodic"]
    for per in (False, True):
        got, _ = _eval(6, GOLD["incircle_in"][p == per], per)
        assert np.array_equal(got, GOLD["incircle_sos"][p == per]), per
