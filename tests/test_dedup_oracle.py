"""Duplicate removal of a pick list: the FP64 oracle
(oracle/gisspa_oracle.cpp, or_remove_repeats, restating
python_scripts/remove_repeat_particles_from_list.py:9-145), pinned by
  * the orientation distance against an independent rotation-angle
    computation (numpy ZXZ matrices, scipy's rotation magnitude);
  * closed forms: identical orientations, in-plane turns, 180-degree turns
    about z (kept as duplicates by the script's folding) and about x (a pair
    the script skips through its quaternion branches);
  * hand-worked walks of the script's loop (ties, chains, a dropped pick that
    still drops later ones, groups as runs of equal paths);
  * for scores descending in list order, the textbook greedy suppression.
CPU only."""
import numpy as np
import pytest
from scipy.spatial.transform import Rotation

from oracle import oracle as orc


def zxz(alt, az, phi):
    """Euler matrix Rz(phi) Rx(alt) Rz(az) (EMAN convention), degrees."""
    def rz(t):
        c, s = np.cos(np.radians(t)), np.sin(np.radians(t))
        return np.array([[c, s, 0], [-s, c, 0], [0, 0, 1]])

    def rx(t):
        c, s = np.cos(np.radians(t)), np.sin(np.radians(t))
        return np.array([[1, 0, 0], [0, c, s], [0, -s, c]])
    return rz(phi) @ rx(alt) @ rz(az)


def test_euler_distance_is_the_relative_rotation_angle():
    rng = np.random.default_rng(0)
    for _ in range(400):
        e1 = (rng.uniform(0, 180), rng.uniform(0, 360), rng.uniform(0, 360))
        e2 = (rng.uniform(0, 180), rng.uniform(0, 360), rng.uniform(0, 360))
        d = orc.euler_distance(e1, e2)
        assert d is not None
        theta = np.degrees(Rotation.from_matrix(zxz(*e2) @ zxz(*e1).T).magnitude())
        # the script's quaternion sign is not fixed: 2 acos(+-cos(theta/2))
        assert min(abs(d - theta), abs(d - (360 - theta))) < 1e-6, (e1, e2, d, theta)


def test_euler_distance_closed_forms():
    assert orc.euler_distance((30, 40, 50), (30, 40, 50)) == pytest.approx(0, abs=1e-6)
    for dphi in (1.0, 5.0, 33.0, 90.0):
        d = orc.euler_distance((30, 40, 50), (30, 40, 50 + dphi))
        assert min(abs(d - dphi), abs(d - (360 - dphi))) < 1e-8
    # 180 degrees about z: computed (q0 = 0)
    assert orc.euler_distance((0, 0, 0), (0, 0, 180)) == pytest.approx(180, abs=1e-6)
    # 180 degrees about x: trace -1, x-dominant, then the z branch with t = 0
    # -> the script skips the pair
    assert orc.euler_distance((0, 0, 0), (180, 0, 0)) is None


def _rm(x, y, alt=0.0, az=0.0, phi=0.0, score=1.0, group=0):
    return dict(group=group, x=x, y=y, alt=alt, az=az, phi=phi, score=score)


def _run(rows, center=20.0, thr=10.0):
    cols = {k: [r[k] for r in rows] for k in rows[0]}
    return orc.remove_repeats(cols["group"], cols["x"], cols["y"], cols["alt"], cols["az"],
                              cols["phi"], cols["score"], center, thr).tolist()


def test_pair_rules():
    # same place, same orientation: the lower score goes, the earlier on a tie
    assert _run([_rm(100, 100, score=2), _rm(101, 100, score=1)]) == [True, False]
    assert _run([_rm(100, 100, score=1), _rm(101, 100, score=2)]) == [False, True]
    assert _run([_rm(100, 100, score=1), _rm(101, 100, score=1)]) == [False, True]
    # centre distance is strict: 12-16-20 triangle, distance exactly 20
    assert _run([_rm(0, 0, score=2), _rm(12, 16, score=1)]) == [True, True]
    assert _run([_rm(0, 0, score=2), _rm(12, 15.9, score=1)]) == [True, False]
    # orientation within the threshold, or near a 180-degree in-plane turn
    assert _run([_rm(0, 0, phi=0, score=2), _rm(1, 0, phi=5, score=1)]) == [True, False]
    assert _run([_rm(0, 0, phi=0, score=2), _rm(1, 0, phi=15, score=1)]) == [True, True]
    assert _run([_rm(0, 0, phi=0, score=2), _rm(1, 0, phi=175, score=1)]) == [True, False]
    assert _run([_rm(0, 0, phi=0, score=2), _rm(1, 0, phi=169, score=1)]) == [True, True]
    # a 180-degree flip of the tilt axis is never a duplicate (skipped pair)
    assert _run([_rm(0, 0, alt=0, score=2), _rm(1, 0, alt=180, score=1)]) == [True, True]


def test_walk_chain_and_dropped_pick_still_drops():
    # A(3) ~ B(2) ~ C(1), A !~ C: A drops B; B is skipped at its turn, so C stays
    rows = [_rm(0, 0, score=3), _rm(15, 0, score=2), _rm(30, 0, score=1)]
    assert _run(rows) == [True, False, True]
    # A(1) ~ B(2), A ~ C(0.5), B !~ C: at A's turn B drops A, yet A goes on and
    # drops C; at B's turn C is already gone -> only B stays (a textbook
    # greedy suppression would keep B and C)
    rows = [_rm(15, 0, score=1), _rm(0, 0, score=2), _rm(30, 0, score=0.5)]
    assert _run(rows) == [False, True, False]


def test_groups_are_runs_of_equal_paths():
    a, b = _rm(0, 0, score=2), _rm(1, 0, score=1)
    assert _run([dict(a, group=0), dict(b, group=1)]) == [True, True]
    # a path repeated after another one starts a new group (:25-36)
    rows = [dict(a, group=7), _rm(500, 500, group=8), dict(b, group=7)]
    assert _run(rows) == [True, True, True]
    rows = [dict(a, group=7), dict(b, group=7), _rm(500, 500, group=8)]
    assert _run(rows) == [True, False, True]


def _dup(p, q, center, thr):
    if np.hypot(p[0] - q[0], p[1] - q[1]) >= center:
        return False
    d = orc.euler_distance(p[2:5], q[2:5])
    if d is None:
        return False
    d2 = d % 180
    if d2 > thr + 1:
        d2 = 180 - d2
    return d < thr or d2 < thr


def test_descending_scores_give_greedy_suppression():
    rng = np.random.default_rng(1)
    n = 120
    centres = rng.uniform(0, 300, (15, 2))
    pick = rng.integers(0, 15, n)
    xy = centres[pick] + rng.normal(0, 6, (n, 2))
    views = rng.uniform(0, 30, (4, 2))
    v = rng.integers(0, 4, n)
    alt, az = views[v, 0], views[v, 1]
    phi = rng.choice([0.0, 5.0, 10.0, 180.0], n)
    score = np.sort(rng.uniform(5, 20, n))[::-1]
    center, thr = 15.0, 8.0
    keep = orc.remove_repeats(np.zeros(n, np.int32), xy[:, 0], xy[:, 1], alt, az, phi, score,
                              center, thr)
    pts = [(xy[i, 0], xy[i, 1], alt[i], az[i], phi[i]) for i in range(n)]
    kept = []
    for i in range(n):
        if not any(_dup(pts[k], pts[i], center, thr) for k in kept):
            kept.append(i)
    expect = np.zeros(n, bool)
    expect[kept] = True
    assert 0 < expect.sum() < n
    assert np.array_equal(keep, expect)
