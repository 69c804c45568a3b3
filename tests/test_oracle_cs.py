"""Pins of the cooperative-signalling oracle (oracle/cs_oracle.py).

Independent of the oracle's own formulas: the reference's known-answer tests,
the VSWRM oracle's per-object decisions (pinned in test_oracle_vf.py by an
exact-rational re-derivation of every bin), and hand-derived cases whose
expected values are worked out in the comments from the reference's text.
"""
import math

import numpy as np
import pytest

from oracle import cs_oracle as co
from oracle import vf_oracle as vo

PI = np.pi


# --------------------------------------------------- reference known answers
def test_reference_kat_projection_field():
    # test_cs_supcalc.py:143-158
    out = co.projection_field((-PI, PI), 8, np.array([-1, -1]), 1, 0, [np.array([0, -1])], None)
    assert out.shape == (1, 8)
    assert np.all(out == [[0., 0., 0., 0., 1., 1., 0., 0.]])


@pytest.mark.parametrize("v1, v2, expected", [
    ([1, 0], [1, 0], 0), ([1, 0], [0, 1], -np.pi / 2), ([1, 0], [-1, 0], -np.pi)])
def test_reference_kat_closed_angle(v1, v2, expected):
    # test_cs_supcalc.py:161-169
    assert np.isclose(co.calculate_closed_angle(np.array(v1, float), np.array(v2, float)), expected)


# ------------------------------------------ agreement with the VSWRM decisions
def _raster(start, end, R):
    # the wrap rules of cs_supcalc.py:277-284 as sets of raw rays
    rays = set()
    if start < 0:
        rays |= set(range(R + start, R))
        start = 0
    if end >= R:
        rays |= set(range(0, end - (R - 1)))
        end = R
    return rays | set(range(start, end))


@pytest.mark.parametrize("seed", range(6))
@pytest.mark.parametrize("R", [8, 64, 512, 1200])
def test_rows_follow_vf_decisions(seed, R):
    # With walls, full FOV, no size limit and no meters the two projection
    # fields share centres (object size = focal radius), closed angle, bin and
    # half-width; the CS field draws every object (ca always lies in [-pi, pi])
    # where the VSWRM one applies its pixel test.
    rng = np.random.RandomState(seed)
    n = 40
    pos = np.array([rng.uniform(0, 300), rng.uniform(0, 300)])
    th = rng.uniform(0, 2 * PI)
    r = float(rng.choice([1, 5, 10]))
    objs = [np.array([rng.uniform(0, 300), rng.uniform(0, 300)]) for _ in range(n)]
    objs[3] = pos.copy()                                   # coincident: skipped
    objs[7] = pos + np.array([2.5 * r, 0.0])               # close: wide projection
    recs, _ = vo.projection_records((-PI, PI), R, pos, r, th, [o.copy() for o in objs])
    cs = co.projection_field((-PI, PI), R, pos, r, th, objs)
    for i, rec in enumerate(recs):
        raw = set(np.flatnonzero(np.flip(cs[i])))
        if rec.skipped:
            assert raw == set()
        else:
            assert raw == _raster(rec.proj_start, rec.proj_end, R), (i, rec)


# --------------------------------------------------------- hand-derived cases
def _facing_right(R, fov, objs, r=10.0, **kw):
    # focal at (100, 100) facing +x (orientation 0); centre (110, 110)
    return co.projection_field(fov, R, np.array([100.0, 100.0]), r, 0.0,
                               [np.array(o, float) for o in objs], **kw)


def test_fov_closed_angle_and_mask():
    R = 1200
    fov = (-0.45 * PI, 0.45 * PI)
    ahead = [160.0, 100.0]      # centre (170, 110): dead ahead, ca = 0
    behind = [40.0, 100.0]      # dead behind: angle_between = pi -> ca = -pi
    left = [100.0, 40.0]        # screen up (-y) = +90 deg -> ca = +pi/2 > 0.45 pi
    rows = _facing_right(R, fov, [ahead, behind, left])
    phis = np.linspace(-PI, PI, R)
    # the mask zeroes every post position with phis outside the FOV
    outside = (phis < fov[0]) | (phis > fov[1])
    assert not rows[:, outside].any()
    assert rows[0].sum() > 0
    assert not rows[1].any() and not rows[2].any()
    # ca = 0 ties between linspace entries 599 and 600 (first index wins ->
    # raw 599); size = 2 atan(10/60)/(2 pi) * 1200 = 63.33 -> half 31:
    # raw rays [568, 630) -> post positions 1199 - ray
    half = math.floor((2 * math.atan(10 / 60) / (2 * PI)) * R / 2)
    assert half == 31
    expect = {R - 1 - k for k in range(599 - half, 599 + half)}
    assert set(np.flatnonzero(rows[0])) == expect


def test_full_fov_masks_nothing_and_draws_behind():
    R = 64
    rows = _facing_right(R, (-PI, PI), [[40.0, 100.0]])
    # dead behind: ca = -pi -> bin 0, half of 2 atan(10/60)/(2 pi)*64 = 3.38 -> 1:
    # start -1, end 1 -> raw rays {63} (wrap) and [0, 1) -> post positions {0, 63}
    assert set(np.flatnonzero(rows[0])) == {0, 63}


def test_max_proj_size_filter():
    R = 1200
    size = (2 * math.atan(10 / 60) / (2 * PI)) * R      # 63.33...
    obj = [[160.0, 100.0]]
    assert _facing_right(R, (-PI, PI), obj, max_proj_size=64.0).any()
    assert not _facing_right(R, (-PI, PI), obj, max_proj_size=63.0).any()
    assert _facing_right(R, (-PI, PI), obj, max_proj_size=size).any()        # <= is inclusive
    assert _facing_right(R, (-PI, PI), obj, max_proj_size=None).any()


def test_meters_scale_rows():
    R = 256
    objs = [[160.0, 100.0], [100.0, 160.0], [30.0, 30.0]]
    base = _facing_right(R, (-PI, PI), objs)
    met = [0.25, 0.7, 1.5]
    scaled = _facing_right(R, (-PI, PI), objs, object_meters=met)
    for i in range(3):
        assert np.array_equal(scaled[i], base[i] * met[i])
    # a NaN meter turns the whole (unmasked) row into NaN: row *= nan
    nan_row = _facing_right(R, (-PI, PI), objs[:1], object_meters=[float("nan")])
    assert np.isnan(nan_row).all()


def test_combines():
    R = 128
    rng = np.random.RandomState(3)
    objs = [[rng.uniform(0, 300), rng.uniform(0, 300)] for _ in range(12)]
    pos, r, th = np.array([150.0, 150.0]), 10.0, 1.0
    rows = co.projection_field((-PI, PI), R, pos, r, th, [np.array(o) for o in objs],
                               max_proj_size=0.5 * R)
    # crowd: integer counts per ray divided by len(agents) (self included)
    crowd = co.crowd_density_proj((-PI, PI), R, pos, r, th, [np.array(o) for o in objs], 13, 0.5 * R)
    counts = (rows != 0).sum(axis=0)
    assert np.array_equal(crowd, counts / 13)
    # signalling: maximum of the meter-scaled rows; zeros when nobody signals
    met = list(rng.uniform(0, 1, 12))
    sig = co.signaling_density_proj((-PI, PI), R, pos, r, th, [np.array(o) for o in objs], met)
    expect = np.zeros(R)
    for i in range(12):
        expect = np.where(rows[i] != 0, np.maximum(expect, met[i]), expect)
    unbounded = co.projection_field((-PI, PI), R, pos, r, th, [np.array(o) for o in objs])
    assert np.array_equal(sig, np.max(unbounded * np.array(met)[:, None], axis=0))
    assert np.array_equal(co.signaling_density_proj((-PI, PI), R, pos, r, th, [], []), np.zeros(R))


def test_signal_memory_trace():
    # cs_agent.py:106-146 with depth 3: entries are inserted at the front,
    # entry n is scaled by 1 - n/3, and the loop stops (popping the last
    # entry) once n reaches 3, after having appended that entry.
    m = co.SignalMemory()
    a, b, c = np.array([1.0, 1.0]), np.array([2.0, 2.0]), np.array([3.0, 3.0])
    m.push([a], [0.9])                     # step 1
    pos, met = m.extended(3)
    assert [list(p) for p in pos] == [[1, 1]] and met == [0.9]
    m.push([], [])                         # step 2: nobody signals
    pos, met = m.extended(3)
    assert met == [0.9 * (1 - 1 / 3)]
    m.push([b, c], [0.5, 0.4])             # step 3
    pos, met = m.extended(3)
    assert met == [0.5, 0.4, 0.9 * (1 - 2 / 3)]
    m.push([], [])                         # step 4: entries n=0..3, n=3 has decay 0
    pos, met = m.extended(3)
    assert met == [0.5 * (1 - 1 / 3), 0.4 * (1 - 1 / 3), 0.9 * 0.0]
    assert len(m.entries) == 3             # the oldest was dropped
    m.push([], [])                         # step 5
    pos, met = m.extended(3)
    assert met == [0.5 * (1 - 2 / 3), 0.4 * (1 - 2 / 3)]
    m.push([], [])                         # step 6: the b/c entry reaches n = 3
    pos, met = m.extended(3)
    assert met == [0.5 * 0.0, 0.4 * 0.0] and len(m.entries) == 3
