"""Pins of the visual-flocking oracle (oracle/vf_oracle.py).

Every check here is independent of the oracle's own formulas: the
reference's known-answer tests, hand-derived raster cases (expected bits
worked out from the reference's text in the comments), an exact-arithmetic
re-derivation of every ray-bin decision by a different method, exhaustive
enumeration for the edge rule, and closed forms for integrals and motion.
"""
import itertools
import math
from fractions import Fraction

import numpy as np
import pytest

from oracle import vf_oracle as vo

PI = np.pi


def raw_bits(rows):
    """Oracle rows are flipped along rays; return the raw raster bit sets."""
    return [set(np.flatnonzero(np.flip(r))) for r in np.atleast_2d(rows)]


# --------------------------------------------------- reference known answers
def test_reference_kat_projection_field():
    # test_cs_supcalc.py:143-158 (cs_supcalc.projection_field is the same
    # algorithm as vf_supcalc.projection_field in walls mode, full FOV)
    out = vo.projection_field((-PI, PI), 8, np.array([-1, -1]), 1, 0, [np.array([0, -1])], None)
    assert out.shape == (1, 8)
    assert np.all(out == [[0., 0., 0., 0., 1., 1., 0., 0.]])


@pytest.mark.parametrize("v1, v2, expected", [
    ([1, 0], [1, 0], 0), ([1, 0], [0, 1], -np.pi / 2), ([1, 0], [-1, 0], -np.pi)])
def test_reference_kat_closed_angle(v1, v2, expected):
    # test_cs_supcalc.py:161-168 (identical function, vf_supcalc.py:142-158)
    assert np.isclose(vo.closed_angle(np.array(v1, float), np.array(v2, float)), expected)


def test_reference_kat_rng_stream():
    # test_cs_supcalc.py:10-34: legacy MT19937 seeded 42, uniform(lo, hi)
    np.random.seed(42)
    assert np.random.uniform(-0.3, 0.3) == -0.0752759286915825
    assert np.random.uniform(-0.5, 0.3) == 0.26057144512793295
    assert np.random.uniform(-0.3, 0.3) == 0.13919636508684308


# ----------------------------------------------------------------- grids
@pytest.mark.parametrize("R", [2, 3, 8, 16, 255, 256, 512, 600, 1023, 1024, 1200, 2048, 3000])
def test_linspace_closed_form(R):
    # numpy linspace: fl(fl(k*step) + start), last element = stop
    step = (PI - (-PI)) / (R - 1)
    mine = np.array([(k * step) + (-PI) for k in range(R)])
    mine[-1] = PI
    assert np.array_equal(np.linspace(-PI, PI, R), mine)


def test_arange_length_quirk():
    # PHI = np.arange(-pi, pi, 2pi/R) (vf_agent.py:44) has R+1 entries for R=3000
    assert len(np.arange(-PI, PI, 2 * PI / 3000)) == 3001
    for R in (8, 256, 512, 1024, 1200, 2048):
        assert len(np.arange(-PI, PI, 2 * PI / R)) == R


@pytest.mark.parametrize("R, idx", [(8, 3), (256, 127), (512, 256), (1024, 511), (1200, 599),
                                    (2048, 1023)])
def test_dead_ahead_tie_index(R, idx):
    # argmin |linspace + 0.0| takes the first index on ties
    assert vo.find_nearest(np.linspace(-PI, PI, R), -0.0) == idx


# ------------------------------------------------------ hand-derived rasters
FOCAL = np.array([10.0, 10.0])     # focal top-left, radius 1 -> centre (11, 11), heading theta=0


def one(objpos, R=16, boundary="walls", fov=(-PI, PI), W=None, H=None, size=1):
    return vo.projection_field(fov, R, FOCAL, 1, 0.0, [np.array(objpos, float)], [size],
                               boundary, W, H)


def test_raster_wrap_at_bin0():
    # object centre 2 px straight behind: v2 = (-2, 0) -> angle pi -> a'=pi ->
    # VF inclusive rule -> ca = -pi -> t = 0.  ps/2 = atan(1/2)*16/(2pi) = 1.18
    # -> h = 1 -> start=-1, end=1: visible (0<1<15); start<0 paints [15,16),
    # then [0,1) -> raw bits {0, 15}
    assert raw_bits(one([8.0, 10.0])) == [{0, 15}]


def test_raster_wrap_at_last_bin_paints_extra_bit():
    # centre (9, 10.99): v2 = (-2, -0.01): cross < 0 -> a = -(pi - 0.005) ->
    # a' = pi + 0.005 -> ca = pi - 0.005 -> t = 15; h = 1 -> start=14, end=16
    # -> end >= R paints [0, 16-15) = {0} and end=R -> [14,16): raw {0, 14, 15}
    assert raw_bits(one([8.0, 9.99])) == [{0, 14, 15}]


def test_raster_below_two_rays_is_empty():
    # d = 3: ps/2 = atan(1/3)*16/(2pi) = 0.82 -> h = 0 -> empty slice
    assert raw_bits(one([13.0, 10.0])) == [set()]
    # d = 2 ahead: ps/2 = 1.18 -> h = 1; ca = -0.0 -> t = 7 (first of the 7/8 tie?)
    # the tie is broken by rounding of linspace; both neighbours give 2 bits
    b = raw_bits(one([12.0, 10.0]))[0]
    assert len(b) == 2 and max(b) - min(b) == 1


def test_raster_side_object_and_limited_fov():
    # v2 = 2*(cos a, sin a), a = pi/2 + 0.3: cross > 0 -> angle a -> ca = -a
    # = -1.8708 -> (ca+pi)/(2pi/15) = 3.03 -> t = 3, h = 1 -> start 2, end 4.
    a = PI / 2 + 0.3
    pos = [10.0 + 2 * math.cos(a), 10.0 + 2 * math.sin(a)]
    assert raw_bits(one(pos)) == [{2, 3}]
    # FOV (-pi/2, pi/2): fov_px = [4, 11] (nearest linspace points); neither
    # 4<2<11 nor 4<4<11 holds -> not drawn
    assert raw_bits(one(pos, fov=(-PI / 2, PI / 2))) == [set()]


def test_raster_torus_image():
    # walls: object centre (109, 11) is 98 px ahead -> h = 0 -> empty
    assert raw_bits(one([108.0, 10.0], boundary="walls", W=100, H=100)) == [set()]
    # torus, W=100: |98| > 50 and centre_x < object_x -> object shifted to 9:
    # 2 px behind -> same as test_raster_wrap_at_bin0
    assert raw_bits(one([108.0, 10.0], boundary="infinite", W=100, H=100)) == [{0, 15}]


def test_coincident_object_skipped():
    out = vo.projection_field((-PI, PI), 16, FOCAL, 1, 0.0, [FOCAL.copy()], [1])
    assert not out.any()


# ------------------------- independent re-derivation of every bin decision
def _exact_decisions(pos, r, theta, objs, sizes, R, torus, W, H):
    """Ray bin t and half-width h by exact rational geometry and atan2:
    a different route from the reference's unit vectors + arccos + sign +
    remainder + mapping.  Returns (t, h, margin_t, margin_h) per object."""
    phis = [Fraction(float(p)) for p in np.linspace(-PI, PI, R)]
    cx, cy = Fraction(pos[0]) + r, Fraction(pos[1]) + r
    ex = Fraction(pos[0] + (1 + np.cos(theta)) * r)
    ey = Fraction(pos[1] + (1 - np.sin(theta)) * r)
    v1x, v1y = ex - Fraction(float(pos[0] + r)), ey - Fraction(float(pos[1] + r))
    out = []
    for (ox, oy), s in zip(objs, sizes):
        ocx, ocy = Fraction(ox) + s, Fraction(oy) + s
        dx, dy = ocx - cx, ocy - cy
        if torus:
            if abs(dx) > Fraction(W, 2):
                ocx += -W if cx < ocx else W
            if abs(dy) > Fraction(H, 2):
                ocy += -H if cy < ocy else H
            dx, dy = ocx - cx, ocy - cy
        cross = v1x * dy - v1y * dx
        dot = v1x * dx + v1y * dy
        ang = math.atan2(float(cross), float(dot))          # signed angle v1 -> v2
        ca = -ang if ang != -math.pi else -math.pi          # VF mapping, exact math
        caf = Fraction(ca)
        dist = [abs(p - caf) for p in phis]
        t = min(range(R), key=lambda k: dist[k])
        srt = sorted(dist)
        margin_t = float(srt[1] - srt[0])
        d = math.sqrt(float(dx * dx + dy * dy))
        half = math.atan(s / d) * R / (2 * math.pi)
        h = math.floor(half)
        out.append((t, h, margin_t, min(half - h, h + 1 - half)))
    return out


@pytest.mark.parametrize("torus", [False, True])
def test_bin_decisions_match_exact_geometry(torus):
    rng = np.random.RandomState(7 + torus)
    R, W, H = 64, 200, 200
    checked = 0
    for trial in range(40):
        pos = rng.uniform(30, 230, 2)
        theta = rng.uniform(0, 2 * PI)
        objs = [rng.uniform(30, 230, 2) for _ in range(12)]
        sizes = [int(s) for s in rng.randint(1, 8, 12)]
        recs, _ = vo.projection_records((-PI, PI), R, pos, 5, theta, [o.copy() for o in objs],
                                        sizes, "infinite" if torus else "walls", W, H)
        ex = _exact_decisions(pos, 5, theta, objs, sizes, R, torus, W, H)
        for rec, (t, h, mt, mh) in zip(recs, ex):
            if mt > 1e-9:
                assert rec.phi_target == t
                checked += 1
            if mh > 1e-9:
                assert math.floor(rec.proj_size / 2) == h
    assert checked > 400


# ------------------------------------------------------------ dPhi rule
@pytest.mark.parametrize("R", range(2, 13))
def test_dphi_exhaustive(R):
    # closed form (vf_supcalc.py:266-274): with e_k = V_k xor V_{k-1 mod R},
    # out_k = V_k - V_{k-1} when V_0 = 1 and V_{R-1} = 0, else V_{k+1} - V_k.
    Phi = np.arange(-PI, PI, 2 * PI / R)[:R]
    for bits in itertools.product([0.0, 1.0], repeat=R):
        V = np.array(bits)
        got = vo.dphi_v_of(Phi, V)
        if V[0] == 1 and V[-1] == 0:
            exp = V - np.roll(V, 1)
        else:
            exp = np.roll(V, -1) - V
        assert np.array_equal(got, exp)


# ------------------------------------------------------ integral closed forms
P = vo.VFParams(GAM=0.1, V0=1.0, ALP0=1.3, ALP1=0.09, BET0=0.7, BET1=0.09)


def test_empty_field_closed_form():
    R = 512
    Phi = np.arange(-PI, PI, 2 * PI / R)
    dv, dpsi = vo.vswrm_state_variables(0.4, Phi, np.zeros(R), P)
    assert dv == P.GAM * (P.V0 - 0.4) and dpsi == 0.0


@pytest.mark.parametrize("a, b", [(100, 140), (3, 9), (250, 400), (1, 511)])
def test_single_run_closed_form(a, b):
    # V = 1 on rays [a, b): trapz(-cos Phi) ~ -(sin(Phi_{b-1}+d/2) - sin(Phi_a-d/2))
    # (midpoint-rule error O(d^2)); V_0 = 0 so out_k^2 = e_{k+1}: the edges at a
    # and b land on k = a-1 and k = b-1 -> spike sum cos(Phi_{a-1}) + cos(Phi_{b-1}).
    R = 512
    Phi = np.arange(-PI, PI, 2 * PI / R)
    d = 2 * PI / R
    V = np.zeros(R)
    V[a:b] = 1
    _, _, ab, ae, bb, be = vo.vswrm_state_variables(0.0, Phi, V, P, verbose=True)
    lo, hi = Phi[a] - d / 2, Phi[b - 1] + d / 2
    assert abs(ab - P.ALP0 * -(math.sin(hi) - math.sin(lo))) < 5 * d * d
    assert abs(bb - P.BET0 * (math.cos(hi) - math.cos(lo))) < 5 * d * d
    assert math.isclose(ae, P.ALP0 * P.ALP1 * (math.cos(Phi[a - 1]) + math.cos(Phi[b - 1])),
                        rel_tol=1e-12, abs_tol=1e-15)
    assert math.isclose(be, P.BET0 * P.BET1 * (math.sin(Phi[a - 1]) + math.sin(Phi[b - 1])),
                        rel_tol=1e-12, abs_tol=1e-15)


def test_run_through_ray0_uses_other_alignment():
    # V_0 = 1, V_{R-1} = 0 -> out_k = V_k - V_{k-1}: edges at k=0? no (V_{-1}=V_{R-1}=0
    # -> e_0 = 1) and at k=b -> spike sum cos(Phi_0) + cos(Phi_b)
    R = 64
    Phi = np.arange(-PI, PI, 2 * PI / R)
    V = np.zeros(R)
    V[0:10] = 1
    _, _, _, ae, _, _ = vo.vswrm_state_variables(0.0, Phi, V, P, verbose=True)
    assert math.isclose(ae, P.ALP0 * P.ALP1 * (math.cos(Phi[0]) + math.cos(Phi[10])), rel_tol=1e-12)


# ----------------------------------------------------------- motion
def _agent(x=100.0, y=100.0, th=0.0, v=0.0, boundary="infinite"):
    return vo.VFAgentRec(id=0, radius=5, position=np.array([x, y]), orientation=th, velocity=v,
                         v_field_res=64, FOV=(-PI, PI), WIDTH=200, HEIGHT=200, window_pad=30,
                         boundary_cond=boundary)


def test_lone_agent_closed_form():
    # no neighbours -> V = 0 -> dv = GAM*(V0 - v), dpsi = 0; x += v' cos(0)
    ag = _agent(v=0.25)
    vo.agent_update(ag, [ag], P)
    v1 = 0.25 + P.GAM * (P.V0 - 0.25)
    assert ag.velocity == v1 and ag.orientation == 0.0
    assert ag.position[0] == 100.0 + v1 and ag.position[1] == 100.0


def test_orientation_one_shot_wrap():
    ag = _agent(th=0.1)
    vo.update_agent_position(ag, -0.2, 0.0)
    assert ag.orientation == 2 * PI + (0.1 - 0.2)
    ag = _agent(th=6.2)
    vo.update_agent_position(ag, 0.2, 0.0)
    assert ag.orientation == (6.2 + 0.2) - 2 * PI


def test_teleport_snaps_not_wraps():
    # centre x < 30 -> pos_x = 230 - r; centre x > 230 -> pos_x = 30 + r (snap)
    ag = _agent(x=20.0)
    vo.teleport_infinite_arena(ag)
    assert ag.position[0] == 230 - 5
    ag = _agent(x=230.0)
    vo.teleport_infinite_arena(ag)
    assert ag.position[0] == 30 + 5


def test_walls_rotate_quarter_turn():
    ag = _agent(x=20.0, th=3 * PI / 4, boundary="walls")
    vo.reflect_from_walls(ag)
    assert ag.position[0] == 30 - 5 and ag.orientation == 3 * PI / 4 - PI / 2
    ag = _agent(y=240.0, th=5 * PI / 4, boundary="walls")
    vo.reflect_from_walls(ag)
    assert ag.position[1] == 230 - 5 - 1 and ag.orientation == 5 * PI / 4 - PI / 2


# ---------------------------------------------------- invariants of a step
def test_torus_translation_invariance():
    # shifting every agent by the same vector on a torus (without crossing a wall)
    # leaves every field unchanged
    rng = np.random.RandomState(3)
    N, R, W = 12, 128, 200
    x = rng.uniform(60, 160, N)
    y = rng.uniform(60, 160, N)
    th = rng.uniform(0, 2 * PI, N)
    p = vo.VFParams(BOUNDARY="infinite")
    a = vo.step_arrays(x, y, th, np.zeros(N), 5, R, (-PI, PI), W, W, 30, p)
    b = vo.step_arrays(x + 16, y - 8, th, np.zeros(N), 5, R, (-PI, PI), W, W, 30, p)
    assert np.array_equal(a[4], b[4])


def test_sequential_order_differs_from_synchronous():
    # agent 1 sees agent 0's new pose in the reference order
    x = np.array([100.0, 112.0])
    y = np.array([100.0, 100.0])
    th = np.array([0.0, PI])
    v = np.array([3.0, 0.0])
    p = vo.VFParams(BOUNDARY="walls")
    seq = vo.step_arrays(x, y, th, v, 5, 256, (-PI, PI), 400, 400, 30, p, order="sequential")
    syn = vo.step_arrays(x, y, th, v, 5, 256, (-PI, PI), 400, 400, 30, p, order="synchronous")
    assert np.array_equal(seq[4][0], syn[4][0])
    assert not np.array_equal(seq[4][1], syn[4][1])


# ------------------------------------------- aggregation (vf_agent.py:216-237)
def _rec(aid, pos, th=0.0, R=16):
    return vo.VFAgentRec(id=aid, radius=1, position=np.array(pos, dtype=float), orientation=th,
                         velocity=0.0, v_field_res=R, FOV=(-PI, PI), WIDTH=400, HEIGHT=400,
                         window_pad=30, boundary_cond="walls")


def test_aggregation_is_an_or_and_filters_by_id():
    # Focal id 0 at (10, 10), heading 0, radius 1, R = 16 (phis step 2pi/15).
    #  * id 1 at (10, 12): v2 = (0, 2), angle +pi/2 -> ca = -pi/2 -> t = 4;
    #    vis = 2 atan(1/2) -> s = 2.36, floor(s/2) = 1 -> rays [3, 5).
    #  * id 2 at angle +2.0, distance 2: ca = -2.0 -> t = 3 -> rays [2, 4).
    #    Rays 3 is covered twice: the sum is binarised, so it stays 1.
    #  * id 3 at the focal's own position: skipped by coordinate equality
    #    (vf_supcalc.py:57-58).
    #  * a second object with id 0 at (10, 8) would draw rays [10, 12)
    #    (ca = +pi/2 -> t = 11) but the id filter (vf_agent.py:220) drops it.
    # Raw rays {2, 3, 4}; soc_v_field is in the flipped order -> {13, 12, 11}.
    focal = _rec(0, (10.0, 10.0))
    others = [_rec(1, (10.0, 12.0)), _rec(2, (10.0 + 2 * math.cos(2.0), 10.0 + 2 * math.sin(2.0))),
              _rec(3, (10.0, 10.0)), _rec(0, (10.0, 8.0))]
    f = vo.calc_soc_v_proj(focal, [focal] + others)
    assert set(np.flatnonzero(f)) == {11, 12, 13}
    assert f.max() == 1.0
    # the same object under another id is drawn
    others[3] = _rec(4, (10.0, 8.0))
    f = vo.calc_soc_v_proj(focal, [focal] + others)
    assert set(np.flatnonzero(f)) == {11, 12, 13} | {15 - 10, 15 - 11}
