"""Pins of the foraging oracle (oracle/forage_oracle.py).

Expected values are worked out by hand from the reference's text (the
comments give the derivation), or come from the reference's own known
answers; none is produced by the oracle itself."""
import math

import numpy as np
import pytest

from oracle import forage_oracle as fo

PI = np.pi


def agent(pos=(10.0, 10.0), th=0.0, R=16, fov=(-PI, PI), vis_ex=False, patchwise=True,
          vision_range=1e9, radius=1, params=None, aid=0):
    return fo.ForageAgent(aid, radius, pos, th, R, fov, 500, 500, 30, vision_range, vis_ex,
                          patchwise, params or fo.ForageParams())


class Obj:
    def __init__(self, pos, aid=99):
        self.position = np.array(pos, dtype=np.float64)
        self.id = aid


def out_bits(field):
    return set(np.flatnonzero(field))


# ------------------------------------------------------------ the field
def test_reference_kat_input_is_invisible_to_foragers():
    # test_cs_supcalc.py:143-158 input: object dead ahead -> angle 0 -> the
    # strict rule 0 < a' < pi fails -> ca = 2*pi -> outside (-pi, pi) -> zeros
    ag = agent(pos=(-1.0, -1.0), R=8)
    assert not fo.projection_field(ag, [Obj((0.0, -1.0))]).any()


def test_object_straight_behind_is_invisible():
    # a' = pi -> ca = 2*pi - pi = pi -> fails the strict FOV test
    assert not fo.projection_field(agent(), [Obj((8.0, 10.0))]).any()


def test_truncated_half_width_side_object():
    # v2 = (0, 2): angle +pi/2 -> ca = -pi/2 -> t = 4 (R=16); vis = 2 atan(1/2),
    # ps/2 = 1.18 -> start = int(2.82) = 2, end = int(5.18) = 5 -> rays {2,3,4}
    # -> flipped output {13, 12, 11} (VF's floor rule would give two rays)
    assert out_bits(fo.projection_field(agent(), [Obj((10.0, 12.0))])) == {11, 12, 13}


def test_start_truncates_toward_zero():
    # angle 2.7 with cross > 0 -> ca = -2.7 -> t = round(1.054) = 1; ps/2 = 1.18
    # -> start = int(-0.18) = 0 (no wrap), end = int(2.18) = 2 -> rays {0, 1}
    # -> output {15, 14}
    a = 2.7
    pos = (10.0 + 2 * math.cos(a), 10.0 + 2 * math.sin(a))
    assert out_bits(fo.projection_field(agent(), [Obj(pos)])) == {14, 15}


def test_limited_fov_masks_after_flip():
    # side object at ca = -pi/2 with FOV (-pi/4, pi/4): the record fails
    # fov0 < ca < fov1 -> nothing drawn
    ag = agent(fov=(-PI / 4, PI / 4))
    assert not fo.projection_field(ag, [Obj((10.0, 12.0))]).any()


# ------------------------------------------------------- exclusion rules
def rec(s, e, d, vis=0.1):
    return {"proj_start": s, "proj_end": e, "proj_start_ex": s, "proj_end_ex": e,
            "distance": d, "vis_angle": vis, "proj_size": e - s, "proj_size_ex": e - s,
            "phi_target": 0, "is_social_cue": True}


@pytest.mark.parametrize("f, i, exp", [
    ((3, 9), (5, 12), (3, 5)),     # rule 1: occluder start inside -> end cut to 5
    ((3, 9), (1, 5), (5, 9)),      # rule 2: occluder end inside -> start moved to 5
    ((3, 9), (2, 10), (0, 0)),     # rule 3: fully covered -> (0, 0)
    ((3, 9), (4, 6), (3, 4)),      # inner occluder: rule 1 cuts the tail, rule 2 then fails
])
def test_exclusion_single(f, i, exp):
    data = fo.exclude_source_data({0: rec(*i, d=1.0), 1: rec(*f, d=2.0)})
    assert (data[1]["proj_start_ex"], data[1]["proj_end_ex"]) == exp
    assert (data[0]["proj_start_ex"], data[0]["proj_end_ex"]) == i     # rank 0 untouched


def test_exclusion_sequential_and_ties():
    # f=[3,9) d=3; i1=[5,12) d=1 -> [3,5); i2=[1,4) d=2 -> rule 2: 3<=4<=5 -> [4,5)
    data = fo.exclude_source_data({7: rec(3, 9, 3.0), 2: rec(5, 12, 1.0), 5: rec(1, 4, 2.0)})
    assert (data[7]["proj_start_ex"], data[7]["proj_end_ex"]) == (4, 5)
    # equal distances never exclude each other
    data = fo.exclude_source_data({0: rec(0, 5, 1.0), 1: rec(2, 9, 1.0)})
    assert (data[1]["proj_start_ex"], data[1]["proj_end_ex"]) == (2, 9)


def test_raster_negative_end_counts_from_the_end():
    # after exclusion f = [-3, -2): start<0 paints [R-3, R) then start=0; the
    # main slice [0:-2] is Python's [0, R-2) -> every ray set
    R = 16
    d = {0: rec(-3, -2, 1.0)}
    assert np.all(fo.raster_records(d, R) == 1)


def test_raster_end_wrap_leaves_last_ray_to_main_slice():
    # [14, 18): end>=R paints [0, 2), end = R-1 -> main [14, 15) -> {0,1,14}
    assert out_bits(fo.raster_records({0: rec(14, 18, 1.0)}, 16)) == {0, 1, 14}


# --------------------------------------------------- decisions and modes
def test_decision_closed_form():
    p = fo.ForageParams(Eps_w=0.5, g_w=0.085, B_w=0.0, S_uw=0.01, Eps_u=3.0, g_u=0.085, B_u=0.0,
                        S_wu=0.25, T_w=0.5, T_u=0.5)
    ag = agent(params=p, R=8)
    ag.soc_v_field = np.array([1, 1, 0, 0, 0, 0, 0, 0], dtype=float)   # mean 0.25
    ag.w, ag.u, ag.I_priv = 0.6, 0.7, 0.2
    fo.update_decision_processes(ag)
    # u_p = 0.7 (> T_u), w_p = 0.6 (> T_w)
    # dw = 0.5*0.25 - 0.085*0.6 - 0.7*0.01 = 0.067 (no clamp)
    assert ag.w == 0.6 + ((0.5 * 0.25 - 0.085 * (0.6 - 0.0)) - 0.7 * 0.01)
    # du = 3*0.2 - 0.085*0.7 - 0.6*0.25 = 0.3905 -> 1.0905 clamped to u_max = 1
    assert ag.u == 1.0


def test_tr_u_uses_T_w():
    ag = agent(params=fo.ForageParams(T_w=0.5, T_u=0.9))
    ag.u = 0.6
    assert ag.tr_u()        # 0.6 > T_w even though 0.6 < T_u


def test_explore_draws_from_the_global_stream():
    # not tr_w, not tr_u -> random_walk: dv = exp_vel_max, dtheta = uniform(lo, hi)
    p = fo.ForageParams(exp_theta_min=-0.3, exp_theta_max=0.3, exp_vel_max=1.0, Eps_w=0, Eps_u=0)
    ag = agent(params=p, pos=(200.0, 200.0), th=1.0)
    np.random.seed(42)
    fo.agent_update(ag, [ag])
    assert ag.mode == "explore" and ag.overriding_mode is None
    # test_cs_supcalc.py:10-19: uniform(-0.3, 0.3) after seed 42
    assert ag.orientation == 1.0 + -0.0752759286915825
    assert ag.velocity == 1.0
    assert ag.position[0] == 200.0 + 1.0 * np.cos(ag.orientation)


def test_exploit_and_relocate_branches():
    p = fo.ForageParams(Eps_w=0, Eps_u=0, g_w=0, g_u=0, exp_stop_ratio=0.08, exp_vel_max=2.0,
                        reloc_theta_max=0.5)
    ag = agent(params=p, R=8, pos=(200.0, 200.0))
    ag.w, ag.u, ag.env_status, ag.velocity = 0.9, 0.9, 1, 1.5
    fo.agent_update(ag, [ag])                 # tw and tu on a patch -> exploit
    assert ag.mode == "exploit" and ag.get_mode() == "exploit"
    assert ag.velocity == 1.5 + (-1.5 * 0.08)
    ag2 = agent(params=p, R=8, pos=(200.0, 200.0))
    ag2.w, ag2.u, ag2.env_status, ag2.velocity = 0.9, 0.9, 0, 1.5
    fo.agent_update(ag2, [ag2])               # tw and tu off patch -> relocate, no draw
    assert ag2.mode == "relocate" and ag2.get_mode() == "relocate"
    assert ag2.velocity == 1.5 + (2.0 - 1.5)  # F_reloc_LR: v_desired - v, empty field -> theta 0


def test_reloc_left_right_difference():
    p = fo.ForageParams(reloc_theta_max=0.5)
    ag = agent(params=p, R=8)
    V = np.array([1, 1, 1, 0, 0, 0, 0, 1], dtype=float)   # left mean 3/4, right mean 1/4
    dv, th = fo.f_reloc_lr(ag, 0.3, V, 1.0)
    assert dv == 1.0 - 0.3 and th == (0.75 - 0.25) * 0.5


def test_prove_velocity_only_in_explore():
    ag = agent()
    ag.velocity, ag.max_exp_vel = 1.7, 1.2
    fo.prove_velocity(ag)
    assert ag.velocity == 1.2
    ag.velocity, ag.w = 1.7, 0.9               # get_mode() == "relocate"
    fo.prove_velocity(ag)
    assert ag.velocity == 1.7


def test_collide_mode_freezes_increments():
    ag = agent(pos=(200.0, 200.0), th=0.5)
    ag.set_mode("collide")
    ag.velocity = 1.0
    fo.agent_update(ag, [ag])
    assert ag.orientation == 0.5 and ag.velocity == 1.0 and ag.mode == "collide"


# ------------------------------------------------------ cue selection
def test_cue_selection_patchwise_and_empty_patch():
    # focal on patch 3; agent 1 exploits patch 3 (same patch -> occluder only),
    # agent 2 exploits patch 5 (social cue), agent 3 exploits with id -1 (dropped)
    p = fo.ForageParams()
    ags = [agent(pos=(100.0, 100.0), aid=i, R=64, params=p, radius=10) for i in range(4)]
    ags[0].exploited_patch_id = 3
    ags[1].position = np.array([100.0, 140.0]); ags[1].set_mode("exploit"); ags[1].exploited_patch_id = 3
    ags[2].position = np.array([100.0, 60.0]); ags[2].set_mode("exploit"); ags[2].exploited_patch_id = 5
    ags[3].position = np.array([60.0, 100.0]); ags[3].set_mode("exploit"); ags[3].exploited_patch_id = -1
    fo.calc_social_V_proj(ags[0], ags)
    only2 = fo.projection_field(ags[0], [ags[2]])
    assert np.array_equal(ags[0].soc_v_field, only2) and only2.any()


def test_notify_and_deplete():
    ag = agent()
    ag.novelty = np.zeros(4)
    fo.notify_agent(ag, 1, 7)
    assert ag.env_status == 1 and ag.exploited_patch_id == 7 and list(ag.novelty) == [1, 0, 0, 0]
    fo.notify_agent(ag, -1)
    assert ag.exploited_patch_id == -1 and list(ag.novelty) == [0, 1, 0, 0]
    r = {"quality": 0.6, "left": 1.0}
    assert fo.deplete(r, 1) == (0.6, False)
    assert fo.deplete(r, 1) == (0.4, True)


def _raw_words(g, n):
    t = np.random.RandomState()
    t.set_state(g.get_state())
    return t.randint(0, 2 ** 32, size=n, dtype=np.uint32)


def test_word_draws_follow_numpys_randint_and_uniform():
    # the placement draws (sims.py:349-359) restated over raw MT19937 outputs
    # must give numpy's own randint / uniform values and consume as many words
    bounds = [(30, 440), (0, 1), (5, 6), (-30, 470), (200, 1000), (0, 2 ** 31), (7, 2 ** 32 - 1)]
    for seed in range(20):
        g = np.random.RandomState(seed)
        g.random_sample(seed * 13)
        words = _raw_words(g, 4000)
        i = 0
        for k in range(60):
            lo, hi = bounds[(seed + k) % len(bounds)]
            v, i = fo.randint_from_words(words, i, lo, hi)
            assert v == g.randint(lo, hi), (seed, k, lo, hi)
            q, i = fo.uniform_from_words(words, i, 0.1, 0.9)
            assert q == g.uniform(0.1, 0.9)
        assert np.array_equal(_raw_words(g, 3), words[i:i + 3])   # same stream position


def _world(agents, W=300, n_resc=3, radius=30, lo=3, hi=6, **kw):
    return fo.ForageWorld(agents, W, W, 30, n_resc, radius, lo, hi, 0.5, 1.0, **kw)


def _agents(seed, N=12, W=300):
    p = fo.ForageParams(Eps_u=1.0, F_N=0.5, F_R=0.5)
    rng = np.random.RandomState(seed)
    return [fo.ForageAgent(i, 10, (rng.randint(20, W + 20), rng.randint(20, W + 20)),
                           rng.uniform(0, 2 * np.pi), 64, (-np.pi, np.pi), W, W, 30, 150.0, False,
                           True, p) for i in range(N)]


def test_world_patches_never_overlap_and_keep_ids():
    # placement rule (proove_sprite with collide_circle): integer centres more
    # than r1 + r2 apart, at creation and after every regeneration; a
    # regenerated patch keeps the killed one's id
    np.random.seed(4)
    ags = _agents(4, N=30)
    w = _world(ags, n_resc=5)
    ids0 = sorted(p.id for p in w.patches)
    assert ids0 == [1, 2, 3, 4, 5]
    for a in ags:                                   # everyone exploits
        a.u, a.env_status = 1.0, 1
        a.set_mode("exploit")
    moved = 0
    for step in range(60):
        pos = {p.id: tuple(p.position) for p in w.patches}
        w.step()
        moved += sum(pos[p.id] != tuple(p.position) for p in w.patches)
        assert sorted(p.id for p in w.patches) == ids0
        for i, p in enumerate(w.patches):
            for q in w.patches[i + 1:]:
                d2 = (p.center[0] - q.center[0]) ** 2 + (p.center[1] - q.center[1]) ** 2
                assert d2 > (p.radius + q.radius) ** 2
    assert moved > 0


def test_world_conserves_resources():
    # what the patches lose each step is what the agents collect
    np.random.seed(8)
    ags = _agents(8, N=25)
    w = _world(ags, n_resc=4, lo=20, hi=40)
    for a in ags:
        a.u, a.env_status = 1.0, 1
        a.set_mode("exploit")
    for step in range(40):
        left = {p.id: p.resc_left for p in w.patches}
        got = sum(a.collected_r for a in ags)
        w.step()
        new = {p.id: p.resc_left for p in w.patches}
        # a killed-and-regenerated patch gave all it had left
        lost = sum(left[i] - (new[i] if new[i] <= left[i] else 0.0) for i in left)
        assert abs(sum(a.collected_r for a in ags) - got - lost) <= 1e-9


def test_world_single_exploiter_known_answers():
    # one agent on one patch: 2.5 units, quality 1, consumption 1 -> it
    # collects 1, 1, 0.5; the third depletion kills the patch, the agent is
    # told -1 (twice rolled in: once as the depleting agent, once by the
    # notify-all), and the patch is regenerated with the same id
    np.random.seed(2)
    ag = _agents(2, N=1)[0]
    w = _world([ag], n_resc=1)
    p = w.patches[0]
    p.resc_units = p.resc_left = 2.5
    p.unit_per_timestep = 1.0
    ag.position = np.array([p.center[0] - 10.0, p.center[1] - 10.0])   # centre on the patch centre
    ag.u, ag.env_status, ag.velocity = 1.0, 1, 0.0
    ag.set_mode("exploit")
    ag.novelty = np.zeros(ag.novelty.shape)
    collected = []
    for step in range(3):
        before = ag.collected_r
        ag.position = np.array([w.patches[0].center[0] - 10.0, w.patches[0].center[1] - 10.0])
        ag.set_mode("exploit")
        w.step()
        collected.append(ag.collected_r - before)
    assert collected == [1.0, 1.0, 0.5]
    assert w.patches[0].id == 1 and w.patches[0].resc_left == w.patches[0].resc_units
    assert ag.env_status == -1 and ag.exploited_patch_id == -1
    assert ag.novelty[0] == 0 and ag.novelty[1] == 0
