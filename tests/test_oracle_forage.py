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


# ------------------------------------------- private information (I_priv)
def test_calc_I_priv_weights_max_novelty_and_new_units():
    # agent.py:168-175: I_priv = F_N * max(novelty) + F_R * (collected_r -
    # collected_r_before).  Two novelty entries are 1, so max (1) differs from
    # sum (2) and mean (0.2); F_N != F_R, so swapped weights show.
    ag = agent(params=fo.ForageParams(F_N=2.0, F_R=0.5))
    ag.novelty = np.array([0, 1, 0, 1, 0, 0, 0, 0, 0, 0], dtype=float)
    ag.collected_r, ag.collected_r_before = 5.5, 3.0
    fo.calc_I_priv(ag)
    assert ag.I_priv == 2.0 * 1 + 0.5 * 2.5          # 3.25
    ag.novelty[:] = 0
    fo.calc_I_priv(ag)
    assert ag.I_priv == 1.25                          # only the new units count


def test_I_priv_feeds_u_and_resets_collected_before():
    # one Agent.update: I_priv from the novelty ring and the units collected
    # since the last step, du = Eps_u*I_priv - g_u*(u - B_u) - w_p*S_wu
    # (agent.py:201), then collected_r_before = collected_r (agent.py:283)
    p = fo.ForageParams(F_N=1.0, F_R=0.25, Eps_u=0.5, g_u=0.1, B_u=0.0, Eps_w=0.0, g_w=0.0,
                        T_w=0.5, T_u=0.5, u_max=10.0)
    ag = agent(params=p, pos=(200.0, 200.0))
    ag.novelty = np.array([0, 0, 1, 0, 0, 0, 0, 0, 0, 0], dtype=float)
    ag.collected_r, ag.collected_r_before = 4.0, 2.0
    ag.u, ag.w = 0.2, 0.0
    ag.set_mode("collide")                            # no draw, no motion
    fo.agent_update(ag, [ag])
    assert ag.I_priv == 1.0 + 0.25 * 2.0             # 1.5
    # w_p = 0 (w = 0 < T_w): du = 0.5*1.5 - 0.1*0.2 = 0.73
    assert ag.u == 0.2 + (0.5 * 1.5 - 0.1 * (0.2 - 0.0) - 0 * p.S_wu)
    assert math.isclose(ag.u, 0.93, rel_tol=1e-15)
    assert ag.collected_r_before == 4.0


# ------------------------------------------- keep_distance_info amplitude
def _kd_agent(vision_range):
    return agent(pos=(10.0, 10.0), th=0.0, R=16, vision_range=vision_range)


def test_keep_distance_uses_the_last_loop_distance():
    # agent.py:587-590 weighs every main slice with 1 - distance/vision_range,
    # where `distance` is the loop variable left by the LAST obstacle the loop
    # measured (agent.py:528), visible or not.  Obstacle A at (10, 12): d = 2,
    # rays [2, 5) -> output {11, 12, 13} (see test_truncated_half_width_side_object).
    # Obstacle B at (6, 10) is straight behind (a' = pi -> invisible), d = 4.
    # Order [A, B]: amplitude 1 - 4/10 = 0.6; order [B, A]: 1 - 2/10 = 0.8.
    a, b = Obj((10.0, 12.0)), Obj((6.0, 10.0))
    f = fo.projection_field(_kd_agent(10.0), [a, b], keep_distance_info=True)
    assert out_bits(f) == {11, 12, 13} and np.all(f[[11, 12, 13]] == 1 - 4.0 / 10.0)
    f = fo.projection_field(_kd_agent(10.0), [b, a], keep_distance_info=True)
    assert out_bits(f) == {11, 12, 13} and np.all(f[[11, 12, 13]] == 1 - 2.0 / 10.0)


def test_keep_distance_leaves_the_wrapped_piece_at_one():
    # angle 3.1 with cross > 0 -> ca = -3.1 -> t = 0 (phis[0] = -pi is nearest);
    # d = 2 -> ps/2 = 1.18 -> start = int(-1.18) = -1, end = int(1.18) = 1.
    # start < 0 paints V[15:16] = 1 (not the amplitude, agent.py:579-581),
    # then the main slice V[0:1] = 1 - d/vision_range.  Flipped: out[0] = 1,
    # out[15] = 1 - d/10.
    a = 3.1
    ob = Obj((10.0 + 2 * math.cos(a), 10.0 + 2 * math.sin(a)))
    f = fo.projection_field(_kd_agent(10.0), [ob], keep_distance_info=True)
    assert out_bits(f) == {0, 15}
    assert f[0] == 1.0
    assert math.isclose(f[15], 0.8, rel_tol=0, abs_tol=1e-15)


# ------------------------------------------- vision range (own radii)
@pytest.mark.parametrize("x,visible", [(58.0, True), (58.5, False), (47.0, True)])
def test_vision_range_uses_each_agents_own_radius(x, visible):
    # supcalc.distance (supcalc.py:73-78) measures centres with each agent's
    # OWN radius: focal radius 10 at (0, 0) -> centre (10, 10); cue radius 2
    # at (x, 8) -> centre (x + 2, 10): d = x - 8.  vision_range 50 and the
    # test is <= (agent.py:400): x = 58 -> d = 50 exactly -> in range;
    # x = 58.5 -> out.  With the focal radius for both centres d would be x
    # (58 -> out).  The cue exploits patch 5, the focal none.
    ags = [agent(pos=(0.0, 0.0), aid=0, R=64, radius=10, vision_range=50.0),
           agent(pos=(x, 8.0), aid=1, R=64, radius=2, vision_range=50.0)]
    ags[1].set_mode("exploit")
    ags[1].exploited_patch_id = 5
    fo.calc_social_V_proj(ags[0], ags)
    assert bool(ags[0].soc_v_field.any()) == visible


# ------------------------------------------- resource loop branches
def _one_patch_world(ag_list, **kw):
    np.random.seed(3)
    w = _world(ag_list, n_resc=1, lo=50, hi=60, **kw)
    p = w.patches[0]
    p.position = np.array([200.0, 200.0])
    p.center = (230.0, 230.0)
    return w, p


def test_pooling_time_notifies_only_pooled_relocators():
    # sims.py:816-822, 847-853 with pooling_time > 0: on a patch only agents
    # in "pool"/"relocate" with pool_success are told 1 (and the patch id);
    # off the patch those are told -1, and so are exploiters; an explorer is
    # left alone.  pool_success and the env status are touched only by
    # notify_agent, which the update after the loop never calls.
    ags = _agents(5, N=6)
    for a in ags:
        a.pooling_time = 2
        a.is_moved_with_cursor = 1
        a.novelty = np.zeros(a.novelty.shape)
    w, p = _one_patch_world(ags)
    on = lambda a: a.__setattr__("position", np.array([225.0, 222.0]))   # centre (235, 232)
    # 0: relocating, pooled -> told 1 on the patch
    on(ags[0]); ags[0].w = 0.9; ags[0].pool_success = 1
    # 1: relocating, not pooled -> nothing
    on(ags[1]); ags[1].w = 0.9; ags[1].pool_success = 0
    # 2: exploring on the patch -> nothing
    on(ags[2]); ags[2].w = 0.0; ags[2].pool_success = 1
    # 3: off the patch, relocating and pooled -> told -1
    ags[3].position = np.array([40.0, 40.0]); ags[3].w = 0.9; ags[3].pool_success = 1
    # 4: off the patch, exploiting -> told -1
    ags[4].position = np.array([40.0, 80.0]); ags[4].set_mode("exploit"); ags[4].exploited_patch_id = 1
    ags[4].env_status = 1
    # 5: off the patch, exploring -> nothing
    ags[5].position = np.array([80.0, 40.0]); ags[5].w = 0.0; ags[5].pool_success = 0
    ags[5].env_status = 1
    w.step()
    assert (ags[0].env_status, ags[0].exploited_patch_id, ags[0].novelty[0]) == (1, p.id, 1)
    assert (ags[1].env_status, ags[1].exploited_patch_id, ags[1].pool_success) == (0, -1, 0)
    assert (ags[2].env_status, ags[2].novelty[0]) == (0, 0)
    assert (ags[3].env_status, ags[3].exploited_patch_id) == (-1, -1)
    assert (ags[4].env_status, ags[4].exploited_patch_id) == (-1, -1)
    assert ags[5].env_status == 1 and ags[5].pool_success == 0


def test_teleport_and_bias_towards_patch_centre():
    # sims.py:544-552: cl = (atan2(dy, dx) + theta) % 2pi with (dx, dy) from
    # the agent centre to the patch centre, theta += (cl - pi) * 0.02; then
    # notify and, with TELEPORT_TO_MIDDLE, position = patch position + patch
    # radius - agent radius (sims.py:823-824).
    # Agent centre 5 px left of the patch centre: dx = 5, dy = 0, atan2 = 0;
    # theta = 0 -> cl = 0 -> theta = -0.02 pi.  A second agent 5 px below
    # (dy = -5, atan2 = -pi/2) heading pi/2: cl = 0 -> pi/2 - 0.02 pi; a third
    # 5 px above (dy = 5, atan2 = pi/2) heading pi/2: cl = pi -> no turn.
    ags = _agents(6, N=3)
    for a in ags:
        a.is_moved_with_cursor = 1
    w, p = _one_patch_world(ags, teleport_exploit=True)
    ags[0].position = np.array([215.0, 220.0]); ags[0].orientation = 0.0
    ags[1].position = np.array([220.0, 225.0]); ags[1].orientation = PI / 2
    ags[2].position = np.array([220.0, 215.0]); ags[2].orientation = PI / 2
    w.step()
    assert ags[0].orientation == (0.0 - PI) * 0.02
    assert ags[1].orientation == PI / 2 + (0.0 - PI) * 0.02
    assert ags[2].orientation == PI / 2
    for a in ags:
        assert np.array_equal(a.position, [220.0, 220.0])      # 200 + 30 - 10
        assert a.env_status == 1 and a.exploited_patch_id == p.id


@pytest.mark.parametrize("overlap", [False, True])
def test_patch_placement_ranges_and_draw_order(overlap):
    # sims.py:349-359: x, y = randint over [pad - r, W + pad - r) with border
    # overlap, else [pad, W + pad - 2r); then units = randint(min, max) and
    # quality = uniform(min_q, max_q), in that order, from the global stream.
    W, r, pad = 300, 30, 30
    xs = []
    for seed in range(40):
        np.random.seed(seed)
        w = fo.ForageWorld([], W, W, pad, 1, r, 5, 9, 0.2, 0.7, allow_border_overlap=overlap)
        g = np.random.RandomState(seed)
        lo, hi = (pad - r, W + pad - r) if overlap else (pad, W + pad - 2 * r)
        x, y = g.randint(lo, hi), g.randint(lo, hi)
        units, q = g.randint(5, 9), g.uniform(0.2, 0.7)
        p = w.patches[0]
        assert (p.position[0], p.position[1], p.resc_units, p.unit_per_timestep) == (x, y, units, q)
        xs += [x, y]
    if overlap:
        assert min(xs) < pad           # reaches over the left/top border
    else:
        assert min(xs) >= pad and max(xs) <= W + pad - 2 * r - 1


# ------------------------------------------- wall reflection (8 branches)
def _wall_agent(pos, th):
    a = agent(pos=pos, th=th, radius=1)
    return a


# (position, heading, expected heading, expected position).  The arena is
# [30, 530]^2 with radius 1.  Each expected heading is the mirror image of
# the heading at that wall, except at the branch edges where the reference's
# quadrant bounds decide (agent.py:356-394).
WALL_CASES = [
    # left wall: pi/2 <= th < pi -> th - pi/2; pi <= th <= 3pi/2 -> th + pi/2
    ((20.0, 200.0), 3 * PI / 4, PI / 4, (29.0, 200.0)),
    ((20.0, 200.0), 5 * PI / 4, 7 * PI / 4, (29.0, 200.0)),
    ((20.0, 200.0), PI, 3 * PI / 2, (29.0, 200.0)),           # edge: second branch
    ((20.0, 200.0), PI / 4, PI / 4, (29.0, 200.0)),           # moving away: unchanged
    # right wall: 3pi/2 <= th < 2pi -> th - pi/2; 0 <= th <= pi/2 -> th + pi/2
    ((540.0, 200.0), 7 * PI / 4, 5 * PI / 4, (528.0, 200.0)),
    ((540.0, 200.0), PI / 4, 3 * PI / 4, (528.0, 200.0)),
    ((540.0, 200.0), PI / 2, PI, (528.0, 200.0)),             # edge: second branch
    # top wall: pi/2 <= th <= pi -> th + pi/2; 0 <= th < pi/2 -> th - pi/2 (+2pi)
    ((200.0, 20.0), 3 * PI / 4, 5 * PI / 4, (200.0, 29.0)),
    ((200.0, 20.0), PI / 4, 7 * PI / 4, (200.0, 29.0)),
    ((200.0, 20.0), PI / 2, PI, (200.0, 29.0)),               # edge: first branch
    # bottom wall: 3pi/2 <= th <= 2pi -> th + pi/2 (-2pi); pi <= th < 3pi/2 -> th - pi/2
    ((200.0, 540.0), 7 * PI / 4, PI / 4, (200.0, 528.0)),
    ((200.0, 540.0), 5 * PI / 4, 3 * PI / 4, (200.0, 528.0)),
    ((200.0, 540.0), 3 * PI / 2, 2 * PI, (200.0, 528.0)),     # edge: 2pi is kept (not > 2pi)
]


@pytest.mark.parametrize("pos,th,exp_th,exp_pos", WALL_CASES)
def test_reflect_from_walls_branches(pos, th, exp_th, exp_pos):
    a = _wall_agent(pos, th)
    fo.reflect_from_walls(a)
    assert tuple(a.position) == exp_pos
    assert math.isclose(a.orientation, exp_th, rel_tol=0, abs_tol=1e-12)


def test_reflect_corner_applies_both_walls():
    # left and top at once, heading up-left (3pi/4): the left wall turns it to
    # pi/4 (up-right), then the top wall (0 <= pi/4 < pi/2) to -pi/4 -> 7pi/4
    a = _wall_agent((20.0, 20.0), 3 * PI / 4)
    fo.reflect_from_walls(a)
    assert tuple(a.position) == (29.0, 29.0)
    assert math.isclose(a.orientation, 7 * PI / 4, abs_tol=1e-12)


def test_bias_multiplies_heading_differences_by_1_02_per_step():
    # bias_agent_towards_res_center (sims.py:544-552) is
    #   theta' = theta + ((atan2(dy, dx) + theta) mod 2pi - pi) * 0.02,
    # so d theta' / d theta = 1.02 away from the mod wrap: a heading
    # difference of an agent that stays on a patch grows by 2 % per step.
    # This is why a last-bit difference of a transcendental function
    # between two hosts eventually changes discrete state in long runs.
    # Two copies 1e-9 rad apart, agent parked on the patch (not moved by the
    # update), 50 steps: the difference is 1e-9 * 1.02^50.
    thetas = []
    for d in (0.0, 1e-9):
        ags = _agents(6, N=1)
        ags[0].is_moved_with_cursor = 1
        w, p = _one_patch_world(ags)
        ags[0].position = np.array([215.0, 220.0])    # centre 5 px left of the patch centre
        # atan2(0, 5) = 0: the fixed point is theta = pi; start 0.1 above it
        ags[0].orientation = PI + 0.1 + d
        for _ in range(50):
            w.step()
        thetas.append(ags[0].orientation)
    assert math.isclose(thetas[1] - thetas[0], 1e-9 * 1.02 ** 50, rel_tol=1e-5)
    assert math.isclose(thetas[0] - PI, 0.1 * 1.02 ** 50, rel_tol=1e-9)


# ------------------------------------------------ agent-agent collisions
def _collide_world(positions, thetas=None, R=1200, **kw):
    p = fo.ForageParams()
    thetas = thetas or [0.0] * len(positions)
    ags = [fo.ForageAgent(i, 10, pos, th, R, (-PI, PI), 500, 500, 30, 150.0, False, True, p)
           for i, (pos, th) in enumerate(zip(positions, thetas))]
    for a in ags:
        a.velocity = 0.5
    w = fo.ForageWorld(ags, 500, 500, 30, 0, 30, 3, 6, 0.5, 1.0, collide_agents=True, **kw)
    return w, ags


@pytest.mark.parametrize("p1,hit", [((124.0, 100.0), True), ((124.999, 100.0), True),
                                    ((125.0, 100.0), False), ((116.0, 117.0), True),
                                    ((117.0, 117.0), False), ((100.0, 76.0), True)])
def test_collisions_use_truncated_rect_centres_and_radii_grown_by_2(p1, hit):
    # sims.py:738-747 + pygame collide_circle: rect centres int(position) + r,
    # radii 10 + 2 each: squared distance <= 24^2 = 576 (17^2 + 17^2 = 578 > 576)
    w, (a, b) = _collide_world([(100.0, 100.0), p1])
    w.collide()
    assert (a.get_mode() == "collide") == hit and (b.get_mode() == "collide") == hit
    assert (a.env_status == -1) == hit   # collided agents in collide mode are told -1


def test_object_dead_ahead_is_invisible_so_the_tie_turns_by_plus_0_2():
    # the other agent 24 px straight ahead: angle 0 makes the closed angle
    # 2 pi (agent.py:520-523), outside the strict FOV test, so the field is
    # empty: left and right excitations tie, np.sign gives 0, taken as -1
    # (sims.py:449-450): theta += 0.2; the front is clear: velocity =
    # max_exp_vel
    w, (a, b) = _collide_world([(100.0, 100.0), (124.0, 100.0)])
    w.collide()
    assert a.orientation == 0.2 and a.velocity == a.max_exp_vel


def test_object_ahead_off_axis_turns_away_and_stops():
    # agent 1 at (124, 99): rect centres 24 and 1 apart, 577 > 576, no
    # collision of its own, but 24.02 px away (inside 2 r + 20); agent 2 right
    # behind (invisible, closed angle pi) collides with agent 0.  Agent 1's
    # closed angle 0.0416 -> bin 607, proj_size 150.7 -> rays [531, 682),
    # flipped 518..668: 82 rays left of 600, 69 right: D = +1, theta -= 0.2;
    # the front window [500, 700) is covered: velocity 0
    w, ags = _collide_world([(100.0, 100.0), (124.0, 99.0), (76.0, 100.0)])
    w.collide()
    assert ags[0].orientation == -0.2 and ags[0].velocity == 0
    assert ags[1].get_mode() != "collide"


@pytest.mark.parametrize("p1,dth,stop", [((117.0, 83.0), -0.2, True),   # ahead, screen up: left half
                                         ((117.0, 117.0), 0.2, True),   # mirror image: right half
                                         ((83.0, 83.0), -0.2, False)])  # behind, up: left half, front clear
def test_proximity_turn_sign_and_front_window(p1, dth, stop):
    # the other agent 17 px off in x and y (centres 24.04 px apart, inside
    # 2 r + 20 = 40, rect centres 17, 17 -> 578 > 576: no collision of its
    # own), so agent 0 collides with agent 2 straight behind it at 24 px and
    # takes one response with agents 1 and 2 in its field
    w, ags = _collide_world([(100.0, 100.0), p1, (76.0, 100.0)])
    w.collide()
    a = ags[0]
    # agent 2 behind (closed angle pi) is outside the strict FOV test, so the
    # field holds agent 1 only: +pi/4 (up) -> flipped rays 376..526 (left
    # half, front window reached), 3 pi/4 (behind, up) -> rays 76..226
    assert a.orientation == pytest.approx(dth, abs=1e-15)
    assert (a.velocity == 0) == stop
    if not stop:
        assert a.velocity == a.max_exp_vel


def test_two_partners_give_two_responses():
    # agent 0 collides with agent 1 straight ahead and agent 2 straight
    # behind, both invisible: two responses (one per groupcollide partner),
    # each a tie turning by +0.2
    w, ags = _collide_world([(100.0, 100.0), (124.0, 100.0), (76.0, 100.0)])
    w.collide()
    assert ags[0].orientation == pytest.approx(0.4, abs=1e-15)
    assert ags[1].orientation == 0.2 and ags[2].orientation == 0.2


@pytest.mark.parametrize("ghost,teleport", [(True, False), (False, False), (True, True)])
def test_ghost_mode_and_collided_list(ghost, teleport):
    # agent 1 exploits (sims.py:431-436, 760-784).  Ghost mode: no response
    # for the pair; agent 0 is collided only with teleport (that branch adds
    # every non-exploiting agent), so its old collide mode ends (explore)
    # without it and is told -1 with it.  Without ghost mode: agent 0 takes
    # the (tie) response and is told -1; the exploiter keeps its heading and
    # mode, and with the front clear its speed too.
    w, (a, b) = _collide_world([(100.0, 100.0), (124.0, 100.0)], ghost_mode=ghost, teleport_exploit=teleport)
    a.set_mode("collide")
    b.set_mode("exploit")
    w.collide()
    if ghost:
        assert a.orientation == 0.0
        if teleport:
            assert a.get_mode() == "collide" and a.env_status == -1
        else:
            assert a.get_mode() == "explore" and a.env_status == 0
    else:
        assert a.get_mode() == "collide" and a.orientation == 0.2 and a.env_status == -1
        assert b.get_mode() == "exploit" and b.orientation == 0.0 and b.velocity == 0.5
    assert b.env_status == 0
