"""The drop-in classes driven like the reference's host loops, in lockstep
with the oracle: every step compares agent state, fields, modes, patches and
the position of numpy's global random stream."""
import numpy as np
import pytest

from oracle import forage_oracle as fo
from oracle import vf_oracle as vo

pytestmark = pytest.mark.gpu
PI = np.pi


def rel(a, b, s=1.0):
    return float(np.max(np.abs(np.asarray(a, float) - np.asarray(b, float)) /
                        np.maximum(np.abs(np.asarray(b, float)), s)))


def same_stream(a, b):
    return a[2] == b[2] and np.array_equal(a[1], b[1])


def test_vf_simulation_lockstep(gpu):
    from abm_amd.dropin import VFParams, VFSimulation
    p = VFParams(GAM=0.1, V0=1.0, ALP0=1.0, ALP1=0.09, BET0=0.5, BET1=0.09, BOUNDARY="walls")
    np.random.seed(42)
    sim = VFSimulation(N=25, T=30, v_field_res=256, width=400, height=400, agent_radius=5,
                       agent_fov=0.5, vf_params=p)
    sim.create_agents()
    ags = list(sim.agents)
    assert sim.v_field_res == 512
    ags[3].ALP0 = 2.5                       # per-agent override (vf_agent.py:24-26)
    ags[5].verbose_supcalc = True
    op = vo.VFParams(GAM=p.GAM, V0=p.V0, ALP0=p.ALP0, ALP1=p.ALP1, BET0=p.BET0, BET1=p.BET1,
                     BOUNDARY="walls")
    orc = vo.make_agents([a.position[0] for a in ags], [a.position[1] for a in ags],
                         [a.orientation for a in ags], np.zeros(25), 5, 512, sim.agent_fov, 400,
                         400, 30, op)
    orc[3].ALP0 = 2.5
    orc[5].verbose_supcalc = True
    for step in range(30):
        sim.step_sim()
        vo.group_update(orc, op)
        for a, o in zip(ags, orc):
            assert np.array_equal(a.soc_v_field, o.soc_v_field), (step, a.id)
            assert rel(a.position, o.position) < 1e-9 and abs(a.orientation - o.orientation) < 1e-9
            assert abs(a.velocity - o.velocity) < 1e-9
        assert np.allclose([ags[5].dv, ags[5].dphi, ags[5].ablob, ags[5].aedge, ags[5].bblob,
                            ags[5].bedge], orc[5].diag, rtol=1e-9, atol=1e-12)
    assert isinstance(ags[0].position, list)
    # the paused branch recomputes fields through the projection probe
    sim.is_paused = True
    sim.step_sim()
    for a, o in zip(ags, orc):
        assert np.array_equal(a.soc_v_field, vo.calc_soc_v_proj(o, orc))


@pytest.mark.parametrize("vis_ex,collide", [(False, False), (True, False), (False, True)])
def test_forage_simulation_lockstep(gpu, vis_ex, collide):
    # C1 shape: N=10, r=10, R=1200, 10 patches of radius 30, quality U[0.5, 1],
    # 500x500 walls, vision range 2000; .env decision/movement values.  With
    # agent-agent collisions on (sims.py:736-787) the arena is 150x150 with 3
    # patches, so that agents collide every few steps.
    from abm_amd.dropin import DecisionParams, MovementParams, Simulation
    env = {"DEC_TW": 0.5, "DEC_EPSW": 3, "DEC_GW": 0.085, "DEC_BW": 0, "DEC_WMAX": 1, "DEC_TU": 0.5,
           "DEC_EPSU": 1, "DEC_GU": 0.085, "DEC_BU": 0, "DEC_UMAX": 1, "DEC_SWU": 0.25,
           "DEC_SUW": 0.01, "DEC_TAU": 10, "DEC_FN": 0.5, "DEC_FR": 0.5, "MOV_EXP_VEL_MAX": 3,
           "MOV_EXP_TH_MIN": -0.5, "MOV_EXP_TH_MAX": 0.5, "MOV_REL_TH_MAX": 1.8,
           "CONS_STOP_RATIO": 0.175}
    env = {k: str(v) for k, v in env.items()}
    dp, mp = DecisionParams.from_env(env), MovementParams.from_env(env)
    np.random.seed(3 + vis_ex + 2 * collide)
    side, n_resc = (150, 3) if collide else (500, 10)
    sim = Simulation(N=10, T=0, v_field_res=1200, width=side, height=side, pooling_time=0,
                     agent_radius=10, N_resc=n_resc, min_resc_perpatch=100, max_resc_perpatch=-1,
                     min_resc_quality=0.5, max_resc_quality=1.0, patch_radius=30,
                     teleport_exploit=False, vision_range=2000, agent_fov=1.0,
                     visual_exclusion=vis_ex, collide_agents=collide, decision_params=dp,
                     movement_params=mp)
    sim.create_agents()
    sim.create_resources()
    ags = sim.agents.sprites()
    op = fo.ForageParams(T_w=dp.T_w, Eps_w=dp.Eps_w, g_w=dp.g_w, B_w=dp.B_w, w_max=dp.w_max,
                         T_u=dp.T_u, Eps_u=dp.Eps_u, g_u=dp.g_u, B_u=dp.B_u, u_max=dp.u_max,
                         S_wu=dp.S_wu, S_uw=dp.S_uw, Tau=dp.Tau, F_N=dp.F_N, F_R=dp.F_R,
                         exp_vel_max=mp.exp_vel_max, exp_theta_min=mp.exp_theta_min,
                         exp_theta_max=mp.exp_theta_max, reloc_theta_max=mp.reloc_theta_max,
                         exp_stop_ratio=mp.exp_stop_ratio)
    orc = [fo.ForageAgent(a.id, a.radius, a.position.copy(), a.orientation, 1200, sim.agent_fov,
                          side, side, 30, 2000, vis_ex, True, op) for a in ags]
    world = fo.ForageWorld.__new__(fo.ForageWorld)
    world.agents, world.W, world.H, world.pad = orc, side, side, 30
    world.collide_agents, world.ghost_mode = collide, True
    world.resc_radius, world.min_units, world.max_units = 30, 100, 101
    world.min_quality, world.max_quality = 0.5, 1.0
    world.regenerate, world.teleport_exploit, world.allow_border_overlap = True, False, False
    world.consumption = 1
    world.patches = [fo.Patch(r.id, r.radius, r.position.copy(), r.resc_units, r.unit_per_timestep)
                     for r in sim.rescources]
    modes_seen, collected = set(), 0.0
    for step in range(150):
        g0 = np.random.get_state()
        world.step()
        g_o = np.random.get_state()
        np.random.set_state(g0)
        sim.step()
        assert same_stream(np.random.get_state(), g_o), step
        for a, o in zip(ags, orc):
            assert a.mode == o.mode and a.overriding_mode == o.overriding_mode, (step, a.id)
            assert a.exploited_patch_id == o.exploited_patch_id and a.env_status == o.env_status
            assert np.array_equal(a.soc_v_field, o.soc_v_field), (step, a.id)
            assert a.collected_r == o.collected_r
            for att in ("orientation", "velocity", "w", "u", "I_priv"):
                assert abs(getattr(a, att) - getattr(o, att)) <= 1e-9 * max(1.0, abs(getattr(o, att))), att
            assert rel(a.position, o.position) <= 1e-9
            modes_seen.add(a.mode)
        assert [(r.id, r.resc_left) for r in sim.rescources] == [(p.id, p.resc_left) for p in world.patches]
        collected = sum(a.collected_r for a in ags)
    assert {"explore", "exploit"} <= modes_seen and collected > 0
    if collide:
        assert "collide" in modes_seen
