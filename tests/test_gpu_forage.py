"""GPU parity of the foraging step (abm_forage_step_f64) and the
Agent.projection_field probe against the oracle.

Bar: fields, modes, overriding modes and the numpy RNG stream position
identical; continuous state (x, y, theta, v, w, u, I_priv) within 1e-12
relative after each step from identical state."""
import numpy as np
import pytest

from oracle import forage_oracle as fo

pytestmark = pytest.mark.gpu
PI = np.pi


def rel(a, b, s=1.0):
    return float(np.max(np.abs(np.asarray(a) - np.asarray(b)) / np.maximum(np.abs(b), s)))


class Obs:
    def __init__(self, p):
        self.position = np.array(p, dtype=np.float64)


def test_projection_probe_matches_oracle(gpu):
    from abm_amd import forage
    rng = np.random.RandomState(4)
    for trial in range(60):
        R = int(rng.choice([8, 16, 64, 200, 1200]))
        fov = (-PI, PI) if trial % 3 else (-0.5 * PI, 0.5 * PI)
        ag = fo.ForageAgent(0, int(rng.randint(3, 12)), rng.uniform(40, 160, 2), rng.uniform(0, 2 * PI),
                            R, fov, 200, 200, 30, 90.0, bool(trial % 2), True, fo.ForageParams())
        n_soc = int(rng.randint(0, 9))
        n_non = int(rng.randint(0, 6))
        soc = [Obs(rng.uniform(30, 170, 2)) for _ in range(n_soc)]
        if trial % 7 == 0 and n_soc > 1:
            soc[1].position = soc[0].position.copy()       # coincident obstacles
        if trial % 5 == 0:
            soc.append(Obs(ag.position.copy()))               # coincident with the focal
        non = [Obs(rng.uniform(30, 170, 2)) for _ in range(n_non)] if trial % 4 else None
        keep = trial % 6 == 1
        exp = fo.projection_field(ag, soc, keep_distance_info=keep, non_expl_agents=non)
        got = forage.projection_field(ag, soc, keep_distance_info=keep, non_expl_agents=non)
        assert np.array_equal(got, exp), (trial, R, np.flatnonzero(got != exp)[:10])
    # the reference's test input is invisible to a forager (dead ahead)
    ag = fo.ForageAgent(0, 1, np.array([-1.0, -1.0]), 0, 8, (-PI, PI), 200, 200, 30, 1e9, False,
                        True, fo.ForageParams())
    assert not forage.projection_field(ag, [Obs([0.0, -1.0])]).any()


def make_world(seed, N=10, R=1200, W=500, vis_ex=False, vision=2000.0, mixed=True, params=None):
    """Simulation.create_agents-style start (sims.py:526-532) plus a mixed
    decision state so that every branch of the mode table is exercised."""
    p = params or fo.ForageParams(Eps_w=3.0, Eps_u=1.0, S_wu=0.25, S_uw=0.01, F_N=0.5, F_R=0.5,
                                  exp_vel_max=3.0, exp_theta_min=-0.5, exp_theta_max=0.5,
                                  reloc_theta_max=1.8, exp_stop_ratio=0.175)
    rng = np.random.RandomState(seed)
    ags = []
    for i in range(N):
        x = rng.randint(30 - 10, W + 30 - 10)
        y = rng.randint(30 - 10, W + 30 - 10)
        th = rng.uniform(0, 2 * PI)
        ags.append(fo.ForageAgent(i, 10, (x, y), th, R, (-PI, PI), W, W, 30, vision, vis_ex, True, p))
    if mixed:
        for a in ags:
            a.w = float(rng.uniform(-0.2, 1.0))
            a.u = float(rng.uniform(-0.2, 1.0))
            a.env_status = int(rng.choice([-1, 0, 1]))
            a.exploited_patch_id = int(rng.choice([-1, 0, 1, 2]))
            a.velocity = float(rng.uniform(0, 3))
            a.collected_r = float(rng.uniform(0, 5))
            a.collected_r_before = a.collected_r - float(rng.choice([0.0, 0.5]))
            a.novelty = (rng.rand(p.Tau) < 0.2).astype(np.float64)
            m = rng.choice(["explore", "exploit", "relocate", "collide"], p=[0.4, 0.35, 0.15, 0.1])
            a.set_mode(m)
    cfg_kw = dict(width=W, height=W, window_pad=30, fov=(-PI, PI), vision_range=vision,
                  visual_exclusion=vis_ex, patchwise_exclusion=True, T_w=p.T_w, Eps_w=p.Eps_w,
                  g_w=p.g_w, B_w=p.B_w, w_max=p.w_max, T_u=p.T_u, Eps_u=p.Eps_u, g_u=p.g_u,
                  B_u=p.B_u, u_max=p.u_max, S_wu=p.S_wu, S_uw=p.S_uw, Tau=p.Tau, F_N=p.F_N,
                  F_R=p.F_R, exp_vel_max=p.exp_vel_max, exp_theta_min=p.exp_theta_min,
                  exp_theta_max=p.exp_theta_max, reloc_theta_max=p.reloc_theta_max,
                  exp_stop_ratio=p.exp_stop_ratio)
    return ags, cfg_kw


def load_swarm(sw, ags, sl):
    arr = fo.agents_arrays(ags)
    st = sw.get_state()
    for k in ("x", "y", "theta", "v", "w", "u", "I_priv", "collected_r", "collected_r_before",
              "mode", "overriding_mode", "env_status", "exploited_patch_id"):
        st[k][sl] = arr[k]
    nov = st["novelty"].reshape(-1, sw.tau)
    nov[sl] = arr["novelty"]
    st["novelty"] = nov.reshape(-1)
    sw.set_state(**{k: v for k, v in st.items() if k != "radius"})


def compare(sw, b, ags, fields=True):
    sl = sw.replica_slice(b)
    st = sw.get_state()
    arr = fo.agents_arrays(ags)
    for k in ("mode", "overriding_mode"):
        assert np.array_equal(st[k][sl], arr[k]), k
    if fields:
        assert np.array_equal(sw.fields_numpy(b), (arr["soc_v_field"] != 0).astype(np.uint8))
    for k in ("x", "y", "theta", "v", "w", "u", "I_priv", "collected_r_before"):
        assert rel(st[k][sl], arr[k]) <= 1e-12, k


@pytest.mark.parametrize("vis_ex", [False, True])
def test_step_matches_oracle_with_global_stream(gpu, vis_ex):
    # C1-like: N=10, R=1200, 500x500 walls, vision range 2000; 12 steps with the
    # host changing patch status between steps
    from abm_amd.forage import ForageConfig, ForageSwarm
    ags, kw = make_world(7 + vis_ex, vis_ex=vis_ex)
    sw = ForageSwarm([ForageConfig(n_agents=10, n_rays=1200, **kw)],
                     [a.position[0] for a in ags], [a.position[1] for a in ags],
                     [a.orientation for a in ags], radius=10)
    load_swarm(sw, ags, sw.replica_slice(0))
    host = np.random.RandomState(99)
    np.random.seed(1234)
    for step in range(12):
        g0 = np.random.get_state()
        fo.group_update(ags)
        g_oracle = np.random.get_state()
        np.random.set_state(g0)
        sw.step(fields=True)
        g_gpu = np.random.get_state()
        assert g_gpu[2] == g_oracle[2] and np.array_equal(g_gpu[1], g_oracle[1]), step
        compare(sw, 0, ags)
        # host side: notify a few agents of patch status, as the resource loop does
        for a in ags:
            r = host.rand()
            if r < 0.2:
                fo.notify_agent(a, 1, int(host.randint(0, 3)))
            elif r < 0.35:
                fo.notify_agent(a, -1)
            if host.rand() < 0.1:
                a.collected_r += 0.5
        load_swarm(sw, ags, sw.replica_slice(0))


def test_batched_replicas_with_own_streams(gpu):
    from abm_amd.forage import ForageConfig, ForageSwarm
    worlds, cfgs, rngs = [], [], []
    for b, (N, R, vis_ex, vision) in enumerate([(10, 1200, False, 2000.0), (25, 256, True, 150.0),
                                                 (3, 64, True, 2000.0), (40, 512, False, 120.0)]):
        ags, kw = make_world(300 + b, N=N, R=R, vis_ex=vis_ex, vision=vision)
        worlds.append(ags)
        cfgs.append(ForageConfig(n_agents=N, n_rays=R, **kw))
        rngs.append(np.random.RandomState(5000 + b))
    allags = [a for w in worlds for a in w]
    sw = ForageSwarm(cfgs, [a.position[0] for a in allags], [a.position[1] for a in allags],
                     [a.orientation for a in allags], radius=10, rngs=rngs)
    for b, ags in enumerate(worlds):
        load_swarm(sw, ags, sw.replica_slice(b))
    oracle_rngs = [np.random.RandomState(5000 + b) for b in range(len(worlds))]
    for step in range(6):
        with_fields = step % 2 == 0   # without: wave 0 rasters alone where there is no exclusion
        sw.step(fields=with_fields)
        for b, ags in enumerate(worlds):
            np.random.set_state(oracle_rngs[b].get_state())
            fo.group_update(ags)
            oracle_rngs[b].set_state(np.random.get_state())
            assert oracle_rngs[b].get_state()[2] == rngs[b].get_state()[2]
            assert np.array_equal(oracle_rngs[b].get_state()[1], rngs[b].get_state()[1])
            compare(sw, b, ags, fields=with_fields)


@pytest.mark.parametrize("vis_ex", [False, True])
def test_eight_wave_pipeline_matches_oracle(gpu, vis_ex):
    # 240 agents: the step runs eight waves per replica (six classify), and
    # over 64 exploiters see each other, so without exclusion the listed
    # social records span two 64-record rounds of the wave raster
    from abm_amd.forage import ForageConfig, ForageSwarm
    N = 240
    ags, kw = make_world(900 + vis_ex, N=N, R=256, W=800, vis_ex=vis_ex)
    rng = np.random.RandomState(77)
    sw = ForageSwarm([ForageConfig(n_agents=N, n_rays=256, **kw)], [a.position[0] for a in ags],
                     [a.position[1] for a in ags], [a.orientation for a in ags], radius=10, rngs=[rng])
    load_swarm(sw, ags, sw.replica_slice(0))
    assert sum(a.mode == "exploit" for a in ags) > 64
    orng = np.random.RandomState(77)
    for step in range(3):
        with_fields = step == 0
        sw.step(fields=with_fields)
        np.random.set_state(orng.get_state())
        fo.group_update(ags)
        orng.set_state(np.random.get_state())
        assert orng.get_state()[2] == rng.get_state()[2]
        assert np.array_equal(orng.get_state()[1], rng.get_state()[1])
        compare(sw, 0, ags, fields=with_fields)


def test_replica_size_limit(gpu):
    # a replica lives in one workgroup's LDS (172 bytes per agent plus a byte
    # per ray, 160 KB on gfx950): 920 agents at R = 1200 fit; 940 do not and
    # are updated by sweeps instead (test_gpu_sweep.py), which refuse visual
    # exclusion
    from abm_amd import _lib as L
    from abm_amd.forage import ForageConfig, ForageSwarm
    W = 3000.0
    for N, vis_ex in [(920, True), (940, False), (940, True)]:
        rng = np.random.RandomState(3)
        args = ([ForageConfig(n_agents=N, n_rays=1200, width=W, height=W, vision_range=200.0,
                              visual_exclusion=vis_ex)],
                rng.uniform(30, 30 + W - 20, N), rng.uniform(30, 30 + W - 20, N),
                rng.uniform(0, 2 * PI, N))
        if N == 940 and vis_ex:
            with pytest.raises(L.AbmUnsupported, match="LDS"):
                ForageSwarm(*args, radius=10, rngs=[np.random.RandomState(1)])
            continue
        sw = ForageSwarm(*args, radius=10, rngs=[np.random.RandomState(1)])
        sw.step()
        assert np.isfinite(sw.get_state("x")["x"]).all()


@pytest.mark.parametrize("fov_ratio", [1.0, 0.5])
def test_guard_bands_of_the_float32_records(gpu, fov_ratio):
    # every agent exploits its own patch, so each is a cue for every other;
    # around agent 0 the obstacles sit 1e-9 inside and 1e-5 outside each guard
    # band of the float32 record (bin ties, integer half-widths, dead ahead and
    # behind, the FOV ends): inside, the exact record decides; outside, the
    # float32 one.  Fields, and the one step, must match the oracle.
    from abm_amd.forage import ForageConfig, ForageSwarm
    R, r, W = 1200, 10, 500
    fov = (-fov_ratio * PI, fov_ratio * PI)
    step = 2 * PI / (R - 1)
    th0 = 0.3
    px = py = 250.0
    u1 = np.array([np.cos(th0), -np.sin(th0)])

    def at(a, d):
        # object whose centre lies at signed angle a (cross > 0 for a > 0) from the heading
        u2 = np.array([u1[0] * np.cos(a) - u1[1] * np.sin(a), u1[0] * np.sin(a) + u1[1] * np.cos(a)])
        return px + d * u2[0], py + d * u2[1]

    pos = []
    for t in (100, 333, 600, 850, 1100):          # bin ties: ca = -pi + (t + 1/2) step
        for e in (1e-9, -1e-9, 1e-5, -1e-5):
            pos.append(at(-(-PI + (t + 0.5) * step) + e, 80.0 + 7 * len(pos) % 90))
    for k in (1, 2, 3, 5):                          # half-width = k exactly at d_k
        dk = r / np.tan(2 * PI * k / R)
        for e in (1e-9, -1e-9, 1e-5, -1e-5):
            pos.append(at(0.4 + 0.1 * k, dk * (1 + e)))
    for a in (1e-9, -1e-9, 1e-5, -1e-5, PI - 1e-9, -(PI - 1e-9), PI - 1e-5, -(PI - 1e-5)):
        pos.append(at(a, 60.0))                     # ahead and behind
    for e in (1e-9, -1e-9, 1e-5, -1e-5):            # the FOV ends: ca = fov0, fov1
        pos.append(at(-(fov[1] - e), 120.0))
        pos.append(at(-(fov[0] + e), 140.0))
    p = fo.ForageParams(Eps_w=3.0, Eps_u=1.0, F_N=0.5, F_R=0.5)
    ags = [fo.ForageAgent(0, r, (px, py), th0, R, fov, W, W, 30, 2000.0, False, True, p)]
    for j, (x, y) in enumerate(pos):
        ags.append(fo.ForageAgent(j + 1, r, (x, y), 0.1 * j, R, fov, W, W, 30, 2000.0, False, True, p))
    for i, a in enumerate(ags):
        a.w, a.u, a.env_status, a.exploited_patch_id = 0.9, 0.9, 1, i
        a.set_mode("exploit")
    N = len(ags)
    kw = dict(width=W, height=W, window_pad=30, fov=fov, vision_range=2000.0, visual_exclusion=False,
              patchwise_exclusion=True, T_w=p.T_w, Eps_w=p.Eps_w, g_w=p.g_w, B_w=p.B_w, w_max=p.w_max,
              T_u=p.T_u, Eps_u=p.Eps_u, g_u=p.g_u, B_u=p.B_u, u_max=p.u_max, S_wu=p.S_wu,
              S_uw=p.S_uw, Tau=p.Tau, F_N=p.F_N, F_R=p.F_R, exp_vel_max=p.exp_vel_max,
              exp_theta_min=p.exp_theta_min, exp_theta_max=p.exp_theta_max,
              reloc_theta_max=p.reloc_theta_max, exp_stop_ratio=p.exp_stop_ratio)
    sw = ForageSwarm([ForageConfig(n_agents=N, n_rays=R, **kw)], [a.position[0] for a in ags],
                     [a.position[1] for a in ags], [a.orientation for a in ags], radius=r,
                     rngs=[np.random.RandomState(3)])
    load_swarm(sw, ags, sw.replica_slice(0))
    np.random.seed(3)
    fo.group_update(ags)
    sw.step(fields=True)
    assert ags[0].soc_v_field.any()
    compare(sw, 0, ags)
