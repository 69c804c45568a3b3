"""GPU parity of the device resource loop (abm_forage_world_run through
abm_amd.forage.ResourceWorld) against the oracle's ForageWorld, the loop body
of Simulation.start with agent collisions off (sims.py:789-861) and patch
creation / regeneration (sims.py:321-368, 538-542), each replica on its own
numpy RandomState.

Bar: patches (ids, positions, units, resources left, quality), modes,
environment status, patch ids, novelty rings, pool_success, collected
resources, fields and every generator's stream position identical after each
step; continuous state (x, y, theta, v, w, u, I_priv, collected_r_before)
within 1e-12 relative.  The state is reloaded from the oracle after each step
so that every step starts from identical values (the bias towards the patch
centre uses atan2, whose device and host results may differ in the last
ulp)."""
import numpy as np
import pytest

from oracle import forage_oracle as fo

pytestmark = pytest.mark.gpu
PI = np.pi

DEC = dict(Eps_w=3.0, Eps_u=1.0, S_wu=0.25, S_uw=0.01, F_N=0.5, F_R=0.5, exp_vel_max=3.0,
           exp_theta_min=-0.5, exp_theta_max=0.5, reloc_theta_max=1.8, exp_stop_ratio=0.175)

# (N agents, arena, R, WorldConfig kwargs, mixed start)
WORLDS = [
    (30, 300, 256, dict(N_resc=4, patch_radius=30, min_resc_perpatch=3, max_resc_perpatch=8,
                        min_resc_quality=0.5, max_resc_quality=1.0, teleport_exploit=True), False),
    (20, 300, 128, dict(N_resc=5, patch_radius=25, min_resc_perpatch=2, max_resc_perpatch=4,
                        min_resc_quality=0.2, max_resc_quality=0.6, teleport_exploit=False,
                        regenerate_patches=False, agent_consumption=0.5), True),
    (25, 250, 256, dict(N_resc=8, patch_radius=30, min_resc_perpatch=4, max_resc_perpatch=6,
                        min_resc_quality=0.7, max_resc_quality=-1.0, teleport_exploit=True,
                        allow_border_patch_overlap=True), True),
    (1, 200, 64, dict(N_resc=1, patch_radius=40, min_resc_perpatch=1, max_resc_perpatch=2,
                      min_resc_quality=0.4, max_resc_quality=0.9, teleport_exploit=True), False),
    # a positive pooling time: only relocating agents that were notified
    # before (pool_success) hear of patches
    (15, 250, 128, dict(N_resc=3, patch_radius=35, min_resc_perpatch=2, max_resc_perpatch=5,
                        pooling_time=2, teleport_exploit=False), True),
]


def rel(a, b, s=1.0):
    return float(np.max(np.abs(np.asarray(a) - np.asarray(b)) / np.maximum(np.abs(b), s)))


def make_agents(seed, N, W, R, mixed, pooling_time):
    p = fo.ForageParams(**DEC)
    rng = np.random.RandomState(seed)
    ags = []
    for i in range(N):
        x = rng.randint(30 - 10, W + 30 - 10)
        y = rng.randint(30 - 10, W + 30 - 10)
        ags.append(fo.ForageAgent(i, 10, (x, y), rng.uniform(0, 2 * PI), R, (-PI, PI), W, W, 30,
                                  150.0, False, True, p, pooling_time=pooling_time))
    if mixed:
        for a in ags:
            a.w = float(rng.uniform(-0.2, 1.0))
            a.u = float(rng.uniform(-0.2, 1.0))
            a.velocity = float(rng.uniform(0, 3))
            a.set_mode(str(rng.choice(["explore", "exploit", "relocate"], p=[0.5, 0.3, 0.2])))
            a.pool_success = int(rng.rand() < 0.5)
    return ags, p


def forage_config(N, W, R, p, vis_ex=False):
    from abm_amd.forage import ForageConfig
    return ForageConfig(n_agents=N, n_rays=R, width=W, height=W, window_pad=30, fov=(-PI, PI),
                        vision_range=150.0, visual_exclusion=vis_ex, patchwise_exclusion=True,
                        T_w=p.T_w, Eps_w=p.Eps_w, g_w=p.g_w, B_w=p.B_w, w_max=p.w_max, T_u=p.T_u,
                        Eps_u=p.Eps_u, g_u=p.g_u, B_u=p.B_u, u_max=p.u_max, S_wu=p.S_wu,
                        S_uw=p.S_uw, Tau=p.Tau, F_N=p.F_N, F_R=p.F_R, exp_vel_max=p.exp_vel_max,
                        exp_theta_min=p.exp_theta_min, exp_theta_max=p.exp_theta_max,
                        reloc_theta_max=p.reloc_theta_max, exp_stop_ratio=p.exp_stop_ratio)


def oracle_world(ags, W, wk):
    from abm_amd.forage import WorldConfig
    c = WorldConfig(**wk)
    return fo.ForageWorld(ags, W, W, 30, c.N_resc, c.patch_radius, c.min_resc_perpatch,
                          c.max_resc_perpatch, c.min_resc_quality, c.max_resc_quality,
                          regenerate=c.regenerate_patches, teleport_exploit=c.teleport_exploit,
                          allow_border_overlap=c.allow_border_patch_overlap,
                          agent_consumption=c.agent_consumption, collide_agents=c.collide_agents,
                          ghost_mode=c.ghost_mode)


def load_agents(sw, rw, ags, b):
    sl = sw.replica_slice(b)
    arr = fo.agents_arrays(ags)
    st = sw.get_state()
    for k in ("x", "y", "theta", "v", "w", "u", "I_priv", "collected_r", "collected_r_before",
              "mode", "overriding_mode", "env_status", "exploited_patch_id"):
        st[k][sl] = arr[k]
    nov = st["novelty"].reshape(-1, sw.tau)
    nov[sl] = arr["novelty"]
    st["novelty"] = nov.reshape(-1)
    sw.set_state(**{k: v for k, v in st.items() if k != "radius"})
    ps = rw.pool_success.cpu().numpy()
    ps[sl] = [a.pool_success for a in ags]
    rw.pool_success.copy_(rw.pool_success.new_tensor(ps))


def compare(sw, rw, world, b, where):
    sl = sw.replica_slice(b)
    st = sw.get_state()
    arr = fo.agents_arrays(world.agents)
    for k in ("mode", "overriding_mode", "env_status", "exploited_patch_id", "collected_r"):
        assert np.array_equal(st[k][sl], arr[k]), (where, k)
    assert np.array_equal(st["novelty"].reshape(-1, sw.tau)[sl], arr["novelty"]), where
    assert np.array_equal(rw.pool_success.cpu().numpy()[sl], [a.pool_success for a in world.agents])
    for k in ("x", "y", "theta", "v", "w", "u", "I_priv", "collected_r_before"):
        assert rel(st[k][sl], arr[k]) <= 1e-12, (where, k)
    got = rw.patches(b)
    exp = dict(id=[p.id for p in world.patches], x=[p.position[0] for p in world.patches],
               y=[p.position[1] for p in world.patches], radius=[p.radius for p in world.patches],
               units=[p.resc_units for p in world.patches], left=[p.resc_left for p in world.patches],
               quality=[p.unit_per_timestep for p in world.patches])
    for k, v in exp.items():
        assert np.array_equal(got[k], np.asarray(v, dtype=got[k].dtype)), (where, k, got[k], v)


def same_stream(a, b):
    return a.get_state()[2] == b.get_state()[2] and np.array_equal(a.get_state()[1], b.get_state()[1])


@pytest.mark.parametrize("regen_words,device_rng", [(32, False), (1, False), (32, True)])
def test_world_matches_oracle(gpu, regen_words, device_rng):
    from abm_amd.forage import ForageSwarm, ResourceWorld, WorldConfig
    worlds, cfgs, wcfgs, allags = [], [], [], []
    for b, (N, W, R, wk, mixed) in enumerate(WORLDS):
        ags, p = make_agents(70 + b, N, W, R, mixed, wk.get("pooling_time", 0))
        worlds.append((ags, W, wk))
        cfgs.append(forage_config(N, W, R, p))
        wcfgs.append(WorldConfig(**wk))
        allags += ags
    rngs = [np.random.RandomState(900 + b) for b in range(len(WORLDS))]
    orngs = [np.random.RandomState(900 + b) for b in range(len(WORLDS))]
    sw = ForageSwarm(cfgs, [a.position[0] for a in allags], [a.position[1] for a in allags],
                     [a.orientation for a in allags], radius=10, rngs=rngs)
    rw = ResourceWorld(sw, wcfgs, regen_words=regen_words, device_rng=device_rng)
    oracle = []
    for b, (ags, W, wk) in enumerate(worlds):
        load_agents(sw, rw, ags, b)
        np.random.set_state(orngs[b].get_state())
        oracle.append(oracle_world(ags, W, wk))
        orngs[b].set_state(np.random.get_state())
    print(f"create_resources regen_words={regen_words} device_rng={device_rng}", flush=True)
    rw.create_resources()
    rw.sync_rng()
    for b, world in enumerate(oracle):
        assert same_stream(rngs[b], orngs[b]), b
        compare(sw, rw, world, b, ("create", b))
    kills = [0] * len(oracle)
    first = [[(p.id, p.position[0], p.position[1]) for p in w.patches] for w in oracle]
    for step in range(80):
        # odd steps without field output: wave 0 rasters and counts alone
        # where there is no visual exclusion
        with_fields = step % 2 == 0
        rw.step(fields=with_fields)
        rw.sync_rng()
        for b, world in enumerate(oracle):
            before = [(p.id, p.position[0], p.position[1]) for p in world.patches]
            np.random.set_state(orngs[b].get_state())
            world.step()
            orngs[b].set_state(np.random.get_state())
            after = [(p.id, p.position[0], p.position[1]) for p in world.patches]
            kills[b] += len(set(before) - set(after))
            assert same_stream(rngs[b], orngs[b]), (step, b)
            compare(sw, rw, world, b, (step, b))
            if with_fields:
                assert np.array_equal(sw.fields_numpy(b),
                                      (fo.agents_arrays(world.agents)["soc_v_field"] != 0).astype(np.uint8))
            load_agents(sw, rw, world.agents, b)
    # depletion, regeneration (ids kept) and removal all happened
    assert kills[0] > 0 and kills[2] > 0 and kills[3] > 0, kills
    assert len(oracle[1].patches) < len(first[1])
    assert sorted(p.id for p in oracle[0].patches) == sorted(i for i, _, _ in first[0])
    if regen_words == 1:
        assert rw.resumes > 0
    if device_rng:
        assert rw.resumes == 0


def test_world_placement_timeout_raises(gpu):
    # two patches of radius 30 whose centres must lie in a 40 x 40 square
    # always collide: the reference gives up after 10000 retries
    from abm_amd.forage import ForageSwarm, ResourceWorld, WorldConfig, TIMEOUT_MESSAGE
    ags, p = make_agents(5, 1, 100, 16, False, 0)
    wk = dict(N_resc=2, patch_radius=30, min_resc_perpatch=5, max_resc_perpatch=10)
    np.random.seed(11)
    with pytest.raises(Exception, match=TIMEOUT_MESSAGE):
        oracle_world(ags, 100, wk)
    np.random.seed(11)
    sw = ForageSwarm([forage_config(1, 100, 16, p)], [a.position[0] for a in ags],
                     [a.position[1] for a in ags], [a.orientation for a in ags], radius=10)
    rw = ResourceWorld(sw, [WorldConfig(**wk)])
    with pytest.raises(Exception, match=TIMEOUT_MESSAGE):
        rw.create_resources()
    # with the generator on the device the timeout shows when the status is read
    np.random.seed(11)
    rw = ResourceWorld(sw, [WorldConfig(**wk)], device_rng=True)
    rw.create_resources()
    with pytest.raises(Exception, match=TIMEOUT_MESSAGE):
        rw.check()


def test_world_rejects_bad_patch_count(gpu):
    from abm_amd import _lib as L
    from abm_amd.forage import ForageSwarm, ResourceWorld, WorldConfig
    ags, p = make_agents(6, 3, 200, 16, False, 0)
    sw = ForageSwarm([forage_config(3, 200, 16, p)], [a.position[0] for a in ags],
                     [a.position[1] for a in ags], [a.orientation for a in ags], radius=10,
                     rngs=[np.random.RandomState(1)])
    rw = ResourceWorld(sw, [WorldConfig(N_resc=2)])
    rw.p["count"][0] = 3
    with pytest.raises(L.AbmError, match="patch count"):
        rw.step()
    with pytest.raises(L.AbmError):
        ResourceWorld(sw, [WorldConfig(N_resc=65)])
    with pytest.raises(L.AbmError):
        ResourceWorld(sw, [WorldConfig(N_resc=2, patch_radius=120)])   # randint(30, -10)


C1_PARAMS = dict(T_w=0.5, Eps_w=3.0, g_w=0.085, B_w=0.0, w_max=1.0, T_u=0.5, Eps_u=1.0,
                 g_u=0.085, B_u=0.0, u_max=1.0, S_wu=0.25, S_uw=0.01, Tau=10, F_N=0.5,
                 F_R=0.5, exp_vel_max=3.0, exp_theta_min=-0.5, exp_theta_max=0.5,
                 reloc_theta_max=1.8, exp_stop_ratio=0.175)
C1_WORLD = dict(N_resc=10, patch_radius=30, min_resc_perpatch=200, max_resc_perpatch=1000,
                min_resc_quality=0.5, max_resc_quality=1.0, teleport_exploit=False)


def c1_replicas(B, seed0=4000):
    """B replicas of the reference's CPU case (N=10, radius 10, R=1200, 10
    patches of radius 30 with quality U[0.5, 1], 500x500 walls, the .env
    decision values), created as Simulation.create_agents does
    (sims.py:526-532), plus matching device state and generators."""
    from abm_amd.forage import ForageSwarm, ResourceWorld, WorldConfig
    p = fo.ForageParams(**C1_PARAMS)
    N, R, W = 10, 1200, 500
    worlds, allags, rngs, orngs = [], [], [], []
    for b in range(B):
        g = np.random.RandomState(seed0 + b)
        ags = []
        for i in range(N):
            x = g.randint(30 - 10, W + 30 - 10)
            y = g.randint(30 - 10, W + 30 - 10)
            ags.append(fo.ForageAgent(i, 10, (x, y), g.uniform(0, 2 * PI), R, (-PI, PI), W, W, 30,
                                      2000.0, False, True, p, pooling_time=0))
        worlds.append(ags)
        allags += ags
        for lst in (rngs, orngs):
            lst.append(np.random.RandomState())
            lst[-1].set_state(g.get_state())
    cfg = forage_config(N, W, R, p)
    cfg.vision_range = 2000.0
    sw = ForageSwarm([cfg] * B, [a.position[0] for a in allags], [a.position[1] for a in allags],
                     [a.orientation for a in allags], radius=10, rngs=rngs)
    rw = ResourceWorld(sw, [WorldConfig(**C1_WORLD)] * B)
    oracle = []
    for b in range(B):
        np.random.set_state(orngs[b].get_state())
        oracle.append(oracle_world(worlds[b], W, C1_WORLD))
        orngs[b].set_state(np.random.get_state())
    rw.create_resources()
    return sw, rw, oracle, rngs, orngs


def oracle_step(world, g):
    np.random.set_state(g.get_state())
    world.step()
    g.set_state(np.random.get_state())


DISCRETE = ("mode", "overriding_mode", "env_status", "exploited_patch_id")


def test_world_c1_one_step_parity_over_1000_steps(gpu):
    # The C1 run for its full 1000 steps along the oracle's trajectory: every
    # step starts the device from the oracle's state (agents and generator),
    # so every step's discrete state must be identical and the continuous
    # state within 1e-12.  Patches are never reloaded: they must stay
    # identical on their own.
    B, steps = 12, 1000
    sw, rw, oracle, rngs, orngs = c1_replicas(B)
    for b in range(B):
        assert same_stream(rngs[b], orngs[b])
        compare(sw, rw, oracle[b], b, ("create", b))
    print(f"C1 one-step parity: {B} replicas x {steps} steps", flush=True)
    for step in range(steps):
        rw.step()
        for b in range(B):
            oracle_step(oracle[b], orngs[b])
        st = sw.get_state()
        for b in range(B):
            assert same_stream(rngs[b], orngs[b]), (step, b)
            sl = sw.replica_slice(b)
            arr = fo.agents_arrays(oracle[b].agents)
            for k in DISCRETE + ("collected_r",):
                assert np.array_equal(st[k][sl], arr[k]), (step, b, k)
            for k in ("x", "y", "theta", "v", "w", "u", "I_priv", "collected_r_before"):
                assert rel(st[k][sl], arr[k]) <= 1e-12, (step, b, k)
            assert np.array_equal(st["novelty"].reshape(-1, sw.tau)[sl], arr["novelty"])
            got = rw.patches(b) if step % 50 == 0 else None
            if got is not None:
                assert np.array_equal(got["id"], [q.id for q in oracle[b].patches])
                assert np.array_equal(got["left"], [q.resc_left for q in oracle[b].patches])
            # the next step starts from the oracle's continuous values
            for k in ("x", "y", "theta", "v", "w", "u", "I_priv", "collected_r_before"):
                st[k][sl] = arr[k]
        sw.set_state(**{k: st[k] for k in ("x", "y", "theta", "v", "w", "u", "I_priv",
                                           "collected_r_before")})
        if step % 200 == 0:
            print(f"  step {step} ok", flush=True)
    assert sum(a.collected_r for w in oracle for a in w.agents) > 0


def test_world_c1_free_run(gpu):
    # C1 run free for its full 1000 steps on device and oracle from the same
    # start and streams, nothing reloaded.  The device's libm and the host's
    # (glibc, numpy) differ in the last bit of atan2 / cos / sin for about
    # one call in a thousand, and the reference's bias towards the patch
    # centre multiplies a heading difference by 1.02 per step an agent stays
    # on a patch (sims.py:551-552; test_bias_multiplies_heading_differences_
    # by_1_02_per_step), so a last-bit difference grows about tenfold per 100
    # steps.  Bar: every replica's discrete state and generator position
    # identical for the first 600 steps (the growth needs about 800 steps
    # from one ulp to a changed decision), and the ensemble of collected
    # resource after 1000 steps equal by Welch's t-test at 99.9 %.
    from scipy import stats
    B, steps, exact_until = 12, 1000, 600
    sw, rw, oracle, rngs, orngs = c1_replicas(B)
    first_diff = [None] * B
    for step in range(steps):
        rw.step()
        for b in range(B):
            oracle_step(oracle[b], orngs[b])
        if step < exact_until or step % 50 == 0:
            st = sw.get_state(*DISCRETE, "collected_r")
            for b in range(B):
                if first_diff[b] is not None:
                    continue
                sl = sw.replica_slice(b)
                arr = fo.agents_arrays(oracle[b].agents)
                if not (same_stream(rngs[b], orngs[b]) and all(
                        np.array_equal(st[k][sl], arr[k]) for k in DISCRETE + ("collected_r",))):
                    first_diff[b] = step
    got = sw.get_state("collected_r")["collected_r"].reshape(B, 10).sum(axis=1)
    exp = np.array([sum(a.collected_r for a in oracle[b].agents) for b in range(B)])
    print(f"collected resource after {steps} steps: device {got.mean():.2f} +- {got.std():.2f}, "
          f"oracle {exp.mean():.2f} +- {exp.std():.2f}; first discrete difference per replica "
          f"{first_diff}", flush=True)
    assert all(f is None or f >= exact_until for f in first_diff), first_diff
    assert exp.sum() > 0
    assert np.array_equal(got, exp) or stats.ttest_ind(got, exp, equal_var=False).pvalue > 1e-3


def test_world_membership_over_several_words(gpu):
    # 170 agents: a patch's membership spans three 64-agent words, so the
    # group walk, notify order and the depletion broadcast cross words.  Six
    # large patches (radius 60) on a 400 px arena with small stocks, so
    # patches are emptied and regenerated while crowded; teleport on.
    from abm_amd.forage import ForageSwarm, ResourceWorld, WorldConfig
    N, W, R = 170, 400, 256
    wk = dict(N_resc=6, patch_radius=60, min_resc_perpatch=3, max_resc_perpatch=9,
              min_resc_quality=0.5, max_resc_quality=1.0, teleport_exploit=True)
    ags, p = make_agents(1170, N, W, R, True, 0)
    rng, orng = np.random.RandomState(1171), np.random.RandomState(1171)
    sw = ForageSwarm([forage_config(N, W, R, p)], [a.position[0] for a in ags],
                     [a.position[1] for a in ags], [a.orientation for a in ags], radius=10,
                     rngs=[rng])
    rw = ResourceWorld(sw, [WorldConfig(**wk)])
    load_agents(sw, rw, ags, 0)
    np.random.set_state(orng.get_state())
    world = oracle_world(ags, W, wk)
    orng.set_state(np.random.get_state())
    rw.create_resources()
    compare(sw, rw, world, 0, "create")
    kills = 0
    crowded = 0
    for step in range(30):
        rw.step(fields=step % 3 == 0)
        for q in world.patches:
            words = {a.id // 64 for a in world.agents if fo.distance(q, a) < q.radius}
            crowded = max(crowded, len(words))
        before = {q.id: tuple(q.position) for q in world.patches}
        np.random.set_state(orng.get_state())
        world.step()
        orng.set_state(np.random.get_state())
        kills += sum(before[q.id] != tuple(q.position) for q in world.patches if q.id in before)
        assert same_stream(rng, orng), step
        compare(sw, rw, world, 0, step)
        load_agents(sw, rw, world.agents, 0)
    assert kills > 0 and crowded == 3, (kills, crowded)   # a patch held agents of all three words


def test_world_overlapping_caller_patches(gpu):
    # patches a caller loads may overlap (the loop never places such): an
    # agent in two groups is walked twice, in Group order, as in the
    # reference's loop (sims.py:803-845); the kernel detects the shared
    # agent and walks that replica's groups in one thread.  A second replica
    # with separate patches takes the parallel walk in the same launch.
    from abm_amd.forage import ForageSwarm, ResourceWorld, WorldConfig
    N, W, R = 40, 300, 256
    wk = dict(N_resc=3, patch_radius=40, min_resc_perpatch=3, max_resc_perpatch=9,
              min_resc_quality=0.5, max_resc_quality=1.0, teleport_exploit=False)
    layouts = [[(1, (80, 80)), (2, (120, 80)), (3, (220, 220))],     # 1 and 2 overlap
               [(1, (40, 40)), (2, (140, 80)), (3, (220, 220))]]     # apart
    worlds, rngs, orngs = [], [], []
    for b, layout in enumerate(layouts):
        ags, p = make_agents(2200 + b, N, W, R, True, 0)
        r0 = np.random.RandomState(2210 + b)
        for a in ags:   # crowd the agents over the first two patches
            a.position = np.array([r0.randint(100, 160), r0.randint(90, 130)], dtype=np.float64)
        world = fo.ForageWorld(ags, W, W, 30, 0, 40, 3, 9, 0.5, 1.0)
        world.patches = [fo.Patch(i, 40, pos, 5, 0.75) for i, pos in layout]
        worlds.append(world)
        rngs.append(np.random.RandomState(2220 + b))
        orngs.append(np.random.RandomState(2220 + b))
    cfg = forage_config(N, W, R, p)
    allags = [a for w in worlds for a in w.agents]
    sw = ForageSwarm([cfg, cfg], [a.position[0] for a in allags], [a.position[1] for a in allags],
                     [a.orientation for a in allags], radius=10, rngs=rngs)
    rw = ResourceWorld(sw, [WorldConfig(**wk)] * 2)
    for b, world in enumerate(worlds):
        load_agents(sw, rw, world.agents, b)
        pt = world.patches
        rw.set_patches(b, [q.id for q in pt], [q.position[0] for q in pt], [q.position[1] for q in pt],
                       [q.radius for q in pt], [q.resc_units for q in pt], [q.resc_left for q in pt],
                       [q.unit_per_timestep for q in pt])
    shared = 0
    for step in range(12):
        print(f"  overlap step {step}", flush=True)
        q1, q2 = worlds[0].patches[0], worlds[0].patches[1]
        shared += sum(fo.distance(q1, a) < q1.radius and fo.distance(q2, a) < q2.radius
                      for a in worlds[0].agents)
        rw.step()
        for b, world in enumerate(worlds):
            np.random.set_state(orngs[b].get_state())
            world.step()
            orngs[b].set_state(np.random.get_state())
            assert same_stream(rngs[b], orngs[b]), (step, b)
            compare(sw, rw, world, b, (step, b))
            load_agents(sw, rw, world.agents, b)
    assert shared > 0


# (N agents, arena, R, WorldConfig kwargs, visual exclusion): crowded arenas
# with agent-agent collisions on
COLLIDE_WORLDS = [
    (40, 220, 256, dict(N_resc=3, patch_radius=30, min_resc_perpatch=3, max_resc_perpatch=8,
                        min_resc_quality=0.5, max_resc_quality=1.0, teleport_exploit=True,
                        collide_agents=True, ghost_mode=True), False),
    (30, 180, 1200, dict(N_resc=2, patch_radius=25, min_resc_perpatch=2, max_resc_perpatch=5,
                         min_resc_quality=0.5, max_resc_quality=1.0, teleport_exploit=False,
                         collide_agents=True, ghost_mode=False), False),
    (35, 200, 512, dict(N_resc=3, patch_radius=30, min_resc_perpatch=3, max_resc_perpatch=6,
                        min_resc_quality=0.5, max_resc_quality=1.0, teleport_exploit=False,
                        collide_agents=True, ghost_mode=True), True),
    # collisions off next to them
    (20, 250, 256, dict(N_resc=3, patch_radius=30, min_resc_perpatch=3, max_resc_perpatch=6,
                        teleport_exploit=True), False),
]


@pytest.mark.parametrize("device_rng", [False, True])
def test_world_collisions_match_oracle(gpu, device_rng):
    # sims.py:736-787 (groupcollide on rect centres, the proximity response,
    # the collided list) before the resource loop, every step from the
    # oracle's state: discrete state identical, continuous <= 1e-12
    from abm_amd.forage import ForageSwarm, ResourceWorld, WorldConfig
    worlds, cfgs, wcfgs, allags = [], [], [], []
    for b, (N, W, R, wk, vis_ex) in enumerate(COLLIDE_WORLDS):
        ags, p = make_agents(300 + b, N, W, R, True, 0)
        for a in ags:
            a.visual_exclusion = vis_ex
        worlds.append((ags, W, wk))
        cfgs.append(forage_config(N, W, R, p, vis_ex))
        wcfgs.append(WorldConfig(**wk))
        allags += ags
    rngs = [np.random.RandomState(950 + b) for b in range(len(worlds))]
    orngs = [np.random.RandomState(950 + b) for b in range(len(worlds))]
    sw = ForageSwarm(cfgs, [a.position[0] for a in allags], [a.position[1] for a in allags],
                     [a.orientation for a in allags], radius=10, rngs=rngs)
    rw = ResourceWorld(sw, wcfgs, device_rng=device_rng)
    oracle = []
    for b, (ags, W, wk) in enumerate(worlds):
        load_agents(sw, rw, ags, b)
        np.random.set_state(orngs[b].get_state())
        oracle.append(oracle_world(ags, W, wk))
        orngs[b].set_state(np.random.get_state())
    rw.create_resources()
    rw.sync_rng()
    collide_seen = [0] * len(worlds)
    for step in range(40):
        print(f"  collisions step {step} device_rng={device_rng}", flush=True)
        rw.step(fields=step % 4 == 0)
        rw.sync_rng()
        for b, world in enumerate(oracle):
            np.random.set_state(orngs[b].get_state())
            world.step()
            orngs[b].set_state(np.random.get_state())
            collide_seen[b] += sum(a.mode == "collide" for a in world.agents)
            assert same_stream(rngs[b], orngs[b]), (step, b)
            compare(sw, rw, world, b, (step, b))
            load_agents(sw, rw, world.agents, b)
    assert all(c > 0 for c in collide_seen[:3]) and collide_seen[3] == 0, collide_seen


def test_world_collisions_crowded_and_size_limits(gpu):
    # more than 256 agents within 2 r + 20 of a colliding agent: the replica
    # is reported and not stepped; collisions above 2048 agents are refused
    from abm_amd import AbmUnsupported
    from abm_amd.forage import ForageSwarm, ResourceWorld, WorldConfig
    N, W, R = 300, 400, 64
    ags, p = make_agents(5, N, W, R, False, 0)
    rng = np.random.RandomState(6)
    for a in ags:
        a.position = np.array([200.0 + rng.randint(0, 20), 200.0 + rng.randint(0, 20)])
    rngs = [np.random.RandomState(7)]
    sw = ForageSwarm([forage_config(N, W, R, p)], [a.position[0] for a in ags],
                     [a.position[1] for a in ags], [a.orientation for a in ags], radius=10, rngs=rngs)
    rw = ResourceWorld(sw, [WorldConfig(N_resc=1, collide_agents=True)])
    load_agents(sw, rw, ags, 0)
    rw.create_resources()
    with pytest.raises(AbmUnsupported, match="2 r \\+ 20"):
        rw.step()
    big, _ = make_agents(8, 2049, 2000, 64, False, 0)
    sw2 = ForageSwarm([forage_config(2049, 2000, 64, p)], [a.position[0] for a in big],
                      [a.position[1] for a in big], [a.orientation for a in big], radius=10,
                      rngs=[np.random.RandomState(9)])
    with pytest.raises(AbmUnsupported, match="collisions"):
        ResourceWorld(sw2, [WorldConfig(N_resc=1, collide_agents=True)])


def test_collision_response_on_mirror_images(gpu):
    # the proximity response's left/right decision where the two halves hold
    # mirror images (two neighbours at +-dy, 2 r + 20 apart at most, no
    # collision of their own), so the means differ only by the order of
    # numpy's pairwise sums; a colliding partner straight behind (invisible)
    # triggers the response.  R = 100 puts the front window's start before 0
    # (a Python slice from the end).
    from abm_amd.forage import ForageSwarm, ResourceWorld, WorldConfig
    cases = [(dx, dy, R) for R in (1200, 1024, 256, 100) for dx, dy in ((10, 18), (21, 12), (5, 25), (16, 17))]
    worlds, allags, cfgs = [], [], []
    p = None
    for b, (dx, dy, R) in enumerate(cases):
        ags, p = make_agents(400 + b, 4, 300, R, False, 0)
        pos = [(100.0, 100.0), (76.0, 100.0), (100.0 + dx, 100.0 - dy), (100.0 + dx, 100.0 + dy)]
        for a, q in zip(ags, pos):
            a.position = np.array(q)
            a.orientation = 0.0
            a.velocity = 0.5
        worlds.append(ags)
        cfgs.append(forage_config(4, 300, R, p))
        allags += ags
    rngs = [np.random.RandomState(1000 + b) for b in range(len(cases))]
    orngs = [np.random.RandomState(1000 + b) for b in range(len(cases))]
    sw = ForageSwarm(cfgs, [a.position[0] for a in allags], [a.position[1] for a in allags],
                     [a.orientation for a in allags], radius=10, rngs=rngs)
    wk = dict(N_resc=1, patch_radius=20, min_resc_perpatch=3, max_resc_perpatch=6,
              collide_agents=True, ghost_mode=True, teleport_exploit=False)
    rw = ResourceWorld(sw, [WorldConfig(**wk)] * len(cases))
    oracle = []
    for b, ags in enumerate(worlds):
        load_agents(sw, rw, ags, b)
        np.random.set_state(orngs[b].get_state())
        oracle.append(oracle_world(ags, 300, wk))
        orngs[b].set_state(np.random.get_state())
    rw.create_resources()   # the same draws as the oracle's patch
    for b in range(len(cases)):
        # the patch away from the agents
        oracle[b].patches[0].position = np.array([250.0, 250.0])
        oracle[b].patches[0].center = (270.0, 270.0)
        q = oracle[b].patches[0]
        rw.set_patches(b, [q.id], [250.0], [250.0], [q.radius], [q.resc_units], [q.resc_left],
                       [q.unit_per_timestep])
    rw.step()
    for b, world in enumerate(oracle):
        np.random.set_state(orngs[b].get_state())
        world.step()
        orngs[b].set_state(np.random.get_state())
        assert world.agents[0].mode == "collide"
        compare(sw, rw, world, b, cases[b])
