"""GPU parity of the foraging update by sweeps (replicas beyond one
workgroup's LDS; any replica without visual exclusion on request).

Agent.update runs in id order (sims.py:861); the sweeps reach the same
result as a fixed point.  Bar: bit-identical to the step kernel (which keeps
the replica in LDS) on the same inputs, and to the oracle as the other
foraging tests: fields, modes and generator positions identical, continuous
state within 1e-12."""
import numpy as np
import pytest

from oracle import forage_oracle as fo
from test_gpu_forage import compare, load_swarm, make_world

pytestmark = pytest.mark.gpu
PI = np.pi
KEYS = ("x", "y", "theta", "v", "w", "u", "I_priv", "collected_r_before", "mode", "overriding_mode")


def _swarm(worlds, seed, path=None):
    from abm_amd.forage import ForageConfig, ForageSwarm
    cfgs = [ForageConfig(n_agents=len(ags), n_rays=R, **kw) for ags, R, kw in worlds]
    allags = [a for ags, _, _ in worlds for a in ags]
    rngs = [np.random.RandomState(seed + b) for b in range(len(cfgs))]
    sw = ForageSwarm(cfgs, [a.position[0] for a in allags], [a.position[1] for a in allags],
                     [a.orientation for a in allags], radius=10, rngs=rngs)
    for b, (ags, _, _) in enumerate(worlds):
        load_swarm(sw, ags, sw.replica_slice(b))
    if path:
        sw.set_path(path)
    return sw, rngs


def test_sweeps_equal_the_step_kernel_bit_for_bit(gpu):
    # four replicas (one with visual exclusion, which stays on the step
    # kernel): the same state and streams through both paths, eight steps
    worlds = []
    for b, (N, R, vis_ex, vision) in enumerate([(60, 256, False, 2000.0), (10, 1200, False, 2000.0),
                                                 (45, 512, True, 150.0), (130, 1024, False, 300.0)]):
        ags, kw = make_world(3100 + b, N=N, R=R, W=int(100 * np.sqrt(N)) + 200, vis_ex=vis_ex,
                             vision=vision)
        worlds.append((ags, R, kw))
    a, ra = _swarm(worlds, 3200)
    s, rs = _swarm(worlds, 3200, path="sweep")
    for step in range(8):
        print(f"  step {step}", flush=True)
        a.step(fields=True)
        s.step(fields=True)
        ga, gs = a.get_state(), s.get_state()
        for k in KEYS + ("novelty",):
            assert np.array_equal(ga[k], gs[k]), (step, k)
        assert np.array_equal(a.fields.cpu().numpy(), s.fields.cpu().numpy()), step
        for b in range(len(worlds)):
            assert ra[b].get_state()[2] == rs[b].get_state()[2], (step, b)


@pytest.mark.parametrize("device_rng", [False, True])
def test_swept_world_matches_oracle(gpu, device_rng):
    # the loop body with the resource patches, every replica swept
    from abm_amd.forage import ResourceWorld, WorldConfig
    from test_gpu_world import compare as wcompare, oracle_world, same_stream
    wk = dict(N_resc=6, patch_radius=40, min_resc_perpatch=3, max_resc_perpatch=9,
              min_resc_quality=0.5, max_resc_quality=1.0, teleport_exploit=True)
    worlds = []
    for b, (N, R) in enumerate([(40, 256), (12, 1200), (90, 512)]):
        ags, kw = make_world(3300 + b, N=N, R=R, W=400, vision=2000.0, mixed=False)
        worlds.append((ags, R, kw))
    sw, rngs = _swarm(worlds, 3400, path="sweep")
    orngs = [np.random.RandomState(3400 + b) for b in range(len(worlds))]
    rw = ResourceWorld(sw, [WorldConfig(**wk)] * len(worlds), device_rng=device_rng)
    oracle = []
    for b, (ags, _, _) in enumerate(worlds):
        np.random.set_state(orngs[b].get_state())
        oracle.append(oracle_world(ags, 400, wk))
        orngs[b].set_state(np.random.get_state())
    rw.create_resources()
    rw.sync_rng()
    if device_rng:
        rw.rng_to_device()
    for step in range(25):
        rw.step(fields=step % 5 == 0)
        rw.sync_rng()
        for b, w in enumerate(oracle):
            np.random.set_state(orngs[b].get_state())
            w.step()
            orngs[b].set_state(np.random.get_state())
            assert same_stream(rngs[b], orngs[b]), (step, b)
            wcompare(sw, rw, w, b, (step, b))
            if step % 5 == 0:
                assert np.array_equal(sw.fields_numpy(b),
                                      (fo.agents_arrays(w.agents)["soc_v_field"] != 0).astype(np.uint8))
        if device_rng:
            rw.rng_to_device()


def test_replica_beyond_the_lds_limit_matches_oracle(gpu):
    # 1500 agents at R = 1200 exceed one workgroup's LDS: swept automatically
    N, R, W = 1500, 1200, 3900
    ags, kw = make_world(3500, N=N, R=R, W=W, vision=300.0)
    print(f"foraging N={N} R={R} (sweeps)", flush=True)
    sw, rngs = _swarm([(ags, R, kw)], 3600)
    orng = np.random.RandomState(3600)
    for step in range(2):
        sw.step(fields=step == 0)
        np.random.set_state(orng.get_state())
        fo.group_update(ags)
        orng.set_state(np.random.get_state())
        assert orng.get_state()[2] == rngs[0].get_state()[2]
        compare(sw, 0, ags, fields=step == 0)


def test_visual_exclusion_beyond_the_lds_limit_is_refused(gpu):
    from abm_amd import _lib as L
    from abm_amd.forage import ForageConfig, ForageSwarm
    N, W = 1500, 3900.0
    rng = np.random.RandomState(3)
    with pytest.raises(L.AbmUnsupported, match="visual exclusion"):
        ForageSwarm([ForageConfig(n_agents=N, n_rays=1200, width=W, height=W, vision_range=200.0,
                                  visual_exclusion=True)],
                    rng.uniform(30, 30 + W - 20, N), rng.uniform(30, 30 + W - 20, N),
                    rng.uniform(0, 2 * PI, N), radius=10, rngs=[np.random.RandomState(1)])
