"""GPU parity at the largest sizes the benchmark configurations reach.

* uneven batch: the largest flocking replicas (N = 256, 512 at R = 1024,
  side 100 sqrt(N)) with every agent's field bit-exact against the oracle on
  the poses it saw (a sample of agents across the id range) and verify mode
  at zero disagreements; the largest foraging replicas (N = 500 at R = 256
  with and without visual exclusion, N = 512 at R = 1024) step by step; the
  920-agent LDS limit for one step; a ragged batch of foraging and flocking
  replicas stepped together on one stream, each half against its oracle;
* batched sweep at full size: 4096 replicas x N = 50 x R = 512 in verify mode
  (no float32 decision differs from the float64 one), and the float32 and
  float64 kernels' synchronous-order fields identical.

The reference paths: vf_sims.py:302 (VFAgent.update in id order) and
sims.py:789-861 (resource loop, then Agent.update in id order)."""
import numpy as np
import pytest

from oracle import forage_oracle as fo
from oracle import vf_oracle as vo
from test_gpu_forage import compare as fcompare, load_swarm, make_world
from test_gpu_vf_f32 import per_agent_check, uniform_start, vfparams

pytestmark = pytest.mark.gpu
PI = np.pi


def sample_agents(N, k=20, seed=0):
    rng = np.random.RandomState(seed)
    picks = {0, 1, N // 3, N // 2, N - 2, N - 1}
    picks |= set(rng.choice(N, size=min(N, k), replace=False).tolist())
    return sorted(picks)


@pytest.mark.parametrize("N", [256, 512])
def test_largest_flocking_replicas_of_the_uneven_batch(gpu, N):
    from abm_amd.vf import VFConfig, VFSwarm
    side = 100.0 * np.sqrt(N)
    cfg = VFConfig(n_agents=N, n_rays=1024, width=side, height=side, boundary="infinite",
                   ALP0=1.0, BET0=0.5)
    # one small replica in the same batch, so the launch is a ragged batch
    small = VFConfig(n_agents=9, n_rays=256, width=300.0, height=300.0, boundary="infinite")
    st = uniform_start(N, side, side, 30, seed=21000 + N, v=0.8)
    st2 = uniform_start(9, 300, 300, 30, seed=5, v=0.8)
    cat = [np.concatenate([a, b]) for a, b in zip(st, st2)]
    sw = VFSwarm([cfg, small], *cat, radius=5, precision="f32")
    print(f"flocking N={N} R=1024 side={side:.0f}", flush=True)
    sw.step(fields=True, diag=True, verify=True)
    after = sw.state_numpy()
    sl = sw.replica_slice(0)
    per_agent_check(sw, 0, cfg, st, tuple(a[sl] for a in after), agents=sample_agents(N))
    pairs, rechecked, disagree = sw.fallback_counts()
    assert disagree == 0 and pairs == N * (N - 1) + 9 * 8


def _forage_swarm(worlds_cfgs, seed):
    from abm_amd.forage import ForageConfig, ForageSwarm
    cfgs = [ForageConfig(n_agents=len(ags), n_rays=R, **kw) for ags, R, kw in worlds_cfgs]
    allags = [a for ags, _, _ in worlds_cfgs for a in ags]
    rngs = [np.random.RandomState(seed + b) for b in range(len(cfgs))]
    sw = ForageSwarm(cfgs, [a.position[0] for a in allags], [a.position[1] for a in allags],
                     [a.orientation for a in allags], radius=10, rngs=rngs)
    for b, (ags, _, _) in enumerate(worlds_cfgs):
        load_swarm(sw, ags, sw.replica_slice(b))
    return sw, rngs, [np.random.RandomState(seed + b) for b in range(len(cfgs))]


def _forage_steps(sw, worlds, rngs, orngs, steps, field_steps=(0,)):
    for step in range(steps):
        with_fields = step in field_steps
        print(f"  forage step {step} fields={with_fields}", flush=True)
        sw.step(fields=with_fields)
        for b, ags in enumerate(worlds):
            np.random.set_state(orngs[b].get_state())
            fo.group_update(ags)
            orngs[b].set_state(np.random.get_state())
            assert orngs[b].get_state()[2] == rngs[b].get_state()[2], (step, b)
            assert np.array_equal(orngs[b].get_state()[1], rngs[b].get_state()[1]), (step, b)
            fcompare(sw, b, ags, fields=with_fields)


@pytest.mark.parametrize("vis_ex", [False, True])
def test_500_agent_foraging_replica(gpu, vis_ex):
    # above 384 agents the six classifying waves (stride 384) take two passes
    # per focal agent while two of them also build headings and geometry;
    # vision range 150 on a 2236 px arena (density 1e-4)
    N, R, W = 500, 256, 2236
    ags, kw = make_world(1500 + vis_ex, N=N, R=R, W=W, vis_ex=vis_ex, vision=150.0)
    print(f"foraging N={N} R={R} vis_ex={vis_ex}", flush=True)
    sw, rngs, orngs = _forage_swarm([(ags, R, kw)], 600 + vis_ex)
    _forage_steps(sw, [ags], rngs, orngs, 3)


def test_largest_foraging_replica_of_the_uneven_batch(gpu):
    # N = 512 at R = 1024 on a 2263 px arena, vision range 2000 (the bench's
    # value), with a small replica beside it
    N, R, W = 512, 1024, 2263
    ags, kw = make_world(1600, N=N, R=R, W=W, vision=2000.0)
    ags2, kw2 = make_world(1601, N=12, R=256, W=346, vision=2000.0)
    print(f"foraging N={N} R={R} + N=12 R=256", flush=True)
    sw, rngs, orngs = _forage_swarm([(ags, R, kw), (ags2, 256, kw2)], 700)
    _forage_steps(sw, [ags, ags2], rngs, orngs, 2)


def test_foraging_replica_at_the_lds_limit(gpu):
    # 920 agents at R = 1200 (the largest replica one workgroup's LDS holds):
    # one step against the oracle, fields included
    N, R, W = 920, 1200, 3000
    ags, kw = make_world(1700, N=N, R=R, W=W, vision=300.0)
    print(f"foraging N={N} R={R}", flush=True)
    sw, rngs, orngs = _forage_swarm([(ags, R, kw)], 800)
    _forage_steps(sw, [ags], rngs, orngs, 1)


def test_mixed_forage_and_flock_batch(gpu):
    # a reduced copy of the uneven batch: 24 replicas with N log-uniform in
    # [8, 128], R in {256, 512, 1024}, side 100 sqrt(N); odd replicas flock
    # (fused float32), even replicas forage with the resource loop on the
    # device; both halves step on one stream, then each is checked against
    # its oracle (flocking: every agent's field and pose; foraging: the
    # loop body of ForageWorld, patches included)
    import torch
    from abm_amd.forage import ForageConfig, ForageSwarm, ResourceWorld, WorldConfig
    from abm_amd.vf import VFConfig, VFSwarm
    from test_gpu_world import compare as wcompare, forage_config, oracle_world, same_stream
    rng = np.random.RandomState(20000)
    vf_cfgs, vf_st, fg = [], [], []
    for b in range(24):
        n = int(round(np.exp(rng.uniform(np.log(8), np.log(128)))))
        r = int(rng.choice([256, 512, 1024]))
        side = float(100.0 * np.sqrt(n))
        if b % 2:
            vf_cfgs.append(VFConfig(n_agents=n, n_rays=r, width=side, height=side,
                                    boundary="infinite", ALP0=1.0, BET0=0.5))
            vf_st.append(uniform_start(n, side, side, 30, seed=20000 + b, v=0.5))
        else:
            ags, kw = make_world(20000 + b, N=n, R=r, W=int(side), vision=2000.0, mixed=False)
            fg.append((ags, r, kw))
    cat = [np.concatenate([s[i] for s in vf_st]) for i in range(4)]
    vsw = VFSwarm(vf_cfgs, *cat, radius=5, precision="f32")
    fcfgs = [ForageConfig(n_agents=len(a), n_rays=r, **kw) for a, r, kw in fg]
    allags = [a for ags, _, _ in fg for a in ags]
    rngs = [np.random.RandomState(30000 + i) for i in range(len(fg))]
    orngs = [np.random.RandomState(30000 + i) for i in range(len(fg))]
    fsw = ForageSwarm(fcfgs, [a.position[0] for a in allags], [a.position[1] for a in allags],
                      [a.orientation for a in allags], radius=10, rngs=rngs)
    wk = dict(N_resc=10, patch_radius=30, min_resc_quality=0.5, max_resc_quality=1.0,
              teleport_exploit=False)
    world = ResourceWorld(fsw, [WorldConfig(**wk)] * len(fg), device_rng=True)
    oracle = []
    for i, (ags, _, kw) in enumerate(fg):
        np.random.set_state(orngs[i].get_state())
        oracle.append(oracle_world(ags, kw["width"], wk))
        orngs[i].set_state(np.random.get_state())
    world.create_resources()
    for step in range(3):
        print(f"  mixed step {step}", flush=True)
        before = vsw.state_numpy()
        vsw.step(fields=True, diag=True)
        world.step(fields=True)
        torch.cuda.synchronize()
        after = vsw.state_numpy()
        for b, cfg in enumerate(vf_cfgs):
            sl = vsw.replica_slice(b)
            per_agent_check(vsw, b, cfg, tuple(a[sl] for a in before), tuple(a[sl] for a in after),
                            agents=sample_agents(cfg.n_agents, k=8, seed=b))
        world.sync_rng()
        for i, w in enumerate(oracle):
            np.random.set_state(orngs[i].get_state())
            w.step()
            orngs[i].set_state(np.random.get_state())
            assert same_stream(rngs[i], orngs[i]), (step, i)
            wcompare(fsw, world, w, i, (step, i))
            assert np.array_equal(fsw.fields_numpy(i),
                                  (fo.agents_arrays(w.agents)["soc_v_field"] != 0).astype(np.uint8))
        world.rng_to_device()


def test_full_size_sweep_verify_and_float64_fields(gpu):
    # the benchmark's sweep at its full size: 4096 replicas x 50 x 512 on the
    # 707 px torus.  Verify mode decides every pair both ways for two
    # sequential steps (0 disagreements); then one synchronous step of the
    # float32 and float64 kernels from the same state gives identical fields
    # and a one-step pose error <= 1e-5.
    import torch
    from abm_amd.vf import VFSwarm
    from bench import make_workload
    cfgs, x, y, th = make_workload()
    sw = VFSwarm(cfgs, x, y, th, 0.5, radius=5, precision="f32")
    for _ in range(2):
        sw.step(verify=True)
    pairs, rechecked, disagree = sw.fallback_counts()
    assert pairs == 2 * 4096 * 50 * 49 and disagree == 0
    s32 = VFSwarm(cfgs, x, y, th, 0.5, radius=5, precision="f32")
    s64 = VFSwarm(cfgs, x, y, th, 0.5, radius=5, precision="f64")
    s32.step(order="synchronous", fields=True)
    s64.step(order="synchronous", fields=True)
    torch.cuda.synchronize()
    assert torch.equal(s32.fields, s64.fields)
    a32, a64 = s32.state_numpy(), s64.state_numpy()
    W = 707.0
    dx = np.abs((a32[0] - a64[0] + W / 2) % W - W / 2) / np.maximum(np.abs(a64[0]), 1.0)
    dy = np.abs((a32[1] - a64[1] + W / 2) % W - W / 2) / np.maximum(np.abs(a64[1]), 1.0)
    dth = np.abs((a32[2] - a64[2] + PI) % (2 * PI) - PI) / np.maximum(np.abs(a64[2]), 1.0)
    dv = np.abs(a32[3] - a64[3]) / np.maximum(np.abs(a64[3]), 1.0)
    assert max(dx.max(), dy.max(), dth.max(), dv.max()) <= 1e-5
