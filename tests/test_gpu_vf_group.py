"""GPU parity of the grouped fused float32 kernel (G = 1, 2, 4 replicas per
wave) against the oracle and against the per-replica fused kernel.

Bar (as for the per-replica kernel): every agent's field bit-exact against the
oracle on the poses it saw; zero float32/float64 decision disagreements in
verify mode; one-step relative pose error <= 1e-5; the diagnostics within
rtol 1e-4 / atol 2e-6; fields identical to the per-replica kernel's."""
import numpy as np
import pytest

from oracle import vf_oracle as vo
from test_gpu_vf_f32 import TOL, per_agent_check, pose_err, uniform_start, vfparams

pytestmark = pytest.mark.gpu


def swarm(cfgs, states, radius=5, **kw):
    from abm_amd.vf import VFSwarm
    cat = [np.concatenate([s[c] for s in states]) for c in range(4)]
    return VFSwarm(cfgs, *cat, radius=radius, precision="f32", **kw)


def sweep_batch(Ns, R=512, side=707.0, boundary="infinite", seed=0, **cfgkw):
    from abm_amd.vf import VFConfig
    rng = np.random.RandomState(seed)
    cfgs, states = [], []
    for b, N in enumerate(Ns):
        cfgs.append(VFConfig(n_agents=N, n_rays=R, width=side, height=side, boundary=boundary,
                             ALP0=float(rng.uniform(0.1, 3)), BET0=float(rng.uniform(0.1, 3)),
                             **cfgkw))
        states.append(uniform_start(N, side, side, 30, seed=seed * 100 + b, v=rng.uniform(0, 1.5)))
    return cfgs, states


@pytest.mark.parametrize("G", [1, 2, 4])
def test_ragged_groups_per_agent_parity(gpu, G):
    # replicas of one class with uneven N share waves; empty slots in the last group
    Ns = [50, 50, 37, 12, 50, 1, 2, 64, 49, 3, 50]
    cfgs, states = sweep_batch(Ns, seed=G)
    sw = swarm(cfgs, states)
    sw.set_grouping(G)
    sw.step(fields=True, diag=True, verify=True)
    after = sw.state_numpy()
    for b, cfg in enumerate(cfgs):
        sl = sw.replica_slice(b)
        per_agent_check(sw, b, cfg, states[b], tuple(a[sl] for a in after), max_agents=20)
    pairs, _, disagree = sw.fallback_counts()
    assert pairs == sum(n * (n - 1) for n in Ns) and disagree == 0


@pytest.mark.parametrize("G", [1, 2, 4])
@pytest.mark.parametrize("R", [100, 256, 1024])
def test_ray_spans_for_every_resolution(gpu, G, R):
    # R = 100 leaves lanes without rays; R = 1024 gives 64 rays per lane at G = 4
    cfgs, states = sweep_batch([30, 24, 30, 17, 30], R=R, side=500.0, seed=R + G)
    sw = swarm(cfgs, states)
    sw.set_grouping(G)
    sw.step(fields=True, diag=True, verify=True)
    after = sw.state_numpy()
    for b, cfg in enumerate(cfgs):
        sl = sw.replica_slice(b)
        per_agent_check(sw, b, cfg, states[b], tuple(a[sl] for a in after), max_agents=12)
    assert sw.fallback_counts()[2] == 0


@pytest.mark.parametrize("G", [1, 2, 4])
def test_walls_fov_limits(gpu, G):
    from abm_amd.vf import VFConfig
    cfgs, states = [], []
    for b in range(6):
        cfgs.append(VFConfig.from_env_resolution(40, 512, agent_fov=0.5, width=300, height=300,
                                                 boundary="walls", LIMIT_MOVEMENT=True,
                                                 MAX_VEL=1.2, MAX_TH=0.05, ALP0=2.0 + 0.1 * b,
                                                 BET0=1.5))
        states.append(uniform_start(40, 300, 300, 30, seed=30 + b, v=1.0))
    sw = swarm(cfgs, states)
    sw.set_grouping(G)
    sw.step(fields=True, diag=True, verify=True)
    after = sw.state_numpy()
    for b, cfg in enumerate(cfgs):
        sl = sw.replica_slice(b)
        per_agent_check(sw, b, cfg, states[b], tuple(a[sl] for a in after), max_agents=16)
    assert sw.fallback_counts()[2] == 0


@pytest.mark.parametrize("G", [1, 2, 4])
def test_synchronous_order(gpu, G):
    cfgs, states = sweep_batch([50, 50, 50], R=512, side=400.0, seed=11)
    sw = swarm(cfgs, states)
    sw.set_grouping(G)
    sw.step(order="synchronous", fields=True)
    after = sw.state_numpy()
    for b, cfg in enumerate(cfgs):
        ox, oy, oth, ov, obits = vo.step_arrays(*states[b], 5, 512, cfg.fov, 400, 400, 30,
                                                vfparams(cfg), order="synchronous")
        assert np.array_equal(sw.fields_numpy(b), obits)
        sl = sw.replica_slice(b)
        assert pose_err(tuple(a[sl] for a in after), (ox, oy, oth, ov), cfg) <= TOL


def test_classes_never_share_a_wave(gpu):
    # two arena sizes and two ALP1 values in one batch: groups split by class
    from abm_amd.vf import VFConfig
    cfgs, states = [], []
    for b in range(10):
        side = 500.0 if b % 2 else 600.0
        cfgs.append(VFConfig(n_agents=20, n_rays=256, width=side, height=side, boundary="infinite",
                             ALP1=0.09 if b < 5 else 0.2, ALP0=1.0, BET0=1.0))
        states.append(uniform_start(20, side, side, 30, seed=500 + b, v=0.7))
    sw = swarm(cfgs, states)
    sw.set_grouping(4)
    sw.step(fields=True, diag=True)
    after = sw.state_numpy()
    for b, cfg in enumerate(cfgs):
        sl = sw.replica_slice(b)
        per_agent_check(sw, b, cfg, states[b], tuple(a[sl] for a in after), max_agents=8)


@pytest.mark.parametrize("G", [1, 2, 4])
def test_grouped_equals_per_replica_kernel_over_steps(gpu, G):
    # 64 sweep replicas for 20 steps: fields identical to the per-replica kernel
    # on the same input poses at every step; poses agree to float32 rounding
    cfgs, states = sweep_batch([50] * 64, seed=5)
    a = swarm(cfgs, states)
    b = swarm(cfgs, states)
    a.set_grouping(G)
    b.set_launch(1)
    for step in range(20):
        b.x.copy_(a.x); b.y.copy_(a.y); b.theta.copy_(a.theta); b.v.copy_(a.v)
        a.step(fields=True)
        b.step(fields=True)
        for r in range(0, 64, 7):
            assert np.array_equal(a.fields_numpy(r), b.fields_numpy(r)), (step, r)
        sa, sb = a.state_numpy(), b.state_numpy()
        assert pose_err(sa, sb, cfgs[0]) <= 2e-6, step


def test_autotune_and_grouping_errors(gpu):
    from abm_amd import _lib as L
    cfgs, states = sweep_batch([50] * 16, seed=6)
    sw = swarm(cfgs, states)
    chosen, ms = sw.autotune(trial_steps=2)
    assert chosen in (1, 2, 4, -1, -2, -4) and ms > 0
    with pytest.raises(L.AbmError):
        sw.set_grouping(3)
    # two resolutions in one batch: no grouped kernel
    from abm_amd.vf import VFConfig
    mixed = [VFConfig(n_agents=5, n_rays=256), VFConfig(n_agents=5, n_rays=512)]
    sw2 = swarm(mixed, [uniform_start(5, 500, 500, 30, seed=s) for s in (1, 2)])
    with pytest.raises(L.AbmUnsupported):
        sw2.set_grouping(2)
    sw2.set_grouping(-1)
    sw2.step()


@pytest.mark.parametrize("G", [1, 4])
def test_close_pairs_long_intervals(gpu, G):
    # a dense cluster: many pairs closer than 25 px draw 32+ rays (R = 512)
    from abm_amd.vf import VFConfig
    cfgs, states = [], []
    for b in range(4):
        cfgs.append(VFConfig(n_agents=40, n_rays=512, width=400, height=400, boundary="infinite",
                             ALP0=1.0, BET0=1.0))
        rng = np.random.RandomState(900 + b)
        x = rng.uniform(200, 260, 40).astype(np.float32).astype(np.float64)
        y = rng.uniform(200, 260, 40).astype(np.float32).astype(np.float64)
        th = rng.uniform(0, 2 * np.pi, 40).astype(np.float32).astype(np.float64)
        states.append((x, y, th, np.full(40, 0.5)))
    sw = swarm(cfgs, states)
    sw.set_grouping(G)
    sw.step(fields=True, diag=True, verify=True)
    after = sw.state_numpy()
    for b, cfg in enumerate(cfgs):
        sl = sw.replica_slice(b)
        per_agent_check(sw, b, cfg, states[b], tuple(a[sl] for a in after), max_agents=15)
    assert sw.fallback_counts()[2] == 0


def seam_batch(B=64, N=50, side=707.0, R=512, seed=1234):
    """Agents crowded into thin bands along the periodic seams (and their
    corner): many pairs are images across the seam at a few px."""
    from abm_amd.vf import VFConfig
    rng = np.random.RandomState(seed)
    cfgs, states = [], []
    for b in range(B):
        cfgs.append(VFConfig(n_agents=N, n_rays=R, width=side, height=side, boundary="infinite",
                             ALP0=1.0, BET0=1.0))
        lo, hi = 30.0 - 5.0, 30.0 + side - 5.0          # positions of centres at the seam
        x = np.where(rng.rand(N) < 0.5, rng.uniform(lo - 4, lo + 4, N), rng.uniform(hi - 4, hi + 4, N))
        y = np.where(rng.rand(N) < 0.5, rng.uniform(lo - 4, lo + 4, N), rng.uniform(30, 30 + side, N))
        th = rng.uniform(0, 2 * np.pi, N)
        f32 = lambda a: a.astype(np.float32).astype(np.float64)
        states.append((f32(x), f32(y), f32(th), np.full(N, 0.5)))
    return cfgs, states


@pytest.mark.parametrize("shape", [("launch", 1), ("group", 1), ("group", 4)])
def test_seam_pairs_decided_like_float64(gpu, shape):
    # wrapped pairs at a few px: the periodic image difference must be exact
    # enough that no float32 decision outside the guard bands differs from the
    # float64 one (verify mode), over several steps
    cfgs, states = seam_batch()
    sw = swarm(cfgs, states)
    sw.set_launch(shape[1]) if shape[0] == "launch" else sw.set_grouping(shape[1])
    for _ in range(3):
        sw.step(verify=True)
    pairs, rechecked, disagree = sw.fallback_counts()
    assert pairs > 0 and disagree == 0, (pairs, rechecked, disagree)
    # and the first step's fields against the oracle for a few replicas
    sw2 = swarm(cfgs[:3], states[:3])
    sw2.set_launch(shape[1]) if shape[0] == "launch" else sw2.set_grouping(shape[1])
    sw2.step(fields=True, diag=True)
    after = sw2.state_numpy()
    for b in range(3):
        sl = sw2.replica_slice(b)
        per_agent_check(sw2, b, cfgs[b], states[b], tuple(a[sl] for a in after), max_agents=10)


def tie_batch(B=256, N=50, R=512, side=707.0, seed=99):
    """Agent 0 of each replica sees the other agents at closed angles within
    ~1e-6 rad of bin ties (midpoints of linspace(-pi, pi, R)), so nearly all
    its pairs fall inside the float32 tie band."""
    from abm_amd.vf import VFConfig
    rng = np.random.RandomState(seed)
    phis = np.linspace(-np.pi, np.pi, R)
    mids = 0.5 * (phis[:-1] + phis[1:])
    cfgs, states = [], []
    for b in range(B):
        cfgs.append(VFConfig(n_agents=N, n_rays=R, width=side, height=side, boundary="infinite",
                             ALP0=1.0, BET0=1.0))
        th0 = float(np.float32(rng.uniform(0, 2 * np.pi)))
        c0 = np.array([rng.uniform(200, 500), rng.uniform(200, 500)])     # centre of agent 0
        ca = rng.choice(mids[5:-5], N - 1) + rng.uniform(-1e-6, 1e-6, N - 1)
        phi = ca + th0                                                     # world angle, y up
        d = rng.uniform(20, 150, N - 1)
        cx = c0[0] + d * np.cos(phi)
        cy = c0[1] - d * np.sin(phi)
        x = np.concatenate([[c0[0]], cx]) - 5.0
        y = np.concatenate([[c0[1]], cy]) - 5.0
        th = np.concatenate([[th0], rng.uniform(0, 2 * np.pi, N - 1)])
        f32 = lambda a: a.astype(np.float32).astype(np.float64)
        states.append((f32(x), f32(y), f32(th), np.full(N, 0.5)))
    return cfgs, states


@pytest.mark.parametrize("shape", [("launch", 1), ("group", 1), ("group", 4)])
def test_tie_pairs_decided_like_float64(gpu, shape):
    # the float64 side test of tie-flagged pairs against the float64 replica of
    # the reference arithmetic, thousands of times (verify mode), and the fields
    # of the agents that see the ties against the oracle
    cfgs, states = tie_batch()
    sw = swarm(cfgs, states)
    sw.set_launch(shape[1]) if shape[0] == "launch" else sw.set_grouping(shape[1])
    sw.step(verify=True)
    pairs, rechecked, disagree = sw.fallback_counts()
    assert rechecked > 256 * 30 and disagree == 0, (pairs, rechecked, disagree)
    sw2 = swarm(cfgs[:4], states[:4])
    sw2.set_launch(shape[1]) if shape[0] == "launch" else sw2.set_grouping(shape[1])
    sw2.step(fields=True, diag=True)
    after = sw2.state_numpy()
    for b in range(4):
        sl = sw2.replica_slice(b)
        per_agent_check(sw2, b, cfgs[b], states[b], tuple(a[sl] for a in after), max_agents=3)


@pytest.mark.parametrize("G", [1, 2, 4])
def test_plain_step_equals_step_with_outputs(gpu, G):
    # a step that asks for no field, diagnostics or counters runs the kernel's
    # plain instantiation (the output tests compiled out); its poses and
    # speeds equal those of the same kernel writing fields, and of a
    # verify-mode step, bit for bit, step after step (ragged N, ALP0 / BET0
    # per replica, torus).  (Diagnostics total the four sums separately, a
    # rounding apart from the two integrals, so they are not compared here.)
    import torch
    cfgs, states = sweep_batch([50, 37, 64, 12, 50, 61, 3, 50] * 8, seed=9)
    plain, full, ver = swarm(cfgs, states), swarm(cfgs, states), swarm(cfgs, states)
    for s in (plain, full, ver):
        s.set_grouping(G)
    for step in range(6):
        print(f"  plain vs outputs G={G} step {step}", flush=True)
        plain.step()
        full.step(fields=True)
        ver.step(verify=True)
        assert plain.last_kernels() == {"grouped"} and full.last_kernels() == {"grouped"}
        torch.cuda.synchronize()
        for s in (full, ver):
            for p, q in zip(plain.state_numpy(), s.state_numpy()):
                assert np.array_equal(p, q), step
    _, _, disagree = ver.fallback_counts()
    assert disagree == 0
