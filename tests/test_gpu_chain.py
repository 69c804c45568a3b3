"""The chain step: single flocks above 2048 agents in the reference (id) order
(vf_sims.py:302 -> vf_agent.py:52-77).

Each agent's pairs with the agents after it are decided at their start poses
over the whole GPU first; the agents are then moved in id order in segments of
256 (one workgroup), and each segment's new poses are decided in the fields of
every later agent over the whole GPU (through the grid of start poses).  Every
decision is exact and the segments sum and move as the push kernel does, so
with the chain step forced on a flock the push kernel also holds (up to 2048
agents) the two agree bit for bit: fields, poses, speeds, the integrals.
Against the 16-wave kernel, which adds the same ray values in another order,
the poses agree to float32 rounding."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu
PI = np.pi
f32 = lambda a: np.asarray(a, dtype=np.float64).astype(np.float32).astype(np.float64)


def flock(N, R, side, seed, boundary="infinite", seam=False, radius=2, params=False):
    from abm_amd.vf import VFConfig
    rng = np.random.RandomState(seed)
    if seam:   # half the agents within 4 px of the periodic seam in x
        lo, hi = 25.0, 25.0 + side
        x = np.where(rng.rand(N) < 0.5, rng.uniform(lo - 4, lo + 4, N), rng.uniform(hi - 4, hi + 4, N))
    else:
        x = rng.uniform(30, 30 + side, N)
    y = rng.uniform(30, 30 + side, N)
    th, v = rng.uniform(0, 2 * PI, N), rng.uniform(0, 1.5, N)
    if radius == "mixed":
        radius = rng.randint(1, 7, N)
    extra = {}
    if params:   # per-agent ALP0 / BET0 / V0
        extra = dict(alp0=f32(rng.uniform(0.5, 1.5, N)), bet0=f32(rng.uniform(0.2, 0.8, N)),
                     v0=f32(rng.uniform(0.5, 1.5, N)))
    cfg = VFConfig(n_agents=N, n_rays=R, width=side, height=side, boundary=boundary, ALP0=1.0, BET0=0.5)
    return cfg, (f32(x), f32(y), f32(th), f32(v)), radius, extra


PUSH_CASES = {
    # (sizes the push kernel holds: every field in one CU's LDS)
    # two segments on a grid (torus, three cells a side)
    "grid": dict(N=450, R=2048, side=2000.0, seed=11),
    # agents crowded on the seam of a 3 x 3-cell torus, four segments
    "seam": dict(N=800, R=1024, side=1500.0, seed=5, seam=True, radius=3),
    # walls: no grid, every later agent tested
    "walls": dict(N=1200, R=512, side=1400.0, seed=8, boundary="walls", radius=5),
    # a torus too small for three cells a side: no grid
    "small_torus": dict(N=400, R=2048, side=1000.0, seed=9),
    # radii 1..6 (the grid's reach follows the largest) and per-agent parameters
    "mixed": dict(N=600, R=1536, side=4500.0, seed=12, radius="mixed", params=True),
    # one partial segment
    "short": dict(N=300, R=1024, side=600.0, seed=4, radius=4),
}


@pytest.mark.parametrize("diag", [False, True])
@pytest.mark.parametrize("case", list(PUSH_CASES))
def test_chain_matches_push_kernel(gpu, case, diag):
    from abm_amd.vf import VFSwarm
    cfg, st, radius, extra = flock(**PUSH_CASES[case])
    a = VFSwarm([cfg], *st, radius=radius, precision="f32", **extra)
    b = VFSwarm([cfg], *st, radius=radius, precision="f32", **extra)
    a.set_launch(64)
    b.set_launch(32)
    for step in range(3):
        print(f"  chain {case} diag={diag} step {step}", flush=True)
        a.step(order="sequential", fields=True, diag=diag)
        b.step(order="sequential", fields=True, diag=diag)
        assert a.last_kernels() == {"chain"} and b.last_kernels() == {"push"}
        assert np.array_equal(a.fields_numpy(0), b.fields_numpy(0)), step
        for p, q in zip(a.state_numpy(), b.state_numpy()):
            assert np.array_equal(p, q), step
        if diag:
            assert np.array_equal(a.diag_numpy(), b.diag_numpy()), step


BIG_CASES = {
    "c5": dict(N=8192, R=2048, side=4000.0, seed=11),
    "walls": dict(N=3000, R=512, side=2400.0, seed=8, boundary="walls", radius=5),
    "mixed": dict(N=4096, R=1536, side=3000.0, seed=12, radius="mixed", params=True),
}


@pytest.mark.parametrize("case", list(BIG_CASES))
def test_chain_agrees_with_16_wave_kernel(gpu, case):
    # one step from the same state: the fields equal but for the few agents
    # whose bins the slightly different poses of earlier agents tip (one ray
    # changes an integral by about R^-1); the other agents' poses within
    # float32 rounding of the sums' order (relative 1e-5)
    from abm_amd.vf import VFSwarm
    from test_gpu_vf_f32 import pose_err
    cfg, st, radius, extra = flock(**BIG_CASES[case])
    a = VFSwarm([cfg], *st, radius=radius, precision="f32", **extra)
    b = VFSwarm([cfg], *st, radius=radius, precision="f32", **extra)
    b.set_launch(16)
    a.step(order="sequential", fields=True)
    b.step(order="sequential", fields=True)
    assert a.last_kernels() == {"chain"} and b.last_kernels() == {"big"}
    same = np.all(a.fields_numpy(0) == b.fields_numpy(0), axis=1)
    print(f"  {case}: {int((~same).sum())} of {cfg.n_agents} fields differ", flush=True)
    assert (~same).sum() <= max(2, cfg.n_agents // 1000)
    sa, sb = a.state_numpy(), b.state_numpy()
    assert pose_err(tuple(q[same] for q in sa), tuple(q[same] for q in sb), cfg) <= 1e-5


@pytest.mark.parametrize("case", ["c5", "walls"])
def test_chain_verify_mode(gpu, case):
    # counters and verify mode take the full scans: every pair is counted
    # once, every float32 decision and every dropped pair agrees with float64,
    # and the result equals the default path's
    from abm_amd.vf import VFSwarm
    cfg, st, radius, extra = flock(**BIG_CASES[case])
    a = VFSwarm([cfg], *st, radius=radius, precision="f32", **extra)
    b = VFSwarm([cfg], *st, radius=radius, precision="f32", **extra)
    a.step(order="sequential", fields=True, verify=True)
    b.step(order="sequential", fields=True)
    assert a.last_kernels() == {"chain"}
    pairs, _, disagree = a.fallback_counts()
    N = cfg.n_agents
    assert disagree == 0 and pairs == N * (N - 1)
    assert np.array_equal(a.fields_numpy(0), b.fields_numpy(0))
    for p, q in zip(a.state_numpy(), b.state_numpy()):
        assert np.array_equal(p, q)


def test_chain_sampled_agents_match_oracle(gpu):
    # N = 2500 > 2048 on a 2400 px torus (grid path): sampled agents' fields
    # and poses against the oracle on the poses each agent saw, among them
    # both sides of segment boundaries
    from abm_amd.vf import VFSwarm
    from test_gpu_vf_f32 import per_agent_check
    cfg, st, radius, extra = flock(N=2500, R=2048, side=2400.0, seed=3)
    sw = VFSwarm([cfg], *st, radius=radius, precision="f32")
    sw.step(order="sequential", fields=True, diag=True)
    assert sw.last_kernels() == {"chain"}
    agents = sorted({0, 1, 2, 255, 256, 257, 511, 512, 1249, 2303, 2304, 2497, 2498, 2499} |
                    set(np.random.RandomState(1).choice(2500, 12, replace=False).tolist()))
    per_agent_check(sw, 0, cfg, st, sw.state_numpy(), radius=2, agents=agents)


def test_chain_beyond_the_16_wave_kernels_lds(gpu):
    # 12,000 agents (R = 1024, C5's density): more poses than the 16-wave
    # kernel holds in LDS (about 8,900), which the chain step does not need
    # (its fields live in global memory).  One step in verify mode (every
    # pair counted once, none disagreeing with float64), then one with fields
    # whose sampled agents match the oracle on the poses they saw.
    from abm_amd import _lib as L
    from abm_amd.vf import VFSwarm
    from test_gpu_vf_f32 import per_agent_check
    cfg, st, radius, extra = flock(N=12000, R=1024, side=4800.0, seed=21)
    big = VFSwarm([cfg], *st, radius=radius, precision="f32")
    big.set_launch(16)
    with pytest.raises(L.AbmUnsupported, match="LDS"):
        big.step(order="sequential")
    sw = VFSwarm([cfg], *st, radius=radius, precision="f32")
    sw.step(order="sequential", verify=True)
    assert sw.last_kernels() == {"chain"}
    pairs, _, disagree = sw.fallback_counts()
    assert disagree == 0 and pairs == 12000 * 11999
    before = sw.state_numpy()
    sw.step(order="sequential", fields=True, diag=True)
    agents = [0, 1, 255, 256, 6000, 11775, 11776, 11998, 11999]
    per_agent_check(sw, 0, cfg, before, sw.state_numpy(), radius=2, agents=agents)


def test_chain_batch_of_replicas(gpu):
    # the chain step runs a batch replica by replica, reusing its grid and
    # field buffers: a batch of three forced onto it equals the push kernel,
    # and two flocks above 2048 agents in one batch equal each flock alone
    # (one R per batch: the summing lanes' ray spans follow the batch's
    # largest R)
    from abm_amd.vf import VFSwarm
    parts = [flock(N=300, R=1024, side=700.0, seed=41, radius=4),
             flock(N=450, R=1024, side=2000.0, seed=42),
             flock(N=129, R=1024, side=500.0, seed=43, boundary="walls", radius=3)]
    cfgs = [p[0] for p in parts]
    cat = [np.concatenate([p[1][i] for p in parts]) for i in range(4)]
    rad = np.concatenate([np.full(p[0].n_agents, p[2]) for p in parts])
    a = VFSwarm(cfgs, *cat, radius=rad, precision="f32")
    b = VFSwarm(cfgs, *cat, radius=rad, precision="f32")
    a.set_launch(64)
    b.set_launch(32)
    for step in range(2):
        a.step(order="sequential", fields=True)
        b.step(order="sequential", fields=True)
        assert a.last_kernels() == {"chain"} and b.last_kernels() == {"push"}
        for r in range(3):
            assert np.array_equal(a.fields_numpy(r), b.fields_numpy(r)), (step, r)
        for p, q in zip(a.state_numpy(), b.state_numpy()):
            assert np.array_equal(p, q), step

    big = [flock(N=2100, R=2048, side=2200.0, seed=44), flock(N=2600, R=2048, side=2500.0, seed=45)]
    cat = [np.concatenate([p[1][i] for p in big]) for i in range(4)]
    both = VFSwarm([p[0] for p in big], *cat, radius=2, precision="f32")
    alone = [VFSwarm([p[0]], *p[1], radius=2, precision="f32") for p in big]
    both.step(order="sequential", fields=True)
    for s in alone:
        s.step(order="sequential", fields=True)
    assert both.last_kernels() == {"chain"}
    st = both.state_numpy()
    for r, s in enumerate(alone):
        assert np.array_equal(both.fields_numpy(r), s.fields_numpy(0)), r
        sl = both.replica_slice(r)
        for p, q in zip(st, s.state_numpy()):
            assert np.array_equal(p[sl], q), r
