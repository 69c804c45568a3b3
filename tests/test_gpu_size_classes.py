"""Launch shapes of ragged VSWRM batches: the size classes (replicas above
64 agents on the push kernel, the rest on the per-replica kernel), the
16-wave kernel for every replica and the per-replica kernel alone.  Each
decides every pair as float64 does (verify mode, three steps, no
disagreement), and each agent of a sample matches the oracle on the poses
it saw (reference order, vf_sims.py:302): field bit-exact, one-step pose
error within 1e-5.  (Across shapes the sums run in other orders, so later
agents may see float32 poses a rounding apart; the shapes are compared
with the oracle, not with each other.)"""
import numpy as np
import pytest

from test_gpu_vf_f32 import per_agent_check, uniform_start

pytestmark = pytest.mark.gpu


def ragged(seed=41):
    from abm_amd.vf import VFConfig
    rng = np.random.RandomState(seed)
    cfgs, states = [], []
    for b, n in enumerate([8, 40, 129, 200, 300, 512, 64, 150]):
        r = int(rng.choice([256, 512, 1024]))
        side = float(100.0 * np.sqrt(n))
        cfgs.append(VFConfig(n_agents=n, n_rays=r, width=side, height=side, boundary="infinite",
                             ALP0=1.0, BET0=0.5))
        states.append(uniform_start(n, side, side, 30, seed=seed * 100 + b, v=0.6))
    return cfgs, states


def test_size_classes_and_launch_shapes_match_oracle(gpu):
    from abm_amd.vf import VFSwarm
    cfgs, states = ragged()
    cat = [np.concatenate([s[i] for s in states]) for i in range(4)]
    for shape in (0, 16, 4):
        sw = VFSwarm(cfgs, *cat, radius=5, precision="f32")
        sw.set_launch(shape)
        for step in range(3):
            sw.step(fields=True, diag=True, verify=True)
        pairs, _, disagree = sw.fallback_counts()
        assert disagree == 0 and pairs == 3 * sum(c.n_agents * (c.n_agents - 1) for c in cfgs), shape
    # one step per shape from the same start, each agent of a sample against the oracle
    for shape in (0, 16, 4):
        sw = VFSwarm(cfgs, *cat, radius=5, precision="f32")
        sw.set_launch(shape)
        sw.step(fields=True, diag=True)
        after = sw.state_numpy()
        for b, cfg in enumerate(cfgs):
            sl = sw.replica_slice(b)
            n = cfg.n_agents
            agents = sorted({0, n // 2, n - 1} | set(np.random.RandomState(b).choice(n, min(n, 5), replace=False).tolist()))
            per_agent_check(sw, b, cfg, states[b], tuple(a[sl] for a in after), agents=agents)
