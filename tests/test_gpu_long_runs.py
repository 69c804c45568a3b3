"""The benchmark configurations run for their stated lengths, with parity
checked along the way (the reference's order, vf_sims.py:302).

* C2, one flock N = 100, R = 1024, 1000 px torus, 5000 steps: every 500
  steps the float64 kernel's next step equals the oracle's from the same
  state (fields bit-exact, poses within 1e-12), the float32 kernel's next
  synchronous step equals the float64 kernel's fields bit for bit with a pose
  error within 1e-5 (default and Kahan sums), and the free-running float32
  kernel decides every pair of a step as float64 does (verify mode).
* C3, the 4096-replica sweep, 2000 steps: every 400 steps verify mode and
  the float32 / float64 synchronous comparison over all replicas; at the end
  sampled agents of two replicas against the oracle on the poses they saw.
* C4's flocking half, 1000 steps of the 1000 ragged replicas: verify mode
  every 250 steps, and the largest replica's sampled agents against the
  oracle at the end.
* C5, the 8192-agent flock in the reference order for 1000 steps, verify
  mode at the end and sampled agents against the oracle."""
import numpy as np
import pytest

from oracle import vf_oracle as vo
from test_gpu_vf_f32 import per_agent_check, pose_err, vfparams

pytestmark = pytest.mark.gpu
PI = np.pi
f32r = lambda a: np.asarray(a).astype(np.float32).astype(np.float64)


def sync_f32_vs_f64(cfgs, st, radius, kahan=False):
    import torch
    from abm_amd.vf import VFSwarm
    r = [f32r(a) for a in st]
    a = VFSwarm(cfgs, *r, radius=radius, precision="f32")
    b = VFSwarm(cfgs, *r, radius=radius, precision="f64")
    a.step(order="synchronous", fields=True, kahan=kahan)
    b.step(order="synchronous", fields=True)
    torch.cuda.synchronize()
    assert torch.equal(a.fields, b.fields)
    return a.state_numpy(), b.state_numpy()


def test_c2_5000_steps(gpu):
    from abm_amd.vf import VFConfig, VFSwarm
    N, R, W = 100, 1024, 1000.0
    rng = np.random.RandomState(42)
    x, y = f32r(rng.uniform(30, 30 + W, N)), f32r(rng.uniform(30, 30 + W, N))
    th = f32r(rng.uniform(0, 2 * PI, N))
    cfg = VFConfig(n_agents=N, n_rays=R, width=W, height=W, boundary="infinite", ALP0=1.0, BET0=0.5)
    p = vfparams(cfg)
    s64 = VFSwarm([cfg], x, y, th, 0.0, radius=5, precision="f64")
    s32 = VFSwarm([cfg], x, y, th, 0.0, radius=5, precision="f32")
    for t in range(5001):
        if t % 500 == 0:
            print(f"  C2 step {t}", flush=True)
            st = s64.state_numpy()
            ox, oy, oth, ov, obits = vo.step_arrays(*st, 5, R, cfg.fov, W, W, 30, p)
            one = VFSwarm([cfg], *st, radius=5, precision="f64")
            one.step(fields=True)
            assert np.array_equal(one.fields_numpy(0), obits), t
            assert pose_err(one.state_numpy(), (ox, oy, oth, ov), cfg) <= 1e-12, t
            for kahan in (False, True):
                a, b = sync_f32_vs_f64([cfg], st, 5, kahan)
                assert pose_err(a, b, cfg) <= 1e-5, (t, kahan)
            s32.step(verify=True)
        else:
            s32.step()
        s64.step()
    pairs, _, disagree = s32.fallback_counts()
    assert disagree == 0 and pairs > 0


def test_c3_2000_steps(gpu):
    from abm_amd.vf import VFSwarm
    from bench import make_workload
    cfgs, x, y, th = make_workload()
    sw = VFSwarm(cfgs, x, y, th, 0.5, radius=5, precision="f32")
    for t in range(2000):
        if t % 400 == 0:
            print(f"  C3 step {t}", flush=True)
            a, b = sync_f32_vs_f64(cfgs, sw.state_numpy(), 5)
            assert pose_err(a, b, cfgs[0]) <= 1e-5, t
            sw.step(verify=True)
        else:
            sw.step()
    pairs, _, disagree = sw.fallback_counts()
    assert disagree == 0 and pairs == 5 * 4096 * 50 * 49
    before = sw.state_numpy()
    sw.step(fields=True, diag=True)
    after = sw.state_numpy()
    for b in (0, 4095):
        sl = sw.replica_slice(b)
        per_agent_check(sw, b, cfgs[b], tuple(q[sl] for q in before), tuple(q[sl] for q in after))


def test_c4_flocking_half_1000_steps(gpu):
    from abm_amd.vf import VFConfig, VFSwarm
    from bench import mixed_shapes
    cfgs, states = [], []
    rng = np.random.RandomState(31)
    for b, (n, r) in enumerate(mixed_shapes()):
        if b % 2 == 0:
            continue
        side = float(100.0 * np.sqrt(n))
        cfgs.append(VFConfig(n_agents=n, n_rays=r, width=side, height=side, boundary="infinite",
                             ALP0=1.0, BET0=0.5))
        states.append((f32r(rng.uniform(30, 30 + side, n)), f32r(rng.uniform(30, 30 + side, n)),
                       f32r(rng.uniform(0, 2 * PI, n)), np.full(n, 0.5)))
    cat = [np.concatenate([s[i] for s in states]) for i in range(4)]
    sw = VFSwarm(cfgs, *cat, radius=5, precision="f32")
    for t in range(1000):
        if t % 250 == 0:
            print(f"  C4 step {t} kernels {sw.last_kernels()}", flush=True)
            sw.step(verify=True)
        else:
            sw.step()
    pairs, _, disagree = sw.fallback_counts()
    assert disagree == 0 and pairs == 4 * sum(c.n_agents * (c.n_agents - 1) for c in cfgs)
    big = int(np.argmax([c.n_agents for c in cfgs]))
    before = sw.state_numpy()
    sw.step(fields=True, diag=True)
    after = sw.state_numpy()
    sl = sw.replica_slice(big)
    n = cfgs[big].n_agents
    agents = sorted({0, n // 2, n - 1} | set(np.random.RandomState(3).choice(n, 8, replace=False).tolist()))
    per_agent_check(sw, big, cfgs[big], tuple(q[sl] for q in before), tuple(q[sl] for q in after),
                    agents=agents)


def test_c5_flock_1000_steps_reference_order(gpu):
    from abm_amd.vf import VFConfig, VFSwarm
    N, R, W = 8192, 2048, 4000.0
    rng = np.random.RandomState(7)
    x, y = f32r(rng.uniform(30, 30 + W, N)), f32r(rng.uniform(30, 30 + W, N))
    th = f32r(rng.uniform(0, 2 * PI, N))
    cfg = VFConfig(n_agents=N, n_rays=R, width=W, height=W, boundary="infinite", ALP0=1.0, BET0=0.5)
    sw = VFSwarm([cfg], x, y, th, 0.5, radius=2, precision="f32")
    for t in range(999):
        if t % 100 == 0:
            print(f"  C5 step {t}", flush=True)
        sw.step()
    before = sw.state_numpy()
    sw.step(fields=True, diag=True, verify=True)
    pairs, _, disagree = sw.fallback_counts()
    assert disagree == 0 and pairs == N * (N - 1)
    after = sw.state_numpy()
    agents = [0, 1, 4096, 8190, 8191]
    per_agent_check(sw, 0, cfg, before, after, radius=2, agents=agents)
