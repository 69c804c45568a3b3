"""Single large flocks (more agents than the per-replica kernel holds).

Reference (id) order runs one workgroup per replica with every pose in LDS
(vf_sims.py:302 -> vf_agent.py:52-77): every float32 decision and every pair
dropped by the visibility cut-off agree with float64 (verify counters), and
sampled agents match the oracle on the poses they saw (agents < k moved).
Synchronous order takes the wide kernel: the same checks on the pre-step
poses."""
import numpy as np
import pytest

from oracle import vf_oracle as vo

pytestmark = pytest.mark.gpu
PI = np.pi


def flock(N, R, W, seed, r=2):
    from abm_amd.vf import VFConfig, VFSwarm
    rng = np.random.RandomState(seed)
    f = lambda a: a.astype(np.float32).astype(np.float64)
    x, y = f(rng.uniform(30, 30 + W, N)), f(rng.uniform(30, 30 + W, N))
    th, v = f(rng.uniform(0, 2 * PI, N)), f(rng.uniform(0, 1.5, N))
    cfg = VFConfig(n_agents=N, n_rays=R, width=W, height=W, boundary="infinite", ALP0=1.0, BET0=0.5)
    return cfg, (x, y, th, v), VFSwarm([cfg], x, y, th, v, radius=r, precision="f32")


def test_reference_order_large_flock_matches_oracle(gpu):
    # N = 2500 > 2048, R = 2048, radius 2, 2400 px torus: one sequential step
    # in verify mode, then every sampled agent's field and pose against the
    # oracle on the poses that agent saw
    from test_gpu_vf_f32 import per_agent_check
    cfg, st, sw = flock(2500, 2048, 2400, 3)
    sw.step(order="sequential", fields=True, diag=True, verify=True)
    pairs, rechecked, disagree = sw.fallback_counts()
    assert disagree == 0 and pairs == 2500 * 2499
    agents = sorted({0, 1, 1249, 2498, 2499} | set(np.random.RandomState(1).choice(2500, 12, replace=False).tolist()))
    per_agent_check(sw, 0, cfg, st, sw.state_numpy(), radius=2, agents=agents)


def test_reference_order_c5_flock_sampled_agents(gpu):
    # the large flock itself (N = 8192, R = 2048, 4000 px torus) in the
    # reference order for two steps; on the second, sampled agents against
    # the oracle on the poses they saw
    from test_gpu_vf_f32 import per_agent_check
    cfg, st, sw = flock(8192, 2048, 4000, 11)
    sw.step(order="sequential")
    before = sw.state_numpy()
    sw.step(order="sequential", fields=True, diag=True)
    agents = sorted({0, 4096, 8191} | set(np.random.RandomState(2).choice(8192, 5, replace=False).tolist()))
    per_agent_check(sw, 0, cfg, before, sw.state_numpy(), radius=2, agents=agents)


def test_large_flock_sampled_agents_match_oracle(gpu):
    cfg, (x, y, th, v), sw = flock(3000, 2048, 2400, 7)
    sw.step(order="synchronous", fields=True, diag=True, verify=True)
    pairs, rechecked, disagree = sw.fallback_counts()
    assert disagree == 0 and pairs == 3000 * 2999
    bits = sw.fields_numpy(0)
    nx, ny, nth, nv = sw.state_numpy()
    p = vo.VFParams(ALP0=1.0, BET0=0.5, BOUNDARY="infinite")
    for k in np.random.RandomState(0).choice(3000, 12, replace=False):
        ox, oy, oth, ov, obits, _ = vo.agent_view_step(x, y, th, v, 2, k, 2048, cfg.fov, 2400, 2400,
                                                       30, p)
        assert np.array_equal(bits[k], obits), k
        d = [abs((nx[k] - ox + 1200) % 2400 - 1200) / max(abs(ox), 1),
             abs((ny[k] - oy + 1200) % 2400 - 1200) / max(abs(oy), 1),
             abs((nth[k] - oth + PI) % (2 * PI) - PI) / max(abs(oth), 1),
             abs(nv[k] - ov) / max(abs(ov), 1)]
        assert max(d) <= 1e-5, (k, d)


def test_c5_flock_verify(gpu):
    # N=8192, radius 2, 4000x4000 torus, R=2048: every decision checked
    cfg, st, sw = flock(8192, 2048, 4000, 11)
    for _ in range(3):
        sw.step(order="synchronous", verify=True)
    pairs, rechecked, disagree = sw.fallback_counts()
    assert disagree == 0
    assert pairs == 3 * 8192 * 8191


def test_wide_kernel_seam_pairs(gpu):
    # the wide (synchronous, N > 2048) kernel shares the pair decision: crowd
    # agents at the periodic seams and check every decision in verify mode
    from abm_amd.vf import VFConfig, VFSwarm
    N, side = 2304, 1500.0
    rng = np.random.RandomState(77)
    lo, hi = 25.0, 25.0 + side
    x = np.where(rng.rand(N) < 0.5, rng.uniform(lo - 4, lo + 4, N), rng.uniform(hi - 4, hi + 4, N))
    y = rng.uniform(30, 30 + side, N)
    th = rng.uniform(0, 2 * np.pi, N)
    f32 = lambda a: a.astype(np.float32).astype(np.float64)
    cfg = VFConfig(n_agents=N, n_rays=1024, width=side, height=side, boundary="infinite")
    sw = VFSwarm([cfg], f32(x), f32(y), f32(th), 0.5, radius=3, precision="f32")
    sw.step(order="synchronous", verify=True)
    pairs, rechecked, disagree = sw.fallback_counts()
    assert pairs == N * (N - 1) and disagree == 0


@pytest.mark.parametrize("N,R,side,r,seam", [(8192, 2048, 4000.0, 2, False), (2304, 1024, 1500.0, 3, True),
                                             (4096, 512, 3000.0, 5, False)])
def test_wide_grid_path_matches_full_scan(gpu, N, R, side, r, seam):
    # without counters or verify mode the wide step takes only the focal
    # agent's 3 x 3 grid cells (cells wider than the cut-off); fields and
    # poses must equal the full scan's (verify mode) bit for bit, here over
    # three steps, also with agents crowded on the periodic seams (3 cells)
    from abm_amd.vf import VFConfig, VFSwarm
    rng = np.random.RandomState(5 + N)
    f32 = lambda a: a.astype(np.float32).astype(np.float64)
    if seam:
        lo, hi = 25.0, 25.0 + side
        x = np.where(rng.rand(N) < 0.5, rng.uniform(lo - 4, lo + 4, N), rng.uniform(hi - 4, hi + 4, N))
    else:
        x = rng.uniform(30, 30 + side, N)
    y = rng.uniform(30, 30 + side, N)
    th, v = rng.uniform(0, 2 * PI, N), rng.uniform(0, 1.5, N)
    cfg = VFConfig(n_agents=N, n_rays=R, width=side, height=side, boundary="infinite", ALP0=1.0, BET0=0.5)
    a = VFSwarm([cfg], f32(x), f32(y), f32(th), f32(v), radius=r, precision="f32")
    b = VFSwarm([cfg], f32(x), f32(y), f32(th), f32(v), radius=r, precision="f32")
    for step in range(3):
        a.step(order="synchronous", fields=True, verify=True)
        b.step(order="synchronous", fields=True)
        assert np.array_equal(a.fields_numpy(0), b.fields_numpy(0)), step
        for p, q in zip(a.state_numpy(), b.state_numpy()):
            assert np.array_equal(p, q), step
    assert a.fallback_counts()[2] == 0
