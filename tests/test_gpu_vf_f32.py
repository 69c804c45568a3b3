"""GPU parity of the fused float32 VSWRM path.

Bar:
* every float32 ray-bin / half-width decision equals the float64 one
  (verify mode counts disagreements over whole runs: must be 0);
* each agent's field is bit-exact against the oracle evaluated on the exact
  poses that agent saw (agents < k already moved, float32 poses promoted);
* one-step relative pose error <= 1e-5 (scale 1 px for x, y; 1 rad for
  theta, compared modulo 2 pi; V0 for v) against the float64 oracle.
"""
import numpy as np
import pytest

from oracle import vf_oracle as vo

pytestmark = pytest.mark.gpu
PI = np.pi
TOL = 1e-5


def uniform_start(N, W, H, pad, seed, v=0.0):
    rng = np.random.RandomState(seed)
    return (rng.uniform(pad, pad + W, N).astype(np.float32).astype(np.float64),
            rng.uniform(pad, pad + H, N).astype(np.float32).astype(np.float64),
            rng.uniform(0, 2 * PI, N).astype(np.float32).astype(np.float64),
            np.full(N, v, dtype=np.float32).astype(np.float64))


def vfparams(cfg):
    return vo.VFParams(GAM=cfg.GAM, V0=cfg.V0, ALP0=cfg.ALP0, ALP1=cfg.ALP1, BET0=cfg.BET0,
                       BET1=cfg.BET1, BOUNDARY=cfg.boundary, LIMIT_MOVEMENT=cfg.LIMIT_MOVEMENT,
                       MAX_VEL=cfg.MAX_VEL, MAX_TH=cfg.MAX_TH)


def pose_err(g, o, cfg):
    """Max one-step relative error over (x, y, theta, v); torus positions and
    angles compared on the circle."""
    gx, gy, gth, gv = g
    ox, oy, oth, ov = o
    dx, dy = np.asarray(gx) - ox, np.asarray(gy) - oy
    if cfg.boundary == "infinite":
        dx = (dx + cfg.width / 2) % cfg.width - cfg.width / 2
        dy = (dy + cfg.height / 2) % cfg.height - cfg.height / 2
    dth = (np.asarray(gth) - oth + PI) % (2 * PI) - PI
    return max(np.max(np.abs(dx) / np.maximum(np.abs(ox), 1.0)),
               np.max(np.abs(dy) / np.maximum(np.abs(oy), 1.0)),
               np.max(np.abs(dth) / np.maximum(np.abs(oth), 1.0)),
               np.max(np.abs(np.asarray(gv) - ov) / np.maximum(np.abs(ov), cfg.V0)))


def per_agent_check(sw, b, cfg, before, after, radius=5, overrides=None, max_agents=None,
                    agents=None):
    """Each agent k against the oracle on the poses it saw."""
    bits = sw.fields_numpy(b)
    diag = sw.diag_numpy()[sw.replica_slice(b)]
    x0, y0, t0, v0 = before
    x1, y1, t1, v1 = after
    N = len(x0)
    p = vfparams(cfg)
    worst = 0.0
    ov = overrides or {}
    if agents is None:
        agents = range(N if max_agents is None else min(N, max_agents))
    for k in agents:
        vx = np.concatenate([x1[:k], x0[k:]])
        vy = np.concatenate([y1[:k], y0[k:]])
        vt = np.concatenate([t1[:k], t0[k:]])
        vv = np.concatenate([v1[:k], v0[k:]])
        ox, oy, oth, ovel, obits, odiag = vo.agent_view_step(
            vx, vy, vt, vv, radius, k, cfg.n_rays, cfg.fov, cfg.width, cfg.height,
            cfg.window_pad, p, **ov)
        assert np.array_equal(bits[k], obits), f"agent {k}: field differs"
        e = pose_err(([x1[k]], [y1[k]], [t1[k]], [v1[k]]), (ox, oy, oth, ovel), cfg)
        worst = max(worst, e)
        assert np.allclose(diag[k], odiag, rtol=1e-4, atol=2e-6), (k, diag[k], odiag)
    assert worst <= TOL, worst
    return worst


def run_one(cfgs, states, radius=5, order="sequential", verify=True, **kw):
    from abm_amd.vf import VFSwarm
    x = np.concatenate([s[0] for s in states])
    y = np.concatenate([s[1] for s in states])
    th = np.concatenate([s[2] for s in states])
    v = np.concatenate([s[3] for s in states])
    sw = VFSwarm(cfgs, x, y, th, v, radius=radius, precision="f32", **kw)
    sw.step(order=order, fields=True, diag=True, verify=verify)
    return sw


def test_c3_shape_per_agent_parity(gpu):
    # one replica of the batched sweep: N=50, R=512, 707x707 periodic
    from abm_amd.vf import VFConfig
    cfg = VFConfig(n_agents=50, n_rays=512, width=707, height=707, boundary="infinite",
                   ALP0=1.5, BET0=2.0)
    st = uniform_start(50, 707, 707, 30, seed=10000, v=0.8)
    sw = run_one([cfg], [st])
    after = sw.state_numpy()
    per_agent_check(sw, 0, cfg, st, after)
    pairs, rechecked, disagree = sw.fallback_counts()
    assert pairs == 50 * 49 and disagree == 0


def test_c2_shape_per_agent_parity(gpu):
    from abm_amd.vf import VFConfig
    cfg = VFConfig(n_agents=100, n_rays=1024, width=1000, height=1000, boundary="infinite",
                   ALP0=1.0, BET0=0.5)
    st = uniform_start(100, 1000, 1000, 30, seed=42, v=0.5)
    sw = run_one([cfg], [st])
    per_agent_check(sw, 0, cfg, st, sw.state_numpy(), max_agents=40)
    assert sw.fallback_counts()[2] == 0


@pytest.mark.parametrize("nw", [1, 2, 4])
def test_walls_fov_limits_all_launch_shapes(gpu, nw):
    from abm_amd.vf import VFConfig, VFSwarm
    cfg = VFConfig.from_env_resolution(60, 512, agent_fov=0.5, width=300, height=300,
                                       boundary="walls", LIMIT_MOVEMENT=True, MAX_VEL=1.2,
                                       MAX_TH=0.05, ALP0=2.0, BET0=1.5)
    st = uniform_start(60, 300, 300, 30, seed=3, v=1.0)
    sw = VFSwarm([cfg], *st, radius=5, precision="f32")
    sw.set_launch(nw)
    sw.step(fields=True, diag=True, verify=True)
    per_agent_check(sw, 0, cfg, st, sw.state_numpy())
    assert sw.fallback_counts()[2] == 0


def test_ragged_batch_with_overrides(gpu):
    from abm_amd.vf import VFConfig
    rng = np.random.RandomState(19)
    cfgs, states = [], []
    for b in range(9):
        N = int(rng.choice([1, 2, 8, 31, 64, 65, 130]))
        R = int(rng.choice([256, 512, 1000, 2048]))
        side = float(np.sqrt(N / 1e-4)) if b % 3 else 150.0
        cfgs.append(VFConfig(n_agents=N, n_rays=R, width=side, height=side,
                             boundary="infinite" if b % 2 else "walls",
                             ALP0=float(rng.uniform(0.1, 3)), BET0=float(rng.uniform(0.1, 3))))
        states.append(uniform_start(N, side, side, 30, seed=200 + b, v=rng.uniform(0, 1.5)))
    n = sum(c.n_agents for c in cfgs)
    radius = rng.randint(3, 8, n)
    ov = {"alp0": rng.uniform(0.1, 3, n).astype(np.float32).astype(np.float64),
          "bet0": rng.uniform(0.1, 3, n).astype(np.float32).astype(np.float64),
          "v0": rng.uniform(0.5, 2, n).astype(np.float32).astype(np.float64)}
    sw = run_one(cfgs, states, radius=radius, **ov)
    after = sw.state_numpy()
    for b, cfg in enumerate(cfgs):
        sl = sw.replica_slice(b)
        st = states[b]
        aft = tuple(a[sl] for a in after)
        per_agent_check(sw, b, cfg, st, aft, radius=radius[sl],
                        overrides={k: v[sl] for k, v in ov.items()}, max_agents=24)
    assert sw.fallback_counts()[2] == 0


def test_synchronous_order_matches_oracle(gpu):
    from abm_amd.vf import VFConfig
    cfg = VFConfig(n_agents=50, n_rays=512, width=400, height=400, boundary="infinite")
    st = uniform_start(50, 400, 400, 30, seed=8, v=1.0)
    sw = run_one([cfg], [st], order="synchronous")
    ox, oy, oth, ov, obits = vo.step_arrays(*st, 5, 512, cfg.fov, 400, 400, 30, vfparams(cfg),
                                            order="synchronous")
    assert np.array_equal(sw.fields_numpy(0), obits)
    assert pose_err(sw.state_numpy(), (ox, oy, oth, ov), cfg) <= TOL


def test_verify_mode_over_a_sweep_run(gpu):
    # 256 replicas of the ALP0 x BET0 sweep for 40 steps: every float32 decision
    # equals the float64 one; re-checks stay rare
    from abm_amd.vf import VFConfig, VFSwarm
    a = np.linspace(0.1, 3, 16)
    cfgs, xs = [], []
    for i in range(16):
        for j in range(16):
            cfgs.append(VFConfig(n_agents=50, n_rays=512, width=707, height=707,
                                 boundary="infinite", ALP0=float(a[i]), BET0=float(a[j])))
            xs.append(uniform_start(50, 707, 707, 30, seed=10000 + 16 * i + j))
    cat = [np.concatenate([s[c] for s in xs]) for c in range(4)]
    sw = VFSwarm(cfgs, *cat, radius=5, precision="f32")
    for _ in range(40):
        sw.step(verify=True)
    pairs, rechecked, disagree = sw.fallback_counts()
    assert pairs == 40 * 256 * 50 * 49
    assert disagree == 0
    assert rechecked / pairs < 0.01
