"""GPU parity of the plain float64 VSWRM path against the oracle.

Bar: visual fields bit-exact; poses within 1e-12 relative (scale 1 px for
x, y, 1 rad for theta, V0 for v) after one step from identical state."""
import numpy as np
import pytest

from oracle import vf_oracle as vo

pytestmark = pytest.mark.gpu
PI = np.pi


def rel_err(a, b, scale):
    return float(np.max(np.abs(np.asarray(a) - np.asarray(b)) / np.maximum(np.abs(b), scale)))


def uniform_start(N, W, H, pad, seed, v=0.0):
    rng = np.random.RandomState(seed)
    x = rng.uniform(pad, pad + W, N)
    y = rng.uniform(pad, pad + H, N)
    th = rng.uniform(0, 2 * PI, N)
    return x, y, th, np.full(N, v)


def vfparams(cfg):
    return vo.VFParams(GAM=cfg.GAM, V0=cfg.V0, ALP0=cfg.ALP0, ALP1=cfg.ALP1, BET0=cfg.BET0,
                       BET1=cfg.BET1, BOUNDARY=cfg.boundary, LIMIT_MOVEMENT=cfg.LIMIT_MOVEMENT,
                       MAX_VEL=cfg.MAX_VEL, MAX_TH=cfg.MAX_TH)


def check_step(cfgs, states, radius=5, order="sequential", overrides=None, tol=1e-12):
    from abm_amd.vf import VFSwarm
    x = np.concatenate([s[0] for s in states])
    y = np.concatenate([s[1] for s in states])
    th = np.concatenate([s[2] for s in states])
    v = np.concatenate([s[3] for s in states])
    ov = overrides or {}
    sw = VFSwarm(cfgs, x, y, th, v, radius=radius, precision="f64", **ov)
    sw.step(order=order, fields=True, diag=True)
    gx, gy, gth, gv = sw.state_numpy()
    diag = sw.diag_numpy()
    for b, cfg in enumerate(cfgs):
        sl = sw.replica_slice(b)
        s = states[b]
        rad = radius if np.isscalar(radius) else np.asarray(radius)[sl]
        kw = {k: np.asarray(val)[sl] for k, val in ov.items()}
        ox, oy, oth, ov_, obits, odiag = vo.step_arrays(
            s[0], s[1], s[2], s[3], rad, cfg.n_rays, cfg.fov, cfg.width, cfg.height,
            cfg.window_pad, vfparams(cfg), order=order, verbose=True, **kw)
        gbits = sw.fields_numpy(b)
        bad = np.flatnonzero((gbits != obits).any(axis=1))
        assert bad.size == 0, f"replica {b}: field mismatch for agents {bad[:10]}"
        assert rel_err(gx[sl], ox, 1.0) <= tol
        assert rel_err(gy[sl], oy, 1.0) <= tol
        assert rel_err(gth[sl], oth, 1.0) <= tol
        assert rel_err(gv[sl], ov_, cfg.V0) <= tol
        assert rel_err(diag[sl], odiag, 1e-3) <= 1e-10
    return sw


def test_projection_field_probe_matches_oracle(gpu):
    from abm_amd import vf
    rng = np.random.RandomState(11)
    for trial in range(30):
        R = int(rng.choice([8, 16, 64, 255, 512, 1024]))
        torus = trial % 2 == 1
        fov = (-PI, PI) if trial % 3 else (-0.6 * PI, 0.6 * PI)
        pos = rng.uniform(30, 230, 2)
        th = rng.uniform(0, 2 * PI)
        objs = [rng.uniform(30, 230, 2) for _ in range(25)]
        objs.append(pos.copy())          # coincident object is skipped
        sizes = [int(s) for s in rng.randint(1, 12, len(objs))]
        kw = dict(boundary_cond="infinite" if torus else "walls", arena_width=200,
                  arena_height=200)
        got = vf.projection_field(fov, R, pos, 5, th, objs, sizes, **kw)
        exp = vo.projection_field(fov, R, pos, 5, th, [o.copy() for o in objs], sizes, **kw)
        assert np.array_equal(got, exp), (trial, R)
    # the reference's known-answer input
    got = vf.projection_field((-PI, PI), 8, np.array([-1, -1]), 1, 0, [np.array([0, -1])], None)
    assert np.all(got == [[0., 0., 0., 0., 1., 1., 0., 0.]])


def test_state_variables_probe_matches_oracle(gpu):
    from abm_amd import vf
    rng = np.random.RandomState(5)
    p = vo.VFParams(GAM=0.1, V0=1.0, ALP0=1.2, ALP1=0.09, BET0=0.6, BET1=0.09)
    for R in (8, 64, 512, 1024, 1200):
        Phi = np.arange(-PI, PI, 2 * PI / R)
        for density in (0.0, 0.05, 0.5, 1.0):
            V = (rng.rand(R) < density).astype(np.float64)
            vel = rng.uniform(-1, 2)
            exp = vo.vswrm_state_variables(vel, Phi, V, p, verbose=True)
            got = vf.VSWRM_flocking_state_variables(vel, Phi, V, p, verbose=True)
            assert np.allclose(got, exp, rtol=1e-12, atol=1e-14), (R, density)
            assert np.array_equal(vf.dPhi_V_of(Phi, V), vo.dphi_v_of(Phi, V))


def test_step_c2_shape_torus(gpu):
    # N=100, R=1024, 1000x1000 periodic, GAM .1 V0 1 ALP0 1 ALP1 .09 BET0 .5 BET1 .09
    from abm_amd.vf import VFConfig
    cfg = VFConfig(n_agents=100, n_rays=1024, width=1000, height=1000, boundary="infinite",
                   ALP0=1.0, BET0=0.5)
    st = uniform_start(100, 1000, 1000, 30, seed=42, v=0.3)
    check_step([cfg], [st])


def test_step_walls_limited_fov_limit_movement(gpu):
    from abm_amd.vf import VFConfig
    cfg = VFConfig.from_env_resolution(40, 256, agent_fov=0.5, width=300, height=300,
                                       boundary="walls", LIMIT_MOVEMENT=True, MAX_VEL=1.2,
                                       MAX_TH=0.05, ALP0=2.0, BET0=1.5)
    assert cfg.n_rays == 512
    st = uniform_start(40, 300, 300, 30, seed=3, v=1.0)
    check_step([cfg], [st])


def test_step_ragged_batch_with_overrides(gpu):
    from abm_amd.vf import VFConfig
    rng = np.random.RandomState(9)
    cfgs, states = [], []
    for b in range(7):
        N = int(rng.choice([1, 2, 9, 33, 64, 70]))
        R = int(rng.choice([64, 256, 512, 1000]))
        side = float(rng.choice([120, 400]))
        cfgs.append(VFConfig(n_agents=N, n_rays=R, width=side, height=side,
                             boundary="infinite" if b % 2 else "walls",
                             ALP0=float(rng.uniform(0.1, 3)), BET0=float(rng.uniform(0.1, 3))))
        states.append(uniform_start(N, side, side, 30, seed=100 + b, v=rng.uniform(0, 1.5)))
    n = sum(c.n_agents for c in cfgs)
    radius = rng.randint(3, 8, n)
    overrides = {"alp0": rng.uniform(0.1, 3, n), "bet0": rng.uniform(0.1, 3, n),
                 "v0": rng.uniform(0.5, 2, n)}
    check_step(cfgs, states, radius=radius, overrides=overrides)


def test_step_synchronous_order(gpu):
    from abm_amd.vf import VFConfig
    cfg = VFConfig(n_agents=50, n_rays=512, width=400, height=400, boundary="infinite")
    st = uniform_start(50, 400, 400, 30, seed=8, v=1.0)
    check_step([cfg], [st], order="synchronous")


def test_step_phi_length_mismatch_moves_without_increments(gpu):
    # R = 3000: len(arange) = 3001 != R -> dv = dpsi = 0 (vf_agent.py:282-284)
    from abm_amd.vf import VFConfig
    cfg = VFConfig(n_agents=6, n_rays=3000, width=200, height=200, boundary="walls")
    st = uniform_start(6, 200, 200, 30, seed=2, v=0.7)
    sw = check_step([cfg], [st])
    assert np.allclose(sw.state_numpy()[3], 0.7)


def test_multi_step_trajectory(gpu):
    from abm_amd.vf import VFConfig, VFSwarm
    cfg = VFConfig(n_agents=30, n_rays=512, width=300, height=300, boundary="infinite",
                   ALP0=1.0, BET0=0.5)
    x, y, th, v = uniform_start(30, 300, 300, 30, seed=21, v=0.5)
    sw = VFSwarm([cfg], x, y, th, v, radius=5, precision="f64")
    p = vfparams(cfg)
    for _ in range(10):
        sw.step()
        x, y, th, v, _ = vo.step_arrays(x, y, th, v, 5, cfg.n_rays, cfg.fov, cfg.width,
                                        cfg.height, cfg.window_pad, p)
    gx, gy, gth, gv = sw.state_numpy()
    assert rel_err(gx, x, 1.0) < 1e-9 and rel_err(gy, y, 1.0) < 1e-9
    assert rel_err(gth, th, 1.0) < 1e-9 and rel_err(gv, v, 1.0) < 1e-9
