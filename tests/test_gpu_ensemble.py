"""Long runs: float32 against float64 by ensemble statistics.

Both kernels start 384 replicas of the same VSWRM parameters from the same
states and run 800 reference-order steps.  The chaotic trajectories
decorrelate, so individual states are not compared; the ensemble means of
polarisation and milling must agree within the 99.9% confidence interval of
their difference (Welch), and so must those of the float64 kernel started
from a second, independent set of states (a control for the test's power)."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def batch(n, seed0, precision, N=40, R=512, W=500):
    from abm_amd.vf import VFConfig, VFSwarm
    cfgs = [VFConfig(n_agents=N, n_rays=R, width=W, height=W, boundary="infinite", ALP0=1.0,
                     BET0=0.5) for _ in range(n)]
    xs, ys, ts = [], [], []
    for b in range(n):
        rng = np.random.RandomState(seed0 + b)
        xs.append(rng.uniform(30, 30 + W, N)); ys.append(rng.uniform(30, 30 + W, N))
        ts.append(rng.uniform(0, 2 * np.pi, N))
    f = lambda v: np.concatenate(v).astype(np.float32).astype(np.float64)
    return VFSwarm(cfgs, f(xs), f(ys), f(ts), 0.0, radius=5, precision=precision)


def welch_ok(a, b, z=3.29):
    se = np.sqrt(a.var(ddof=1) / len(a) + b.var(ddof=1) / len(b))
    return abs(a.mean() - b.mean()) <= z * se + 1e-12, (a.mean(), b.mean(), se)


def test_float32_and_float64_ensembles_agree(gpu):
    from abm_amd import stats
    runs = {"f32": batch(384, 777, "f32"), "f64": batch(384, 777, "f64"),
            "f64_other": batch(384, 99777, "f64")}
    for _ in range(800):
        for sw in runs.values():
            sw.step()
    pol = {k: stats.polarization(v) for k, v in runs.items()}
    mil = {k: stats.milling(v) for k, v in runs.items()}
    for k in ("f32", "f64_other"):
        ok, info = welch_ok(pol[k], pol["f64"])
        assert ok, ("polarisation", k, info)
        ok, info = welch_ok(mil[k], mil["f64"])
        assert ok, ("milling", k, info)
    spd = {k: np.array([float(v.v[v.replica_slice(b)].abs().mean()) for b in range(384)])
           for k, v in runs.items()}
    ok, info = welch_ok(spd["f32"], spd["f64"])
    assert ok, ("speed", info)
    # the trajectories did decorrelate: the comparison is between ensembles,
    # not between nearly identical states
    x32 = runs["f32"].state_numpy()[0]
    x64 = runs["f64"].state_numpy()[0]
    assert np.median(np.abs(x32 - x64)) > 1.0
