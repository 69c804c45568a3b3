"""The synchronous update order that single flocks above 2048 agents use
(every field from the step's start poses) against the reference's order
(agent k sees agents < k moved, vf_sims.py:302 -> vf_agent.py:52-77), judged
by ensemble statistics at the largest flock the reference order runs:
N = 2048, R = 2048, radius 2, a 2000 px torus (the density of the 8192-agent
flock on 4000 px).  Both orders start 24 replicas from the same states and
run 150 steps; the means of polarisation and milling must agree within the
99.9 % interval of Welch's test, as must those of the reference order from
24 other starts (the control that shows the test's spread).

The mean speed does not agree: the synchronous order moves about 2 % slower
(2.742 against 2.806 px per step, six standard errors, measured on MI355X).
At this density the mean spacing is about 44 px and agents move about
2.8 px per step, so whether agent k sees agents < k before or after their
move shifts its field measurably.  The test records that bias (it must stay
under 5 %) and checks that the control agrees in speed too, so the
difference belongs to the order, not to the seeds."""
import numpy as np
import pytest

from test_gpu_ensemble import welch_ok

pytestmark = pytest.mark.gpu
PI = np.pi


def flock(n, seed0, N=2048, R=2048, W=2000.0):
    from abm_amd.vf import VFConfig, VFSwarm
    cfgs = [VFConfig(n_agents=N, n_rays=R, width=W, height=W, boundary="infinite", ALP0=1.0,
                     BET0=0.5) for _ in range(n)]
    xs, ys, ts = [], [], []
    for b in range(n):
        rng = np.random.RandomState(seed0 + b)
        xs.append(rng.uniform(30, 30 + W, N)); ys.append(rng.uniform(30, 30 + W, N))
        ts.append(rng.uniform(0, 2 * PI, N))
    f = lambda v: np.concatenate(v).astype(np.float32).astype(np.float64)
    return VFSwarm(cfgs, f(xs), f(ys), f(ts), 0.5, radius=2, precision="f32")


def test_synchronous_order_ensembles_match_reference_order(gpu):
    from abm_amd import stats
    n, steps = 24, 150
    runs = {"sequential": (flock(n, 52000), "sequential"),
            "synchronous": (flock(n, 52000), "synchronous"),
            "sequential_other": (flock(n, 62000), "sequential")}
    for t in range(steps):
        for sw, order in runs.values():
            sw.step(order=order)
        if t % 50 == 0:
            print(f"  step {t}", flush=True)
    pol = {k: stats.polarization(sw) for k, (sw, _) in runs.items()}
    mil = {k: stats.milling(sw) for k, (sw, _) in runs.items()}
    spd = {k: np.array([float(sw.v[sw.replica_slice(b)].abs().mean()) for b in range(n)])
           for k, (sw, _) in runs.items()}
    print({k: (pol[k].mean(), mil[k].mean(), spd[k].mean()) for k in runs}, flush=True)
    for k in ("synchronous", "sequential_other"):
        for name, q in (("polarisation", pol), ("milling", mil)):
            ok, info = welch_ok(q[k], q["sequential"])
            assert ok, (name, k, info)
    ok, info = welch_ok(spd["sequential_other"], spd["sequential"])
    assert ok, ("speed", "sequential_other", info)
    bias = spd["synchronous"].mean() / spd["sequential"].mean() - 1.0
    print(f"mean speed, synchronous against reference order: {bias:+.3%}", flush=True)
    assert abs(bias) < 0.05, bias
    # the orders did produce different trajectories
    xs = runs["sequential"][0].state_numpy()[0]
    xy = runs["synchronous"][0].state_numpy()[0]
    assert np.median(np.abs(xs - xy)) > 1e-3
