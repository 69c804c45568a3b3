"""Steps that ask for no outputs (no fields, diagnostics or counters) run
each kernel's plain instantiation, whose output paths are compiled out.  They
must move every agent exactly as the same kernel does when it writes fields:
the per-replica, push and 16-wave kernels and the chain step on a ragged batch
(reference order, vf_sims.py:302), the per-replica kernel in synchronous
order, and for a flock above 2048 agents the chain step and the wide kernel
(synchronous order).  (The grouped kernel's plain step is checked in
test_gpu_vf_group.py.)"""
import numpy as np
import pytest

from test_gpu_push import SPEC, batch

pytestmark = pytest.mark.gpu


def pair(cfgs, states, launch):
    from abm_amd.vf import VFSwarm
    cat = [np.concatenate([s[i] for s in states]) for i in range(4)]
    a = VFSwarm(cfgs, *cat, radius=5, precision="f32")
    b = VFSwarm(cfgs, *cat, radius=5, precision="f32")
    a.set_launch(launch)
    b.set_launch(launch)
    return a, b


def run(a, b, order, steps, kernels):
    for step in range(steps):
        print(f"  plain vs fields {sorted(kernels)} {order} step {step}", flush=True)
        a.step(order=order)
        b.step(order=order, fields=True)
        assert a.last_kernels() == kernels and b.last_kernels() == kernels
        for p, q in zip(a.state_numpy(), b.state_numpy()):
            assert np.array_equal(p, q), step


@pytest.mark.parametrize("launch, order, kernels", [
    (4, "sequential", {"per_replica"}),
    (1, "synchronous", {"per_replica"}),
    (32, "sequential", {"push"}),
    (16, "sequential", {"big"}),
    (64, "sequential", {"chain"}),
    (0, "sequential", {"per_replica", "push"}),
])
def test_plain_step_equals_step_with_fields(gpu, launch, order, kernels):
    cfgs, states = batch(SPEC[1:], 66)
    a, b = pair(cfgs, states, launch)
    run(a, b, order, 3, kernels)


@pytest.mark.parametrize("order, kernels", [("sequential", {"chain"}), ("synchronous", {"wide"})])
def test_plain_step_of_a_large_flock(gpu, order, kernels):
    from abm_amd.vf import VFConfig
    from test_gpu_vf_f32 import uniform_start
    cfg = VFConfig(n_agents=2500, n_rays=1024, width=2400.0, height=2400.0, boundary="infinite",
                   ALP0=1.0, BET0=0.5)
    st = uniform_start(2500, 2400.0, 2400.0, 30, seed=67, v=0.8)
    a, b = pair([cfg], [st], 0)
    run(a, b, order, 2, kernels)
