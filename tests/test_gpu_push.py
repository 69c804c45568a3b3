"""The push kernel: the reference order (vf_sims.py:302 -> vf_agent.py:52-77)
with each agent's pairs decided off the id-order chain, as soon as their
object has moved, into per-agent fields in LDS.

* It makes the one-wave per-replica kernel's decisions, rasters, sums and
  moves, so fields, poses, speeds and diagnostics equal that kernel's bit for
  bit, step after step (torus and walls, limited FOV and movement limits,
  R = 256..2048).
* Every pair is decided as float64 decides it (verify mode), with the pair
  count of every non-coincident pair.
* Sampled agents match the oracle on the poses they saw.
* The automatic launch sends the large size class to it, and replicas whose
  fields do not fit in LDS to the 16-wave kernel."""
import numpy as np
import pytest

from test_gpu_vf_f32 import per_agent_check, uniform_start

pytestmark = pytest.mark.gpu


def batch(spec, seed):
    from abm_amd.vf import VFConfig
    cfgs, states = [], []
    for b, (n, r, boundary, fovr, limit) in enumerate(spec):
        side = float(100.0 * np.sqrt(n))
        cfgs.append(VFConfig(n_agents=n, n_rays=r, width=side, height=side, boundary=boundary,
                             fov=VFConfig.fov_from_ratio(fovr), LIMIT_MOVEMENT=limit, ALP0=1.0, BET0=0.5))
        states.append(uniform_start(n, side, side, 30, seed=seed * 100 + b, v=0.7))
    return cfgs, states


SPEC = [(600, 1024, "infinite", 1.0, False), (512, 1024, "infinite", 1.0, False),
        (300, 512, "walls", 0.6, True), (200, 2048, "infinite", 1.0, False),
        (129, 256, "walls", 1.0, False), (64, 512, "infinite", 1.0, False), (8, 256, "walls", 1.0, False)]


def swarm(cfgs, states, launch):
    from abm_amd.vf import VFSwarm
    cat = [np.concatenate([s[i] for s in states]) for i in range(4)]
    sw = VFSwarm(cfgs, *cat, radius=5, precision="f32")
    sw.set_launch(launch)
    return sw


@pytest.mark.parametrize("diag", [False, True])
def test_push_kernel_equals_the_one_wave_kernel(gpu, diag):
    import torch
    cfgs, states = batch(SPEC, 61)
    a, b = swarm(cfgs, states, 32), swarm(cfgs, states, 1)
    for step in range(4):
        print(f"  push vs one-wave step {step} diag={diag}", flush=True)
        a.step(fields=True, diag=diag)
        b.step(fields=True, diag=diag)
        assert a.last_kernels() == {"push"} and b.last_kernels() == {"per_replica"}
        torch.cuda.synchronize()
        assert torch.equal(a.fields, b.fields), step
        assert torch.equal(a.diag, b.diag), step
        for p, q in zip(a.state_numpy(), b.state_numpy()):
            assert np.array_equal(p, q), step


def test_push_kernel_decides_every_pair_as_float64(gpu):
    cfgs, states = batch(SPEC, 62)
    sw = swarm(cfgs, states, 32)
    for step in range(3):
        sw.step(verify=True)
    pairs, _, disagree = sw.fallback_counts()
    assert disagree == 0 and pairs == 3 * sum(c.n_agents * (c.n_agents - 1) for c in cfgs)


def test_push_kernel_agents_match_oracle(gpu):
    cfgs, states = batch(SPEC, 63)
    sw = swarm(cfgs, states, 0)
    sw.step(fields=True, diag=True)
    assert sw.last_kernels() == {"per_replica", "push"}
    after = sw.state_numpy()
    for b, cfg in enumerate(cfgs):
        if cfg.n_agents <= 64:
            continue
        n = cfg.n_agents
        sl = sw.replica_slice(b)
        agents = sorted({0, 1, n // 2, n - 2, n - 1} |
                        set(np.random.RandomState(b).choice(n, 6, replace=False).tolist()))
        per_agent_check(sw, b, cfg, states[b], tuple(x[sl] for x in after), agents=agents)


def test_large_class_beyond_lds_takes_the_16_wave_kernel(gpu):
    # 1500 agents at R = 2048: the fields need about 0.5 MB, so the large
    # class runs on the 16-wave kernel; the push launch shape is refused
    from abm_amd import AbmUnsupported
    cfgs, states = batch([(1500, 2048, "infinite", 1.0, False), (40, 256, "walls", 1.0, False)], 64)
    sw = swarm(cfgs, states, 0)
    sw.step()
    assert sw.last_kernels() == {"per_replica", "big"}
    sw.set_launch(32)
    with pytest.raises(AbmUnsupported, match="push kernel"):
        sw.step()


def test_synchronous_order_stays_off_the_reference_order_kernels(gpu):
    # the push and 16-wave kernels and the chain step move agents in id
    # order only: a synchronous step of a batch with a large size class runs
    # on the per-replica kernel, with the float64 kernel's synchronous fields
    # bit for bit; a forced reference-order shape refuses it
    import torch
    from abm_amd import AbmUnsupported
    from abm_amd.vf import VFSwarm
    cfgs, states = batch(SPEC[1:], 65)
    cat = [np.concatenate([s[i] for s in states]) for i in range(4)]
    a = VFSwarm(cfgs, *cat, radius=5, precision="f32")
    b = VFSwarm(cfgs, *[x.astype(np.float32).astype(np.float64) for x in cat], radius=5, precision="f64")
    a.step(order="synchronous", fields=True)
    b.step(order="synchronous", fields=True)
    assert a.last_kernels() == {"per_replica"}
    torch.cuda.synchronize()
    assert torch.equal(a.fields, b.fields)
    for shape in (16, 32, 64):
        a.set_launch(shape)
        with pytest.raises(AbmUnsupported, match="reference order only"):
            a.step(order="synchronous")


def test_size_class_rule_follows_the_cus(gpu):
    # the automatic launch sends the large class (replicas above 64 agents)
    # to the push kernel, one workgroup per CU, only while the class's agents
    # spread over the CUs do not outlast its largest replica's chain; a class
    # of many similar replicas stays on the per-replica kernel
    import torch
    from abm_amd.vf import VFConfig, VFSwarm
    n_cu = torch.cuda.get_device_properties(0).multi_processor_count
    for n_rep, kernels in ((max(1, n_cu // 2), {"push"}), (2 * n_cu, {"per_replica"})):
        cfg = VFConfig(n_agents=80, n_rays=256, width=900.0, height=900.0, boundary="infinite",
                       ALP0=1.0, BET0=0.5)
        st = uniform_start(80 * n_rep, 900.0, 900.0, 30, seed=68, v=0.5)
        sw = VFSwarm([cfg] * n_rep, *st, radius=5, precision="f32")
        sw.step()
        assert sw.last_kernels() == kernels, (n_rep, sw.last_kernels())
