"""Kahan-compensated ray sums (ABM_KAHAN_SUMS, grouped kernel): fields equal
the float64 kernel's and every agent of a sample matches the oracle on the
poses it saw, with the same float32 bar as the default sums (one-step
relative pose error <= 1e-5); batches the grouped kernel does not take are
refused."""
import numpy as np
import pytest

from test_gpu_vf_f32 import per_agent_check, pose_err, uniform_start

pytestmark = pytest.mark.gpu


def test_kahan_sums_match_float64_and_oracle(gpu):
    import torch
    from abm_amd.vf import VFConfig, VFSwarm
    a = np.linspace(0.1, 3, 8)
    cfgs, sts = [], []
    for i in range(8):
        for j in range(8):
            cfgs.append(VFConfig(n_agents=50, n_rays=512, width=707, height=707, boundary="infinite",
                                 ALP0=float(a[i]), BET0=float(a[j])))
            sts.append(uniform_start(50, 707, 707, 30, seed=10000 + 8 * i + j, v=0.5))
    cat = [np.concatenate([s[c] for s in sts]) for c in range(4)]
    s64 = VFSwarm(cfgs, *cat, radius=5, precision="f64")
    s64.step(order="synchronous", fields=True)
    errs = {}
    for kahan in (False, True):
        sw = VFSwarm(cfgs, *cat, radius=5, precision="f32")
        sw.step(order="synchronous", fields=True, kahan=kahan)
        torch.cuda.synchronize()
        assert torch.equal(sw.fields, s64.fields), kahan
        errs[kahan] = pose_err(sw.state_numpy(), s64.state_numpy(), cfgs[0])
    print(f"one-step relative pose error vs float64: default {errs[False]:.3e}, Kahan {errs[True]:.3e}",
          flush=True)
    assert errs[True] <= 1e-5
    # reference order, agent by agent against the oracle (diagnostics included)
    sw = VFSwarm(cfgs, *cat, radius=5, precision="f32")
    sw.step(fields=True, diag=True, kahan=True)
    after = sw.state_numpy()
    for b in (0, 27, 63):
        sl = sw.replica_slice(b)
        per_agent_check(sw, b, cfgs[b], sts[b], tuple(x[sl] for x in after))


def test_kahan_refused_without_the_grouped_kernel(gpu):
    from abm_amd import AbmUnsupported
    from abm_amd.vf import VFConfig, VFSwarm
    cfgs = [VFConfig(n_agents=20, n_rays=r, width=300, height=300, boundary="infinite") for r in (256, 512)]
    sts = [uniform_start(20, 300, 300, 30, seed=s) for s in (1, 2)]
    cat = [np.concatenate([s[c] for s in sts]) for c in range(4)]
    sw = VFSwarm(cfgs, *cat, radius=5, precision="f32")
    with pytest.raises(AbmUnsupported, match="KAHAN"):
        sw.step(kahan=True)
