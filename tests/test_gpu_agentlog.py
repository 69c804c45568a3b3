"""On-device agent records (abm_log_record) against the oracle's
restatement of ifdb.save_agent_data_RAM over the same states and fields.

Bar: the rebuilt dicts (and the JSON text) identical to the oracle's."""
import json

import numpy as np
import pytest

from oracle import log_oracle as lo

pytestmark = pytest.mark.gpu
PI = np.pi


class Rec:
    pass


def _recs(ids, x, y, th, v, fields, extra=None):
    out = []
    for c, i in enumerate(ids):
        r = Rec()
        r.id, r.position, r.orientation, r.velocity = i, np.array([x[c], y[c]]), th[c], v[c]
        r.w = r.u = r.I_priv = r.collected_r = 0
        r.mode, r.exploited_patch_id = "explore", -1
        if extra is not None:
            r.w, r.u, r.I_priv, r.collected_r, r.mode, r.exploited_patch_id = (e[c] for e in extra)
        r.soc_v_field = fields[c].astype(np.float64)
        out.append(r)
    return out


@pytest.mark.parametrize("precision", ["f32", "f64"])
def test_vf_log_matches_oracle(gpu, precision, tmp_path):
    from abm_amd.vf import VFConfig, VFSwarm
    from abm_amd.agentlog import AgentLog
    rng = np.random.RandomState(8)
    sizes, R, W = (12, 30, 5), 64, 200.0
    cfgs = [VFConfig(n_agents=n, n_rays=R, width=W, height=W, boundary="infinite", ALP0=1.5, BET0=1.0)
            for n in sizes]
    n = sum(sizes)
    sw = VFSwarm(cfgs, rng.uniform(30, 230, n), rng.uniform(30, 230, n), rng.uniform(0, 2 * PI, n),
                 0.5, radius=5, precision=precision)
    log = AgentLog(n, 7, n_rays=R)
    ref = [{} for _ in sizes]
    for step in range(7):
        sw.step(fields=True)
        log.record(sw)
        x, y, th, v = sw.state_numpy()
        for b in range(len(sizes)):
            sl = sw.replica_slice(b)
            soc = np.flip(sw.fields_numpy(b), axis=1)          # VSWRM soc_v_field = flip(raster)
            lo.save_agent_data_RAM(ref[b], _recs(range(sizes[b]), x[sl], y[sl], th[sl], v[sl], soc))
    for b in range(len(sizes)):
        sl = sw.replica_slice(b)
        got = log.agents_dict(range(sl.start, sl.stop))
        assert got == ref[b]
        path = log.save_json(str(tmp_path / f"r{b}"), range(sl.start, sl.stop))
        assert open(path).read() == json.dumps(ref[b])


def test_forage_log_matches_oracle(gpu):
    from abm_amd.forage import ForageConfig, ForageSwarm
    from abm_amd.agentlog import AgentLog
    rng = np.random.RandomState(9)
    sizes, R = (6, 9), 80
    cfgs = [ForageConfig(n_agents=s, n_rays=R, width=300.0, height=300.0) for s in sizes]
    n = sum(sizes)
    sw = ForageSwarm(cfgs, rng.uniform(40, 300, n), rng.uniform(40, 300, n), rng.uniform(0, 2 * PI, n),
                     radius=10, rngs=[np.random.RandomState(70 + b) for b in range(len(sizes))])
    log = AgentLog(n, 5, n_rays=R, foraging=True)
    ref = [{} for _ in sizes]
    for step in range(5):
        sw.step(fields=True)
        log.record(sw)
        st = {k: sw.t[k].cpu().numpy() for k in ("x", "y", "theta", "v", "w", "u", "I_priv",
                                                  "collected_r", "mode", "exploited_patch_id")}
        for b in range(len(sizes)):
            sl = sw.replica_slice(b)
            extra = [st[k][sl] for k in ("w", "u", "I_priv", "collected_r", "mode", "exploited_patch_id")]
            lo.save_agent_data_RAM(ref[b], _recs(range(sizes[b]), st["x"][sl], st["y"][sl],
                                                 st["theta"][sl], st["v"][sl], sw.fields_numpy(b), extra))
    for b in range(len(sizes)):
        sl = sw.replica_slice(b)
        assert log.agents_dict(range(sl.start, sl.stop)) == ref[b]


def test_log_errors(gpu):
    from abm_amd.vf import VFConfig, VFSwarm
    from abm_amd.agentlog import AgentLog
    from abm_amd import _lib as L
    sw = VFSwarm([VFConfig(n_agents=4, n_rays=32), VFConfig(n_agents=3, n_rays=64)],
                 np.full(7, 50.0) + np.arange(7), np.full(7, 60.0), np.zeros(7), 0.1, radius=5,
                 precision="f32")
    sw.step(fields=True)
    with pytest.raises(L.AbmError):          # mixed n_rays cannot log fields
        AgentLog(7, 2, n_rays=32).record(sw)
    log = AgentLog(7, 1)
    log.record(sw)
    with pytest.raises(L.AbmError):          # full
        log.record(sw)
    with pytest.raises(L.AbmError):          # agent count
        AgentLog(5, 1).record(sw)
