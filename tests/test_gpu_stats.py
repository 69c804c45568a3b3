"""GPU ensemble statistics (abm_ensemble_stats) against the oracle's
restatement of data_loader.py on the logged (int-truncated) positions.

Bar: nearest-neighbour values exact (minima of identical distances, one
division); polarisation and mean distances within 1e-12 relative (their sums
run in a different order)."""
import numpy as np
import pytest
import torch

from oracle import stats_oracle as so

pytestmark = pytest.mark.gpu
PI = np.pi


def _swarm(precision, boundary, seed, sizes=(50, 1, 2, 17, 130)):
    from abm_amd.vf import VFConfig, VFSwarm
    rng = np.random.RandomState(seed)
    W = 707.3
    cfgs = [VFConfig(n_agents=n, n_rays=256, width=W, height=W, boundary=boundary) for n in sizes]
    n = sum(sizes)
    x, y = rng.uniform(30, 30 + W, n), rng.uniform(30, 30 + W, n)
    th = rng.uniform(0, 2 * PI, n)
    if precision == "f32":
        x, y, th = (a.astype(np.float32).astype(np.float64) for a in (x, y, th))
    return VFSwarm(cfgs, x, y, th, 0.0, radius=5, precision=precision), W


@pytest.mark.parametrize("precision", ["f32", "f64"])
@pytest.mark.parametrize("boundary", ["walls", "infinite"])
def test_stats_match_oracle(gpu, precision, boundary):
    from abm_amd import stats
    sw, W = _swarm(precision, boundary, 3)
    for _ in range(3):
        sw.step()
    got = stats.ensemble_stats(sw).cpu().numpy()
    x, y, th = (t.cpu().numpy().astype(np.float64) for t in (sw.x, sw.y, sw.theta))
    for b in range(len(sw.batch.configs)):
        sl = sw.replica_slice(b)
        px, py = so.logged_positions(x[sl], y[sl])
        iid = so.iid_matrix(px, py, boundary == "infinite", W, W)
        n = len(px)
        assert got[b, stats.POLARIZATION] == pytest.approx(so.polarization(th[sl]), rel=1e-12, abs=1e-15)
        if n == 1:
            assert np.isnan(got[b, 1:]).all()
            continue
        assert got[b, stats.MEAN_IID] == pytest.approx(so.mean_iid(iid), rel=1e-12)
        assert got[b, stats.MEAN_NN] == pytest.approx(so.mean_nn(iid), rel=1e-13)
        assert got[b, stats.REFERENCE_NN] == so.reference_mean_nn(iid)


def test_polarization_helper_and_errors(gpu):
    from abm_amd import stats, _lib as L
    sw, _ = _swarm("f32", "infinite", 4, sizes=(10, 10))
    sw.theta.fill_(0.7)
    assert np.allclose(stats.polarization(sw), 1.0, atol=1e-7)
    dev = torch.device("cuda")
    off = torch.tensor([0, 5, 4], dtype=torch.int32, device=dev)
    z = torch.zeros(4, dtype=torch.float64, device=dev)
    with pytest.raises(L.AbmError):
        stats.ensemble_stats_arrays(off, z, z, z)
    with pytest.raises(L.AbmError):
        stats.ensemble_stats_arrays(torch.tensor([0, 4], dtype=torch.int32, device=dev), z, z,
                                    z.float())
