"""GPU parity of the cooperative-signalling projections (abm_cs_projection)
against the oracle.

Bar: rows, crowding sums and signalling maxima equal to the oracle's
float64 values (every row value is 0 or a meter; a crowding value is an
integer count over len(agents); the device libm's last-ulp differences in
acos / atan move a bin only within an ulp of a tie, which these random cases
do not reach)."""
import numpy as np
import pytest
import torch

from oracle import cs_oracle as co

pytestmark = pytest.mark.gpu
PI = np.pi


class Ag:
    def __init__(self, position, orientation, radius=10, R=1200, fov=(-PI, PI), meter=0.0,
                 is_signaling=False):
        self.position = np.array(position, dtype=np.float64)
        self.orientation = float(orientation)
        self.radius = radius
        self.v_field_res = R
        self.FOV = fov
        self.meter = meter
        self.is_signaling = is_signaling


def _case(rng, R, n, fov, with_meters, max_size):
    pos = rng.uniform(0, 400, 2)
    th = rng.uniform(0, 2 * PI)
    r = float(rng.choice([1, 5, 10]))
    objs = [rng.uniform(0, 400, 2) for _ in range(n)]
    if n > 3:
        objs[1] = pos.copy()                                  # coincident: skipped
        objs[2] = pos + np.array([1.5 * r, -0.5 * r])         # very close: wide or filtered
        objs[3] = objs[0].copy()                              # duplicate object
    met = list(rng.uniform(0, 1, n)) if with_meters else None
    return pos, r, th, objs, met


@pytest.mark.parametrize("R", [8, 64, 600, 1200, 4096])
def test_rows_match_oracle(gpu, R):
    from abm_amd import cs
    rng = np.random.RandomState(R)
    for trial in range(12):
        fov = [(-PI, PI), (-0.45 * PI, 0.45 * PI), (-0.9 * PI, 0.2 * PI)][trial % 3]
        n = int(rng.choice([0, 1, 5, 40, 300]))              # 300 > one 256-object chunk
        max_size = [None, 0.5 * R, 0.02 * R][trial % 3]
        pos, r, th, objs, met = _case(rng, R, n, fov, trial % 2 == 1, max_size)
        got = cs.projection_field(fov, R, pos, r, th, objs, met, max_size)
        ref = co.projection_field(fov, R, pos, r, th, objs, met, max_size)
        assert got.shape == ref.shape == (n, R)
        assert np.array_equal(got, ref), (R, trial)


def test_reference_kat_on_gpu(gpu):
    # test_cs_supcalc.py:143-158
    from abm_amd import cs
    out = cs.projection_field((-PI, PI), 8, np.array([-1, -1]), 1, 0, [np.array([0, -1])], None)
    assert np.array_equal(out, [[0., 0., 0., 0., 1., 1., 0., 0.]])


def test_nan_meter_row(gpu):
    from abm_amd import cs
    args = ((-0.5 * PI, 0.5 * PI), 64, np.array([100.0, 100.0]), 10, 0.0, [np.array([160.0, 100.0])])
    got = cs.projection_field(*args, [float("nan")])
    ref = co.projection_field(*args, [float("nan")])
    assert np.array_equal(np.isnan(got), np.isnan(ref))
    assert np.array_equal(np.nan_to_num(got, nan=7.0), np.nan_to_num(ref, nan=7.0))


def test_batched_sum_and_max_match_oracle(gpu):
    # many focal agents in one launch, each with its own object list
    from abm_amd import cs
    rng = np.random.RandomState(11)
    R, fov = 1200, (-0.45 * PI, 0.45 * PI)
    proj = cs.CSProjection(R, fov, 0.5 * R)
    proj_nomax = cs.CSProjection(R, fov, None)
    focal, offs, objs, mets, norms, cases = [], [0], [], [], [], []
    for f in range(150):
        n = int(rng.choice([0, 1, 3, 20, 80, 270]))
        pos, r, th, o, m = _case(rng, R, n, fov, True, None)
        focal.append([pos[0], pos[1], r, th])
        objs += o
        mets += m
        offs.append(offs[-1] + n)
        norms.append(float(n + 1))
        cases.append((pos, r, th, o, m, n + 1))
    dev = torch.device("cuda")
    F = torch.tensor(focal, dtype=torch.float64, device=dev)
    O = torch.tensor(np.array(objs).reshape(-1, 2), dtype=torch.float64, device=dev)
    off = torch.tensor(offs, dtype=torch.int64, device=dev)
    M = torch.tensor(mets, dtype=torch.float64, device=dev)
    N = torch.tensor(norms, dtype=torch.float64, device=dev)
    crowd = proj.sum(F, off, O, N).cpu().numpy()
    sig = proj_nomax.max(F, off, O, M).cpu().numpy()
    rows = proj_nomax.rows(F, off, O, M).cpu().numpy()
    for f, (pos, r, th, o, m, n_agents) in enumerate(cases):
        ref_c = co.crowd_density_proj(fov, R, pos, r, th, o, n_agents, 0.5 * R)
        ref_s = co.signaling_density_proj(fov, R, pos, r, th, o, m)
        assert np.array_equal(crowd[f], ref_c), f
        assert np.array_equal(sig[f], ref_s), f
        ref_rows = co.projection_field(fov, R, pos, r, th, o, m) if o else np.zeros((0, R))
        assert np.array_equal(rows[offs[f]:offs[f + 1]], ref_rows), f


def test_agent_level_dropins(gpu):
    # CSAgent.calc_crowing_density_proj / calc_others_signaling_density_proj
    # over a few steps of memory (cs_agent.py:83-146)
    from abm_amd import cs
    rng = np.random.RandomState(5)
    agents = [Ag(rng.uniform(0, 300, 2), rng.uniform(0, 2 * PI), fov=(-0.45 * PI, 0.45 * PI))
              for _ in range(9)]
    me = agents[0]
    mem_gpu, mem_ref = cs.SignalingMemory(), co.SignalMemory()
    for step in range(8):
        for a in agents:
            a.position = a.position + rng.uniform(-3, 3, 2)
            a.is_signaling = bool(rng.rand() < 0.4)
            a.meter = float(rng.uniform(0, 1))
        crowd = cs.calc_crowing_density_proj(me, agents)
        others = [np.array(a.position) for a in agents if a is not me]
        ref_c = co.crowd_density_proj(me.FOV, me.v_field_res, me.position, me.radius, me.orientation,
                                      others, len(agents), me.v_field_res * 0.5)
        assert np.array_equal(crowd, ref_c)
        assert me.soc_v_field is crowd
        sig = cs.calc_others_signaling_density_proj(me, agents, mem_gpu, memory_depth=3)
        mem_ref.push([np.array(a.position) for a in agents if a.is_signaling and a is not me],
                     [a.meter for a in agents if a.is_signaling and a is not me])
        pos, met = mem_ref.extended(3)
        ref_s = co.signaling_density_proj(me.FOV, me.v_field_res, me.position, me.radius,
                                          me.orientation, pos, met)
        assert np.array_equal(sig, ref_s), step


def test_argument_errors(gpu):
    from abm_amd import cs, _lib as L
    dev = torch.device("cuda")
    proj = cs.CSProjection(64)
    F = torch.zeros((2, 4), dtype=torch.float64, device=dev)
    O = torch.zeros((3, 2), dtype=torch.float64, device=dev)
    with pytest.raises(L.AbmError):     # offsets past the object array
        proj.sum(F, torch.tensor([0, 2, 5], dtype=torch.int64, device=dev), O)
    with pytest.raises(L.AbmError):     # decreasing offsets
        proj.sum(F, torch.tensor([0, 3, 2], dtype=torch.int64, device=dev), O)
    with pytest.raises(L.AbmError):
        cs.CSProjection(1).rows(F, torch.tensor([0, 1, 3], dtype=torch.int64, device=dev), O)


def test_ordered_combines_match_numpy(gpu):
    # the object-order fold: a SUM with meters, and a MAX where a meter is
    # negative, -0.0 or NaN (numpy's reduce semantics, NaN propagating)
    from abm_amd import cs
    rng = np.random.RandomState(23)
    R, fov = 600, (-0.8 * PI, 0.8 * PI)
    proj = cs.CSProjection(R, fov, None)
    dev = torch.device("cuda")
    focal, offs, objs, mets, cases = [], [0], [], [], []
    for f in range(40):
        n = int(rng.choice([1, 4, 30, 260]))
        pos, r, th, o, m = _case(rng, R, n, fov, True, None)
        if f % 4 == 1:
            m[0] = -0.3
        elif f % 4 == 2:
            m[min(2, n - 1)] = -0.0
        elif f % 4 == 3:
            m[n // 2] = float("nan")
        focal.append([pos[0], pos[1], r, th])
        objs += o
        mets += m
        offs.append(offs[-1] + n)
        cases.append((pos, r, th, o, m))
    F = torch.tensor(focal, dtype=torch.float64, device=dev)
    O = torch.tensor(np.array(objs).reshape(-1, 2), dtype=torch.float64, device=dev)
    off = torch.tensor(offs, dtype=torch.int64, device=dev)
    M = torch.tensor(mets, dtype=torch.float64, device=dev)
    N = torch.tensor([7.0] * len(cases), dtype=torch.float64, device=dev)
    got_sum = proj.sum(F, off, O, N, M).cpu().numpy()
    got_max = proj.max(F, off, O, M).cpu().numpy()
    for f, (pos, r, th, o, m) in enumerate(cases):
        rows = co.projection_field(fov, R, pos, r, th, o, m)
        for got, ref in ((got_sum[f], rows.sum(axis=0) / 7.0), (got_max[f], rows.max(axis=0))):
            assert np.array_equal(np.isnan(got), np.isnan(ref)), f
            assert np.array_equal(np.nan_to_num(got, nan=3.0), np.nan_to_num(ref, nan=3.0)), f
