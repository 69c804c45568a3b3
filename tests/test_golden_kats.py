"""The reference's known-answer vectors (tests/golden/reference_kats.json,
extracted from test_cs_supcalc.py by tests/golden/make_reference_kats.py)
against the oracles (CPU) and the GPU probes."""
import json
import os

import numpy as np
import pytest

from oracle import cs_oracle as co
from oracle import vf_oracle as vo

KATS = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "reference_kats.json")))


def val(v):
    if isinstance(v, dict) and "pi" in v:
        return v["pi"] * np.pi
    if isinstance(v, list):
        return [val(e) for e in v]
    return v


def _pf_case(c):
    return (tuple(val(c["fov"])), c["v_field_resolution"], np.array(c["position"], float), c["radius"],
            c["orientation"], [np.array(p, float) for p in c["object_positions"]])


@pytest.mark.parametrize("case", KATS["projection_field"])
def test_projection_field_oracles(case):
    args = _pf_case(case)
    assert np.array_equal(co.projection_field(*args, case["object_meters"]), case["true_projection"])
    # vf_supcalc.projection_field is the same algorithm in walls mode, full FOV
    assert np.array_equal(vo.projection_field(*args), case["true_projection"])


@pytest.mark.parametrize("case", KATS["closed_angle"])
def test_closed_angle_oracles(case):
    v1, v2 = np.array(case["v1"], float), np.array(case["v2"], float)
    assert np.isclose(co.calculate_closed_angle(v1, v2), val(case["expected_output"]))
    assert np.isclose(vo.closed_angle(v1, v2), val(case["expected_output"]))


def test_random_walk_stream():
    np.random.seed(42)
    for c in KATS["random_walk_seed42"]:
        assert np.random.uniform(c["lo"], c["hi"]) == c["dtheta"]


@pytest.mark.gpu
@pytest.mark.parametrize("case", KATS["projection_field"])
def test_projection_field_on_gpu(gpu, case):
    from abm_amd import cs, vf
    args = _pf_case(case)
    assert np.array_equal(cs.projection_field(*args, case["object_meters"]), case["true_projection"])
    assert np.array_equal(vf.projection_field(*args), case["true_projection"])


def test_random_walk_stream_native():
    # the library's MT19937 peek (host code, no GPU) gives the same draws from
    # numpy's legacy generator state, as lo + (hi - lo) * u
    import ctypes as C
    from abm_amd import _lib as L
    lib = L.load()
    rng = np.random.RandomState(42)
    n = len(KATS["random_walk_seed42"])
    addr = np.array([rng._bit_generator.ctypes.state_address], dtype=np.uint64)
    cnt = np.array([n], dtype=np.int64)
    out = np.empty(n)
    L.check(lib.abm_mt19937_peek(addr.ctypes.data_as(C.c_void_p), 1, cnt.ctypes.data_as(C.c_void_p),
                                 out.ctypes.data_as(C.c_void_p)))
    for c, u in zip(KATS["random_walk_seed42"], out):
        assert c["lo"] + (c["hi"] - c["lo"]) * u == c["dtheta"]
