"""Pins of the ensemble-statistics oracle (oracle/stats_oracle.py):
closed forms, scipy's pdist, brute-force torus images and the nanmin of a
hand-built iid matrix."""
import math

import numpy as np
import pytest
from scipy.spatial.distance import pdist

from oracle import stats_oracle as so


def test_polarization_closed_forms():
    assert so.polarization(np.full(7, 1.3)) == pytest.approx(1.0, abs=1e-15)
    assert so.polarization(np.array([0.2, 0.2 + math.pi])) == pytest.approx(0.0, abs=1e-15)
    # four headings 90 degrees apart cancel; a NaN heading is skipped by
    # nansum but still counts in N
    assert so.polarization(np.array([0.0, math.pi / 2, math.pi, 1.5 * math.pi])) < 1e-15
    assert so.polarization(np.array([0.5, 0.5, np.nan, np.nan])) == pytest.approx(0.5, abs=1e-15)


def test_logged_positions_truncate_toward_zero():
    px, py = so.logged_positions([3.9, -2.7, 10.0], [0.5, 7.99, -0.2])
    assert list(px) == [3, -2, 10] and list(py) == [0, 7, 0]


def test_iid_open_arena_against_pdist():
    rng = np.random.RandomState(1)
    px, py = so.logged_positions(rng.uniform(0, 500, 20), rng.uniform(0, 500, 20))
    iid = so.iid_matrix(px, py)
    d = pdist(np.stack([px, py], 1))
    iu = np.triu_indices(20, k=1)
    assert np.allclose(iid[iu], d, rtol=1e-15, atol=0)
    assert np.all(iid[np.tril_indices(20)] == 0)
    assert so.mean_iid(iid) == pytest.approx(d.mean(), rel=1e-14)
    # 3-4-5 triangle
    t = so.iid_matrix(np.array([0.0, 3.0, 0.0]), np.array([0.0, 0.0, 4.0]))
    assert t[0, 1] == 3 and t[0, 2] == 4 and t[1, 2] == 5
    assert so.mean_iid(t) == 4.0


def test_iid_torus_minimum_image():
    rng = np.random.RandomState(2)
    W, H = 707, 640
    px, py = so.logged_positions(rng.uniform(0, W, 15), rng.uniform(0, H, 15))
    iid = so.iid_matrix(px, py, True, W + 0.9, H + 0.2)    # int() of the arena size
    for i in range(15):
        for j in range(i + 1, 15):
            best = min(math.hypot(px[j] - px[i] + a * W, py[j] - py[i] + b * H)
                       for a in (-1, 0, 1) for b in (-1, 0, 1))
            assert iid[i, j] == pytest.approx(best, rel=1e-14)


def test_reference_nn_quirk_and_true_nn():
    # agents on a line at 0, 10, 13, 30: iid holds d(i, j) above the diagonal;
    # row i's nanmin sees zeros at j < i, so rows 1..3 give 0 and row 0 gives
    # min(10, 13, 30) = 10 -> mean 10 / 4
    px = np.array([0.0, 10.0, 13.0, 30.0])
    iid = so.iid_matrix(px, np.zeros(4))
    assert so.reference_mean_nn(iid) == pytest.approx(2.5)
    # true nearest neighbours: 10, 3, 3, 17
    assert so.mean_nn(iid) == pytest.approx((10 + 3 + 3 + 17) / 4)
    one = so.iid_matrix(np.array([1.0]), np.array([1.0]))
    assert np.isnan(so.reference_mean_nn(one)) and np.isnan(so.mean_nn(one))
