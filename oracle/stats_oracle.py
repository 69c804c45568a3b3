"""Oracle: ensemble statistics of logged trajectories, numpy float64.

TEST INFRASTRUCTURE ONLY -- the checker for ``abm_amd.stats``; never imported
by the product path.

Restates, for one time step of one replica, what abm/loader/data_loader.py
computes from the logged arrays (positions logged as ``int(position)``,
abm/monitoring/ifdb.py:83-84; orientations as floats):

* ``polarization``      calculate_polarization, data_loader.py:1761-1836
* ``iid_matrix``        calculate_interindividual_distance, :1367-1459
                        (distances only for i < j, zeros elsewhere;
                        supcalc.distance_coords / distance_torus,
                        abm/agent/supcalc.py:49-70)
* ``mean_iid``          the mean over the upper triangle (:1448-1458)
* ``reference_mean_nn`` calculate_mean_NN_dist, :1461-1486 (nanmin over the
                        second agent axis of that matrix)
* ``mean_nn``           the mean nearest-neighbour distance itself (not a
                        reference function; the quantity the name describes)

Parity pins (tests/test_oracle_stats.py): closed forms (aligned and opposed
headings, a 3-4-5 triangle), scipy's pdist for the open arena, brute-force
minimum images for the torus, and the iid matrix's zero lower triangle worked
out by hand for the nanmin.
"""
from __future__ import annotations

import warnings

import numpy as np


def logged_positions(x, y):
    """int(position) per agent, as ifdb.save_agent_data_RAM logs it."""
    return (np.array([int(v) for v in x], dtype=np.float64),
            np.array([int(v) for v in y], dtype=np.float64))


def polarization(theta):
    """|nansum (cos th, sin th)| / N with numpy's axis norm (data_loader.py:1815-1823)."""
    u = np.zeros((2, len(theta)))
    u[0] = np.array([np.cos(a) for a in theta])
    u[1] = np.array([np.sin(a) for a in theta])
    s = np.nansum(u, axis=-1)
    return np.linalg.norm(s, axis=0) / len(theta)


def iid_matrix(px, py, periodic=False, width=None, height=None):
    """iid[i, j] for j > i, zeros elsewhere (data_loader.py:1409-1433)."""
    n = len(px)
    iid = np.zeros((n, n))
    for i in range(n):
        for j in range(n):
            if j > i:
                if not periodic:
                    c1 = np.array([px[i], py[i]])
                    c2 = np.array([px[j], py[j]])
                    iid[i, j] = np.linalg.norm(c2 - c1, axis=0)
                else:
                    p1 = np.array([[px[i], py[i]]])
                    p2 = np.array([[px[j], py[j]]])
                    dim = np.array([[int(width), int(height)]])
                    delta = np.abs(p1 - p2)
                    delta = np.where(delta > 0.5 * dim, delta - dim, delta)
                    iid[i, j] = np.sqrt((delta ** 2).sum(axis=1))[0]
    return iid


def mean_iid(iid):
    n = iid.shape[0]
    iu = np.triu_indices(n, k=1)
    return np.mean(iid[iu[0], iu[1]])


def reference_mean_nn(iid):
    m = iid.copy()
    for i in range(m.shape[0]):
        m[i, i] = np.nan
    with warnings.catch_warnings():
        warnings.simplefilter("ignore", RuntimeWarning)
        nearest = np.nanmin(m, axis=-1)        # the second agent axis (data_loader.py:1482)
    return np.mean(nearest)


def mean_nn(iid):
    full = iid + iid.T
    n = full.shape[0]
    np.fill_diagonal(full, np.inf)
    return float(np.mean(full.min(axis=1))) if n > 1 else float("nan")
