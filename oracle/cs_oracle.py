"""Oracle: cooperative-signalling visual projections, numpy float64.

TEST INFRASTRUCTURE ONLY -- the checker for ``abm_amd.cs``; never imported by
the product path.

A CPU restatement of the cooperative-signalling agent's perception
(scioip34/ABM, abm/projects/cooperative_signaling/cs_agent/):

* ``projection_field``          cs_supcalc.py:204-295
* ``validate_projection_size``  cs_supcalc.py:298-307
* ``calculate_closed_angle``    cs_supcalc.py:310-326 (+ abm/agent/supcalc.py:19-34)
* ``crowd_density_proj``        cs_agent.py:83-104 (sum over rows / len(agents))
* ``signaling_density_proj``    cs_agent.py:106-131 (max over rows)
* ``SignalMemory``              cs_agent.py:106-116, 133-146 (memory of
                                signalling positions with linear decay)

The same numpy calls are made on the same float64 operands as in the
reference, so on a given host this oracle produces the reference's bits.

Parity pins (tests/test_oracle_cs.py): the reference's own known-answer tests
(cs_agent/tests/test_cs_supcalc.py:143-169: the R=8 field and the closed
angles of three vector pairs), agreement with the VSWRM oracle's rows
(already pinned by exact-rational geometry) wherever the two algorithms
coincide, hand-derived FOV-mask, size-filter, meter and combine cases, and
the memory bookkeeping traced by hand over six steps.
"""
from __future__ import annotations

from typing import List, Optional, Sequence, Tuple

import numpy as np

from .vf_oracle import angle_between, find_nearest


def calculate_closed_angle(v1, v2):
    """Closed angle in [-pi, pi]: 0 <= a' <= pi maps to -a', else 2 pi - a'
    (cs_supcalc.py:310-326)."""
    a = angle_between(v1, v2) % (2 * np.pi)
    if 0 <= a <= np.pi:
        return -a
    return 2 * np.pi - a


def validate_projection_size(max_proj_size, proj_size):
    """None accepts every size; otherwise proj_size <= max (cs_supcalc.py:298-307)."""
    return max_proj_size is None or proj_size <= max_proj_size


def projection_field(fov, v_field_resolution, position, radius, orientation,
                     object_positions, object_meters=None, max_proj_size=None):
    """Per-object rows, flipped and FOV-masked (cs_supcalc.py:204-295)."""
    R = v_field_resolution
    rows = np.zeros((len(object_positions), R))
    phis = np.linspace(-np.pi, np.pi, R)
    position = np.asarray(position, dtype=np.float64)
    centre = position + radius
    edge = position + np.array([1 + np.cos(orientation), 1 - np.sin(orientation)]) * radius
    v1 = edge - centre
    for i, p in enumerate(object_positions):
        p = np.asarray(p, dtype=np.float64)
        if p[0] == position[0] and p[1] == position[1]:                # :240
            continue
        oc = p + radius                                                # :244 focal radius
        v2 = oc - centre
        ca = calculate_closed_angle(v1, v2)
        dist = np.linalg.norm(oc - centre)
        vis = 2 * np.arctan(radius / (1 * dist))
        t = find_nearest(phis, ca)
        if not (fov[0] <= ca <= fov[1]):                               # :262
            continue
        size = (vis / (2 * np.pi)) * R
        if not validate_projection_size(max_proj_size, size):          # :270
            continue
        s = int(t - np.floor(size / 2))
        e = int(t + np.floor(size / 2))
        if s < 0:                                                      # :277-283
            rows[i, R + s:R] = 1
            s = 0
        if e >= R:
            rows[i, 0:e - (R - 1)] = 1
            e = R
        rows[i, s:e] = 1
        if object_meters is not None:                                  # :286-287
            rows[i] *= object_meters[i]
    post = np.flip(rows, axis=1)                                       # :291
    post[:, phis < fov[0]] = 0
    post[:, phis > fov[1]] = 0
    return post


def crowd_density_proj(fov, v_field_resolution, position, radius, orientation,
                       other_positions, n_agents, max_proj_size):
    """calc_crowing_density_proj: rows of all other agents (no meters) summed
    over objects and divided by len(agents), self included (cs_agent.py:83-104)."""
    field = projection_field(fov, v_field_resolution, position, radius, orientation,
                             other_positions, None, max_proj_size)
    return field.sum(axis=0) / n_agents


def signaling_density_proj(fov, v_field_resolution, position, radius, orientation,
                           positions, meters):
    """The combine of calc_others_signaling_density_proj: zeros when nobody
    signalled, else the maximum over meter-scaled rows (cs_agent.py:117-131)."""
    if len(positions) == 0:
        return np.zeros(v_field_resolution)
    field = projection_field(fov, v_field_resolution, position, radius, orientation,
                             positions, meters)
    return field.max(axis=0)


class SignalMemory:
    """One agent's memory of signalling agents (cs_agent.py:106-116, 133-146).

    ``push`` inserts the current (positions, meters) at the front;
    ``extended`` returns the positions and decayed meters of entries
    n = 0, 1, ... with decay 1 - n / depth, dropping the oldest entry and
    stopping once n reaches depth (the reference's loop pops inside the loop).
    """

    def __init__(self):
        self.entries: List[Tuple[list, list]] = []

    def push(self, positions, meters):
        self.entries.insert(0, (list(positions), list(meters)))

    def extended(self, depth) -> Tuple[list, list]:
        pos, met = [], []
        for n, (p, m) in enumerate(self.entries):
            decay = 1 - n / depth
            if p:
                pos.extend(p)
                met.extend([x * decay for x in m])
            if n >= depth:
                self.entries.pop(-1)
                break
        return pos, met
