"""Per-step ensemble statistics of batched runs on the GPU.

The reference summarises runs offline from the logged trajectories
(abm/loader/data_loader.py): polarisation (calculate_polarization,
:1761-1836), the mean inter-individual distance
(calculate_interindividual_distance, :1367-1459) and the mean nearest-
neighbour distance (calculate_mean_NN_dist, :1461-1486), with positions logged
as ``int(position)`` (abm/monitoring/ifdb.py:83-84).  ``ensemble_stats``
computes the same quantities for every replica of a swarm from the state in
HBM through ``abm_ensemble_stats``, so a long batched run can be tracked step
by step without copying trajectories to the host.

Columns of the result (one row per replica):

* ``POLARIZATION``  |sum_i (cos th_i, sin th_i)| / N
* ``MEAN_IID``      mean distance over pairs i < j
* ``MEAN_NN``       mean over agents of the nearest-neighbour distance
* ``REFERENCE_NN``  the value calculate_mean_NN_dist gives for the step:
  its nanmin runs over an iid matrix filled only above the diagonal (zeros
  below), so it equals min_j>0 d(0, j) / N

``milling`` (not in the reference; defined here) is the milling index
|sum_i r_i x v_i| / N about the replica's centre of mass with unit vectors
r_i (centre -> agent) and v_i (heading), the arena unwrapped on the torus by
the circular mean.  Screen coordinates: the heading vector of orientation th
is (cos th, -sin th).
"""
from __future__ import annotations

import ctypes as C

import numpy as np
import torch

from . import _lib as L

POLARIZATION, MEAN_IID, MEAN_NN, REFERENCE_NN = range(4)

def ensemble_stats_arrays(agent_offset, x, y, theta, periodic=False, width=0.0, height=0.0,
                          stream=None):
    """Statistics of replicas given as device tensors: ``agent_offset``
    (int32, n_replicas + 1), ``x``, ``y``, ``theta`` (float32 or float64, all
    of one dtype).  Returns a float64 (n_replicas, 4) device tensor."""
    lib = L.load()
    n = int(agent_offset.shape[0]) - 1
    if n < 0 or agent_offset.dtype != torch.int32:
        raise L.AbmError(L.ABM_ERR_INVALID, "agent_offset must be int32 (n_replicas + 1,)")
    if not (x.dtype == y.dtype == theta.dtype) or x.dtype not in (torch.float32, torch.float64):
        raise L.AbmError(L.ABM_ERR_INVALID, "x, y, theta must share one float dtype")
    n_agents = int(x.shape[0])
    if y.shape[0] != n_agents or theta.shape[0] != n_agents:
        raise L.AbmError(L.ABM_ERR_INVALID, "x, y, theta must have one length")
    # the offsets bound every state read: check them on the host
    off = agent_offset.cpu().numpy()
    if n >= 0 and (off[0] != 0 or off[-1] != n_agents or np.any(np.diff(off) < 0)):
        raise L.AbmError(L.ABM_ERR_INVALID, "agent_offset must rise from 0 to the agent count")
    out = torch.empty((max(n, 0), 4), dtype=torch.float64, device=x.device)
    if n == 0:
        return out
    L.check(lib.abm_ensemble_stats(n, L.ptr(agent_offset), L.ptr(x), L.ptr(y), L.ptr(theta),
                                   1 if x.dtype == torch.float32 else 0, 1 if periodic else 0,
                                   float(width), float(height), L.ptr(out), L.stream_ptr(stream)))
    return out


def ensemble_stats(swarm, periodic=None, stream=None):
    """Statistics of every replica of a ``VFSwarm`` or ``ForageSwarm`` (one
    row per replica, float64 device tensor).  ``periodic`` defaults to the
    swarm's boundary (torus distances on int(width) x int(height) need one
    arena size across the batch)."""
    if hasattr(swarm, "batch"):                  # VFSwarm
        x, y, th, ao = swarm.x, swarm.y, swarm.theta, swarm.batch.agent_offset
    else:                                        # ForageSwarm
        x, y, th, ao = swarm.t["x"], swarm.t["y"], swarm.t["theta"], swarm.agent_offset
    offs = getattr(swarm, "_stats_offsets", None)
    if offs is None:
        offs = torch.as_tensor(np.asarray(ao, dtype=np.int32), device=x.device)
        swarm._stats_offsets = offs
    cfgs = getattr(getattr(swarm, "batch", swarm), "configs", None) or []
    if periodic is None:
        periodic = bool(cfgs) and all(getattr(c, "boundary", "walls") == "infinite" for c in cfgs)
    W = H = 0.0
    if periodic:
        sizes = {(float(c.width), float(c.height)) for c in cfgs}
        if len(sizes) != 1:
            raise L.AbmError(L.ABM_ERR_INVALID, "torus statistics need one arena size per batch")
        W, H = sizes.pop()
    return ensemble_stats_arrays(offs, x, y, th, periodic, W, H, stream)


def polarization(swarm):
    """Per-replica polarisation (float64 numpy array), from abm_ensemble_stats."""
    return ensemble_stats(swarm)[:, POLARIZATION].cpu().numpy()

def _segments(swarm):
    off = torch.as_tensor(swarm.batch.agent_offset, device=swarm.x.device, dtype=torch.long)
    n = off[1:] - off[:-1]
    seg = torch.repeat_interleave(torch.arange(len(n), device=off.device), n)
    return seg, n.to(torch.float64)


def milling(swarm):
    """Per-replica milling index (float64 numpy array)."""
    seg, n = _segments(swarm)
    dev = swarm.x.device
    cfg = swarm.batch.configs
    W = torch.as_tensor([c.width for c in cfg], dtype=torch.float64, device=dev)[seg]
    H = torch.as_tensor([c.height for c in cfg], dtype=torch.float64, device=dev)[seg]
    torus = torch.as_tensor([c.boundary == "infinite" for c in cfg], device=dev)[seg]
    x = swarm.x.to(torch.float64)
    y = swarm.y.to(torch.float64)

    def centre(p, L):
        # circular mean on a torus, plain mean in a box
        a = 2 * np.pi * p / L
        cs = torch.zeros(len(n), dtype=torch.float64, device=dev).index_add_(0, seg, torch.cos(a))
        sn = torch.zeros(len(n), dtype=torch.float64, device=dev).index_add_(0, seg, torch.sin(a))
        circ = torch.atan2(sn, cs)[seg] * L / (2 * np.pi)
        mean = (torch.zeros(len(n), dtype=torch.float64, device=dev).index_add_(0, seg, p) / n)[seg]
        return torch.where(torus, circ, mean)

    def delta(p, c, L):
        d = p - c
        return torch.where(torus, torch.remainder(d + L / 2, L) - L / 2, d)

    dx, dy = delta(x, centre(x, W), W), delta(y, centre(y, H), H)
    r = torch.sqrt(dx * dx + dy * dy).clamp_min(1e-12)
    th = swarm.theta.to(torch.float64)
    vx, vy = torch.cos(th), -torch.sin(th)
    cross = (dx / r) * vy - (dy / r) * vx
    s = torch.zeros(len(n), dtype=torch.float64, device=dev).index_add_(0, seg, cross)
    return (torch.abs(s) / n).cpu().numpy()
