"""Ensemble statistics of replica batches, on the device.

* polarisation: |sum_i (cos th_i, sin th_i)| / N per replica, the formula of
  ExperimentLoader.calculate_polarization (abm/loader/data_loader.py:1761-1836);
* milling index: |sum_i r_i x v_i| / N about the replica's centre of mass with
  unit vectors r_i (centre -> agent) and v_i (heading), the arena unwrapped on
  the torus by the circular mean (not in the reference; defined here).

Screen coordinates: the heading vector of orientation th is (cos th, -sin th).
"""
from __future__ import annotations

import numpy as np
import torch


def _segments(swarm):
    off = torch.as_tensor(swarm.batch.agent_offset, device=swarm.x.device, dtype=torch.long)
    n = off[1:] - off[:-1]
    seg = torch.repeat_interleave(torch.arange(len(n), device=off.device), n)
    return seg, n.to(torch.float64)


def polarization(swarm):
    """Per-replica polarisation (float64 numpy array)."""
    seg, n = _segments(swarm)
    th = swarm.theta.to(torch.float64)
    sx = torch.zeros(len(n), dtype=torch.float64, device=th.device).index_add_(0, seg, torch.cos(th))
    sy = torch.zeros(len(n), dtype=torch.float64, device=th.device).index_add_(0, seg, torch.sin(th))
    return (torch.sqrt(sx * sx + sy * sy) / n).cpu().numpy()


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
