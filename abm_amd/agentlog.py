"""Agent records of a batched run, kept on the GPU.

``AgentLog`` holds, for every agent and step, what the reference's
``ifdb.save_agent_data_RAM`` appends to its dict of lists
(abm/monitoring/ifdb.py:55-96): ``int(position)``, orientation, velocity, the
foraging state and the edges of the agent's ``soc_v_field``
(``vfield_up`` / ``vfield_down``, np.where(np.roll(f, 1) < f)[0] and > f).
``record`` appends one step with one launch of ``abm_log_record``;
``agents_dict`` rebuilds the reference's dict for one replica (the edge lists
as the same numpy-formatted strings) and ``save_json`` writes it as
``save_ifdb_as_csv(use_ram=True, as_zar=False)`` does (ifdb.py:413-434).
"""
from __future__ import annotations

import ctypes as C
import json
import os
from typing import Optional, Sequence

import numpy as np
import torch

from . import _lib as L


def pad_to_n_digits(number, n=3):
    """ifdb.pad_to_n_digits (ifdb.py:38-52)."""
    s = str(number)
    return (n - len(s)) * "0" + s if n > len(s) else s


class AgentLog:
    """Per-step agent records of ``n_agents`` agents for up to ``steps`` steps."""

    def __init__(self, n_agents: int, steps: int, n_rays: Optional[int] = None,
                 foraging: bool = False, device=None):
        self.lib = L.load()
        self.n, self.T = int(n_agents), int(steps)
        self.R = None if n_rays is None else int(n_rays)
        self.foraging = bool(foraging)
        dev = torch.device(device) if device is not None else torch.device("cuda")
        self.device = dev
        shape = (self.T, self.n)
        i32 = lambda: torch.zeros(shape, dtype=torch.int32, device=dev)
        f64 = lambda: torch.zeros(shape, dtype=torch.float64, device=dev)
        self.posx, self.posy, self.orientation, self.velocity = i32(), i32(), f64(), f64()
        self.w = self.u = self.I_priv = self.collected_r = self.mode = self.expl_patch_id = None
        if self.foraging:
            self.w, self.u, self.I_priv, self.collected_r = f64(), f64(), f64(), f64()
            self.mode, self.expl_patch_id = i32(), i32()
        self.up = self.down = None
        if self.R is not None:
            words = (self.R + 31) // 32
            self.up = torch.zeros((self.T, self.n, words), dtype=torch.int32, device=dev)
            self.down = torch.zeros((self.T, self.n, words), dtype=torch.int32, device=dev)
        p = lambda t: None if t is None else t.data_ptr()
        self._buf = L.LogBuffers(p(self.posx), p(self.posy), p(self.orientation), p(self.velocity),
                                 p(self.w), p(self.u), p(self.I_priv), p(self.collected_r),
                                 p(self.mode), p(self.expl_patch_id), p(self.up), p(self.down))
        self.t = 0

    def record(self, swarm, fields: bool = True, stream=None):
        """Append the swarm's current state (and, with ``fields``, the fields
        of its last step, which needs ``step(fields=True)``) as the next step."""
        if self.t >= self.T:
            raise L.AbmError(L.ABM_ERR_INVALID, f"the log holds {self.T} steps")
        forage = hasattr(swarm, "t")
        if forage:
            st = swarm.t
            x, y, th, v = st["x"], st["y"], st["theta"], st["v"]
            extra = [st["w"], st["u"], st["I_priv"], st["collected_r"], st["mode"],
                     st["exploited_patch_id"]]
        else:
            x, y, th, v = swarm.x, swarm.y, swarm.theta, swarm.v
            extra = [None] * 6
        if x.numel() != self.n:
            raise L.AbmError(L.ABM_ERR_INVALID, "the swarm's agent count differs from the log's")
        bits, R = None, 0
        if fields and self.R is not None:
            cfgs = swarm.configs if forage else swarm.batch.configs
            if any(c.n_rays != self.R for c in cfgs):
                raise L.AbmError(L.ABM_ERR_INVALID, "field logging needs one n_rays across the batch")
            if swarm.fields.numel() < self.n * ((self.R + 31) // 32):
                raise L.AbmError(L.ABM_ERR_INVALID, "the swarm holds no fields of every agent")
            bits, R = swarm.fields, self.R
        L.check(self.lib.abm_log_record(
            C.byref(self._buf), self.t, self.T, self.n, L.ptr(x), L.ptr(y), L.ptr(th), L.ptr(v),
            1 if x.dtype == torch.float32 else 0, *[L.ptr(a) for a in extra], L.ptr(bits), R,
            0 if forage else 1, L.stream_ptr(stream)))
        self.t += 1

    def agents_dict(self, agents: Sequence[int], ids: Optional[Sequence[int]] = None):
        """The reference's agents_dict for the agents ``agents`` (global
        indices, e.g. one replica's slice), keyed by ``ids`` (default 0..)."""
        idx = list(agents)
        ids = list(range(len(idx))) if ids is None else list(ids)
        T = self.t
        cpu = lambda t: None if t is None else t[:T, idx].cpu().numpy()
        px, py, ori, vel = cpu(self.posx), cpu(self.posy), cpu(self.orientation), cpu(self.velocity)
        extra = {k: cpu(getattr(self, k)) for k in ("w", "u", "I_priv", "mode", "collected_r",
                                                    "expl_patch_id")}
        up = None if self.up is None else self.up[:T, idx].cpu().numpy().view(np.uint32)
        down = None if self.down is None else self.down[:T, idx].cpu().numpy().view(np.uint32)
        out = {}
        for c, aid in enumerate(ids):
            d = {"agent_name": f"agent-{pad_to_n_digits(aid, n=2)}",
                 "posx": [int(a) for a in px[:, c]], "posy": [int(a) for a in py[:, c]],
                 "orientation": [float(a) for a in ori[:, c]],
                 "velocity": [float(a) for a in vel[:, c]]}
            # a VSWRM agent keeps Agent's initial foraging state, logged as
            # constants (abm/agent/agent.py:69-126: w = u = I_priv = 0,
            # mode "explore" -> 0, collected_r = 0, exploited_patch_id = -1)
            for key, name, conv, const in (("w", "w", float, 0.0), ("u", "u", float, 0.0),
                                           ("I_priv", "Ipriv", float, 0.0), ("mode", "mode", int, 0),
                                           ("collected_r", "collectedr", float, 0.0),
                                           ("expl_patch_id", "expl_patch_id", int, -1)):
                d[name] = ([conv(a) for a in extra[key][:, c]] if extra[key] is not None
                           else [const] * T)
            if up is not None:
                d["vfield_up"] = [f"{_edges(up[t, c], self.R)}" for t in range(T)]
                d["vfield_down"] = [f"{_edges(down[t, c], self.R)}" for t in range(T)]
            out[aid] = d
        return out

    def save_json(self, save_dir: str, agents: Sequence[int], ids=None, filename="agent_data"):
        """The agent JSON file of save_ifdb_as_csv(use_ram=True, as_zar=False)."""
        os.makedirs(save_dir, exist_ok=True)
        path = os.path.join(save_dir, f"{filename}.json")
        with open(path, "w") as f:
            json.dump(self.agents_dict(agents, ids), f)
        return path


def _edges(words, R):
    """Indices of the set bits as the int64 array np.where returns."""
    bits = np.unpackbits(np.asarray(words, dtype="<u4").view(np.uint8), bitorder="little")[:R]
    return np.flatnonzero(bits)
