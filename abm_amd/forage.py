"""Collective foraging on the GPU: batched replicas of ``Agent.update``.

``ForageSwarm`` holds the agent state of many independent foraging
simulations on the device and advances all of them with one launch of
``abm_forage_step_f64`` (the reference's ``self.agents.update(self.agents)``,
abm/simulation/sims.py:861, i.e. ``Agent.update`` in id order,
abm/agent/agent.py:212-283).

Random numbers: the explore branch draws ``np.random.uniform(lo, hi)`` from
numpy's legacy MT19937 stream (abm/agent/supcalc.py:45-46).  Before a step the
host copies each replica's generator, draws N uniforms ahead into a device
buffer, and after the step advances the real generator by exactly the number
the kernel consumed, so the stream stays where the reference would leave it.

``projection_field`` mirrors ``Agent.projection_field`` (agent.py:457-597)
for one focal agent through ``abm_forage_projection``.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass, field
from typing import List, Optional, Sequence

import numpy as np
import torch

from . import _lib as L

MODES = L.MODES


@dataclass
class ForageConfig:
    """One foraging simulation: Simulation kwargs (sims.py:60-161) and the
    decision / movement constants (decision_params.py:13-42,
    movement_params.py:13-24)."""
    n_agents: int
    n_rays: int
    width: float = 500.0
    height: float = 500.0
    window_pad: float = 30.0
    fov: tuple = (-np.pi, np.pi)
    vision_range: float = 150.0
    visual_exclusion: bool = False
    patchwise_exclusion: bool = True
    T_w: float = 0.5
    Eps_w: float = 3.0
    g_w: float = 0.085
    B_w: float = 0.0
    w_max: float = 1.0
    T_u: float = 0.5
    Eps_u: float = 3.0
    g_u: float = 0.085
    B_u: float = 0.0
    u_max: float = 1.0
    S_wu: float = 0.25
    S_uw: float = 0.01
    Tau: int = 10
    F_N: float = 2.0
    F_R: float = 1.0
    exp_vel_max: float = 1.0
    exp_theta_min: float = -0.3
    exp_theta_max: float = 0.3
    reloc_theta_max: float = 0.5
    exp_stop_ratio: float = 0.08

    def to_c(self) -> L.ForageReplica:
        return L.ForageReplica(
            n_agents=int(self.n_agents), n_rays=int(self.n_rays),
            visual_exclusion=int(bool(self.visual_exclusion)),
            patchwise_exclusion=int(bool(self.patchwise_exclusion)), tau=int(self.Tau), pad_=0,
            fov0=float(self.fov[0]), fov1=float(self.fov[1]), width=float(self.width),
            height=float(self.height), window_pad=float(self.window_pad),
            vision_range=float(self.vision_range),
            T_w=self.T_w, Eps_w=self.Eps_w, g_w=self.g_w, B_w=self.B_w, w_max=self.w_max,
            T_u=self.T_u, Eps_u=self.Eps_u, g_u=self.g_u, B_u=self.B_u, u_max=self.u_max,
            S_wu=self.S_wu, S_uw=self.S_uw, F_N=self.F_N, F_R=self.F_R,
            exp_vel_max=self.exp_vel_max, exp_theta_min=self.exp_theta_min,
            exp_theta_max=self.exp_theta_max, reloc_theta_max=self.reloc_theta_max,
            exp_stop_ratio=self.exp_stop_ratio)


_F64 = ("x", "y", "theta", "v", "w", "u", "I_priv", "collected_r", "collected_r_before")
_I32 = ("mode", "overriding_mode", "env_status", "exploited_patch_id", "radius")


class ForageSwarm:
    """Device state of many foraging replicas.

    ``rngs``: one ``np.random.RandomState`` per replica, or ``None`` to use
    numpy's global generator (only for a single replica, as in the
    reference's one-simulation-per-process runs)."""

    def __init__(self, configs: Sequence[ForageConfig], x, y, theta, radius=10, rngs=None,
                 device=None):
        self.lib = L.load()
        self.configs = list(configs)
        self.device = torch.device(device) if device is not None else torch.device("cuda")
        arr = (L.ForageReplica * len(self.configs))(*[c.to_c() for c in self.configs])
        h = C.c_void_p()
        with torch.cuda.device(self.device):
            L.check(self.lib.abm_forage_batch_create(arr, len(self.configs), C.byref(h)))
        self.handle = h
        info = L.BatchInfo()
        L.check(self.lib.abm_forage_batch_get_info(h, C.byref(info)))
        self.info = info
        nb = len(self.configs)
        self.agent_offset = np.zeros(nb + 1, dtype=np.int32)
        self.field_offset = np.zeros(nb + 1, dtype=np.int64)
        L.check(self.lib.abm_forage_batch_get_offsets(
            h, self.agent_offset.ctypes.data_as(C.c_void_p),
            self.field_offset.ctypes.data_as(C.c_void_p)))
        taus = {c.Tau for c in self.configs}
        if len(taus) != 1:
            raise ValueError("all replicas of a batch must share Tau (novelty length)")
        self.tau = taus.pop()
        if rngs is None and nb != 1:
            raise ValueError("give one RandomState per replica (the global stream serves one)")
        self.rngs = rngs
        n = int(info.n_agents_total)
        host = {k: np.zeros(n) for k in _F64}
        host["x"], host["y"], host["theta"] = (np.asarray(a, dtype=np.float64).reshape(n)
                                                for a in (x, y, theta))
        ints = {k: np.zeros(n, dtype=np.int32) for k in _I32}
        ints["overriding_mode"][:] = -1
        ints["exploited_patch_id"][:] = -1
        ints["radius"][:] = np.broadcast_to(np.asarray(radius, dtype=np.int32), (n,))
        self.t = {k: torch.as_tensor(v, device=self.device) for k, v in host.items()}
        self.t.update({k: torch.as_tensor(v, device=self.device) for k, v in ints.items()})
        self.t["novelty"] = torch.zeros(n * self.tau, dtype=torch.float64, device=self.device)
        self.fields = torch.zeros(max(1, int(info.field_words)), dtype=torch.int32, device=self.device)
        self.u_off = torch.as_tensor(self.agent_offset[:-1].astype(np.int64), device=self.device)
        self.u_cnt = torch.as_tensor(np.diff(self.agent_offset).astype(np.int32), device=self.device)
        self.u_used = torch.zeros(nb, dtype=torch.int32, device=self.device)
        # uniforms drawn ahead: N per replica, staged in pinned host memory
        self._counts = np.diff(self.agent_offset).astype(np.int64)
        self._u_host = torch.zeros(n, dtype=torch.float64, pin_memory=True)
        self.u_dev = torch.zeros(n, dtype=torch.float64, device=self.device)
        self._addr, self._addr_of = None, []
        self._check_rng_layout()

    def close(self):
        if getattr(self, "handle", None):
            self.lib.abm_forage_batch_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # ---------------------------------------------------------------- state
    def set_state(self, **arrays):
        """Overwrite state arrays from host values (e.g. after the host's
        resource interaction)."""
        for k, v in arrays.items():
            t = self.t[k]
            t.copy_(torch.as_tensor(np.asarray(v).reshape(t.shape), dtype=t.dtype))

    def get_state(self, *names):
        names = names or tuple(self.t)
        return {k: self.t[k].cpu().numpy().copy() for k in names}

    def replica_slice(self, b):
        return slice(int(self.agent_offset[b]), int(self.agent_offset[b + 1]))

    # ---------------------------------------------------------------- step
    def _state_addrs(self):
        """Addresses of the replicas' MT19937 states (numpy's mt19937_state,
        the layout of abm_mt19937_state)."""
        gens = self.rngs if self.rngs is not None else [np.random.mtrand._rand]
        if self.rngs is None or self._addr is None or any(
                g is not c for g, c in zip(gens, self._addr_of)):
            self._addr_of = list(gens)
            self._addr = np.array([g._bit_generator.ctypes.state_address for g in gens],
                                  dtype=np.uint64)
        return self._addr

    def _draw_ahead(self):
        """Uniform doubles each replica would draw next, without consuming them."""
        addrs = self._state_addrs()
        L.check(self.lib.abm_mt19937_peek(addrs.ctypes.data_as(C.c_void_p), len(addrs),
                                          self._counts.ctypes.data_as(C.c_void_p),
                                          self._u_host.data_ptr()))
        return self._u_host

    def _advance(self, used):
        used = np.ascontiguousarray(used, dtype=np.int64)
        addrs = self._state_addrs()
        L.check(self.lib.abm_mt19937_skip(addrs.ctypes.data_as(C.c_void_p), len(addrs),
                                          used.ctypes.data_as(C.c_void_p)))

    def _check_rng_layout(self):
        """numpy's generator state must be the published mt19937_state struct:
        compare a peek with draws from a copy of the first generator."""
        g = self.rngs[0] if self.rngs is not None else np.random.mtrand._rand
        tmp = np.random.RandomState()
        tmp.set_state(g.get_state())
        want = tmp.random_sample(4)
        got = np.zeros(4)
        addr = np.array([g._bit_generator.ctypes.state_address], dtype=np.uint64)
        cnt = np.array([4], dtype=np.int64)
        L.check(self.lib.abm_mt19937_peek(addr.ctypes.data_as(C.c_void_p), 1,
                                          cnt.ctypes.data_as(C.c_void_p),
                                          got.ctypes.data_as(C.c_void_p)))
        if not np.array_equal(got, want):
            raise L.AbmError(L.ABM_ERR_UNSUPPORTED, "numpy's MT19937 state layout is not the "
                             "expected mt19937_state struct")

    def step(self, fields=False, stream=None):
        """One Agent.update for every agent of every replica; returns the
        number of uniforms each replica consumed."""
        with torch.cuda.device(self.device):
            s = stream if stream is not None else torch.cuda.current_stream()
            with torch.cuda.stream(s):
                # the pinned staging buffer is free again: the previous step
                # ended with a synchronising read of u_used on this stream
                self.u_dev.copy_(self._draw_ahead(), non_blocking=True)
                st = L.ForageState(**{k: self.t[k].data_ptr() for k in
                                      _F64 + ("novelty",) + _I32})
                fb = L.ptr(self.fields) if fields else None
                L.check(self.lib.abm_forage_step_f64(self.handle, C.byref(st), L.ptr(self.u_dev),
                                                     L.ptr(self.u_off), L.ptr(self.u_cnt),
                                                     L.ptr(self.u_used), fb, L.stream_ptr(s)))
                used = self.u_used.cpu().numpy()
        self._advance(used)
        return used

    def fields_numpy(self, b):
        """Replica b's soc_v_field of every agent, (N, R) uint8, in the
        returned (flipped, FOV-masked) order."""
        from .vf import unpack_bits
        cfg = self.configs[b]
        words = (cfg.n_rays + 31) // 32
        lo = int(self.field_offset[b])
        w = self.fields[lo: lo + cfg.n_agents * words].cpu().numpy().view(np.uint32)
        return unpack_bits(w.reshape(cfg.n_agents, words), cfg.n_rays)


# ------------------------------------------------------- reference API
def projection_field(agent, obstacles, keep_distance_info=False, non_expl_agents=None, fov=None):
    """Agent.projection_field (agent.py:457-597) on the GPU.

    ``agent`` needs position, radius, orientation, v_field_res, FOV,
    visual_exclusion and vision_range; obstacles need position."""
    lib = L.load()
    if not torch.cuda.is_available():
        raise L.AbmError(L.ABM_ERR_UNSUPPORTED, "abm_amd needs a ROCm GPU")
    dev = torch.device("cuda")
    fov = agent.FOV if fov is None else fov
    pos = [np.asarray(o.position, dtype=np.float64) for o in obstacles]
    n_social = len(pos)
    n_non = -1
    if non_expl_agents is not None:
        pos += [np.asarray(o.position, dtype=np.float64) for o in non_expl_agents]
        n_non = len(non_expl_agents)
    P = np.asarray(pos, dtype=np.float64).reshape(-1, 2)
    ox = torch.as_tensor(P[:, 0].copy() if len(P) else np.zeros(1), device=dev)
    oy = torch.as_tensor(P[:, 1].copy() if len(P) else np.zeros(1), device=dev)
    R = int(agent.v_field_res)
    out = torch.zeros(R, dtype=torch.float64, device=dev)
    L.check(lib.abm_forage_projection(
        R, float(fov[0]), float(fov[1]), int(bool(agent.visual_exclusion)),
        float(agent.vision_range), int(bool(keep_distance_info)), float(agent.position[0]),
        float(agent.position[1]), float(agent.orientation), float(agent.radius), L.ptr(ox),
        L.ptr(oy), n_social, n_non, L.ptr(out), L.stream_ptr()))
    return out.cpu().numpy()
