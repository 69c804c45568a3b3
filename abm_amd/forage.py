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

    def set_path(self, path: str = "auto"):
        """How replicas are updated: "auto" (the step kernel for replicas that
        fit one workgroup's LDS, sweeps for larger ones) or "sweep" (sweeps for
        every replica without visual exclusion; the same results, used to
        compare the two).  Sweeps read a few control words per round, so a
        step with swept replicas synchronises the stream."""
        code = {"auto": L.FORAGE_PATH_AUTO, "sweep": L.FORAGE_PATH_SWEEP}[path]
        L.check(self.lib.abm_forage_set_path(self.handle, code))

    def sweep_counts(self):
        """(replica updates made by sweeps, sweeps they took) since creation."""
        u, s = C.c_int64(), C.c_int64()
        L.check(self.lib.abm_forage_sweep_counts(self.handle, C.byref(u), C.byref(s)))
        return u.value, s.value

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


# ------------------------------------------------ resources on the device
@dataclass
class WorldConfig:
    """Resource patches of one foraging replica: the Simulation kwargs of the
    resource loop (sims.py:60-68, 160-177) with their defaults, except
    ``pooling_time``, which defaults to 0 as in the reference's .env (with a
    positive pooling time the reference reads ``pool_success`` before any
    notify has set it; here it starts at 0)."""
    N_resc: int = 10
    patch_radius: int = 30
    min_resc_perpatch: int = 200
    max_resc_perpatch: int = 1000
    min_resc_quality: float = 0.1
    max_resc_quality: float = 1.0
    pooling_time: int = 0
    regenerate_patches: bool = True
    agent_consumption: float = 1.0
    teleport_exploit: bool = True
    allow_border_patch_overlap: bool = False
    collide_agents: bool = False
    ghost_mode: bool = True

    def to_c(self) -> L.ForageWorldParams:
        max_q = self.min_resc_quality if self.max_resc_quality < 0 else self.max_resc_quality
        max_u = self.min_resc_perpatch + 1 if self.max_resc_perpatch < 0 else self.max_resc_perpatch
        return L.ForageWorldParams(
            n_patches=int(self.N_resc), patch_radius=int(self.patch_radius),
            min_units=int(self.min_resc_perpatch), max_units=int(max_u),
            pooling_time=int(self.pooling_time), regenerate=int(bool(self.regenerate_patches)),
            teleport_exploit=int(bool(self.teleport_exploit)),
            allow_border_overlap=int(bool(self.allow_border_patch_overlap)),
            collide_agents=int(bool(self.collide_agents)), ghost_mode=int(bool(self.ghost_mode)),
            consumption=float(self.agent_consumption), min_quality=float(self.min_resc_quality),
            max_quality=float(max_q))


_PATCH_F64 = ("x", "y", "radius", "units", "left", "quality")
_MT_WORDS = 625                       # abm_mt19937_state: 624 key words and the position
_MT_BYTES = 4 * _MT_WORDS
TIMEOUT_MESSAGE = "Reached timeout while trying to create resources without overlap!"
CROWDED_MESSAGE = ("a colliding agent has more than 256 agents within 2 r + 20 (the collision "
                   "kernel's proximity limit); the replica was not stepped")


class ResourceWorld:
    """Resource patches of every replica of a ``ForageSwarm`` and the whole
    loop body of ``Simulation.start`` around them (sims.py:736-861) on the
    device: ``create_resources`` places the initial patches (sims.py:538-542),
    ``step`` runs the agent-agent collisions of replicas whose
    ``WorldConfig.collide_agents`` is set (sims.py:736-787), the agent-resource
    interaction, the regeneration of depleted patches and ``Agent.update`` for
    all replicas.

    Random numbers: each replica's numpy generator is peeked as raw 32-bit
    words (patch placement draws randint and uniform from them, the random
    walk two per uniform) and advanced afterwards by exactly the words used.
    A replica whose placements need more than ``regen_words`` words is
    resumed with a larger buffer; nothing is drawn twice.

    With ``device_rng=True`` the generators' states move to device memory
    instead and the kernels draw from them there: a step is two launches and
    no host work or synchronisation.  The numpy generators are then stale
    until ``sync_rng()`` writes the device states back into them (call it
    before anything else draws from them); ``check()`` reads the replicas'
    status and raises the reference's exception for a placement timeout, or
    the crowded-collision error.  Steps in this mode do not read the status:
    a replica whose placements timed out (or whose collisions were too
    crowded) is skipped, and stays frozen on every later step, until
    ``check()`` or ``sync_rng()`` raises for it; call ``check()`` every few
    steps in long runs."""

    def __init__(self, swarm: ForageSwarm, configs: Sequence[WorldConfig], regen_words: int = 32,
                 device_rng: bool = False):
        self.swarm, self.lib = swarm, swarm.lib
        self.configs = list(configs)
        nb = len(swarm.configs)
        if len(self.configs) != nb:
            raise ValueError("give one WorldConfig per replica")
        arr = (L.ForageWorldParams * nb)(*[c.to_c() for c in self.configs])
        h = C.c_void_p()
        with torch.cuda.device(swarm.device):
            L.check(self.lib.abm_forage_world_create(swarm.handle, arr, C.byref(h)))
        self.handle = h
        self.patch_offset = np.zeros(nb + 1, dtype=np.int64)
        L.check(self.lib.abm_forage_world_get_offsets(h, self.patch_offset.ctypes.data_as(C.c_void_p)))
        dev = swarm.device
        slots = max(1, int(self.patch_offset[-1]))
        self.p = {k: torch.zeros(slots, dtype=torch.float64, device=dev) for k in _PATCH_F64}
        self.p["id"] = torch.zeros(slots, dtype=torch.int32, device=dev)
        self.p["count"] = torch.zeros(nb, dtype=torch.int32, device=dev)
        self.pool_success = torch.zeros(int(swarm.info.n_agents_total), dtype=torch.int32, device=dev)
        self._ret = torch.zeros((2, nb), dtype=torch.int32, device=dev)   # used, status
        self._woff = torch.zeros(nb, dtype=torch.int64, device=dev)
        self._wcnt = torch.zeros(nb, dtype=torch.int32, device=dev)
        self._words_host = torch.zeros(0, dtype=torch.int32).pin_memory()
        self._words = torch.zeros(1, dtype=torch.int32, device=dev)
        self.regen_words = int(regen_words)
        self.resumes = 0      # phases resumed for want of words
        self._n = np.diff(swarm.agent_offset).astype(np.int64)
        self.device_rng = bool(device_rng)
        self._mt = None
        if self.device_rng:
            self._mt = torch.zeros(nb * _MT_WORDS, dtype=torch.int32, device=dev)
            self.rng_to_device()

    # ---------------------------------------------------------------- generators
    def rng_to_device(self):
        """Copy every replica's numpy generator state to the device."""
        addrs = self.swarm._state_addrs()
        host = np.empty((len(addrs), _MT_WORDS), dtype=np.uint32)
        for b, a in enumerate(addrs):
            C.memmove(host[b].ctypes.data, int(a), _MT_BYTES)
        self._mt.copy_(torch.from_numpy(host.view(np.int32).reshape(-1)))

    def sync_rng(self):
        """Write the device generator states back into the numpy generators,
        so they stand where the reference's would (after ``check()``)."""
        if not self.device_rng:
            return
        self.check()
        host = self._mt.cpu().numpy().view(np.uint32).reshape(-1, _MT_WORDS)
        for b, a in enumerate(self.swarm._state_addrs()):
            C.memmove(int(a), host[b].ctypes.data, _MT_BYTES)

    def check(self):
        """Raise for a replica whose placements timed out (the reference's
        exception) or whose patch count is invalid."""
        status = self._ret[1].cpu().numpy()
        if np.any(status == L.WORLD_TIMEOUT) or np.any(status == L.WORLD_NEED_WORDS):
            raise Exception(TIMEOUT_MESSAGE)
        if np.any(status == L.WORLD_BAD_COUNT):
            raise L.AbmError(L.ABM_ERR_INVALID, "a patch count is outside [0, N_resc]")
        if np.any(status == L.WORLD_CROWDED):
            raise L.AbmUnsupported(L.ABM_ERR_UNSUPPORTED, CROWDED_MESSAGE)

    def close(self):
        if getattr(self, "handle", None):
            self.lib.abm_forage_world_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # ---------------------------------------------------------------- patches
    def patches(self, b):
        """Replica b's live patches in Group order, as a dict of numpy arrays."""
        lo = int(self.patch_offset[b])
        n = int(self.p["count"][b].item())
        return {k: self.p[k][lo:lo + n].cpu().numpy().copy() for k in ("id",) + _PATCH_F64}

    def set_patches(self, b, id, x, y, radius, units, left, quality):
        """Overwrite replica b's patch list (Group order)."""
        lo, cap = int(self.patch_offset[b]), int(self.patch_offset[b + 1] - self.patch_offset[b])
        n = len(id)
        if n > cap:
            raise L.AbmError(L.ABM_ERR_INVALID, f"replica {b} holds at most {cap} patches")
        vals = dict(id=id, x=x, y=y, radius=radius, units=units, left=left, quality=quality)
        for k, v in vals.items():
            t = self.p[k]
            t[lo:lo + n] = torch.as_tensor(np.asarray(v, dtype=np.int32 if k == "id" else np.float64))
        self.p["count"][b] = n

    # ---------------------------------------------------------------- phases
    def create_resources(self, stream=None):
        """create_resources (sims.py:538-542): fill every replica's list."""
        self._run(L.WORLD_FILL, False, stream)

    def step(self, fields=False, stream=None):
        """One loop body for every replica: interaction, regeneration, update."""
        self._run(L.WORLD_STEP, fields, stream)

    def _run(self, phase, fields, stream):
        if self.device_rng:
            return self._run_dev(phase, fields, stream)
        sw = self.swarm
        nb = len(sw.configs)
        reserve = np.zeros(nb, dtype=np.int64) if phase == L.WORLD_FILL else 2 * self._n
        active = np.ones(nb, dtype=bool)
        budget, ph = self.regen_words, phase
        st = L.ForageState(**{k: sw.t[k].data_ptr() for k in _F64 + ("novelty",) + _I32})
        ps = L.PatchState(**{k: self.p[k].data_ptr() for k in ("count", "id") + _PATCH_F64})
        fb = L.ptr(sw.fields) if fields else None
        addrs = sw._state_addrs()
        with torch.cuda.device(sw.device):
            s = stream if stream is not None else torch.cuda.current_stream()
            with torch.cuda.stream(s):
                while True:
                    cnt = np.where(active, reserve + budget, 0).astype(np.int64)
                    off = np.concatenate([[0], np.cumsum(cnt)[:-1]]).astype(np.int64)
                    total = int(cnt.sum())
                    if self._words_host.numel() < total:
                        self._words_host = torch.zeros(total, dtype=torch.int32).pin_memory()
                    if self._words.numel() < total:
                        self._words = torch.zeros(total, dtype=torch.int32, device=sw.device)
                    L.check(self.lib.abm_mt19937_peek32(addrs.ctypes.data_as(C.c_void_p), nb,
                                                        cnt.ctypes.data_as(C.c_void_p),
                                                        self._words_host.data_ptr()))
                    self._words[:total].copy_(self._words_host[:total], non_blocking=True)
                    self._woff.copy_(torch.from_numpy(off))
                    self._wcnt.copy_(torch.from_numpy(cnt.astype(np.int32)))
                    L.check(self.lib.abm_forage_world_run(
                        self.handle, ph, C.byref(st), L.ptr(self.pool_success), C.byref(ps),
                        L.ptr(self._words), L.ptr(self._woff), L.ptr(self._wcnt), L.ptr(self._ret[0]),
                        L.ptr(self._ret[1]), fb, L.stream_ptr(s)))
                    used, status = self._ret.cpu().numpy()
                    L.check(self.lib.abm_mt19937_skip32(addrs.ctypes.data_as(C.c_void_p), nb,
                                                        used.astype(np.int64).ctypes.data_as(C.c_void_p)))
                    if np.any(status[active] == L.WORLD_TIMEOUT):
                        raise Exception(TIMEOUT_MESSAGE)
                    if np.any(status == L.WORLD_BAD_COUNT):
                        raise L.AbmError(L.ABM_ERR_INVALID, "a patch count is outside [0, N_resc]")
                    if np.any(status == L.WORLD_CROWDED):
                        raise L.AbmUnsupported(L.ABM_ERR_UNSUPPORTED, CROWDED_MESSAGE)
                    need = status == L.WORLD_NEED_WORDS
                    if not need.any():
                        return
                    budget = max(16 * budget, 64)
                    if budget > (1 << 24):
                        raise L.AbmError(L.ABM_ERR_UNSUPPORTED, "patch placement needs more than 2^24 words")
                    ph, active = L.WORLD_RESUME, need
                    self.resumes += 1

    def _run_dev(self, phase, fields, stream):
        sw = self.swarm
        st = L.ForageState(**{k: sw.t[k].data_ptr() for k in _F64 + ("novelty",) + _I32})
        ps = L.PatchState(**{k: self.p[k].data_ptr() for k in ("count", "id") + _PATCH_F64})
        fb = L.ptr(sw.fields) if fields else None
        with torch.cuda.device(sw.device):
            s = stream if stream is not None else torch.cuda.current_stream()
            L.check(self.lib.abm_forage_world_run_dev(
                self.handle, phase, C.byref(st), L.ptr(self.pool_success), C.byref(ps), L.ptr(self._mt),
                L.ptr(self._ret[0]), L.ptr(self._ret[1]), fb, L.stream_ptr(s)))


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
