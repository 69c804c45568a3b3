"""Visual flocking (VSWRM) on the GPU: batched replicas and the reference API.

Two layers:

* ``VFSwarm`` -- many independent replicas (ragged: each with its own N, R,
  arena, boundary and VSWRM constants) stepped by one kernel launch per step
  through ``abm_vf_step_f64`` (plain float64) or ``abm_vf_step_f32`` (fused
  float32).  State lives on the device as structure-of-arrays torch tensors.
* Drop-in functions with the names and argument order of
  ``abm/projects/visual_flocking/vf_agent/vf_supcalc.py``:
  ``projection_field`` (:20-138), ``VSWRM_flocking_state_variables``
  (:161-254) and ``dPhi_V_of`` (:257-277), and the agent/group classes in
  ``abm_amd.dropin``.

Every computation goes through the HIP library; there is no CPU path.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass, field
from typing import List, Optional, Sequence

import numpy as np
import torch

from . import _lib as L

_PI = np.pi


@dataclass
class VFConfig:
    """One replica: VFSimulation + vf_params (vf_params.py:12-23, sims.py:60-161).

    ``n_rays`` is the resolution actually used by the agents, i.e. after the
    ``int(R / FOV)`` scaling that VFSimulation applies (vf_sims.py:40-44);
    ``VFConfig.from_env_resolution`` applies it.
    """
    n_agents: int
    n_rays: int
    width: float = 500.0
    height: float = 500.0
    window_pad: float = 30.0
    boundary: str = "walls"
    fov: tuple = (-_PI, _PI)
    GAM: float = 0.1
    V0: float = 1.0
    ALP0: float = 1.0
    ALP1: float = 0.09
    BET0: float = 1.0
    BET1: float = 0.09
    LIMIT_MOVEMENT: bool = False
    MAX_VEL: float = 3.0
    MAX_TH: float = 0.1

    @staticmethod
    def fov_from_ratio(ratio: float):
        return (-ratio * np.pi, ratio * np.pi)                 # sims.py:161

    @classmethod
    def from_env_resolution(cls, n_agents, v_field_res, agent_fov=1.0, **kw):
        """Apply VFSimulation's resolution scaling R <- int(R * (1/FOV))."""
        r = int(v_field_res * (1 / agent_fov))                 # vf_sims.py:42-43
        return cls(n_agents=n_agents, n_rays=r, fov=cls.fov_from_ratio(agent_fov), **kw)

    def to_c(self) -> L.VFReplica:
        if self.boundary not in L.BOUNDARY:
            raise ValueError(f"boundary must be 'walls' or 'infinite', got {self.boundary!r}")
        return L.VFReplica(
            n_agents=int(self.n_agents), n_rays=int(self.n_rays),
            boundary=L.BOUNDARY[self.boundary], limit_movement=int(bool(self.LIMIT_MOVEMENT)),
            fov0=float(self.fov[0]), fov1=float(self.fov[1]),
            width=float(self.width), height=float(self.height), window_pad=float(self.window_pad),
            gam=float(self.GAM), v0=float(self.V0), alp0=float(self.ALP0), alp1=float(self.ALP1),
            bet0=float(self.BET0), bet1=float(self.BET1), max_vel=float(self.MAX_VEL),
            max_th=float(self.MAX_TH))


class VFBatch:
    """Owns an ``abm_vf_batch`` handle (replica descriptors + ray tables)."""

    def __init__(self, configs: Sequence[VFConfig], device=None):
        self.lib = L.load()
        self.configs = list(configs)
        self.device = torch.device(device) if device is not None else torch.device("cuda")
        arr = (L.VFReplica * len(self.configs))(*[c.to_c() for c in self.configs])
        h = C.c_void_p()
        with torch.cuda.device(self.device):
            L.check(self.lib.abm_vf_batch_create(arr, len(self.configs), C.byref(h)))
        self.handle = h
        info = L.BatchInfo()
        L.check(self.lib.abm_vf_batch_get_info(h, C.byref(info)))
        self.info = info
        n = len(self.configs)
        self.agent_offset = np.zeros(n + 1, dtype=np.int32)
        self.field_offset = np.zeros(n + 1, dtype=np.int64)
        L.check(self.lib.abm_vf_batch_get_offsets(
            h, self.agent_offset.ctypes.data_as(C.c_void_p),
            self.field_offset.ctypes.data_as(C.c_void_p)))

    def close(self):
        if getattr(self, "handle", None):
            self.lib.abm_vf_batch_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class VFSwarm:
    """Device state of many replicas and their per-step update.

    Positions are sprite top-left corners as in the reference.  ``precision``
    selects the state dtype and entry point: ``"f64"`` (plain float64 kernel)
    or ``"f32"`` (fused float32 kernel).
    """

    def __init__(self, configs: Sequence[VFConfig], x, y, theta, v=None, radius=5,
                 precision: str = "f64", alp0=None, bet0=None, v0=None, device=None):
        if precision not in ("f64", "f32"):
            raise ValueError("precision must be 'f64' or 'f32'")
        self.precision = precision
        self.batch = VFBatch(configs, device)
        self.device = self.batch.device
        n = int(self.batch.info.n_agents_total)
        dt = torch.float64 if precision == "f64" else torch.float32

        def dev(a, dtype=dt, fill=None):
            if a is None:
                return None
            t = torch.as_tensor(np.asarray(a, dtype=np.float64 if dtype != torch.int32 else np.int32))
            t = t.reshape(-1)
            if t.numel() == 1 and n > 1:
                t = t.expand(n)
            if t.numel() != n:
                raise ValueError(f"state array has {t.numel()} entries, batch has {n} agents")
            return t.to(dtype=dtype).contiguous().to(self.device)

        self.x = dev(x)
        self.y = dev(y)
        self.theta = dev(theta)
        self.v = dev(np.zeros(n) if v is None else v)
        self.radius = dev(radius, torch.int32)
        self.alp0 = dev(alp0)
        self.bet0 = dev(bet0)
        self.v0 = dev(v0)
        self.fields = torch.zeros(max(1, int(self.batch.info.field_words)), dtype=torch.int32,
                                  device=self.device)
        self.diag = torch.zeros(6 * n, dtype=dt, device=self.device)

    def load(self, x=None, y=None, theta=None, v=None, radius=None, alp0=None, bet0=None, v0=None):
        """Overwrite state arrays in place from host values (None keeps an
        array; per-agent overrides are replaced when given)."""
        for name, val in (("x", x), ("y", y), ("theta", theta), ("v", v), ("radius", radius)):
            if val is not None:
                t = getattr(self, name)
                t.copy_(torch.as_tensor(np.asarray(val).reshape(t.shape), dtype=t.dtype))
        n = self.x.numel()
        dt = self.x.dtype
        for name, val in (("alp0", alp0), ("bet0", bet0), ("v0", v0)):
            if val is not None:
                setattr(self, name, torch.as_tensor(np.asarray(val, dtype=np.float64).reshape(n),
                                                    dtype=dt).to(self.device))

    # -------------------------------------------------------------- stepping
    def _state(self):
        S = L.VFStateF64 if self.precision == "f64" else L.VFStateF32
        p = lambda t: None if t is None else t.data_ptr()
        return S(x=p(self.x), y=p(self.y), theta=p(self.theta), v=p(self.v), radius=p(self.radius),
                 alp0=p(self.alp0), bet0=p(self.bet0), v0=p(self.v0))

    def step(self, order: str = "sequential", fields: bool = False, diag: bool = False,
             count: bool = False, verify: bool = False, kahan: bool = False, stream=None):
        """Advance every replica by one step (one kernel launch).

        ``count`` / ``verify`` (fused kernel) keep the cumulative pair,
        float64 re-check and float32-vs-float64 disagreement counters;
        ``count`` on the float64 kernel keeps its near-tie count.  ``kahan``
        (fused kernel, batches the grouped kernel takes): Kahan-compensated
        ray sums and group sums in float64."""
        if order not in ("sequential", "synchronous"):
            raise ValueError(order)
        flags = L.ORDER_SEQUENTIAL if order == "sequential" else L.ORDER_SYNCHRONOUS
        if diag:
            flags |= L.WANT_DIAG
        if count:
            flags |= L.COUNT_RECHECKS
        if verify:
            flags |= L.VERIFY_FAST_PATH
        if kahan:
            flags |= L.KAHAN_SUMS
        st = self._state()
        fb = L.ptr(self.fields) if fields else None
        dg = L.ptr(self.diag) if diag else None
        fn = self.batch.lib.abm_vf_step_f64 if self.precision == "f64" else self.batch.lib.abm_vf_step_f32
        with torch.cuda.device(self.device):
            L.check(fn(self.batch.handle, C.byref(st), fb, dg, flags, L.stream_ptr(stream)))

    def fallback_counts(self, stream=None):
        """(pairs, float64 re-checks, float32/float64 disagreements), cumulative."""
        tot, rec, dis = C.c_int64(), C.c_int64(), C.c_int64()
        L.check(self.batch.lib.abm_vf_f32_fallback_count(self.batch.handle, C.byref(tot),
                                                         C.byref(rec), C.byref(dis),
                                                         L.stream_ptr(stream)))
        return tot.value, rec.value, dis.value

    def near_ties(self, stream=None):
        """Pairs the float64 step decided within 1e-12 of a bin or half-width
        tie, cumulative over steps run with ``count=True``."""
        n = C.c_int64()
        L.check(self.batch.lib.abm_vf_f64_near_ties(self.batch.handle, C.byref(n), L.stream_ptr(stream)))
        return n.value

    def autotune(self, trial_steps=5, stream=None):
        if self.precision != "f32":
            raise ValueError("autotune applies to the fused float32 kernel")
        st = self._state()
        w, ms = C.c_int32(), C.c_float()
        L.check(self.batch.lib.abm_vf_autotune(self.batch.handle, C.byref(st), trial_steps,
                                               C.byref(w), C.byref(ms), L.stream_ptr(stream)))
        return w.value, ms.value

    def set_launch(self, waves_per_group: int):
        """Per-replica fused kernel with this many waves per replica (1, 2, 4),
        the 16-wave kernel (one workgroup per replica) for every replica (16),
        the push kernel (one workgroup per replica, pairs off the id-order
        chain) for every replica (32), the chain step (segments of agents
        moved in id order, pairs decided over the whole GPU) for every replica
        in the reference order (64), or the automatic choice (0: single flocks
        above 2048 agents on the chain step; otherwise replicas above 64
        agents on the push kernel where their fields fit in LDS and the
        class's agents over the CUs do not outlast its largest replica's
        chain, the 16-wave kernel where the fields do not fit, the rest on
        the per-replica kernel)."""
        L.check(self.batch.lib.abm_vf_set_launch(self.batch.handle, waves_per_group))

    def last_kernels(self):
        """Names of the kernels the last step launched."""
        m = C.c_uint32()
        L.check(self.batch.lib.abm_vf_last_kernels(self.batch.handle, C.byref(m)))
        names = ("grouped", "per_replica", "wide", "big", "push", "f64", "chain")
        return {n for i, n in enumerate(names) if m.value >> i & 1}

    def set_grouping(self, replicas_per_wave: int):
        """Grouped fused kernel with 1, 2 or 4 replicas per wave; -1 selects the
        per-replica kernel, 0 the automatic choice."""
        L.check(self.batch.lib.abm_vf_set_grouping(self.batch.handle, replicas_per_wave))

    # ----------------------------------------------------------- read back
    def replica_slice(self, b):
        return slice(int(self.batch.agent_offset[b]), int(self.batch.agent_offset[b + 1]))

    def state_numpy(self):
        return tuple(t.detach().cpu().numpy().astype(np.float64)
                     for t in (self.x, self.y, self.theta, self.v))

    def fields_numpy(self, b):
        """Replica b's last fields as (N, R) uint8, raw raster order (un-flipped)."""
        cfg = self.batch.configs[b]
        R, N = cfg.n_rays, cfg.n_agents
        words = (R + 31) // 32
        lo = int(self.batch.field_offset[b])
        w = self.fields[lo: lo + N * words].cpu().numpy().view(np.uint32).reshape(N, words)
        return unpack_bits(w, R)

    def diag_numpy(self):
        return self.diag.detach().cpu().numpy().astype(np.float64).reshape(-1, 6)


def unpack_bits(words: np.ndarray, R: int) -> np.ndarray:
    """uint32 words (..., ceil(R/32)) -> (..., R) uint8, bit k of the field = ray k."""
    w = np.ascontiguousarray(words, dtype=np.uint32)
    b = np.unpackbits(w.view(np.uint8), axis=-1, bitorder="little")
    return b[..., :R]


def pack_bits(bits: np.ndarray) -> np.ndarray:
    """(..., R) 0/1 -> (..., ceil(R/32)) uint32 (inverse of unpack_bits)."""
    b = np.asarray(bits, dtype=np.uint8)
    R = b.shape[-1]
    words = (R + 31) // 32
    padded = np.zeros(b.shape[:-1] + (words * 32,), dtype=np.uint8)
    padded[..., :R] = b
    return np.packbits(padded, axis=-1, bitorder="little").view(np.uint32)


# ------------------------------------------------------- reference API
def _dev():
    if not torch.cuda.is_available():
        raise L.AbmError(L.ABM_ERR_UNSUPPORTED, "abm_amd needs a ROCm GPU")
    return torch.device("cuda")


def projection_field(fov, v_field_resolution, position, radius, orientation, object_positions,
                     object_sizes=None, boundary_cond="walls", arena_width=None,
                     arena_height=None, vision_range=None, ag_id=0):
    """vf_supcalc.projection_field (vf_supcalc.py:20-138) on the GPU.

    Returns float64 ``(n_objects, R)`` 0/1 rows flipped along rays, like the
    reference.  ``vision_range`` is honoured on the host from the distances
    (the VF agents never set it)."""
    lib = L.load()
    dev = _dev()
    R = int(v_field_resolution)
    n = len(object_positions)
    pos = np.asarray(position, dtype=np.float64)
    if n == 0:
        return np.zeros((0, R))
    op = np.asarray([np.asarray(p, dtype=np.float64) for p in object_positions]).reshape(n, 2)
    sizes = np.full(n, float(radius)) if object_sizes is None else np.asarray(object_sizes, dtype=np.float64)
    boundary = 1 if boundary_cond == "infinite" else 0
    W = float(arena_width) if arena_width is not None else 0.0
    H = float(arena_height) if arena_height is not None else 0.0
    ox = torch.as_tensor(op[:, 0].copy(), device=dev)
    oy = torch.as_tensor(op[:, 1].copy(), device=dev)
    os_ = torch.as_tensor(sizes.copy(), device=dev)
    words = (R + 31) // 32
    rows = torch.zeros(n * words, dtype=torch.int32, device=dev)
    rec = torch.zeros(n * 5, dtype=torch.int32, device=dev)
    L.check(lib.abm_vf_projection_rows(R, float(fov[0]), float(fov[1]), boundary, W, H,
                                       float(pos[0]), float(pos[1]), float(orientation),
                                       float(radius), L.ptr(ox), L.ptr(oy), L.ptr(os_), n,
                                       L.ptr(rows), L.ptr(rec), L.stream_ptr()))
    bits = unpack_bits(rows.cpu().numpy().view(np.uint32).reshape(n, words), R).astype(np.float64)
    if vision_range is not None:
        c = pos + radius
        d = np.array([np.linalg.norm(op[i] + sizes[i] - c) for i in range(n)])
        bits[d > vision_range] = 0.0
    return np.flip(bits, axis=1)


def _state_variables(V_rows, vel, Phi, params, ALP0, BET0, V0):
    lib = L.load()
    dev = _dev()
    V_rows = np.atleast_2d(np.asarray(V_rows))
    n, R = V_rows.shape
    if len(Phi) != R:
        raise ValueError("Phi and V_now differ in length")
    bits = torch.as_tensor(pack_bits(V_rows != 0).view(np.int32).reshape(-1).copy(), device=dev)
    col = lambda a: torch.as_tensor(np.broadcast_to(np.asarray(a, dtype=np.float64), (n,)).copy(),
                                    device=dev)
    # keep every argument tensor referenced until the launch is enqueued: a
    # temporary freed early returns its block to the caching allocator, and the
    # next temporary would alias it
    t_vel, t_a0, t_b0, t_v0 = col(vel), col(ALP0), col(BET0), col(V0)
    out = torch.zeros(6 * n, dtype=torch.float64, device=dev)
    dphi = torch.zeros(n * R, dtype=torch.int8, device=dev)
    L.check(lib.abm_vf_state_variables(R, n, L.ptr(bits), L.ptr(t_vel), L.ptr(t_a0),
                                       L.ptr(t_b0), L.ptr(t_v0), float(params.GAM),
                                       float(params.ALP1), float(params.BET1), L.ptr(out),
                                       L.ptr(dphi), L.stream_ptr()))
    return out.cpu().numpy().reshape(n, 6), dphi.cpu().numpy().reshape(n, R)


def VSWRM_flocking_state_variables(vel_now, Phi, V_now, vf_params, t_now=None, V_prev=None,
                                   t_prev=None, verbose=False, ALP0=None, BET0=None, V0=None):
    """vf_supcalc.VSWRM_flocking_state_variables (vf_supcalc.py:161-254) on the GPU.

    ``V_now`` must be binary (the reference only ever passes a binarised
    field).  The computation uses the reference's grid for this length R
    (np.arange(-pi, pi, 2*pi/R)); ``Phi`` is checked against it."""
    Phi = np.asarray(Phi, dtype=np.float64)
    V = np.asarray(V_now, dtype=np.float64)
    if not np.all((V == 0) | (V == 1)):
        raise ValueError("V_now must be a binary field")
    R = len(V)
    if len(Phi) != R or not np.array_equal(Phi, np.arange(-np.pi, np.pi, 2 * np.pi / R)):
        raise ValueError("Phi must be np.arange(-pi, pi, 2*pi/R) of the field's length")
    a0 = vf_params.ALP0 if ALP0 is None else ALP0
    b0 = vf_params.BET0 if BET0 is None else BET0
    v0 = vf_params.V0 if V0 is None else V0
    out, _ = _state_variables(V[None, :], vel_now, Phi, vf_params, a0, b0, v0)
    o = [float(x) for x in out[0]]
    if not verbose:
        return o[0], o[1]
    return tuple(o)


def dPhi_V_of(Phi, V):
    """vf_supcalc.dPhi_V_of (vf_supcalc.py:257-277) for a binary field, on the GPU."""
    V = np.asarray(V, dtype=np.float64)
    if not np.all((V == 0) | (V == 1)):
        raise ValueError("V must be a binary field")
    R = len(V)

    class _P:
        GAM = 0.0; ALP1 = 0.0; BET1 = 0.0
    _, d = _state_variables(V[None, :], 0.0, np.arange(-np.pi, np.pi, 2 * np.pi / R)[:R],
                            _P, 0.0, 0.0, 0.0)
    return d[0].astype(np.float64)
