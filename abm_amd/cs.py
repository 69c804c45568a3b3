"""Cooperative-signalling perception on the GPU.

The cooperative-signalling agent perceives the others through its own
``projection_field`` (abm/projects/cooperative_signaling/cs_agent/
cs_supcalc.py:204-295) and combines the per-object rows two ways
(cs_agent.py:83-131): the crowding density (rows summed, divided by
``len(agents)``) and the signalling density (maximum over meter-scaled rows
of the signalling agents and the decayed memory of earlier signallers).

Everything here goes through ``abm_cs_projection`` in ``libabm_hip.so``
(float64, the reference's operation order); the memory of signalling agents
is host bookkeeping, as in the reference.

* ``projection_field``            -> cs_supcalc.projection_field
* ``CSProjection``                -> the batched form: many focal agents, each
                                     with its own object list, in one launch
* ``calc_crowing_density_proj``   -> CSAgent.calc_crowing_density_proj
* ``calc_others_signaling_density_proj``
                                  -> CSAgent.calc_others_signaling_density_proj
* ``SignalingMemory``             -> CSAgent.memory_signaling and
                                     extend_signaling_agent_positions_with_memory
"""
from __future__ import annotations

import ctypes as C
from typing import List, Optional, Sequence, Tuple

import numpy as np
import torch

from . import _lib as L

# cs_params defaults (abm/projects/cooperative_signaling/cs_contrib/cs_params.py:13-19)
MAX_PROJ_SIZE_PERCENTAGE = 0.5
MEMORY_DEPTH = 5


def _device():
    if not torch.cuda.is_available():
        raise L.AbmError(L.ABM_ERR_UNSUPPORTED, "abm_amd needs a ROCm GPU")
    return torch.device("cuda")


class CSProjection:
    """Batched cooperative-signalling projections for one field geometry.

    ``focal``: (n_focal, 4) float64 device tensor of (x, y, radius,
    orientation); ``obj_offset``: (n_focal + 1,) int64 CSR offsets into
    ``obj``, an (n_obj, 2) float64 tensor of object positions; ``meters``:
    optional (n_obj,) float64.  Positions are sprite top-left corners.
    """

    def __init__(self, v_field_resolution: int, fov=(-np.pi, np.pi), max_proj_size=None,
                 check_offsets: bool = True):
        self.R = int(v_field_resolution)
        # the offsets bound every object read of the kernel; a caller that
        # built them itself may skip the host check (a device-to-host copy)
        self.check_offsets = check_offsets
        self.params = L.CSParams(self.R, 0 if max_proj_size is None else 1, float(fov[0]),
                                 float(fov[1]), 0.0 if max_proj_size is None else float(max_proj_size))
        self.lib = L.load()

    def _run(self, mode, focal, obj_offset, obj, meters, norm, out, stream=None):
        n_focal = int(focal.shape[0])
        if n_focal == 0:
            return out
        if tuple(focal.shape) != (n_focal, 4) or tuple(obj_offset.shape) != (n_focal + 1,):
            raise L.AbmError(L.ABM_ERR_INVALID, "focal must be (n, 4) and obj_offset (n + 1,)")
        if focal.dtype != torch.float64 or obj.dtype != torch.float64 or obj_offset.dtype != torch.int64:
            raise L.AbmError(L.ABM_ERR_INVALID, "focal/obj must be float64 and obj_offset int64")
        n_obj = int(obj.shape[0])
        if obj.dim() != 2 or obj.shape[1] != 2:
            raise L.AbmError(L.ABM_ERR_INVALID, "obj must be (n_obj, 2)")
        if meters is not None and (meters.dtype != torch.float64 or tuple(meters.shape) != (n_obj,)):
            raise L.AbmError(L.ABM_ERR_INVALID, "meters must be float64 (n_obj,)")
        if norm is not None and (norm.dtype != torch.float64 or tuple(norm.shape) != (n_focal,)):
            raise L.AbmError(L.ABM_ERR_INVALID, "norm must be float64 (n_focal,)")
        if self.check_offsets:
            off = obj_offset.cpu()
            if int(off[0]) != 0 or int(off[-1]) != n_obj or bool((off[1:] < off[:-1]).any()):
                raise L.AbmError(L.ABM_ERR_INVALID, "obj_offset must rise from 0 to n_obj")
        if mode == L.CS_ROWS and n_obj == 0:
            return out                       # no rows to write
        obj_ptr = L.ptr(obj) if n_obj > 0 else None
        L.check(self.lib.abm_cs_projection(C.byref(self.params), n_focal, L.ptr(focal),
                                           L.ptr(obj_offset), obj_ptr, L.ptr(meters), L.ptr(norm),
                                           mode, L.ptr(out), L.stream_ptr(stream)))
        return out

    def _out(self, rows, focal, out):
        if out is None:
            return torch.empty((rows, self.R), dtype=torch.float64, device=focal.device)
        if out.dtype != torch.float64 or tuple(out.shape) != (rows, self.R) or not out.is_contiguous():
            raise L.AbmError(L.ABM_ERR_INVALID, f"out must be a contiguous float64 ({rows}, {self.R})")
        return out

    def rows(self, focal, obj_offset, obj, meters=None, out=None, stream=None):
        """Per-object rows as projection_field returns them (flipped, FOV-masked)."""
        out = self._out(int(obj.shape[0]), focal, out)
        return self._run(L.CS_ROWS, focal, obj_offset, obj, meters, None, out, stream)

    def sum(self, focal, obj_offset, obj, norm=None, meters=None, out=None, stream=None):
        """Rows summed over each focal agent's objects, divided by norm."""
        out = self._out(int(focal.shape[0]), focal, out)
        return self._run(L.CS_SUM, focal, obj_offset, obj, meters, norm, out, stream)

    def max(self, focal, obj_offset, obj, meters=None, out=None, stream=None):
        """Maximum over each focal agent's rows (zeros without objects)."""
        out = self._out(int(focal.shape[0]), focal, out)
        return self._run(L.CS_MAX, focal, obj_offset, obj, meters, None, out, stream)


def _focal_tensor(position, radius, orientation, dev):
    return torch.tensor([[float(position[0]), float(position[1]), float(radius), float(orientation)]],
                        dtype=torch.float64, device=dev)


def _objects(object_positions, dev):
    P = np.asarray([np.asarray(p, dtype=np.float64) for p in object_positions],
                   dtype=np.float64).reshape(-1, 2)
    return torch.as_tensor(P, device=dev)


def projection_field(fov, v_field_resolution, position, radius, orientation,
                     object_positions, object_meters=None, max_proj_size=None):
    """cs_supcalc.projection_field (cs_supcalc.py:204-295) on the GPU: rows of
    shape (len(object_positions), v_field_resolution)."""
    dev = _device()
    proj = CSProjection(v_field_resolution, fov, max_proj_size)
    obj = _objects(object_positions, dev)
    off = torch.tensor([0, obj.shape[0]], dtype=torch.int64, device=dev)
    meters = None
    if object_meters is not None:
        meters = torch.as_tensor(np.asarray(object_meters, dtype=np.float64).reshape(-1), device=dev)
    return proj.rows(_focal_tensor(position, radius, orientation, dev), off, obj, meters).cpu().numpy()


class SignalingMemory:
    """An agent's memory of signalling agents (cs_agent.py:106-116, 133-146):
    each step's (positions, meters) is inserted at the front; the extended
    list takes entry n with meters scaled by 1 - n / depth, and the oldest
    entry is dropped once n reaches depth."""

    def __init__(self):
        self.entries: List[Tuple[list, list]] = []

    def push(self, positions, meters):
        self.entries.insert(0, (list(positions), list(meters)))

    def extended(self, depth: int):
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


def calc_crowing_density_proj(agent, agents, max_proj_size_percentage=MAX_PROJ_SIZE_PERCENTAGE):
    """CSAgent.calc_crowing_density_proj (cs_agent.py:83-104): the rows of every
    other agent summed and divided by len(agents).  ``agent`` needs FOV,
    v_field_res, position, radius and orientation; ``agents`` position.  Sets
    ``agent.soc_v_field`` like the reference."""
    dev = _device()
    proj = CSProjection(agent.v_field_res, agent.FOV, agent.v_field_res * max_proj_size_percentage)
    obj = _objects([np.array(a.position) for a in agents if a is not agent], dev)
    off = torch.tensor([0, obj.shape[0]], dtype=torch.int64, device=dev)
    norm = torch.tensor([float(len(agents))], dtype=torch.float64, device=dev)
    out = proj.sum(_focal_tensor(agent.position, agent.radius, agent.orientation, dev), off, obj,
                   norm)[0].cpu().numpy()
    agent.soc_v_field = out
    return out


def calc_others_signaling_density_proj(agent, agents, memory: SignalingMemory,
                                       memory_depth=MEMORY_DEPTH):
    """CSAgent.calc_others_signaling_density_proj (cs_agent.py:106-131): the
    positions and meters of the signalling others now and in memory, combined
    by the maximum over rows.  ``agents`` need position, is_signaling, meter."""
    cur_pos = [np.array(a.position) for a in agents if a.is_signaling if a is not agent]
    cur_met = [a.meter for a in agents if a.is_signaling if a is not agent]
    memory.push(cur_pos, cur_met)
    pos, met = memory.extended(memory_depth)
    R = int(agent.v_field_res)
    if len(pos) == 0:
        return np.zeros(R)
    dev = _device()
    proj = CSProjection(R, agent.FOV, None)
    obj = _objects(pos, dev)
    off = torch.tensor([0, obj.shape[0]], dtype=torch.int64, device=dev)
    meters = torch.as_tensor(np.asarray(met, dtype=np.float64), device=dev)
    return proj.max(_focal_tensor(agent.position, agent.radius, agent.orientation, dev), off, obj,
                    meters)[0].cpu().numpy()
