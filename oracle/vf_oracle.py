"""Oracle: visual-flocking (VSWRM) perception and motion, numpy float64.

TEST INFRASTRUCTURE ONLY -- the checker for ``abm_amd``; never imported by
the product path.

A CPU restatement of the reference's per-agent VSWRM step (scioip34/ABM):

* ``projection_field``      abm/projects/visual_flocking/vf_agent/vf_supcalc.py:20-138
* ``closed_angle``          vf_supcalc.py:142-158 (+ abm/agent/supcalc.py:8-34)
* ``dphi_v_of``             vf_supcalc.py:257-277
* ``vswrm_state_variables`` vf_supcalc.py:161-254
* ``calc_soc_v_proj``       abm/projects/visual_flocking/vf_agent/vf_agent.py:216-237
* ``perform_action``        vf_agent.py:263-287
* ``update_agent_position`` vf_agent.py:289-326 (+ abm/agent/agent.py:605-610)
* ``teleport_infinite_arena`` vf_agent.py:188-204
* ``reflect_from_walls``    abm/agent/agent.py:347-395 (VFAgent without guide lines)
* ``group_update``          pygame ``Group.update`` in id order
                            (abm/projects/visual_flocking/vf_simulation/vf_sims.py:302)

The same numpy calls are made on the same float64 operands as in the
reference (``np.linalg.norm``, ``np.dot``, ``np.arccos``, ``np.arctan``,
``np.linspace``/``np.arange`` grids, ``argmin`` ties), so on a given host this
oracle produces the reference's bits.  Float intermediates are host dependent
(BLAS dot and SVML transcendentals); integer raster decisions are not, except
within an ulp of a tie.

Parity pins (tests/test_oracle_vf.py): the reference's own known-answer tests
(abm/projects/cooperative_signaling/cs_agent/tests/test_cs_supcalc.py:143-168,
same algorithm), hand-derived raster cases (dead ahead, wrap at bins 0/R-1,
the s<2 threshold, torus images, limited FOV), an exact-rational geometry
cross-check of every bin decision, exhaustive dPhi checks for R<=12, closed
forms of the integrals and of the pose update, and torus/translation
invariants.
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field
from typing import List, Optional, Sequence

import numpy as np

_trapezoid = getattr(np, "trapezoid", None) or getattr(np, "trapz")


# ---------------------------------------------------------------- parameters
@dataclass
class VFParams:
    """VSWRM constants (abm/projects/visual_flocking/vf_contrib/vf_params.py:12-23).

    ``GAM`` is read from the env key ``VF_GAMMA`` in the reference while the
    env files set ``VF_GAM``, so in practice it is always the default 0.1.
    """
    GAM: float = 0.1
    V0: float = 1.0
    ALP0: float = 1.0
    ALP1: float = 0.09
    ALP2: float = 0.0
    BET0: float = 1.0
    BET1: float = 0.09
    BET2: float = 0.0
    BOUNDARY: str = "walls"
    LIMIT_MOVEMENT: bool = False
    MAX_VEL: float = 3.0
    MAX_TH: float = 0.1


@dataclass
class VFAgentRec:
    """The attributes of a ``VFAgent`` that the step reads or writes
    (abm/agent/agent.py:20-150, vf_agent.py:13-50)."""
    id: int
    radius: int
    position: object            # np.ndarray at start, list after a move
    orientation: float
    velocity: float
    v_field_res: int
    FOV: tuple
    WIDTH: int
    HEIGHT: int
    window_pad: int
    boundary_cond: str = "walls"
    limit_movement: bool = False
    max_vel: float = 3.0
    max_th: float = 0.1
    ALP0: Optional[float] = None
    BET0: Optional[float] = None
    V0: Optional[float] = None
    PHI: np.ndarray = None
    soc_v_field: np.ndarray = None
    verbose_supcalc: bool = False
    diag: tuple = None

    def __post_init__(self):
        if self.PHI is None:
            # vf_agent.py:44
            self.PHI = np.arange(-np.pi, np.pi, (2 * np.pi) / self.v_field_res)
        if self.soc_v_field is None:
            self.soc_v_field = np.zeros(self.v_field_res)

    @property
    def boundaries_x(self):
        return [self.window_pad, self.window_pad + self.WIDTH]

    @property
    def boundaries_y(self):
        return [self.window_pad, self.window_pad + self.HEIGHT]


# ----------------------------------------------------------- small helpers
def find_nearest(array, value):
    """First index of the minimum of |array - value| (abm/agent/supcalc.py:8-11)."""
    return int(np.abs(np.asarray(array) - value).argmin())


def _unit(v):
    # supcalc.py:14-16 -- norm is np.linalg.norm (BLAS dot + sqrt)
    return v / np.linalg.norm(v)


def angle_between(v1, v2):
    """Signed angle: arccos of the clipped dot of unit vectors, negated when the
    2-D cross product is negative (supcalc.py:19-34)."""
    a = _unit(np.asarray(v1, dtype=np.float64) if not isinstance(v1, np.ndarray) else v1)
    b = _unit(np.asarray(v2, dtype=np.float64) if not isinstance(v2, np.ndarray) else v2)
    ang = np.arccos(np.clip(np.dot(a, b), -1.0, 1.0))
    if a[0] * b[1] - a[1] * b[0] < 0:
        ang = -ang
    return ang


def closed_angle(v1, v2):
    """VF closed angle in [-pi, pi): inclusive 0 <= a' <= pi maps to -a'
    (vf_supcalc.py:142-158)."""
    a = angle_between(v1, v2) % (2 * np.pi)
    if 0 <= a <= np.pi:
        return -a
    return 2 * np.pi - a


# ------------------------------------------------------- projection field
@dataclass
class ObjectRecord:
    """Per-object raster decision of ``projection_field`` (for diagnostics)."""
    index: int
    skipped: bool = False
    closed_angle: float = float("nan")
    distance: float = float("nan")
    vis_angle: float = float("nan")
    phi_target: int = -1
    proj_size: float = float("nan")
    proj_start: int = 0
    proj_end: int = 0
    drawn: bool = False


def projection_records(fov, v_field_resolution, position, radius, orientation,
                       object_positions, object_sizes=None, boundary_cond="walls",
                       arena_width=None, arena_height=None, vision_range=None):
    """Per-object geometry and raster decisions of vf_supcalc.projection_field
    (vf_supcalc.py:38-119), plus the filled (un-flipped) raster rows."""
    R = v_field_resolution
    n = len(object_positions)
    rows = np.zeros((n, R))
    phis = np.linspace(-np.pi, np.pi, R)                               # :39
    position = np.asarray(position) if not isinstance(position, np.ndarray) else position
    centre = position + radius                                         # :42
    edge = position + np.array([1 + np.cos(orientation),
                                1 - np.sin(orientation)]) * radius     # :45-46
    v1 = edge - centre                                                 # :49
    recs = []
    for i, opos in enumerate(object_positions):                        # :55
        rec = ObjectRecord(index=i)
        recs.append(rec)
        if opos[0] == position[0] and opos[1] == position[1]:          # :57
            rec.skipped = True
            continue
        size = radius if object_sizes is None else object_sizes[i]
        ocentre = opos + size                                          # :61-64
        v2 = ocentre - centre                                          # :67
        if boundary_cond == "infinite":                                # :70-83
            if np.abs(v2[0]) > arena_width / 2:
                if centre[0] < ocentre[0]:
                    ocentre[0] -= arena_width
                elif centre[0] > ocentre[0]:
                    ocentre[0] += arena_width
            if np.abs(v2[1]) > arena_height / 2:
                if centre[1] < ocentre[1]:
                    ocentre[1] -= arena_height
                elif centre[1] > ocentre[1]:
                    ocentre[1] += arena_height
            v2 = ocentre - centre
        ca = closed_angle(v1, v2)                                      # :86
        dist = np.linalg.norm(ocentre - centre)                        # :88
        rec.closed_angle, rec.distance = float(ca), float(dist)
        if vision_range is not None and dist > vision_range:           # :91-93
            rec.skipped = True
            continue
        vis = 2 * np.arctan(size / (1 * dist))                         # :96-99
        t = find_nearest(phis, ca)                                     # :102
        fov_px = [find_nearest(phis, fov[0]), find_nearest(phis, fov[1])]  # :105
        ps = (vis / (2 * np.pi)) * R                                   # :114
        start = int(t - np.floor(ps / 2))                              # :116
        end = int(t + np.floor(ps / 2))                                # :117
        rec.vis_angle, rec.phi_target, rec.proj_size = float(vis), t, float(ps)
        rec.proj_start, rec.proj_end = start, end
        if fov_px[0] < start < fov_px[1] or fov_px[0] < end < fov_px[1]:   # :119
            rec.drawn = True
            if start < 0:                                              # :122-124
                rows[i, R + start:R] = 1
                start = 0
            if end >= R:                                               # :125-127
                rows[i, 0:end - (R - 1)] = 1
                end = R
            rows[i, start:end] = 1                                     # :129
    return recs, rows


def projection_field(fov, v_field_resolution, position, radius, orientation,
                     object_positions, object_sizes=None, boundary_cond="walls",
                     arena_width=None, arena_height=None, vision_range=None, ag_id=0):
    """Per-object binary projection rows, flipped along rays (vf_supcalc.py:20-138)."""
    _, rows = projection_records(fov, v_field_resolution, position, radius, orientation,
                                 [np.array(p, dtype=np.float64) if not isinstance(p, np.ndarray)
                                  else p.copy() for p in object_positions],
                                 object_sizes, boundary_cond, arena_width, arena_height,
                                 vision_range)
    return np.flip(rows, axis=1)                                       # :134


# --------------------------------------------------------------- integrals
def dphi_v_of(Phi, V):
    """Circular first difference with the edge-alignment rule (vf_supcalc.py:257-277)."""
    raw = np.diff(np.pad(V, (1, 1), "wrap"))
    if raw[0] > 0 and raw[-1] > 0:
        return raw[0:-1]
    return raw[1:, ...]


def vswrm_state_variables(vel_now, Phi, V_now, params: VFParams, verbose=False,
                          ALP0=None, BET0=None, V0=None):
    """(dvel, dpsi) or the verbose 6-tuple (vf_supcalc.py:161-254).

    The temporal-derivative field is identically zero in the reference
    (:199), so the ALP2/BET2 terms contribute nothing; the spike sums carry no
    dPhi factor (:235, :238)."""
    ALP0 = params.ALP0 if ALP0 is None else ALP0
    BET0 = params.BET0 if BET0 is None else BET0
    V0 = params.V0 if V0 is None else V0
    dt_V = np.zeros(len(Phi))
    dPhi_V = dphi_v_of(Phi, V_now)
    G_vel = -V_now + params.ALP2 * dt_V
    G_vel_spike = np.square(dPhi_V)
    G_psi = -V_now + params.BET2 * dt_V
    G_psi_spike = np.square(dPhi_V)
    a_blob = ALP0 * _trapezoid(np.cos(Phi) * G_vel, Phi)
    a_edge = ALP0 * params.ALP1 * np.sum(np.cos(Phi) * G_vel_spike)
    b_blob = BET0 * _trapezoid(np.sin(Phi) * G_psi, Phi)
    b_edge = BET0 * params.BET1 * np.sum(np.sin(Phi) * G_psi_spike)
    dvel = params.GAM * (V0 - vel_now) + a_blob + a_edge
    dpsi = b_blob + b_edge
    if not verbose:
        return dvel, dpsi
    return dvel, dpsi, a_blob, a_edge, b_blob, b_edge


# ------------------------------------------------------------- agent step
def calc_soc_v_proj(agent: VFAgentRec, agents: Sequence[VFAgentRec]):
    """Binarised sum of the per-object rows, in flipped ray order
    (vf_agent.py:216-237)."""
    others = [ag for ag in agents if ag.id != agent.id]
    vf = projection_field(
        fov=agent.FOV, v_field_resolution=agent.v_field_res,
        position=np.array(agent.position), radius=agent.radius,
        orientation=agent.orientation,
        object_positions=[np.array(ag.position) for ag in others],
        object_sizes=[ag.radius for ag in others],
        boundary_cond=agent.boundary_cond,
        arena_width=agent.WIDTH, arena_height=agent.HEIGHT, ag_id=agent.id)
    s = vf.sum(axis=0)
    s[s > 0] = 1
    return s


def prove_orientation(agent):
    """One-shot wrap into [0, 2pi] (abm/agent/agent.py:605-610)."""
    if agent.orientation < 0:
        agent.orientation = 2 * np.pi + agent.orientation
    if agent.orientation > np.pi * 2:
        agent.orientation = agent.orientation - 2 * np.pi


def _prove_turning(theta, theta_lim):
    # vf_agent.py:311-317
    sgn = np.sign(theta)
    if sgn == 0:
        sgn = +1
    if np.abs(theta) > theta_lim:
        theta = theta_lim * sgn
    return theta


def _prove_velocity_vf(agent, velocity_limit):
    # vf_agent.py:320-326
    sgn = np.sign(agent.velocity)
    if sgn == 0:
        sgn = +1
    if np.abs(agent.velocity) > velocity_limit:
        agent.velocity = velocity_limit * sgn


def update_agent_position(agent: VFAgentRec, theta, vel):
    """Euler step, dt = 1 (vf_agent.py:289-309)."""
    if agent.limit_movement:
        theta = _prove_turning(theta, agent.max_th)
    agent.orientation += theta
    prove_orientation(agent)
    agent.velocity += vel
    if agent.limit_movement:
        _prove_velocity_vf(agent, agent.max_vel)
    new_pos = (agent.position[0] + agent.velocity * np.cos(agent.orientation),
               agent.position[1] - agent.velocity * np.sin(agent.orientation))
    agent.position = list(new_pos)


def perform_action(agent: VFAgentRec, params: VFParams):
    """VSWRM motion from the current field (vf_agent.py:263-287).  A PHI/field
    length mismatch leaves dv = dphi = 0 but still moves the agent."""
    if len(agent.PHI) == len(agent.soc_v_field):
        out = vswrm_state_variables(agent.velocity, agent.PHI, np.flip(agent.soc_v_field),
                                    params, verbose=agent.verbose_supcalc,
                                    ALP0=agent.ALP0, BET0=agent.BET0, V0=agent.V0)
        if agent.verbose_supcalc:
            agent.diag = tuple(float(x) for x in out)
        dv, dphi = out[0], out[1]
    else:
        dv, dphi = 0, 0
        if agent.verbose_supcalc:
            agent.diag = (0.0,) * 6
    update_agent_position(agent, dphi, dv)


def teleport_infinite_arena(agent: VFAgentRec):
    """Periodic boundary by snapping the centre to the opposite wall
    (vf_agent.py:188-204)."""
    x = agent.position[0] + agent.radius
    y = agent.position[1] + agent.radius
    bx, by = agent.boundaries_x, agent.boundaries_y
    if x < bx[0]:
        agent.position[0] = bx[1] - agent.radius
    elif x > bx[1]:
        agent.position[0] = bx[0] + agent.radius
    if y < by[0]:
        agent.position[1] = by[1] - agent.radius
    elif y > by[1]:
        agent.position[1] = by[0] + agent.radius


def reflect_from_walls(agent):
    """Clamp to the wall and turn by +-pi/2 by quadrant (abm/agent/agent.py:347-395)."""
    x = agent.position[0] + agent.radius
    y = agent.position[1] + agent.radius
    bx, by = agent.boundaries_x, agent.boundaries_y
    if x < bx[0]:
        agent.position[0] = bx[0] - agent.radius
        if np.pi / 2 <= agent.orientation < np.pi:
            agent.orientation -= np.pi / 2
        elif np.pi <= agent.orientation <= 3 * np.pi / 2:
            agent.orientation += np.pi / 2
        prove_orientation(agent)
    if x > bx[1]:
        agent.position[0] = bx[1] - agent.radius - 1
        if 3 * np.pi / 2 <= agent.orientation < 2 * np.pi:
            agent.orientation -= np.pi / 2
        elif 0 <= agent.orientation <= np.pi / 2:
            agent.orientation += np.pi / 2
        prove_orientation(agent)
    if y < by[0]:
        agent.position[1] = by[0] - agent.radius
        if np.pi / 2 <= agent.orientation <= np.pi:
            agent.orientation += np.pi / 2
        elif 0 <= agent.orientation < np.pi / 2:
            agent.orientation -= np.pi / 2
        prove_orientation(agent)
    if y > by[1]:
        agent.position[1] = by[1] - agent.radius - 1
        if 3 * np.pi / 2 <= agent.orientation <= 2 * np.pi:
            agent.orientation += np.pi / 2
        elif np.pi <= agent.orientation < 3 * np.pi / 2:
            agent.orientation -= np.pi / 2
        prove_orientation(agent)


def agent_update(agent: VFAgentRec, agents: Sequence[VFAgentRec], params: VFParams):
    """One ``VFAgent.update`` (vf_agent.py:52-77), drawing and the discarded
    line-following probe omitted."""
    agent.soc_v_field = calc_soc_v_proj(agent, agents)
    perform_action(agent, params)
    if agent.boundary_cond == "walls":
        reflect_from_walls(agent)
    elif agent.boundary_cond == "infinite":
        teleport_infinite_arena(agent)


def group_update(agents: List[VFAgentRec], params: VFParams, order: str = "sequential"):
    """One simulation step.

    ``sequential`` is the reference: ``Group.update`` calls each agent in id
    order and in place, so agent k sees agents < k at their new poses.
    ``synchronous`` (not in the reference) computes every field from the
    pre-step poses first."""
    if order == "sequential":
        for ag in agents:
            agent_update(ag, agents, params)
    elif order == "synchronous":
        snapshot = [VFAgentRec(**{**ag.__dict__, "position": list(ag.position)}) for ag in agents]
        fields = [calc_soc_v_proj(ag, snapshot) for ag in agents]
        for ag, fld in zip(agents, fields):
            ag.soc_v_field = fld
            perform_action(ag, params)
            if ag.boundary_cond == "walls":
                reflect_from_walls(ag)
            elif ag.boundary_cond == "infinite":
                teleport_infinite_arena(ag)
    else:
        raise ValueError(order)


# ------------------------------------------------------- array conversions
def make_agents(x, y, theta, v, radius, R, fov, W, H, pad, params: VFParams,
                alp0=None, bet0=None, v0=None) -> List[VFAgentRec]:
    """Build agent records from state arrays (positions are sprite top-left)."""
    out = []
    n = len(x)
    rad = np.broadcast_to(np.asarray(radius), (n,))
    for i in range(n):
        out.append(VFAgentRec(
            id=i, radius=int(rad[i]), position=np.array([x[i], y[i]], dtype=np.float64),
            orientation=float(theta[i]), velocity=float(v[i]), v_field_res=int(R),
            FOV=tuple(fov), WIDTH=W, HEIGHT=H, window_pad=pad,
            boundary_cond=params.BOUNDARY, limit_movement=params.LIMIT_MOVEMENT,
            max_vel=params.MAX_VEL, max_th=params.MAX_TH,
            ALP0=None if alp0 is None else float(alp0[i]),
            BET0=None if bet0 is None else float(bet0[i]),
            V0=None if v0 is None else float(v0[i])))
    return out


def agents_state(agents: Sequence[VFAgentRec]):
    """(x, y, theta, v) float64 arrays."""
    x = np.array([float(a.position[0]) for a in agents])
    y = np.array([float(a.position[1]) for a in agents])
    th = np.array([float(a.orientation) for a in agents])
    v = np.array([float(a.velocity) for a in agents])
    return x, y, th, v


def raster_bits(agents: Sequence[VFAgentRec]):
    """Each agent's field in raw raster order (un-flipped) as a 0/1 uint8 array."""
    return np.stack([np.flip(a.soc_v_field).astype(np.uint8) for a in agents])


def step_arrays(x, y, theta, v, radius, R, fov, W, H, pad, params: VFParams,
                order="sequential", alp0=None, bet0=None, v0=None, verbose=False):
    """One step from state arrays; returns new arrays, the raster-order fields
    and, if ``verbose``, the per-agent (dv, dpsi, a_blob, a_edge, b_blob, b_edge)."""
    ags = make_agents(x, y, theta, v, radius, R, fov, W, H, pad, params, alp0, bet0, v0)
    for a in ags:
        a.verbose_supcalc = verbose
    group_update(ags, params, order)
    nx, ny, nth, nv = agents_state(ags)
    out = (nx, ny, nth, nv, raster_bits(ags))
    if verbose:
        out = out + (np.array([a.diag for a in ags]),)
    return out


def agent_view_step(x, y, theta, v, radius, k, R, fov, W, H, pad, params: VFParams,
                    alp0=None, bet0=None, v0=None):
    """Update agent ``k`` alone against the given poses of every agent.

    Used to check a kernel agent by agent on the exact neighbour poses it saw
    (agents < k already moved, agents > k not yet).  Returns the new
    (x, y, theta, v), the raster-order field and the verbose 6-tuple."""
    ags = make_agents(x, y, theta, v, radius, R, fov, W, H, pad, params, alp0, bet0, v0)
    ag = ags[k]
    ag.verbose_supcalc = True
    agent_update(ag, ags, params)
    return (float(ag.position[0]), float(ag.position[1]), float(ag.orientation),
            float(ag.velocity), np.flip(ag.soc_v_field).astype(np.uint8), ag.diag)


# ------------------------------------------------------ ensemble statistics
def polarization(theta):
    """|sum_i (cos th_i, sin th_i)| / N (abm/loader/data_loader.py:1761-1836)."""
    th = np.asarray(theta, dtype=np.float64)
    return float(np.hypot(np.cos(th).sum(), np.sin(th).sum()) / len(th))
