"""Drop-in agent, group and headless simulation classes.

The reference's host loops call ``self.agents.update(self.agents)`` on a
pygame ``Group`` of agents every step (abm/simulation/sims.py:861,
abm/projects/visual_flocking/vf_simulation/vf_sims.py:302).  Here the agents
are attribute-compatible plain objects and ``AgentGroup.update(agents)`` runs
that whole step -- every agent in id order, in place -- as one GPU launch,
then writes back the attributes the data path reads (position, orientation,
velocity, soc_v_field, w, u, I_priv, mode, collected_r, exploited_patch_id;
abm/monitoring/ifdb.py:55-96).  The reference's RAM/InfluxDB logging and
metarunner can therefore keep reading the agents unchanged.

``VFSimulation`` and ``Simulation`` are headless versions of the reference's
loops (no pygame drawing or event handling) for running without pygame:
the visual-flocking loop (vf_sims.py:212-228, 291-366) and the foraging
loop with agent-resource interaction (sims.py:321-374, 526-552, 702-870;
agent-agent collisions are not supported).

Configuration follows the reference's env-file keys (``read_env``,
``vf_params_from_env``, ``decision_params_from_env``,
``movement_params_from_env``).
"""
from __future__ import annotations

import math
import os
from dataclasses import dataclass, fields
from typing import Iterable, List, Optional

import numpy as np

from . import _lib as L

_PI = np.pi


# ------------------------------------------------------------------ config
def read_env(path) -> dict:
    """KEY=VALUE lines of a .env file (comments, blank lines and surrounding
    quotes handled), like dotenv_values."""
    out = {}
    if not path or not os.path.isfile(path):
        return out
    with open(path) as f:
        for line in f:
            line = line.strip()
            if not line or line.startswith("#") or "=" not in line:
                continue
            k, v = line.split("=", 1)
            k = k.strip()
            if k.startswith("export "):
                k = k[7:].strip()
            v = v.strip()
            if len(v) >= 2 and v[0] == v[-1] and v[0] in "'\"":
                v = v[1:-1]
            out[k] = v
    return out


@dataclass
class VFParams:
    """vf_params.py:12-23.  ``from_env`` reads GAM from the key VF_GAMMA, as the
    reference does (the env files set VF_GAM, so GAM stays 0.1)."""
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

    @classmethod
    def from_env(cls, envconf: dict):
        g = envconf.get
        return cls(GAM=float(g("VF_GAMMA", 0.1)), V0=float(g("VF_V0", 1)), ALP0=float(g("VF_ALP0", 1)),
                   ALP1=float(g("VF_ALP1", 0.09)), ALP2=float(g("VF_ALP2", 0)),
                   BET0=float(g("VF_BET0", 1)), BET1=float(g("VF_BET1", 0.09)),
                   BET2=float(g("VF_BET2", 0)), BOUNDARY=g("BOUNDARY", "walls"),
                   LIMIT_MOVEMENT=bool(float(g("VF_LIMIT_MOVEMENT", "0"))),
                   MAX_VEL=float(g("VF_MAX_VEL", "3")), MAX_TH=float(g("VF_MAX_TH", "0.1")))


@dataclass
class DecisionParams:
    """decision_params.py:13-42."""
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

    @classmethod
    def from_env(cls, e: dict):
        g = e.get
        return cls(T_w=float(g("DEC_TW", 0.5)), Eps_w=float(g("DEC_EPSW", 3)), g_w=float(g("DEC_GW", 0.085)),
                   B_w=float(g("DEC_BW", 0)), w_max=float(g("DEC_WMAX", 1)), T_u=float(g("DEC_TU", 0.5)),
                   Eps_u=float(g("DEC_EPSU", 3)), g_u=float(g("DEC_GU", 0.085)), B_u=float(g("DEC_BU", 0)),
                   u_max=float(g("DEC_UMAX", 1)), S_wu=float(g("DEC_SWU", 0.25)),
                   S_uw=float(g("DEC_SUW", 0.01)), Tau=int(float(g("DEC_TAU", 10))),
                   F_N=float(g("DEC_FN", 2)), F_R=float(g("DEC_FR", 1)))


@dataclass
class MovementParams:
    """movement_params.py:13-24."""
    exp_vel_min: float = 1.0
    exp_vel_max: float = 1.0
    exp_theta_min: float = -0.3
    exp_theta_max: float = 0.3
    reloc_des_vel: float = 1.0
    reloc_theta_max: float = 0.5
    exp_stop_ratio: float = 0.08

    @classmethod
    def from_env(cls, e: dict):
        g = e.get
        return cls(exp_vel_min=float(g("MOV_EXP_VEL_MIN", 1)), exp_vel_max=float(g("MOV_EXP_VEL_MAX", 1)),
                   exp_theta_min=float(g("MOV_EXP_TH_MIN", -0.3)),
                   exp_theta_max=float(g("MOV_EXP_TH_MAX", 0.3)),
                   reloc_des_vel=float(g("MOV_REL_DES_VEL", 1)),
                   reloc_theta_max=float(g("MOV_REL_TH_MAX", 0.5)),
                   exp_stop_ratio=float(g("CONS_STOP_RATIO", 0.08)))


# ------------------------------------------------------------------ agents
class Agent:
    """Foraging agent with the reference's constructor (agent.py:20-166) and
    the attributes its update reads and writes; no drawing."""

    def __init__(self, id, radius, position, orientation, env_size, color=None, v_field_res=800,
                 FOV=(-_PI, _PI), window_pad=30, pooling_time=3, pooling_prob=0.05, consumption=1,
                 vision_range=150, visual_exclusion=False, patchwise_exclusion=True,
                 behave_params=None, decision_params: DecisionParams = None,
                 movement_params: MovementParams = None):
        dp = decision_params or DecisionParams()
        mp = movement_params or MovementParams()
        self.exclude_agents_same_patch = patchwise_exclusion
        self.id = id
        self.radius = radius
        self.position = np.array(position, dtype=np.float64)
        self.orientation = orientation
        self.color = color
        self.v_field_res = v_field_res
        self.pooling_time = pooling_time
        self.pooling_prob = pooling_prob
        self.consumption = consumption
        self.vision_range = vision_range
        self.visual_exclusion = visual_exclusion
        self.FOV = FOV
        self.velocity = 0
        self.collected_r = 0
        self.collected_r_before = 0
        self.exploited_patch_id = -1
        self.mode = "explore"
        self.soc_v_field = np.zeros(self.v_field_res)
        self.vis_field_source_data = {}
        self.is_moved_with_cursor = 0
        self.overriding_mode = None
        b = behave_params
        src = (lambda k, d: b[k]) if b is not None else (lambda k, d: d)
        self.S_wu = src("S_wu", dp.S_wu); self.T_w = src("T_w", dp.T_w); self.w = 0
        self.Eps_w = src("Eps_w", dp.Eps_w); self.g_w = src("g_w", dp.g_w)
        self.B_w = src("B_w", dp.B_w); self.w_max = src("w_max", dp.w_max)
        self.I_priv = 0
        self.novelty = np.zeros(src("Tau", dp.Tau))
        self.S_uw = src("S_uw", dp.S_uw); self.T_u = src("T_u", dp.T_u); self.u = 0
        self.Eps_u = src("Eps_u", dp.Eps_u); self.g_u = src("g_u", dp.g_u)
        self.B_u = src("B_u", dp.B_u); self.u_max = src("u_max", dp.u_max)
        self.F_N = src("F_N", dp.F_N); self.F_R = src("F_R", dp.F_R)
        self.max_exp_vel = src("exp_vel_max", mp.exp_vel_max)
        self.exp_stop_ratio = src("exp_stop_ratio", mp.exp_stop_ratio)
        self.movement_params = mp
        self.time_spent_pooling = 0
        self.env_status_before = 0
        self.env_status = 0
        self.pool_success = 0
        self.WIDTH, self.HEIGHT = env_size
        self.window_pad = window_pad
        self.boundaries_x = [self.window_pad, self.window_pad + self.WIDTH]
        self.boundaries_y = [self.window_pad, self.window_pad + self.HEIGHT]

    # agent.py:645-693
    def tr_w(self):
        return self.w > self.T_w

    def tr_u(self):
        return self.u > self.T_w

    def get_mode(self):
        if self.overriding_mode is None:
            return "relocate" if self.tr_w() else "explore"
        return self.overriding_mode

    def set_mode(self, mode):
        if mode in ("explore", "relocate"):
            self.overriding_mode = None
        elif mode in ("collide", "exploit", "pool"):
            self.overriding_mode = mode
        self.mode = mode

    def projection_field(self, obstacles, keep_distance_info=False, non_expl_agents=None, fov=None):
        """Agent.projection_field (agent.py:457-597) on the GPU."""
        from . import forage
        return forage.projection_field(self, obstacles, keep_distance_info, non_expl_agents, fov)

    def calc_social_V_proj(self, agents):
        """Cue selection (agent.py:396-419) and the projection on the GPU; used
        by the paused branch of the host loop (sims.py:867-870)."""
        def dist(a, b):
            c1 = np.array([a.position[0] + a.radius, a.position[1] + a.radius])
            c2 = np.array([b.position[0] + b.radius, b.position[1] + b.radius])
            return np.linalg.norm(c2 - c1)
        inr = [ag for ag in agents if dist(self, ag) <= self.vision_range]
        expl = [ag for ag in inr if ag.id != self.id and ag.get_mode() == "exploit"]
        non_expl = [ag for ag in inr if ag not in expl]
        if self.exclude_agents_same_patch:
            non_expl.extend([ag for ag in expl if ag.exploited_patch_id == self.exploited_patch_id])
            expl = [ag for ag in expl if ag.exploited_patch_id != self.exploited_patch_id]
        expl = [ag for ag in expl if ag.exploited_patch_id != -1]
        if self.visual_exclusion:
            self.soc_v_field = self.projection_field(expl, False, non_expl_agents=non_expl)
        else:
            self.soc_v_field = self.projection_field(expl, False)

    def update(self, agents):
        raise L.AbmError(L.ABM_ERR_UNSUPPORTED,
                         "agents are updated together on the GPU: call AgentGroup.update(agents)")


class VFAgent(Agent):
    """Visual-flocking agent (vf_agent.py:12-50) with the reference's
    attributes; ``vf_params`` plays the role of the vf_params module."""

    def __init__(self, *args, vf_params: VFParams = None, **kw):
        super().__init__(*args, **kw)
        p = vf_params or VFParams()
        self.vf_params = p
        self.agent_state = "flocking"
        self.ALP0 = None
        self.BET0 = None
        self.V0 = None
        self.lines = []
        self.verbose_supcalc = False
        self.boundary_cond = p.BOUNDARY
        self.limit_movement = p.LIMIT_MOVEMENT
        self.max_vel = p.MAX_VEL
        self.max_th = p.MAX_TH
        self.PHI = np.arange(-np.pi, np.pi, (2 * np.pi) / self.v_field_res)
        self.soc_v_field_proj = np.zeros(self.v_field_res)
        self.coll_times = []

    def calc_soc_v_proj(self, agents):
        """vf_agent.py:216-237 on the GPU."""
        from . import vf
        others = [ag for ag in agents if ag.id != self.id]
        field = vf.projection_field(
            fov=self.FOV, v_field_resolution=self.v_field_res, position=np.array(self.position),
            radius=self.radius, orientation=self.orientation,
            object_positions=[np.array(ag.position) for ag in others],
            object_sizes=[ag.radius for ag in others], boundary_cond=self.boundary_cond,
            arena_width=self.WIDTH, arena_height=self.HEIGHT, ag_id=self.id)
        s = field.sum(axis=0)
        s[s > 0] = 1
        return s

    def update_social_info(self, agents):
        """vf_agent.py:206-214 (the paused branch, vf_sims.py:306-310)."""
        self.soc_v_field = self.calc_soc_v_proj(agents)


# ------------------------------------------------------------------ group
class AgentGroup:
    """pygame Group stand-in: insertion-ordered agents; ``update(agents)`` is
    one GPU step of all of them in that order (Group.update semantics).

    ``precision``: "f64" (reference arithmetic) or "f32" (fused kernel) for
    visual-flocking agents; foraging agents always step in float64."""

    def __init__(self, *agents, precision="f64"):
        self._agents: List[Agent] = []
        self.precision = precision
        self._engine = None
        self._key = None
        self.add(*agents)

    # container protocol used by the reference's loops
    def add(self, *agents):
        for a in agents:
            if isinstance(a, (list, tuple)):
                self.add(*a)
            elif a not in self._agents:
                self._agents.append(a)
        self._key = None

    def remove(self, *agents):
        for a in agents:
            if a in self._agents:
                self._agents.remove(a)
        self._key = None

    def sprites(self):
        return list(self._agents)

    def __iter__(self):
        return iter(list(self._agents))

    def __len__(self):
        return len(self._agents)

    def __contains__(self, a):
        return a in self._agents

    def __bool__(self):
        return bool(self._agents)

    # one step
    def update(self, agents=None):
        if not self._agents:
            return
        if all(isinstance(a, VFAgent) for a in self._agents):
            self._update_vf()
        elif not any(isinstance(a, VFAgent) for a in self._agents):
            self._update_forage()
        else:
            raise L.AbmError(L.ABM_ERR_INVALID, "a group must hold only VFAgent or only Agent objects")

    def _update_vf(self):
        from .vf import VFConfig, VFSwarm
        ags = self._agents
        a0 = ags[0]
        p = a0.vf_params
        key = ("vf", len(ags), a0.v_field_res, a0.WIDTH, a0.HEIGHT, a0.window_pad, tuple(a0.FOV),
               a0.boundary_cond, a0.limit_movement, a0.max_vel, a0.max_th,
               (p.GAM, p.V0, p.ALP0, p.ALP1, p.BET0, p.BET1), self.precision)
        x = np.array([float(a.position[0]) for a in ags])
        y = np.array([float(a.position[1]) for a in ags])
        th = np.array([float(a.orientation) for a in ags])
        v = np.array([float(a.velocity) for a in ags])
        rad = np.array([int(a.radius) for a in ags], dtype=np.int32)
        ov = lambda name, dflt: (None if all(getattr(a, name) is None for a in ags) else
                                 np.array([dflt if getattr(a, name) is None else getattr(a, name)
                                           for a in ags], dtype=np.float64))
        alp0, bet0, v0 = ov("ALP0", p.ALP0), ov("BET0", p.BET0), ov("V0", p.V0)
        verbose = any(a.verbose_supcalc for a in ags)
        if key != self._key or self._engine is None:
            cfg = VFConfig(n_agents=len(ags), n_rays=a0.v_field_res, width=a0.WIDTH, height=a0.HEIGHT,
                           window_pad=a0.window_pad, boundary=a0.boundary_cond, fov=tuple(a0.FOV),
                           GAM=p.GAM, V0=p.V0, ALP0=p.ALP0, ALP1=p.ALP1, BET0=p.BET0, BET1=p.BET1,
                           LIMIT_MOVEMENT=a0.limit_movement, MAX_VEL=a0.max_vel, MAX_TH=a0.max_th)
            self._engine = VFSwarm([cfg], x, y, th, v, radius=rad, precision=self.precision,
                                   alp0=alp0, bet0=bet0, v0=v0)
            self._key = key
        else:
            self._engine.load(x=x, y=y, theta=th, v=v, radius=rad, alp0=alp0, bet0=bet0, v0=v0)
        eng = self._engine
        eng.step(fields=True, diag=verbose)
        nx, ny, nth, nv = eng.state_numpy()
        bits = eng.fields_numpy(0)
        diag = eng.diag_numpy() if verbose else None
        for i, a in enumerate(ags):
            a.position = [float(nx[i]), float(ny[i])]          # a list after a move (vf_agent.py:309)
            a.orientation = float(nth[i])
            a.velocity = float(nv[i])
            a.soc_v_field = np.flip(bits[i]).astype(np.float64)
            if a.verbose_supcalc:
                a.dv, a.dphi, a.ablob, a.aedge, a.bblob, a.bedge = (float(q) for q in diag[i])

    def _update_forage(self):
        from .forage import ForageConfig, ForageSwarm
        ags = self._agents
        a0 = ags[0]
        mp = a0.movement_params
        key = ("forage", len(ags), a0.v_field_res, a0.WIDTH, a0.HEIGHT, a0.window_pad, tuple(a0.FOV),
               a0.vision_range, a0.visual_exclusion, a0.exclude_agents_same_patch, len(a0.novelty),
               tuple((a.T_w, a.Eps_w, a.g_w, a.B_w, a.w_max, a.T_u, a.Eps_u, a.g_u, a.B_u, a.u_max,
                      a.S_wu, a.S_uw, a.F_N, a.F_R, a.max_exp_vel, a.exp_stop_ratio) for a in ags),
               (mp.exp_theta_min, mp.exp_theta_max, mp.reloc_theta_max))
        if len(set(key[-2])) != 1:
            raise L.AbmError(L.ABM_ERR_UNSUPPORTED,
                             "agents of one group must share their decision parameters")
        if key != self._key or self._engine is None:
            cfg = ForageConfig(n_agents=len(ags), n_rays=a0.v_field_res, width=a0.WIDTH,
                               height=a0.HEIGHT, window_pad=a0.window_pad, fov=tuple(a0.FOV),
                               vision_range=a0.vision_range, visual_exclusion=a0.visual_exclusion,
                               patchwise_exclusion=a0.exclude_agents_same_patch,
                               T_w=a0.T_w, Eps_w=a0.Eps_w, g_w=a0.g_w, B_w=a0.B_w, w_max=a0.w_max,
                               T_u=a0.T_u, Eps_u=a0.Eps_u, g_u=a0.g_u, B_u=a0.B_u, u_max=a0.u_max,
                               S_wu=a0.S_wu, S_uw=a0.S_uw, Tau=len(a0.novelty), F_N=a0.F_N,
                               F_R=a0.F_R, exp_vel_max=a0.max_exp_vel,
                               exp_theta_min=mp.exp_theta_min, exp_theta_max=mp.exp_theta_max,
                               reloc_theta_max=mp.reloc_theta_max, exp_stop_ratio=a0.exp_stop_ratio)
            self._engine = ForageSwarm([cfg], [a.position[0] for a in ags],
                                       [a.position[1] for a in ags], [a.orientation for a in ags],
                                       radius=[a.radius for a in ags])
            self._key = key
        eng = self._engine
        modes = L.MODES
        eng.set_state(
            x=[a.position[0] for a in ags], y=[a.position[1] for a in ags],
            theta=[a.orientation for a in ags], v=[a.velocity for a in ags],
            w=[a.w for a in ags], u=[a.u for a in ags], I_priv=[a.I_priv for a in ags],
            collected_r=[a.collected_r for a in ags],
            collected_r_before=[a.collected_r_before for a in ags],
            novelty=np.concatenate([np.asarray(a.novelty, dtype=np.float64) for a in ags]),
            mode=np.array([modes.index(a.mode) for a in ags], dtype=np.int32),
            overriding_mode=np.array([-1 if a.overriding_mode is None else modes.index(a.overriding_mode)
                                      for a in ags], dtype=np.int32),
            env_status=np.array([a.env_status for a in ags], dtype=np.int32),
            exploited_patch_id=np.array([a.exploited_patch_id for a in ags], dtype=np.int32),
            radius=np.array([a.radius for a in ags], dtype=np.int32))
        eng.step(fields=True)
        st = eng.get_state()
        bits = eng.fields_numpy(0)
        for i, a in enumerate(ags):
            a.position = np.array([st["x"][i], st["y"][i]])
            a.orientation = float(st["theta"][i])
            a.velocity = float(st["v"][i])
            a.w, a.u, a.I_priv = float(st["w"][i]), float(st["u"][i]), float(st["I_priv"][i])
            a.collected_r_before = float(st["collected_r_before"][i])
            a.mode = modes[int(st["mode"][i])]
            o = int(st["overriding_mode"][i])
            a.overriding_mode = None if o < 0 else modes[o]
            a.soc_v_field = bits[i].astype(np.float64)


# ------------------------------------------------------- headless loops
def _notify_agent(agent, status, res_id=None):
    """sims.py:29-42."""
    agent.env_status_before = agent.env_status
    agent.env_status = status
    agent.novelty = np.roll(agent.novelty, 1)
    novelty = agent.env_status - agent.env_status_before
    agent.novelty[0] = 1 if novelty > 0 else 0
    agent.pool_success = 1
    agent.exploited_patch_id = -1 if res_id is None else res_id


class Rescource:
    """Resource patch (abm/environment/rescource.py:15-56, 118-133), no drawing."""

    def __init__(self, id, radius, position, env_size, color=None, window_pad=30, resc_units=None,
                 quality=1):
        self.resc_units = radius if resc_units is None else resc_units
        self.id = id
        self.radius = radius
        self.resc_left = self.resc_units
        self.position = np.array(position, dtype=np.float64)
        self.center = (self.position[0] + self.radius, self.position[1] + self.radius)
        self.unit_per_timestep = quality
        self.WIDTH, self.HEIGHT = env_size
        self.window_pad = window_pad

    def deplete(self, rescource_units):
        if rescource_units > self.unit_per_timestep:
            rescource_units = self.unit_per_timestep
        if self.resc_left >= rescource_units:
            self.resc_left -= rescource_units
            depleted = rescource_units
        else:
            depleted = self.resc_left
            self.resc_left = 0
        return depleted, not (self.resc_left > 0)


class VFSimulation:
    """Headless visual-flocking loop (vf_sims.py:16-49, 212-228, 291-366)."""

    def __init__(self, N, T, v_field_res=800, width=600, height=480, window_pad=30,
                 agent_radius=10, agent_fov=1.0, vf_params: VFParams = None, precision="f64",
                 **_ignored):
        self.N, self.T, self.t = N, T, 0
        self.WIDTH, self.HEIGHT, self.window_pad = width, height, window_pad
        self.agent_radii = agent_radius
        self.fov_ratio = agent_fov
        self.agent_fov = (-self.fov_ratio * np.pi, self.fov_ratio * np.pi)
        self.v_field_res = int(v_field_res * (1 / self.fov_ratio))      # vf_sims.py:42-43
        self.vf_params = vf_params or VFParams()
        self.agents = AgentGroup(precision=precision)
        self.is_paused = False

    def add_new_agent(self, id, x, y, orient):
        self.agents.add(VFAgent(id=id, radius=self.agent_radii, position=(x, y), orientation=orient,
                                env_size=(self.WIDTH, self.HEIGHT), v_field_res=self.v_field_res,
                                FOV=self.agent_fov, window_pad=self.window_pad,
                                vf_params=self.vf_params))

    def create_agents(self):
        """Agents on a disc around the arena centre (vf_sims.py:212-228)."""
        for i in range(self.N):
            orient = np.random.uniform(0, 2 * np.pi)
            dist = np.random.rand() * (self.HEIGHT / 2 - 2 * self.window_pad - self.agent_radii)
            x = np.cos(orient) * dist + self.WIDTH / 2
            y = np.sin(orient) * dist + self.HEIGHT / 2
            self.add_new_agent(i, x, y, orient)

    def step_sim(self):
        if not self.is_paused:
            self.agents.update(self.agents)
            self.t += 1
        else:
            for ag in self.agents:
                ag.update_social_info(self.agents)

    def start(self):
        self.create_agents()
        while self.t < self.T:
            self.step_sim()


class Simulation:
    """Headless foraging loop (sims.py:60-221 parameters, 321-374 resources,
    526-552 creation, 702-870 step), agent-agent collisions included
    (sims.py:421-467, 736-787)."""

    def __init__(self, N, T, v_field_res=800, width=600, height=480, framerate=25, window_pad=30,
                 with_visualization=False, show_vis_field=False, show_vis_field_return=False,
                 pooling_time=3, pooling_prob=0.05, agent_radius=10, N_resc=10,
                 min_resc_perpatch=200, max_resc_perpatch=1000, min_resc_quality=0.1,
                 max_resc_quality=1, patch_radius=30, regenerate_patches=True, agent_consumption=1,
                 teleport_exploit=True, vision_range=150, agent_fov=1.0, visual_exclusion=False,
                 show_vision_range=False, use_ifdb_logging=False, use_ram_logging=False,
                 save_csv_files=False, ghost_mode=True, patchwise_exclusion=True, parallel=False,
                 use_zarr=True, allow_border_patch_overlap=False, agent_behave_param_list=None,
                 collide_agents=True, decision_params: DecisionParams = None,
                 movement_params: MovementParams = None):
        self.collide_agents = collide_agents
        self.N, self.T, self.t = N, T, 0
        self.WIDTH, self.HEIGHT, self.window_pad = width, height, window_pad
        self.allow_border_patch_overlap = allow_border_patch_overlap
        self.agent_behave_param_list = agent_behave_param_list
        self.agent_radii = agent_radius
        self.v_field_res = v_field_res
        self.pooling_time, self.pooling_prob = pooling_time, pooling_prob
        self.agent_consumption = agent_consumption
        self.teleport_exploit = teleport_exploit
        self.vision_range = vision_range
        self.fov_ratio = agent_fov
        self.agent_fov = (-self.fov_ratio * np.pi, self.fov_ratio * np.pi)
        self.visual_exclusion = visual_exclusion
        self.ghost_mode = ghost_mode
        self.patchwise_exclusion = patchwise_exclusion
        self.N_resc = N_resc
        self.resc_radius = patch_radius
        self.min_resc_units, self.max_resc_units = min_resc_perpatch, max_resc_perpatch
        self.min_resc_quality, self.max_resc_quality = min_resc_quality, max_resc_quality
        if self.max_resc_quality < 0:
            self.max_resc_quality = self.min_resc_quality
        if self.max_resc_units < 0:
            self.max_resc_units = self.min_resc_units + 1
        self.regenerate_resources = regenerate_patches
        self.decision_params = decision_params or DecisionParams()
        self.movement_params = movement_params or MovementParams()
        self.agents = AgentGroup()
        self.rescources: List[Rescource] = []
        self.is_paused = False

    # creation (sims.py:321-374, 526-542)
    def add_new_agent(self, id, x, y, orient, behave_params=None):
        self.agents.add(Agent(id=id, radius=self.agent_radii, position=(x, y), orientation=orient,
                              env_size=(self.WIDTH, self.HEIGHT), v_field_res=self.v_field_res,
                              FOV=self.agent_fov, window_pad=self.window_pad,
                              pooling_time=self.pooling_time, pooling_prob=self.pooling_prob,
                              consumption=self.agent_consumption, vision_range=self.vision_range,
                              visual_exclusion=self.visual_exclusion,
                              patchwise_exclusion=self.patchwise_exclusion,
                              behave_params=behave_params, decision_params=self.decision_params,
                              movement_params=self.movement_params))

    def create_agents(self):
        for i in range(self.N):
            x = np.random.randint(self.window_pad - self.agent_radii, self.WIDTH + self.window_pad - self.agent_radii)
            y = np.random.randint(self.window_pad - self.agent_radii, self.HEIGHT + self.window_pad - self.agent_radii)
            orient = np.random.uniform(0, 2 * np.pi)
            bp = None if self.agent_behave_param_list is None else self.agent_behave_param_list[i]
            self.add_new_agent(i, x, y, orient, bp)

    def _collides(self, a, b):
        # pygame.sprite.collide_circle on the patches' integer rect centres
        dx = int(a.center[0]) - int(b.center[0])
        dy = int(a.center[1]) - int(b.center[1])
        return dx * dx + dy * dy <= (a.radius + b.radius) ** 2

    def add_new_resource_patch(self, force_id=None):
        max_retries = 10000
        if force_id is None:
            pid = max([r.id for r in self.rescources]) if self.rescources else 0
        else:
            pid = force_id
        retries = 0
        while True:
            if retries > max_retries:
                raise Exception("Reached timeout while trying to create resources without overlap!")
            r = self.resc_radius
            if self.allow_border_patch_overlap:
                x = np.random.randint(self.window_pad - r, self.WIDTH + self.window_pad - r)
                y = np.random.randint(self.window_pad - r, self.HEIGHT + self.window_pad - r)
            else:
                x = np.random.randint(self.window_pad, self.WIDTH + self.window_pad - 2 * r)
                y = np.random.randint(self.window_pad, self.HEIGHT + self.window_pad - 2 * r)
            units = np.random.randint(self.min_resc_units, self.max_resc_units)
            quality = np.random.uniform(self.min_resc_quality, self.max_resc_quality)
            res = Rescource(pid + 1 if force_id is None else pid, r, (x, y), (self.WIDTH, self.HEIGHT),
                            window_pad=self.window_pad, resc_units=units, quality=quality)
            retries += 1
            if not any(self._collides(o, res) for o in self.rescources):
                break
        self.rescources.append(res)
        return res.id

    def create_resources(self):
        for _ in range(self.N_resc):
            self.add_new_resource_patch()

    def kill_resource(self, resource):
        rid = resource.id
        self.rescources.remove(resource)
        if self.regenerate_resources:
            self.add_new_resource_patch(force_id=rid)

    @staticmethod
    def _distance(a, b):
        c1 = np.array([a.position[0] + a.radius, a.position[1] + a.radius])
        c2 = np.array([b.position[0] + b.radius, b.position[1] + b.radius])
        return np.linalg.norm(c2 - c1)

    def bias_agent_towards_res_center(self, agent, resc, relative_speed=0.02):
        x1, y1 = agent.position + agent.radius
        x2, y2 = resc.center
        cl_ang = (math.atan2(y2 - y1, x2 - x1) + agent.orientation) % (np.pi * 2)
        agent.orientation += (cl_ang - np.pi) * relative_speed

    def agent_agent_collision_proximity(self, agent1, agent2):
        """Collision response from the proximity field (sims.py:421-467); the
        field on the GPU (Agent.projection_field with keep_distance_info)."""
        agents2 = agent2 if isinstance(agent2, list) else [agent2]
        for agent2 in agents2:
            do_collision = True
            if self.ghost_mode:
                do_collision = agent2.get_mode() != "exploit" and agent1.get_mode() != "exploit"
            if not do_collision:
                continue
            if agent2.get_mode() != "exploit":
                agent2.set_mode("collide")
            near = [a for a in self.agents
                    if self._distance(a, agent2) < 2 * agent2.radius + 20 and a is not agent2]
            field = agent2.projection_field(near, keep_distance_info=True, fov=(-np.pi, np.pi))
            n = len(field)
            d_lr = np.sign(np.mean(field[0:int(n / 2)]) - np.mean(field[int(n / 2)::]))
            if d_lr == 0:
                d_lr = -1
            if agent2.get_mode() != "exploit":
                agent2.orientation -= d_lr * 0.2
            if np.mean(field[int(n / 2) - 100:int(n / 2) + 100]) > 0:
                agent2.velocity = 0
            elif agent2.get_mode() != "exploit":
                agent2.velocity = agent2.max_exp_vel

    def collide(self):
        """Agent-agent collisions (sims.py:736-787): groupcollide of the agents
        with themselves, collide_circle on the rect centres (int(position) +
        radius, pygame's Rect holding int() of the position) with every radius
        grown by 2; the responses in group order; collide modes ended or
        notified."""
        agents = self.agents.sprites()
        cent = [(int(a.position[0]) + a.radius, int(a.position[1]) + a.radius) for a in agents]
        groups = []
        for i, a in enumerate(agents):
            hits = [b for j, b in enumerate(agents) if j != i and
                    (cent[i][0] - cent[j][0]) ** 2 + (cent[i][1] - cent[j][1]) ** 2
                    <= (a.radius + 2 + b.radius + 2) ** 2]
            if hits:
                groups.append((a, hits))
        collided = []
        for agent1, agents2 in groups:
            self.agent_agent_collision_proximity(agent1, agents2)
            for agent2 in agents2:
                if self.teleport_exploit:
                    if agent1.get_mode() != "exploit":
                        collided.append(agent1)
                    if agent2.get_mode() != "exploit":
                        collided.append(agent2)
                elif not self.ghost_mode:
                    collided += [agent1, agent2]
                elif agent1.get_mode() != "exploit" and agent2.get_mode() != "exploit":
                    collided += [agent1, agent2]
        for a in agents:
            hit = any(a is c for c in collided)
            if not hit and a.get_mode() == "collide":
                a.set_mode("explore")
            if hit and a.get_mode() == "collide":
                _notify_agent(a, -1)

    def step(self):
        """One step of the loop body (sims.py:736-870)."""
        if self.is_paused:
            for ag in self.agents:
                ag.calc_social_V_proj(self.agents)
            return
        if self.collide_agents:
            self.collide()
        agents = self.agents.sprites()
        groups = [(r, [a for a in agents if self._distance(r, a) < r.radius]) for r in self.rescources]
        on_patch = []
        for resc, here in groups:
            if not here:
                continue
            destroy = False
            for agent in here:
                self.bias_agent_towards_res_center(agent, resc)
                if destroy:
                    _notify_agent(agent, -1)
                else:
                    if (agent.get_mode() in ["pool", "relocate"] and agent.pool_success) or agent.pooling_time == 0:
                        _notify_agent(agent, 1, resc.id)
                        if self.teleport_exploit:
                            agent.position = resc.position + resc.radius - agent.radius
                    if agent.get_mode() == "exploit":
                        dep, destroy = resc.deplete(agent.consumption)
                        agent.collected_r_before = agent.collected_r
                        agent.collected_r += dep
                        if destroy:
                            for other in here:
                                _notify_agent(other, -1)
                on_patch.append(agent)
            if destroy:
                self.kill_resource(resc)
        for agent in agents:
            if agent not in on_patch:
                if (agent.get_mode() in ["pool", "relocate"] and agent.pool_success) or agent.pooling_time == 0:
                    _notify_agent(agent, -1)
                elif agent.get_mode() == "exploit":
                    _notify_agent(agent, -1)
        self.agents.update(self.agents)
        self.t += 1

    def start(self):
        self.create_agents()
        self.create_resources()
        while self.t < self.T:
            self.step()
