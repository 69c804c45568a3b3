"""Oracle: per-step agent records, as the reference logs them in RAM.

TEST INFRASTRUCTURE ONLY -- the checker for ``abm_amd.agentlog``; never
imported by the product path.

``save_agent_data_RAM`` restates abm/monitoring/ifdb.py:55-96 for a sequence
of agent-like records (attributes id, position, orientation, velocity, w, u,
I_priv, mode, collected_r, exploited_patch_id, soc_v_field), and
``pad_to_n_digits`` / ``mode_to_int`` restate ifdb.py:38-52 and :197-206.

Parity pins (tests/test_oracle_log.py): hand-worked edge lists of small
fields (including the wrap-around through np.roll), the padded names and the
key order of the dict.
"""
from __future__ import annotations

import numpy as np


def pad_to_n_digits(number, n=3):
    s = str(number)
    return (n - len(s)) * "0" + s if n > len(s) else s


def mode_to_int(mode):
    return {"explore": 0, "exploit": 1, "relocate": 2, "collide": 3}.get(mode)


def save_agent_data_RAM(agents_dict, agents):
    """Append one step of every agent to agents_dict (a dict of dicts of lists)."""
    for a in agents:
        if a.id not in agents_dict:
            agents_dict[a.id] = {"agent_name": f"agent-{pad_to_n_digits(a.id, n=2)}"}
            for k in ("posx", "posy", "orientation", "velocity", "w", "u", "Ipriv", "mode",
                      "collectedr", "expl_patch_id", "vfield_up", "vfield_down"):
                agents_dict[a.id][k] = []
        d = agents_dict[a.id]
        d["posx"].append(int(a.position[0]))
        d["posy"].append(int(a.position[1]))
        d["orientation"].append(float(a.orientation))
        d["velocity"].append(float(a.velocity))
        d["w"].append(float(a.w))
        d["u"].append(float(a.u))
        d["Ipriv"].append(float(a.I_priv))
        d["mode"].append(int(a.mode if isinstance(a.mode, (int, np.integer)) else mode_to_int(a.mode)))
        d["collectedr"].append(float(a.collected_r))
        d["expl_patch_id"].append(int(a.exploited_patch_id))
        f = a.soc_v_field
        d["vfield_up"].append(f"{np.where(np.roll(f, 1) < f)[0]}")
        d["vfield_down"].append(f"{np.where(np.roll(f, 1) > f)[0]}")
    return agents_dict
