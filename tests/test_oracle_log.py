"""Pins of the agent-record oracle (oracle/log_oracle.py): hand-worked edge
lists, names and key order."""
import numpy as np

from oracle import log_oracle as lo


class A:
    def __init__(self, i, pos, f, mode="explore"):
        self.id, self.position, self.orientation, self.velocity = i, np.array(pos), 1.25, 0.5
        self.w = self.u = self.I_priv = self.collected_r = 0
        self.mode, self.exploited_patch_id, self.soc_v_field = mode, -1, np.array(f, dtype=float)


def test_edges_and_names():
    d = lo.save_agent_data_RAM({}, [A(3, [10.7, -2.5], [0, 1, 1, 0, 0, 1])])
    r = d[3]
    # roll(f, 1) = [1, 0, 1, 1, 0, 0]: rises at k = 1, 5 and falls at k = 0, 3
    assert r["vfield_up"] == ["[1 5]"] and r["vfield_down"] == ["[0 3]"]
    assert r["agent_name"] == "agent-03" and r["posx"] == [10] and r["posy"] == [-2]
    assert list(r) == ["agent_name", "posx", "posy", "orientation", "velocity", "w", "u", "Ipriv",
                       "mode", "collectedr", "expl_patch_id", "vfield_up", "vfield_down"]
    # an all-zero field has no edges; a full field neither
    d = lo.save_agent_data_RAM(d, [A(3, [0, 0], [0] * 6)])
    d = lo.save_agent_data_RAM(d, [A(3, [0, 0], [1] * 6, mode="exploit")])
    assert d[3]["vfield_up"][1:] == ["[]", "[]"] and d[3]["mode"] == [0, 0, 1]
    # numpy pads multi-digit entries to one width
    f = np.zeros(120)
    f[[5, 100]] = 1
    e = lo.save_agent_data_RAM({}, [A(12, [0, 0], f)])[12]["vfield_up"][0]
    assert e == "[  5 100]"
    assert lo.pad_to_n_digits(123, n=2) == "123"
