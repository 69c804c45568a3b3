"""abm_amd -- MI355X (gfx950) perception and motion step for scioip34/ABM agents.

The per-step hot path of the reference (visual projection fields, VSWRM
integrals, Euler pose update with boundary handling, foraging decisions) runs
in hand-written HIP kernels behind the C ABI of ``include/abm_amd.h``; this
package is the Python host side that mirrors the reference's interfaces.

* ``abm_amd.vf``      -- visual flocking: ``VFSwarm`` (batched replicas) and
  ``projection_field`` / ``VSWRM_flocking_state_variables`` / ``dPhi_V_of``.
* ``abm_amd.forage``  -- collective foraging step.
* ``abm_amd.dropin``  -- ``VFAgent``/``Agent``-compatible classes and a
  Group whose ``update(agents)`` runs one GPU step.
* ``abm_amd.build``   -- in-tree build of ``libabm_hip.so``.
"""
from ._lib import AbmError, AbmUnsupported, load  # noqa: F401

__version__ = "0.1.0"
