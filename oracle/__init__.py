"""CPU oracle for the ABM perception/motion hot path.

TEST INFRASTRUCTURE ONLY.  Nothing under ``oracle/`` is part of the product:
only ``tests/``, ``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of
``bench.py`` may import it, and only as the checker (or as the timed CPU
baseline).  The product path (``abm_amd``) never imports this package and
fails loudly when its HIP library is missing.

The modules restate, step by step in numpy float64, the reference's
per-step perception and motion code (scioip34/ABM):

* ``vf_oracle``      -- visual flocking (VSWRM): ``vf_supcalc.projection_field``,
  ``calculate_closed_angle``, ``dPhi_V_of``, ``VSWRM_flocking_state_variables``,
  ``VFAgent`` pose update and boundary handling, and the in-place
  (Gauss-Seidel) group update.
* ``forage_oracle``  -- collective foraging: ``Agent.projection_field`` with
  visual exclusion, ``calc_social_V_proj``, the w/u decision processes, mode
  selection, random walk / relocation and wall reflection.

Each function cites the reference file:line it follows.
"""
