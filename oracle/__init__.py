"""CPU oracle for the ANI energy/force/MD path (test infrastructure only).

This package is the checker.  It restates, in NumPy (FP64 unless stated),
what FeNNol computes on the path neighbour list -> angular triplets -> ANI AEV
-> species MLP -> energy sum -> forces / virial -> BAOAB Langevin step.  Only
``tests/``, ``__graft_entry__.smoke()`` and ``bench.py``'s ``cpu_baseline`` leg
may import it, and only to check or time; the product path (``fennol_amd``)
never calls into it and fails loudly when its HIP library is missing.

Each function cites the reference file:line it follows.  How each one is
pinned (closed forms, brute force, finite differences, invariances) is listed
in the module docstring of :mod:`oracle.ani_oracle`.
"""

from .ani_oracle import *  # noqa: F401,F403
