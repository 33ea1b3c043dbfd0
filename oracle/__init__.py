"""CPU oracle for the magnification hot path — TEST INFRASTRUCTURE ONLY.

Only ``tests/``, ``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of
``bench.py`` may import this package, and only as the checker: nothing in
``caustics_amd`` imports it, and no product path falls back to it.

Contents
  ea.py        ctypes wrapper of ``libea_oracle.so`` (plain-C FP64 restatement of
               the reference's Ehrlich-Aberth solver, ``ea_oracle.c``).
  ref_op.py    loader for the reference's own CPU op built from its sources
               into ``_ref/`` (present only where it was built).
  lens.py      NumPy FP64 restatement of coefficients, lens equation, image
               filter, point-source magnification and its gradient.
  multipole.py NumPy restatement of the hexadecapole and the gate.
  extended.py  NumPy restatement of the limb sampling / stitching / contour
               integration pipeline.
  linalg.py    NumPy restatement of the flux-marginalised likelihood.
  trajectory.py NumPy restatement of the annual-parallax trajectory.
  mp_lens.py   the point-source route evaluated in 40 significant digits, the
               truth for clustered roots that FP64 cannot resolve.
"""
