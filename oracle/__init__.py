"""CPU oracle for the frequency-domain inverse solve.

TEST INFRASTRUCTURE ONLY.  This package is a numpy/scipy restatement of the
reference's algorithms (LuisCorralR/InverseAcoustics: ``CUDA/*.cuh`` and the
``MATLAB/*.m`` specifications), written to check the HIP implementation in
``inverseacoustics_amd``.  Only ``tests/``, ``__graft_entry__.smoke()`` and the
``cpu_baseline`` leg of ``bench.py`` may import it, and only as the checker or
the timed CPU baseline; the product never routes through it.

Every function cites the reference file:line it restates.  How each one is
pinned (closed forms, scipy/numpy library routines, brute force, invariants)
is listed in the module docstrings and exercised in ``tests/test_oracle*.py``.
Functions whose exact numerics depend on closed-source code (MATLAB's
``fminbnd``/``svd``/``\\``, cuSOLVER phases, GSL last-ulp behaviour) are pinned
only through invariants; see each docstring.
"""
