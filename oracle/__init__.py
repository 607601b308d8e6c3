"""FP64 CPU oracle for the HOPS propagation path.

TEST INFRASTRUCTURE ONLY.  This package restates the reference MesoHOPS
algorithm (system dictionary derivation, triangular hierarchy, hierarchy
super-operators, LINEAR / NONLINEAR / NORMALIZED NONLINEAR derivative with the
low-temperature correction, classical RK4, FFT-filtered noise, the adaptive
basis, and dyadic trajectories with their absorption / fluorescence pathways)
in NumPy/SciPy so that the HIP engine in ``mesohops_amd`` can be checked
against it.  Only ``tests/``, ``__graft_entry__.smoke()`` and the
``cpu_baseline`` leg of ``bench.py`` may import it, and only as the checker or
as the timed CPU baseline; the product path never calls it.

Every function cites the reference file:line it follows.  The pins that make
each part trustworthy are listed in the header of ``oracle/hops_oracle.py``
and tested in ``tests/test_oracle_pins.py`` and ``tests/test_dyadic_host.py``.
"""
