"""CPU oracle for the Fourier-propagation path (test infrastructure only).

This package is the checker, never the product: only ``tests/``, ``__graft_entry__.smoke()`` and
the ``cpu_baseline`` leg of ``bench.py`` may import it.  The product path (``hcipy_amd``) never
routes through it.

``oracle.hcipy_oracle`` restates, in numpy/scipy FP64 (and FP32 where the reference computes in
complex64), the algorithm of the reference HCIPy files cited in each function's docstring.
How each function is pinned (golden vectors, closed forms, brute force, invariants) is listed
in ``oracle/hcipy_oracle.py``'s module docstring.
"""
