"""CPU oracle for the MAC shifted-Stokes stage solve -- TEST INFRASTRUCTURE ONLY.

This package restates, in NumPy/SciPy, the reference's DST/DCT + pressure Schur-CG
backend (`solver/mac/tensor_stokes.py:118-364`), its sparse-direct comparator
(`solver/mac/stokes.py:42-86`), the MAC operators (`solver/mac/operators.py:6-29`)
and the matrix-free kernels (`solver/mac/kernels.py:6-132`).  It is the checker:
only `tests/`, `__graft_entry__.smoke()` and the `cpu_baseline` leg of `bench.py`
may import it.  The product package never imports, links or calls anything here,
and never falls back to it.

Parity status of each function is written in its docstring ("pinned by ...").
"""
