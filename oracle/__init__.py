"""CPU oracle for the compact-difference operator path.

TEST INFRASTRUCTURE ONLY: imported by ``tests/``, ``__graft_entry__.smoke()``
and the CPU-baseline leg of ``bench.py``, always as the checker and never as
the thing measured or shipped.  ``pade_dense`` is the dense LAPACK
restatement, ``banded`` the C restatement of the reference's banded sweeps.
"""
