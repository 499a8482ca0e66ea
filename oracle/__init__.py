"""CPU oracle for the LAPW H+O setup — test infrastructure only.

Imported by ``tests/``, ``__graft_entry__.smoke`` and the CPU-baseline leg of
``bench.py``, and nowhere else: the library (``exciting_amd``) never calls it.
``specfun`` restates the special functions, ``lapw_ref`` the setup, the
eigensolve, the IBS force and the setup's vector-Jacobian product.
"""
