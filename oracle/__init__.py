"""CPU FP64 oracle: a NumPy restatement of the reference BSSN evolution path.

TEST INFRASTRUCTURE ONLY.  Only ``tests/``, ``__graft_entry__.smoke()`` and
the ``cpu_baseline`` leg of ``bench.py`` may import this package, and only as
the checker or the timed CPU baseline; the product (``jax_bssn_amd``) never
imports, links or executes anything from it.
"""
