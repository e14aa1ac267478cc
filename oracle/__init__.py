"""ORACLE -- test infrastructure, not part of the product.

CPU restatement of trench01's per-batch charging path (see each module's
header for the reference lines it follows).  Only ``tests/``,
``__graft_entry__.smoke()`` and ``bench.py``'s ``cpu_baseline`` leg may import
it, and only as the checker; the ``trench01_amd`` package never does.
"""
