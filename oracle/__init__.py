"""CPU oracle for the kspaceFirstOrder2D/3D step loop.

TEST INFRASTRUCTURE ONLY.  Nothing in the product (the ``kwhip`` package and
its HIP library) imports, links or calls anything under ``oracle/``.  Only
``tests/``, ``__graft_entry__.smoke()`` and ``bench.py``'s CPU-baseline leg use
it, and only as the checker / the timed CPU baseline.
"""
