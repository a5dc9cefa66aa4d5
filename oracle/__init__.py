"""CPU FP64 oracle of the pressure projection -- test infrastructure only.

Only ``tests/``, ``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of
``bench.py`` may import this package.  The engine never calls it.
"""
