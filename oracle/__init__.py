"""TEST INFRASTRUCTURE ONLY: ctypes wrapper of the CPU binary64 oracle.

Only tests/, ``__graft_entry__.smoke()`` and the cpu_baseline leg of
``bench.py`` may import this package; the product never does.
"""
from .pyoracle import Oracle, build, load  # noqa: F401
