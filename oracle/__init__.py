"""CPU oracle (NumPy) for the shear_fdtd hot path.

TEST INFRASTRUCTURE: only ``tests/``, ``__graft_entry__.smoke()`` and the
``cpu_baseline`` leg of ``bench.py`` may import this package, as the checker.
The product path never routes through it.
"""

from . import solver, strain  # noqa: F401
